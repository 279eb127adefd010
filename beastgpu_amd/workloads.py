"""Seeded synthetic problem sets: many independent assemblies of perturbed
icospheres, and a Poisson stream of assemble-and-apply requests.

The benchmark and the GPU tests draw their problems from here, so both see
the same meshes, wavenumbers and right-hand sides.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import meshes, spaces
from .operators import Maxwell3D, Operator
from .spaces import Space


@dataclasses.dataclass
class Problem:
    level: int          # icosphere subdivision level
    kind: str           # "efie" (Maxwell single layer) or "mfie" (double layer, n x K)
    amplitude: float    # radial perturbation (fraction of the radius)
    op: Operator
    X: Space


def _base(levels):
    base = {L: meshes.icosphere(L) for L in levels}
    return base, {L: spaces.raviartthomas(base[L]) for L in levels}, {L: {} for L in levels}


def batched_efie_mfie(n_prob: int = 64, seed: int = 4) -> list[Problem]:
    """``n_prob`` EFIE and MFIE (n x K) problems on RWG, alternating: icosphere
    level uniform in 2..4 (320-5,120 triangles), radial perturbation by 8
    seeded spherical harmonics with amplitude U[0, 5%], wavenumber U[0.5, 8]
    capped so that the longest edge is at most a wavelength / 8."""
    rng = np.random.default_rng(seed)
    base, bsp, cache = _base((2, 3, 4))
    out = []
    for i in range(n_prob):
        L = int(rng.integers(2, 5))
        amp = float(rng.uniform(0.0, 0.05))
        m = meshes.perturb_radial(base[L], seed=400 + i, amplitude=amp, cache=cache[L])
        k = min(float(rng.uniform(0.5, 8.0)), 2 * np.pi / (8 * m.max_edge()))
        X = Space(m, bsp[L].kind, bsp[L].idx, bsp[L].coeff, bsp[L].nfun)
        if i % 2 == 0:
            out.append(Problem(L, "efie", amp, Maxwell3D.singlelayer(wavenumber=k), X))
        else:
            out.append(Problem(L, "mfie", amp, Maxwell3D.doublelayer(wavenumber=k), X))
    return out


@dataclasses.dataclass
class StreamRequest:
    arrival_s: float    # seconds after the start of the stream
    level: int
    op: Operator
    X: Space

    def rhs(self, index: int, nrhs: int = 4) -> np.ndarray:
        """Right-hand sides of request ``index``: entries uniform in the complex
        unit disk (seeded by the request's index)."""
        r = np.random.default_rng([55, index])
        n = self.X.nfun
        return np.sqrt(r.uniform(size=(n, nrhs))) * np.exp(2j * np.pi * r.uniform(size=(n, nrhs)))


def poisson_stream(n_req: int, rate: float = 20.0, seed: int = 5) -> list[StreamRequest]:
    """Assemble-and-apply requests with Poisson arrivals (mean ``rate`` per
    second): EFIE on RWG of perturbed icospheres, level uniform in 2..5
    (log-uniform triangle count 320-20,480), amplitude U[0, 5%],
    wavenumber U[0.5, 6]."""
    rng = np.random.default_rng(seed)
    levels = rng.integers(2, 6, size=n_req)
    ks = rng.uniform(0.5, 6.0, size=n_req)
    amps = rng.uniform(0.0, 0.05, size=n_req)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    base, bsp, cache = _base((2, 3, 4, 5))
    out = []
    for i in range(n_req):
        L = int(levels[i])
        m = meshes.perturb_radial(base[L], seed=10000 + i, amplitude=float(amps[i]), cache=cache[L])
        X = Space(m, bsp[L].kind, bsp[L].idx, bsp[L].coeff, bsp[L].nfun)
        out.append(StreamRequest(float(arrivals[i]), L, Maxwell3D.singlelayer(wavenumber=float(ks[i])), X))
    return out


def unit_disk_vector(n: int, seed: int) -> np.ndarray:
    """Seeded vector with entries uniform in the complex unit disk."""
    rng = np.random.default_rng(seed)
    return np.sqrt(rng.uniform(size=n)) * np.exp(2j * np.pi * rng.uniform(size=n))
