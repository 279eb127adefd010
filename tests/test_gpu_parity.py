"""GPU parity: libbmi.so against the CPU FP64 oracle on the same inputs."""
import numpy as np
import pytest

from beastgpu_amd import meshes, spaces, Helmholtz3D, Maxwell3D
from beastgpu_amd.api import Assembler, GPUConfiguration
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12   # relative Frobenius, FP64 path (summation order differs)
MIXED_TOL = 1e-6   # relative Frobenius, mixed path


def relfro(A, B):
    return np.linalg.norm(A - B) / np.linalg.norm(B)


def _case(kind, level, k, amp=0.0, seed=1):
    m = meshes.icosphere(level)
    if amp:
        m = meshes.perturb_radial(m, seed, amp)
    if kind == "efie":
        return Maxwell3D.singlelayer(wavenumber=k), spaces.raviartthomas(m)
    return Helmholtz3D.singlelayer(wavenumber=k), spaces.lagrangec0d1(m)


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
@pytest.mark.parametrize("level,k,amp", [(1, 1.0, 0.0), (2, 2.0, 0.03), (3, 1.0, 0.0)])
def test_fp64_matches_oracle(gpu_ctx, kind, level, k, amp):
    op, X = _case(kind, level, k, amp)
    print(f"case {kind} L{level} k={k} amp={amp} nfun={X.nfun}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    Zref, counts = O.assemble(op, X, X)
    st = a.stats()
    a.close()
    err = relfro(Z, Zref)
    print("rel err", err, st)
    assert [st["pairs"][n] for n in ("far", "common_vertex", "common_edge", "common_face", "near")] == list(counts)
    assert err <= FP64_TOL


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
def test_mixed_within_tolerance(gpu_ctx, kind):
    op, X = _case(kind, 3, 2.0, 0.03, seed=7)
    print(f"case mixed {kind}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(precision="mixed"), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zref, _ = O.assemble(op, X, X)
    err = relfro(Z, Zref)
    print("mixed rel err", err)
    assert err <= MIXED_TOL
