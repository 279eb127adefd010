"""Boundary integral operators, with BEAST's parameter conventions.

Green's function G(R) = exp(-gamma R) / (4 pi R), gamma = i k for a real
wavenumber k (the reference's kernels use exp(-R*gamma) with biop.gamma,
doubleQuadRuleMainCase.jl:143-146).

* ``Helmholtz3D.singlelayer(wavenumber=k, alpha=1)``: HH3D single layer,
  Z_mn = alpha <f_m, G f_n>; the operator the reference assembles
  (run.jl:49-52, struct at assamble_chunk_gpu.jl:44-47).
* ``Maxwell3D.singlelayer(wavenumber=k)``: EFIE single layer,
  Z_mn = alpha <f_m, G f_n> + beta <div f_m, G div f_n> with BEAST's defaults
  alpha = -gamma = -ik and beta = -1/gamma = 1/(-ik).
* ``Maxwell3D.doublelayer(wavenumber=k)``: the double layer K tested with
  rotated RWG functions, the n x K block of the MFIE (BEAST's
  ``assemble(K, n x X, X)``; the identity part 1/2 <n x f_m, f_n> is a sparse
  local Gram matrix outside this path):
  Z_mn = alpha <n x f_m, grad_x G x f_n>, alpha = 1 by default.
"""
from __future__ import annotations

import dataclasses

HH3D_SINGLELAYER = 1
MW_SINGLELAYER = 2
MW_DOUBLELAYER = 3


@dataclasses.dataclass(frozen=True)
class Operator:
    kind: int
    alpha: complex
    beta: complex
    gamma: complex

    @property
    def wavenumber(self) -> float:
        return float(self.gamma.imag)


def _gamma(gamma, wavenumber):
    if (gamma is None) == (wavenumber is None):
        raise ValueError("give exactly one of gamma or wavenumber")
    return complex(gamma) if gamma is not None else complex(0.0, float(wavenumber))


class Helmholtz3D:
    @staticmethod
    def singlelayer(*, wavenumber=None, gamma=None, alpha=None) -> Operator:
        g = _gamma(gamma, wavenumber)
        a = complex(1.0 if alpha is None else alpha)
        return Operator(HH3D_SINGLELAYER, a, 0j, g)


class Maxwell3D:
    @staticmethod
    def singlelayer(*, wavenumber=None, gamma=None, alpha=None, beta=None) -> Operator:
        g = _gamma(gamma, wavenumber)
        if g == 0:
            raise ValueError("the Maxwell single layer needs gamma != 0 (beta = -1/gamma)")
        a = complex(-g if alpha is None else alpha)
        b = complex(-1.0 / g if beta is None else beta)
        return Operator(MW_SINGLELAYER, a, b, g)

    @staticmethod
    def doublelayer(*, wavenumber=None, gamma=None, alpha=None) -> Operator:
        g = _gamma(gamma, wavenumber)
        return Operator(MW_DOUBLELAYER, complex(1.0 if alpha is None else alpha), 0j, g)
