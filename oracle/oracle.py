"""ctypes wrapper of the CPU FP64 oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the CPU-baseline leg of bench.py, always as the checker.  The product
library never loads it.

Parity status of each oracle function, and what pins it (tests named):

* far-pair product rule (doubleQuadRuleMainCase.jl:83-103) -- pinned by exact
  polynomial kernels against closed-form integrals computed independently in
  numpy (tests/test_oracle.py::test_far_rule_polynomial_kernels), by rule
  exactness (test_triangle_rules_exact) and by convergence to a high-order
  brute-force cubature for well separated pairs (test_far_rule_vs_bruteforce).
* Sauter-Schwab CV/CE/CF (SauterSchwab.jl:265-298, pulledback_integrand_gpu.jl)
  -- pinned by exact polynomial kernels (test_ss_polynomial_kernels), by the
  static 1/R integrals against the analytic potential of a triangle (Wilton et
  al. 1984) plus an outer Gauss rule (test_ss_static_vs_analytic), and by
  exponential convergence in the Gauss-Legendre order (test_ss_convergence).
* classifier (quadrule_determine.jl:54-77) -- pinned by the touching-pair
  counts the reference recorded for meshcuboid(1,1,1,0.5/24)
  (data/GPU/Saut-Swab.txt:50-55), by Euler closed forms on closed meshes and by
  a brute-force near test (test_classifier_*).
* scatter (store_with_kernel.jl:88-97) -- pinned by renumbering invariance and
  by the G = 1 closed form for whole matrices (test_scatter_*).
* double layer with rotated test functions (the n x K block of the MFIE; the
  reference has no such operator) -- pinned by polynomial stand-ins of grad G
  for every pair class against numpy cubature (test_dl_ss_polynomial_kernels),
  the constant-gradient whole-matrix closed form
  (test_dl_closed_form_constant_gradient), the gradient factor against central
  differences of G (test_dl_gradient_factor_by_finite_differences), the
  vanishing of the operator on a flat sheet (test_dl_vanishes_on_a_flat_sheet)
  and convergence in the Sauter-Schwab order (test_dl_ss_convergence).  BEAST's
  own MWDoubleLayer3D integrand is not importable here: its sign and
  normalisation conventions are parity unpinned.
* near pairs (class 4, which the reference drops: doubleQuadRuleMainCase.jl:154)
  of the single layers by singularity subtraction -- the closed-form triangle
  integrals of 1/R, (y - xp)/R, R and (y - xp) R are pinned against the
  independent closed form of Wilton et al. and a composite Duffy cubature at
  points above, in and beside the plane (test_triangle_potentials_vs_cubature);
  the static P1 block's row and column sums against the closed-form potential
  and cubature (test_near_static_vs_analytic); whole blocks, lossless and
  lossy, RWG and P1, against the same outer rule with a composite Duffy inner
  integral, with the error falling in the order (test_near_vs_bruteforce).
* BEAST conventions that only an import of BEAST 2.6.0 / CompScienceMeshes
  0.9.3 / SauterSchwabQuadrature 2.4.1 could fix (triangle-rule tables and
  point order, RT orientation signs and edge numbering, the exact orders and
  expansion terms of BEAST's WiltonSE near rule) are parity unpinned; the
  symmetric rules used here make the point order irrelevant up to rounding.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

KERNEL_GREEN, KERNEL_ONE, KERNEL_LINEAR, KERNEL_QUAD = 0, 1, 2, 3
FAR, CV, CE, CF, NEAR = 0, 1, 2, 3, 4


class _Op(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32),
        ("kernel", ctypes.c_int32),
        ("alpha", ctypes.c_double * 2),
        ("beta", ctypes.c_double * 2),
        ("gamma", ctypes.c_double * 2),
        ("lin", ctypes.c_double * 3),
    ]


class _Space(ctypes.Structure):
    _fields_ = [
        ("xyz", ctypes.c_void_p),
        ("nv", ctypes.c_int64),
        ("tri", ctypes.c_void_p),
        ("nt", ctypes.c_int64),
        ("kind", ctypes.c_int32),
        ("idx", ctypes.c_void_p),
        ("coeff", ctypes.c_void_p),
        ("nfun", ctypes.c_int64),
    ]


class _Quad(ctypes.Structure):
    _fields_ = [
        ("far_test", ctypes.c_int32),
        ("far_trial", ctypes.c_int32),
        ("ss_order", ctypes.c_int32 * 3),
        ("near_order", ctypes.c_int32),
        ("custom", ctypes.c_int32),
        ("n_test", ctypes.c_int32),
        ("n_trial", ctypes.c_int32),
        ("test_uv", (ctypes.c_double * 2) * 16),
        ("test_w", ctypes.c_double * 16),
        ("trial_uv", (ctypes.c_double * 2) * 16),
        ("trial_w", ctypes.c_double * 16),
        ("gl_n", ctypes.c_int32 * 3),
        ("gl_x", (ctypes.c_double * 64) * 3),
        ("gl_w", (ctypes.c_double * 64) * 3),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_assemble.argtypes = [ctypes.POINTER(_Op), ctypes.POINTER(_Space), ctypes.POINTER(_Space),
                                   ctypes.POINTER(_Quad), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                   ctypes.c_int, ctypes.c_void_p]
        L.orc_assemble.restype = ctypes.c_int
        L.orc_pair_block.argtypes = [ctypes.POINTER(_Op), ctypes.POINTER(_Space), ctypes.POINTER(_Space),
                                     ctypes.POINTER(_Quad), ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
        L.orc_pair_block.restype = ctypes.c_int
        L.orc_classify_all.argtypes = [ctypes.POINTER(_Op), ctypes.POINTER(_Space), ctypes.POINTER(_Space),
                                       ctypes.c_void_p, ctypes.c_int]
        L.orc_far_rule.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_collapsed_rule.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_gauss_legendre01.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_gauss_legendre01.restype = None
        L.orc_tri_potentials.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.orc_tri_potentials.restype = None
        _lib = L
    return _lib


def _op(op, kernel=KERNEL_GREEN, lin=(0.0, 0.0, 0.0)) -> _Op:
    o = _Op()
    o.op = int(op.kind)
    o.kernel = int(kernel)
    for name in ("alpha", "beta", "gamma"):
        z = complex(getattr(op, name))
        getattr(o, name)[0] = z.real
        getattr(o, name)[1] = z.imag
    for i in range(3):
        o.lin[i] = float(lin[i])
    return o


def _space(X, keep):
    m = X.mesh
    keep.extend([m.vertices, m.faces, X.idx, X.coeff])
    s = _Space()
    s.xyz = m.vertices.ctypes.data
    s.nv = m.nv
    s.tri = m.faces.ctypes.data
    s.nt = m.nt
    s.kind = int(X.kind)
    s.idx = X.idx.ctypes.data
    s.coeff = X.coeff.ctypes.data
    s.nfun = int(X.nfun)
    return s


def _quad(far_test=2, far_trial=3, ss_order=(5, 5, 5), near_order=4, rules=None) -> _Quad:
    """rules: optional tables (test_uv (n, 2), test_weights (n,), trial_uv,
    trial_weights, gauss_legendre = three (nodes, weights) pairs on [0, 1] for
    CV, CE, CF), given as an object with those attributes."""
    q = _Quad()
    q.far_test = far_test
    q.far_trial = far_trial
    for i in range(3):
        q.ss_order[i] = int(ss_order[i])
    q.near_order = near_order
    if rules is not None:
        q.custom = 1
        tu, tw = np.asarray(rules.test_uv, float).reshape(-1, 2), np.asarray(rules.test_weights, float).ravel()
        bu, bw = np.asarray(rules.trial_uv, float).reshape(-1, 2), np.asarray(rules.trial_weights, float).ravel()
        q.n_test, q.n_trial = tu.shape[0], bu.shape[0]
        for i in range(q.n_test):
            q.test_uv[i][0], q.test_uv[i][1], q.test_w[i] = tu[i, 0], tu[i, 1], tw[i]
        for i in range(q.n_trial):
            q.trial_uv[i][0], q.trial_uv[i][1], q.trial_w[i] = bu[i, 0], bu[i, 1], bw[i]
        for g, (x, w) in enumerate(rules.gauss_legendre):
            x, w = np.asarray(x, float).ravel(), np.asarray(w, float).ravel()
            q.gl_n[g] = x.size
            for i in range(x.size):
                q.gl_x[g][i], q.gl_w[g][i] = x[i], w[i]
    return q


def assemble(op, X, Y, *, rows=None, ss_order=(5, 5, 5), far_test=2, far_trial=3, near_order=4,
             kernel=KERNEL_GREEN, lin=(0.0, 0.0, 0.0), nthreads=0, rules=None):
    """Z (complex128, rows x Y.nfun) and per-class pair counts (far, CV, CE, CF, near);
    rules: caller quadrature tables in place of far_test / far_trial / ss_order."""
    keep: list = []
    o, sx, sy = _op(op, kernel, lin), _space(X, keep), _space(Y, keep)
    q = _quad(far_test, far_trial, ss_order, near_order, rules)
    if rows is not None:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        nout = rows.size
    else:
        nout = X.nfun
    Z = np.zeros((Y.nfun, nout), dtype=np.complex128)  # column-major nout x nfun
    counts = np.zeros(5, dtype=np.int64)
    rc = lib().orc_assemble(ctypes.byref(o), ctypes.byref(sx), ctypes.byref(sy), ctypes.byref(q),
                            rows.ctypes.data if rows is not None else None, nout, Z.ctypes.data,
                            int(nthreads), counts.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle assembly failed")
    return Z.T, counts


def pair_block(op, X, Y, p, q, *, ss_order=(5, 5, 5), far_test=2, far_trial=3, near_order=4,
               kernel=KERNEL_GREEN, lin=(0.0, 0.0, 0.0), rules=None):
    """Local 3x3 element matrix (before coefficients) of test cell p, trial cell q."""
    keep: list = []
    o, sx, sy = _op(op, kernel, lin), _space(X, keep), _space(Y, keep)
    qd = _quad(far_test, far_trial, ss_order, near_order, rules)
    out = np.zeros(18)
    cls = lib().orc_pair_block(ctypes.byref(o), ctypes.byref(sx), ctypes.byref(sy), ctypes.byref(qd),
                               int(p), int(q), out.ctypes.data)
    if cls < 0:
        raise RuntimeError("oracle pair block failed")
    return (out[0::2] + 1j * out[1::2]).reshape(3, 3), cls


def classify_all(op, X, Y, nthreads=0):
    keep: list = []
    o, sx, sy = _op(op), _space(X, keep), _space(Y, keep)
    out = np.zeros((X.mesh.nt, Y.mesh.nt), dtype=np.int8)
    lib().orc_classify_all(ctypes.byref(o), ctypes.byref(sx), ctypes.byref(sy), out.ctypes.data, int(nthreads))
    return out


def far_rule(order):
    b1, b2, w = np.zeros(512), np.zeros(512), np.zeros(512)
    n = lib().orc_far_rule(order, b1.ctypes.data, b2.ctypes.data, w.ctypes.data)
    if n < 0:
        raise ValueError(order)
    return b1[:n], b2[:n], w[:n]


def collapsed_rule(order):
    b1, b2, w = np.zeros(512), np.zeros(512), np.zeros(512)
    n = lib().orc_collapsed_rule(order, b1.ctypes.data, b2.ctypes.data, w.ctypes.data)
    return b1[:n], b2[:n], w[:n]


def gauss_legendre01(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss_legendre01(n, x.ctypes.data, w.ctypes.data)
    return x, w


def tri_potentials(v, x):
    """Closed-form integrals over the flat triangle v (3, 3) at the points x
    (m, 3), R = |x - y|, xp the projection of x onto the triangle's plane:
    returns (int 1/R, int (y - xp)/R, int R, int (y - xp) R)."""
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(3, 3)
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
    out = np.zeros((x.shape[0], 8))
    lib().orc_tri_potentials(v.ctypes.data, x.ctypes.data, x.shape[0], out.ctypes.data)
    return out[:, 0], out[:, 1:4], out[:, 4], out[:, 5:8]
