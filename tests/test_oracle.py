"""Pinning tests of the CPU oracle (no GPU).

Each oracle function is checked against something that does not share its
code: closed-form integrals, independent high-order cubature, the analytic
potential of a triangle, mathematical invariants, or the counts the reference
recorded for its own meshes.
"""
import math

import numpy as np
import pytest

from beastgpu_amd import Helmholtz3D, Maxwell3D, meshes, spaces
from oracle import oracle as O
from tests import _indep as ind


# ------------------------------------------------------------------ rules

@pytest.mark.parametrize("order,degree", [(1, 1), (2, 2), (3, 3)])
def test_triangle_rules_exact(order, degree):
    b1, b2, w = O.far_rule(order)
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(w * b1 ** a * b2 ** b) - exact) < 1e-15
    # one degree higher is not integrated exactly by the 3- and 4-point rules
    if order in (2, 3):
        errs = [abs(np.sum(w * b1 ** a * b2 ** (degree + 1 - a)) -
                    math.factorial(a) * math.factorial(degree + 1 - a) / math.factorial(degree + 3))
                for a in range(degree + 2)]
        assert max(errs) > 1e-6


@pytest.mark.parametrize("n", [1, 2, 4, 5, 8, 12])
def test_gauss_legendre_exact(n):
    x, w = O.gauss_legendre01(n)
    xs, ws = ind.gl01(n)
    assert np.allclose(x, xs, atol=1e-15, rtol=0) and np.allclose(w, ws, atol=1e-15, rtol=0)
    for k in range(2 * n):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-14


def test_collapsed_rule_exact():
    b1, b2, w = O.collapsed_rule(4)
    for a in range(6):
        for b in range(6 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(w * b1 ** a * b2 ** b) - exact) < 1e-15


# -------------------------------------------------- polynomial kernels

def _poly_reference(op, X, Y, p, q, kernel, lin, n=10):
    """int_K int_L P(x - y) f_i(x) . g_j(y) with the high-order collapsed rule
    (exact for these polynomial integrands), independent of the oracle."""
    tv = X.mesh.vertices[X.mesh.faces[p]]
    sv = Y.mesh.vertices[Y.mesh.faces[q]]
    b1, b2, w = ind.collapsed(n)
    x = ind.cell_points(tv, b1, b2)
    y = ind.cell_points(sv, b1, b2)
    At, As = ind.area(tv), ind.area(sv)
    wx, wy = w * 2 * At, w * 2 * As
    d = x[:, None, :] - y[None, :, :]
    if kernel == O.KERNEL_ONE:
        G = np.ones(d.shape[:2])
    elif kernel == O.KERNEL_LINEAR:
        G = 1.0 + d @ np.asarray(lin)
    else:
        G = np.einsum("abk,abk->ab", d, d)
    W = wx[:, None] * wy[None, :] * G
    Z = np.zeros((3, 3), dtype=complex)
    if X.kind == spaces.P1:
        fx = np.stack([b1, b2, 1 - b1 - b2], axis=1)
        Z = op.alpha * fx.T @ W @ fx
    else:
        for i in range(3):
            fi = (x - tv[i]) / (2 * At)
            for j in range(3):
                gj = (y - sv[j]) / (2 * As)
                Z[i, j] = op.alpha * np.sum(W * (fi @ gj.T)) + op.beta * np.sum(W) / (At * As)
    return Z


def _pairs_by_class(op, X):
    cls = O.classify_all(op, X, X)
    out = {}
    for c in (O.FAR, O.CV, O.CE, O.CF):
        p, q = np.argwhere(cls == c)[0]
        out[c] = (int(p), int(q))
    return out


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
@pytest.mark.parametrize("kernel", [O.KERNEL_ONE, O.KERNEL_LINEAR, O.KERNEL_QUAD])
def test_ss_polynomial_kernels(kind, kernel):
    m = meshes.perturb_radial(meshes.icosphere(1), 5, 0.1)
    if kind == "efie":
        op, X = Maxwell3D.singlelayer(wavenumber=1.3), spaces.raviartthomas(m)
    else:
        op, X = Helmholtz3D.singlelayer(wavenumber=1.3, alpha=0.7 - 0.2j), spaces.lagrangec0d1(m)
    lin = (0.3, -1.1, 0.7)
    pairs = _pairs_by_class(op, X)
    for c in (O.CV, O.CE, O.CF):
        p, q = pairs[c]
        Zo, cls = O.pair_block(op, X, X, p, q, kernel=kernel, lin=lin, ss_order=(8, 8, 8))
        assert cls == c
        Zr = _poly_reference(op, X, X, p, q, kernel, lin)
        assert np.abs(Zo - Zr).max() <= 1e-13 * np.abs(Zr).max(), (c, Zo, Zr)


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
@pytest.mark.parametrize("kernel", [O.KERNEL_ONE, O.KERNEL_LINEAR])
def test_far_rule_polynomial_kernels(kind, kernel):
    """Far rule: linear shapes times a kernel of degree <= 1 is within the
    3-point (degree 2) and 4-point (degree 3) rules' exactness."""
    m = meshes.perturb_radial(meshes.icosphere(1), 5, 0.1)
    if kind == "efie":
        op, X = Maxwell3D.singlelayer(wavenumber=1.3), spaces.raviartthomas(m)
    else:
        op, X = Helmholtz3D.singlelayer(wavenumber=1.3), spaces.lagrangec0d1(m)
    p, q = _pairs_by_class(op, X)[O.FAR]
    lin = (0.3, -1.1, 0.7)
    Zo, cls = O.pair_block(op, X, X, p, q, kernel=kernel, lin=lin)
    assert cls == O.FAR
    Zr = _poly_reference(op, X, X, p, q, kernel, lin)
    assert np.abs(Zo - Zr).max() <= 1e-13 * np.abs(Zr).max()


def _green_reference(op, X, p, q, n=24):
    tv = X.mesh.vertices[X.mesh.faces[p]]
    sv = X.mesh.vertices[X.mesh.faces[q]]
    b1, b2, w = ind.collapsed(n)
    x, y = ind.cell_points(tv, b1, b2), ind.cell_points(sv, b1, b2)
    At, As = ind.area(tv), ind.area(sv)
    R = np.linalg.norm(x[:, None, :] - y[None, :, :], axis=2)
    W = (w * 2 * At)[:, None] * (w * 2 * As)[None, :] * np.exp(-op.gamma * R) / (4 * np.pi * R)
    Z = np.zeros((3, 3), dtype=complex)
    for i in range(3):
        fi = (x - tv[i]) / (2 * At)
        for j in range(3):
            gj = (y - sv[j]) / (2 * As)
            Z[i, j] = op.alpha * np.sum(W * (fi @ gj.T)) + op.beta * np.sum(W) / (At * As)
    return Z


def test_far_rule_vs_bruteforce():
    """Well separated pair: the 3x4 rule is within its quadrature error of a
    576x576-point reference; a wrong weight, point or shape would be O(1)."""
    m = meshes.icosphere(2)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=2.0)
    c = m.vertices[m.faces].mean(axis=1)
    p = 0
    q = int(np.argmax(np.linalg.norm(c - c[p], axis=1)))  # antipodal cell
    Zo, cls = O.pair_block(op, X, X, p, q)
    assert cls == O.FAR
    Zr = _green_reference(op, X, p, q)
    assert np.abs(Zo - Zr).max() <= 2e-4 * np.abs(Zr).max()


# ------------------------------------------------ static singular integrals

def _hh_static(m):
    X = spaces.lagrangec0d1(m)
    op = Helmholtz3D.singlelayer(gamma=0.0)
    return op, X


@pytest.mark.parametrize("cls_want", [O.CV, O.CE, O.CF])
def test_ss_static_vs_analytic(cls_want):
    """gamma = 0, P1: sum_j Z_ij = (1/4pi) int_K lambda_i(x) Phi_L(x) dx with
    Phi_L the closed-form potential of the trial cell, integrated by a
    composite rule; likewise sum_i Z_ij with the roles swapped.  This pins
    the vertex correspondence, the region maps and the Jacobians."""
    m = meshes.perturb_radial(meshes.icosphere(1), 3, 0.05)
    op, X = _hh_static(m)
    p, q = _pairs_by_class(op, X)[cls_want]
    Zo, cls = O.pair_block(op, X, X, p, q, ss_order=(8, 8, 8))
    assert cls == cls_want
    tv = m.vertices[m.faces[p]]
    sv = m.vertices[m.faces[q]]

    def outer(a, b, level):
        # composite rule; its error falls as 4^-level (log-singular derivative
        # of the potential at the shared vertex/edge), so two levels are
        # Richardson-extrapolated
        b1, b2, w = ind.subdivided(10, level)
        pts = ind.cell_points(a, b1, b2)
        phi = ind.triangle_potential(b, pts)
        lam = np.stack([b1, b2, 1 - b1 - b2], axis=1)
        return (w * 2 * ind.area(a) * phi) @ lam / (4 * np.pi)

    row = (4 * outer(tv, sv, 4) - outer(tv, sv, 3)) / 3
    col = (4 * outer(sv, tv, 4) - outer(sv, tv, 3)) / 3
    assert np.allclose(Zo.sum(axis=1).real, row, rtol=3e-8, atol=0), (Zo.sum(axis=1), row)
    assert np.allclose(Zo.sum(axis=0).real, col, rtol=3e-8, atol=0), (Zo.sum(axis=0), col)
    assert np.abs(Zo.imag).max() == 0.0


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
def test_ss_convergence(kind):
    """Exponential convergence in the Gauss-Legendre order towards n = 12."""
    m = meshes.perturb_radial(meshes.icosphere(1), 4, 0.05)
    if kind == "efie":
        op, X = Maxwell3D.singlelayer(wavenumber=3.0), spaces.raviartthomas(m)
    else:
        op, X = Helmholtz3D.singlelayer(wavenumber=3.0), spaces.lagrangec0d1(m)
    pairs = _pairs_by_class(op, X)
    for c in (O.CV, O.CE, O.CF):
        p, q = pairs[c]
        ref, _ = O.pair_block(op, X, X, p, q, ss_order=(12, 12, 12))
        errs = []
        for n in (3, 5, 8):
            Z, _ = O.pair_block(op, X, X, p, q, ss_order=(n, n, n))
            errs.append(np.abs(Z - ref).max() / np.abs(ref).max())
        assert errs[2] < errs[1] < errs[0]
        assert errs[1] < 5e-4 and errs[2] < 5e-6, errs


# ------------------------------------------------------ near pairs (class 4)

def _potential_points(v):
    nrm = np.cross(v[0] - v[2], v[1] - v[2])
    nrm = nrm / np.linalg.norm(nrm)
    c = v.mean(axis=0)
    return np.array([
        c + 0.1 * nrm,                               # above the centroid
        c,                                           # in the plane, inside
        c - 0.37 * nrm,                              # below
        0.3 * v[0] + 0.7 * v[1] + 0.4 * (v[1] - v[2]),  # in the plane, outside
        2.0 * v[1] - v[0],                           # on an edge line, outside the edge
        2.0 * v[1] - v[0] + 0.3 * nrm,               # above that point
        v[0] + np.array([0.01, 0.02, 0.03]),         # near a vertex
        v[2] + 1.5 * (v[2] - c) - 0.2 * nrm,         # beyond a vertex
    ])


def test_triangle_potentials_vs_cubature():
    """Closed-form int 1/R, int (y - xp)/R, int R and int (y - xp) R over a
    triangle against the independent closed form of 1/R (Wilton et al. 1984)
    and against a composite Duffy rule collapsed at the point's projection,
    at points above, in and beside the plane, on an edge line and near a
    vertex."""
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.3, 0.9, 0.05]])
    xs = _potential_points(v)
    I0, I1, K1, K1v = O.tri_potentials(v, xs)
    assert np.allclose(I0, ind.triangle_potential(v, xs), rtol=1e-13, atol=0)
    nrm = np.cross(v[0] - v[2], v[1] - v[2])
    nrm = nrm / np.linalg.norm(nrm)
    for k, x in enumerate(xs):
        P, W = ind.split_rule(v, x, 40)
        R = np.linalg.norm(P - x, axis=1)
        xp = x - ((x - v[0]) @ nrm) * nrm
        q0 = np.sum(W / R)
        q1 = (W / R) @ (P - xp)
        k1 = np.sum(W * R)
        k1v = (W * R) @ (P - xp)
        assert abs(I0[k] - q0) <= 1e-8 * abs(q0), (k, I0[k], q0)
        assert np.abs(I1[k] - q1).max() <= 1e-8 * np.abs(q1).max(), (k, I1[k], q1)
        assert abs(K1[k] - k1) <= 1e-8 * abs(k1), (k, K1[k], k1)
        assert np.abs(K1v[k] - k1v).max() <= 1e-8 * np.abs(k1v).max(), (k, K1v[k], k1v)
        # the vector integrals lie in the plane
        assert abs(I1[k] @ nrm) <= 1e-14 * np.abs(I1[k]).max() + 1e-300
        assert abs(K1v[k] @ nrm) <= 1e-14 * np.abs(K1v[k]).max() + 1e-300


def _shells(scale):
    """A coarse sphere and a scaled copy: cell q of the copy faces cell q of
    the sphere across a gap of (scale - 1) times the radius."""
    a = meshes.icosphere(1)
    return a, meshes.Mesh(a.vertices * scale, a.faces)


@pytest.mark.parametrize("scale", [1.02, 1.1])
def test_near_static_vs_analytic(scale):
    """gamma = 0, P1, a near pair: the row sums of the block are the outer
    rule applied to the closed-form potential of the trial cell (exact up to
    rounding), the column sums the outer rule applied to int_L lambda_j / R
    by a composite Duffy rule."""
    a, b = _shells(scale)
    X, Y = spaces.lagrangec0d1(a), spaces.lagrangec0d1(b)
    op = Helmholtz3D.singlelayer(gamma=0.0)
    tv, sv = a.vertices[a.faces[0]], b.vertices[b.faces[0]]
    Zo, cls = O.pair_block(op, X, Y, 0, 0)
    assert cls == O.NEAR
    b1, b2, w = ind.collapsed(4)
    xs = ind.cell_points(tv, b1, b2)
    wx = w * 2 * ind.area(tv)
    lam = np.stack([b1, b2, 1 - b1 - b2], axis=1)
    row = (wx * ind.triangle_potential(sv, xs)) @ lam / (4 * np.pi)
    col = np.zeros(3)
    for I, x in enumerate(xs):
        P, W = ind.split_rule(sv, x, 40)
        R = np.linalg.norm(P - x, axis=1)
        col += wx[I] * (W / R) @ ind.barycentric(sv, P) / (4 * np.pi)
    assert np.allclose(Zo.sum(axis=1).real, row, rtol=1e-13, atol=0), (Zo.sum(axis=1), row)
    assert np.allclose(Zo.sum(axis=0).real, col, rtol=1e-11, atol=0), (Zo.sum(axis=0), col)
    assert np.abs(Zo.imag).max() == 0.0


def _near_reference(op, kind, tv, sv, n, inner):
    """The near block with the outer collapsed n^2 rule on the test cell and
    the inner integral by the composite Duffy rule (inner = "split") or by the
    plain collapsed n^2 rule (inner = "plain"), in numpy."""
    b1, b2, w = ind.collapsed(n)
    xs = ind.cell_points(tv, b1, b2)
    wx = w * 2 * ind.area(tv)
    At, As = ind.area(tv), ind.area(sv)
    Z = np.zeros((3, 3), dtype=complex)
    for I, x in enumerate(xs):
        if inner == "split":
            P, W = ind.split_rule(sv, x, 40)
        else:
            P, W = ind.cell_points(sv, b1, b2), w * 2 * As
        R = np.linalg.norm(P - x, axis=1)
        G = W * np.exp(-op.gamma * R) / (4 * np.pi * R)
        if kind == "hh3d":
            fx = np.array([b1[I], b2[I], 1 - b1[I] - b2[I]])
            Z += wx[I] * op.alpha * np.outer(fx, G @ ind.barycentric(sv, P))
        else:
            for i in range(3):
                fi = (x - tv[i]) / (2 * At)
                for j in range(3):
                    gj = (P - sv[j]) / (2 * As)
                    Z[i, j] += wx[I] * (op.alpha * (G @ (gj @ fi)) + op.beta * G.sum() / (At * As))
    return Z


@pytest.mark.parametrize("kind", ["efie", "efie-lossy", "hh3d"])
def test_near_vs_bruteforce(kind):
    """Near blocks against the same outer rule with the inner integral by a
    composite Duffy rule: the subtraction leaves only the smooth remainder to
    the inner rule, so the error is small at near_order 4 and falls fast with
    the order, where the plain product rule is off by tens of percent for the
    facing cells."""
    a, b = _shells(1.02)
    if kind == "efie":
        op, X, Y = Maxwell3D.singlelayer(wavenumber=3.0), spaces.raviartthomas(a), spaces.raviartthomas(b)
    elif kind == "efie-lossy":
        op, X, Y = Maxwell3D.singlelayer(gamma=0.8 + 3.0j), spaces.raviartthomas(a), spaces.raviartthomas(b)
    else:
        op = Helmholtz3D.singlelayer(wavenumber=3.0, alpha=0.7 - 0.2j)
        X, Y = spaces.lagrangec0d1(a), spaces.lagrangec0d1(b)
    k = "hh3d" if kind == "hh3d" else "efie"
    tv = a.vertices[a.faces[0]]
    for q in (0, 1, 5):
        sv = b.vertices[b.faces[q]]
        for n, tol in ((4, 2e-5), (8, 1e-6)):
            Zo, cls = O.pair_block(op, X, Y, 0, q, near_order=n)
            assert cls == O.NEAR
            Zr = _near_reference(op, k, tv, sv, n, "split")
            err = np.abs(Zo - Zr).max() / np.abs(Zr).max()
            assert err <= tol, (q, n, err)
            if q == 0 and n == 4:
                Zp = _near_reference(op, k, tv, sv, n, "plain")
                assert np.abs(Zp - Zr).max() / np.abs(Zr).max() > 1e4 * err


# ------------------------------------------------------------- classifier

def test_classifier_euler_counts():
    for level in (1, 2, 3):
        m = meshes.icosphere(level)
        X = spaces.raviartthomas(m)
        op = Maxwell3D.singlelayer(wavenumber=1.0)
        cls = O.classify_all(op, X, X)
        deg = m.vertex_degrees().astype(np.int64)
        assert (cls == O.CF).sum() == m.nt
        assert (cls == O.CE).sum() == 3 * m.nt
        assert (cls == O.CV).sum() == int((deg ** 2).sum() - 9 * m.nt)
        assert (cls == O.NEAR).sum() == 0


@pytest.mark.slow
def test_classifier_reference_meshcuboid_counts():
    """Touching-pair counts the reference recorded for meshcuboid(1,1,1,0.5/24)
    (data/GPU/Saut-Swab.txt:50-55): CV 248,784, CE 82,944, CF 27,648."""
    m = meshes.meshcuboid(1.0, 1.0, 1.0, 0.5 / 24)
    X = spaces.lagrangec0d1(m)
    op = Helmholtz3D.singlelayer(wavenumber=1.0)
    cls = O.classify_all(op, X, X)
    assert (cls == O.CF).sum() == 27648
    assert (cls == O.CE).sum() == 82944
    assert (cls == O.CV).sum() == 248784
    assert (cls == O.NEAR).sum() == 0


def test_classifier_near_bruteforce():
    """Two parallel, slightly offset copies of a coarse sphere give near pairs;
    compare with a direct evaluation of max(d^2 k^2, d^2/(16 A)) < 0.2^2."""
    a = meshes.icosphere(1)
    b = meshes.Mesh(a.vertices * 1.02 + np.array([0.0, 0.0, 0.001]), a.faces)
    X = spaces.lagrangec0d1(a)
    Y = spaces.lagrangec0d1(b)
    for k in (0.5, 4.0):
        op = Helmholtz3D.singlelayer(wavenumber=k)
        cls = O.classify_all(op, X, Y)
        tv = a.vertices[a.faces]
        sv = b.vertices[b.faces]
        d2 = ((tv[:, None, :, None, :] - sv[None, :, None, :, :]) ** 2).sum(-1).reshape(a.nt, b.nt, 9).min(-1)
        A = b.areas()[None, :]
        near = np.maximum(d2 * k * k, d2 / (16 * A)) < 0.04
        assert near.any()
        assert np.array_equal(cls == O.NEAR, near)


# ---------------------------------------------------------------- scatter

def test_scatter_closed_form_constant_kernel():
    """G = 1: Z_mn = alpha (int f_m).(int f_n) + beta (int div f_m)(int div f_n),
    with int_K f_i = (c_K - v_i)/2 and int div f_i = 1 per cell."""
    m = meshes.perturb_radial(meshes.icosphere(2), 9, 0.05)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=1.0)
    Z, _ = O.assemble(op, X, X, kernel=O.KERNEL_ONE)
    v = m.vertices[m.faces]
    c = v.mean(axis=1)
    u = np.zeros((X.nfun, 3))
    dv = np.zeros(X.nfun)
    for K in range(m.nt):
        for i in range(3):
            f = X.idx[K, i]
            if f >= 0:
                u[f] += X.coeff[K, i] * (c[K] - v[K, i]) / 2
                dv[f] += X.coeff[K, i]
    Zr = op.alpha * (u @ u.T) + op.beta * np.outer(dv, dv)
    # the beta terms of the two cells of an edge cancel: scale the bound by |beta|
    assert np.abs(Z - Zr).max() <= 1e-13 * (np.abs(Zr).max() + abs(op.beta))
    assert np.abs(dv).max() < 1e-12  # RWG divergence integrates to zero


def test_scatter_renumbering_invariance():
    m = meshes.perturb_radial(meshes.icosphere(1), 2, 0.05)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=1.5)
    Z, _ = O.assemble(op, X, X)
    rng = np.random.default_rng(0)
    pc = rng.permutation(m.nt)
    pf = rng.permutation(X.nfun)
    m2 = meshes.Mesh(m.vertices, m.faces[pc])
    inv = np.empty_like(pf)
    inv[pf] = np.arange(X.nfun)
    idx2 = np.where(X.idx[pc] >= 0, inv[np.maximum(X.idx[pc], 0)], -1)
    X2 = spaces.Space(m2, spaces.RWG, idx2, X.coeff[pc], X.nfun)
    Z2, _ = O.assemble(op, X2, X2)
    assert np.abs(Z2 - Z[np.ix_(pf, pf)]).max() <= 1e-14 * np.abs(Z).max()


def test_rows_subset_matches_full():
    m = meshes.icosphere(2)
    X = spaces.lagrangec0d1(m)
    op = Helmholtz3D.singlelayer(wavenumber=2.0)
    Z, _ = O.assemble(op, X, X)
    rows = np.array([3, 77, 12, 150])
    Zr, _ = O.assemble(op, X, X, rows=rows)
    # same terms; the colouring (hence the summation order) differs
    assert np.abs(Zr - Z[rows]).max() <= 1e-15 * np.abs(Z).max()


# ------------------------------------------------- double layer (n x K)

def _dl_poly_reference(op, X, p, q, kernel, lin, n=10):
    """alpha int_K int_L (n_K x f_i) . (P(x - y) x g_j) with the high-order
    collapsed rule, P the polynomial stand-in of grad G (exact here)."""
    tv = X.mesh.vertices[X.mesh.faces[p]]
    sv = X.mesh.vertices[X.mesh.faces[q]]
    b1, b2, w = ind.collapsed(n)
    x, y = ind.cell_points(tv, b1, b2), ind.cell_points(sv, b1, b2)
    At, As = ind.area(tv), ind.area(sv)
    nrm = np.cross(tv[0] - tv[2], tv[1] - tv[2])
    nrm /= np.linalg.norm(nrm)
    d = x[:, None, :] - y[None, :, :]
    lin = np.asarray(lin)
    if kernel == O.KERNEL_ONE:
        P = np.broadcast_to(lin, d.shape)
    elif kernel == O.KERNEL_LINEAR:
        P = lin + d
    else:
        P = np.einsum("abk,abk->ab", d, d)[..., None] * lin
    W = (w * 2 * At)[:, None] * (w * 2 * As)[None, :]
    Z = np.zeros((3, 3), dtype=complex)
    for i in range(3):
        u = np.cross(nrm, (x - tv[i]) / (2 * At))           # (na, 3)
        for j in range(3):
            g = (y - sv[j]) / (2 * As)                       # (nb, 3)
            c = np.cross(P, np.broadcast_to(g[None, :, :], P.shape))
            Z[i, j] = op.alpha * np.sum(W * np.einsum("ak,abk->ab", u, c))
    return Z


@pytest.mark.parametrize("kernel", [O.KERNEL_ONE, O.KERNEL_LINEAR, O.KERNEL_QUAD])
def test_dl_ss_polynomial_kernels(kernel):
    m = meshes.perturb_radial(meshes.icosphere(1), 5, 0.1)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.doublelayer(wavenumber=1.3, alpha=0.8 + 0.3j)
    lin = (0.3, -1.1, 0.7)
    pairs = _pairs_by_class(op, X)
    for c in (O.CV, O.CE, O.CF, O.FAR):
        p, q = pairs[c]
        Zo, cls = O.pair_block(op, X, X, p, q, kernel=kernel, lin=lin, ss_order=(8, 8, 8))
        assert cls == c
        if c == O.FAR and kernel == O.KERNEL_QUAD:
            continue  # degree 4 exceeds the 3-point rule
        Zr = _dl_poly_reference(op, X, p, q, kernel, lin)
        assert np.abs(Zo - Zr).max() <= 1e-13 * np.abs(Zr).max(), (c, Zo, Zr)


def test_dl_closed_form_constant_gradient():
    """grad G -> c: Z_mn = alpha (int n x f_m) . (c x int f_n), int_K f_i = (c_K - v_i)/2."""
    m = meshes.perturb_radial(meshes.icosphere(2), 9, 0.05)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.doublelayer(wavenumber=1.0)
    cvec = np.array([0.4, -0.9, 1.3])
    Z, _ = O.assemble(op, X, X, kernel=O.KERNEL_ONE, lin=tuple(cvec))
    v = m.vertices[m.faces]
    cen = v.mean(axis=1)
    nrm = np.cross(v[:, 0] - v[:, 2], v[:, 1] - v[:, 2])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    a = np.zeros((X.nfun, 3))
    b = np.zeros((X.nfun, 3))
    for K in range(m.nt):
        for i in range(3):
            f = X.idx[K, i]
            if f >= 0:
                fi = X.coeff[K, i] * (cen[K] - v[K, i]) / 2
                a[f] += np.cross(nrm[K], fi)
                b[f] += np.cross(cvec, fi)
    Zr = op.alpha * (a @ b.T)
    assert np.abs(Z - Zr).max() <= 1e-13 * np.abs(Zr).max()


def test_dl_gradient_factor_by_finite_differences():
    """Centroid rules on a far pair: Z_ij = A_K A_L (n x f_i(c_K)) . (grad G(c_K - c_L) x g_j(c_L)),
    grad G by central differences of exp(-gamma R)/(4 pi R)."""
    m = meshes.perturb_radial(meshes.icosphere(2), 4, 0.05)
    X = spaces.raviartthomas(m)
    g = complex(0.3, 2.1)
    op = Maxwell3D.doublelayer(gamma=g)
    p, q = _pairs_by_class(op, X)[O.FAR]
    Zo, cls = O.pair_block(op, X, X, p, q, far_test=1, far_trial=1)
    assert cls == O.FAR
    tv = m.vertices[m.faces[p]]
    sv = m.vertices[m.faces[q]]
    cK, cL = tv.mean(axis=0), sv.mean(axis=0)
    AK, AL = ind.area(tv), ind.area(sv)
    nrm = np.cross(tv[0] - tv[2], tv[1] - tv[2])
    nrm /= np.linalg.norm(nrm)

    def G(r):
        R = np.linalg.norm(r)
        return np.exp(-g * R) / (4 * np.pi * R)

    h = 1e-5
    r0 = cK - cL
    grad = np.array([(G(r0 + h * e) - G(r0 - h * e)) / (2 * h) for e in np.eye(3)])
    Zr = np.zeros((3, 3), dtype=complex)
    for i in range(3):
        u = np.cross(nrm, (cK - tv[i]) / (2 * AK))
        for j in range(3):
            w_ = (cL - sv[j]) / (2 * AL)
            Zr[i, j] = AK * AL * np.dot(u, np.cross(grad, w_))
    assert np.abs(Zo - Zr).max() <= 1e-8 * np.abs(Zr).max()


def test_dl_vanishes_on_a_flat_sheet():
    """Coplanar cells: grad G x f_n is normal to the plane, n x f_m lies in it."""
    nx = 6
    xs = np.linspace(0.0, 1.0, nx + 1)
    V = np.array([(x, y, 0.0) for y in xs for x in xs])
    F = []
    for j in range(nx):
        for i in range(nx):
            a, b, c, d = j * (nx + 1) + i, j * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i, (j + 1) * (nx + 1) + i + 1
            F += [(a, b, d), (a, d, c)]
    m = meshes.Mesh(V, np.array(F, dtype=np.int32))
    X = spaces.raviartthomas(m)
    op = Maxwell3D.doublelayer(wavenumber=3.0)
    Z, counts = O.assemble(op, X, X)
    assert counts[1] > 0 and counts[2] > 0 and counts[3] > 0
    assert np.abs(Z).max() <= 1e-14


def test_dl_ss_convergence():
    m = meshes.perturb_radial(meshes.icosphere(1), 3, 0.1)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.doublelayer(wavenumber=2.0)
    pairs = _pairs_by_class(op, X)
    far, _ = O.pair_block(op, X, X, *pairs[O.FAR])
    for c in (O.CV, O.CE):
        p, q = pairs[c]
        ref, _ = O.pair_block(op, X, X, p, q, ss_order=(14, 14, 14))
        e5 = np.abs(O.pair_block(op, X, X, p, q, ss_order=(5, 5, 5))[0] - ref).max()
        e8 = np.abs(O.pair_block(op, X, X, p, q, ss_order=(8, 8, 8))[0] - ref).max()
        scale = np.abs(ref).max()
        assert e8 <= 1e-6 * scale and e8 < e5, (c, e5, e8, scale)
    # a cell against itself is coplanar: its block vanishes
    cf, _ = O.pair_block(op, X, X, *pairs[O.CF])
    assert np.abs(cf).max() <= 1e-14 * np.abs(far).max() * 1e3


# ------------------------------------------------------ caller rule tables

def _table(test_order, trial_order, gl):
    """Quadrature tables from the oracle's own rules (for the identity check)."""
    from beastgpu_amd.api import QuadratureRules

    tb1, tb2, tw = O.far_rule(test_order)
    bb1, bb2, bw = O.far_rule(trial_order)
    return QuadratureRules(np.stack([tb1, tb2], 1), tw, np.stack([bb1, bb2], 1), bw,
                           tuple(O.gauss_legendre01(n) for n in gl))


def _edge_midpoint_tables(gl=(6, 4, 7)):
    """Test cells: the edge-midpoint rule (degree 2; not the interior 3-point
    rule of the defaults).  Trial cells: the 2 x 2 collapsed Gauss product
    rule (degree 3, no symmetry).  Gauss-Legendre from numpy's nodes."""
    from beastgpu_amd.api import QuadratureRules

    cb1, cb2, cw = O.collapsed_rule(2)
    gls = []
    for n in gl:
        x, w = np.polynomial.legendre.leggauss(n)
        gls.append((0.5 * (x + 1.0), 0.5 * w))
    return QuadratureRules([[0.5, 0.5], [0.0, 0.5], [0.5, 0.0]], [1.0 / 6] * 3, np.stack([cb1, cb2], 1), cw,
                           tuple(gls))


def test_rule_tables_equal_to_the_builtin_rules_give_the_same_blocks():
    m = meshes.perturb_radial(meshes.icosphere(1), 5, 0.1)
    op, X = Maxwell3D.singlelayer(wavenumber=1.3), spaces.raviartthomas(m)
    rules = _table(2, 3, (5, 6, 4))
    for c, (p, q) in _pairs_by_class(op, X).items():
        Zd, cd = O.pair_block(op, X, X, p, q, ss_order=(5, 6, 4))
        Zt, ct = O.pair_block(op, X, X, p, q, rules=rules)
        assert cd == ct == c and np.array_equal(Zd, Zt), c


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
@pytest.mark.parametrize("kernel", [O.KERNEL_ONE, O.KERNEL_LINEAR])
def test_rule_tables_far_pair_polynomial_kernels(kind, kernel):
    """The caller's far tables (edge midpoints on the test cell, a collapsed
    product rule on the trial cell) integrate linear shapes times a kernel of
    degree <= 1 exactly: against the closed form."""
    m = meshes.perturb_radial(meshes.icosphere(1), 5, 0.1)
    if kind == "efie":
        op, X = Maxwell3D.singlelayer(wavenumber=1.3), spaces.raviartthomas(m)
    else:
        op, X = Helmholtz3D.singlelayer(wavenumber=1.3), spaces.lagrangec0d1(m)
    p, q = _pairs_by_class(op, X)[O.FAR]
    lin = (0.3, -1.1, 0.7)
    Zo, cls = O.pair_block(op, X, X, p, q, kernel=kernel, lin=lin, rules=_edge_midpoint_tables())
    assert cls == O.FAR
    Zr = _poly_reference(op, X, X, p, q, kernel, lin)
    assert np.abs(Zo - Zr).max() <= 1e-13 * np.abs(Zr).max()
    # a mistake in the table's weights or points shows (the rule is not the default)
    Zd, _ = O.pair_block(op, X, X, p, q, kernel=O.KERNEL_GREEN)
    Zt, _ = O.pair_block(op, X, X, p, q, kernel=O.KERNEL_GREEN, rules=_edge_midpoint_tables())
    assert np.abs(Zd - Zt).max() > 1e-9 * np.abs(Zd).max()


def test_rule_tables_gauss_legendre_match_the_builtin_orders():
    """numpy's Gauss-Legendre nodes as caller tables reproduce the oracle's
    own n-point rules in the Sauter-Schwab integrals to rounding."""
    m = meshes.perturb_radial(meshes.icosphere(1), 5, 0.1)
    op, X = Maxwell3D.singlelayer(wavenumber=1.3), spaces.raviartthomas(m)
    t = _edge_midpoint_tables(gl=(6, 4, 7))
    for c, (p, q) in _pairs_by_class(op, X).items():
        if c == O.FAR:
            continue
        Zd, _ = O.pair_block(op, X, X, p, q, ss_order=(6, 4, 7))
        Zt, _ = O.pair_block(op, X, X, p, q, rules=t)
        assert np.abs(Zd - Zt).max() <= 1e-14 * np.abs(Zd).max(), c
