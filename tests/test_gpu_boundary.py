"""The drop-in boundary beyond the defaults: caller-supplied quadrature tables
(bmi_plan_create_rules), rectangular matrices with test and trial spaces on
different meshes, and the refusal of meshes that overlap."""
import numpy as np
import pytest

from beastgpu_amd import Helmholtz3D, Maxwell3D, _bmi, meshes, spaces
from beastgpu_amd.api import Assembler, GPUConfiguration, QuadratureRules
from oracle import oracle as O
from tests.test_oracle import _edge_midpoint_tables

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-6


def relfro(A, B):
    return np.linalg.norm(A - B) / np.linalg.norm(B)


def _op_space(kind, m, k=1.9):
    if kind == "efie":
        return Maxwell3D.singlelayer(wavenumber=k), spaces.raviartthomas(m)
    if kind == "mfie":
        return Maxwell3D.doublelayer(wavenumber=k), spaces.raviartthomas(m)
    return Helmholtz3D.singlelayer(wavenumber=k), spaces.lagrangec0d1(m)


@pytest.mark.parametrize("kind", ["efie", "hh3d", "mfie"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_caller_rule_tables_match_oracle(gpu_ctx, kind, precision):
    """Edge-midpoint test rule, collapsed 2 x 2 trial rule, Gauss-Legendre
    with 6 / 4 / 7 points: the device and the oracle fed the same tables."""
    m = meshes.perturb_radial(meshes.icosphere(2), 17, 0.03)
    op, X = _op_space(kind, m)
    rules = _edge_midpoint_tables(gl=(6, 4, 7))
    print(f"case rules {kind} {precision}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(precision=precision, rules=rules), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zref, _ = O.assemble(op, X, X, rules=rules)
    err = relfro(Z, Zref)
    print("rules rel err", err)
    assert err <= (MIXED_TOL if precision == "mixed" else FP64_TOL)
    # the tables are used: the default rules give another matrix
    Zdef, _ = O.assemble(op, X, X)
    assert relfro(Zdef, Zref) > 1e-6


def test_default_tables_run_the_default_path(gpu_ctx):
    """Tables equal to the default rules (3 x 4 points, 5-point Gauss-Legendre)
    give bitwise the matrix of the default plan."""
    m = meshes.perturb_radial(meshes.icosphere(3), 4, 0.03)
    op, X = _op_space("efie", m)
    tb1, tb2, tw = O.far_rule(2)
    bb1, bb2, bw = O.far_rule(3)
    gl = O.gauss_legendre01(5)
    rules = QuadratureRules(np.stack([tb1, tb2], 1), tw, np.stack([bb1, bb2], 1), bw, (gl, gl, gl))
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    b = Assembler(op, X, X, GPUConfiguration(rules=rules), ctx=gpu_ctx)
    Za, Zb = a.assemble().to_host(), b.assemble().to_host()
    a.close()
    b.close()
    assert np.array_equal(Za, Zb)


def test_reference_rule_from_cell_points():
    """BEAST's per-cell form (Cartesian points, weights times 2A) maps back to
    the reference rule."""
    v = np.array([[0.3, 0.1, 0.2], [1.1, 0.4, -0.2], [0.2, 0.9, 0.5]])
    uv = np.array([[0.5, 0.5], [0.0, 0.5], [0.5, 0.0]])
    w = np.array([1.0, 2.0, 3.0]) / 6
    pts = v[2] + uv[:, :1] * (v[0] - v[2]) + uv[:, 1:] * (v[1] - v[2])
    two_a = np.linalg.norm(np.cross(v[0] - v[2], v[1] - v[2]))
    uv2, w2 = QuadratureRules.reference_rule(v, pts, w * two_a)
    assert np.allclose(uv2, uv, atol=1e-14) and np.allclose(w2, w, rtol=1e-14)


def _split_sphere(level):
    """The two hemispheres (by cell centroid height) of an icosphere as two
    open meshes: cells touch across the cut (common edges and vertices)."""
    s = meshes.perturb_radial(meshes.icosphere(level), 6, 0.02)
    up = s.vertices[s.faces].mean(axis=1)[:, 2] >= 0.0
    out = []
    for sel in (up, ~up):
        f = s.faces[sel]
        used, inv = np.unique(f.ravel(), return_inverse=True)
        out.append(meshes.Mesh(s.vertices[used], inv.reshape(-1, 3).astype(np.int32)))
    return out


@pytest.mark.parametrize("geometry", ["shells", "hemispheres"])
@pytest.mark.parametrize("kind", ["efie", "hh3d", "mfie"])
def test_rectangular_two_meshes_match_oracle(gpu_ctx, geometry, kind):
    """Test space on one mesh, trial space on another: two concentric shells
    3 % apart at different refinement (near pairs), or the two halves of a
    sphere (touching pairs across the cut).  Rectangular Z against the oracle,
    FP64 at 1e-12, the streamed host path bitwise equal.  Mixed: 1e-6 for the
    shells; the hemispheres' block has no singular part, every entry is a
    sum of far contributions of the two cells of each function that cancel
    to about h / R of their size, so the FP32 evaluation error of the far
    pairs is not masked by larger entries (measured 2.5e-6): 1e-5."""
    if geometry == "shells":
        mx, my = meshes.icosphere(2), meshes.icosphere(3, radius=1.03)
    else:
        mx, my = _split_sphere(3)
    op, X = _op_space(kind, mx)
    _, Y = _op_space(kind, my)
    assert X.nfun != Y.nfun
    print(f"case rectangular {geometry} {kind} {X.nfun} x {Y.nfun}", flush=True)
    Zref, counts = O.assemble(op, X, Y)
    if geometry == "shells":
        assert counts[4] > 0
    else:
        assert counts[1] > 0 and counts[2] > 0 and counts[3] == 0
    for precision, tol in (("fp64", FP64_TOL), ("mixed", MIXED_TOL if geometry == "shells" else 1e-5)):
        a = Assembler(op, X, Y, GPUConfiguration(precision=precision), ctx=gpu_ctx)
        Z = a.assemble().to_host()
        st = a.stats()
        Zs = a.assemble_to_host(budget=2 * 3 * 64 * Y.nfun * 16)
        a.close()
        assert Z.shape == (X.nfun, Y.nfun)
        assert np.array_equal(Z, Zs)
        err = relfro(Z, Zref)
        print(precision, "rel err", err, st["pairs"], flush=True)
        assert err <= tol
        assert [st["pairs"][c] for c in ("common_vertex", "common_edge", "common_face", "near")] == list(counts[1:])


def test_overlapping_meshes_are_refused(gpu_ctx):
    """A trial mesh that refines the test mesh (midpoint subdivision, its new
    vertices on the coarse edges) is refused with BMI_EUNSUPPORTED, either way
    round; a separate copy of the same mesh is accepted."""
    c = meshes.icosphere(1)
    v, f = c.vertices, c.faces
    verts, mid, faces = list(v), {}, []

    def m(a, b):
        key = (min(a, b), max(a, b))
        if key not in mid:
            mid[key] = len(verts)
            verts.append(0.5 * (v[a] + v[b]))
        return mid[key]

    for a, b, cc in f:
        ab, bc, ca = m(a, b), m(b, cc), m(cc, a)
        faces += [(a, ab, ca), (ab, b, bc), (ca, bc, cc), (ab, bc, ca)]
    fine = meshes.Mesh(np.array(verts), np.array(faces, dtype=np.int32))
    op = Maxwell3D.singlelayer(wavenumber=1.0)
    X, Y = spaces.raviartthomas(c), spaces.raviartthomas(fine)
    for A, B in ((X, Y), (Y, X)):
        with pytest.raises(_bmi.BMIError) as e:
            Assembler(op, A, B, GPUConfiguration(), ctx=gpu_ctx)
        assert e.value.status == _bmi.EUNSUPPORTED
    copy = meshes.Mesh(c.vertices.copy(), c.faces.copy())
    Y2 = spaces.raviartthomas(copy)
    a = Assembler(op, X, Y2, GPUConfiguration(), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zref, _ = O.assemble(op, X, X)
    assert relfro(Z, Zref) <= FP64_TOL
