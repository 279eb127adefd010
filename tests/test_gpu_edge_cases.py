"""Small and ragged inputs through the device path, against the oracle, and
the errors the boundary reports."""
import numpy as np
import pytest

from beastgpu_amd import Helmholtz3D, Maxwell3D, _bmi, meshes, spaces
from beastgpu_amd.api import Assembler, DeviceMatrix, GPUConfiguration
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _square():
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.2]], dtype=float)
    F = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    return meshes.Mesh(V, F)


def _plate(n=5):
    xs = np.linspace(0.0, 1.0, n + 1)
    V = np.array([(x, y, 0.05 * np.sin(3 * x + 2 * y)) for y in xs for x in xs])
    F = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = j * (n + 1) + i, j * (n + 1) + i + 1, (j + 1) * (n + 1) + i, (j + 1) * (n + 1) + i + 1
            F += [(a, b, d), (a, d, c)]
    return meshes.Mesh(V, np.array(F, dtype=np.int32))


def _check(gpu_ctx, op, X, cfg=None, tol=1e-12):
    a = Assembler(op, X, X, cfg or GPUConfiguration(), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zr, _ = O.assemble(op, X, X)
    scale = max(np.linalg.norm(Zr), 1e-300)
    assert np.linalg.norm(Z - Zr) <= tol * scale
    return Z


@pytest.mark.parametrize("kind", ["p1", "rwg-efie", "rwg-mfie"])
def test_two_triangles(gpu_ctx, kind):
    m = _square()
    print("case two triangles", kind, flush=True)
    if kind == "p1":
        _check(gpu_ctx, Helmholtz3D.singlelayer(wavenumber=1.0), spaces.lagrangec0d1(m))
    else:
        X = spaces.raviartthomas(m)
        assert X.nfun == 1
        op = Maxwell3D.singlelayer(wavenumber=1.0) if kind == "rwg-efie" else Maxwell3D.doublelayer(wavenumber=1.0)
        _check(gpu_ctx, op, X)


def test_single_triangle_p1(gpu_ctx):
    m = meshes.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=float), np.array([[0, 1, 2]], dtype=np.int32))
    print("case one triangle", flush=True)
    _check(gpu_ctx, Helmholtz3D.singlelayer(wavenumber=2.0), spaces.lagrangec0d1(m))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_open_plate(gpu_ctx, precision):
    """Open surface: boundary edges carry no RWG function."""
    m = _plate()
    X = spaces.raviartthomas(m)
    print("case plate", precision, X.nfun, flush=True)
    _check(gpu_ctx, Maxwell3D.singlelayer(wavenumber=3.0), X, GPUConfiguration(precision=precision),
           tol=1e-12 if precision == "fp64" else 1e-6)


def test_uncovered_functions_give_zero_rows(gpu_ctx):
    """A space whose last functions have no local shape (an isolated vertex)."""
    m = meshes.icosphere(1)
    V = np.vstack([m.vertices, [[3.0, 0.0, 0.0]]])
    m2 = meshes.Mesh(V, m.faces.copy())
    X = spaces.lagrangec0d1(m2)
    assert X.nfun == m2.nv
    print("case uncovered", flush=True)
    Z = _check(gpu_ctx, Helmholtz3D.singlelayer(wavenumber=1.0), X)
    assert np.all(Z[-1, :] == 0) and np.all(Z[:, -1] == 0)


def test_boundary_errors(gpu_ctx):
    m = meshes.icosphere(1)
    P1, RWG = spaces.lagrangec0d1(m), spaces.raviartthomas(m)
    with pytest.raises(_bmi.BMIError) as e:
        Assembler(Maxwell3D.singlelayer(wavenumber=1.0), P1, P1, ctx=gpu_ctx)
    assert e.value.status == _bmi.EUNSUPPORTED
    with pytest.raises(_bmi.BMIError) as e:
        Assembler(Helmholtz3D.singlelayer(wavenumber=1.0), RWG, RWG, ctx=gpu_ctx)
    assert e.value.status == _bmi.EUNSUPPORTED
    a = Assembler(Maxwell3D.singlelayer(wavenumber=1.0), RWG, RWG, ctx=gpu_ctx)
    wrong = DeviceMatrix(gpu_ctx, RWG.nfun + 1, RWG.nfun)
    with pytest.raises(_bmi.BMIError) as e:
        _bmi.check(_bmi.lib().bmi_assemble(a.plan.h, wrong.h))
    assert e.value.status == _bmi.EINVAL
    with pytest.raises(_bmi.BMIError):
        Assembler(Maxwell3D.singlelayer(wavenumber=1.0), RWG, RWG, GPUConfiguration(far_test_order=5), ctx=gpu_ctx)
    with pytest.raises(_bmi.BMIError):
        a.matrix.apply(np.zeros((RWG.nfun, 9)))
    wrong.close()
    a.close()


def test_stats_before_and_after_the_first_assembly(gpu_ctx):
    """Pair counts asked for before any assembly (the plan then classifies
    once) equal the oracle's and the counts after assembling."""
    m = meshes.ellipsoid(cells=(16, 10, 10))
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=3.0)
    names = ("far", "common_vertex", "common_edge", "common_face", "near")
    print("case stats before assembly", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    before = [a.stats()["pairs"][n] for n in names]
    Z = a.assemble().to_host()
    after = [a.stats()["pairs"][n] for n in names]
    a.close()
    Zr, counts = O.assemble(op, X, X)
    assert before == after == list(counts)
    assert np.linalg.norm(Z - Zr) <= 1e-12 * np.linalg.norm(Zr)
