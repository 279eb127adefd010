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


@pytest.mark.parametrize("output", ["c32", "c64"])
def test_mixed_icosphere_l3_k4(gpu_ctx, output):
    """The unperturbed level-3 icosphere (1,920 RWG) at k = 4 on the mixed
    path: relative Frobenius error and relative Z x error (x uniform in the
    complex unit disk, seed 33) against the FP64 oracle."""
    m = meshes.icosphere(3)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=4.0)
    print(f"case mixed L3 k=4 {output}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(precision="mixed", output=output), ctx=gpu_ctx)
    M = a.assemble()
    Z = M.to_host().astype(np.complex128)
    rng = np.random.default_rng(33)
    x = np.sqrt(rng.uniform(size=X.nfun)) * np.exp(2j * np.pi * rng.uniform(size=X.nfun))
    y = M.apply(x)
    a.close()
    Zref, _ = O.assemble(op, X, X)
    yref = Zref @ x
    err = relfro(Z, Zref)
    zx = np.linalg.norm(y - yref) / np.linalg.norm(yref)
    print("mixed L3 k=4 rel err", err, "zx", zx)
    assert err <= MIXED_TOL and zx <= MIXED_TOL


@pytest.mark.parametrize("kind", ["efie", "mfie", "hh3d"])
def test_level4_full_matrix_matches_oracle(gpu_ctx, kind):
    """A perturbed level-4 icosphere (5,120 triangles; 7,680 RWG or 2,562 P1
    functions, several panels and trial colours per launch): the whole FP64
    matrix against the oracle at 1e-12 and the pair counts, then the mixed
    path within 1e-6."""
    m = meshes.perturb_radial(meshes.icosphere(4), 13, 0.04)
    if kind == "hh3d":
        op, X = Helmholtz3D.singlelayer(wavenumber=2.5), spaces.lagrangec0d1(m)
    elif kind == "mfie":
        op, X = Maxwell3D.doublelayer(wavenumber=2.5), spaces.raviartthomas(m)
    else:
        op, X = Maxwell3D.singlelayer(wavenumber=2.5), spaces.raviartthomas(m)
    print(f"case L4 full {kind} nfun={X.nfun}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    st = a.stats()
    a.close()
    Zref, counts = O.assemble(op, X, X)
    err = relfro(Z, Zref)
    print("L4 fp64 rel err", err, flush=True)
    assert [st["pairs"][n] for n in ("far", "common_vertex", "common_edge", "common_face", "near")] == list(counts)
    assert err <= FP64_TOL
    del Z
    b = Assembler(op, X, X, GPUConfiguration(precision="mixed"), ctx=gpu_ctx)
    Zm = b.assemble().to_host()
    b.close()
    errm = relfro(Zm, Zref)
    print("L4 mixed rel err", errm, flush=True)
    assert errm <= MIXED_TOL


@pytest.mark.parametrize("kind", ["efie", "mfie", "hh3d"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_complexf32_output_matches_oracle(gpu_ctx, kind, precision):
    """ComplexF32 matrices (the FP64 path rounded once at the write-back, and
    the mixed path) of every operator on a perturbed level-3 mesh: relative
    Frobenius and Z x errors against the FP64 oracle within 1e-6."""
    m = meshes.perturb_radial(meshes.icosphere(3), 11, 0.04)
    if kind == "hh3d":
        op, X = Helmholtz3D.singlelayer(wavenumber=3.0), spaces.lagrangec0d1(m)
    elif kind == "mfie":
        op, X = Maxwell3D.doublelayer(wavenumber=3.0), spaces.raviartthomas(m)
    else:
        op, X = Maxwell3D.singlelayer(wavenumber=3.0), spaces.raviartthomas(m)
    print(f"case c32 {kind} {precision}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(precision=precision, output="c32"), ctx=gpu_ctx)
    M = a.assemble()
    Z = M.to_host()
    assert Z.dtype == np.complex64
    rng = np.random.default_rng(33)
    x = np.sqrt(rng.uniform(size=X.nfun)) * np.exp(2j * np.pi * rng.uniform(size=X.nfun))
    y = M.apply(x)
    a.close()
    Zref, _ = O.assemble(op, X, X)
    yref = Zref @ x
    err = relfro(Z.astype(np.complex128), Zref)
    zx = np.linalg.norm(y - yref) / np.linalg.norm(yref)
    print("c32 rel err", err, "zx", zx, flush=True)
    assert err <= MIXED_TOL and zx <= MIXED_TOL
    if precision == "fp64":
        # one or two roundings to ComplexF32 per entry (read-modify-written
        # columns are rounded again after their second part is added)
        assert err <= 2e-7


@pytest.mark.parametrize("d,precision", [(4, "fp64"), (6, "fp64"), (6, "mixed")])
def test_meshcuboid_hh3d_multi_panel(gpu_ctx, d, precision):
    """The reference's own workload shape, meshcuboid(1,1,1,0.5/d) with the
    Helmholtz single layer on P1, at sizes with several 128-row panels (d=4:
    386 functions, d=6: 866): class counts equal to the oracle's, whole matrix
    within 1e-12 (FP64) or 1e-6 (mixed)."""
    m = meshes.meshcuboid(1.0, 1.0, 1.0, 0.5 / d)
    X = spaces.lagrangec0d1(m)
    assert X.nfun > 2 * 128
    op = Helmholtz3D.singlelayer(wavenumber=1.0)
    print(f"case meshcuboid d={d} {precision} nfun={X.nfun}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(precision=precision), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    st = a.stats()
    a.close()
    Zref, counts = O.assemble(op, X, X)
    err = relfro(Z, Zref)
    print("meshcuboid rel err", err, st)
    assert [st["pairs"][n] for n in ("far", "common_vertex", "common_edge", "common_face", "near")] == \
        [int(counts[0]) - (st["pairs"]["far_fp64"] if precision == "mixed" else 0)] + list(counts[1:])
    assert err <= (MIXED_TOL if precision == "mixed" else FP64_TOL)


def test_apply_and_compare_match_host(gpu_ctx):
    op, X = _case("efie", 3, 1.5, 0.02, seed=3)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    M = a.assemble()
    Z = M.to_host()
    rng = np.random.default_rng(1)
    x = rng.standard_normal((X.nfun, 4)) + 1j * rng.standard_normal((X.nfun, 4))
    y = M.apply(x)
    assert np.abs(y - Z @ x).max() <= 1e-12 * np.abs(Z @ x).max()
    b = Assembler(op, X, X, GPUConfiguration(precision="mixed", output="c32"), ctx=gpu_ctx)
    Mb = b.assemble()
    cmp = Mb.compare(M)
    Zb = Mb.to_host().astype(np.complex128)
    assert abs(cmp["rel_fro"] - relfro(Zb, Z)) <= 1e-6 * relfro(Zb, Z)
    assert abs(cmp["max_abs"] - np.abs(Zb - Z).max()) <= 1e-9 * np.abs(Z).max()
    a.close()
    b.close()


def test_repeat_assembly_bitwise_deterministic(gpu_ctx):
    op, X = _case("efie", 3, 2.0, 0.03, seed=4)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    Z1 = a.assemble().to_host()
    Z2 = a.assemble().to_host()
    a.close()
    assert np.array_equal(Z1, Z2)


def test_icosphere_l5_sampled_rows_and_mixed(gpu_ctx):
    """The full-size perturbed level-5 icosphere (20,480 triangles, 30,720
    RWG): 48 rows against the oracle at 1e-12, the mixed path against FP64 on
    the whole matrix, and the class counts against the Euler closed forms."""
    m = meshes.perturb_radial(meshes.icosphere(5), seed=2, amplitude=0.03)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=2.0)
    print("case icosphere L5 EFIE", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    M = a.assemble()
    st = a.stats()
    deg = m.vertex_degrees().astype(np.int64)
    assert st["pairs"]["common_face"] == m.nt and st["pairs"]["common_edge"] == 3 * m.nt
    assert st["pairs"]["common_vertex"] == int((deg ** 2).sum() - 9 * m.nt)
    rows = np.concatenate([np.arange(0, 16), np.arange(15000, 16000, 64)[:16], np.arange(X.nfun - 16, X.nfun)])
    Zr, _ = O.assemble(op, X, X, rows=rows)
    rng = np.random.default_rng(2)
    # rows of the device matrix through apply with unit vectors would be slow; copy the matrix
    Z = M.to_host()
    err = relfro(Z[rows], Zr)
    print("icosphere L5 sampled-row err", err, flush=True)
    assert err <= FP64_TOL
    del Z
    b = Assembler(op, X, X, GPUConfiguration(precision="mixed"), ctx=gpu_ctx)
    cmp = b.assemble().compare(M)
    print("icosphere L5 mixed", cmp, b.stats(), flush=True)
    assert cmp["rel_fro"] <= MIXED_TOL
    x = np.sqrt(rng.uniform(size=X.nfun)) * np.exp(2j * np.pi * rng.uniform(size=X.nfun))
    y64, ym = M.apply(x), b.matrix.apply(x)
    assert np.linalg.norm(ym - y64) <= MIXED_TOL * np.linalg.norm(y64)
    a.close()
    b.close()


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
@pytest.mark.parametrize("far_test,far_trial", [(3, 2), (1, 3), (2, 1)])
def test_nondefault_far_rules_match_oracle(gpu_ctx, kind, far_test, far_trial):
    """Far-rule orders other than the default 3 x 4 points run the kernels
    that read the rule at run time."""
    op, X = _case(kind, 2, 1.5, 0.02, seed=5)
    print(f"case rules {kind} {far_test}x{far_trial}", flush=True)
    cfg = GPUConfiguration(far_test_order=far_test, far_trial_order=far_trial)
    a = Assembler(op, X, X, cfg, ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zref, _ = O.assemble(op, X, X, far_test=far_test, far_trial=far_trial)
    err = relfro(Z, Zref)
    print("rel err", err)
    assert err <= FP64_TOL


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("mixed", MIXED_TOL)])
def test_lossy_operator_matches_oracle(gpu_ctx, kind, precision, tol):
    """Complex gamma (attenuating medium): the exp(-Re(gamma) R) factor."""
    m = meshes.perturb_radial(meshes.icosphere(2), 6, 0.02)
    g = complex(0.4, 1.7)
    if kind == "efie":
        op, X = Maxwell3D.singlelayer(gamma=g), spaces.raviartthomas(m)
    else:
        op, X = Helmholtz3D.singlelayer(gamma=g), spaces.lagrangec0d1(m)
    print(f"case lossy {kind} {precision}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(precision=precision), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zref, _ = O.assemble(op, X, X)
    err = relfro(Z, Zref)
    print("rel err", err)
    assert err <= tol


def test_ellipsoid_60k_full_size(gpu_ctx):
    """The 60,000-triangle ellipsoid (90,000 RWG, k = 4): FP64 matrix (129.6 GB)
    and mixed ComplexF32 matrix (64.8 GB) resident together.  Checks: class
    counts against the Euler closed forms; (Z x) on 24 sampled rows against
    the oracle's rows times x at 1e-12; the mixed matrix against the FP64 one
    (relative Frobenius and Z x) at 1e-6."""
    import ctypes
    from beastgpu_amd import _bmi

    free, total = ctypes.c_int64(), ctypes.c_int64()
    _bmi.check(_bmi.lib().bmi_device_memory(gpu_ctx.h, ctypes.byref(free), ctypes.byref(total)))
    if free.value < 210e9:
        pytest.skip("needs about 200 GB of free device memory")
    m = meshes.ellipsoid()
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=4.0)
    print("case ellipsoid-60k", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    M = a.assemble()
    st = a.stats()
    deg = m.vertex_degrees().astype(np.int64)
    assert st["pairs"]["common_face"] == m.nt and st["pairs"]["common_edge"] == 3 * m.nt
    assert st["pairs"]["common_vertex"] == int((deg ** 2).sum() - 9 * m.nt)
    rng = np.random.default_rng(33)
    x = np.sqrt(rng.uniform(size=X.nfun)) * np.exp(2j * np.pi * rng.uniform(size=X.nfun))
    y64 = M.apply(x)
    rows = np.sort(rng.choice(X.nfun, size=24, replace=False))
    Zr, _ = O.assemble(op, X, X, rows=rows)
    yr = Zr @ x
    err_rows = np.linalg.norm(y64[rows] - yr) / np.linalg.norm(yr)
    print("ellipsoid sampled-row Zx err", err_rows, st, flush=True)
    assert err_rows <= FP64_TOL
    b = Assembler(op, X, X, GPUConfiguration(precision="mixed", output="c32"), ctx=gpu_ctx)
    cmp = b.assemble().compare(M)
    ym = b.matrix.apply(x)
    zx = np.linalg.norm(ym - y64) / np.linalg.norm(y64)
    print("ellipsoid mixed c32", cmp, "zx", zx, b.stats(), flush=True)
    b.close()
    a.close()
    assert cmp["rel_fro"] <= MIXED_TOL
    assert zx <= MIXED_TOL


@pytest.mark.parametrize("level,k,amp", [(1, 1.0, 0.0), (2, 2.0, 0.03), (3, 1.5, 0.02)])
def test_double_layer_fp64_matches_oracle(gpu_ctx, level, k, amp):
    """MFIE block n x K on RWG, FP64 path."""
    m = meshes.icosphere(level)
    if amp:
        m = meshes.perturb_radial(m, 3, amp)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.doublelayer(wavenumber=k)
    print(f"case dl L{level} k={k}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    st = a.stats()
    a.close()
    Zref, counts = O.assemble(op, X, X)
    err = relfro(Z, Zref)
    print("dl rel err", err)
    assert [st["pairs"][n] for n in ("far", "common_vertex", "common_edge", "common_face", "near")] == list(counts)
    assert err <= FP64_TOL


@pytest.mark.parametrize("variant", ["mixed", "lossy", "rules"])
def test_double_layer_variants(gpu_ctx, variant):
    m = meshes.perturb_radial(meshes.icosphere(2), 8, 0.03)
    X = spaces.raviartthomas(m)
    if variant == "lossy":
        op = Maxwell3D.doublelayer(gamma=complex(0.5, 2.0), alpha=0.7 - 0.4j)
    else:
        op = Maxwell3D.doublelayer(wavenumber=2.5)
    cfg = GPUConfiguration(precision="mixed" if variant == "mixed" else "fp64",
                           far_test_order=3 if variant == "rules" else 2,
                           far_trial_order=2 if variant == "rules" else 3)
    print(f"case dl {variant}", flush=True)
    a = Assembler(op, X, X, cfg, ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zref, _ = O.assemble(op, X, X, far_test=cfg.far_test_order, far_trial=cfg.far_trial_order)
    err = relfro(Z, Zref)
    print("dl rel err", err)
    assert err <= (MIXED_TOL if variant == "mixed" else FP64_TOL)


@pytest.mark.parametrize("kind", ["efie", "hh3d", "mfie"])
def test_nondefault_ss_orders_match_oracle(gpu_ctx, kind):
    """Sauter-Schwab orders other than 5 run the kernels that read the order
    at run time."""
    m = meshes.perturb_radial(meshes.icosphere(2), 12, 0.03)
    if kind == "efie":
        op, X = Maxwell3D.singlelayer(wavenumber=1.7), spaces.raviartthomas(m)
    elif kind == "mfie":
        op, X = Maxwell3D.doublelayer(wavenumber=1.7), spaces.raviartthomas(m)
    else:
        op, X = Helmholtz3D.singlelayer(wavenumber=1.7), spaces.lagrangec0d1(m)
    orders = (4, 6, 3)
    print(f"case ss orders {kind}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(ss_order=orders), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zref, _ = O.assemble(op, X, X, ss_order=orders)
    err = relfro(Z, Zref)
    print("rel err", err)
    assert err <= FP64_TOL


def _near_mesh(which):
    """Meshes with class-4 (near) pairs: two concentric spheres 3% apart (one
    mesh, two closed components; every cell is near its copy), or the
    jittered ellipsoid, whose thin cells come near cells they do not touch."""
    if which == "shells":
        a = meshes.icosphere(2)
        nv = a.vertices.shape[0]
        return meshes.Mesh(np.vstack([a.vertices, 1.03 * a.vertices]), np.vstack([a.faces, a.faces + nv]))
    return meshes.ellipsoid(cells=(16, 10, 10))


@pytest.mark.parametrize("which", ["shells", "ellipsoid"])
@pytest.mark.parametrize("kind,precision,near_order", [
    ("efie", "fp64", 4), ("hh3d", "fp64", 4), ("mfie", "fp64", 4), ("efie-lossy", "fp64", 4),
    ("efie", "fp64", 7), ("hh3d", "mixed", 4)])
def test_near_pairs_match_oracle(gpu_ctx, which, kind, precision, near_order):
    """Near pairs: closed-form singular terms plus the product rule on the
    remainder for the single layers, the plain product rule for the double
    layer, against the oracle."""
    m = _near_mesh(which)
    if kind == "efie":
        op, X = Maxwell3D.singlelayer(wavenumber=3.0), spaces.raviartthomas(m)
    elif kind == "efie-lossy":
        op, X = Maxwell3D.singlelayer(gamma=complex(0.8, 3.0)), spaces.raviartthomas(m)
    elif kind == "mfie":
        op, X = Maxwell3D.doublelayer(wavenumber=3.0), spaces.raviartthomas(m)
    else:
        op, X = Helmholtz3D.singlelayer(wavenumber=3.0, alpha=0.7 - 0.2j), spaces.lagrangec0d1(m)
    print(f"case near {which} {kind} {precision} order {near_order} nfun={X.nfun}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(precision=precision, near_order=near_order), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    st = a.stats()
    a.close()
    Zref, counts = O.assemble(op, X, X, near_order=near_order)
    err = relfro(Z, Zref)
    print("near rel err", err, st)
    assert st["pairs"]["near"] == counts[4] > 0
    assert err <= (MIXED_TOL if precision == "mixed" else FP64_TOL)
