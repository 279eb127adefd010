"""Request loop on the device: results alone and in a concurrent mix are the
same, resident matrices serve later applies, admission under a small memory
budget still completes every request."""
import numpy as np
import pytest

from beastgpu_amd import Helmholtz3D, Maxwell3D, meshes, spaces
from beastgpu_amd.api import Assembler, GPUConfiguration
from beastgpu_amd.serving import Server

pytestmark = pytest.mark.gpu


def _requests(n=10, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        level = int(rng.integers(1, 5))
        m = meshes.perturb_radial(meshes.icosphere(level), seed=100 + i, amplitude=float(rng.uniform(0, 0.05)))
        k = float(rng.uniform(0.5, 6.0))
        if i % 3 == 2:
            op, X = Helmholtz3D.singlelayer(wavenumber=k), spaces.lagrangec0d1(m)
        else:
            op, X = Maxwell3D.singlelayer(wavenumber=k), spaces.raviartthomas(m)
        cfg = GPUConfiguration(precision="mixed" if i % 4 == 3 else "fp64")
        x = np.sqrt(rng.uniform(size=(X.nfun, 4))) * np.exp(2j * np.pi * rng.uniform(size=(X.nfun, 4)))
        out.append((op, X, cfg, x))
    return out


def test_alone_equals_in_mix_and_matches_assembler(gpu_ctx):
    reqs = _requests()
    srv = Server(slots=4)
    alone = []
    for op, X, cfg, x in reqs:
        print("alone", X.nfun, flush=True)
        res, y = srv.wait(srv.submit_assemble_apply(op, X, x, cfg))
        assert res["entries"] == X.nfun ** 2
        alone.append(y)
    ids = [srv.submit_assemble_apply(op, X, x, cfg) for op, X, cfg, x in reqs]
    mixed = [srv.wait(i)[1] for i in ids]
    cur, peak = srv.memory()
    srv.close()
    # what stays resident is the workers' matrix buffers: at most one of the
    # largest matrix per slot
    biggest = max(X.nfun for _, X, _, _ in reqs) ** 2 * 16
    assert 0 < cur <= 4 * biggest and peak >= cur
    for ya, ym in zip(alone, mixed):
        assert np.array_equal(ya, ym)
    # the same request through the single-assembly path
    for (op, X, cfg, x), ya in zip(reqs[:4], alone[:4]):
        a = Assembler(op, X, X, cfg, ctx=gpu_ctx)
        yr = a.assemble().apply(x)
        a.close()
        assert np.array_equal(ya, yr)


def test_resident_matrix_serves_applies_and_release():
    op = Maxwell3D.singlelayer(wavenumber=2.0)
    X = spaces.raviartthomas(meshes.icosphere(3))
    rng = np.random.default_rng(3)
    xs = [rng.standard_normal((X.nfun, 2)) + 1j * rng.standard_normal((X.nfun, 2)) for _ in range(3)]
    srv = Server(slots=3)
    mid = srv.submit_assemble(op, X)
    app = [srv.submit_apply(mid, x) for x in xs]  # queued before the matrix exists
    res, _ = srv.wait(mid)
    ys = [srv.wait(i)[1] for i in app]
    assert srv.memory()[0] == X.nfun ** 2 * 16  # the kept matrix
    y_direct = srv.wait(srv.submit_assemble_apply(op, X, xs[0]))[1]
    assert np.array_equal(ys[0], y_direct)
    held = srv.memory()[0]  # kept matrix + one worker buffer
    srv.release(mid)
    assert srv.memory()[0] == held - X.nfun ** 2 * 16
    with pytest.raises(Exception):
        srv.submit_apply(mid, xs[0])
    srv.close()


def test_small_budget_serialises_but_completes():
    reqs = _requests(n=6, seed=5)
    biggest = max(X.nfun for _, X, _, _ in reqs)
    # room for about one request at a time
    srv = Server(slots=4, memory_budget=int(biggest ** 2 * 16 * 1.5 + (200 << 20)))
    ids = [srv.submit_assemble_apply(op, X, x, cfg) for op, X, cfg, x in reqs]
    out = [srv.wait(i) for i in ids]
    srv.close()
    for (op, X, cfg, x), (res, y) in zip(reqs, out):
        assert y.shape == (X.nfun, 4) and np.all(np.isfinite(y))


@pytest.mark.parametrize("which,precision", [("ellipsoid", "mixed"), ("ellipsoid", "fp64"), ("shells", "mixed"),
                                             ("icosphere4", "fp64")])
def test_admission_bound_covers_the_plan(gpu_ctx, which, precision):
    """The server admits an assembly against bmi_request_bytes: its plan
    bound is at least what the plan allocates (jittered ellipsoid with many
    FP64-kept far pairs, two concentric shells with near pairs)."""
    import ctypes

    from beastgpu_amd import _bmi
    from beastgpu_amd.api import _c_operator

    if which == "ellipsoid":
        m = meshes.ellipsoid(cells=(40, 28, 28))
    elif which == "shells":
        s = meshes.icosphere(3)
        m = meshes.Mesh(np.vstack([s.vertices, 1.03 * s.vertices]), np.vstack([s.faces, s.faces + s.nv]))
    else:
        m = meshes.icosphere(4)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=3.0)
    cfg = GPUConfiguration(precision=precision)
    print(f"case admission bound {which} {precision} nt={m.nt}", flush=True)
    a = Assembler(op, X, X, cfg, ctx=gpu_ctx)
    st = a.stats()
    a.close()
    keep = [np.ascontiguousarray(m.vertices), np.ascontiguousarray(m.faces, dtype=np.int32),
            np.ascontiguousarray(X.idx), np.ascontiguousarray(X.coeff)]
    r = _bmi.Request()
    r.kind = _bmi.REQ_ASSEMBLE_APPLY
    r.xyz, r.nv, r.tri, r.nt = keep[0].ctypes.data, m.nv, keep[1].ctypes.data, m.nt
    r.space_kind, r.gidx, r.coeff, r.nfun = _bmi.SPACE_RWG, keep[2].ctypes.data, keep[3].ctypes.data, X.nfun
    r.op, r.opts, r.dtype = _c_operator(op), cfg.options(), cfg.dtype()
    mb, pb = ctypes.c_int64(), ctypes.c_int64()
    _bmi.check(_bmi.lib().bmi_request_bytes(ctypes.byref(r), ctypes.byref(mb), ctypes.byref(pb)))
    print("plan bytes", st["plan_bytes"], "bound", pb.value, st["pairs"], flush=True)
    assert 0 < st["plan_bytes"] <= pb.value
