"""Far-pair tile shapes other than the default (bmi_space_create_shaped):
panels of fewer rows and patches of fewer columns assemble the oracle's
matrix, alone and in a batch, and shapes outside the kernels' tile are
refused; the tile autotuner (tile="auto") picks the fastest of its
candidate shapes."""
import numpy as np
import pytest

from beastgpu_amd import Helmholtz3D, Maxwell3D, _bmi, api, meshes, spaces
from beastgpu_amd.api import Assembler, Batch, DeviceMesh, DeviceSpace, GPUConfiguration
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-6
CLASSES = ("far", "common_vertex", "common_edge", "common_face", "near")


def relfro(A, B):
    return np.linalg.norm(A - B) / np.linalg.norm(B)


def _case(kind, seed=5):
    m = meshes.perturb_radial(meshes.icosphere(2), seed, 0.03)
    if kind == "efie":
        return Maxwell3D.singlelayer(wavenumber=2.0), spaces.raviartthomas(m)
    if kind == "mfie":
        return Maxwell3D.doublelayer(wavenumber=2.0), spaces.raviartthomas(m)
    return Helmholtz3D.singlelayer(wavenumber=2.0), spaces.lagrangec0d1(m)


@pytest.mark.parametrize("kind,tile", [("efie", (48, 28)), ("efie", (32, 24)), ("efie", (17, 9)),
                                       ("mfie", (40, 20)), ("hh3d", (96, 12)), ("hh3d", (64, 20)),
                                       ("hh3d", (33, 7))])
def test_tile_shapes_match_oracle(gpu_ctx, kind, tile):
    op, X = _case(kind)
    print(f"case tile {kind} {tile} nfun={X.nfun}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(tile=tile), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    st = a.stats()
    a.close()
    Zref, counts = O.assemble(op, X, X)
    err = relfro(Z, Zref)
    print("rel err", err, flush=True)
    assert [st["pairs"][n] for n in CLASSES] == list(counts)
    assert err <= FP64_TOL


@pytest.mark.parametrize("kind,tile", [("efie", (32, 24)), ("hh3d", (64, 12))])
def test_tile_shapes_mixed_within_tolerance(gpu_ctx, kind, tile):
    op, X = _case(kind, seed=6)
    print(f"case tile mixed {kind} {tile}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(tile=tile, precision="mixed"), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    a.close()
    Zref, _ = O.assemble(op, X, X)
    err = relfro(Z, Zref)
    print("mixed rel err", err, flush=True)
    assert err <= MIXED_TOL


@pytest.mark.parametrize("kind,tile", [("efie", (65, 0)), ("efie", (0, 41)), ("efie", (0, 2)), ("efie", (-1, 0)),
                                       ("hh3d", (129, 0)), ("hh3d", (0, 21))])
def test_tile_shape_out_of_range_refused(gpu_ctx, kind, tile):
    _, X = _case(kind)
    dm = DeviceMesh(gpu_ctx, X.mesh)
    with pytest.raises(_bmi.BMIError) as e:
        DeviceSpace(dm, X, tile)
    assert e.value.status == _bmi.EINVAL
    dm.close()


def test_batch_of_shaped_spaces_bitwise(gpu_ctx):
    """A batch whose problems have different tile shapes gives bitwise each
    problem's own assembly under every schedule."""
    probs = [("efie", (0, 0), "fp64"), ("efie", (48, 28), "fp64"), ("efie", (32, 40), "mixed"),
             ("mfie", (40, 20), "fp64"), ("hh3d", (64, 12), "fp64"), ("hh3d", (0, 0), "fp64")]
    asms, others = [], []
    for i, (kind, tile, prec) in enumerate(probs):
        op, X = _case(kind, seed=20 + i)
        cfg = GPUConfiguration(tile=tile, precision=prec)
        asms.append(Assembler(op, X, X, cfg, ctx=gpu_ctx))
        # another wavenumber on the same space: overwrites the matrix between replays
        op2 = Helmholtz3D.singlelayer(wavenumber=2.7) if kind == "hh3d" else Maxwell3D.singlelayer(wavenumber=2.7)
        others.append(Assembler(op2, X, X, cfg, ctx=gpu_ctx))
    ref = [a.assemble().to_host() for a in asms]
    print("case batch of shaped spaces", flush=True)
    b = Batch(asms)
    st = b.stats()
    print("batch", st, flush=True)
    schedules = [(None, 0), ("persistent", 1), ("per-problem", 0), ("per-colour", 0)]
    for mode, w in schedules:
        if mode is not None:
            for bk in range(st["buckets"]):
                b.set_schedule(bk, mode, w)
        for a, o in zip(asms, others):
            _bmi.check(_bmi.lib().bmi_assemble(o.plan.h, a.matrix.h))
        b.assemble()
        for i, (a, r) in enumerate(zip(asms, ref)):
            assert np.array_equal(a.matrix.to_host(), r), (mode, i)
    b.close()
    for a in asms + others:
        a.close()


@pytest.mark.parametrize("kind", ["efie", "hh3d"])
def test_auto_tile_chooses_fastest_candidate_and_matches_oracle(gpu_ctx, kind):
    """tile="auto" times every candidate shape on the problem, keeps the
    fastest (the default unless beaten by more than 3 %), assembles the
    oracle's matrix with it, and a second problem of the same signature
    reuses the choice without timing again."""
    api._TILE_CACHE.clear()
    op, X = _case(kind, seed=7)
    print(f"case auto tile {kind} nfun={X.nfun}", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(tile="auto"), ctx=gpu_ctx)
    print("tile", a.tile, "times", a.tile_times, flush=True)
    cands = api.TILE_CANDIDATES[X.kind]
    assert a.tile in cands and set(a.tile_times) == set(cands)
    # the fastest candidate, or the default when none beats it by 3 %
    best = min(a.tile_times, key=a.tile_times.get)
    assert a.tile == (best if a.tile_times[best] < 0.97 * a.tile_times[(0, 0)] else (0, 0))
    assert all(0 < t < 1e4 for t in a.tile_times.values())
    Z = a.assemble().to_host()
    st = a.stats()
    a.close()
    Zref, counts = O.assemble(op, X, X)
    assert [st["pairs"][n] for n in CLASSES] == list(counts)
    assert relfro(Z, Zref) <= FP64_TOL
    op2, X2 = _case(kind, seed=8)
    b = Assembler(op2, X2, X2, GPUConfiguration(tile="auto"), ctx=gpu_ctx)
    assert b.tile == a.tile and b.tile_times is a.tile_times and len(api._TILE_CACHE) == 1
    Z2 = b.assemble().to_host()
    b.close()
    Zref2, _ = O.assemble(op2, X2, X2)
    assert relfro(Z2, Zref2) <= FP64_TOL


def test_auto_tile_without_resident_matrix_takes_default(gpu_ctx):
    op, X = _case("efie", seed=9)
    a = Assembler(op, X, X, GPUConfiguration(tile="auto"), ctx=gpu_ctx, resident=False)
    assert a.tile == (0, 0) and a.tile_times is None and a.matrix is None
    a.close()
