"""Batched assembly (bmi_batch_*): one replayed graph over many problems gives
bitwise the matrices of the single assemblies, on every replay, with other
work on the context between replays."""
import numpy as np
import pytest

from beastgpu_amd import Helmholtz3D, Maxwell3D, _bmi, meshes, spaces
from beastgpu_amd.api import Assembler, Batch, GPUConfiguration

pytestmark = pytest.mark.gpu


def _problems():
    rng = np.random.default_rng(4)
    out = []
    for i in range(12):
        level = 1 + i % 3
        m = meshes.perturb_radial(meshes.icosphere(level), seed=40 + i, amplitude=float(rng.uniform(0, 0.05)))
        k = float(rng.uniform(0.5, 4.0))
        kind = i % 4
        if kind == 0:
            op, X = Maxwell3D.singlelayer(wavenumber=k), spaces.raviartthomas(m)
        elif kind == 1:
            op, X = Maxwell3D.doublelayer(wavenumber=k), spaces.raviartthomas(m)
        elif kind == 2:
            op, X = Helmholtz3D.singlelayer(wavenumber=k), spaces.lagrangec0d1(m)
        else:
            op, X = Maxwell3D.singlelayer(gamma=complex(0.2, k)), spaces.raviartthomas(m)  # run-time rule kernels
        cfg = GPUConfiguration(precision="mixed" if i % 3 == 1 else "fp64", output="c32" if i % 5 == 4 else "c64")
        out.append((op, X, cfg))
    return out


def test_batch_bitwise_equals_single_assemblies_on_every_replay(gpu_ctx):
    probs = _problems()
    asms = [Assembler(op, X, X, cfg, ctx=gpu_ctx) for op, X, cfg in probs]
    ref = []
    for a in asms:
        ref.append(a.assemble().to_host())
    # a second plan per problem (other wavenumber) to overwrite the matrices between replays
    others = []
    for (op, X, cfg), a in zip(probs, asms):
        if op.kind == 1:
            op2 = Helmholtz3D.singlelayer(wavenumber=op.wavenumber + 0.7)
        elif op.kind == 3:
            op2 = Maxwell3D.doublelayer(wavenumber=op.wavenumber + 0.7)
        else:
            op2 = Maxwell3D.singlelayer(wavenumber=op.wavenumber + 0.7)
        others.append(Assembler(op2, X, X, cfg, ctx=gpu_ctx))
    print("case batch of", len(asms), flush=True)
    b = Batch(asms)
    st = b.stats()
    print("batch", st, flush=True)
    assert st["problems"] == len(asms) and st["buckets"] >= 2
    for rep in range(3):
        for a, o in zip(asms, others):
            _bmi.check(_bmi.lib().bmi_assemble(o.plan.h, a.matrix.h))
        changed = sum(not np.array_equal(a.matrix.to_host(), r) for a, r in zip(asms, ref))
        assert changed == len(asms)
        b.assemble()
        assert b.stats()["last_ms"] > 0
        for i, (a, r) in enumerate(zip(asms, ref)):
            assert np.array_equal(a.matrix.to_host(), r), (rep, i)
    # every schedule of every bucket: persistent with 1 and with the most
    # work-groups per CU, per-problem launches, then per-colour
    for bk in range(st["buckets"]):
        for mode, w in (("persistent", 1), ("persistent", 3), ("per-problem", 0), ("per-colour", 0)):
            try:
                b.set_schedule(bk, mode, w)
            except _bmi.BMIError as e:  # 3 work-groups per CU beyond the FP64 double layer's occupancy
                assert e.status == _bmi.EINVAL and w == 3
                b.set_schedule(bk, mode, 2)
            for a, o in zip(asms, others):
                _bmi.check(_bmi.lib().bmi_assemble(o.plan.h, a.matrix.h))
            b.assemble()
            s2 = b.stats()
            assert s2["bucket_mode"][bk] == mode
            for i, (a, r) in enumerate(zip(asms, ref)):
                assert np.array_equal(a.matrix.to_host(), r), (bk, mode, w, i)
    with pytest.raises(_bmi.BMIError):
        b.set_schedule(st["buckets"], "per-colour")
    with pytest.raises(_bmi.BMIError):
        b.set_schedule(0, "persistent", 0)
    b.close()
    for a in asms + others:
        a.close()
