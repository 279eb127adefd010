"""GPU parity on the full batched and serving workloads: the 64-problem
EFIE/MFIE batch (icosphere L2-L4, mixed precision) and a time-compressed
Poisson stream of assemble-and-apply requests with level-5 meshes.  Each
result is checked bitwise against the single-assembly path and against the
CPU oracle."""
import threading
import time

import numpy as np
import pytest

from beastgpu_amd import _bmi, workloads
from beastgpu_amd.api import Assembler, Batch, DeviceMatrix, GPUConfiguration
from beastgpu_amd.serving import Server
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12   # relative, FP64 path against the FP64 oracle
MIXED_TOL = 1e-6   # relative, mixed path against the FP64 oracle / FP64 path


def relfro(A, B):
    return np.linalg.norm(A - B) / np.linalg.norm(B)


def _sample_rows(n, count, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, size=min(count, n), replace=False))


def test_batched_64_problems_bitwise_oracle_and_mixed_error(gpu_ctx):
    """Every matrix of the replayed batch equals its own bmi_assemble bitwise;
    against the oracle: whole matrices of the L2/L3 problems, 48 sampled rows
    of the L4 ones, at 1e-6; the mixed batch against FP64 single assemblies
    (relative Frobenius and Z x) at 1e-6."""
    probs = workloads.batched_efie_mfie()
    cfg = GPUConfiguration(precision="mixed")
    asms = [Assembler(p.op, p.X, p.X, cfg, ctx=gpu_ctx) for p in probs]
    singles = []
    for a in asms:
        s = DeviceMatrix(gpu_ctx, a.matrix.rows, a.matrix.cols)
        _bmi.check(_bmi.lib().bmi_assemble(a.plan.h, s.h))
        singles.append(s)
    print("case batched 64 problems", flush=True)
    b = Batch(asms)
    b.assemble()
    b.assemble()  # a second replay over the first one's results
    st = b.stats()
    print("batch", st, flush=True)
    assert st["problems"] == len(probs)
    for i, (a, s) in enumerate(zip(asms, singles)):
        cmp = a.matrix.compare(s)
        assert cmp["diff_fro"] == 0.0 and cmp["max_abs"] == 0.0, (i, cmp)
        s.close()
    worst = {"oracle": 0.0, "fro": 0.0, "zx": 0.0}
    for i, (p, a) in enumerate(zip(probs, asms)):
        Z = a.matrix.to_host()
        if p.level <= 3:
            Zr, _ = O.assemble(p.op, p.X, p.X)
            err = relfro(Z, Zr)
        else:
            rows = _sample_rows(p.X.nfun, 48, 100 + i)
            Zr, _ = O.assemble(p.op, p.X, p.X, rows=rows)
            err = relfro(Z[rows], Zr)
        del Z
        f = Assembler(p.op, p.X, p.X, GPUConfiguration(), ctx=gpu_ctx)
        M64 = f.assemble()
        cmp = a.matrix.compare(M64)
        x = workloads.unit_disk_vector(p.X.nfun, 33)
        y64 = M64.apply(x)
        zx = float(np.linalg.norm(a.matrix.apply(x) - y64) / np.linalg.norm(y64))
        f.close()
        print(f"problem {i} L{p.level} {p.kind} k={p.op.wavenumber:.3f}: oracle {err:.2e}, "
              f"mixed vs fp64 fro {cmp['rel_fro']:.2e} zx {zx:.2e}", flush=True)
        worst["oracle"] = max(worst["oracle"], err)
        worst["fro"] = max(worst["fro"], cmp["rel_fro"])
        worst["zx"] = max(worst["zx"], zx)
        assert err <= MIXED_TOL and cmp["rel_fro"] <= MIXED_TOL and zx <= MIXED_TOL, i
    print("worst", worst, flush=True)
    b.close()
    for a in asms:
        a.close()


def test_poisson_stream_with_level5_requests(gpu_ctx):
    """48 requests of the Poisson stream (12 of them level 5, 30,720 RWG)
    arriving ten times faster than in the benchmark, on 4 slots: each y is
    bitwise the single-assembly path's Z x, and 4 sampled rows of each agree
    with the oracle's rows times x at 1e-12."""
    reqs = workloads.poisson_stream(48)
    assert sum(r.level == 5 for r in reqs) >= 8
    xs = [r.rhs(i) for i, r in enumerate(reqs)]
    srv = Server(slots=4)
    ids = [None] * len(reqs)
    print("case poisson stream 48 requests", flush=True)
    t0 = time.perf_counter()

    def submitter():
        for i, r in enumerate(reqs):
            dt = t0 + r.arrival_s / 10.0 - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids[i] = srv.submit_assemble_apply(r.op, r.X, xs[i])

    th = threading.Thread(target=submitter)
    th.start()
    ys = []
    lat = []
    for i in range(len(reqs)):
        while ids[i] is None:
            time.sleep(0.001)
        res, y = srv.wait(ids[i])
        assert res["entries"] == reqs[i].X.nfun ** 2
        ys.append(y)
        lat.append(res["total_ms"])
    th.join()
    srv.close()
    print("stream latency p50 %.1f ms p99 %.1f ms" % (np.percentile(lat, 50), np.percentile(lat, 99)), flush=True)
    for i, (r, x, y) in enumerate(zip(reqs, xs, ys)):
        a = Assembler(r.op, r.X, r.X, GPUConfiguration(), ctx=gpu_ctx)
        yr = a.assemble().apply(x)
        a.close()
        assert np.array_equal(y, yr), i
        rows = _sample_rows(r.X.nfun, 4, 500 + i)
        Zr, _ = O.assemble(r.op, r.X, r.X, rows=rows)
        yo = Zr @ x
        err = np.linalg.norm(y[rows] - yo) / np.linalg.norm(yo)
        assert err <= FP64_TOL, (i, err)
