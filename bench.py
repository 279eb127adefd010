"""Benchmark: dense BEM matrix assembly on one MI355X.

Default workload: Maxwell single layer (EFIE) on RWG functions of a level-5
icosphere (20,480 triangles, 30,720 RWG) radially perturbed by 8 seeded real
spherical harmonics (degree <= 6, 3 % of the radius, seed 2), k = 2.0,
ComplexF64 output, FP64 far pairs.  One step = one full assembly
(classification, Sauter-Schwab, far pairs, write-back) with mesh, spaces and
output matrix resident in HBM.

Also reported on the same mesh: the mixed-precision path (FP32 far pairs,
ComplexF32 and ComplexF64 outputs) with its relative Frobenius and Z.x errors
against the FP64 path, the roofline of the far-pair kernel, and the CPU oracle
timed on a bounded sample of the same workload; then the 64-problem batched
EFIE/MFIE workload with each problem's mixed-vs-FP64 error, and the
2,000-request serving stream.  Other workloads (--workload): efie-L3,
efie-L3-k4, efie-L4, ellipsoid-60k, hh3d-cubeD, batched, serving;
--tile-sweep times the far-pair tile shapes and the autotuner's choice.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from beastgpu_amd import Helmholtz3D, Maxwell3D, meshes, spaces, workloads  # noqa: E402
from beastgpu_amd.api import Assembler, Context, GPUConfiguration  # noqa: E402

METRIC = "matrix entries/s, mixed vs FP64 GPU (rel err ≤1e-6); serving p50/p99 latency"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix) peak
HBM_PEAK_GBS = 8000.0

# algorithmic flops per far pair (12 Green evaluations + epilogue), per
# Sauter-Schwab hypercube node; counting FMA = 2, add/mul = 1,
# div/sqrt/sin/cos/exp = 1 each
FLOPS_FAR_PAIR = {"efie": 12 * 52 + 300, "hh3d": 12 * 62 + 27}
FLOPS_SS_NODE = {"efie": 107, "hh3d": 117}
FAR_KERNEL = {"efie": "k_panels_rr<MW, FP64, fixed 3x4 rule, C64> (far pairs + write-back)",
              "hh3d": "k_panels_rr<HH3D, FP64, fixed 3x4 rule, C64> (far pairs + write-back)"}
# HBM bytes per far-kernel launch (FETCH_SIZE + WRITE_SIZE over the launches of
# one assembly, divided by the launches, each counter scaled by its reading on a
# device copy of known size); from rocprofv3 --pmc runs of this script, kept
# in profiles/k_panels_rr_efie_L5_pmc.md
TRAFFIC = {
    "efie-L5": {"bytes_per_launch": 4.598e9, "read_bytes_per_launch": 1.128e9, "write_bytes_per_launch": 3.470e9,
                "algorithmic_bytes_per_launch": 2.517e9, "fetch_scale": 2.0, "write_scale": 1.0,
                "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, FP64 C64 far-pair kernel, 6 launches"},
}


def workload(name: str):
    if name == "efie-L5":
        m = meshes.perturb_radial(meshes.icosphere(5), seed=2, amplitude=0.03)
        return dict(op=Maxwell3D.singlelayer(wavenumber=2.0), X=spaces.raviartthomas(m), kind="efie",
                    desc="EFIE (Maxwell single layer) on RWG, icosphere L5 radially perturbed 3% (8 SH, l<=6, seed 2), "
                         "k=2.0, 20,480 triangles, 30,720 RWG")
    if name == "efie-L4":
        m = meshes.perturb_radial(meshes.icosphere(4), seed=2, amplitude=0.03)
        return dict(op=Maxwell3D.singlelayer(wavenumber=2.0), X=spaces.raviartthomas(m), kind="efie",
                    desc="EFIE on RWG, icosphere L4 radially perturbed 3% (seed 2), k=2.0, 5,120 triangles, 7,680 RWG")
    if name == "efie-L3":
        m = meshes.icosphere(3)
        return dict(op=Maxwell3D.singlelayer(wavenumber=1.0), X=spaces.raviartthomas(m), kind="efie",
                    desc="EFIE on RWG, icosphere L3, k=1.0, 1,280 triangles, 1,920 RWG")
    if name == "efie-L3-k4":
        m = meshes.icosphere(3)
        return dict(op=Maxwell3D.singlelayer(wavenumber=4.0), X=spaces.raviartthomas(m), kind="efie",
                    desc="EFIE on RWG, icosphere L3, k=4.0, 1,280 triangles, 1,920 RWG")
    if name == "ellipsoid-60k":
        m = meshes.ellipsoid()
        return dict(op=Maxwell3D.singlelayer(wavenumber=4.0), X=spaces.raviartthomas(m), kind="efie",
                    desc="EFIE on RWG, triaxial ellipsoid 1.0/0.7/0.5 (cube-sphere grid 95x60x60, seed 3 jitter), "
                         "k=4.0, 60,000 triangles, 90,000 RWG")
    if name.startswith("hh3d-cube"):
        d = int(name[len("hh3d-cube"):])
        m = meshes.meshcuboid(1.0, 1.0, 1.0, 0.5 / d)
        return dict(op=Helmholtz3D.singlelayer(wavenumber=1.0), X=spaces.lagrangec0d1(m), kind="hh3d",
                    desc=f"Helmholtz single layer on P1, meshcuboid(1,1,1,0.5/{d}), k=1.0, {m.nt} triangles, "
                         f"{m.nv} functions")
    raise SystemExit(f"unknown workload {name}")


def time_steps(asm, steps, warmup):
    for _ in range(warmup):
        asm.assemble()
    # the first stats request counts the pairs by class on the host (a
    # diagnostic, not part of an assembly): outside the timed steps
    asm.stats()
    t0 = time.perf_counter()
    far_ms = special_ms = classify_ms = 0.0
    for _ in range(steps):
        asm.assemble()
        st = asm.stats()
        far_ms += st["far_ms"]
        special_ms += st["special_ms"]
        classify_ms += st["classify_ms"]
    dt = time.perf_counter() - t0
    return dt / steps, far_ms / steps, special_ms / steps, classify_ms / steps, asm.stats()


def cpu_baseline(w, seconds=15.0, full_limit=45.0):
    """The oracle (oracle/oracle.c: the project's plain-C FP64 restatement of
    the reference algorithm, OpenMP over all host cores; not BEAST itself,
    which needs Julia) timed on the same workload: the whole matrix when a
    probe projects it to finish within ``full_limit`` seconds, else the rows
    of a compact run of test functions (about ``seconds`` of work), with the
    full-assembly time extrapolated by the cell pairs the sample integrated."""
    from oracle import oracle as O

    X, op = w["X"], w["op"]
    cores = os.cpu_count() or 1
    nt = X.mesh.nt
    rows = np.arange(0, 32)
    Z, counts = O.assemble(op, X, X, rows=rows, nthreads=cores)  # warm-up and rate probe
    t0 = time.perf_counter()
    Z, counts = O.assemble(op, X, X, rows=rows, nthreads=cores)
    dt = time.perf_counter() - t0
    pairs = int(counts.sum())
    rate = pairs / dt
    full_pairs = nt * nt
    entries = X.nfun * X.nfun
    label = "project CPU restatement (oracle/oracle.c, C + OpenMP), not BEAST"
    if full_pairs / rate <= full_limit:
        t0 = time.perf_counter()
        Z, counts = O.assemble(op, X, X, nthreads=cores)
        t_full = time.perf_counter() - t0
        del Z
        return {
            "value": entries / t_full,
            "unit": "matrix entries/s",
            "cores": cores,
            "kind": "port",
            "sample": f"{label}: whole matrix ({X.nfun} x {X.nfun}, {full_pairs} cell pairs) in {t_full:.1f} s",
            "pairs_per_s": full_pairs / t_full,
            "full_assembly_s": t_full,
        }
    want = int(min(max(rate * seconds / (2 * nt), 4), X.nfun))
    rows = np.arange(0, want)
    t0 = time.perf_counter()
    Z, counts = O.assemble(op, X, X, rows=rows, nthreads=cores)
    dt = time.perf_counter() - t0
    pairs = int(counts.sum())
    t_full = dt * full_pairs / pairs
    return {
        "value": entries / t_full,
        "unit": "matrix entries/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{label}: rows 0..{want - 1} of {X.nfun} ({pairs} cell pairs of {full_pairs}, {dt:.1f} s); "
                  f"full assembly extrapolated to {t_full:.0f} s by cell pairs",
        "pairs_per_s": pairs / dt,
        "full_assembly_s": t_full,
    }


def calibrate_copy(mib: int, reps: int = 3):
    """Device copies of a known size: what FETCH_SIZE / WRITE_SIZE read for
    them scales the counters of the assembly kernels to bytes."""
    import torch

    n = mib * (1 << 20)
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n, "reps": reps}), flush=True)


def serving_run(n_req: int, rate: float = 20.0, seed: int = 5, slots: int = 4, precision: str = "fp64",
                output: str = "c64") -> dict:
    """Poisson stream of assemble-and-apply requests through the request loop:
    EFIE on RWG of perturbed icospheres, level 2..5 uniform (log-uniform
    triangle count 320..20,480), amplitude U[0, 5%], k ~ U[0.5, 6], 4 random
    right-hand sides with entries uniform in the unit disk.  Latency is
    arrival -> result, measured against the scheduled arrival time."""
    import threading

    from beastgpu_amd.serving import Server

    reqs = workloads.poisson_stream(n_req, rate=rate, seed=seed)
    levels = np.array([r.level for r in reqs])
    cfg = GPUConfiguration(precision=precision, output=output)
    base = {L: spaces.raviartthomas(meshes.icosphere(L)) for L in range(2, 6)}

    srv = Server(slots=slots)
    solo = {}
    for rep in range(2):  # warm-up, then one request per level alone (untimed stream)
        for L, X in base.items():
            x0 = workloads.StreamRequest(0.0, L, None, X).rhs(0)
            res, _ = srv.wait(srv.submit_assemble_apply(Maxwell3D.singlelayer(wavenumber=1.0), X, x0, cfg))
            if rep == 1:
                solo[f"L{L}"] = {k: res[k] for k in ("total_ms", "prep_ms", "device_ms")}
    ids = [None] * n_req
    lag = np.zeros(n_req)
    t0 = time.perf_counter()

    def submitter():
        for i, r in enumerate(reqs):
            x = r.rhs(i)
            dt = t0 + r.arrival_s - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            lag[i] = time.perf_counter() - (t0 + r.arrival_s)
            ids[i] = srv.submit_assemble_apply(r.op, r.X, x, cfg)

    th = threading.Thread(target=submitter)
    th.start()
    lat = np.zeros(n_req)
    parts = np.zeros((n_req, 3))
    entries = 0
    for i in range(n_req):
        while ids[i] is None:
            time.sleep(0.001)
        res, y = srv.wait(ids[i])
        lat[i] = lag[i] * 1e3 + res["total_ms"]
        parts[i] = (res["queue_ms"], res["prep_ms"], res["device_ms"])
        entries += res["entries"]
        if (i + 1) % 200 == 0:
            print(f"[bench] serving {i + 1}/{n_req} requests, p50 so far {np.percentile(lat[:i + 1], 50):.1f} ms",
                  file=sys.stderr, flush=True)
    t_end = time.perf_counter()
    th.join()
    _, peak = srv.memory()
    srv.close()
    makespan = t_end - t0
    per_level = {}
    for L in base:
        sel = levels == L
        if sel.any():
            per_level[f"L{L}"] = {"n": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                                  "p99_ms": float(np.percentile(lat[sel], 99))}
    return {
        "requests": n_req,
        "arrival_rate_per_s": rate,
        "slots": slots,
        "precision": precision,
        "output": output,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "max_ms": float(lat.max()),
        "entries_per_s": entries / makespan,
        "makespan_s": makespan,
        "mean_queue_ms": float(parts[:, 0].mean()),
        "mean_prep_ms": float(parts[:, 1].mean()),
        "mean_device_ms": float(parts[:, 2].mean()),
        "max_submit_lag_ms": float(lag.max() * 1e3),
        "peak_resident_gb": peak / 1e9,
        "per_level": per_level,
        "alone_ms": solo,
    }


def batched_run(steps: int, warmup: int, n_prob: int = 64, seed: int = 4, errors: bool = True,
                tile=(0, 0)) -> dict:
    """64 independent EFIE and MFIE (n x K) assemblies, mixed precision,
    ComplexF64 output: perturbed icospheres of level 2..4 (amplitude
    U[0, 5%]), k ~ U[0.5, 8] capped so that the longest edge is at most a
    wavelength / 8.  Times the batch (one replayed graph) against the same
    plans assembled one after another; checks every batched matrix bitwise
    against its own single assembly and reports each problem's error against
    an FP64 assembly (relative Frobenius and relative Z x, x uniform in the
    complex unit disk, seed 33).  ``tile``: the far-pair tile shape of every
    problem, "auto" = the tile autotuner per problem signature (its timing
    is part of assemblers_setup_s, its choices are reported per level and
    operator in schedule.tiles)."""
    from beastgpu_amd.api import Batch, DeviceMatrix

    from beastgpu_amd import _bmi as _b

    probs = workloads.batched_efie_mfie(n_prob, seed)
    ctx = Context(0)
    t0 = time.perf_counter()
    asms = [Assembler(p.op, p.X, p.X, GPUConfiguration(precision="mixed", tile=tile), ctx=ctx) for p in probs]
    t_setup = time.perf_counter() - t0
    tiles = {}
    for p, a in zip(probs, asms):
        tiles.setdefault(f"L{p.level} {p.kind}", set()).add(a.tile)
    entries = sum(a.matrix.rows * a.matrix.cols for a in asms)
    for _ in range(warmup):
        for a in asms:
            a.assemble()
    t0 = time.perf_counter()
    for _ in range(steps):
        for a in asms:
            a.assemble()
    t_seq = (time.perf_counter() - t0) / steps
    # each plan's own assembly, kept for the bitwise check
    singles = []
    for a in asms:
        s = DeviceMatrix(ctx, a.matrix.rows, a.matrix.cols)
        _b.check(_b.lib().bmi_assemble(a.plan.h, s.h))
        singles.append(s)
    t0 = time.perf_counter()
    b = Batch(asms)
    t_create = time.perf_counter() - t0
    for _ in range(warmup):
        b.assemble()
    t0 = time.perf_counter()
    dev = 0.0
    for _ in range(steps):
        b.assemble()
        dev += b.stats()["last_ms"]
    t_b = (time.perf_counter() - t0) / steps
    same = all(a.matrix.compare(s)["diff_fro"] == 0.0 for a, s in zip(asms, singles))
    for s in singles:
        s.close()
    st = b.stats()
    per = []
    if errors:
        for p, a in zip(probs, asms):
            f = Assembler(p.op, p.X, p.X, GPUConfiguration(), ctx=ctx)
            M64 = f.assemble()
            cmp = a.matrix.compare(M64)
            x = workloads.unit_disk_vector(p.X.nfun, 33)
            y64 = M64.apply(x)
            zx = float(np.linalg.norm(a.matrix.apply(x) - y64) / np.linalg.norm(y64))
            f.close()
            per.append([p.level, p.kind, round(p.op.wavenumber, 4), cmp["rel_fro"], zx])
    b.close()
    for a in asms:
        a.close()
    out = {
        "problems": n_prob,
        "entries": int(entries),
        "batched_ms": t_b * 1e3,
        "batched_device_ms": dev / steps,
        "sequential_ms": t_seq * 1e3,
        "speedup_vs_sequential": t_seq / t_b,
        "entries_per_s": entries / t_b,
        "create_and_tune_s": t_create,
        "assemblers_setup_s": t_setup,
        "bitwise_equal_to_single_assemblies": bool(same),
        "checked_problems": n_prob,
        "schedule": dict(st, tile=str(tile), tiles={k: sorted(map(list, v)) for k, v in sorted(tiles.items())}),
        "levels": {f"L{L}": sum(1 for p in probs if p.level == L) for L in (2, 3, 4)},
    }
    if per:
        out["mixed_vs_fp64"] = {
            "max_rel_fro": max(r[3] for r in per),
            "max_rel_zx": max(r[4] for r in per),
            "per_problem": per,
            "columns": ["level", "operator", "k", "rel_fro", "rel_zx"],
        }
    return out


TILE_SHAPES = {"efie": [(0, 0), (48, 40), (32, 40), (64, 28), (32, 24)],
               "hh3d": [(0, 0), (96, 20), (64, 20), (128, 12)]}


def tile_sweep(steps: int, warmup: int, names=("efie-L3", "efie-L4", "efie-L5", "hh3d-cube24")) -> list:
    """Far-pair tile shapes (panel rows, patch columns; (0, 0) = the default
    (64, 40) on RWG, (128, 20) on Lagrange spaces) on each workload: FP64 and
    mixed C32 time per assembly, and each shape's FP64 matrix against the
    default shape's (relative Frobenius)."""
    ctx = Context(0)
    rows = []
    for name in names:
        w = workload(name)
        ref = None
        for shape in TILE_SHAPES[w["kind"]]:
            print(f"[bench] tile sweep {name} shape {shape}", file=sys.stderr, flush=True)
            a = Assembler(w["op"], w["X"], w["X"], GPUConfiguration(tile=shape), ctx=ctx)
            t, far, _, _, _ = time_steps(a, steps, warmup)
            if ref is None:
                ref, rel = a, 0.0
            else:
                rel = a.matrix.compare(ref.matrix)["rel_fro"]
            m = Assembler(w["op"], w["X"], w["X"], GPUConfiguration(tile=shape, precision="mixed", output="c32"),
                          ctx=ctx)
            tm, farm, _, _, _ = time_steps(m, steps, warmup)
            m.close()
            rows.append({"workload": name, "tile": list(shape), "fp64_ms": round(t * 1e3, 3),
                         "fp64_far_ms": round(far, 3), "mixed_c32_ms": round(tm * 1e3, 3),
                         "mixed_c32_far_ms": round(farm, 3), "fp64_rel_fro_vs_default": rel})
            if a is not ref:
                a.close()
        print(f"[bench] tile sweep {name} auto", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        a = Assembler(w["op"], w["X"], w["X"], GPUConfiguration(tile="auto"), ctx=ctx)
        tune_s = time.perf_counter() - t0
        t, far, _, _, _ = time_steps(a, steps, warmup)
        rows.append({"workload": name, "tile": "auto", "chosen": list(a.tile), "tune_s": round(tune_s, 3),
                     "candidate_ms": {f"{r}x{c}": round(v, 3) for (r, c), v in a.tile_times.items()},
                     "fp64_ms": round(t * 1e3, 3), "fp64_far_ms": round(far, 3),
                     "fp64_rel_fro_vs_default": a.matrix.compare(ref.matrix)["rel_fro"]})
        a.close()
        ref.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="efie-L5")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-to-host", dest="to_host", action="store_false",
                    help="skip timing plan + assembly + the copy into a host matrix")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serving-requests", type=int, default=2000,
                    help="requests of the serving stream in the default run (0: skip)")
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--no-batched", dest="batched", action="store_false",
                    help="skip the 64-problem batched sample (--workload batched runs it alone)")
    ap.add_argument("--serving-total", type=int, default=2000, help="requests of --workload serving")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    ap.add_argument("--tile-sweep", action="store_true",
                    help="only time the far-pair tile shapes and tile=\"auto\" on efie-L3/L4/L5 and hh3d-cube24, then exit")
    args = ap.parse_args()
    if args.tile_sweep:
        print(json.dumps({"tile_sweep": tile_sweep(args.steps, args.warmup),
                          "columns": "ms per assembly (wall clock, mean over the steps); far ms from HIP events; tile (0, 0) = default"}),
              flush=True)
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: multi-GPU runs are not part of this build "
              "(assembly shards by test rows with no exchange)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        bt = batched_run(args.steps, args.warmup)
        print(json.dumps({"metric": "matrix entries/s, 64 batched EFIE + MFIE assemblies (mixed)",
                          "value": bt["entries_per_s"], "unit": "matrix entries/s", "n_gpus": 1,
                          "higher_is_better": True, "data": "synthetic", "dtype": "f32/f64 mixed",
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": bt["batched_ms"],
                          "config": {"workload": "64 EFIE/MFIE assemblies, icosphere L2-L4 perturbed U[0,5%] "
                                                 "(seed 4), k ~ U[0.5, 8] capped at lambda/8 >= max edge, "
                                                 "mixed precision, C64; step = one batch"},
                          "batched": bt}), flush=True)
        return
    if args.workload == "serving":
        sv = serving_run(args.serving_total, slots=args.serving_slots)
        print(json.dumps({"metric": "serving p50/p99 latency (ms), assemble-and-apply requests",
                          "value": sv["p50_ms"], "unit": "ms", "n_gpus": 1, "higher_is_better": False,
                          "data": "synthetic", "dtype": "f64",
                          "config": {"workload": f"Poisson arrivals 20 req/s (seed 5), {args.serving_total} EFIE assemble-and-apply "
                                                 "requests on perturbed icospheres L2-L5, k in [0.5, 6], 4 RHS"},
                          "serving": sv}), flush=True)
        return
    w = workload(args.workload)
    X, op, kind = w["X"], w["op"], w["kind"]
    n = X.nfun
    entries = n * n
    print(f"[bench] workload {args.workload}: {w['desc']}", file=sys.stderr, flush=True)
    ctx = Context(0)
    asm = Assembler(op, X, X, GPUConfiguration(), ctx=ctx)
    t_step, far_ms, special_ms, classify_ms, st = time_steps(asm, args.steps, args.warmup)
    print(f"[bench] fp64 step {t_step * 1e3:.2f} ms (far {far_ms:.2f}, special {special_ms:.2f}, "
          f"classify {classify_ms:.2f}) {st}", file=sys.stderr, flush=True)
    pairs = st["pairs"]
    far_pairs = pairs["far"]
    far_flops = far_pairs * FLOPS_FAR_PAIR[kind]
    achieved = far_flops / (far_ms * 1e-3) / 1e12

    result = {
        "metric": METRIC,
        "value": entries / t_step,
        "unit": "matrix entries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": w["desc"] + ", ComplexF64 output, FP64 far pairs; step = one full assembly, inputs resident",
            "matrix": [n, n],
            "precision": "fp64",
            "quadrature": "far 3x4-point triangle rules, Sauter-Schwab 5-point Gauss-Legendre",
        },
        "pairs_per_s": X.mesh.nt ** 2 / t_step,
        "green_evals_per_s": (far_pairs * 12) / t_step,
        "phases_ms": {"classify": classify_ms, "special": special_ms, "far_blocks": far_ms},
        "pair_counts": pairs,
        "roofline": {
            "bound": "valu",
            "pipe": "VALU FP64 (no MFMA: sqrt/sincos evaluation); peak = MI355X FP64 78.6 TFLOP/s",
            "kernel": FAR_KERNEL[kind],
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": TRAFFIC.get(args.workload),
            "flops_per_far_pair": FLOPS_FAR_PAIR[kind],
            "far_launches_per_step": st["far_launches"],
            "avg_launch_ms": far_ms / max(1, st["far_launches"]),
            "flops_per_launch": far_flops / max(1, st["far_launches"]),
            "write_gbs": entries * 16 / (far_ms * 1e-3) / 1e9,
        },
    }

    # end to end as a caller sees it: mesh, space, plan and matrix set up,
    # then one assembly (plan + assembly); then also the copy of the whole
    # matrix into a host ComplexF64 array, as the reference's store-on-GPU
    # path returns it; and the streamed assembly straight into host memory
    # (the store-on-CPU path: row blocks through device staging, copies
    # overlapped with the next block's kernels).  An untimed run first sizes
    # the device memory pool and touches the host array's pages.
    a2 = Assembler(op, X, X, GPUConfiguration(), ctx=ctx)
    a2.assemble()
    a2.close()
    t0 = time.perf_counter()
    a2 = Assembler(op, X, X, GPUConfiguration(), ctx=ctx)
    a2.assemble()
    t_plan_asm = time.perf_counter() - t0
    e2e = {"plan_and_assembly_ms": t_plan_asm * 1e3}
    if args.to_host:
        host = np.empty((n, n), dtype=np.complex128, order="F")
        from beastgpu_amd import _bmi as _b
        _b.check(_b.lib().bmi_matrix_to_host(a2.matrix.h, host.ctypes.data, _b.C64))  # touch the pages
        a2.close()
        t0 = time.perf_counter()
        a2 = Assembler(op, X, X, GPUConfiguration(), ctx=ctx)
        a2.assemble()
        t1 = time.perf_counter()
        _b.check(_b.lib().bmi_matrix_to_host(a2.matrix.h, host.ctypes.data, _b.C64))
        t2 = time.perf_counter()
        e2e["plan_assembly_and_host_copy_ms"] = (t2 - t0) * 1e3
        e2e["host_copy_gbs"] = entries * 16 / (t2 - t1) / 1e9
        # streamed: plan without a device matrix, staging blocks under the
        # reference's default pipeline budget (24 GiB)
        a2.close()
        t0 = time.perf_counter()
        a3 = Assembler(op, X, X, GPUConfiguration(writeBackStrategy="cpu"), ctx=ctx, resident=False)
        a3.assemble_to_host(out=host)
        t3 = time.perf_counter()
        e2e["streamed_plan_assembly_to_host_ms"] = (t3 - t0) * 1e3
        e2e["streamed_device_ms"] = a3.stats()["total_ms"]
        a3.close()
        print(f"[bench] end to end {e2e}", file=sys.stderr, flush=True)
        del host
    else:
        a2.close()
    result["end_to_end"] = e2e
    if not args.no_mixed:
        rng = np.random.default_rng(33)
        x = np.sqrt(rng.uniform(size=n)) * np.exp(2j * np.pi * rng.uniform(size=n))
        y64 = asm.matrix.apply(x)
        mixed = {}
        for out in ("c32", "c64"):
            am = Assembler(op, X, X, GPUConfiguration(precision="mixed", output=out), ctx=ctx)
            tm, fm, sm, cm, stm = time_steps(am, args.steps, args.warmup)
            cmp = am.matrix.compare(asm.matrix)
            ym = am.matrix.apply(x)
            mixed[out] = {
                "value": entries / tm,
                "ms_per_step": tm * 1e3,
                "speedup_vs_fp64": t_step / tm,
                "rel_fro_err_vs_fp64": cmp["rel_fro"],
                "rel_zx_err_vs_fp64": float(np.linalg.norm(ym - y64) / np.linalg.norm(y64)),
                "phases_ms": {"classify": cm, "special": sm, "far_blocks": fm},
                "far_fp64_pairs": stm["pairs"]["far_fp64"],
            }
            print(f"[bench] mixed {out}: {mixed[out]}", file=sys.stderr, flush=True)
            am.close()
        result["mixed"] = mixed

    asm.close()
    if args.batched and args.workload == "efie-L5":
        bt = batched_run(3, 1)
        print(f"[bench] batched {bt}", file=sys.stderr, flush=True)
        result["batched"] = bt
    if args.serving_requests > 0 and args.workload == "efie-L5":
        sv = serving_run(args.serving_requests, slots=args.serving_slots)
        print(f"[bench] serving {sv}", file=sys.stderr, flush=True)
        result["serving"] = sv
    if not args.no_cpu:
        cb = cpu_baseline(w, args.cpu_seconds)
        print(f"[bench] cpu baseline {cb}", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cb
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
