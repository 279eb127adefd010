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
timed on a bounded sample of the same workload.

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

from beastgpu_amd import Helmholtz3D, Maxwell3D, meshes, spaces  # noqa: E402
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
    "efie-L5": {"bytes_per_launch": 4.939e9, "read_bytes_per_launch": 1.469e9, "write_bytes_per_launch": 3.470e9,
                "algorithmic_bytes_per_launch": 2.517e9, "fetch_scale": 2.0, "write_scale": 1.0,
                "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, FP64 C64 far-pair kernel, 6 launches"},
}


def workload(name: str):
    if name == "efie-L5":
        m = meshes.perturb_radial(meshes.icosphere(5), seed=2, amplitude=0.03)
        return dict(op=Maxwell3D.singlelayer(wavenumber=2.0), X=spaces.raviartthomas(m), kind="efie",
                    desc="EFIE (Maxwell single layer) on RWG, icosphere L5 radially perturbed 3% (8 SH, l<=6, seed 2), "
                         "k=2.0, 20,480 triangles, 30,720 RWG")
    if name == "efie-L3":
        m = meshes.icosphere(3)
        return dict(op=Maxwell3D.singlelayer(wavenumber=1.0), X=spaces.raviartthomas(m), kind="efie",
                    desc="EFIE on RWG, icosphere L3, k=1.0, 1,280 triangles, 1,920 RWG")
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


def cpu_baseline(w, seconds=15.0):
    """The oracle (C, OpenMP over all host cores) on a sample of test cells:
    the rows of a compact run of test functions.  Full-assembly time is
    extrapolated by the number of cell pairs the sample integrated."""
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
    want = int(min(max(rate * seconds / (2 * nt), 4), X.nfun))
    rows = np.arange(0, want)
    t0 = time.perf_counter()
    Z, counts = O.assemble(op, X, X, rows=rows, nthreads=cores)
    dt = time.perf_counter() - t0
    pairs = int(counts.sum())
    full_pairs = nt * nt
    t_full = dt * full_pairs / pairs
    entries = X.nfun * X.nfun
    return {
        "value": entries / t_full,
        "unit": "matrix entries/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle rows 0..{want - 1} of {X.nfun} ({pairs} cell pairs of {full_pairs}, {dt:.1f} s); "
                  f"full assembly extrapolated to {t_full:.0f} s by cell pairs",
        "pairs_per_s": pairs / dt,
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="efie-L5")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: multi-GPU runs are not part of this build "
              "(assembly shards by test rows with no exchange)", file=sys.stderr)
        sys.exit(2)

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
            "bound": "mfma",
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
    if not args.no_cpu:
        cb = cpu_baseline(w, args.cpu_seconds)
        print(f"[bench] cpu baseline {cb}", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cb
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
