#!/usr/bin/env python3
"""Benchmark: N=32 S=1/2 Kitaev-Heisenberg H.v on one MI355X (HBM GB/s vs roofline).

Workload (the single-GPU configuration the headline metric is quoted on):
honeycomb 4x4 PBC cluster (N=32, 48 bonds), MODEL 20 with phi = 0.7 pi (J = cos phi,
K = 2 sin phi), [111] field h = 0.1, complex128 state of dim 2^32 (64 GiB per vector,
three resident vectors as a Lanczos run holds them), seeded Gaussian start vector.

A step is one H.v in the reference's accumulate form, p_mltply_v3: v0 += H v1 with
<v1|H|v1> (S_one_half/Hybrid_TPS_parallel.f90:190-263).  Algorithmic bytes per H.v are
3*16*2^32 (read v1, read v0, write v0).  ms_per_step and the roofline figure come from
HIP events recorded on the library's own stream around the timed H.v calls; the wall
clock around the same region (barrier + synchronize on both sides) gives `value`.

cpu_baseline: the oracle's restatement of the reference's per-element pass structure
(OpenMP, all host threads) on a bounded sample -- output rows [0, 2^24) of the same
N=32 operator -- reported as algorithmic H.v GB/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--cfg N32|N24]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=1 << 24)
    ap.add_argument("--cfg", default="N32")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune", default="", help="fixed plan K,L,threads (skips the autotuner)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device copies of a 4 GiB vector (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(every measured figure is for one GPU)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pned_amd import gpu, lattice, model

    if args.calibrate_copy:
        op = model.build_operator(lattice.torus(2, 7), 20, 126.0)   # N=28: 4 GiB per vector
        ctx = gpu.Context(op)
        a, b = ctx.vector(), ctx.vector()
        ctx.set_initial_vec(a, 1)
        ctx.sync()
        for _ in range(3):
            b.copy_from(a)
        ctx.sync()
        print(json.dumps({"calibrate_copy_bytes": 16 * ctx.dim, "copies": 3}))
        return

    if args.cfg == "N32":
        cl = lattice.torus(4, 4)
        phi_deg, h, nres = 126.0, 0.1, 3
        workload = "S=1/2 KH honeycomb 4x4 PBC N=32, phi=0.7pi, [111] h=0.1, complex128 dim 2^32, H.v accumulate"
    else:
        cl = lattice.reference_cluster(24)
        phi_deg, h, nres = 54.0, 0.0, 3
        workload = "S=1/2 KH reference 24-site PBC, phi=0.3pi, complex128 dim 2^24, H.v accumulate"
    op = model.build_operator(cl, 20, phi_deg, field=h)
    ctx = gpu.Context(op)
    vecs = [ctx.vector() for _ in range(nres)]
    v1, v0 = vecs[0], vecs[1]
    ctx.set_initial_vec(v1, 20261202)
    tune = None
    if args.tune:
        k, l, t = (int(x) for x in args.tune.split(","))
        ctx.tune(k, l, t)
        tune = {"tile_bits": k, "low_bits": l, "threads": t}
    elif not args.no_autotune:
        tune = ctx.autotune(v1, v0, reps=1)
    v0.zero()
    npass, _ = ctx.plan_info()

    stream = torch.cuda.ExternalStream(ctx.stream())
    for _ in range(args.warmup):
        ctx.p_mltply_v3(v1, v0)
    ctx.sync()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    prd = None
    for _ in range(args.steps):
        prd = ctx.p_mltply_v3(v1, v0)
    e1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1) / args.steps
    ms = wall * 1e3 / args.steps
    D = ctx.dim
    alg_bytes = 3 * 16 * D
    value = alg_bytes / (ms * 1e-3) / 1e9
    achieved = alg_bytes / (ev_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        import oracle
        dt, rows, threads = oracle.cpu_rows_sample(op, args.cpu_rows)
        cpu = {"value": round(3 * 16 * rows / dt / 1e9, 4), "unit": "GB/s", "cores": int(threads),
               "kind": "port",
               "sample": f"oracle reference-pass-structure H.v (OpenMP) on output rows [0,{rows}) of the "
                         f"same N=32 operator, {dt:.2f} s"}

    out = {
        "metric": "N=32 H·v HBM GB/s (% peak); batched/mixed solves/s + p99 vs plain @1e-10",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded Philox Gaussian start vector)",
        "config": {"workload": workload, "n_sites": cl.n, "dim": D, "bonds": cl.ibond,
                   "resident_vectors": nres, "passes": npass, "tune": tune,
                   "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 2),
                   "prdct_re": float(np.real(prd))},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
