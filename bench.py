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

One H.v is P launches of the tile-pass kernel (P = passes of the plan) plus one small
reduction; `roofline.achieved` is the algorithmic bytes of the H.v over the event time of
those launches, `roofline.per_launch` the same per pass launch.  `roofline.traffic` is the
measured HBM bytes per H.v from the committed counter summary
(profiles/hv_s12_N32_counters.json: FETCH_SIZE and WRITE_SIZE per launch, calibrated on
device copies of known size) when that summary was taken with the same tile plan, else null.

cpu_baseline: the oracle's restatement of the reference's per-element pass structure
(OpenMP, all host threads) on a bounded sample -- output rows [0, 2^27) of the same
N=32 operator, about 10-15 s on 16 cores -- reported as algorithmic H.v GB/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--cfg N32|N24|TPS|SERVE|...]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=1 << 27)
    ap.add_argument("--cfg", default="N32",
                    help="N32 (default), N24, TPS, SERVE, LANCZOS, LANCZOS32, SPIN1, SPINHV, GRAPHS")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune", default="", help="fixed plan K,L,threads (skips the autotuner)")
    ap.add_argument("--requests", type=int, default=120, help="SERVE: number of requests")
    ap.add_argument("--serve-big", action="store_true", help="SERVE: include the N=28 (2x7) size")
    ap.add_argument("--no-extras", action="store_true",
                    help="default run without the batched-mTPQ and serving objects (profiling runs)")
    ap.add_argument("--batched-steps", type=int, default=40,
                    help="mTPQ steps of both paths in the default run's batched object")
    ap.add_argument("--serve-requests", type=int, default=40,
                    help="requests per mode in the default run's serving object")
    ap.add_argument("--calibrate-spin", action="store_true",
                    help="SPINHV: the same cluster with diagonal-only bond matrices -- the same tile passes "
                         "and runs with no partner reads: a 3-stream copy of known bytes for counter calibration "
                         "and the streaming ceiling of these tile layouts")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device copies and 3 elementwise y += x of 4 GiB vectors "
                         "(counter calibration and the plain-streaming reference)")
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
        # plain 3-stream kernel of known size: y += x over 2^28 complex128
        x = torch.zeros(1 << 28, dtype=torch.complex128, device="cuda")
        y = torch.zeros_like(x)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        y.add_(x)
        e0.record()
        for _ in range(3):
            y.add_(x)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        print(json.dumps({"calibrate_copy_bytes": 16 * ctx.dim, "copies": 3,
                          "rmw_bytes": 3 * 16 * (1 << 28), "rmw_ms": round(ms, 4),
                          "rmw_GBps": round(3 * 16 * (1 << 28) / (ms * 1e-3) / 1e9, 1)}))
        return

    if args.cfg == "TPS":
        return bench_tps(args)
    if args.cfg == "SPINHV":
        return bench_spin_hv(args)
    if args.cfg == "LANCZOS":
        return bench_lanczos(args)
    if args.cfg == "LANCZOS32":
        return bench_lanczos32(args)
    if args.cfg == "SPIN1":
        return bench_spin1(args)
    if args.cfg == "SERVE":
        return bench_serve(args)
    if args.cfg == "GRAPHS":
        return bench_graphs(args)

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

    stream = torch.cuda.ExternalStream(ctx.stream())
    for _ in range(max(args.warmup, 1)):
        ctx.p_mltply_v3(v1, v0)
    ctx.sync()
    npass, _ = ctx.plan_info()   # the plan the timed calls use (built by the first call)
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

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hv_s12_N32_counters.json")
    if args.cfg == "N32" and os.path.exists(tfile):
        with open(tfile) as fh:
            prof = json.load(fh)
        plan = [tune.get("tile_bits"), tune.get("low_bits"), tune.get("threads")] if tune else None
        for entry in prof.get("plans", []):
            if (plan is not None and entry.get("tune") == plan and prof.get("kernel") == "hv_s12_tm"
                    and entry.get("launches_per_hv") == npass and entry.get("layout") == [int(b) for b in ctx.layout()]):
                traffic = int(entry["hbm_bytes_per_hv"])

    extras = {}
    if not args.no_extras and args.cfg == "N32":
        for v in vecs:
            v.free()
        ctx.close()
        extras["batched"] = batched_block(args)
        extras["serving"] = serving_block(args)
        extras["graphs"] = graphs_block(tps_cases=((2, 4, 8, "c128"),), lnc1_cases=("N=16",))

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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "hv_s12_tm", "launches_per_unit": npass,
                     "per_launch": round(achieved * npass, 1),
                     "achieved_is": "algorithmic bytes 3*16*dim per H.v over its HIP-event time",
                     "dram_GBps": round(traffic / (ev_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "dram_frac": round(traffic / (ev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None},
        "cpu_baseline": cpu,
    }
    out["config"]["pct_hbm_peak_is"] = "value (algorithmic GB/s from wall time) over 8 TB/s"
    out.update(extras)
    print(json.dumps(out), flush=True)


def batched_block(args):
    """The batched-mTPQ half of the metric at its configured size: 3x4 torus (N=24), MODEL 20
    phi = 0.5 pi (J = 0, K1 = +2 in H = sum K S^g S^g: antiferromagnetic Kitaev; the
    ferromagnetic K = -2 is phi = 1.5 pi and has the same spectrum), h = 0, R = 64 samples,
    (l - H/N)^k with l = 1.2 max|bond sum| / N (per-bond shift 1.2 max_b ||h_b||).  Both
    paths run the same k steps on the same 64 start vectors, both measured: fused complex64
    (one H.v launch per pass for all columns, FP64 reductions) and the plain FP64 GPU path
    (complex128, one sample at a time, the reference's step structure).  Errors of the
    complex64 run against the FP64 path: E/N and T of every step, and the cTPQ C(T) of all
    64 samples on the temperatures k steps resolve."""
    import numpy as np

    from pned_amd import gpu, lattice, model

    cl = lattice.torus(3, 4)
    lb = model.bond_norm_shift(cl, 20, 90.0)
    op = model.build_operator(cl, 20, 90.0, tps=True, tps_lb=lb)
    L = op.shift_L
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    R, k = 64, args.batched_steps
    seeds = [20261202 + 7919 * c for c in range(R)]
    tv1, tv0 = ctx.vector(R, gpu.C64), ctx.vector(R, gpu.C64)
    for c in range(R):
        ctx.set_initial_column(tv1, c, seeds[c])
    tune64 = ctx.autotune(tv1, tv0, reps=1)
    tv1.free()
    tv0.free()
    v1, va, vb = (ctx.vector(R, gpu.C64) for _ in range(3))
    for c in range(R):
        ctx.set_initial_column(v1, c, seeds[c])
    ctx.sync()
    t = time.perf_counter()
    a64, b64, c64 = ctx.tps_fused(v1, va, vb, L, k)
    ctx.sync()
    t64 = time.perf_counter() - t
    for v in (v1, va, vb):
        v.free()
    _progress(f"batched: fused complex64, {R} samples x {k} steps in {t64:.3f} s")
    p1, p0, pg, pt = (ctx.vector(1, gpu.C128) for _ in range(4))
    ctx.set_initial_vec(p1, seeds[0])
    tune128 = ctx.autotune(p1, p0, reps=1)
    ap_, bp_, cp_ = [], [], []
    tplain = 0.0
    for c in range(R):
        ctx.set_initial_vec(p1, seeds[c])
        ctx.sync()
        t = time.perf_counter()
        a, b, co = ctx.tps(p1, p0, pg, pt, L, k)
        ctx.sync()
        tplain += time.perf_counter() - t
        ap_.append(np.asarray(a).reshape(-1))
        bp_.append(np.asarray(b).reshape(-1))
        cp_.append(np.asarray(co).reshape(k, -1).T)
    for v in (p1, p0, pg, pt):
        v.free()
    ctx.close()
    _progress(f"batched: plain complex128, {R} samples x {k} steps in {tplain:.3f} s")
    A64, B64, C64 = gpu.tps_arrays(a64, b64, c64, R, c64.shape[-1])
    AP, BP, CP = np.array(ap_), np.array(bp_), np.array(cp_)
    n = cl.n
    ks = np.arange(1, k + 1)
    E64, EP = A64 / n, AP / n
    T64 = (L - A64) / (2.0 * ks)
    TP = (L - AP) / (2.0 * ks)
    # cTPQ temperatures the k-step series resolves: (beta L)^(2k)/(2k)! small at the top term
    t_lo = 2.0 * L / k
    cq64 = gpu.ctpq(A64, B64, C64, t_lo, 20.0 * t_lo, 20)
    cqp = gpu.ctpq(AP, BP, CP, t_lo, 20.0 * t_lo, 20)
    return {
        "metric": "batched mTPQ samples x steps per second, fused complex64 over the plain FP64 GPU path",
        "speedup": round(tplain / t64, 3),
        "fused_c64_s": round(t64, 3), "plain_c128_s": round(tplain, 3),
        "sample_steps_per_s": {"fused_c64": round(R * k / t64, 1), "plain_c128": round(R * k / tplain, 1)},
        "workload": f"mTPQ 3x4 torus N=24, phi=0.5pi (K1=+2, AFM in H=sum K S S), h=0, R={R}, {k} steps both "
                    f"paths (measured), l=1.2 max|bond sum|/N={L / n:.4f}",
        "tune": {"fused_c64": tune64, "plain_c128": tune128},
        "max_abs_err_E_per_site": float(np.max(np.abs(E64 - EP))),
        "max_rel_err_T": float(np.max(np.abs(T64 - TP) / np.abs(TP))),
        "ctpq_T_range": [round(t_lo, 4), round(20.0 * t_lo, 4)],
        "ctpq_max_abs_err_C": float(np.max(np.abs(cq64[:, 1] - cqp[:, 1]))),
        "ctpq_max_C": float(np.max(np.abs(cqp[:, 1]))),
        "error_reference": "plain FP64 path (complex128), which the GPU tests hold to the oracle at 1e-10",
    }


def serving_block(args):
    """The serving half of the metric on the configured size mix {2x4, 3x3, 2x5, 3x4, 2x7}
    (N = 16..28), phi ~ U[0, 2pi), h ~ U[0, 0.5], tol 1e-10: closed-loop capacity of
    continuous batching and of sequential plain solves, then both on the same Poisson
    arrivals at 80 % of the sequential capacity (p50/p99 latency)."""
    from pned_amd import serving

    specs = ["2x4", "3x3", "2x5", "3x4", "2x7"]
    nreq = args.serve_requests
    warm = serving.make_requests(len(specs), specs, seed=1)
    serving.Server(tol=1e-10).run(warm)
    serving.Server(tol=1e-10, sequential=True).run(warm)

    def capacity(sequential):
        reqs = serving.make_requests(nreq, specs, seed=20261202)
        t = time.perf_counter()
        res = serving.Server(tol=1e-10, sequential=sequential).run(reqs)
        return len(res) / (time.perf_counter() - t), res

    cap_b, res_b = capacity(False)
    cap_s, res_s = capacity(True)
    _progress(f"serving: capacity {cap_b:.3f} (continuous batching) vs {cap_s:.3f} (sequential) solves/s")
    rate = 0.8 * cap_s

    def open_loop(sequential):
        reqs = serving.make_requests(nreq, specs, seed=20261203, rate=rate)
        t = time.perf_counter()
        res = serving.Server(tol=1e-10, sequential=sequential).run(reqs)
        return serving.summarize(res, time.perf_counter() - t)

    lb = open_loop(False)
    ls = open_loop(True)
    same = max(abs(a.e0 - b.e0) / abs(b.e0) for a, b in zip(res_b, res_s))
    for d in (lb, ls):
        d.pop("per_size", None)
    return {"metric": "mixed-size ground-state serving, continuous batching vs sequential plain solves",
            "sizes": specs, "requests_per_run": nreq, "tol": 1e-10,
            "capacity_solves_per_s": {"batching": round(cap_b, 3), "sequential": round(cap_s, 3),
                                      "ratio": round(cap_b / cap_s, 3)},
            "arrival_rate_per_s": round(rate, 3),
            "latency": {"batching": {"p50_s": lb["p50_s"], "p99_s": lb["p99_s"]},
                        "sequential": {"p50_s": ls["p50_s"], "p99_s": ls["p99_s"]}},
            "max_rel_diff_e0_batching_vs_sequential": same,
            "note": f"{nreq} requests per run, not the configured 2000 (one 2000-request run per mode "
                    "would take about an hour at these rates); p99 of a short run is close to its maximum"}


def _ground_state(ctx, gpu, v1, v0, x, seed, max_itr):
    """Lanczos until the lowest eigenvalue converges (the reference's criterion), then
    3x as many fixed steps so the Ritz vector's residual is at round-off, and the Ritz
    vector itself (a second Lanczos from the same start)."""
    ctx.set_initial_vec(v1, seed)
    t = time.perf_counter()
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, max_itr=max_itr, nvec=1)
    ctx.set_initial_vec(v1, seed)
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, fixed_steps=min(3 * itr, max_itr), nvec=1)
    coef = gpu.vec12(alpha, beta, [ene[0]], 1)[0]
    ctx.set_initial_vec(v1, seed)
    ctx.ritz_vector(v1, v0, x, alpha, beta, coef)
    ctx.sync()
    dt = time.perf_counter() - t
    res = ctx.residual(x, ene[0], v0)
    return itr, float(ene[0]), res, dt


def bench_spin_hv(args):
    """The spin-S H.v alone on the spin-1 workload (reference N=18 cluster, 3^18 states,
    MODEL 20 phi=108 deg, complex128): warmup + timed H.v (v0 += H v1), algorithmic bytes
    3 * 16 * 3^18 per pass-equivalent stream set."""
    import torch  # noqa: F401

    from pned_amd import gpu, lattice, model

    import dataclasses

    import numpy as np

    cl = lattice.reference_cluster(18)
    op = model.build_operator(cl, 20, 108.0, q=3)
    if args.calibrate_spin:
        # keep each bond matrix's diagonal only: the planner covers the same bonds with the
        # same tiles, and every pass streams v1, v0 -> v0 without partner reads
        eye = np.eye(op.q * op.q, dtype=bool)
        op = dataclasses.replace(op, xlocal=np.where(eye, op.xlocal, 0.0), llocal=np.where(eye, op.llocal, True))
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 1)
    ctx.hv_fresh(v1, v0)
    for _ in range(args.warmup):
        ctx.p_mltply_v3(v1, v0)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(args.steps):
        ctx.p_mltply_v3(v1, v0)
    ctx.sync()
    ms = (time.perf_counter() - t) / args.steps * 1e3
    npass = ctx.plan_info()[0]
    alg = 3 * 16 * ctx.dim
    if args.calibrate_spin:
        print(json.dumps({"calibrate_spin": "diagonal-only spin-1 N=18 passes (copy-like, same tile runs)",
                          "ms_per_hv": round(ms, 3), "passes": npass,
                          "bytes_per_pass": {"fetch": 2 * 16 * ctx.dim, "write": 16 * ctx.dim},
                          "per_pass_GBps": round(alg * npass / ms / 1e6, 1)}), flush=True)
        return
    print(json.dumps({"metric": "spin-1 N=18 H·v GB/s (algorithmic 48 B/state)", "value": round(alg / ms / 1e6, 2),
                      "unit": "GB/s", "higher_is_better": True, "ms_per_step": round(ms, 3), "steps": args.steps,
                      "warmup": args.warmup, "dtype": "c128",
                      "config": {"workload": "spin-1 KH honeycomb N=18 reference cluster, phi=0.6pi, complex128 "
                                             "dim 3^18", "dim": ctx.dim, "passes": npass,
                                 "per_pass_GBps": round(alg * npass / ms / 1e6, 1)}}), flush=True)


def bench_lanczos(args):
    """Plain FP64 Lanczos at N=24 (reference 24-site cluster and the 3x4 torus), MODEL 20
    phi=54 deg, h=0, complex128: 300 fixed steps (timed), then the converged ground state
    and its residual |Hx - E0 x| / |x|."""
    import numpy as np  # noqa: F401
    import torch  # noqa: F401

    from pned_amd import gpu, lattice, model

    rows = {}
    for spec in (24, "3x4"):
        cl = lattice.get_cluster(spec)
        op = model.build_operator(cl, 20, 54.0)
        ctx = gpu.Context(op)
        v1, v0, x = ctx.vector(), ctx.vector(), ctx.vector()
        ctx.set_initial_vec(v1, 1)
        tune = ctx.autotune(v1, v0, reps=2)
        ctx.set_initial_vec(v1, 20261202)
        ctx.sync()
        t = time.perf_counter()
        itr, alpha, beta, ene = ctx.lnc1(v1, v0, fixed_steps=300, nvec=1)
        ctx.sync()
        t300 = time.perf_counter() - t
        itc, e0, res, tgs = _ground_state(ctx, gpu, v1, v0, x, 20261202, 2000)
        rows[str(spec)] = {"n_sites": cl.n, "dim": ctx.dim, "tune": tune, "ms_per_step_300": round(t300 / 300 * 1e3, 4),
                           "E0_per_site_300": float(ene[0]) / cl.n, "converged_steps": itc,
                           "E0_per_site": e0 / cl.n, "residual": res, "ground_state_s": round(tgs, 3)}
        _progress(f"Lanczos {spec}: {rows[str(spec)]}")
        for v in (v1, v0, x):
            v.free()
        ctx.close()
    out = {"metric": "plain FP64 Lanczos ms per step (N=24, complex128)", "value": rows["24"]["ms_per_step_300"],
           "unit": "ms/step", "higher_is_better": False, "config": rows}
    print(json.dumps(out), flush=True)


def bench_lanczos32(args):
    """Plain FP64 Lanczos at the single-GPU limit: S=1/2 KH honeycomb 4x4 PBC (N=32),
    MODEL 20 phi=126 deg, [111] field h=0.1, complex128 (2^32 states, three resident
    64 GiB vectors), seeded Gaussian start, 200 timed steps; then the Ritz vector of the
    lowest Ritz value (a second 200-step run from the same start) and its residual."""
    import numpy as np  # noqa: F401
    import torch  # noqa: F401

    from pned_amd import gpu, lattice, model

    cl = lattice.torus(4, 4)
    op = model.build_operator(cl, 20, 126.0, field=0.1)
    ctx = gpu.Context(op)
    v1, v0, x = ctx.vector(), ctx.vector(), ctx.vector()
    seed, steps = 20261202, 200
    ctx.set_initial_vec(v1, seed)
    ctx.sync()
    _progress(f"N=32 Lanczos: {steps} steps")
    t = time.perf_counter()
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, fixed_steps=steps, nvec=1)
    ctx.sync()
    dt = time.perf_counter() - t
    ms = dt / steps * 1e3
    _progress(f"N=32 Lanczos: {ms:.2f} ms per step, lowest Ritz value / N {ene[0] / cl.n:.13f}")
    coef = gpu.vec12(alpha, beta, [ene[0]], 1)[0]
    ctx.set_initial_vec(v1, seed)
    t = time.perf_counter()
    ctx.ritz_vector(v1, v0, x, alpha, beta, coef)
    ctx.sync()
    tr = time.perf_counter() - t
    res = ctx.residual(x, ene[0], v0)
    _progress(f"N=32 Ritz vector {tr:.1f} s, residual {res:.3e}")
    npass, _ = ctx.plan_info()
    for v in (v1, v0, x):
        v.free()
    ctx.close()
    alg = 3 * 16 * 2.0 ** 32
    out = {"metric": "plain FP64 Lanczos ms per step (N=32, complex128)", "value": round(ms, 3), "unit": "ms/step",
           "higher_is_better": False, "steps": steps, "dtype": "c128",
           "config": {"workload": "S=1/2 KH honeycomb 4x4 PBC N=32, phi=0.7pi, [111] h=0.1, complex128 dim 2^32, "
                                  "seeded Gaussian start, 200 Lanczos steps",
                      "passes": npass, "hv_equivalent_GBps": round(alg / (ms * 1e-3) / 1e9, 1),
                      "E0_per_site_200": float(ene[0]) / cl.n, "ritz_s": round(tr, 2), "residual": res}}
    print(json.dumps(out), flush=True)


def bench_spin1(args):
    """Spin-1 N=18 reference cluster (3^18 states), MODEL 20 phi=108 deg, complex128:
    ground state to 1e-10 (Lanczos + Ritz vector), S^z(q) DSF at the cluster's 8 momenta
    plus (pi, pi/sqrt 3) (300 continued-fraction steps, eta=0.05, 2000 omegas on [0,5)),
    and a COCG solve of (H - E0 - omega - i eta) x = S^z(q)|0> at omega = 1 checked against
    the continued fraction."""
    import numpy as np
    import torch  # noqa: F401

    from pned_amd import gpu, lattice, model

    cl = lattice.reference_cluster(18)
    op = model.build_operator(cl, 20, 108.0, q=3)
    ctx = gpu.Context(op)
    v1, v0, x, w = (ctx.vector() for _ in range(4))
    ctx.set_initial_vec(v1, 1)
    ctx.sync()
    t = time.perf_counter()
    nhv = 20
    for _ in range(nhv):
        ctx.p_mltply_v3(v1, v0)
    ctx.sync()
    hv_ms = (time.perf_counter() - t) / nhv * 1e3
    _progress(f"spin-1 N=18 H.v {hv_ms:.3f} ms ({3 * 16 * ctx.dim / (hv_ms * 1e-3) / 1e9:.0f} GB/s algorithmic)")
    itc, e0, res, tgs = _ground_state(ctx, gpu, v1, v0, x, 20261202, 3000)
    _progress(f"ground state: E0/N {e0 / cl.n:.12f}, {itc} steps, residual {res:.2e}, {tgs:.2f} s")
    qs = [np.asarray(q, float) for q in cl.q_list] + [np.array([np.pi, np.pi / np.sqrt(3.0)])]
    eta, nomega, split, steps = 0.05, 2000, 2.5e-3, 300
    t = time.perf_counter()
    dsf_rows = []
    cf_at = {}
    for k, q in enumerate(qs):
        cn = ctx.intalv_zz(x, w, q, cl.site_vec)
        sa, sb = ctx.dsflnc1z(w, v0, cn, steps)
        ww, dsf, conv = ctx.renbun(sa, sb, e0, eta, nomega, split, steps, sq=cn)
        dsf_rows.append({"q": [round(float(q[0]), 6), round(float(q[1]), 6)], "S_q": cn,
                         "peak_omega": float(ww[int(np.argmax(dsf))]),
                         # conv[j]: the fraction depth at which omega_j converged (0: not within the steps)
                         "omega_converged": int(np.count_nonzero(conv)),
                         "mean_depth": round(float(conv[conv > 0].mean()), 2) if np.any(conv > 0) else None})
        cf_at[k] = (sa, sb, cn)
        _progress(f"DSF q={dsf_rows[-1]['q']}: S(q)={cn:.6f}, peak at omega={dsf_rows[-1]['peak_omega']:.4f}")
    tdsf = time.perf_counter() - t
    _progress(f"DSF at {len(qs)} momenta in {tdsf:.2f} s")
    # COCG at omega = 1 for the first nonzero momentum, against the continued fraction
    k = 1
    sa, sb, cn = cf_at[k]
    b, xs, r, p, y = w, ctx.vector(), ctx.vector(), ctx.vector(), ctx.vector()
    ctx.intalv_zz(x, b, qs[k], cl.site_vec)
    omega = 1.0
    t = time.perf_counter()
    try:
        it, cres = ctx.cocg(e0 + omega, -eta, b, xs, r, p, y, max_it=4000, tol=1e-10)
    except gpu.PnedError as err:   # not converged within max_it: report what it reached
        it, cres = 4000, float("nan")
        _progress(f"COCG: {err}")
    tcocg = time.perf_counter() - t
    g_cocg = -np.vdot(b.download(), xs.download())
    z = omega + e0 + 1j * eta
    g = 0.0
    for m in range(steps - 1, 0, -1):
        g = sb[m] ** 2 / (z - sa[m] - g)
    g_cf = 1.0 / (z - sa[0] - g)
    rel = abs(g_cocg - g_cf) / abs(g_cf)
    _progress(f"COCG at omega=1: {it} iterations in {tcocg:.2f} s, residual {cres:.1e}, |G_cocg - G_cf|/|G_cf| = {rel:.2e}")
    # the same fraction with more Lanczos steps: the difference above is its truncation
    long_steps = 1000
    ctx.intalv_zz(x, b, qs[k], cl.site_vec)
    sa2, sb2 = ctx.dsflnc1z(b, v0, cn, long_steps)
    g = 0.0
    for m in range(long_steps - 1, 0, -1):
        g = sb2[m] ** 2 / (z - sa2[m] - g)
    g_cf2 = 1.0 / (z - sa2[0] - g)
    rel2 = abs(g_cocg - g_cf2) / abs(g_cf2)
    _progress(f"{long_steps}-step continued fraction vs COCG: {rel2:.2e}")
    out = {"metric": "spin-1 N=18 pipeline (ground state + 9-momentum DSF + COCG check)", "value": round(tgs + tdsf, 3),
           "unit": "s", "higher_is_better": False,
           "config": {"dim": ctx.dim, "hv_ms": round(hv_ms, 4), "E0_per_site": e0 / cl.n, "gs_steps": itc,
                      "gs_residual": res, "gs_s": round(tgs, 3), "dsf_s": round(tdsf, 3), "dsf": dsf_rows,
                      "cocg": {"omega": omega, "eta": eta, "q": dsf_rows[k]["q"], "iterations": it, "residual": cres,
                               "s": round(tcocg, 3), "rel_diff_vs_continued_fraction": rel,
                               "rel_diff_vs_continued_fraction_1000_steps": rel2}}}
    print(json.dumps(out), flush=True)


def bench_tps(args):
    """Batched mTPQ: N=24 reference cluster, MODEL 20 phi=90 deg (J=0, K1=+2 in the
    reference's H = sum K S^g S^g, i.e. antiferromagnetic Kitaev; FM is phi=270 deg),
    R=64 interleaved samples.  Fused complex64 storage (FP64 reductions and scalars) vs
    the plain FP64 GPU path (complex128, one sample at a time, the reference's step
    structure: update pass, H.v, two dots, observable pass).  Errors: E/N and T of every
    step and C(T) of the c64 run against a complex128 run of the same fused recursion on
    the same start vectors."""
    import numpy as np
    import torch  # noqa: F401

    from pned_amd import gpu, lattice, model

    cl = lattice.reference_cluster(24)
    op = model.build_operator(cl, 20, 90.0, tps=True)
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    R = 64
    steps = args.steps if args.steps != 10 else 2000
    L = op.shift_L

    tunes = {}

    def tune_for(nvec, dtype, name):
        # every path gets the tile plan the autotuner picks on vectors of its own shape
        tv1, tv0 = ctx.vector(nvec, dtype), ctx.vector(nvec, dtype)
        ctx.set_initial_vec(tv1, 1)
        tunes[name] = ctx.autotune(tv1, tv0, reps=1)
        tv1.free()
        tv0.free()
        _progress(f"autotuned tile plan for {name}: {tunes[name]}")

    def run_fused(dtype, nsteps):
        tune_for(R, dtype, "fused_c64" if dtype == gpu.C64 else "fused_c128")
        v1, va, vb = (ctx.vector(R, dtype) for _ in range(3))
        ctx.set_initial_vec(v1, 20261202)
        ctx.sync()
        t = time.perf_counter()
        a, b, c = ctx.tps_fused(v1, va, vb, L, nsteps)
        ctx.sync()
        dt = time.perf_counter() - t
        for v in (v1, va, vb):
            v.free()
        return dt, a, b, c

    t64, a64, b64, c64 = run_fused(gpu.C64, steps)
    _progress(f"fused complex64 mTPQ: {steps} steps x {R} samples in {t64:.2f} s")
    t128, a128, b128, c128 = run_fused(gpu.C128, steps)
    _progress(f"fused complex128 mTPQ: {steps} steps x {R} samples in {t128:.2f} s")
    # plain path: complex128, one sample, reference step structure; timed on a bounded sample
    plain_steps = min(200, steps)
    tune_for(1, gpu.C128, "plain_c128")
    v1, v0, vg, vt = (ctx.vector(1, gpu.C128) for _ in range(4))
    ctx.set_initial_vec(v1, 20261202)
    ctx.sync()
    t = time.perf_counter()
    ctx.tps(v1, v0, vg, vt, L, plain_steps)
    ctx.sync()
    tplain = (time.perf_counter() - t) * (steps / plain_steps) * R
    _progress(f"plain complex128 path: {plain_steps} steps of one sample, extrapolated {tplain:.2f} s")
    n = cl.n
    E64, E128 = a64 / n, a128 / n
    T64 = (L - a64) / (2.0 * np.arange(1, steps + 1))[:, None]
    T128 = (L - a128) / (2.0 * np.arange(1, steps + 1))[:, None]

    # C(T) as the reference evaluates it (canonical TPQ over the 64 samples), on the
    # temperatures the k-step series resolves.  (A per-sample difference quotient dE/dT of
    # consecutive steps divides the storage error of E, ~4e-8 per site, by dT ~ T/k and so
    # reaches ~0.03 at 2000 steps: that is differentiation of rounding, not a kernel error.)
    ncoef = c64.shape[-1]
    A64, B64, Cf64 = gpu.tps_arrays(a64, b64, c64, R, ncoef)
    A128, B128, Cf128 = gpu.tps_arrays(a128, b128, c128, R, ncoef)
    t_lo = 2.0 * L / steps
    cq64 = gpu.ctpq(A64, B64, Cf64, t_lo, 20.0 * t_lo, 20)
    cq128 = gpu.ctpq(A128, B128, Cf128, t_lo, 20.0 * t_lo, 20)
    out = {
        "metric": "batched mTPQ speedup (fused complex64, R=64) over the plain FP64 GPU path",
        "value": round(tplain / t64, 3), "unit": "x", "higher_is_better": True,
        "config": {"workload": "mTPQ N=24 ref cluster, phi=90deg (AFM Kitaev K1=+2), R=64, L=ibond*l_b",
                   "steps": steps, "tune": tunes, "c64_fused_s": round(t64, 3), "c128_fused_s": round(t128, 3),
                   "plain_c128_s_extrapolated": round(tplain, 3),
                   "plain_sample": f"1 sample x {plain_steps} steps, scaled by {R} samples and {steps}/{plain_steps}",
                   "max_abs_err_E_per_site": float(np.max(np.abs(E64 - E128))),
                   "max_rel_err_T": float(np.max(np.abs(T64 - T128) / np.abs(T128))),
                   "ctpq_T_range": [round(t_lo, 5), round(20.0 * t_lo, 5)],
                   "ctpq_max_abs_err_C": float(np.max(np.abs(cq64[:, 1] - cq128[:, 1]))),
                   "ctpq_max_C": float(np.max(np.abs(cq128[:, 1]))),
                   "speedup_fused_c128_vs_plain": round(tplain / t128, 3)},
    }
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Mixed-size ground-state serving: continuous batching vs sequential plain solves.

    Capacity of each mode: all requests available at t = 0 (closed loop).  Latency: both
    modes on the same Poisson arrivals at 80 % of the sequential capacity, a load both can
    sustain."""
    import torch  # noqa: F401

    from pned_amd import serving

    specs = ["2x4", "3x3", "2x5", "3x4"] + (["2x7"] if args.serve_big else [])
    nreq = args.requests
    warm = serving.make_requests(8, specs, seed=1)
    serving.Server(tol=1e-10).run(warm)
    serving.Server(tol=1e-10, sequential=True).run(warm)

    def capacity(sequential):
        reqs = serving.make_requests(nreq, specs, seed=20261202)
        srv = serving.Server(tol=1e-10, sequential=sequential)
        t = time.perf_counter()
        res = srv.run(reqs)
        return len(res) / (time.perf_counter() - t)

    cap_b = capacity(False)
    _progress(f"continuous batching closed-loop capacity {cap_b:.2f} solves/s over {nreq} requests")
    cap_s = capacity(True)
    _progress(f"sequential closed-loop capacity {cap_s:.2f} solves/s over {nreq} requests")
    rate = 0.8 * cap_s

    def open_loop(sequential):
        reqs = serving.make_requests(nreq, specs, seed=20261203, rate=rate)
        srv = serving.Server(tol=1e-10, sequential=sequential)
        t = time.perf_counter()
        res = srv.run(reqs)
        return serving.summarize(res, time.perf_counter() - t)

    batched = open_loop(False)
    _progress(f"continuous batching, Poisson {rate:.2f}/s: {batched}")
    plain = open_loop(True)
    _progress(f"sequential, Poisson {rate:.2f}/s: {plain}")
    out = {"metric": "mixed-size serving capacity solves/s (continuous batching) vs sequential plain solves",
           "value": round(cap_b, 3), "unit": "solves/s", "higher_is_better": True,
           "config": {"sizes": specs, "requests": nreq, "tol": 1e-10,
                      "capacity_batched": round(cap_b, 3), "capacity_sequential": round(cap_s, 3),
                      "capacity_ratio": round(cap_b / cap_s, 3), "arrival_rate": round(rate, 3),
                      "batched": batched, "sequential": plain}}
    print(json.dumps(out), flush=True)


GRAPH_TPS_CASES = ((2, 2, 8, "c128"), (2, 4, 8, "c128"), (3, 3, 8, "c128"), (2, 4, 64, "c64"))
GRAPH_LNC1_CASES = ("N=16", "N=20", "N=24")


def graphs_block(tps_cases=GRAPH_TPS_CASES, lnc1_cases=GRAPH_LNC1_CASES):
    """Captured HIP graphs against launch-by-launch steps, where launches and host round
    trips weigh most: fused mTPQ (pned_tps_fused vs pned_tps_fused_graph, 30-step graphs) and
    fixed-step lnc1 (pned_ctx_graphs on/off, 5-step graphs), ms per step over 301 / 300 steps,
    the better of two timed runs each after a warm-up run; both forms must give the same
    alpha/beta (and mTPQ rows) bit for bit."""
    import numpy as np
    import torch  # noqa: F401

    from pned_amd import gpu, lattice, model

    def timed(run, check):
        best, ref = None, None
        for rep in range(3):
            t0 = time.perf_counter()
            got = run()
            dt = time.perf_counter() - t0
            if ref is None:
                ref = got
            else:
                check(got, ref)
                best = dt if best is None else min(best, dt)
        return best, ref

    def same(a, b):
        for x, y in zip(a, b):
            if not np.array_equal(np.asarray(x), np.asarray(y)):
                raise SystemExit("graph and launch-by-launch runs differ")

    out = {"tps_fused": {}, "lnc1": {}}
    for a, b, R, dname in tps_cases:
        dt = gpu.C128 if dname == "c128" else gpu.C64
        cl = lattice.torus(a, b)
        op = model.build_operator(cl, 20, 90.0, tps=True)
        ctx = gpu.Context(op)
        ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
        v1, va, vb = (ctx.vector(R, dt) for _ in range(3))
        steps = 301
        res = {}
        refs = []
        for g in (None, 30):
            def run(g=g):
                ctx.set_initial_vec(v1, 5)
                ctx.sync()
                return ctx.tps_fused(v1, va, vb, op.shift_L, steps, graph_steps=g)
            t, r = timed(run, same)
            refs.append(r)
            res["graph" if g else "launch"] = round(t / steps * 1e3, 4)
        same(refs[0], refs[1])
        res["speedup"] = round(res["launch"] / res["graph"], 3)
        key = f"N={cl.n} R={R} {dname}"
        out["tps_fused"][key] = res
        _progress(f"fused mTPQ {key}: {res}")
        for v in (v1, va, vb):
            v.free()
        ctx.close()
    clusters = {"N=16": lambda: lattice.reference_cluster(16), "N=20": lambda: lattice.torus(2, 5),
                "N=24": lambda: lattice.reference_cluster(24)}
    for name in lnc1_cases:
        cl = clusters[name]()
        op = model.build_operator(cl, 20, 54.0)
        ctx = gpu.Context(op)
        v1, v0 = ctx.vector(), ctx.vector()
        steps = 300
        res = {}
        refs = []
        for on in (False, True):
            def run(on=on):
                ctx.graphs(on)
                ctx.set_initial_vec(v1, 5)
                ctx.sync()
                r = ctx.lnc1(v1, v0, fixed_steps=steps, nvec=1)
                return r[1], r[2]
            t, r = timed(run, same)
            refs.append(r)
            res["graph" if on else "launch"] = round(t / steps * 1e3, 4)
        same(refs[0], refs[1])
        res["speedup"] = round(res["launch"] / res["graph"], 3)
        out["lnc1"][name] = res
        _progress(f"lnc1 {name}: {res}")
        v1.free()
        v0.free()
        ctx.close()
    out["note"] = "ms per step; both forms give the same outputs bit for bit (checked)"
    return out


def bench_graphs(args):
    print(json.dumps({"metric": "ms per Krylov step, captured HIP graphs vs launch-by-launch",
                      "unit": "ms/step", "higher_is_better": False, "config": graphs_block()}), flush=True)


if __name__ == "__main__":
    main()
