#!/usr/bin/env python3
"""Benchmark: SSA velocity solves per second on one MI355X.

Workload (one "step"): the 64-member ensemble on 161 x 161 grids (dx = 8 km,
51,842 unknowns each; every member has its own seeded friction field and a
rate factor drawn log-uniformly in [1e-17, 1e-16]), each member solved from
rest through the full Picard loop (BiCGSTAB + ILU(0) at rtol 1e-8, Picard
tolerance 1e-4, at most 50 iterations) in one batched call of the
mixed-precision path (FP32 stencil SpMV and ILU(0) inside FP64 iterative
refinement, several workgroups per member; --factor bf16 stores the ILU(0)
factors in BF16 with the U rows scaled by their FP32 inverse diagonals).
Inputs are resident in HBM before the timed region; each step restores the
initial state on the device and solves all members again, on the library's
default path (Picard iterations replayed as captured hipGraphs).  The Krylov
kernel's statistics (HIP-event time of every launch, iteration counts) come
from the same solves run eagerly after the timed steps.

The line also reports the plain FP64 GPU path on the same ensemble (batched,
and one member at a time), the per-member relative L2 and max velocity error
of the mixed path against the FP64 path (with the linear solves converged to
rtol 1e-12, and at the reference's rtol 1e-8), the Krylov kernel's HBM
roofline, and the CPU oracle timed on this host's cores on a bounded sample
of the same workload.
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

HBM_PEAK_GBS = 8000.0


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def _finite(v):
    """NaN (a part that was not run) becomes null, so the line stays JSON."""
    if isinstance(v, dict):
        return {k: _finite(x) for k, x in v.items()}
    if isinstance(v, float) and v != v:
        return None
    return v


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fp64_paths(args, base, par_fp64, h, vel):
    """The FP64 GPU path on the ensemble: the converged error reference, the
    batched run at the reference's solver settings and the single-member
    rate."""
    from yelmo_amd.velocity import ResidentBatch

    # ---- error reference: the FP64 path with the linear solves converged to
    # rtol 1e-12, i.e. the Picard solution without the 1e-8 solver tolerance ----
    par_conv = par_fp64.with_(ssa_lis_opt=par_fp64.ssa_lis_opt.replace("1.0e-8", "1.0e-12"))
    probs_c = [p.astype(np.float64) for p in base]
    rbc = ResidentBatch(probs_c, par_conv, handle=h)
    rbc.run()
    rbc.download()
    rbc.close()
    ref = [vel(p) for p in probs_c]
    del probs_c

    # ---- plain FP64 GPU path (reference solver settings), batched ---------------
    probs64 = [p.astype(np.float64) for p in base]
    rb64 = ResidentBatch(probs64, par_fp64, handle=h)
    rb64.run()                      # warm-up
    _sync()
    t0 = time.perf_counter()
    rb64.run()
    _sync()
    t_fp64_batch = time.perf_counter() - t0
    info64 = rb64.download()
    plain = [vel(p) for p in probs64]
    rb64.close()
    del probs64
    err64 = [rel_l2(plain[k], ref[k]) for k in range(args.members)]
    print(f"# fp64 batched: {t_fp64_batch:.3f}s", file=sys.stderr, flush=True)

    # ---- plain FP64 GPU path, one member at a time ------------------------------
    ns = min(args.single_sample, args.members)
    singles = []
    for k in range(ns):
        rb = ResidentBatch([base[k].astype(np.float64)], par_fp64, handle=h)
        rb.run()
        _sync()
        t0 = time.perf_counter()
        rb.run()
        _sync()
        singles.append(time.perf_counter() - t0)
        rb.close()
    t_fp64_single = float(np.mean(singles))
    print(f"# fp64 single: {t_fp64_single:.3f}s per solve", file=sys.stderr, flush=True)

    return ref, plain, err64, t_fp64_batch, t_fp64_single


def run_diva(args):
    """48 DIVA domains of uneven size in one resident batch: mixed precision
    timed, the FP64 path with converged linear solves as the error reference."""
    from yelmo_amd import _lib, generator
    from yelmo_amd.params import SsaParams
    from yelmo_amd.velocity import ResidentBatch
    t_gen = time.perf_counter()
    base = generator.workload_diva_uneven(48, seed=0)
    sides = [p.nx for p in base]
    print(f"# generated 48 DIVA domains in {time.perf_counter() - t_gen:.1f}s", file=sys.stderr,
          flush=True)
    h = _lib.default_handle()

    def vel(p):
        return np.concatenate([p.state["ux_bar"].ravel(), p.state["uy_bar"].ravel()]).astype(np.float64)

    par64 = SsaParams(precision="fp64")
    par_conv = par64.with_(ssa_lis_opt=par64.ssa_lis_opt.replace("1.0e-8", "1.0e-12"))
    ref_probs = [p.astype(np.float64) for p in base]
    rb = ResidentBatch(ref_probs, par_conv, diva=True, handle=h)
    rb.run()
    rb.download()
    rb.close()
    ref = [vel(p) for p in ref_probs]
    del ref_probs

    probs64 = [p.astype(np.float64) for p in base]
    rb = ResidentBatch(probs64, par64, diva=True, handle=h)
    rb.run()
    _sync()
    t0 = time.perf_counter()
    rb.run()
    _sync()
    t_fp64 = time.perf_counter() - t0
    rb.close()

    par = SsaParams(precision="mixed")
    par = par.with_(ssa_lis_opt=par.ssa_lis_opt + f" -ygpu_factor {args.factor}")
    sched = {}
    if not args.no_tune:
        from yelmo_amd.tune import autotune_sizes
        t_tune = time.perf_counter()
        sched = autotune_sizes(base, par, diva=True, handle=h, progress=sys.stderr)
        print(f"# autotuned {len(sched)} sizes in {time.perf_counter() - t_tune:.1f}s: " +
              ", ".join(f"{k[0]}:{v.ilu_blocks}x{v.block}" for k, v in sorted(sched.items())),
              file=sys.stderr, flush=True)
    probs = [p.copy() for p in base]
    rb = ResidentBatch(probs, par, diva=True, handle=h)
    for _ in range(args.warmup):
        rb.run()
    # timed steps on the default (graphed) path; kernel statistics afterwards
    # from the same solves run eagerly
    steps = []
    _sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        rb.run()
        steps.append(time.perf_counter() - ts)
    _sync()
    elapsed = time.perf_counter() - t0
    info = rb.download()
    ns = max(1, min(2, args.steps))
    h.stats_enable(True)
    for _ in range(ns):
        rb.run()
    _sync()
    st = h.stats()
    h.stats_enable(False)
    rb.close()
    errs = [rel_l2(vel(p), ref[k]) for k, p in enumerate(probs)]
    unknowns = sum(2 * p.nx * p.ny for p in base)
    kname = _krylov_kernel_name(st)
    launches = max(st["krylov_launches"], 1)
    avg_ms = st["krylov_ms"] / launches
    achieved = (st["krylov_bytes"] / launches) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    line = {
        "metric": "DIVA solves/s, 48 uneven domains, at <=1e-6 rel vel err vs FP64",
        "value": 48 * args.steps / elapsed, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"DIVA, 48 domains, sides {sorted(set(sides))} spanning 1536 km, "
                               f"nz 21, A(zeta) = 1e-16 exp(-3(1-zeta)), mixed FP32/FP64, "
                               f"ILU factors {args.factor}",
                   "solver": par.ssa_lis_opt,
                   "unknowns_total": unknowns,
                   "schedules": {f"{k[0]}x{k[1]}": [v.ilu_blocks, v.block]
                                 for k, v in sorted(sched.items())},
                   "workgroups_per_member": st["krylov_group"]},
        "max_rel_l2_vs_fp64": max(errs), "error_ok": max(errs) <= 1e-6,
        "fp64_batched_solves_per_s": 48 / t_fp64,
        "speedup_vs_fp64_batched": (48 * args.steps / elapsed) / (48 / t_fp64),
        "step_latency_ms": {"p50": 1e3 * float(np.percentile(steps, 50)),
                            "p99": 1e3 * float(np.percentile(steps, 99))},
        "picard_iters_mean": float(np.mean([i.ssa_iter_now for i in info])),
        "krylov_iters_per_solve": st["krylov_iters"] / (48 * ns),
        "unknown_iters_per_s": st["unknown_iters"] / ns * args.steps / elapsed,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "avg_launch_ms": avg_ms},
        "cpu_baseline": None,
    }
    print(json.dumps(_finite(line)), flush=True)


def _krylov_kernel_name(st):
    """The Krylov kernel the statistics' last launch ran: the member-group
    kernel (several workgroups per member) or the one-workgroup kernel."""
    if st["krylov_group"] > 1:
        return f"yg_krylov_group_kernel<{st['krylov_block']}>"
    return f"yg_krylov_kernel<1, {st['krylov_block']}>"


def _serve_run(srv, seqs, arrivals, record):
    """Drives a server over a stream of requests.  Request k (the solve
    sequence seqs[k]) arrives at arrivals[k] seconds after the start; each
    solve after the first is submitted when its predecessor finishes.
    Returns per-request latency (arrival -> last solve done), per-solve
    latency (submit -> done), the makespan and the peak number of solves
    running at once.  record(k, j, prob) sees every finished solve."""
    n = len(seqs)
    order = np.argsort(arrivals, kind="stable")
    owner, t_submit, req_lat, solve_lat, t_done = {}, {}, np.zeros(n), [], []
    nxt, done, peak, steps = 0, 0, 0, 0
    t0 = time.perf_counter()
    t_note = t0
    while done < n:
        now = time.perf_counter() - t0
        if time.perf_counter() - t_note > 20.0:
            t_note = time.perf_counter()
            print(f"#   {now:.0f}s: {done}/{n} requests done, {len(owner)} in flight",
                  file=sys.stderr, flush=True)
        while nxt < n and arrivals[order[nxt]] <= now:
            k = int(order[nxt])
            t = srv.submit(seqs[k][0], k, last=len(seqs[k]) == 1)
            owner[t] = (k, 0)
            t_submit[t] = time.perf_counter() - t0
            nxt += 1
        if not owner:
            time.sleep(max(0.0, min(arrivals[order[nxt]] - now, 0.01)))
            continue
        nr, _ = srv.step()
        steps += 1
        peak = max(peak, nr)
        for t in srv.poll():
            k, j = owner.pop(t)
            tn = time.perf_counter() - t0
            solve_lat.append(tn - t_submit.pop(t))
            t_done.append(tn)
            info = srv.fetch(t, seqs[k][j])
            record(k, j, seqs[k][j], info)
            if j + 1 < len(seqs[k]):
                t2 = srv.submit(seqs[k][j + 1], k, cont=True, last=j + 2 == len(seqs[k]))
                owner[t2] = (k, j + 1)
                t_submit[t2] = time.perf_counter() - t0
            else:
                req_lat[k] = tn - arrivals[k]
                done += 1
    t90 = float(np.percentile(t_done, 90)) if t_done else 0.0
    return req_lat, np.array(solve_lat), time.perf_counter() - t0, peak, steps, t90


def run_serve(args):
    """Served stream of SSA solve sequences (continuous batching).  Requests
    of lognormal grid side (clipped to 81..641, 1536 km across) each ask for
    5 successive solves with seeded smooth +-10 m thickness increments
    between them, warm-started from the previous solve's device-resident
    state.  Per precision: a closed-loop run (every request queued at the
    start) gives the capacity in solves/s; then requests arrive Poisson at
    70% of that capacity and p50/p99 latency is measured.  The mixed-precision
    velocities of every solve are compared with the FP64 server's."""
    from yelmo_amd import generator
    from yelmo_amd.params import SsaParams
    from yelmo_amd.serve import Server
    nreq = args.requests
    sides = generator.serve_request_sides(nreq, seed=0)
    t_gen = time.perf_counter()
    seqs = [generator.served_sequence(int(s), 1000 + k, generator.SERVE_SOLVES)
            for k, s in enumerate(sides)]
    print(f"# generated {nreq} requests x {generator.SERVE_SOLVES} solves in "
          f"{time.perf_counter() - t_gen:.1f}s", file=sys.stderr, flush=True)
    nsolves = sum(len(s) for s in seqs)
    budget = int(args.budget_gb * (1 << 30))
    lanes = [tuple(int(v) for v in part.split(":")) for part in args.lanes.split(",")]
    vel = {}
    out = {}
    for prec in ("fp64", "mixed"):
        par = SsaParams(precision=prec)
        srv = Server(par, nz_max=seqs[0][0].nz, hbm_budget=budget, lanes=lanes)
        # warm-up: one small sequence through the same server
        warm = generator.served_sequence(81, 99, 2)
        _serve_run(srv, [warm], np.zeros(1), lambda k, j, p, info: None)
        errs, failed = [], []

        def record(k, j, p, info, prec=prec, errs=errs, failed=failed):
            v = np.concatenate([p.state["ux_b"].ravel(), p.state["uy_b"].ravel()]).astype(np.float64)
            if info.status != 0 or not np.isfinite(v).all():
                failed.append((k, j, info.status))
            if prec == "fp64":
                vel[(k, j)] = (v, info.ssa_iter_now)
            else:
                v64, it64 = vel[(k, j)]
                errs.append((rel_l2(v, v64), k, j, info.ssa_iter_now, it64))

        lat_c, slat_c, span_c, peak_c, steps_c, t90_c = _serve_run(srv, seqs, np.zeros(nreq),
                                                                   record)
        # capacity: the completion rate while the queue is full, up to 90 %
        # of all solves (the tail is one large sequence running alone)
        capacity = 0.9 * nsolves / t90_c
        rate = args.load * capacity / generator.SERVE_SOLVES
        rng = np.random.default_rng(50_000)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
        errs_closed = list(errs)
        errs.clear()
        lat, slat, span, peak, steps, _ = _serve_run(srv, seqs, arrivals, record)
        srv.close()
        out[prec] = {
            "capacity_solves_per_s": capacity,
            "closed_loop": {"makespan_s": span_c, "drain_solves_per_s": nsolves / span_c,
                            "t90_s": t90_c, "peak_running": peak_c, "steps": steps_c,
                            "request_latency_s": {"p50": float(np.percentile(lat_c, 50)),
                                                  "p99": float(np.percentile(lat_c, 99))}},
            "poisson": {"arrival_rate_req_per_s": rate, "solves_per_s": nsolves / span,
                        "makespan_s": span, "peak_running": peak, "steps": steps,
                        "request_latency_s": {"p50": float(np.percentile(lat, 50)),
                                              "p99": float(np.percentile(lat, 99))},
                        "solve_latency_s": {"p50": float(np.percentile(slat, 50)),
                                            "p99": float(np.percentile(slat, 99))}},
            "server_device_bytes": srv.device_bytes,
            "failed_solves": len(failed),
        }
        if prec == "mixed":
            allerr = errs_closed + errs
            e = np.array([a[0] for a in allerr], dtype=np.float64)
            out[prec]["max_rel_l2_vs_fp64"] = float(np.max(e)) if np.isfinite(e).all() else float("nan")
            same = [a[0] for a in allerr if a[3] == a[4]]
            # solves whose Picard loops stopped at the same iteration in both
            # precisions; a solve that stops one iteration apart differs by
            # about the Picard tolerance
            out[prec]["max_rel_l2_vs_fp64_same_picard"] = float(np.max(same)) if same else None
            out[prec]["solves_picard_count_differs"] = len(allerr) - len(same)
            w = max(allerr, key=lambda a: a[0])
            out[prec]["worst_solve"] = {"request": w[1], "solve": w[2], "side": int(sides[w[1]]),
                                        "picard_mixed": w[3], "picard_fp64": w[4],
                                        "rel_l2": w[0]}
        print(f"# {prec}: capacity {capacity:.1f} solves/s, p99 {out[prec]['poisson']['request_latency_s']['p99']:.2f}s",
              file=sys.stderr, flush=True)
    mixed = out["mixed"]
    line = {
        "metric": "served SSA solves/s and p50/p99 request latency, continuous batching",
        "value": mixed["poisson"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": nsolves, "warmup": 1,
        "ms_per_step": 1e3 * mixed["poisson"]["makespan_s"] / nsolves,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{nreq} requests x {generator.SERVE_SOLVES} warm-started SSA "
                               f"solves, side lognormal (median 161) in [81, 641], 1536 km "
                               f"across, Poisson arrivals at {args.load:.0%} of the measured "
                               f"capacity, mixed FP32/FP64",
                   "requests": nreq, "sides": [int(s) for s in sides], "lanes": args.lanes,
                   "hbm_budget_gb": args.budget_gb},
        "p50_latency_s": mixed["poisson"]["request_latency_s"]["p50"],
        "p99_latency_s": mixed["poisson"]["request_latency_s"]["p99"],
        # against the plain FP64 server at the reference's rtol 1e-8, whose
        # own distance from converged linear solves is 0.3-3e-6 on these
        # grids (the mixed path is within 3e-8 of converged)
        "max_rel_l2_vs_fp64": mixed["max_rel_l2_vs_fp64"],
        "max_rel_l2_vs_fp64_same_picard": mixed["max_rel_l2_vs_fp64_same_picard"],
        "failed_solves": mixed["failed_solves"] + out["fp64"]["failed_solves"],
        "mixed": mixed, "fp64": out["fp64"],
        "speedup_capacity_vs_fp64": mixed["capacity_solves_per_s"] / out["fp64"]["capacity_solves_per_s"],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(_finite(line)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--side", type=int, default=161)
    ap.add_argument("--cpu-sample", type=int, default=2,
                    help="ensemble members solved by the CPU oracle for the baseline")
    ap.add_argument("--single-sample", type=int, default=4,
                    help="members solved one at a time on the FP64 GPU path")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--factor", choices=("fp32", "bf16", "bf16u"), default="fp32",
                    help="storage of the mixed path's ILU(0) factors: bf16 (with the U rows "
                         "scaled by their FP32 inverse diagonals), bf16u (the U part in BF16, "
                         "the L part in FP32) or fp32")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the schedule autotuner (use the size rule: ceil(nx/64) line "
                         "blocks)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="time the mixed batched solve only (for counter runs): no FP64 "
                         "reference, no CPU baseline, error fields null")
    ap.add_argument("--workload", choices=("ensemble", "diva", "serve"), default="ensemble",
                    help="ensemble: the 64-member 161^2 SSA ensemble (the headline line); "
                         "diva: 48 DIVA domains of uneven size (96..384)^2, nz = 21; "
                         "serve: a served stream of warm-started solve sequences")
    ap.add_argument("--requests", type=int, default=64, help="serve: requests in the stream")
    ap.add_argument("--lanes", default="161:48,321:24,641:8",
                    help="serve: size-class lanes as max_side:slots, sides increasing")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM budget")
    ap.add_argument("--load", type=float, default=0.7,
                    help="serve: Poisson arrival rate as a fraction of the measured capacity")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, run a plain device copy of MIB MiB "
                         "five times (the reference kernel for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the ensemble members are independent and a "
              f"multi-GPU run would only split them (--gpus {args.gpus} not supported)",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibrate:
        import torch
        n = args.copy_calibrate * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibrate_bytes_read": 4 * n, "copy_calibrate_bytes_written": 4 * n}))
        return

    from yelmo_amd import _lib, generator
    from yelmo_amd.params import SsaParams
    from yelmo_amd.velocity import ResidentBatch

    if args.workload == "diva":
        return run_diva(args)
    if args.workload == "serve":
        return run_serve(args)

    t_gen = time.perf_counter()
    base = generator.workload_ensemble(args.members, args.side, seed=0, dtype=np.float32)
    print(f"# generated {args.members} x {args.side}^2 in {time.perf_counter() - t_gen:.1f}s",
          file=sys.stderr, flush=True)
    par_mixed = SsaParams(precision="mixed")
    par_mixed = par_mixed.with_(ssa_lis_opt=par_mixed.ssa_lis_opt +
                                f" -ygpu_factor {args.factor}")
    par_fp64 = SsaParams(precision="fp64")
    h = _lib.default_handle()

    def vel(p):
        return np.concatenate([p.state["ux_b"].ravel(), p.state["uy_b"].ravel()]).astype(np.float64)

    ref = plain = None
    err64 = [float("nan")]
    t_fp64_batch = t_fp64_single = float("nan")
    if not args.mixed_only:
        ref, plain, err64, t_fp64_batch, t_fp64_single = fp64_paths(args, base, par_fp64, h, vel)

    # ---- mixed precision, batched: the measured value ---------------------------
    sched = {}
    if not args.no_tune:
        from yelmo_amd.tune import autotune_sizes
        sched = autotune_sizes(base[:1], par_mixed, handle=h, progress=sys.stderr)
        print("# autotuned schedule " + ", ".join(f"{k[0]}:{v.ilu_blocks} blocks x {v.block} "
                                                   f"threads ({v.solve_ms:.2f} ms)"
                                                   for k, v in sched.items()),
              file=sys.stderr, flush=True)
    probs = [p.copy() for p in base]
    rb = ResidentBatch(probs, par_mixed, handle=h)
    for _ in range(args.warmup):
        rb.run()
    # the timed steps: the library's default path (Picard iterations replayed
    # as captured hipGraphs, one done-flag read per replay), no statistics
    step_times = []
    _sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        rb.run()
        step_times.append(time.perf_counter() - ts)
    _sync()
    elapsed = time.perf_counter() - t0
    info = rb.download()
    # kernel statistics, after the timed steps: the same solves run eagerly
    # with every Krylov launch timed by HIP events on its stream (the
    # roofline) and the iteration counts read back per launch
    ns = max(1, min(2, args.steps))
    h.stats_enable(True)
    _sync()
    te = time.perf_counter()
    for _ in range(ns):
        rb.run()
    _sync()
    eager_sps = args.members * ns / (time.perf_counter() - te)
    st = h.stats()
    h.stats_enable(False)
    # the stencil SpMV over all members in one full-device launch, on the
    # systems of the last Picard iteration: the SpMV roofline of the metric
    spmv = {}
    for name, fp32 in (("fp64", False), ("fp32", True)):
        _, sms, sbytes = rb.spmv(fp32=fp32, reps=50)
        gbs = sbytes / (sms * 1e-3) / 1e9 if sms > 0 else 0.0
        spmv[name] = {"kernel": f"k_spmv_records<{'float' if fp32 else 'double'}>",
                      "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": gbs / HBM_PEAK_GBS, "avg_launch_ms": sms,
                      "bytes_per_launch": sbytes}
    errs, errs_plain, maxabs = [float("nan")], [float("nan")], [float("nan")]
    if ref is not None:
        errs, errs_plain, maxabs = [], [], []
        for k, p in enumerate(probs):
            v = vel(p)
            errs.append(rel_l2(v, ref[k]))
            errs_plain.append(rel_l2(v, plain[k]))
            maxabs.append(float(np.max(np.abs(v - ref[k]))))
    rb.close()
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.members * args.steps / elapsed
    fp64_batch_sps = args.members / t_fp64_batch
    fp64_single_sps = 1.0 / t_fp64_single

    launches = max(st["krylov_launches"], 1)
    avg_ms = st["krylov_ms"] / launches
    bytes_per_launch = st["krylov_bytes"] / launches
    kname = _krylov_kernel_name(st)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    # ---- CPU baseline: the oracle (FP32 fields + FP64 BiCGSTAB/ILU(0), the
    # reference's arithmetic), one thread, on a bounded sample of members -------
    cpu = None
    if not args.no_cpu and not args.mixed_only and args.cpu_sample > 0:
        from oracle import oracle
        par_ref = SsaParams(precision="ref")
        t0 = time.perf_counter()
        n_cpu = min(args.cpu_sample, args.members)
        for k in range(n_cpu):
            oracle.calc_velocity_ssa(base[k], par_ref, dp=False)
        tc = time.perf_counter() - t0
        cpu = {"value": n_cpu / tc, "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"{n_cpu} members of the {args.members} x {args.side}^2 ensemble, "
                         f"full Picard solve each, C oracle (FP32 fields, FP64 BiCGSTAB+ILU(0)), "
                         f"1 thread, {tc:.1f}s"}

    line = {
        "metric": "SSA solves/s at <=1e-6 rel vel err vs plain FP64; SpMV HBM GB/s; p99 latency",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"ensemble {args.members} x {args.side}x{args.side} SSA (dx 8 km, "
                        f"{2 * args.side * args.side} unknowns each), mixed FP32/FP64, batched, "
                        f"ILU factors {args.factor}",
            "members": args.members, "side": args.side,
            "solver": par_mixed.ssa_lis_opt, "picard_tol": par_mixed.ssa_iter_conv,
            "schedule": {f"{k[0]}x{k[1]}": [v.ilu_blocks, v.block] for k, v in sched.items()},
            "workgroups_per_member": st["krylov_group"],
        },
        # per-member velocity error of the mixed path against the FP64 path
        # with converged linear solves (rtol 1e-12), and against the FP64
        # path at the reference's rtol 1e-8, whose own distance from the
        # converged solution is given beside it
        "max_rel_l2_vs_fp64": max(errs),
        "max_abs_vel_err_m_per_yr": max(maxabs),
        "max_rel_l2_vs_fp64_rtol1e-8": max(errs_plain),
        "fp64_rtol1e-8_max_rel_l2_vs_converged": max(err64),
        "error_ok": max(errs) <= 1e-6,
        "fp64_batched_solves_per_s": fp64_batch_sps,
        "fp64_single_solves_per_s": fp64_single_sps,
        "speedup_vs_fp64_single": value / fp64_single_sps,
        "speedup_vs_fp64_batched": value / fp64_batch_sps,
        "eager_stats_solves_per_s": eager_sps,
        "step_latency_ms": {"p50": 1e3 * float(np.percentile(step_times, 50)),
                            "p99": 1e3 * float(np.percentile(step_times, 99))},
        "picard_iters_mean": float(np.mean([i.ssa_iter_now for i in info])),
        "krylov_iters_per_solve": st["krylov_iters"] / max(args.members * ns, 1),
        "unknown_iters_per_s": st["unknown_iters"] / ns * args.steps / elapsed,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "avg_launch_ms": avg_ms,
                     "bytes_per_launch": bytes_per_launch, "launches": launches},
        # y = A x of all 64 members in one launch (HBM-bound; the FP64 records
        # and vectors, 292 MB, exceed the 256 MB Infinity Cache, the FP32 ones
        # do not)
        "spmv": spmv,
        "cpu_baseline": cpu,
    }
    print(json.dumps(_finite(line)), flush=True)


if __name__ == "__main__":
    main()
