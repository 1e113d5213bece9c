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


def baseline_legs(args, base, h):
    """Fairer comparisons for the value: (1) the plain FP64 path with every
    CU busy (4 copies of the ensemble, one workgroup per member on 256 CUs;
    the 64-member FP64 batch leaves 192 CUs idle); (2) Yelmo's production
    defaults (&ydyn: energy assembly, CG + Jacobi at rtol 1e-2, Picard
    tolerance 1e-2), batched, FP64 and mixed."""
    from yelmo_amd.params import SsaParams
    from yelmo_amd.velocity import ResidentBatch
    out = {}
    par64 = SsaParams(precision="fp64")
    probs = [p.astype(np.float64) for p in base for _ in range(4)]
    rb = ResidentBatch(probs, par64, handle=h)
    rb.run()
    _sync()
    t0 = time.perf_counter()
    rb.run()
    _sync()
    t = time.perf_counter() - t0
    rb.close()
    out["fp64_full_device"] = {"members": len(probs), "solves_per_s": len(probs) / t,
                               "note": "4 copies of the ensemble, one workgroup per member"}
    print(f"# fp64 full device ({len(probs)} members): {len(probs) / t:.1f} solves/s",
          file=sys.stderr, flush=True)
    del probs
    # the mixed path on the same 256 members: the like-for-like ratio
    par_m = SsaParams(precision="mixed")
    probs = [p.copy() for p in base for _ in range(4)]
    rb = ResidentBatch(probs, par_m, handle=h)
    rb.run()
    _sync()
    t0 = time.perf_counter()
    rb.run()
    _sync()
    tm = time.perf_counter() - t0
    rb.close()
    out["mixed_full_device"] = {"members": len(probs), "solves_per_s": len(probs) / tm,
                                "speedup_vs_fp64_full_device": t / tm,
                                "note": "the same 256 members, mixed path (default schedule)"}
    print(f"# mixed full device ({len(probs)} members): {len(probs) / tm:.1f} solves/s",
          file=sys.stderr, flush=True)
    del probs
    prod = {}
    for prec in ("fp64", "mixed"):
        par = SsaParams.yelmo_defaults(precision=prec)
        ps = [p.astype(np.float64) if prec == "fp64" else p.copy() for p in base]
        rb = ResidentBatch(ps, par, handle=h)
        rb.run()
        _sync()
        t0 = time.perf_counter()
        rb.run()
        _sync()
        t = time.perf_counter() - t0
        info = rb.download()
        rb.close()
        prod[prec] = {"solves_per_s": len(ps) / t,
                      "picard_iters_mean": float(np.mean([i.ssa_iter_now for i in info])),
                      "krylov_iters_per_solve": float(np.mean([i.ssa_lin_iter for i in info])),
                      "failed": sum(1 for i in info if i.status != 0)}
        print(f"# production defaults (energy, CG + Jacobi 1e-2) {prec}: "
              f"{len(ps) / t:.1f} solves/s", file=sys.stderr, flush=True)
    out["production_defaults"] = {"solver": SsaParams.yelmo_defaults().ssa_lis_opt,
                                  "form": "energy", **prod}
    return out


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
    ref, t_fp64 = None, float("nan")
    if not args.mixed_only:
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
    errs = ([rel_l2(vel(p), ref[k]) for k, p in enumerate(probs)] if ref is not None
            else [float("nan")])
    unknowns = sum(2 * p.nx * p.ny for p in base)
    kname = _krylov_kernel_name(st)
    launches = max(st["krylov_launches"], 1)
    avg_ms = st["krylov_ms"] / launches
    achieved = (st["krylov_bytes"] / launches) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    workload = (f"DIVA, 48 domains, sides {sorted(set(sides))} spanning 1536 km, "
                f"nz 21, A(zeta) = 1e-16 exp(-3(1-zeta)), mixed FP32/FP64, "
                f"ILU factors {args.factor}")
    traffic, traffic_src = counted_traffic(kname, workload)
    line = {
        "metric": "DIVA solves/s, 48 uneven domains, at <=1e-6 rel vel err vs FP64",
        "value": 48 * args.steps / elapsed, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload,
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
                     "traffic": traffic,
                     "traffic_source": traffic_src["source"] if traffic_src else None,
                     "avg_launch_ms": avg_ms,
                     "bytes_per_launch": st["krylov_bytes"] / launches},
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
    latency (submit -> done), the makespan, the peak number of solves running
    at once, the server steps and the time by which 90 % of the solves were
    done.  record(k, j, prob, info) sees every finished solve (prob holds
    its results)."""
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
            # results land in a scratch copy: the request objects stay as
            # submitted (cold), so passes and repeated requests can share them
            q = seqs[k][j].copy()
            info = srv.fetch(t, q)
            record(k, j, q, info)
            del q
            if j + 1 < len(seqs[k]):
                t2 = srv.submit(seqs[k][j + 1], k, cont=True, last=j + 2 == len(seqs[k]))
                owner[t2] = (k, j + 1)
                t_submit[t2] = time.perf_counter() - t0
            else:
                req_lat[k] = tn - arrivals[k]
                done += 1
    t90 = float(np.percentile(t_done, 90)) if t_done else 0.0
    return req_lat, np.array(solve_lat), time.perf_counter() - t0, peak, steps, t90


def _vel(p):
    return np.concatenate([p.state["ux_b"].ravel(), p.state["uy_b"].ravel()]).astype(np.float64)


def serve_stream(seqs, par, lanes, budget, load=None, ref=None, key=None):
    """Serves the request sequences with continuous batching.  A closed-loop
    pass (every request queued at the start) gives the capacity in solves/s
    (completion rate up to 90 % of the solves); with ``load``, a second pass
    has requests arrive Poisson at that fraction of the capacity and gives
    the latencies.  Every pass submits the requests' own cold inputs.
    Returns (figures, results) with results[(k, j)] = (velocity, Picard
    iterations, status) of every solve; with ``ref`` (reference results of
    the distinct sequences, request k being sequence key[k]) each solve is
    compared as it finishes and results[(k, j)] = (rel L2 from the
    reference, Picard iterations, status, reference Picard iterations)."""
    from yelmo_amd.serve import Server
    nreq = len(seqs)
    nsolves = sum(len(s) for s in seqs)
    srv = Server(par, nz_max=seqs[0][0].nz, hbm_budget=budget, lanes=lanes)
    # warm-up: one small sequence through the same server
    from yelmo_amd import generator
    warm = generator.served_sequence(81, 99, 2)
    _serve_run(srv, [warm], np.zeros(1), lambda k, j, p, info: None)
    res_closed, res_poisson = {}, {}

    def recorder(store):
        def rec(k, j, p, info):
            if ref is None:
                store[(k, j)] = (_vel(p), info.ssa_iter_now, info.status)
            else:
                r = ref[(key[k], j)]
                v = _vel(p)
                store[(k, j)] = (rel_l2(v, r[0]) if np.isfinite(v).all() else np.inf,
                                 info.ssa_iter_now, info.status, r[1])
        return rec

    def failed(v):
        return v[2] != 0 or not (np.isfinite(v[0]).all() if ref is None else np.isfinite(v[0]))

    lat_c, slat_c, span_c, peak_c, steps_c, t90_c = _serve_run(srv, seqs, np.zeros(nreq),
                                                               recorder(res_closed))
    capacity = 0.9 * nsolves / t90_c
    out = {"capacity_solves_per_s": capacity,
           "closed_loop": {"makespan_s": span_c, "drain_solves_per_s": nsolves / span_c,
                           "t90_s": t90_c, "peak_running": peak_c, "steps": steps_c,
                           "request_latency_s": {"p50": float(np.percentile(lat_c, 50)),
                                                 "p99": float(np.percentile(lat_c, 99))}},
           "server_device_bytes": srv.device_bytes,
           "failed_solves": sum(1 for v in res_closed.values() if failed(v))}
    if load is not None:
        from yelmo_amd import generator as g
        rate = load * capacity / g.SERVE_SOLVES
        rng = np.random.default_rng(50_000)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
        lat, slat, span, peak, steps, _ = _serve_run(srv, seqs, arrivals, recorder(res_poisson))
        out["poisson"] = {"load": load, "arrival_rate_req_per_s": rate,
                          "solves_per_s": nsolves / span, "makespan_s": span,
                          "peak_running": peak, "steps": steps,
                          "request_latency_s": {"p50": float(np.percentile(lat, 50)),
                                                "p99": float(np.percentile(lat, 99))},
                          "solve_latency_s": {"p50": float(np.percentile(slat, 50)),
                                              "p99": float(np.percentile(slat, 99))}}
        # the same requests served under another arrival pattern: a solve
        # lands at another place of the device arena, which changes the
        # chunking of its dot products (rounding), nothing else
        if ref is None:
            out["max_rel_l2_poisson_vs_closed_loop"] = max(
                rel_l2(res_poisson[kj][0], res_closed[kj][0]) for kj in res_closed)
        else:
            out["poisson"]["vs_reference"] = serve_errors(res_poisson, None)
            out["poisson"]["failed_solves"] = sum(1 for v in res_poisson.values() if failed(v))
    srv.close()
    return out, res_closed


def serve_errors(res, ref):
    """Per-solve relative L2 distance of served velocities from the
    reference run's, the worst solve and its Picard counts (ref None: the
    results already hold the distances and reference counts)."""
    if ref is None:
        errs = [(e, kj, it, rit) for kj, (e, it, _, rit) in res.items()]
    else:
        errs = [(rel_l2(v, ref[kj][0]), kj, it, ref[kj][1]) for kj, (v, it, _) in res.items()]
    w = max(errs, key=lambda a: a[0])
    return {"max_rel_l2": w[0], "median_rel_l2": float(np.median([e[0] for e in errs])),
            "p99_rel_l2": float(np.percentile([e[0] for e in errs], 99)),
            "solves_above_1e-6": sum(1 for e in errs if e[0] > 1e-6), "solves": len(errs),
            "worst_solve": {"request": w[1][0], "solve": w[1][1], "picard": w[2],
                            "picard_reference": w[3], "rel_l2": w[0]}}


def run_serve(args):
    """Served stream of SSA solve sequences (continuous batching).  Requests
    of lognormal grid side (clipped to 81..641, 1536 km across) each ask for
    5 successive solves with seeded smooth +-10 m thickness increments
    between them, warm-started from the previous solve's device-resident
    state.  Per precision (mixed, plain FP64 at the reference's rtol 1e-8):
    the capacity from a closed-loop pass, then p50/p99 latency with Poisson
    arrivals at --load of that capacity; every pass submits the requests'
    cold inputs.  The error reference is the FP64 server with the linear
    solves converged to rtol 1e-12, run first over the distinct sequences.
    With --distinct D (fewer than --requests) the requests are drawn
    (seeded, uniformly) from a pool of D generated sequences, so that a
    stream of thousands of requests needs the host memory of D."""
    from yelmo_amd import generator
    from yelmo_amd.params import SsaParams
    nreq = args.requests
    ndist = args.distinct if 0 < args.distinct < nreq else nreq
    sides = generator.serve_request_sides(ndist, seed=0)
    t_gen = time.perf_counter()
    pool = [generator.served_sequence(int(s), 1000 + k, generator.SERVE_SOLVES)
            for k, s in enumerate(sides)]
    key = (np.arange(nreq) if ndist == nreq
           else np.random.default_rng(4000).integers(0, ndist, size=nreq))
    seqs = [pool[int(k)] for k in key]
    print(f"# generated {ndist} distinct sequences x {generator.SERVE_SOLVES} solves in "
          f"{time.perf_counter() - t_gen:.1f}s; {nreq} requests", file=sys.stderr, flush=True)
    nsolves = sum(len(s) for s in seqs)
    budget = int(args.budget_gb * (1 << 30))
    lanes = [tuple(int(v) for v in part.split(":")) for part in args.lanes.split(",")]
    ref = None
    if not args.serve_mixed_only:
        conv = SsaParams(precision="fp64")
        conv = conv.with_(ssa_lis_opt=conv.ssa_lis_opt.replace("1.0e-8", "1.0e-12"))
        t_ref = time.perf_counter()
        _, ref = serve_stream(pool, conv, lanes, budget, load=None)
        print(f"# converged FP64 reference of {ndist} sequences in "
              f"{time.perf_counter() - t_ref:.1f}s", file=sys.stderr, flush=True)
    out, res = {}, {}
    keep = f" -ygpu_keep_factors {args.keep_factors}"
    fp64_pool = args.serve_fp64 == "pool" and ndist < nreq
    precs = ("mixed",) if args.serve_mixed_only or args.serve_fp64 == "none" else ("mixed", "fp64")
    for prec in precs:
        # the plain FP64 server's Poisson pass (its 641^2 solves on the FP64
        # path take minutes) only on request; with --serve-fp64 pool its
        # closed-loop pass serves the distinct sequences once each
        load = args.load if prec == "mixed" or args.serve_fp64_poisson else None
        par = SsaParams(precision=prec)
        if prec == "mixed":
            par = par.with_(ssa_lis_opt=par.ssa_lis_opt + keep)
        if prec == "fp64" and fp64_pool:
            out[prec], res[prec] = serve_stream(pool, par, lanes, budget, load=load, ref=ref,
                                                key=np.arange(ndist))
            out[prec]["requests_served"] = f"the {ndist} distinct sequences once each"
        else:
            out[prec], res[prec] = serve_stream(seqs, par, lanes, budget, load=load, ref=ref,
                                                key=key)
        if ref is not None:
            out[prec]["vs_fp64_converged"] = serve_errors(res[prec], None)
        p99 = out[prec]["poisson"]["request_latency_s"]["p99"] if out[prec].get("poisson") else None
        print(f"# {prec}: capacity {out[prec]['capacity_solves_per_s']:.1f} solves/s, p99 "
              f"{'-' if p99 is None else f'{p99:.2f}s'}", file=sys.stderr, flush=True)
        # the pass's figures at once, should a later pass not finish
        print(f"# {prec} figures: {json.dumps(_finite(out[prec]))}", file=sys.stderr, flush=True)
    for extra in args.serve_extra:
        # further mixed-precision servers (closed loop) with extra options,
        # against the same reference
        par = SsaParams(precision="mixed")
        par = par.with_(ssa_lis_opt=par.ssa_lis_opt + keep + " " + extra)
        o, _ = serve_stream(seqs, par, lanes, budget, load=None, ref=ref, key=key)
        if ref is not None:
            o["vs_fp64_converged"] = serve_errors(_, None)
        out.setdefault("mixed_variants", {})[extra] = o
        print(f"# mixed {extra!r} figures: {json.dumps(_finite(o))}", file=sys.stderr, flush=True)
    if args.serve_mixed_only:
        mixed = out["mixed"]
        print(json.dumps(_finite({
            "metric": "served SSA solves/s and p50/p99 request latency, continuous batching "
                      "(mixed server only: no FP64 passes, no error reference)",
            "value": mixed["poisson"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
            "keep_factors": args.keep_factors, "requests": nreq, "distinct": ndist,
            "capacity_solves_per_s": mixed["capacity_solves_per_s"],
            "p50_latency_s": mixed["poisson"]["request_latency_s"]["p50"],
            "p99_latency_s": mixed["poisson"]["request_latency_s"]["p99"], "mixed": mixed})),
            flush=True)
        return
    mixed = out["mixed"]
    fp64 = out.get("fp64", {"capacity_solves_per_s": float("nan"), "failed_solves": 0})
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
                               f"capacity, mixed FP32/FP64" +
                               ("" if ndist == nreq else
                                f"; requests drawn from {ndist} distinct sequences"),
                   "requests": nreq, "distinct_sequences": ndist,
                   "sides": [int(s) for s in sides], "lanes": args.lanes,
                   "hbm_budget_gb": args.budget_gb},
        "p50_latency_s": mixed["poisson"]["request_latency_s"]["p50"],
        "p99_latency_s": mixed["poisson"]["request_latency_s"]["p99"],
        # against the FP64 server with linear solves converged to rtol 1e-12
        "max_rel_l2_vs_fp64": mixed["vs_fp64_converged"]["max_rel_l2"],
        "error_ok": mixed["vs_fp64_converged"]["max_rel_l2"] <= 1e-6,
        "failed_solves": mixed["failed_solves"] + fp64["failed_solves"],
        "mixed": mixed, "fp64": out.get("fp64"), "mixed_variants": out.get("mixed_variants"),
        "speedup_capacity_vs_fp64": mixed["capacity_solves_per_s"] / fp64["capacity_solves_per_s"],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(_finite(line)), flush=True)


def served_leg(args):
    """The bounded served stream of the default line: --serve-requests
    requests (sides lognormal in [81, 641]) of 5 warm-started solves,
    mixed precision, capacity and p50/p99 request latency at 70 % load,
    errors against the converged FP64 server."""
    from yelmo_amd import generator
    from yelmo_amd.params import SsaParams
    nreq = args.serve_requests
    sides = generator.serve_request_sides(nreq, seed=0)
    seqs = [generator.served_sequence(int(s), 1000 + k, generator.SERVE_SOLVES)
            for k, s in enumerate(sides)]
    budget = int(args.budget_gb * (1 << 30))
    lanes = [tuple(int(v) for v in part.split(":")) for part in args.lanes.split(",")]
    mixed, res = serve_stream(seqs, SsaParams(precision="mixed"), lanes, budget, load=args.load)
    conv = SsaParams(precision="fp64")
    conv = conv.with_(ssa_lis_opt=conv.ssa_lis_opt.replace("1.0e-8", "1.0e-12"))
    _, ref = serve_stream(seqs, conv, lanes, budget, load=None)
    e = serve_errors(res, ref)
    print(f"# served {nreq} requests: capacity {mixed['capacity_solves_per_s']:.1f} solves/s, "
          f"p99 {mixed['poisson']['request_latency_s']['p99']:.3f}s", file=sys.stderr, flush=True)
    return {"requests": nreq, "solves": sum(len(q) for q in seqs),
            "sides": [int(v) for v in sides],
            "capacity_solves_per_s": mixed["capacity_solves_per_s"],
            "solves_per_s_at_load": mixed["poisson"]["solves_per_s"], "load": args.load,
            "p50_request_latency_s": mixed["poisson"]["request_latency_s"]["p50"],
            "p99_request_latency_s": mixed["poisson"]["request_latency_s"]["p99"],
            "p50_solve_latency_s": mixed["poisson"]["solve_latency_s"]["p50"],
            "p99_solve_latency_s": mixed["poisson"]["solve_latency_s"]["p99"],
            "max_rel_l2_vs_fp64": e["max_rel_l2"], "worst_solve": e["worst_solve"],
            "max_rel_l2_poisson_vs_closed_loop": mixed["max_rel_l2_poisson_vs_closed_loop"],
            "failed_solves": mixed["failed_solves"]}


def config1_single(args, h):
    """One 161 x 161 instance (dx 8 km), the mixed path beside the plain
    FP64 path: solves/s of each (repeated solves of the resident domain),
    rel-L2 and max velocity error against the FP64 path with converged
    linear solves, Picard and Krylov counts, the Krylov kernel's HBM rate."""
    from yelmo_amd import generator
    from yelmo_amd.params import SsaParams
    from yelmo_amd.velocity import ResidentBatch
    p0 = generator.make_problem(161, 8e3, seed=0, dtype=np.float32)
    par_m = SsaParams(precision="mixed")
    par_m = par_m.with_(ssa_lis_opt=par_m.ssa_lis_opt + f" -ygpu_factor {args.factor}")
    par_64 = SsaParams(precision="fp64")
    conv = par_64.with_(ssa_lis_opt=par_64.ssa_lis_opt.replace("1.0e-8", "1.0e-12"))
    out = {"workload": "one 161 x 161 SSA instance (dx 8 km, 51842 unknowns)"}
    rb = ResidentBatch([p0.astype(np.float64)], conv, handle=h)
    rb.run()
    rb.download()
    ref = _vel(rb.problems[0])
    rb.close()
    for name, par, prob in (("mixed", par_m, p0.copy()), ("fp64", par_64, p0.astype(np.float64))):
        rb = ResidentBatch([prob], par, handle=h)
        rb.run()
        _sync()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            rb.run()
        _sync()
        dt = (time.perf_counter() - t0) / reps
        info = rb.download()[0]
        h.stats_enable(True)
        rb.run()
        st = h.stats()
        h.stats_enable(False)
        rb.close()
        v = _vel(prob)
        avg = st["krylov_ms"] / max(st["krylov_launches"], 1)
        out[name] = {"solves_per_s": 1.0 / dt, "rel_l2_vs_fp64_converged": rel_l2(v, ref),
                     "max_abs_vel_err_m_per_yr": float(np.max(np.abs(v - ref))),
                     "picard_iters": info.ssa_iter_now, "krylov_iters": st["krylov_iters"],
                     "krylov_kernel": _krylov_kernel_name(st),
                     "krylov_gb_per_s": (st["krylov_bytes"] / max(st["krylov_launches"], 1)) /
                                        (avg * 1e-3) / 1e9 if avg > 0 else None}
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    print(f"# single 161^2: mixed {out['mixed']['solves_per_s']:.2f}, fp64 "
          f"{out['fp64']['solves_per_s']:.2f} solves/s", file=sys.stderr, flush=True)
    return out


def cpu_baseline(args, base):
    """The oracle (FP32 fields + FP64 BiCGSTAB/ILU(0): the reference's
    arithmetic) on this host's cores: one thread on --cpu-sample members,
    then --cpu-threads threads on as many members at once (ctypes releases
    the GIL around each oracle call)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle
    from yelmo_amd.params import SsaParams
    par_ref = SsaParams(precision="ref")

    def solve(k):
        _, info = oracle.calc_velocity_ssa(base[k % len(base)], par_ref, dp=False)
        return 2.0 * base[k % len(base)].nx * base[k % len(base)].ny * info["ssa_lin_iter"]

    n1 = min(args.cpu_sample, len(base))
    t0 = time.perf_counter()
    ui1 = sum(solve(k) for k in range(n1))
    t1 = time.perf_counter() - t0
    nt = max(1, min(args.cpu_threads, os.cpu_count() or 1))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(nt) as ex:
        uis = list(ex.map(solve, range(nt)))
    tn = time.perf_counter() - t0
    return {"value": nt / tn, "unit": "solves/s", "cores": nt, "kind": "port",
            "unknown_iters_per_s": sum(uis) / tn,
            "single_thread": {"solves_per_s": n1 / t1, "unknown_iters_per_s": ui1 / t1,
                              "members": n1},
            "host_cpus": os.cpu_count(),
            "sample": f"{nt} members of the {args.members} x {args.side}^2 ensemble solved at "
                      f"once on {nt} threads ({tn:.1f}s), and {n1} members on 1 thread "
                      f"({t1:.1f}s); full Picard solve each, C oracle (FP32 fields, FP64 "
                      f"BiCGSTAB+ILU(0))"}


def counted_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` counted by the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded in profiles/counted_traffic.json
    (scaled by the plain-copy calibration stated there), or None when that
    file holds no count for this kernel on this workload."""
    path = os.path.join(ROOT, "profiles", "counted_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        for e in d["counts"]:
            if e.get("kernel") == kernel and e.get("workload") == workload:
                return e["bytes_per_launch"], e
        return None, None
    except (OSError, KeyError, ValueError, TypeError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--side", type=int, default=161)
    ap.add_argument("--cpu-sample", type=int, default=2,
                    help="ensemble members solved by the CPU oracle on one thread")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads (and members) of the all-core CPU baseline (capped at the "
                         "host's CPUs)")
    ap.add_argument("--serve-requests", type=int, default=16,
                    help="requests of the bounded served stream in the default line (0: none)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-instance 161^2 leg of the default line")
    ap.add_argument("--single-sample", type=int, default=4,
                    help="members solved one at a time on the FP64 GPU path")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-baselines", action="store_true",
                    help="skip the full-device FP64 and production-default legs")
    ap.add_argument("--factor", choices=("fp32", "bf16", "bf16u"), default=None,
                    help="storage of the mixed path's ILU(0) factors: bf16 (with the U rows "
                         "scaled by their FP32 inverse diagonals), bf16u (the U part in BF16, "
                         "the L part in FP32) or fp32; default bf16 for the DIVA batch (its "
                         "workload names BF16 factors), fp32 otherwise")
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
    ap.add_argument("--serve-extra", action="append", default=[],
                    help="serve: an extra closed-loop pass of the mixed server with these "
                         "options appended (repeatable), against the same reference")
    ap.add_argument("--serve-fp64", choices=("full", "pool", "none"), default="full",
                    help="serve: the plain FP64 server's closed-loop pass over every request "
                         "(full), over the distinct sequences once each (pool) or not at all")
    ap.add_argument("--distinct", type=int, default=0,
                    help="serve: draw the requests from this many distinct generated sequences "
                         "(0: every request its own)")
    ap.add_argument("--keep-factors", type=int, choices=(0, 1), default=0,
                    help="serve: 1 = a continuation solve starts on the ILU(0) factors its "
                         "predecessor left on the device (-ygpu_keep_factors 1)")
    ap.add_argument("--serve-mixed-only", action="store_true",
                    help="serve: the mixed server's passes only (no FP64 server, no error "
                         "reference)")
    ap.add_argument("--serve-fp64-poisson", action="store_true",
                    help="serve: also time the plain FP64 server under Poisson arrivals")
    ap.add_argument("--lanes", default="161:48,321:24,641:8",
                    help="serve: size-class lanes as max_side:slots, sides increasing")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM budget")
    ap.add_argument("--load", type=float, default=0.7,
                    help="serve: Poisson arrival rate as a fraction of the measured capacity")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, run a plain device copy of MIB MiB "
                         "five times (the reference kernel for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.factor is None:
        args.factor = "bf16" if args.workload == "diva" else "fp32"
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
    single = None
    if not args.mixed_only and not args.no_single:
        single = config1_single(args, h)
    served = None
    if not args.mixed_only and args.serve_requests > 0:
        served = served_leg(args)
    baselines = None
    if not args.mixed_only and not args.no_baselines:
        baselines = baseline_legs(args, base, h)
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.members * args.steps / elapsed
    fp64_batch_sps = args.members / t_fp64_batch
    fp64_single_sps = 1.0 / t_fp64_single

    launches = max(st["krylov_launches"], 1)
    avg_ms = st["krylov_ms"] / launches
    bytes_per_launch = st["krylov_bytes"] / launches
    kname = _krylov_kernel_name(st)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    # ---- CPU baseline: the oracle (the reference's arithmetic) on this host's
    # cores, on a bounded sample of members -------------------------------------
    cpu = None
    if not args.no_cpu and not args.mixed_only and args.cpu_sample > 0:
        cpu = cpu_baseline(args, base)

    workload = (f"ensemble {args.members} x {args.side}x{args.side} SSA (dx 8 km, "
                f"{2 * args.side * args.side} unknowns each), mixed FP32/FP64, batched, "
                f"ILU factors {args.factor}")
    traffic, traffic_src = counted_traffic(kname, workload)

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
            "workload": workload,
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
        "error_reference": "plain FP64 GPU path with the linear solves converged to rtol 1e-12",
        "rel_l2_vs_fp64_per_member": [float(f"{e:.3g}") for e in errs],
        "max_abs_vel_err_per_member_m_per_yr": [float(f"{e:.3g}") for e in maxabs],
        "max_abs_vel_err_m_per_yr": max(maxabs),
        "max_rel_l2_vs_fp64_rtol1e-8": max(errs_plain),
        "fp64_rtol1e-8_max_rel_l2_vs_converged": max(err64),
        "error_ok": max(errs) <= 1e-6,
        "fp64_batched_solves_per_s": fp64_batch_sps,
        "fp64_single_solves_per_s": fp64_single_sps,
        "speedup_vs_fp64_single": value / fp64_single_sps,
        "speedup_vs_fp64_batched": value / fp64_batch_sps,
        # the plain FP64 path with all CUs busy, and Yelmo's production
        # defaults (energy form, CG + Jacobi at rtol 1e-2)
        "baselines": baselines,
        "speedup_vs_fp64_full_device": (None if baselines is None else
                                        value / baselines["fp64_full_device"]["solves_per_s"]),
        "eager_stats_solves_per_s": eager_sps,
        "step_latency_ms": {"p50": 1e3 * float(np.percentile(step_times, 50)),
                            "p99": 1e3 * float(np.percentile(step_times, 99))},
        "picard_iters_mean": float(np.mean([i.ssa_iter_now for i in info])),
        "krylov_iters_per_solve": st["krylov_iters"] / max(args.members * ns, 1),
        "unknown_iters_per_s": st["unknown_iters"] / ns * args.steps / elapsed,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     # counted HBM bytes per launch from the committed counter runs
                     "traffic": traffic,
                     "traffic_source": None if traffic_src is None else traffic_src.get("source"),
                     "avg_launch_ms": avg_ms,
                     "bytes_per_launch": bytes_per_launch, "launches": launches,
                     "refine_sweeps_per_launch": st["refine_sweeps"] / launches},
        # config 1: one 161^2 instance, mixed beside plain FP64
        "single_161": single,
        # bounded served stream (continuous batching), p50/p99 request latency
        "served": served,
        "p50_latency_s": None if served is None else served["p50_request_latency_s"],
        "p99_latency_s": None if served is None else served["p99_request_latency_s"],
        # y = A x of all 64 members in one launch (HBM-bound; the FP64 records
        # and vectors, 292 MB, exceed the 256 MB Infinity Cache, the FP32 ones
        # do not)
        "spmv": spmv,
        "cpu_baseline": cpu,
    }
    print(json.dumps(_finite(line)), flush=True)


if __name__ == "__main__":
    main()
