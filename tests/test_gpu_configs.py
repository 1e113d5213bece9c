"""The benchmark configurations at their real shapes, against the oracle.

  ensemble: 64 members on 161 x 161 (dx 8 km, nz 11), each with its own
    seeded friction field and A log-uniform in [1e-17, 1e-16]; one batched
    call on the library's default path (member groups, Picard iterations
    replayed as captured hipGraphs), mixed precision.
  DIVA batch: 48 domains with sides from {96, 128, 192, 256, 384} spanning
    1536 km, nz = 21, A(zeta) = 1e-16 exp(-3 (1 - zeta)); BF16 ILU(0) factors
    with per-row scaling, schedules from the autotuner, mixed precision.

Bars: mixed precision within 1e-6 relative L2 of the FP64 solution with the
linear solves converged to rtol 1e-12, for every member against the GPU's
FP64 path and for sampled members (the smallest and largest rate factor;
one domain of every DIVA side, the largest 384 x 384) against the FP64
oracle; the GPU's converged FP64 path within 1e-8 of the oracle.
The oracle runs in threads beside the GPU work (ctypes releases the GIL).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle
from tests._helpers import vel_rel_l2
from yelmo_amd import _lib, generator
from yelmo_amd.params import SsaParams
from yelmo_amd.tune import autotune_sizes
from yelmo_amd.velocity import ResidentBatch

pytestmark = pytest.mark.gpu


def _converged(par):
    return par.with_(precision="fp64", ssa_lis_opt=par.ssa_lis_opt.replace("1.0e-8", "1.0e-12"))


def test_ensemble_64x161_mixed_matches_converged_fp64():
    print("case ensemble 64 x 161^2 mixed", flush=True)
    base = generator.workload_ensemble(64, 161, seed=0, dtype=np.float32)
    A = np.array([p.fields["ATT"][0, 0, 0] for p in base])
    sample = sorted({int(np.argmin(A)), int(np.argmax(A)), 0, 37})
    par = SsaParams(precision="mixed")
    conv = _converged(par)
    with ThreadPoolExecutor(len(sample)) as ex:
        futs = {k: ex.submit(oracle.calc_velocity_ssa, base[k].astype(np.float64), conv, True)
                for k in sample}
        probs = [p.copy() for p in base]
        rb = ResidentBatch(probs, par)
        rb.run()
        info = rb.download()
        rb.close()
        print("mixed done, picard", min(i.ssa_iter_now for i in info), "-",
              max(i.ssa_iter_now for i in info), flush=True)
        probs64 = [p.astype(np.float64) for p in base]
        rb = ResidentBatch(probs64, conv)
        rb.run()
        info64 = rb.download()
        rb.close()
        print("fp64 converged done", flush=True)
        ref = {k: f.result() for k, f in futs.items()}
    assert all(i.status == 0 for i in info) and all(i.status == 0 for i in info64)
    errs = [vel_rel_l2(probs[k].state, probs64[k].state) for k in range(64)]
    print("per-member rel L2 vs converged FP64: max", max(errs), "median", np.median(errs),
          flush=True)
    assert max(errs) <= 1e-6
    for k in sample:
        st, ri = ref[k]
        assert ri["rc"] == 0
        e_mixed = vel_rel_l2(probs[k].state, st)
        e_64 = vel_rel_l2(probs64[k].state, st)
        print(f"member {k} A {A[k]:.3e}: mixed vs oracle {e_mixed:.2e}, fp64 vs oracle {e_64:.2e}, "
              f"picard gpu {info64[k].ssa_iter_now} oracle {ri['ssa_iter_now']}", flush=True)
        assert e_mixed <= 1e-6
        assert e_64 <= 1e-8
        assert abs(info64[k].ssa_iter_now - ri["ssa_iter_now"]) <= 1


def _ubar(p):
    return {"ux_b": p.state["ux_bar"], "uy_b": p.state["uy_bar"]}


def test_diva_uneven_48_bf16_autotuned_matches_converged_fp64():
    print("case DIVA 48 uneven, bf16 factors, autotuned", flush=True)
    base = generator.workload_diva_uneven(48, seed=0)
    sides = [p.nx for p in base]
    sample = [sides.index(s) for s in generator.DIVA_SIDES]
    par = SsaParams(precision="mixed")
    par = par.with_(ssa_lis_opt=par.ssa_lis_opt + " -ygpu_factor bf16")
    conv = _converged(SsaParams())
    h = _lib.Handle(0)
    with ThreadPoolExecutor(len(sample)) as ex:
        futs = {k: ex.submit(oracle.calc_velocity_diva, base[k].astype(np.float64), conv, True)
                for k in sample}
        sched = autotune_sizes(base, par, diva=True, reps=1, handle=h)
        print("schedules", {k[0]: (v.ilu_blocks, v.block) for k, v in sched.items()}, flush=True)
        probs = [oracle.diva_state(p.copy()) for p in base]
        rb = ResidentBatch(probs, par, diva=True, handle=h)
        rb.run()
        info = rb.download()
        rb.close()
        print("mixed done", flush=True)
        probs64 = [oracle.diva_state(p.astype(np.float64)) for p in base]
        rb = ResidentBatch(probs64, conv, diva=True, handle=h)
        rb.run()
        info64 = rb.download()
        rb.close()
        print("fp64 converged done", flush=True)
        ref = {k: f.result() for k, f in futs.items()}
    h.close()
    assert all(i.status == 0 for i in info) and all(i.status == 0 for i in info64)
    errs = [vel_rel_l2(_ubar(probs[k]), _ubar(probs64[k])) for k in range(48)]
    print("per-domain rel L2 (ubar) vs converged FP64: max", max(errs), flush=True)
    assert max(errs) <= 1e-6
    for k in sample:
        st, ri = ref[k]
        assert ri["rc"] == 0
        e_mixed = vel_rel_l2(_ubar(probs[k]), {"ux_b": st["ux_bar"], "uy_b": st["uy_bar"]})
        e_64 = vel_rel_l2(_ubar(probs64[k]), {"ux_b": st["ux_bar"], "uy_b": st["uy_bar"]})
        e3 = vel_rel_l2(probs[k].state, st, "ux", "uy")
        print(f"domain {k} side {sides[k]}: mixed vs oracle {e_mixed:.2e} (3D {e3:.2e}), "
              f"fp64 vs oracle {e_64:.2e}, picard gpu {info64[k].ssa_iter_now} "
              f"oracle {ri['ssa_iter_now']}", flush=True)
        assert e_mixed <= 1e-6 and e3 <= 1e-6
        assert e_64 <= 1e-8
        assert abs(info64[k].ssa_iter_now - ri["ssa_iter_now"]) <= 1
