"""Refinement variants of the mixed precision.

Reliable updates (-ygpu_ir reliable): one FP32 BiCGSTAB per linear solve
whose iterated residual is replaced by the FP64 residual b - A x (x
accumulated in FP64) each time it has fallen by the factor -ygpu_ir_delta,
the search direction kept, instead of a new inner solve per FP64 residual.
Loose solves far from Picard convergence (-ygpu_ir_loose): refinement to
-tol itself while a member's Picard residual is large.  The stopping rules
are the refinement's, so the bars are the mixed path's:

  velocity vs converged FP64 oracle      1e-6 rel L2, Picard count +-1
  CSR solve vs converged FP64 oracle      1e-6 rel L2, status 0, FP64 residual <= tol
  graph replays vs eager loop             bitwise (group sums in a fixed order)
  iteration limit                         status 4 after exactly maxiter iterations,
                                          x finite (the FP32 correction folded in)
  loose rule off / never triggered        bitwise equal
"""
import numpy as np
import pytest

from oracle import oracle
from tests._helpers import vel_rel_l2
from yelmo_amd import _lib, generator, velocity
from yelmo_amd.params import SsaParams
from yelmo_amd.velocity import ResidentBatch

pytestmark = pytest.mark.gpu

OPT = "-i bicgstab -p ilu -maxiter 1000 -tol 1.0e-8 -initx_zeros false -ygpu_ir reliable"
STATE = ("ux_b", "uy_b", "visc_eff", "beta_acx", "taub_acx", "ssa_err_acx")


def _converged(par):
    return par.with_(precision="fp64",
                     ssa_lis_opt="-i bicgstab -p ilu -maxiter 3000 -tol 1.0e-12 -initx_zeros false")


def _batch(probs, par, opt, runs=1):
    h = _lib.default_handle()
    rb = ResidentBatch(probs, par.with_(ssa_lis_opt=opt), handle=h)
    out = []
    for _ in range(runs):
        rb.run(reset=True)
        info = rb.download()
        out.append(([{n: p.state[n].copy() for n in STATE} for p in probs],
                    [(i.ssa_iter_now, i.ssa_lin_iter, i.status) for i in info]))
    rb.close()
    return out


@pytest.mark.parametrize("G", [1, 4])
def test_reliable_batch_within_mixed_bar(G):
    """Three members of different sizes; G = 1 runs the one-workgroup
    kernel, G = 4 the member-group kernel."""
    print(f"case reliable batch G={G}", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    sides = ((161, 8e3, 5), (129, 10e3, 6), (97, 14e3, 7))
    probs = [generator.make_problem(n, dx, seed=s, dtype=np.float64) for n, dx, s in sides]
    (states, infos), = _batch([p.copy() for p in probs], par,
                              OPT + f" -ygpu_factor fp32 -ygpu_group {G}")
    for k, p in enumerate(probs):
        ref_st, ref = oracle.calc_velocity_ssa(p.copy(), _converged(par), dp=True)
        err = vel_rel_l2(states[k], ref_st)
        print(k, "picard/krylov/status", infos[k], "oracle picard", ref["ssa_iter_now"],
              "rel L2", err, flush=True)
        assert infos[k][2] == 0
        assert abs(infos[k][0] - ref["ssa_iter_now"]) <= 1
        assert err <= 1e-6


def test_reliable_graph_replays_equal_eager_loop():
    print("case reliable graph", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    probs = [generator.make_problem(41, 32e3, seed=1, dtype=np.float64),
             generator.make_problem(57, 24e3, seed=2, dtype=np.float64)]
    (eager, ie), = _batch([p.copy() for p in probs], par, OPT + " -ygpu_graph 0")
    runs = _batch([p.copy() for p in probs], par, OPT + " -ygpu_graph 4", runs=3)
    for states, infos in runs:
        assert infos == ie
        for k in range(len(probs)):
            for n in STATE:
                assert np.array_equal(states[k][n], eager[k][n]), (k, n)


def _mid_picard_system(n=81, dx=16e3):
    par = SsaParams(visc_method=2, beta_method=4, ssa_iter_max=4)
    p = generator.make_problem(n, dx, seed=0, dtype=np.float64)
    st, _ = oracle.calc_velocity_ssa(p, par, dp=True)
    q = p.copy()
    q.state.update({k: st[k] for k in ("ux_b", "uy_b")})
    return oracle.assemble(q, np.asfortranarray(st["beta_acx"]), np.asfortranarray(st["beta_acy"]),
                           np.asfortranarray(st["visc_eff_int"]))


def test_reliable_csr_solve_and_iteration_limit():
    """The inner boundary in mixed precision with reliable updates, and the
    iteration limit (the correction accumulated so far reaches x)."""
    print("case reliable csr 81", flush=True)
    ap, ai, av, b, x0, _ = _mid_picard_system()
    base = "-i bicgstab -p ilu -maxiter {m} -tol 1.0e-8 -initx_zeros false -ygpu_precision mixed"
    # the error reference: the oracle's solve converged to rtol 1e-12 (its
    # solve at the 1e-8 tolerance is itself ~1e-6 away from the solution)
    ref = oracle.csr_solve(ap, ai, av, b, x0,
                           "-i bicgstab -p ilu -maxiter 3000 -tol 1.0e-12 -initx_zeros false")
    for delta in ("", " -ygpu_ir_delta 0.1"):
        lgs = velocity.Lgs(ap, ai, av, b, x0.copy())
        velocity.linear_solver_matrix_solve(lgs, base.format(m=1000) + " -ygpu_ir reliable" + delta)
        err = float(np.linalg.norm(lgs.x_value - ref["x"]) / np.linalg.norm(ref["x"]))
        print("reliable" + delta, "iters", lgs.lin_iter, "oracle", ref["iters"], "status", lgs.status,
              "resid", lgs.L2_rel_norm, "err", err, flush=True)
        assert lgs.status == 0
        assert lgs.L2_rel_norm <= 1e-8
        assert err <= 1e-6
    lgs = velocity.Lgs(ap, ai, av, b, x0.copy())
    velocity.linear_solver_matrix_solve(lgs, base.format(m=3) + " -ygpu_ir reliable")
    print("maxiter 3: iters", lgs.lin_iter, "status", lgs.status, "resid", lgs.L2_rel_norm, flush=True)
    assert lgs.status == 4
    assert lgs.lin_iter == 3
    assert np.all(np.isfinite(lgs.x_value))
    assert lgs.L2_rel_norm < 1.0


def test_reliable_option_values():
    h = _lib.default_handle()
    ap, ai, av, b, x0, _ = _mid_picard_system(21, 60e3)
    for bad in ("-ygpu_ir sometimes", "-ygpu_ir_delta 1.5", "-ygpu_ir_delta 0"):
        lgs = velocity.Lgs(ap, ai, av, b, x0.copy())
        with pytest.raises(Exception):
            velocity.linear_solver_matrix_solve(lgs, "-i bicgstab -p ilu -ygpu_precision mixed " + bad,
                                                handle=h)


def test_loose_refinement_far_from_picard_convergence():
    """-ygpu_ir_loose R: members whose last Picard residual is above R stop
    refining at -tol with looser inner solves (automatic R: 10
    ssa_iter_conv).  Off (0) and a threshold no residual reaches give the
    same bits; the automatic rule stays within the mixed bar of the
    converged oracle, Picard count +-1, with no more Krylov iterations
    over the members."""
    print("case loose refinement", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    probs = [generator.make_problem(97, 14e3, seed=21, dtype=np.float64),
             generator.make_problem(129, 10e3, seed=22, dtype=np.float64)]
    base = "-i bicgstab -p ilu -maxiter 1000 -tol 1.0e-8 -initx_zeros false -ygpu_factor fp32"
    (s_auto, i_auto), = _batch([p.copy() for p in probs], par, base)
    (s_off, i_off), = _batch([p.copy() for p in probs], par, base + " -ygpu_ir_loose 0")
    (s_never, i_never), = _batch([p.copy() for p in probs], par, base + " -ygpu_ir_loose 1e30")
    assert i_off == i_never
    for k in range(len(probs)):
        for n in STATE:
            assert np.array_equal(s_off[k][n], s_never[k][n])
    for k, p in enumerate(probs):
        ref_st, ref = oracle.calc_velocity_ssa(p.copy(), _converged(par), dp=True)
        e_auto, e_off = vel_rel_l2(s_auto[k], ref_st), vel_rel_l2(s_off[k], ref_st)
        print(k, "auto", i_auto[k], e_auto, "off", i_off[k], e_off, "oracle picard",
              ref["ssa_iter_now"], flush=True)
        assert i_auto[k][2] == 0 and e_auto <= 1e-6 and e_off <= 1e-6
        assert abs(i_auto[k][0] - ref["ssa_iter_now"]) <= 1
    # fewer Krylov iterations over the members (a member's own count can
    # come out a few iterations either way)
    assert sum(i[1] for i in i_auto) <= sum(i[1] for i in i_off)
    h = _lib.default_handle()
    ap, ai, av, b, x0, _ = _mid_picard_system(21, 60e3)
    for bad in ("-ygpu_ir_loose -1", "-ygpu_ir_loose_inner 2"):
        lgs = velocity.Lgs(ap, ai, av, b, x0.copy())
        with pytest.raises(Exception):
            velocity.linear_solver_matrix_solve(lgs, "-i bicgstab -p ilu -ygpu_precision mixed " + bad,
                                                handle=h)
