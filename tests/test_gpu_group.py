"""Member groups (-ygpu_group G): G workgroups, one per CU, share each
member's mixed-precision solve (block-Jacobi ILU(0) line blocks split over
the workgroups, group sums and group barriers between them).  The
preconditioner's line blocks change with G (their count is rounded up to a
multiple of G) and sums are taken in another order, so iteration counts
differ from one workgroup per member; the refinement meets the same FP64
residual criterion and every member stays within the mixed bar of the
converged FP64 oracle:

  velocity vs converged FP64 oracle   1e-6 rel L2 (the mixed-precision bar)
  graph replays vs eager loop         bitwise (group sums have a fixed order)
"""
import numpy as np
import pytest

from oracle import oracle
from tests._helpers import vel_rel_l2
from yelmo_amd import _lib, generator
from yelmo_amd.params import SsaParams
from yelmo_amd.velocity import ResidentBatch

pytestmark = pytest.mark.gpu

OPT = "-i bicgstab -p ilu -maxiter 1000 -tol 1.0e-8 -initx_zeros false"
STATE = ("ux_b", "uy_b", "visc_eff", "beta_acx", "taub_acx", "ssa_err_acx")


def _converged(par):
    return par.with_(precision="fp64",
                     ssa_lis_opt="-i bicgstab -p ilu -maxiter 3000 -tol 1.0e-12 -initx_zeros false")


def _batch(probs, par, opt, runs=1, diva=False):
    h = _lib.default_handle()
    rb = ResidentBatch(probs, par.with_(ssa_lis_opt=opt), diva=diva, handle=h)
    out = []
    for _ in range(runs):
        rb.run(reset=True)
        info = rb.download()
        out.append(([{n: p.state[n].copy() for n in STATE} for p in probs],
                    [(i.ssa_iter_now, i.ssa_lin_iter, i.status) for i in info]))
    rb.close()
    return out


@pytest.mark.parametrize("G,threads", [(2, 512), (4, 512), (8, 512), (8, 256)])
def test_group_members_within_mixed_bar(G, threads):
    """Three members of different sizes (the largest sets the groups' line
    counts), factors in BF16; 256 threads: two workgroups per CU."""
    print(f"case group G={G} threads={threads}", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    sides = ((161, 8e3, 5), (129, 10e3, 6), (97, 14e3, 7))
    probs = [generator.make_problem(n, dx, seed=s, dtype=np.float64) for n, dx, s in sides]
    (states, infos), = _batch([p.copy() for p in probs], par,
                              OPT + f" -ygpu_factor bf16 -ygpu_group {G} -ygpu_group_threads {threads}")
    for k, p in enumerate(probs):
        ref_st, ref = oracle.calc_velocity_ssa(p.copy(), _converged(par), dp=True)
        err = vel_rel_l2(states[k], ref_st)
        print(k, "picard/krylov/status", infos[k], "oracle picard", ref["ssa_iter_now"],
              "rel L2", err, flush=True)
        assert infos[k][2] == 0
        assert err <= 1e-6


def test_group_members_in_turn_and_empty_ranks():
    """40 members of 41 x 41 with G = 8: 32 groups take the first 32
    members and then the rest; 41 lines make 7 line blocks, so one
    workgroup of every group owns no lines.  FP32 factors."""
    print("case group many", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    probs = [generator.make_problem(41, 32e3, seed=100 + k, dtype=np.float64) for k in range(40)]
    (states, infos), = _batch([p.copy() for p in probs], par,
                              OPT + " -ygpu_factor fp32 -ygpu_group 8")
    worst = 0.0
    for k, p in enumerate(probs):
        ref_st, _ = oracle.calc_velocity_ssa(p.copy(), _converged(par), dp=True)
        err = vel_rel_l2(states[k], ref_st)
        worst = max(worst, err)
        assert infos[k][2] == 0, (k, infos[k])
        assert err <= 1e-6, (k, err)
    print("worst rel L2", worst, flush=True)


def test_group_graph_replays_equal_eager_loop():
    print("case group graph", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    probs = [generator.make_problem(n, 10e3, seed=20 + n, dtype=np.float64) for n in (97, 113, 81)]
    opt = OPT + " -ygpu_factor bf16 -ygpu_group 4"
    (eager,) = _batch([p.copy() for p in probs], par, opt + " -ygpu_graph 0")
    runs = _batch([p.copy() for p in probs], par, opt + " -ygpu_graph 3", runs=3)
    for r, (states, infos) in enumerate(runs):
        print(r, infos, flush=True)
        assert infos == eager[1]
        for k, st in enumerate(states):
            for n in STATE:
                assert np.array_equal(st[n], eager[0][k][n]), (r, k, n)


def test_group_diva_within_mixed_bar():
    print("case group diva", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    p = generator.make_problem(96, 16e3, seed=7, dtype=np.float64, nz=21)
    ref_st, _ = oracle.calc_velocity_diva(p.copy(), _converged(par), dp=True)
    g = oracle.diva_state(p.copy())
    (states, infos), = _batch([g], par, OPT + " -ygpu_factor bf16 -ygpu_group 2", diva=True)
    err = vel_rel_l2(g.state, ref_st, "ux_bar", "uy_bar")
    print("picard/krylov/status", infos[0], "rel L2", err, flush=True)
    assert infos[0][2] == 0 and err <= 1e-6


def test_group_energy_formulation_within_mixed_bar():
    """The energy formulation's systems (element Hessians, free-slip
    folding) under member groups: mixed BiCGSTAB + line-block ILU(0) with
    G = 4 against the converged FP64 oracle of the same formulation."""
    print("case group energy", flush=True)
    par = SsaParams(ssa_solver="energy", boundaries="zeros", visc_method=2, beta_method=4,
                    precision="mixed")
    p = generator.make_problem(97, 14e3, seed=12, dtype=np.float64)
    ref_st, _ = oracle.calc_velocity_ssa(p.copy(), _converged(par), dp=True)
    (states, infos), = _batch([p.copy()], par, OPT + " -ygpu_factor fp32 -ygpu_group 4")
    err = vel_rel_l2(states[0], ref_st)
    print("picard/krylov/status", infos[0], "rel L2", err, flush=True)
    assert infos[0][2] == 0 and err <= 1e-6


@pytest.mark.parametrize("G,cols", [(1, 2), (4, 2), (2, 3)])
def test_column_tiles_within_mixed_bar(G, cols):
    """Block-Jacobi ILU(0) over tiles: every line block cut into `cols`
    column blocks (-ygpu_ilu_cols), with one workgroup per member and with
    member groups; the refinement meets the same criterion and the velocity
    stays within the mixed bar of the converged FP64 oracle."""
    print(f"case tiles G={G} cols={cols}", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    sides = ((161, 8e3, 5), (113, 12e3, 8))
    probs = [generator.make_problem(n, dx, seed=s, dtype=np.float64) for n, dx, s in sides]
    for fac in ("fp32", "bf16"):
        (states, infos), = _batch([p.copy() for p in probs], par,
                                  OPT + f" -ygpu_factor {fac} -ygpu_group {G} -ygpu_ilu_cols {cols}")
        for k, p in enumerate(probs):
            ref_st, _ = oracle.calc_velocity_ssa(p.copy(), _converged(par), dp=True)
            err = vel_rel_l2(states[k], ref_st)
            print(fac, k, "picard/krylov/status", infos[k], "rel L2", err, flush=True)
            assert infos[k][2] == 0
            assert err <= 1e-6


@pytest.mark.parametrize("G,cols", [(1, 1), (4, 1), (4, 3), (8, 1)])
def test_reversed_columns_within_mixed_bar(G, cols):
    """Block-Jacobi ILU(0) with the columns of every line block (or tile)
    numbered in reverse (-ygpu_ilu_order reverse: a skew-1 wavefront), FP32
    factors, one workgroup per member and member groups: every member meets
    the refinement criterion within the mixed bar of the converged FP64
    oracle.  With BF16 factors the option is ignored (natural order), so the
    result equals the natural order's bit for bit."""
    print(f"case reversed G={G} cols={cols}", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    sides = ((161, 8e3, 5), (113, 12e3, 8), (41, 32e3, 9))
    probs = [generator.make_problem(n, dx, seed=s, dtype=np.float64) for n, dx, s in sides]
    base = OPT + f" -ygpu_group {G} -ygpu_ilu_cols {cols}"
    (states, infos), = _batch([p.copy() for p in probs], par,
                              base + " -ygpu_factor fp32 -ygpu_ilu_order reverse")
    (nat_states, nat_infos), = _batch([p.copy() for p in probs], par, base + " -ygpu_factor fp32")
    for k, p in enumerate(probs):
        ref_st, _ = oracle.calc_velocity_ssa(p.copy(), _converged(par), dp=True)
        err = vel_rel_l2(states[k], ref_st)
        print(k, "picard/krylov/status", infos[k], "natural", nat_infos[k], "rel L2", err,
              flush=True)
        assert infos[k][2] == 0
        assert err <= 1e-6
    (b_rev, i_rev), = _batch([p.copy() for p in probs[:2]], par,
                             base + " -ygpu_factor bf16 -ygpu_ilu_order reverse")
    (b_nat, i_nat), = _batch([p.copy() for p in probs[:2]], par, base + " -ygpu_factor bf16")
    assert i_rev == i_nat
    for k in range(2):
        for n in STATE:
            assert np.array_equal(b_rev[k][n], b_nat[k][n]), (k, n)


def test_uneven_batch_members_get_their_own_group_sizes():
    """Automatic groups on a batch of uneven sizes: each member gets the
    group size its own grid asks for (57^2: 1, 97^2: 2, 129^2 and 161^2: 4,
    193^2: 8, 257^2: 16), the groups placed by the launch table.  Every
    member is within the mixed bar of the converged FP64 GPU path (the FP64
    path is held to the oracle elsewhere), and a member with a group of two
    or more gives the bits of the same member solved alone (its group size,
    line blocks and order of sums do not depend on the batch)."""
    print("case group uneven", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    sides = ((57, 28e3, 11), (97, 14e3, 12), (129, 10e3, 13), (161, 8e3, 14), (193, 7e3, 15),
             (257, 5e3, 16))
    probs = [generator.make_problem(n, dx, seed=s, dtype=np.float64) for n, dx, s in sides]
    opt = OPT + " -ygpu_factor fp32"
    (mix, imix), = _batch([p.copy() for p in probs], par, opt)
    (ref, iref), = _batch([p.copy() for p in probs], _converged(par),
                          "-i bicgstab -p ilu -maxiter 3000 -tol 1.0e-12 -initx_zeros false")
    for k, p in enumerate(probs):
        err = vel_rel_l2(mix[k], ref[k])
        print(sides[k][0], "picard/krylov/status", imix[k], "fp64", iref[k], "rel L2", err, flush=True)
        assert imix[k][2] == 0
        assert err <= 1e-6
        assert abs(imix[k][0] - iref[k][0]) <= 1
        if sides[k][0] == 57:
            continue
        (alone, ia), = _batch([p.copy()], par, opt)
        assert ia[0] == imix[k], (sides[k][0], ia[0], imix[k])
        for n in STATE:
            assert np.array_equal(alone[0][n], mix[k][n]), (sides[k][0], n)
