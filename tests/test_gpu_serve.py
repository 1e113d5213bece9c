"""Served solve sequences (ygpu_server_*): continuous batching gives each
solve the same result as solving its sequence alone, and a sequence gives
the oracle's result for the same chain of calc_velocity_ssa calls.

Tolerances:
  server (in the mix) vs session (alone), same precision   bitwise equal
  served FP64 sequence vs FP64 oracle chain               1e-8 rel L2, Picard +-1
"""
import numpy as np
import pytest

from oracle import oracle
from tests._helpers import vel_rel_l2
from yelmo_amd import generator
from yelmo_amd.params import SsaParams
from yelmo_amd.serve import Server
from yelmo_amd.velocity import Session

pytestmark = pytest.mark.gpu

# state a solve leaves for the next solve of its sequence
CARRY = ("ux_b", "uy_b", "taub_acx", "taub_acy", "visc_eff", "beta", "beta_acx", "beta_acy")
COMPARE = ("ux_b", "uy_b", "visc_eff", "visc_eff_int", "beta_acx", "beta_acy", "taub_acx",
           "ssa_err_acx")


def _sequences(sides, nsolves, dtype, seed=0):
    return [generator.served_sequence(int(s), seed + 7 * k, nsolves, dtype=dtype)
            for k, s in enumerate(sides)]


def _alone(seq, par):
    """A sequence solved by itself in a device session (warm starts)."""
    out = []
    ses = None
    for k, p in enumerate(seq):
        q = p.copy()
        if ses is None:
            ses = Session(q, par)
        info = ses.solve(q)
        out.append((q.state, info))
    ses.close()
    return out


def _served(seqs, par, *, max_slots=4, budget, arrive_every=1, lanes=None):
    """All sequences through one server; sequence k arrives at step
    k * arrive_every and submits its next solve as soon as the previous one
    is fetched.  Returns results[seq][solve] = (state, info)."""
    srv = Server(par, max_slots=max_slots, max_side=max(s[0].nx for s in seqs),
                 nz_max=seqs[0][0].nz, hbm_budget=budget, lanes=lanes)
    res = [[None] * len(s) for s in seqs]
    owner = {}
    arrived = 0
    steps = 0
    max_running = 0
    while any(r is None for rs in res for r in rs):
        while arrived < len(seqs) and steps >= arrived * arrive_every:
            t = srv.submit(seqs[arrived][0].copy(), arrived, last=len(seqs[arrived]) == 1)
            owner[t] = (arrived, 0)
            arrived += 1
        nr, nq = srv.step()
        max_running = max(max_running, nr)
        steps += 1
        assert steps < 20000, "server made no progress"
        for t in srv.poll():
            k, j = owner.pop(t)
            q = seqs[k][j].copy()
            info = srv.fetch(t, q)
            res[k][j] = (q.state, info)
            if j + 1 < len(seqs[k]):
                t2 = srv.submit(seqs[k][j + 1].copy(), k, cont=True,
                                last=j + 2 == len(seqs[k]))
                owner[t2] = (k, j + 1)
    srv.close()
    return res, max_running


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_served_alone_equals_in_mix(prec):
    """Four sequences of three solves on uneven grids through a server with
    three slots (one sequence must wait for a slot), arriving one step apart:
    every solve equals the same sequence solved alone, bit for bit."""
    print(f"case serve mix {prec}", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision=prec)
    dt = np.float64
    seqs = _sequences((41, 57, 33, 65), 3, dt)
    alone = [_alone(s, par) for s in seqs]
    mixed, max_running = _served(seqs, par, max_slots=3, budget=1 << 30)
    assert max_running >= 2, "solves never overlapped"
    for k in range(len(seqs)):
        for j in range(len(seqs[k])):
            st_a, inf_a = alone[k][j]
            st_m, inf_m = mixed[k][j]
            print(k, j, "iters", inf_a.ssa_iter_now, inf_m.ssa_iter_now, "lin",
                  inf_a.ssa_lin_iter, inf_m.ssa_lin_iter, flush=True)
            assert inf_m.status == inf_a.status == 0
            assert (inf_m.ssa_iter_now, inf_m.ssa_lin_iter) == (inf_a.ssa_iter_now,
                                                               inf_a.ssa_lin_iter)
            for n in COMPARE:
                assert np.array_equal(st_m[n], st_a[n]), (k, j, n)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_served_lanes_alone_equals_in_mix(prec):
    """Two size-class lanes stepping on their own streams (sides up to 45
    with two slots, up to 65 with two slots); small sequences overflow into
    the large lane when theirs is full.  Every solve equals its sequence
    alone, bit for bit."""
    print(f"case serve lanes {prec}", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision=prec)
    seqs = _sequences((33, 65, 41, 57, 45, 37), 3, np.float64, seed=21)
    alone = [_alone(s, par) for s in seqs]
    mixed, max_running = _served(seqs, par, budget=1 << 30, arrive_every=0,
                                 lanes=[(45, 2), (65, 2)])
    assert max_running >= 3
    for k in range(len(seqs)):
        for j in range(len(seqs[k])):
            st_a, inf_a = alone[k][j]
            st_m, inf_m = mixed[k][j]
            assert inf_m.status == inf_a.status == 0
            assert (inf_m.ssa_iter_now, inf_m.ssa_lin_iter) == (inf_a.ssa_iter_now,
                                                               inf_a.ssa_lin_iter), (k, j)
            for n in COMPARE:
                assert np.array_equal(st_m[n], st_a[n]), (k, j, n)


def test_served_admission_by_footprint():
    """A budget that holds one 65x65 domain but not two: the second large
    sequence waits for the first to release its cells, small ones pass it,
    and every result still equals its sequence alone."""
    print("case serve footprint", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    seqs = _sequences((65, 65, 33), 2, np.float64, seed=3)
    # bytes per cell of the mixed server at nz = 11 (FP64 fields): about 1.2 kB
    budget = int(1.6 * 65 * 65 * 1300)
    alone = [_alone(s, par) for s in seqs]
    mixed, _ = _served(seqs, par, max_slots=4, budget=budget, arrive_every=0)
    for k in range(len(seqs)):
        for j in range(len(seqs[k])):
            assert np.array_equal(mixed[k][j][0]["ux_b"], alone[k][j][0]["ux_b"]), (k, j)
            assert np.array_equal(mixed[k][j][0]["visc_eff"], alone[k][j][0]["visc_eff"]), (k, j)


def test_served_sequence_matches_oracle_chain():
    """A served FP64 sequence equals the oracle's calc_velocity_ssa called
    once per solve, each call starting from the previous call's outputs."""
    print("case serve oracle chain", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="fp64")
    seq = _sequences((33,), 3, np.float64, seed=11)[0]
    served, _ = _served([seq], par, max_slots=2, budget=1 << 28)
    prev = None
    for j, p in enumerate(seq):
        q = p.copy()
        if prev is not None:
            q.state.update({n: prev[n].copy() for n in CARRY})
        ref_st, ref = oracle.calc_velocity_ssa(q, par, dp=True)
        st, info = served[0][j]
        err = vel_rel_l2(st, ref_st)
        print(j, "iters", info.ssa_iter_now, ref["ssa_iter_now"], "rel L2", err, flush=True)
        assert ref["rc"] == 0 and info.status == 0
        assert abs(info.ssa_iter_now - ref["ssa_iter_now"]) <= 1
        assert err <= 1e-8
        prev = ref_st


def test_served_group_lane_alone_equals_in_mix():
    """The largest lane runs its members with member groups (here -ygpu_group
    4: four workgroups per member); sequences of that lane give the result of
    the same sequence solved alone in a session with the same option, bit for
    bit, while a small lane of one workgroup per member runs beside it."""
    print("case serve group lane", flush=True)
    base = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    par = base.with_(ssa_lis_opt=base.ssa_lis_opt + " -ygpu_group 4")
    seqs = _sequences((97, 129, 81, 113), 3, np.float64, seed=31)
    alone = [_alone(s, par) for s in seqs]
    small = _sequences((41,), 3, np.float64, seed=77)
    small_alone = _alone(small[0], base.with_(ssa_lis_opt=base.ssa_lis_opt + " -ygpu_group 1"))
    mixed, max_running = _served(seqs + small, par, budget=1 << 30, arrive_every=0,
                                 lanes=[(65, 1), (129, 3)])
    assert max_running >= 3
    for k in range(len(seqs)):
        for j in range(len(seqs[k])):
            st_a, inf_a = alone[k][j]
            st_m, inf_m = mixed[k][j]
            assert inf_m.status == inf_a.status == 0
            assert (inf_m.ssa_iter_now, inf_m.ssa_lin_iter) == (inf_a.ssa_iter_now,
                                                               inf_a.ssa_lin_iter), (k, j)
            for n in COMPARE:
                assert np.array_equal(st_m[n], st_a[n]), (k, j, n)
    for j in range(len(small[0])):
        assert np.array_equal(mixed[len(seqs)][j][0]["ux_b"], small_alone[j][0]["ux_b"]), j


def test_served_group_lanes_alone_equals_in_mix():
    """By default every lane whose domains are large enough takes member
    groups (here 97^2: 2 workgroups per member, 129^2: 4), both group kernels
    in flight at once; each sequence gives the result of the same sequence
    solved alone in a session, whose batch rule picks the same group size,
    bit for bit."""
    print("case serve group lanes", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    seqs = _sequences((91, 129, 97, 129), 3, np.float64, seed=41)
    alone = [_alone(s, par) for s in seqs]
    mixed, max_running = _served(seqs, par, budget=1 << 30, arrive_every=0,
                                 lanes=[(97, 2), (129, 2)])
    assert max_running >= 3
    for k in range(len(seqs)):
        for j in range(len(seqs[k])):
            st_a, inf_a = alone[k][j]
            st_m, inf_m = mixed[k][j]
            print(k, j, "iters", inf_a.ssa_iter_now, inf_m.ssa_iter_now, "lin",
                  inf_a.ssa_lin_iter, inf_m.ssa_lin_iter, flush=True)
            assert inf_m.status == inf_a.status == 0
            assert (inf_m.ssa_iter_now, inf_m.ssa_lin_iter) == (inf_a.ssa_iter_now,
                                                               inf_a.ssa_lin_iter), (k, j)
            for n in COMPARE:
                assert np.array_equal(st_m[n], st_a[n]), (k, j, n)
