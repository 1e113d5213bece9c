"""hipGraph-captured Picard iterations (-ygpu_graph N, N iterations per
captured graph; 0 = the eager loop that reads the done flags after every
iteration).  Every replay must give the eager result bit for bit: the graph
is replayed many times per solve, the batch is solved again from its saved
initial state several times, with host reads and downloads in between."""
import numpy as np
import pytest

from yelmo_amd import _lib, generator
from yelmo_amd.params import SsaParams
from yelmo_amd.velocity import ResidentBatch

pytestmark = pytest.mark.gpu

STATE = ("ux_b", "uy_b", "visc_eff", "beta_acx", "taub_acx", "ssa_err_acx")


def _probs(diva=False):
    ps = [generator.make_problem(41, 32e3, seed=1, dtype=np.float64),
          generator.make_problem(57, 24e3, seed=2, dtype=np.float64),
          generator.make_problem(33, 40e3, seed=3, dtype=np.float64)]
    return [generator.add_diva_state(p) for p in ps] if diva else ps


def _solve(par, diva, chunk, runs=1, stats=False):
    probs = _probs(diva)
    opt = par.ssa_lis_opt + f" -ygpu_graph {chunk}"
    h = _lib.default_handle()
    rb = ResidentBatch(probs, par.with_(ssa_lis_opt=opt), diva=diva, handle=h)
    out = []
    if stats:
        h.stats_enable(True)
    for _ in range(runs):
        rb.run(reset=True)
        info = rb.download()
        out.append(([{n: p.state[n].copy() for n in STATE} for p in probs],
                    [(i.ssa_iter_now, i.ssa_lin_iter, i.status) for i in info]))
    st = h.stats() if stats else None
    if stats:
        h.stats_enable(False)
    rb.close()
    return out, st


CASES = [
    ("fp64", SsaParams(visc_method=1, beta_method=2, precision="fp64"), False),
    ("mixed", SsaParams(visc_method=1, beta_method=2, precision="mixed"), False),
    ("diva-mixed", SsaParams(visc_method=2, beta_method=4, precision="mixed"), True),
    ("energy", SsaParams(ssa_solver="energy", visc_method=2, beta_method=4, precision="fp64",
                         ssa_lis_opt="-i cg -p jacobi -maxiter 4000 -tol 1.0e-8"), False),
]


@pytest.mark.parametrize("name,par,diva", CASES, ids=[c[0] for c in CASES])
def test_graph_replays_equal_eager_loop(name, par, diva):
    print(f"case graph {name}", flush=True)
    eager, _ = _solve(par, diva, 0)
    for chunk in (1, 4, 7):
        runs, _ = _solve(par, diva, chunk, runs=3)
        for r, (states, infos) in enumerate(runs):
            print(chunk, r, infos, flush=True)
            assert infos == eager[0][1], (chunk, r)
            for k, st in enumerate(states):
                for n in STATE:
                    assert np.array_equal(st[n], eager[0][0][k][n]), (chunk, r, k, n)


def test_graph_and_instrumented_eager_runs_agree():
    """Kernel statistics time every Krylov launch and therefore run the eager
    loop; its result equals the graphed run without statistics."""
    print("case graph vs stats", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="mixed")
    (eager,), st = _solve(par, False, 4, stats=True)
    (graphed,), _ = _solve(par, False, 4)
    picard = [i[0] for i in eager[1]]
    assert st["krylov_launches"] == max(picard)
    assert eager[1] == graphed[1]
    for k in range(len(eager[0])):
        for n in STATE:
            assert np.array_equal(eager[0][k][n], graphed[0][k][n])
