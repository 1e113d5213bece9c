"""The reference's namelists also name LIS's MINRES and BiCGSafe
(input/yelmo_defaults.nml:160 for the residual formulation, par/*.nml).
Both are accepted and solve the same system to the same tolerance with
BiCGSTAB: the oracle gives the identical result for the three option
strings, and an unknown solver stays an error.  (The GPU side of the same
strings is in tests/test_gpu_solver_options.py.)"""
import numpy as np

from oracle import oracle
from yelmo_amd import generator
from yelmo_amd.params import SsaParams


def _system():
    par = SsaParams(visc_method=2, beta_method=4, ssa_iter_max=2)
    p = generator.make_problem(33, 40e3, seed=3, dtype=np.float64)
    st, _ = oracle.calc_velocity_ssa(p, par, dp=True)
    q = p.copy()
    q.state.update({k: st[k] for k in ("ux_b", "uy_b")})
    return oracle.assemble(q, np.asfortranarray(st["beta_acx"]), np.asfortranarray(st["beta_acy"]),
                           np.asfortranarray(st["visc_eff_int"]))


def test_minres_and_bicgsafe_solve_like_bicgstab():
    ap, ai, av, b, x0, _ = _system()
    tail = " -p jacobi -maxiter 100 -tol 1.0e-2 -initx_zeros false"
    ref = oracle.csr_solve(ap, ai, av, b, x0, "-i bicgstab" + tail)
    assert ref["rc"] == 0 and ref["status"] == 0
    for name in ("minres", "bicgsafe"):
        out = oracle.csr_solve(ap, ai, av, b, x0, f"-i {name}" + tail)
        assert out["rc"] == 0
        assert (out["iters"], out["status"]) == (ref["iters"], ref["status"])
        assert np.array_equal(out["x"], ref["x"])


def test_unknown_solver_is_an_error():
    ap, ai, av, b, x0, _ = _system()
    out = oracle.csr_solve(ap, ai, av, b, x0, "-i gmres -p jacobi -maxiter 100 -tol 1.0e-2")
    assert out["rc"] != 0
