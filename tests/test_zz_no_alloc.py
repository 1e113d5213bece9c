"""ygpu_csr_solve_noalloc (the entry point that allocates no device memory:
its buffers come from one static per-device scratch array) against
ygpu_csr_solve on the same CSR systems, called through the C-ABI: the same
kernels on the same inputs give the same bits.  Two calls in a row, the
second with a smaller system, check that the shared scratch keeps nothing of
the previous call that matters; a system above the limit is rejected."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle
from yelmo_amd import _lib, generator
from yelmo_amd.params import SsaParams

pytestmark = pytest.mark.gpu

BASE = "-i bicgstab -p ilu -maxiter 1000 -tol 1.0e-8 -initx_zeros false"


def _system(n, dx, seed):
    """The reference CSR of a mid-Picard state of an n x n domain."""
    par = SsaParams(visc_method=2, beta_method=4, ssa_iter_max=3)
    p = generator.make_problem(n, dx, seed=seed, dtype=np.float64)
    st, _ = oracle.calc_velocity_ssa(p, par, dp=True)
    q = p.copy()
    q.state.update({k: st[k] for k in ("ux_b", "uy_b")})
    ap, ai, av, b, x0, _ = oracle.assemble(q, np.asfortranarray(st["beta_acx"]),
                                           np.asfortranarray(st["beta_acy"]),
                                           np.asfortranarray(st["visc_eff_int"]))
    return ap, ai, av, b, x0


def _solve(fn, h, sysm, opts):
    ap, ai, av, b, x0 = [np.ascontiguousarray(a) for a in sysm]
    ap = ap.astype(np.int32)
    ai = ai.astype(np.int32)
    n = len(b)
    x = x0.astype(np.float64).copy()
    it, stt = C.c_int32(), C.c_int32()
    res, secs = C.c_double(), C.c_double()
    rc = fn(h.h, n, int(ap[n] - 1), ap.ctypes.data, ai.ctypes.data, av.ctypes.data,
            b.ctypes.data, x.ctypes.data, opts.encode(), C.byref(it), C.byref(stt), C.byref(res),
            C.byref(secs))
    return rc, x, it.value, stt.value, res.value


@pytest.mark.parametrize("opts", [BASE, BASE + " -ygpu_precision mixed",
                                  BASE.replace("ilu", "jacobi").replace("1000", "3000")])
def test_noalloc_equals_csr_solve_twice_smaller_second(opts):
    print("case noalloc", opts, flush=True)
    h = _lib.Handle(0)
    for n, dx, seed in ((81, 16e3, 0), (41, 32e3, 1)):
        sysm = _system(n, dx, seed)
        rc0, x0, it0, st0, r0 = _solve(h.lib.ygpu_csr_solve, h, sysm, opts)
        rc1, x1, it1, st1, r1 = _solve(h.lib.ygpu_csr_solve_noalloc, h, sysm, opts)
        print(n, "iters", it0, it1, "status", st0, st1, "resid", r0, r1, flush=True)
        assert rc0 == rc1 == 0
        assert st0 == st1 == 0 and it0 == it1
        assert np.array_equal(x0, x1) and r0 == r1
    h.close()


def test_noalloc_rejects_systems_above_the_limit():
    h = _lib.Handle(0)
    cells = 32768 + 2
    n = 2 * cells
    ap = np.arange(1, n + 2, dtype=np.int32)
    ai = np.arange(1, n + 1, dtype=np.int32)
    av = np.ones(n)
    b = np.ones(n)
    rc, _, _, _, _ = _solve(h.lib.ygpu_csr_solve_noalloc, h, (ap, ai, av, b, np.zeros(n)), BASE)
    assert rc == _lib.YGPU_ERR_ARG
    assert "scratch" in h.lib.ygpu_last_error(h.h).decode()
    h.close()
