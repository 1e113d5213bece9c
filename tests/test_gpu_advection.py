"""The implicit thickness advection solve (calc_advec2D "impl-lis",
solver_advection.f90:105-124) through ygpu_csr_solve: the reference's
five-point CSR system (assembled by the oracle's restatement of
linear_solver_matrix_advection_csr_2D) with the reference's option string,
against the oracle's BiCGSTAB + Jacobi on the same arrays.

Bars: status equal, iterations equal +-1, the solutions within 1e-9
relative L2 (both solves at rtol 1e-12); on a periodic domain of dynamic
cells the GPU solution keeps the mass of H + dt F to 1e-10."""
import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from oracle import oracle
from tests._helpers import rel_l2
from yelmo_amd import velocity

pytestmark = pytest.mark.gpu

OPTS = "-i bicgstab -p jacobi -maxiter 1000 -tol 1.0e-12 -initx_zeros false"


def _fields(nx, ny, seed, umax):
    rng = np.random.default_rng(seed)
    sm = lambda a: gaussian_filter(a, 3.0, mode="wrap")
    H = np.maximum(sm(rng.uniform(-500, 2500, (nx, ny))), 0.0)
    ux = sm(rng.standard_normal((nx, ny)))
    uy = sm(rng.standard_normal((nx, ny)))
    s = umax / max(np.abs(ux).max(), np.abs(uy).max())
    F = sm(rng.uniform(-2, 2, (nx, ny)))
    m = np.full((nx, ny), 2, dtype=np.int32)
    m[H == 0] = 0                       # ice-free cells: MASK_ICE_NONE
    m[nx // 3, ny // 2] = 1             # one prescribed cell
    return H, ux * s, uy * s, F, m


CASES = [(41, 37, "zeros", 500.0, 8e3, 1.0), (41, 37, "periodic", 500.0, 8e3, 1.0),
         (48, 40, "MISMIP3D", 2e3, 4e3, 2.0), (48, 40, "infinite", 2e3, 4e3, 2.0),
         (48, 40, "mask", 2e3, 4e3, 2.0), (96, 80, "periodic-x", 1e4, 2e3, 5.0)]


@pytest.mark.parametrize("nx,ny,bnd,umax,dx,dt", CASES)
def test_advection_solve_matches_oracle(nx, ny, bnd, umax, dx, dt):
    print(f"case advection {nx}x{ny} {bnd} cfl {umax * dt / dx:.2f}", flush=True)
    H, ux, uy, F, m = _fields(nx, ny, nx + ny, umax)
    ap, ai, av, b, x0 = oracle.assemble_advection(H, H, ux, uy, F, m, dx, dx, dt, bnd)
    ref = oracle.csr_solve(ap, ai, av, b, x0, OPTS)
    lgs = velocity.Lgs(ap, ai, av, b, x0.copy())
    velocity.linear_solver_matrix_solve(lgs, OPTS)
    err = rel_l2(lgs.x_value, ref["x"])
    print("iters", lgs.lin_iter, ref["iters"], "status", lgs.status, ref["status"], "resid",
          lgs.L2_rel_norm, ref["resid"], "err", err, flush=True)
    assert lgs.status == ref["status"] == 0
    assert abs(lgs.lin_iter - ref["iters"]) <= 1
    assert err <= 1e-9


@pytest.mark.parametrize("opts", [OPTS.replace("false", "true"),
                                  OPTS.replace("jacobi", "none"),
                                  OPTS.replace("1000", "3")])
def test_advection_option_variants(opts):
    print("case advection options", opts, flush=True)
    H, ux, uy, F, m = _fields(41, 37, 7, 2e3)
    ap, ai, av, b, x0 = oracle.assemble_advection(H, H, ux, uy, F, m, 4e3, 4e3, 2.0, "zeros")
    ref = oracle.csr_solve(ap, ai, av, b, x0, opts)
    lgs = velocity.Lgs(ap, ai, av, b, x0.copy())
    velocity.linear_solver_matrix_solve(lgs, opts)
    print("iters", lgs.lin_iter, ref["iters"], "status", lgs.status, ref["status"], flush=True)
    assert lgs.status == ref["status"]
    assert abs(lgs.lin_iter - ref["iters"]) <= 1
    if ref["status"] == 0:
        assert rel_l2(lgs.x_value, ref["x"]) <= 1e-9
    else:   # the iteration limit: the same iterate to rounding
        assert rel_l2(lgs.x_value, ref["x"]) <= 1e-9


def test_advection_641_matches_oracle():
    """The largest served grid (641 x 641, 410,881 unknowns), periodic, every
    cell dynamic and the system assembled in FP64: the solve conserves the
    mass of H + dt F (every column of the matrix sums to one)."""
    print("case advection 641", flush=True)
    H, ux, uy, F, m = _fields(641, 641, 3, 3e3)
    m[:] = 2
    ap, ai, av, b, x0 = oracle.assemble_advection(H, H, ux, uy, F, m, 2.4e3, 2.4e3, 1.0,
                                                  "periodic", dp=True)
    ref = oracle.csr_solve(ap, ai, av, b, x0, OPTS)
    lgs = velocity.Lgs(ap, ai, av, b, x0.copy())
    velocity.linear_solver_matrix_solve(lgs, OPTS)
    err = rel_l2(lgs.x_value, ref["x"])
    print("iters", lgs.lin_iter, ref["iters"], "err", err, "gpu s", lgs.solver_time,
          "oracle s", ref.get("seconds"), flush=True)
    assert lgs.status == ref["status"] == 0
    assert abs(lgs.lin_iter - ref["iters"]) <= 1
    assert err <= 1e-9
    assert abs(lgs.x_value.sum() - b.sum()) <= 1e-10 * abs(b.sum())
