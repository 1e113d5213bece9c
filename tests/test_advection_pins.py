"""Pins of the oracle's implicit advection system (the impl-lis branch of
calc_advec2D, solver_advection.f90:105-124, assembled by
linear_solver_matrix_advection_csr_2D, :171-487), fixed by what the
upwind finite-volume scheme must satisfy rather than by its formula:
  - no velocity: identity rows, b = H + dt F;
  - mass conservation on a periodic domain: every face flux leaves one cell
    and enters its neighbour, so every column of A sums to 1 and the solved
    thickness has the mass of H + dt F;
  - uniform flow: a constant thickness is steady (rows sum to 1) and each
    row couples to its upstream neighbour only, with the CFL number
    c = u dt / dx on it;
  - border rows: zero (identity, b = 0), Neumann (H - H_inner = 0), fixed
    (identity, b = H_ref), and no inflow across an outflow border.
Planted mistakes in the oracle fail these checks."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle


def _dense(ap, ai, av):
    n = len(ap) - 1
    return sp.csr_matrix((av, ai - 1, ap - 1), shape=(n, n)).toarray()


def _case(nx=7, ny=6, seed=0, bnd="periodic", mask=None, u_scale=40.0):
    rng = np.random.default_rng(seed)
    H = rng.uniform(0, 500, (nx, ny))
    ux = rng.uniform(-u_scale, u_scale, (nx, ny))
    uy = rng.uniform(-u_scale, u_scale, (nx, ny))
    F = rng.uniform(-1, 1, (nx, ny))
    m = np.full((nx, ny), 2, dtype=np.int32) if mask is None else mask
    return H, ux, uy, F, m


def test_no_velocity_gives_identity():
    H, ux, uy, F, m = _case()
    ap, ai, av, b, x0 = oracle.assemble_advection(H, H, 0 * ux, 0 * uy, F, m, 1e3, 1e3, 2.0,
                                                  "periodic", dp=True)
    A = _dense(ap, ai, av)
    assert np.array_equal(A, np.eye(A.shape[0]))
    # scalar row order n = i ny + j
    assert np.allclose(b, (H + 2.0 * F).ravel(order="C"), rtol=1e-15)
    assert np.array_equal(x0, H.ravel(order="C"))


def _mass_error(dp=True):
    H, ux, uy, F, m = _case(seed=3)
    ap, ai, av, b, _ = oracle.assemble_advection(H, H, ux, uy, F, m, 1e3, 1.5e3, 5.0,
                                                  "periodic", dp=dp)
    A = _dense(ap, ai, av)
    colsum = A.sum(axis=0)
    x = np.linalg.solve(A, b)
    return max(np.abs(colsum - 1).max(), abs(x.sum() - b.sum()) / abs(b.sum()))


def test_periodic_domain_conserves_mass():
    assert _mass_error() <= 1e-12
    assert _mass_error(dp=False) <= 1e-6
    for mistake in ("adv_inflow", "adv_diag"):
        with oracle.planted(mistake):
            assert _mass_error() > 1e-4, mistake


def test_uniform_flow_is_upwind_with_the_cfl_number():
    nx, ny, dx, dy, dt = 8, 6, 1e3, 2e3, 4.0
    u, v = 30.0, -20.0
    H, _, _, F, m = _case(nx, ny)
    ap, ai, av, b, _ = oracle.assemble_advection(H, H, np.full((nx, ny), u), np.full((nx, ny), v),
                                                  F, m, dx, dy, dt, "periodic", dp=True)
    A = _dense(ap, ai, av)
    assert np.allclose(A.sum(axis=1), 1.0, rtol=0, atol=1e-14)
    cx, cy = u * dt / dx, -v * dt / dy
    n = lambda i, j: (i % nx) * ny + (j % ny)
    for i in range(nx):
        for j in range(ny):
            r = n(i, j)
            assert A[r, r] == pytest.approx(1 + cx + cy, rel=1e-14)
            assert A[r, n(i - 1, j)] == pytest.approx(-cx, rel=1e-14)   # upstream in x
            assert A[r, n(i, j + 1)] == pytest.approx(-cy, rel=1e-14)   # upstream in y (v < 0)
            assert A[r, n(i + 1, j)] == 0.0 and A[r, n(i, j - 1)] == 0.0
    with oracle.planted("adv_inflow"):
        ap, ai, av, _, _ = oracle.assemble_advection(H, H, np.full((nx, ny), u),
                                                      np.full((nx, ny), v), F, m, dx, dy, dt,
                                                      "periodic", dp=True)
        bad = _dense(ap, ai, av)
        assert np.allclose(bad.sum(axis=1), 1.0)   # a uniform flow does not show it ...
    assert _mass_error() <= 1e-12                  # ... the mass balance does (above)


def test_border_rows():
    nx, ny = 6, 5
    H, ux, uy, F, m = _case(nx, ny)
    m = m.copy()
    m[2, 2] = 1                  # fixed cell
    Href = H + 7.0
    # zeros: identity border rows with b = 0; fixed cell: identity, b = H_ref
    ap, ai, av, b, x0 = oracle.assemble_advection(H, Href, ux, uy, F, m, 1e3, 1e3, 1.0, "zeros",
                                                  dp=True)
    A = _dense(ap, ai, av)
    idx = lambda i, j: i * ny + j
    for (i, j) in [(0, 2), (nx - 1, 3), (3, 0), (2, ny - 1)]:
        r = idx(i, j)
        assert A[r, r] == 1.0 and np.count_nonzero(A[r]) == 1 and b[r] == 0.0
    r = idx(2, 2)
    assert A[r, r] == 1.0 and np.count_nonzero(A[r]) == 1 and b[r] == Href[2, 2] == x0[r]
    # mask: Neumann rows H(border) - H(inner) = 0
    ap, ai, av, b, x0 = oracle.assemble_advection(H, H, ux, uy, F, m, 1e3, 1e3, 1.0, "mask",
                                                  dp=True)
    A = _dense(ap, ai, av)
    r = idx(nx - 1, 3)
    assert A[r, r] == 1.0 and A[r, idx(nx - 2, 3)] == -1.0 and b[r] == 0.0
    r = idx(3, 0)
    assert A[r, r] == 1.0 and A[r, idx(3, 1)] == -1.0
    # infinite: outflow borders are inner rows with no inflow from the border cell
    ux_in = np.full((nx, ny), 25.0)          # flow to the right everywhere
    ap, ai, av, b, x0 = oracle.assemble_advection(H, H, ux_in, 0 * uy, F, m, 1e3, 1e3, 1.0,
                                                  "infinite", dp=True)
    A = _dense(ap, ai, av)
    r = idx(1, 2)   # its left face: ux(0) > 0 from the border cell is cut
    assert A[r, idx(0, 2)] == 0.0
    r = idx(nx - 1, 2)   # the right border cell receives its inner neighbour's outflow
    assert A[r, idx(nx - 2, 2)] == pytest.approx(-25.0 * 1.0 / 1e3, rel=1e-14)


def test_oracle_bicgstab_jacobi_solves_the_advection_system():
    """The reference's solver settings (solver_advection.f90:119) on the
    oracle's Krylov solve reach the sparse direct solution."""
    H, ux, uy, F, m = _case(40, 30, seed=5, u_scale=300.0)
    ap, ai, av, b, x0 = oracle.assemble_advection(H, H, ux, uy, F, m, 2e3, 2e3, 10.0, "zeros",
                                                  dp=True)
    r = oracle.csr_solve(ap, ai, av, b, x0, "-i bicgstab -p jacobi -maxiter 1000 -tol 1.0e-12 "
                                            "-initx_zeros false")
    n = len(b)
    A = sp.csr_matrix((av, ai - 1, ap - 1), shape=(n, n))
    xd = spla.spsolve(A.tocsc(), b)
    assert r["status"] == 0 and r["iters"] > 1
    assert np.linalg.norm(r["x"] - xd) / np.linalg.norm(xd) <= 1e-10
