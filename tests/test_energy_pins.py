"""Pins of the oracle's energy-formulation assembly
(linear_solver_matrix_ssa_ac_csr_2D_energy, solver_ssa_ac_energy.f90:70-708).

The module header (:1-46) defines the matrix as the Hessian of the discrete
SSA energy

    E = sum_cells   N_aa (2 u_x^2 + 2 v_y^2 + 2 u_x v_y) dx dy
      + sum_corners 1/2 N_ab (u_y + v_x)^2 dx dy
      + sum_faces   1/2 w (beta + k) u^2 dx dy  -  sum_faces f u

with ghosts beyond the domain (periodic wrap, free-slip copy, zero) and the
reduced unknowns of Dirichlet and tied faces.  The test builds that Hessian
globally, as P^T (D^T W D) P from difference operators D over every velocity
and the map P from the free unknowns to velocities, and compares it with the
oracle's row-by-row assembly: a wrong element coefficient, sign, index,
ghost rule, drag weight or fold breaks the equality.  A second check ties
the inner rows to the (separately pinned) residual assembly,
K_inner = -A_residual dx dy (:36-42).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from yelmo_amd import generator
from yelmo_amd.params import SsaParams

SIDES = {  # 0 right, 1 upper, 2 left, 3 lower: 'p' periodic, 'f' free-slip, 'n' no-slip
    "zeros": "nnnn", "infinite": "ffff", "periodic": "pppp", "periodic-x": "pfpf",
    "periodic-y": "fpfp", "MISMIP3D": "fpnp",
}


def _problem(nx, ny, seed, bnd):
    """A cap with fronts and inner faces on a small grid, random positive
    viscosity and friction, a few prescribed (-1) faces and random velocity
    (the Dirichlet values)."""
    p = generator.make_problem(nx, 150e3, seed=seed, dtype=np.float64, ny=ny, boundaries=bnd)
    p.dy = 0.8 * p.dx      # dx != dy, so a swapped dx / dy shows
    rng = np.random.default_rng(seed)
    st, fl = p.state, p.fields
    st["ux_b"] = np.asfortranarray(rng.uniform(-50, 50, (nx, ny)))
    st["uy_b"] = np.asfortranarray(rng.uniform(-50, 50, (nx, ny)))
    mx, my = st["ssa_mask_acx"].copy(), st["ssa_mask_acy"].copy()
    # every mask code present: 1/2 inner, 3 front, 4 front as inner, 0, -1
    inner = np.argwhere(mx == 1)
    for k, code in zip(range(3), (4, -1, 4)):
        if len(inner) > 3 * k:
            mx[tuple(inner[3 * k])] = code
    inner = np.argwhere(my == 1)
    if len(inner) > 2:
        my[tuple(inner[1])] = -1
    st["ssa_mask_acx"], st["ssa_mask_acy"] = mx, my
    N = np.asfortranarray(10 ** rng.uniform(9, 11, (nx, ny)))
    bx = np.asfortranarray(10 ** rng.uniform(2, 4, (nx, ny)))
    by = np.asfortranarray(10 ** rng.uniform(2, 4, (nx, ny)))
    lim = [np.asfortranarray(rng.uniform(0, 50, (nx, ny))) for _ in range(2)] + \
          [np.asfortranarray(rng.uniform(-1e4, 1e4, (nx, ny))) for _ in range(2)]
    return p, N, bx, by, lim


def _classify(p, bnd):
    """Kind of every unknown (0 free, 1 Dirichlet, 2 tied), its root and its
    Dirichlet value, by the header's rules: mask 0 / -1 first, then the
    domain edges (ux: left, right, lower, upper; uy: lower, upper, left,
    right), tied on free-slip sides, Dirichlet on no-slip sides."""
    nx, ny = p.nx, p.ny
    s = SIDES[bnd]
    mx, my = p.state["ssa_mask_acx"], p.state["ssa_mask_acy"]
    ux, uy = p.state["ux_b"], p.state["uy_b"]
    q = lambda i, j, c: 2 * (i * ny + j) + c
    n = 2 * nx * ny
    kind, part, val = np.zeros(n, int), np.full(n, -1), np.zeros(n)
    for i in range(nx):
        for j in range(ny):
            for c, m, u in ((0, mx, ux), (1, my, uy)):
                k = q(i, j, c)
                if m[i, j] == 0:
                    kind[k] = 1
                    continue
                if m[i, j] == -1:
                    kind[k], val[k] = 1, u[i, j]
                    continue
                edges = [(i == 0, 2, q(1, j, c)), (i == nx - 1, 0, q(nx - 2, j, c)),
                         (j == 0, 3, q(i, 1, c)), (j == ny - 1, 1, q(i, ny - 2, c))]
                if c == 1:
                    edges = edges[2:] + edges[:2]
                for at, side, inner in edges:
                    if at and s[side] != "p":
                        if s[side] == "f":
                            kind[k], part[k] = 2, inner
                        else:
                            kind[k] = 1
                        break
    root = np.arange(n)
    for k in range(n):
        m = k
        while kind[m] == 2:
            m = part[m]
        root[k] = m
    for k in range(n):
        if kind[k] == 2 and kind[root[k]] == 1:
            kind[k], val[k], root[k] = 1, val[root[k]], k
    return kind, root, val


def _energy_system(p, bnd, N, N_ab, bx, by, lim):
    """Hessian and right-hand side of E over the free unknowns."""
    nx, ny, dx, dy = p.nx, p.ny, p.dx, p.dy
    s = SIDES[bnd]
    kind, root, val = _classify(p, bnd)
    free = np.flatnonzero(kind == 0)
    col = {k: a for a, k in enumerate(free)}
    nf = len(free)

    def vel(c, i, j):
        """(row of P, Dirichlet part) of velocity component c at (i, j),
        (i, j) possibly one beyond the domain."""
        for axis, n_, lo, hi in ((0, nx, 2, 0), (1, ny, 3, 1)):
            k = (i, j)[axis]
            if k < 0 or k >= n_:
                side = lo if k < 0 else hi
                if s[side] == "p":
                    k %= n_
                elif s[side] == "f":
                    k = 0 if k < 0 else n_ - 1
                else:
                    return np.zeros(nf), 0.0
                i, j = (k, j) if axis == 0 else (i, k)
        q = 2 * (i * ny + j) + c
        r = np.zeros(nf)
        if kind[q] == 1:
            return r, val[q]
        r[col[root[q]]] = 1.0
        return r, 0.0

    K = np.zeros((nf, nf))
    g = np.zeros(nf)        # gradient of E at v = 0

    def quad(rows, M, scale):
        """adds scale * (G v + g0)^T M (G v + g0), G = stacked rows"""
        nonlocal K, g
        G = np.array([r for r, _ in rows])
        g0 = np.array([c for _, c in rows])
        K += 2.0 * scale * G.T @ M @ G
        g[:] += 2.0 * scale * G.T @ (M @ g0)

    def diff(a, b, h):
        return (a[0] - b[0]) / h, (a[1] - b[1]) / h

    M_cell = np.array([[2.0, 1.0], [1.0, 2.0]])
    for ci in range(nx):
        for cj in range(ny):
            ux_x = diff(vel(0, ci, cj), vel(0, ci - 1, cj), dx)
            vy_y = diff(vel(1, ci, cj), vel(1, ci, cj - 1), dy)
            quad([ux_x, vy_y], M_cell, N[ci, cj] * dx * dy)
            u_y = diff(vel(0, ci, cj + 1), vel(0, ci, cj), dy)
            v_x = diff(vel(1, ci + 1, cj), vel(1, ci, cj), dx)
            quad([(u_y[0] + v_x[0], u_y[1] + v_x[1])], np.eye(1), 0.5 * N_ab[ci, cj] * dx * dy)
    f_ice = p.fields["f_ice"]
    for c, m, beta, k_, r_, tl, td in (
            (0, p.state["ssa_mask_acx"], bx, lim[0], lim[2], p.fields["taul_int_acx"],
             p.fields["taud_acx"]),
            (1, p.state["ssa_mask_acy"], by, lim[1], lim[3], p.fields["taul_int_acy"],
             p.fields["taud_acy"])):
        for i in range(nx):
            for j in range(ny):
                w = 0.5 if m[i, j] in (3, 4) else 1.0
                row = vel(c, i, j)
                quad([row], np.eye(1), 0.5 * w * (beta[i, j] + k_[i, j]) * dx * dy)
                if m[i, j] == 3:
                    ip, jp = ((i + 1) % nx, j) if c == 0 else (i, (j + 1) % ny)
                    sgn = 1.0 if (f_ice[i, j] == 1.0 and f_ice[ip, jp] < 1.0) else -1.0
                    f = sgn * tl[i, j] * (dy if c == 0 else dx)
                else:
                    f = -td[i, j] * dx * dy
                f += w * r_[i, j] * dx * dy
                g[:] -= f * row[0]
    return K, -g, free, kind, root, val


def _csr(ap, ai, av):
    n = len(ap) - 1
    return sp.csr_matrix((av.copy(), (ai - 1).copy(), (ap - 1).copy()), shape=(n, n)).toarray()


@pytest.mark.parametrize("bnd", list(SIDES))
def test_energy_matrix_is_the_hessian_of_the_energy(bnd):
    for seed in range(2):
        p, N, bx, by, lim = _problem(9, 8, seed, bnd)
        ap, ai, av, b, x0, cf = oracle.assemble_energy(p, bx, by, N, boundaries=bnd, lim=lim)
        _, _, _, _, _, N_ab = oracle.assemble(p, bx, by, N, boundaries=bnd, lim=lim)
        K, rhs, free, kind, root, val = _energy_system(p, bnd, N, N_ab, bx, by, lim)
        A = _csr(ap, ai, av)
        # free rows: the Hessian and the right-hand side of E
        Kf = A[np.ix_(free, free)]
        scale = np.abs(K).max()
        assert np.abs(Kf - K).max() <= 1e-12 * scale, bnd
        assert np.abs(A[free][:, np.flatnonzero(kind != 0)]).max() == 0.0
        assert np.allclose(b[free], rhs, rtol=1e-11, atol=1e-9 * np.abs(rhs).max())
        # symmetric
        assert np.abs(A - A.T).max() <= 1e-12 * scale
        # Dirichlet and tied rows: identity; tied take their root's solution
        for q in np.flatnonzero(kind != 0):
            assert A[q, q] == 1.0 and np.count_nonzero(A[q]) == 1
            if kind[q] == 1:
                assert b[q] == val[q] and cf[q] == -1
            else:
                assert b[q] == 0.0 and cf[q] == root[q]
        assert np.all(cf[free] == -1)
        if bnd == "infinite":
            assert np.any(kind == 2)
        if bnd in ("zeros", "periodic"):
            assert not np.any(kind == 2)


def test_energy_inner_rows_are_residual_rows_times_minus_dxdy():
    """K_inner = -A_residual dx dy and b_inner = -taud dx dy at inner faces
    whose row touches only free unknowns (solver_ssa_ac_energy.f90:36-42)."""
    p, N, bx, by, lim = _problem(11, 9, 5, "zeros")
    A_e = _csr(*oracle.assemble_energy(p, bx, by, N, lim=lim)[:3])
    ap, ai, av, b_r, _, _ = oracle.assemble(p, bx, by, N, lim=lim)
    A_r = _csr(ap, ai, av)
    b_e = oracle.assemble_energy(p, bx, by, N, lim=lim)[3]
    kind = _classify(p, "zeros")[0]
    mx, my = p.state["ssa_mask_acx"], p.state["ssa_mask_acy"]
    dxdy = p.dx * p.dy
    checked = 0
    for q in np.flatnonzero(kind == 0):
        n = q // 2
        i, j = divmod(n, p.ny)
        m = (mx if q % 2 == 0 else my)[i, j]
        cols = np.flatnonzero(A_r[q])
        if m not in (1, 2) or np.any(kind[cols] != 0):
            continue
        assert np.allclose(A_e[q], -A_r[q] * dxdy, rtol=1e-12, atol=1e-12 * np.abs(A_e[q]).max())
        assert np.isclose(b_e[q], -b_r[q] * dxdy, rtol=1e-12)
        checked += 1
    assert checked > 50


def test_energy_picard_solution_mirror_symmetric():
    """The Picard loop with the energy assembler and CG on the
    mirror-symmetric domain: ux antisymmetric, uy symmetric."""
    from tests.test_oracle_pins import _mirror_domain
    p = _mirror_domain()
    par = SsaParams(ssa_solver="energy", visc_method=2, beta_method=4, ssa_iter_max=30,
                    ssa_lis_opt="-i cg -p jacobi -maxiter 5000 -tol 1.0e-12")
    st, info = oracle.calc_velocity_ssa(p, par, dp=True)
    assert info["rc"] == 0
    ux, uy = st["ux_b"], st["uy_b"]
    scale = np.abs(ux).max()
    assert scale > 1.0
    assert np.abs(ux[:-1, :] + ux[-2::-1, :]).max() < 1e-8 * scale
    assert np.abs(uy - uy[::-1, :]).max() < 1e-8 * scale
