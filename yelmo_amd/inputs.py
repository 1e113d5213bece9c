"""Per-solve inputs of the velocity solve, computed on the host.

These routines prepare what Yelmo's ``calc_ydyn`` hands to
``calc_velocity_ssa`` (src/yelmo_dynamics.f90:105-165, 464): thickness-derived
geometry, ice-front codes, driving stress, lateral front stress and the SSA
face masks.  They sit upstream of the drop-in boundary, run once per solve and
are plain numpy.  Arrays are ``(nx, ny)`` in Fortran order, ``a[i, j]`` is
the reference's ``a(i+1, j+1)``.
"""
from __future__ import annotations

import numpy as np

# MASK_FRNT_* codes (src/yelmo_defs.f90:76-81)
MASK_FRNT_ICE_FREE = -1
MASK_FRNT_ICE_FREE_LAND = -2
MASK_FRNT_NONE = 0
MASK_FRNT_FLOAT = 1
MASK_FRNT_MARINE = 2
MASK_FRNT_GRND = 3

# boundaries strings in the order of the C-ABI enum YGPU_BND_*
BOUNDARIES = ("zeros", "infinite", "mask", "MISMIP3D", "TROUGH", "periodic",
              "periodic-x", "periodic-y")
LATERAL_BC = ("none", "floating", "marine", "all")


def _wrap(boundaries: str):
    """Which directions wrap in get_neighbor_indices_bc_codes
    (src/yelmo_tools.f90:118-201)."""
    wx = boundaries in ("periodic", "periodic-x")
    wy = boundaries in ("MISMIP3D", "TROUGH", "periodic", "periodic-y")
    return wx, wy


def neighbors(nx: int, ny: int, boundaries: str):
    """Index arrays im1, ip1 (length nx) and jm1, jp1 (length ny), 0-based,
    as get_neighbor_indices_bc_codes (src/yelmo_tools.f90:118-201)."""
    wx, wy = _wrap(boundaries)
    i = np.arange(nx)
    j = np.arange(ny)
    if wx:
        im1, ip1 = (i - 1) % nx, (i + 1) % nx
    else:
        im1, ip1 = np.maximum(i - 1, 0), np.minimum(i + 1, nx - 1)
    if wy:
        jm1, jp1 = (j - 1) % ny, (j + 1) % ny
    else:
        jm1, jp1 = np.maximum(j - 1, 0), np.minimum(j + 1, ny - 1)
    return im1, ip1, jm1, jp1


def calc_H_eff(H_ice, f_ice, set_frac_zero=True):
    """calc_H_eff (src/physics/topography.f90:851-874)."""
    H_eff = np.where(f_ice > 0, H_ice / np.where(f_ice > 0, f_ice, 1), H_ice)
    if set_frac_zero:
        H_eff = np.where(f_ice < 1.0, 0.0, H_eff)
    return H_eff.astype(H_ice.dtype)


def calc_H_grnd(H_ice, f_ice, z_bed, z_sl, rho_ice, rho_sw):
    """calc_H_grnd with use_f_ice (src/physics/topography.f90:876-931)."""
    rho_sw_ice = rho_sw / rho_ice
    H_eff = calc_H_eff(H_ice, f_ice)
    return np.where(z_sl - z_bed > 0, H_eff - rho_sw_ice * (z_sl - z_bed),
                    H_eff + (z_bed - z_sl)).astype(H_ice.dtype)


def calc_z_srf_max(H_ice, f_ice, z_bed, z_sl, rho_ice, rho_sw):
    """calc_z_srf_max (src/physics/topography.f90:820-849)."""
    H_eff = calc_H_eff(H_ice, f_ice)
    return np.maximum(z_bed + H_eff, z_sl + (1.0 - rho_ice / rho_sw) * H_eff)


def calc_ice_front(f_ice, f_grnd, z_bed, z_sl, boundaries="zeros"):
    """calc_ice_front (src/physics/topography.f90:724-818): front codes of
    ice cells with an open neighbour and of the ice-free cells beside them.
    The loop order matters (later cells overwrite the codes of neighbours),
    so it is kept serial as in the reference."""
    nx, ny = f_ice.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    m = np.zeros((nx, ny), dtype=np.int32, order="F")

    def free_code(i, j):
        return MASK_FRNT_ICE_FREE if z_bed[i, j] < z_sl[i, j] else MASK_FRNT_ICE_FREE_LAND

    for j in range(ny):
        for i in range(nx):
            nb = [(im1[i], j), (ip1[i], j), (i, jm1[j]), (i, jp1[j])]
            n_open = sum(1 for (a, b) in nb if f_ice[a, b] < 1.0)
            if f_ice[i, j] == 1.0 and n_open > 0:
                if f_grnd[i, j] > 0.0 and z_sl[i, j] <= z_bed[i, j]:
                    m[i, j] = MASK_FRNT_GRND
                elif f_grnd[i, j] > 0.0:
                    m[i, j] = MASK_FRNT_MARINE
                else:
                    m[i, j] = MASK_FRNT_FLOAT
                for (a, b) in nb:
                    if f_ice[a, b] < 1.0:
                        m[a, b] = free_code(a, b)
    return m


def calc_gradient_acx(var, f_ice, dx, grad_lim, zero_outside=False, boundaries="zeros"):
    """calc_gradient_acx (src/yelmo_tools.f90:341-408), without slope_bg."""
    nx, ny = var.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    V0 = var.copy()
    V1 = var[ip1, :].copy()
    if zero_outside:
        V0 = np.where(f_ice < 1.0, 0.0, V0)
        V1 = np.where(f_ice[ip1, :] < 1.0, 0.0, V1)
    d = ((V1 - V0) / dx).astype(var.dtype)
    if boundaries in ("infinite", "MISMIP3D", "TROUGH", "mask", "periodic-y"):
        d[0, :] = d[1, :]
        d[-1, :] = d[-2, :]
    return np.clip(d, -grad_lim, grad_lim)


def calc_gradient_acy(var, f_ice, dy, grad_lim, zero_outside=False, boundaries="zeros"):
    """calc_gradient_acy (src/yelmo_tools.f90:410-478), without slope_bg."""
    nx, ny = var.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    V0 = var.copy()
    V1 = var[:, jp1].copy()
    if zero_outside:
        V0 = np.where(f_ice < 1.0, 0.0, V0)
        V1 = np.where(f_ice[:, jp1] < 1.0, 0.0, V1)
    d = ((V1 - V0) / dy).astype(var.dtype)
    if boundaries in ("infinite", "mask", "periodic-x"):
        d[:, 0] = d[:, 1]
        d[:, -1] = d[:, -2]
    return np.clip(d, -grad_lim, grad_lim)


def calc_driving_stress(H_ice, f_ice, dzsdx, dzsdy, dx, taud_lim, rho_ice, g,
                        boundaries="zeros"):
    """calc_driving_stress (src/physics/velocity_general.f90:1118-1205):
    tau_d = rho g H_mid ds/dx with the margin-upstream thickness."""
    nx, ny = H_ice.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    rhog = rho_ice * g
    full = f_ice == 1.0

    def hmid(Ha, Hb, fa, fb):
        return np.where(fa & ~fb, 0.5 * Ha, np.where(~fa & fb, 0.5 * Hb, 0.5 * (Ha + Hb)))

    Hx = hmid(H_ice, H_ice[ip1, :], full, full[ip1, :])
    Hy = hmid(H_ice, H_ice[:, jp1], full, full[:, jp1])
    tx = np.clip(rhog * Hx * dzsdx, -taud_lim, taud_lim).astype(H_ice.dtype)
    ty = np.clip(rhog * Hy * dzsdy, -taud_lim, taud_lim).astype(H_ice.dtype)
    return np.asfortranarray(tx), np.asfortranarray(ty)


def calc_lateral_bc_stress(H_ice, z_srf, z_sl, rho_ice, rho_sw, g):
    """calc_lateral_bc_stress (src/physics/velocity_general.f90:1578-1609):
    0.5 rho_i g H^2 - 0.5 rho_sw g H_ocn^2."""
    f_sub = 1.0 - np.minimum((z_srf - z_sl) / H_ice, 1.0)
    H_ocn = H_ice * f_sub
    return 0.5 * rho_ice * g * H_ice ** 2 - 0.5 * rho_sw * g * H_ocn ** 2


def calc_lateral_bc_stress_2D(mask_frnt, H_ice, f_ice, z_srf, z_sl, rho_ice, rho_sw, g,
                              boundaries="zeros"):
    """calc_lateral_bc_stress_2D (src/physics/velocity_general.f90:1453-1576)."""
    nx, ny = H_ice.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    tx = np.zeros((nx, ny), dtype=H_ice.dtype, order="F")
    ty = np.zeros((nx, ny), dtype=H_ice.dtype, order="F")
    for j in range(ny):
        for i in range(nx):
            for (a, b, out) in ((ip1[i], j, tx), (i, jp1[j], ty)):
                ma, mb = mask_frnt[i, j], mask_frnt[a, b]
                if (ma > 0 and mb < 0) or (ma < 0 and mb > 0):
                    ii, jj = (a, b) if ma < 0 else (i, j)
                    H = H_ice[ii, jj]
                    zs = z_srf[ii, jj]
                    zsl = z_sl[ii, jj]
                    if ma == MASK_FRNT_ICE_FREE_LAND or mb == MASK_FRNT_ICE_FREE_LAND:
                        zsl = zs - H
                    out[i, j] = calc_lateral_bc_stress(np.float64(H), np.float64(zs),
                                                       np.float64(zsl), rho_ice, rho_sw, g)
    return tx, ty


def front_face_mask(code_a, code_b, lateral_bc):
    """front_face_mask (src/physics/solver_ssa_ac.f90:950-996)."""
    if code_a > 0 and code_b < 0:
        code_ice, code_free = code_a, code_b
    elif code_a < 0 and code_b > 0:
        code_ice, code_free = code_b, code_a
    else:
        return -1
    if code_free == MASK_FRNT_ICE_FREE_LAND:
        if code_ice == MASK_FRNT_FLOAT:
            return 0
        code_ice = MASK_FRNT_GRND
    if lateral_bc == "none":
        return 4
    if lateral_bc in ("floating", "float"):
        return 3 if code_ice == MASK_FRNT_FLOAT else 4
    if lateral_bc == "marine":
        return 3 if code_ice in (MASK_FRNT_FLOAT, MASK_FRNT_MARINE) else 4
    if lateral_bc == "all":
        return 3
    return -1


def set_ssa_masks(mask_frnt, f_ice, f_grnd, lateral_bc="all", boundaries="zeros"):
    """set_ssa_masks (src/physics/solver_ssa_ac.f90:851-948): per-face action
    codes -1..4 for the acx and acy faces."""
    nx, ny = f_ice.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    mx = np.zeros((nx, ny), dtype=np.int32, order="F")
    my = np.zeros((nx, ny), dtype=np.int32, order="F")
    for j in range(ny):
        for i in range(nx):
            for (a, b, out) in ((ip1[i], j, mx), (i, jp1[j], my)):
                if f_ice[i, j] == 1.0 or f_ice[a, b] == 1.0:
                    out[i, j] = 1 if (f_grnd[i, j] > 0.0 or f_grnd[a, b] > 0.0) else 2
                ml = front_face_mask(int(mask_frnt[i, j]), int(mask_frnt[a, b]), lateral_bc)
                if ml >= 0:
                    out[i, j] = ml
    return mx, my


def calc_zeta(nz_aa: int, zeta_scale: str = "exp", zeta_exp: float = 2.0, dtype=np.float32):
    """calc_zeta (src/yelmo_grid.f90:41-124), the aa levels only."""
    z = np.array([k / (nz_aa - 1) for k in range(nz_aa)], dtype=dtype)
    if zeta_scale == "linear":
        pass
    elif zeta_scale == "exp":
        z = (z ** dtype(zeta_exp)).astype(dtype)
    elif zeta_scale == "tanh":
        z = np.tanh(np.pi * (z - 0.5)).astype(dtype)
        z = z - z.min()
        z = z / z.max()
    else:
        raise ValueError(f"zeta_scale not recognized: {zeta_scale}")
    return z
