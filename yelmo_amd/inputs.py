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
    The reference loops serially and later cells overwrite neighbour codes,
    but every write depends only on the written cell (an ice cell only ever
    receives its own code, an open cell only its free code), so the result
    does not depend on loop order and is formed array-wise."""
    nx, ny = f_ice.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    full = f_ice == 1.0
    open_ = f_ice < 1.0
    n_open = (open_[im1, :].astype(np.int32) + open_[ip1, :] + open_[:, jm1] + open_[:, jp1])
    front = full & (n_open > 0)
    grnd = f_grnd > 0.0
    code = np.where(grnd & (z_sl <= z_bed), MASK_FRNT_GRND,
                    np.where(grnd, MASK_FRNT_MARINE, MASK_FRNT_FLOAT))
    m = np.where(front, code, 0).astype(np.int32)
    # open neighbours of front cells, scattered through the same index maps
    near = np.zeros((nx, ny), dtype=bool)
    fi_, fj_ = np.nonzero(front)
    near[im1[fi_], fj_] = True
    near[ip1[fi_], fj_] = True
    near[fi_, jm1[fj_]] = True
    near[fi_, jp1[fj_]] = True
    free = np.where(z_bed < z_sl, MASK_FRNT_ICE_FREE, MASK_FRNT_ICE_FREE_LAND)
    m = np.where(near & open_, free, m)
    return np.asfortranarray(m.astype(np.int32))


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
    """calc_lateral_bc_stress_2D (src/physics/velocity_general.f90:1453-1576):
    on each face between a front ice cell and an ice-free cell, the front
    stress of the ice cell (the ocean level replaced by the ice base against
    ice-free land).  Faces are independent; evaluated in float64 and stored
    in the array dtype."""
    nx, ny = H_ice.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    H64 = np.asarray(H_ice, dtype=np.float64)
    zs64 = np.asarray(z_srf, dtype=np.float64)
    zsl64 = np.asarray(z_sl, dtype=np.float64)
    out = []
    for sh in ((ip1, slice(None)), (slice(None), jp1)):
        ma, mb = mask_frnt, mask_frnt[sh]
        sel = ((ma > 0) & (mb < 0)) | ((ma < 0) & (mb > 0))
        take_b = ma < 0
        H = np.where(take_b, H64[sh], H64)
        zs = np.where(take_b, zs64[sh], zs64)
        zsl = np.where(take_b, zsl64[sh], zsl64)
        land = (ma == MASK_FRNT_ICE_FREE_LAND) | (mb == MASK_FRNT_ICE_FREE_LAND)
        zsl = np.where(land, zs - H, zsl)
        t = np.zeros((nx, ny), dtype=H_ice.dtype, order="F")
        if sel.any():
            t[sel] = calc_lateral_bc_stress(H[sel], zs[sel], zsl[sel], rho_ice, rho_sw, g)
        out.append(t)
    return out[0], out[1]


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


def _front_face_mask_array(code_a, code_b, lateral_bc):
    """front_face_mask over arrays of face codes (same table as below)."""
    ia = (code_a > 0) & (code_b < 0)
    ai = (code_a < 0) & (code_b > 0)
    code_ice = np.where(ia, code_a, code_b)
    code_free = np.where(ia, code_b, code_a)
    land = code_free == MASK_FRNT_ICE_FREE_LAND
    wall = land & (code_ice == MASK_FRNT_FLOAT)
    code_ice = np.where(land, MASK_FRNT_GRND, code_ice)
    if lateral_bc == "none":
        r = np.full(code_ice.shape, 4)
    elif lateral_bc in ("floating", "float"):
        r = np.where(code_ice == MASK_FRNT_FLOAT, 3, 4)
    elif lateral_bc == "marine":
        r = np.where((code_ice == MASK_FRNT_FLOAT) | (code_ice == MASK_FRNT_MARINE), 3, 4)
    elif lateral_bc == "all":
        r = np.full(code_ice.shape, 3)
    else:
        r = np.full(code_ice.shape, -1)
    r = np.where(wall, 0, r)
    return np.where(ia | ai, r, -1)


def set_ssa_masks(mask_frnt, f_ice, f_grnd, lateral_bc="all", boundaries="zeros"):
    """set_ssa_masks (src/physics/solver_ssa_ac.f90:851-948): per-face action
    codes -1..4 for the acx and acy faces (faces are independent)."""
    nx, ny = f_ice.shape
    im1, ip1, jm1, jp1 = neighbors(nx, ny, boundaries)
    full = f_ice == 1.0
    grnd = f_grnd > 0.0
    mf = np.asarray(mask_frnt, dtype=np.int64)
    out = []
    for sh in ((ip1, slice(None)), (slice(None), jp1)):
        m = np.where(full | full[sh], np.where(grnd | grnd[sh], 1, 2), 0)
        ml = _front_face_mask_array(mf, mf[sh], lateral_bc)
        m = np.where(ml >= 0, ml, m)
        out.append(np.asfortranarray(m.astype(np.int32)))
    return out[0], out[1]


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
