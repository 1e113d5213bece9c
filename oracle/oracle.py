"""ctypes driver of the CPU oracle (oracle/ssa_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker -- never by the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from yelmo_amd.generator import DIVA_STATE_2D, DIVA_STATE_3D, add_diva_state as diva_state

HERE = Path(__file__).resolve().parent
_libs = {}


def build():
    r = subprocess.run(["make", "-C", str(HERE), "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib(dp: bool) -> C.CDLL:
    key = "dp" if dp else "sp"
    if key not in _libs:
        path = HERE / f"liboracle_{key}.so"
        if not path.exists():
            build()
        _libs[key] = C.CDLL(str(path))
        assert _libs[key].oracle_wp_bytes() == (8 if dp else 4)
    return _libs[key]


class OracleParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "ssa_solver", "boundaries", "visc_method", "beta_method", "beta_gl_scale",
        "beta_gl_stag", "vel_lim_method", "ssa_iter_max", "ssa_write_log", "pad0")]
    _fields_ += [(n, C.c_double) for n in (
        "visc_const", "beta_const", "beta_q", "beta_u0", "beta_gl_f", "H_grnd_lim", "beta_min",
        "eps_0", "ssa_vel_lim_tau", "ssa_vel_max", "ssa_iter_rel", "ssa_iter_conv", "rho_ice",
        "rho_sw", "g")]


BND = {"zeros": 0, "infinite": 1, "mask": 2, "MISMIP3D": 3, "TROUGH": 4, "periodic": 5,
       "periodic-x": 6, "periodic-y": 7}


def params_c(par) -> OracleParams:
    c = OracleParams()
    c.ssa_solver = {"residual": 0, "energy": 1}[par.ssa_solver]
    c.boundaries = BND[par.boundaries]
    c.visc_method = par.visc_method
    c.beta_method = par.beta_method
    c.beta_gl_scale = par.beta_gl_scale
    c.beta_gl_stag = par.beta_gl_stag
    c.vel_lim_method = {"clip": 0, "drag": 1}[par.ssa_vel_lim_method]
    c.ssa_iter_max = par.ssa_iter_max
    c.ssa_write_log = int(par.ssa_write_log)
    for n in ("visc_const", "beta_const", "beta_q", "beta_u0", "beta_gl_f", "H_grnd_lim",
              "beta_min", "eps_0", "ssa_vel_lim_tau", "ssa_vel_max", "ssa_iter_rel",
              "ssa_iter_conv", "rho_ice", "rho_sw", "g"):
        setattr(c, n, float(getattr(par, n)))
    return c


def _p(a):
    return C.c_void_p(a.ctypes.data)


def calc_velocity_ssa(prob, par, dp: bool):
    """Runs the oracle Picard loop on a copy of prob (fields cast to the
    build's precision).  Returns (state dict, info dict)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    P = prob.astype(dt)
    st, fl = P.state, P.fields
    out_i = np.zeros(4, dtype=np.int32)
    rh = np.zeros(par.ssa_iter_max, dtype=np.float64)
    lh = np.zeros(par.ssa_iter_max, dtype=np.int32)
    secs = C.c_double()
    W = C.c_double if dp else C.c_float
    L.oracle_calc_velocity_ssa.restype = C.c_int
    args = [st["ux_b"], st["uy_b"], st["taub_acx"], st["taub_acy"], st["visc_eff"],
            st["visc_eff_int"], st["ssa_mask_acx"], st["ssa_mask_acy"], st["ssa_err_acx"],
            st["ssa_err_acy"], out_i, st["beta"], st["beta_acx"], st["beta_acy"],
            fl["c_bed"], fl["f_slide"], fl["taud_acx"], fl["taud_acy"], fl["taul_int_acx"],
            fl["taul_int_acy"], fl["H_ice"], fl["f_ice"], fl["H_grnd"], fl["f_grnd"],
            fl["f_grnd_acx"], fl["f_grnd_acy"], fl["ATT"], fl["zeta_aa"], fl["z_sl"],
            fl["z_bed"], fl["z_srf"]]
    pc = params_c(par)
    rc = L.oracle_calc_velocity_ssa(
        *[_p(a) for a in args], W(P.dx), W(P.dy), W(P.n_glen), C.byref(pc),
        par.ssa_lis_opt.encode(), P.nx, P.ny, P.nz, _p(rh), _p(lh), C.byref(secs))
    info = dict(rc=rc, ssa_iter_now=int(out_i[0]), ssa_lin_iter=int(out_i[1]),
                ssa_lin_fail=int(out_i[2]), ssa_lim_n=int(out_i[3]),
                resid_hist=rh[: max(out_i[0], 0)].copy(), lin_hist=lh[: max(out_i[0], 0)].copy(),
                solve_seconds=secs.value)
    return st, info




# neff_hook of calc_velocity_diva: (ctx, nx, ny, real_bytes, c_bed, ux_b, uy_b)
NEFF_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                        C.c_void_p, C.c_void_p)


def calc_velocity_diva(prob, par, dp: bool, neff_hook=None):
    """The oracle's DIVA Picard loop (calc_velocity_diva) on a copy of prob.
    neff_hook (a NEFF_HOOK) updates c_bed from u_b after every iteration;
    the updated c_bed is returned in the state dict as "c_bed".
    Returns (state dict, info dict)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    P = diva_state(prob.astype(dt))
    st, fl = P.state, P.fields
    for n in DIVA_STATE_3D + DIVA_STATE_2D:
        st[n] = np.asfortranarray(st[n], dtype=dt)
    out_i = np.zeros(4, dtype=np.int32)
    rh = np.zeros(par.ssa_iter_max, dtype=np.float64)
    lh = np.zeros(par.ssa_iter_max, dtype=np.int32)
    secs = C.c_double()
    W = C.c_double if dp else C.c_float
    L.oracle_calc_velocity_diva.restype = C.c_int
    args = [st["ux"], st["uy"], st["ux_bar"], st["uy_bar"], st["ux_b"], st["uy_b"], st["ux_i"],
            st["uy_i"], st["taub_acx"], st["taub_acy"], st["beta"], st["beta_acx"],
            st["beta_acy"], st["beta_eff"], st["visc_eff"], st["visc_eff_int"], st["duxdz"],
            st["duydz"], st["ssa_mask_acx"], st["ssa_mask_acy"], st["ssa_err_acx"],
            st["ssa_err_acy"], out_i, fl["c_bed"], fl["f_slide"], fl["taud_acx"], fl["taud_acy"],
            fl["taul_int_acx"], fl["taul_int_acy"], fl["H_ice"], fl["f_ice"], fl["H_grnd"],
            fl["f_grnd"], fl["f_grnd_acx"], fl["f_grnd_acy"], fl["ATT"], fl["zeta_aa"],
            fl["z_sl"], fl["z_bed"], fl["z_srf"]]
    pc = params_c(par)
    rc = L.oracle_calc_velocity_diva(
        *[_p(a) for a in args], W(P.dx), W(P.dy), W(P.n_glen), C.byref(pc), int(par.no_slip),
        par.ssa_lis_opt.encode(), P.nx, P.ny, P.nz, _p(rh), _p(lh), C.byref(secs),
        neff_hook if neff_hook is not None else NEFF_HOOK(), None)
    st["c_bed"] = fl["c_bed"]
    info = dict(rc=rc, ssa_iter_now=int(out_i[0]), ssa_lin_iter=int(out_i[1]),
                ssa_lin_fail=int(out_i[2]), ssa_lim_n=int(out_i[3]),
                resid_hist=rh[: max(out_i[0], 0)].copy(), lin_hist=lh[: max(out_i[0], 0)].copy(),
                solve_seconds=secs.value)
    return st, info


def F_integral(visc, H_ice, f_ice, zeta, dp=True):
    """calc_F_integral with n = 2 on given fields."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny, nz = visc.shape
    keep = [np.asfortranarray(a, dtype=dt) for a in (visc, H_ice, f_ice, zeta)]
    F = np.zeros((nx, ny), dtype=dt, order="F")
    L.oracle_F_integral(_p(F), *[_p(a) for a in keep], nx, ny, nz)
    return F


def beta_eff(beta, F2, no_slip=False, dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    keep = [np.asfortranarray(beta, dtype=dt), np.asfortranarray(F2, dtype=dt)]
    out = np.zeros_like(keep[0])
    L.oracle_beta_eff(_p(out), _p(keep[0]), _p(keep[1]), int(no_slip), out.size)
    return out


def vel_horizontal_3D(ux_b, uy_b, taub_acx, taub_acy, visc, H_ice, f_ice, zeta,
                      boundaries="zeros", dp=True):
    """calc_vel_horizontal_3D on given fields: (ux, uy) of shape (nx, ny, nz)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny, nz = visc.shape
    keep = [np.asfortranarray(a, dtype=dt) for a in
            (ux_b, uy_b, taub_acx, taub_acy, visc, H_ice, f_ice, zeta)]
    ux = np.zeros((nx, ny, nz), dtype=dt, order="F")
    uy = np.zeros((nx, ny, nz), dtype=dt, order="F")
    L.oracle_vel_horizontal_3D(_p(ux), _p(uy), *[_p(a) for a in keep], nx, ny, nz,
                               BND[boundaries])
    return ux, uy


def csr_solve(a_ptr1, a_idx1, a_val, b, x0, opts: str):
    L = lib(True)
    n = len(b)
    ap = np.ascontiguousarray(a_ptr1, dtype=np.int32)
    ai = np.ascontiguousarray(a_idx1, dtype=np.int32)
    av = np.ascontiguousarray(a_val, dtype=np.float64)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    it, st = C.c_int32(), C.c_int32()
    res, secs = C.c_double(), C.c_double()
    L.oracle_csr_solve.restype = C.c_int
    rc = L.oracle_csr_solve(n, _p(ap), _p(ai), _p(av), _p(bb), _p(x), opts.encode(), C.byref(it),
                            C.byref(st), C.byref(res), C.byref(secs))
    return dict(rc=rc, x=x, iters=it.value, status=st.value, resid=res.value, seconds=secs.value)


def ilu0_apply(a_ptr1, a_idx1, a_val, v, nb=1):
    L = lib(True)
    n = len(v)
    z = np.zeros(n)
    keep = [np.ascontiguousarray(a_ptr1, dtype=np.int32),
            np.ascontiguousarray(a_idx1, dtype=np.int32),
            np.ascontiguousarray(a_val, dtype=np.float64), np.ascontiguousarray(v, dtype=np.float64)]
    rc = L.oracle_ilu0_apply(n, *[_p(a) for a in keep[:3]], nb, _p(keep[3]), _p(z))
    assert rc == 0, rc
    return z


def assemble(prob, beta_acx, beta_acy, N_aa, boundaries="zeros", lim=None, dp=True):
    """The oracle's residual assembly on given fields; returns 1-based CSR
    (a_ptr, a_index, a_value, b, x0) and N_ab."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny = prob.nx, prob.ny
    nc = nx * ny
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=dt))

    class Lgs(C.Structure):
        _fields_ = [("nmax", C.c_int), ("a_ptr", C.c_void_p), ("a_index", C.c_void_p),
                    ("a_value", C.c_void_p), ("b_value", C.c_void_p), ("x_value", C.c_void_p)]

    ap = np.zeros(2 * nc + 1, dtype=np.int32)
    ai = np.zeros(18 * nc, dtype=np.int32)
    av = np.zeros(18 * nc)
    bv = np.zeros(2 * nc)
    xv = np.zeros(2 * nc)
    lgs = Lgs(2 * nc, ap.ctypes.data, ai.ctypes.data, av.ctypes.data, bv.ctypes.data,
              xv.ctypes.data)
    st, fl = prob.state, prob.fields
    args = [F(st["ux_b"]), F(st["uy_b"]), F(beta_acx), F(beta_acy), F(N_aa)]
    masks = [np.asfortranarray(st["ssa_mask_acx"], dtype=np.int32),
             np.asfortranarray(st["ssa_mask_acy"], dtype=np.int32)]
    rest = [F(fl["H_ice"]), F(fl["f_ice"]), F(fl["z_bed"]), F(fl["z_sl"]), F(fl["taud_acx"]),
            F(fl["taud_acy"]), F(fl["taul_int_acx"]), F(fl["taul_int_acy"])]
    W = C.c_double if dp else C.c_float
    Nab = np.zeros((nx, ny), dtype=dt, order="F")
    if lim is None:
        limp = [None] * 4
    else:
        limp = [F(a) for a in lim]
    L.oracle_assemble_ssa.restype = C.c_int
    k = L.oracle_assemble_ssa(C.byref(lgs), *[_p(a) for a in args], *[_p(a) for a in masks],
                              *[_p(a) for a in rest], W(prob.dx), W(prob.dy), BND[boundaries],
                              *[(_p(a) if a is not None else None) for a in limp], _p(Nab), nx, ny)
    return ap, ai[:k], av[:k], bv, xv, Nab


def assemble_energy(prob, beta_acx, beta_acy, N_aa, boundaries="zeros", lim=None, dp=True):
    """The oracle's energy-formulation assembly (solver_ssa_ac_energy.f90)
    on given fields; returns 1-based CSR (a_ptr, a_index, a_value, b, x0)
    and copy_from (0-based root of each tied unknown, -1 elsewhere)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny = prob.nx, prob.ny
    nc = nx * ny
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=dt))

    class Lgs(C.Structure):
        _fields_ = [("nmax", C.c_int), ("a_ptr", C.c_void_p), ("a_index", C.c_void_p),
                    ("a_value", C.c_void_p), ("b_value", C.c_void_p), ("x_value", C.c_void_p)]

    cap = 48 * 2 * nc
    ap = np.zeros(2 * nc + 1, dtype=np.int32)
    ai = np.zeros(cap, dtype=np.int32)
    av = np.zeros(cap)
    bv = np.zeros(2 * nc)
    xv = np.zeros(2 * nc)
    cf = np.zeros(2 * nc, dtype=np.int32)
    lgs = Lgs(2 * nc, ap.ctypes.data, ai.ctypes.data, av.ctypes.data, bv.ctypes.data,
              xv.ctypes.data)
    st, fl = prob.state, prob.fields
    args = [F(st["ux_b"]), F(st["uy_b"]), F(beta_acx), F(beta_acy), F(N_aa)]
    masks = [np.asfortranarray(st["ssa_mask_acx"], dtype=np.int32),
             np.asfortranarray(st["ssa_mask_acy"], dtype=np.int32)]
    rest = [F(fl["H_ice"]), F(fl["f_ice"]), F(fl["z_bed"]), F(fl["z_sl"]), F(fl["taud_acx"]),
            F(fl["taud_acy"]), F(fl["taul_int_acx"]), F(fl["taul_int_acy"])]
    W = C.c_double if dp else C.c_float
    limp = [None] * 4 if lim is None else [F(a) for a in lim]
    L.oracle_assemble_ssa_energy.restype = C.c_int
    k = L.oracle_assemble_ssa_energy(C.byref(lgs), cap, _p(cf), *[_p(a) for a in args],
                                     *[_p(a) for a in masks], *[_p(a) for a in rest],
                                     W(prob.dx), W(prob.dy), BND[boundaries],
                                     *[(_p(a) if a is not None else None) for a in limp], nx, ny)
    if k < 0:
        raise RuntimeError(f"oracle_assemble_ssa_energy failed ({k})")
    return ap, ai[:k], av[:k], bv, xv, cf


def set_ssa_masks(mask_frnt, f_ice, f_grnd, lateral_bc="all", boundaries="zeros", dp=False):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny = f_ice.shape
    mx = np.zeros((nx, ny), dtype=np.int32, order="F")
    my = np.zeros((nx, ny), dtype=np.int32, order="F")
    lat = {"none": 0, "floating": 1, "float": 1, "marine": 2, "all": 3}[lateral_bc]
    keep = [np.asfortranarray(mask_frnt, dtype=np.int32), np.asfortranarray(f_ice, dtype=dt),
            np.asfortranarray(f_grnd, dtype=dt)]
    L.oracle_set_ssa_masks(_p(mx), _p(my), *[_p(a) for a in keep], lat, nx, ny, BND[boundaries])
    return mx, my


def calc_ice_front(f_ice, f_grnd, z_bed, z_sl, boundaries="zeros", dp=False):
    """calc_ice_front (topography.f90:724-818): the int32 front codes."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny = f_ice.shape
    m = np.zeros((nx, ny), dtype=np.int32, order="F")
    keep = [_F(a, dt) for a in (f_ice, f_grnd, z_bed, z_sl)]
    L.oracle_calc_ice_front(_p(m), *[_p(a) for a in keep], nx, ny, BND[boundaries])
    return m


def calc_gradient(var, f_ice, dxy, grad_lim, direction, zero_outside=False, slope_bg=0.0,
                  boundaries="zeros", dp=False):
    """calc_gradient_acx (direction 0) / calc_gradient_acy (1),
    yelmo_tools.f90:341-478."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    W = _W(dp)
    nx, ny = var.shape
    d = np.zeros((nx, ny), dtype=dt, order="F")
    keep = [_F(var, dt), _F(f_ice, dt)]
    L.oracle_calc_gradient(_p(d), _p(keep[0]), _p(keep[1]), W(dxy), W(grad_lim),
                           int(zero_outside), W(slope_bg), direction, nx, ny, BND[boundaries])
    return d


def calc_driving_stress(H_ice, f_ice, dzsdx, dzsdy, taud_lim, rho_ice, g, boundaries="zeros",
                        dp=False):
    """calc_driving_stress (velocity_general.f90:1118-1205)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    W = _W(dp)
    nx, ny = H_ice.shape
    tx = np.zeros((nx, ny), dtype=dt, order="F")
    ty = np.zeros((nx, ny), dtype=dt, order="F")
    keep = [_F(a, dt) for a in (H_ice, f_ice, dzsdx, dzsdy)]
    L.oracle_calc_driving_stress(_p(tx), _p(ty), *[_p(a) for a in keep], W(taud_lim),
                                 W(rho_ice), W(g), nx, ny, BND[boundaries])
    return tx, ty


def lateral_bc_stress(H, z_srf, z_sl, rho_ice, rho_sw, g, dp=False):
    """calc_lateral_bc_stress (velocity_general.f90:1578-1609) at one point."""
    L = lib(dp)
    W = _W(dp)
    fn = L.oracle_lateral_bc_stress
    fn.restype = W
    fn.argtypes = [W] * 6
    return float(fn(H, z_srf, z_sl, rho_ice, rho_sw, g))


def calc_lateral_bc_stress_2D(mask_frnt, H_ice, f_ice, z_srf, z_sl, rho_ice, rho_sw, g,
                              boundaries="zeros", dp=False):
    """calc_lateral_bc_stress_2D (velocity_general.f90:1453-1576)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    W = _W(dp)
    nx, ny = H_ice.shape
    tx = np.zeros((nx, ny), dtype=dt, order="F")
    ty = np.zeros((nx, ny), dtype=dt, order="F")
    mf = np.asfortranarray(mask_frnt, dtype=np.int32)
    keep = [_F(a, dt) for a in (H_ice, f_ice, z_srf, z_sl)]
    L.oracle_calc_lateral_bc_stress_2D(_p(tx), _p(ty), _p(mf), *[_p(a) for a in keep],
                                       W(rho_ice), W(rho_sw), W(g), nx, ny, BND[boundaries])
    return tx, ty


def dyn_inputs(H_ice, f_ice, f_grnd, z_bed, z_sl, z_srf, dx, dy, taud_lim, grad_lim,
               rho_ice, rho_sw, g, lateral_bc="all", slope_bg=(0.0, 0.0), boundaries="zeros",
               dp=False):
    """The per-solve inputs in calc_ydyn's order: ice front, surface slopes,
    driving stress, lateral front stress, SSA face masks.  Returns a dict."""
    mf = calc_ice_front(f_ice, f_grnd, z_bed, z_sl, boundaries, dp)
    sx = calc_gradient(z_srf, f_ice, dx, grad_lim, 0, False, slope_bg[0], boundaries, dp)
    sy = calc_gradient(z_srf, f_ice, dy, grad_lim, 1, False, slope_bg[1], boundaries, dp)
    tx, ty = calc_driving_stress(H_ice, f_ice, sx, sy, taud_lim, rho_ice, g, boundaries, dp)
    lx, ly = calc_lateral_bc_stress_2D(mf, H_ice, f_ice, z_srf, z_sl, rho_ice, rho_sw, g,
                                       boundaries, dp)
    mx, my = set_ssa_masks(mf, f_ice, f_grnd, lateral_bc, boundaries, dp)
    return dict(mask_frnt=mf, dzsdx=sx, dzsdy=sy, taud_acx=tx, taud_acy=ty, taul_int_acx=lx,
                taul_int_acy=ly, ssa_mask_acx=mx, ssa_mask_acy=my)


def visc_eff_3D_aa(ux, uy, ATT, f_ice, dx, dy, n_glen, eps_0, boundaries="zeros", dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    W = C.c_double if dp else C.c_float
    nx, ny, nz = ATT.shape
    visc = np.zeros((nx, ny, nz), dtype=dt, order="F")
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=dt))
    keep = [F(ux), F(uy), F(ATT), F(f_ice)]
    L.oracle_visc_eff_3D_aa(_p(visc), *[_p(a) for a in keep], W(dx), W(dy), W(n_glen), W(eps_0),
                            nx, ny, nz, BND[boundaries])
    return visc


def strain_rate_2D(ux, uy, f_ice, dx, dy, boundaries="zeros", dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    W = C.c_double if dp else C.c_float
    nx, ny = f_ice.shape
    outs = [np.zeros((nx, ny), dtype=dt, order="F") for _ in range(4)]
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=dt))
    keep = [F(ux), F(uy), F(f_ice)]
    L.oracle_strain_rate_2D(*[_p(o) for o in outs], *[_p(a) for a in keep], W(dx), W(dy), nx, ny,
                            BND[boundaries])
    return outs


def ilu0_factors(a_ptr1, a_idx1, a_val, nb=1):
    """ILU(0) factors (0-based sorted CSR, unit L below the diagonal, U on
    and above it)."""
    L = lib(True)
    n = len(a_ptr1) - 1
    nnz = int(a_ptr1[n]) - 1
    fp = np.zeros(n + 1, dtype=np.int32)
    fc = np.zeros(nnz, dtype=np.int32)
    fv = np.zeros(nnz)
    keep = [np.ascontiguousarray(a_ptr1, dtype=np.int32),
            np.ascontiguousarray(a_idx1, dtype=np.int32),
            np.ascontiguousarray(a_val, dtype=np.float64)]
    rc = L.oracle_ilu0_factors(n, *[_p(a) for a in keep], nb, _p(fp), _p(fc), _p(fv))
    assert rc == 0, rc
    m = fp[n]
    return fp, fc[:m], fv[:m]


def visc_eff_int(visc, H_ice, f_ice, zeta, dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny, nz = visc.shape
    out = np.zeros((nx, ny), dtype=dt, order="F")
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=dt))
    keep = [F(visc), F(H_ice), F(f_ice), np.ascontiguousarray(zeta, dtype=dt)]
    L.oracle_visc_eff_int(_p(out), *[_p(a) for a in keep], nx, ny, nz)
    return out


def calc_beta(prob, par, ux_b, uy_b, dp=True, u_stag=None):
    """calc_beta then stagger_beta on the oracle; returns (beta, beta_acx,
    beta_acy).  u_stag: the (ux, uy) pair stagger_beta reads (ux_b, uy_b by
    default, as in velocity_ssa.f90:247; the depth average for DIVA)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny = prob.nx, prob.ny
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=dt))
    fl = prob.fields
    beta = np.zeros((nx, ny), dtype=dt, order="F")
    bx = np.zeros((nx, ny), dtype=dt, order="F")
    by = np.zeros((nx, ny), dtype=dt, order="F")
    pc = params_c(par)
    z = np.zeros((nx, ny))
    g = lambda k: F(fl.get(k, z))   # fields only beta_gl_scale 2 / beta_gl_stag 3, 4 read
    keep = [F(fl["c_bed"]), F(fl["f_slide"]), F(ux_b), F(uy_b), F(fl["H_ice"]), F(fl["f_ice"]),
            F(fl["H_grnd"]), F(fl["f_grnd"]), g("z_bed"), g("z_sl")]
    rc = L.oracle_calc_beta(_p(beta), *[_p(a) for a in keep], C.byref(pc), nx, ny)
    assert rc == 0, rc
    us = (keep[2], keep[3]) if u_stag is None else (F(u_stag[0]), F(u_stag[1]))
    st = [keep[4], keep[5], us[0], us[1], keep[7], g("f_grnd_acx"), g("f_grnd_acy")]
    rc = L.oracle_stagger_beta(_p(bx), _p(by), _p(beta), *[_p(a) for a in st], C.byref(pc),
                               nx, ny)
    assert rc == 0, rc
    return beta, bx, by


# ---- planted mistakes and the small Picard helpers, for the pinning tests ----

PLANTS = {
    "gq_node_pos": 1, "gq_acx_corner": 2, "gq_act_ignored": 3, "glen_shear": 4, "lim_w": 5,
    "lim_r": 6, "lim_cross": 7, "conv_vel_tol": 8, "conv_mask": 9, "conv_du_reg": 10,
    "conv_norm2": 11, "l1rel_tol": 12, "relax_vel": 13, "relax_visc": 14, "clip": 15,
    "gl_upstream": 16, "beta_min_mask": 17, "diva_shear": 18, "beta_q": 19, "gq_att_act": 20,
    "adv_inflow": 21, "adv_diag": 22, "zstar_norm": 23, "zstar_rho": 24, "subgrid_wt": 25,
    "flux_side": 26, "front_free_bed": 27, "grad_border": 28, "taud_margin": 29,
    "taul_land": 30, "taul_rho": 31,
}


class planted:
    """Context manager: switches on one planted mistake in both oracle builds
    (oracle_plant in ssa_oracle.c), so a test can show its check fails."""

    def __init__(self, name):
        self.k = PLANTS[name]

    def __enter__(self):
        for dp in (False, True):
            C.c_int.in_dll(lib(dp), "oracle_plant").value = self.k
        return self

    def __exit__(self, *exc):
        for dp in (False, True):
            C.c_int.in_dll(lib(dp), "oracle_plant").value = 0
        return False


def _W(dp):
    return C.c_double if dp else C.c_float


def _F(a, dt):
    return np.asfortranarray(np.asarray(a, dtype=dt))


def gq2D_nodes(v, kind, i, j, act=None, boundaries="zeros", dp=True):
    """The 4 quadrature-node values (SW, SE, NE, NW) of cell (i, j) of field v
    on the aa ("aa"), acx or acy grid, with active flags act (None: all)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny = v.shape
    vv = _F(v, dt)
    aa = None if act is None else np.asfortranarray(act, dtype=np.uint8)
    n = np.zeros(4, dtype=dt)
    L.oracle_gq2D_nodes(_p(n), _p(vv), None if aa is None else _p(aa),
                        {"aa": 0, "acx": 1, "acy": 2}[kind], nx, ny, i, j, BND[boundaries])
    return n


def visc_eff_3D_nodes(ux, uy, ATT, f_ice, dx, dy, n_glen, eps_0, boundaries="zeros", dp=True):
    """calc_visc_eff_3D_nodes (visc_method 1) on given fields."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    W = _W(dp)
    nx, ny, nz = ATT.shape
    visc = np.zeros((nx, ny, nz), dtype=dt, order="F")
    keep = [_F(ux, dt), _F(uy, dt), _F(ATT, dt), _F(np.zeros((nx, ny)), dt), _F(f_ice, dt),
            np.zeros(nz, dtype=dt)]
    L.oracle_visc_eff_3D_nodes(_p(visc), *[_p(a) for a in keep], W(dx), W(dy), W(n_glen),
                               W(eps_0), nx, ny, nz, BND[boundaries])
    return visc


def visc_eff_3D_nodes_diva(ux, uy, dudz, dvdz, ATT, f_ice, dx, dy, n_glen, eps_0,
                           boundaries="zeros", dp=True):
    """The DIVA form of calc_visc_eff_3D_nodes (velocity_diva.f90:573-710)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    W = _W(dp)
    nx, ny, nz = ATT.shape
    visc = np.zeros((nx, ny, nz), dtype=dt, order="F")
    keep = [_F(ux, dt), _F(uy, dt), _F(dudz, dt), _F(dvdz, dt), _F(ATT, dt), _F(f_ice, dt)]
    L.oracle_visc_eff_3D_nodes_diva(_p(visc), *[_p(a) for a in keep], W(dx), W(dy), W(n_glen),
                                    W(eps_0), nx, ny, nz, BND[boundaries])
    return visc


def vel_lim_drag(ux, uy, mask_acx, mask_acy, u_max, tau_c, boundaries="zeros", dp=True):
    """calc_vel_lim_drag: (k_acx, k_acy, r_acx, r_acy)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    W = _W(dp)
    nx, ny = ux.shape
    outs = [np.zeros((nx, ny), dtype=dt, order="F") for _ in range(4)]
    keep = [_F(ux, dt), _F(uy, dt), np.asfortranarray(mask_acx, dtype=np.int32),
            np.asfortranarray(mask_acy, dtype=np.int32)]
    L.oracle_vel_lim_drag(*[_p(o) for o in outs], *[_p(a) for a in keep], W(u_max), W(tau_c),
                          nx, ny, BND[boundaries])
    return outs


def picard_conv_l2(ux, uy, ux_prev, uy_prev, mask_acx, mask_acy, tol, iter, iter_max, dp=True):
    """picard_calc_convergence_l2: (conv, resid, points checked); conv 1
    converged, 2 iteration limit, 0 not converged, -1 non-finite."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny = ux.shape
    conv = C.c_int()
    resid = _W(dp)()
    keep = [_F(ux, dt), _F(uy, dt), _F(ux_prev, dt), _F(uy_prev, dt),
            np.asfortranarray(mask_acx, dtype=np.int32), np.asfortranarray(mask_acy, dtype=np.int32)]
    n = L.oracle_picard_conv_l2(C.byref(conv), C.byref(resid), *[_p(a) for a in keep],
                                _W(dp)(tol), iter, iter_max, nx, ny)
    return conv.value, float(resid.value), n


def picard_l1rel(ux, uy, ux_prev, uy_prev, dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    keep = [_F(ux, dt), _F(uy, dt), _F(ux_prev, dt), _F(uy_prev, dt)]
    ex, ey = np.zeros_like(keep[0]), np.zeros_like(keep[0])
    L.oracle_picard_l1rel(_p(ex), _p(ey), *[_p(a) for a in keep], keep[0].size)
    return ex, ey


def picard_relax_vel(ux, uy, ux_prev, uy_prev, rel, dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    a, b = _F(ux, dt).copy(), _F(uy, dt).copy()
    keep = [_F(ux_prev, dt), _F(uy_prev, dt)]
    L.oracle_picard_relax_vel(_p(a), _p(b), _p(keep[0]), _p(keep[1]), _W(dp)(rel), a.size)
    return a, b


def picard_relax_visc(visc, visc_prev, rel, dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    v = _F(visc, dt).copy()
    vp = _F(visc_prev, dt)
    L.oracle_picard_relax_visc(_p(v), _p(vp), _W(dp)(rel), v.size)
    return v


def ssa_vel_clip(ux, uy, u_max, dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    a, b = _F(ux, dt).copy(), _F(uy, dt).copy()
    L.oracle_ssa_vel_clip(_p(a), _p(b), _W(dp)(u_max), a.size)
    return a, b


def set_beta_min_grounded(bx, by, mask_acx, mask_acy, beta_min, dp=True):
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    a, b = _F(bx, dt).copy(), _F(by, dt).copy()
    keep = [np.asfortranarray(mask_acx, dtype=np.int32), np.asfortranarray(mask_acy, dtype=np.int32)]
    L.oracle_set_beta_min_grounded(_p(a), _p(b), _p(keep[0]), _p(keep[1]), _W(dp)(beta_min),
                                   a.size)
    return a, b


def stagger_beta(beta, f_ice, f_grnd, par, dp=True, H_ice=None, ux=None, uy=None,
                 f_grnd_acx=None, f_grnd_acy=None):
    """stagger_beta on given aa-node friction: (beta_acx, beta_acy).  H_ice,
    ux, uy and f_grnd_acx/acy are read by beta_gl_stag 3 and 4 only (zeros
    when not given)."""
    L = lib(dp)
    dt = np.float64 if dp else np.float32
    nx, ny = beta.shape
    bx = np.zeros((nx, ny), dtype=dt, order="F")
    by = np.zeros((nx, ny), dtype=dt, order="F")
    z = np.zeros((nx, ny))
    opt = lambda a: _F(z if a is None else a, dt)
    keep = [_F(beta, dt), opt(H_ice), _F(f_ice, dt), opt(ux), opt(uy), _F(f_grnd, dt),
            opt(f_grnd_acx), opt(f_grnd_acy)]
    pc = params_c(par)
    rc = L.oracle_stagger_beta(_p(bx), _p(by), *[_p(a) for a in keep], C.byref(pc), nx, ny)
    assert rc == 0, rc
    return bx, by


def beta_gl_flux_weight(beta_a, u_a, u_b, H_a, H_b, f_grnd_ac, dp=True):
    """calc_beta_gl_flux_weight (basal_dragging.f90:1760-1845): beta_ac for
    grounded node a, floating node b."""
    L = lib(dp)
    W = _W(dp)
    fn = L.oracle_beta_gl_flux_weight
    fn.restype = W
    fn.argtypes = [W] * 6
    return float(fn(beta_a, u_a, u_b, H_a, H_b, f_grnd_ac))


def assemble_advection(H, H_ref, ux, uy, F, mask, dx, dy, dt, boundaries="zeros", dp=False):
    """linear_solver_matrix_advection_csr_2D (impl-lis advection, with
    limit_border_inflow applied to the velocities): 1-based CSR
    (a_ptr, a_idx, a_val), b and x0 in the scalar row order n = i ny + j."""
    L = lib(dp)
    dt_ = np.float64 if dp else np.float32
    W = _W(dp)
    nx, ny = H.shape
    n = nx * ny
    ap = np.zeros(n + 1, dtype=np.int32)
    ai = np.zeros(5 * n, dtype=np.int32)
    av = np.zeros(5 * n)
    b = np.zeros(n)
    x0 = np.zeros(n)
    keep = [_F(H, dt_), _F(H_ref, dt_), _F(ux, dt_), _F(uy, dt_), _F(F, dt_),
            np.asfortranarray(mask, dtype=np.int32)]
    nnz = L.oracle_assemble_advection(_p(ap), _p(ai), _p(av), _p(b), _p(x0),
                                      *[_p(a) for a in keep], W(dx), W(dy), W(dt), nx, ny,
                                      BND[boundaries])
    return ap, ai[:nnz].copy(), av[:nnz].copy(), b, x0
