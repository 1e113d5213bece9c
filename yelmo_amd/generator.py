"""Synthetic SSA/DIVA problems of the benchmark workloads.

One generator serves every workload: a radial ice cap on a seeded smooth bed,
with flotation, ice fronts, driving and front stress derived by the
input-side routines of ``yelmo_amd.inputs`` (restating calc_ydyn's
preparation, src/yelmo_dynamics.f90:105-165).  Seeds and smoothing lengths
are fixed here so every workload is reproducible:

* thickness  H = 3000 m * sqrt(max(0, 1 - (r/550 km)^2)) plus white noise
  smoothed with a 40 km Gaussian and scaled to 50 m standard deviation, where
  H > 0, clipped at 0, without one-cell-wide ice tongues;
* bed        z_bed = -200 m + 500 m * s, s white noise smoothed with a 150 km
  Gaussian, normalised to max|s| = 1;
* friction   c_bed = 1e4 Pa * exp(sigma_ln * g), g white noise smoothed with
  a 64 km Gaussian, normalised to unit standard deviation; 0 where floating;
* flow       ATT = A on every level (or A(zeta) for the DIVA workload);
* levels     zeta = Yelmo's "exp" scale with exponent 2 (calc_zeta).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter

from . import inputs

RHO_ICE = 910.0
RHO_SW = 1028.0
G = 9.81


def _smooth_noise(rng, nx, ny, length_m, dx):
    w = rng.standard_normal((nx, ny))
    s = gaussian_filter(w, sigma=max(length_m / dx, 0.5), mode="reflect")
    return s


@dataclass
class SsaProblem:
    """All inputs of calc_velocity_ssa / calc_velocity_diva for one domain,
    plus the INOUT state, as Fortran-ordered arrays."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    n_glen: float
    fields: dict = field(default_factory=dict)
    state: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return self.nx * self.ny

    @property
    def unknowns(self) -> int:
        return 2 * self.nx * self.ny

    def copy(self) -> "SsaProblem":
        return SsaProblem(self.nx, self.ny, self.nz, self.dx, self.dy, self.n_glen,
                          {k: v.copy(order="K") for k, v in self.fields.items()},
                          {k: v.copy(order="K") for k, v in self.state.items()})

    def astype(self, dtype) -> "SsaProblem":
        out = self.copy()
        for d in (out.fields, out.state):
            for k, v in d.items():
                if v.dtype.kind == "f":
                    d[k] = np.asfortranarray(v.astype(dtype))
        return out


def _geometry(nx, ny, dx, seed, cb_seed=None, sigma_ln=1.0, H_noise_seed=None):
    """Thickness, bed and unmasked friction of one domain (the seeded part)."""
    rng_geo = np.random.default_rng(seed)
    xc = (np.arange(1, nx + 1) - (nx + 1) / 2.0) * dx
    yc = (np.arange(1, ny + 1) - (ny + 1) / 2.0) * dx
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    r = np.hypot(X, Y)
    H = 3000.0 * np.sqrt(np.maximum(0.0, 1.0 - (r / 550e3) ** 2))
    rng_h = rng_geo if H_noise_seed is None else np.random.default_rng(H_noise_seed)
    noise = _smooth_noise(rng_h, nx, ny, 40e3, dx)
    noise *= 50.0 / max(noise.std(), 1e-12)
    H = np.where(H > 0, np.maximum(H + noise, 0.0), 0.0)
    s = _smooth_noise(rng_geo, nx, ny, 150e3, dx)
    s /= max(np.abs(s).max(), 1e-12)
    z_bed = -200.0 + 500.0 * s
    rng_cb = rng_geo if cb_seed is None else np.random.default_rng(cb_seed)
    gn = _smooth_noise(rng_cb, nx, ny, 64e3, dx)
    # a field smoothed over far more than the grid is nearly constant: its
    # spread is then no noise to normalise, and the friction stays uniform
    sd = gn.std()
    gn = gn / sd if sd > 1e-3 * np.abs(gn).max() else np.zeros_like(gn)
    c_raw = 1e4 * np.exp(sigma_ln * gn)
    return H, z_bed, c_raw


def _prune_tongues(H, boundaries):
    """Removes ice cells that have no ice on both sides along x or along y
    (one-cell-wide tongues), repeatedly.  Both faces of such a cell are
    calving fronts whose lateral-stress rows (solver_ssa_ac.f90:402-471)
    are the same equation, so the SSA system would be singular."""
    nx, ny = H.shape
    im1, ip1, jm1, jp1 = inputs.neighbors(nx, ny, boundaries)
    H = H.copy()
    while True:
        ice = H > 0
        strip = ice & ((~ice[im1, :] & ~ice[ip1, :]) | (~ice[:, jm1] & ~ice[:, jp1]))
        if not strip.any():
            return H
        H[strip] = 0.0


def _derive(nx, ny, dx, H, z_bed, c_raw, A, nz, dtype, lateral_bc, boundaries) -> SsaProblem:
    """Every input that follows from thickness, bed and friction, with the
    zero initial state."""
    H = _prune_tongues(H, boundaries)
    z_sl = np.zeros_like(H)
    f_ice = np.where(H > 0, 1.0, 0.0)
    H_grnd = inputs.calc_H_grnd(H, f_ice, z_bed, z_sl, RHO_ICE, RHO_SW)
    f_grnd = np.where(H_grnd > 0, 1.0, 0.0)
    z_srf = inputs.calc_z_srf_max(H, f_ice, z_bed, z_sl, RHO_ICE, RHO_SW)
    mask_frnt = inputs.calc_ice_front(f_ice, f_grnd, z_bed, z_sl, boundaries)
    dzsdx = inputs.calc_gradient_acx(z_srf, f_ice, dx, 0.5, False, boundaries)
    dzsdy = inputs.calc_gradient_acy(z_srf, f_ice, dx, 0.5, False, boundaries)
    taud_acx, taud_acy = inputs.calc_driving_stress(H, f_ice, dzsdx, dzsdy, dx, 2e5,
                                                    RHO_ICE, G, boundaries)
    taul_acx, taul_acy = inputs.calc_lateral_bc_stress_2D(mask_frnt, H, f_ice, z_srf, z_sl,
                                                          RHO_ICE, RHO_SW, G, boundaries)
    c_bed = np.where(f_grnd > 0, c_raw, 0.0)
    zeta = inputs.calc_zeta(nz, "exp", 2.0, dtype=np.float64)
    if callable(A):
        Az = np.array([A(z) for z in zeta])
    else:
        Az = np.full(nz, float(A))
    ATT = np.broadcast_to(Az[None, None, :], (nx, ny, nz)).copy()
    mx, my = inputs.set_ssa_masks(mask_frnt, f_ice, f_grnd, lateral_bc, boundaries)
    im1, ip1, jm1, jp1 = inputs.neighbors(nx, ny, boundaries)
    f_grnd_acx = 0.5 * (f_grnd + f_grnd[ip1, :])
    f_grnd_acy = 0.5 * (f_grnd + f_grnd[:, jp1])

    F = lambda a: np.asfortranarray(np.asarray(a, dtype=dtype))
    fields = dict(
        H_ice=F(H), f_ice=F(f_ice), z_bed=F(z_bed), z_sl=F(z_sl), z_srf=F(z_srf),
        H_grnd=F(H_grnd), f_grnd=F(f_grnd), f_grnd_acx=F(f_grnd_acx),
        f_grnd_acy=F(f_grnd_acy), c_bed=F(c_bed), f_slide=F(np.ones_like(H)),
        taud_acx=F(taud_acx), taud_acy=F(taud_acy), taul_int_acx=F(taul_acx),
        taul_int_acy=F(taul_acy), ATT=F(ATT), zeta_aa=np.asarray(zeta, dtype=dtype),
        mask_frnt=np.asfortranarray(mask_frnt.astype(np.int32)),
    )
    z2 = np.zeros((nx, ny))
    state = dict(
        ux_b=F(z2), uy_b=F(z2), taub_acx=F(z2), taub_acy=F(z2),
        visc_eff=F(np.full((nx, ny, nz), 1e3)), visc_eff_int=F(z2),
        beta=F(z2), beta_acx=F(z2), beta_acy=F(z2),
        ssa_mask_acx=np.asfortranarray(mx), ssa_mask_acy=np.asfortranarray(my),
        ssa_err_acx=F(z2), ssa_err_acy=F(z2),
    )
    return SsaProblem(nx, ny, nz, float(dx), float(dx), 3.0, fields, state)


def make_problem(nx: int, dx: float, seed: int = 0, A=1e-16, nz: int = 11,
                 dtype=np.float32, cb_seed: int | None = None, sigma_ln: float = 1.0,
                 lateral_bc: str = "all", boundaries: str = "zeros",
                 H_noise_seed: int | None = None, ny: int | None = None,
                 H_offset=None) -> SsaProblem:
    """Build one problem.  ``A`` is a scalar or a callable of zeta giving the
    rate factor on each level; ``cb_seed`` reseeds only the friction field
    (ensemble members share geometry); ``H_offset`` adds a thickness change
    where there is ice."""
    ny = nx if ny is None else ny
    H, z_bed, c_raw = _geometry(nx, ny, dx, seed, cb_seed, sigma_ln, H_noise_seed)
    if H_offset is not None:
        H = np.where(H > 0, np.maximum(H + H_offset, 0.0), 0.0)
    return _derive(nx, ny, dx, H, z_bed, c_raw, A, nz, dtype, lateral_bc, boundaries)


# ---- the benchmark workloads -------------------------------------------

def workload_single(side: int = 161, seed: int = 0, dtype=np.float32, nz: int = 11):
    """One instance on a side x side grid spanning 1296 km (81 x 16 km)."""
    dx = 16e3 * 80.0 / (side - 1)
    return make_problem(side, dx, seed=seed, dtype=dtype, nz=nz)


def workload_ensemble(members: int = 64, side: int = 161, seed: int = 0, dtype=np.float32,
                      nz: int = 11):
    """Ensemble of ``members`` problems sharing geometry, each with its own
    seeded c_bed field and A drawn log-uniformly in [1e-17, 1e-16]."""
    dx = 16e3 * 80.0 / (side - 1)
    rng = np.random.default_rng(10_000 + seed)
    A = 10.0 ** rng.uniform(-17.0, -16.0, size=members)
    return [make_problem(side, dx, seed=seed, A=float(A[m]), cb_seed=1_000 + seed * 97 + m,
                         dtype=dtype, nz=nz) for m in range(members)]


DIVA_STATE_3D = ("ux", "uy", "ux_i", "uy_i", "duxdz", "duydz")
DIVA_STATE_2D = ("ux_bar", "uy_bar", "beta_eff")


def add_diva_state(prob: SsaProblem) -> SsaProblem:
    """Adds the INOUT arrays that calc_velocity_diva holds beyond the SSA ones
    (3D velocity and shear, depth-averaged velocity, beta_eff), zero, where
    prob does not hold them yet."""
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    dt = prob.fields["H_ice"].dtype
    for n in DIVA_STATE_3D:
        if n not in prob.state:
            prob.state[n] = np.zeros((nx, ny, nz), dtype=dt, order="F")
    for n in DIVA_STATE_2D:
        if n not in prob.state:
            prob.state[n] = np.zeros((nx, ny), dtype=dt, order="F")
    return prob


DIVA_SIDES = (96, 128, 192, 256, 384)


def workload_diva_uneven(domains: int = 48, seed: int = 0, dtype=np.float32, nz: int = 21,
                         sides=None):
    """Uneven batch of DIVA domains, side drawn from DIVA_SIDES, each spanning
    1536 km; A(zeta) = 1e-16 exp(-3 (1 - zeta))."""
    rng = np.random.default_rng(20_000 + seed)
    if sides is None:
        sides = rng.choice(DIVA_SIDES, size=domains)
    A = lambda z: 1e-16 * np.exp(-3.0 * (1.0 - z))
    return [add_diva_state(make_problem(int(s), 1536e3 / int(s), seed=seed + 31 * k, A=A, nz=nz,
                                        dtype=dtype))
            for k, s in enumerate(sides)]


# ---- served solve sequences ------------------------------------------------

SERVE_SIDE_RANGE = (81, 641)
SERVE_SIDE_MEDIAN = 161.0
SERVE_SIDE_LOGSIGMA = 0.5
SERVE_SOLVES = 5
SERVE_DH = 10.0        # max |thickness increment| between solves (m)


def serve_request_sides(nreq: int, seed: int = 0):
    """Grid sides of a served stream: lognormal (median 161, log-sigma 0.5),
    clipped to [81, 641]."""
    rng = np.random.default_rng(30_000 + seed)
    s = np.exp(rng.normal(np.log(SERVE_SIDE_MEDIAN), SERVE_SIDE_LOGSIGMA, size=nreq))
    return np.clip(np.rint(s), *SERVE_SIDE_RANGE).astype(int)


def served_sequence(side: int, seed: int, nsolves: int = SERVE_SOLVES, dtype=np.float32,
                    nz: int = 11):
    """One served request: ``nsolves`` successive problems on a side x side
    grid spanning 1536 km.  Between solves the thickness takes a seeded
    smooth increment (40 km Gaussian, max |dH| = 10 m) where there is ice,
    and every thickness-derived input is recomputed.  Only the first problem's
    INOUT state is meant to be used: later solves start from the previous
    solve's velocity, viscosity and friction (a warm start)."""
    dx = 1536e3 / side
    H, z_bed, c_raw = _geometry(side, side, dx, seed)
    rng = np.random.default_rng(40_000 + seed)
    out = []
    for k in range(nsolves):
        if k > 0:
            inc = _smooth_noise(rng, side, side, 40e3, dx)
            inc *= SERVE_DH / max(np.abs(inc).max(), 1e-12)
            H = np.where(H > 0, np.maximum(H + inc, 0.0), 0.0)
        out.append(_derive(side, side, dx, H, z_bed, c_raw, 1e-16, nz, dtype, "all", "zeros"))
    return out
