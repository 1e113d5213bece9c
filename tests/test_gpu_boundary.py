"""The outer boundary called the way a Fortran host calls it: the
argument-order drop-ins ygpu_velocity_ssa (calc_velocity_ssa,
velocity_ssa.f90:60-63) and ygpu_velocity_diva (calc_velocity_diva,
velocity_diva.f90:59-64) with wp = sp (float32) arrays, on the 81 x 81,
dx = 16 km workload, against the FP32 oracle; and the optional neff_hook of
calc_velocity_diva (velocity_diva.f90:99-107, 396-400), the same host
function handed to the oracle and to the library.

Tolerances: FP32 fields (the reference's wp = sp) against the FP32 oracle
1e-5 relative L2 with Picard counts equal +-1 (rounding of the FP32 field
arithmetic); FP64 with a hook 1e-8 and Picard +-1.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle
from tests._helpers import rel_l2, vel_rel_l2
from yelmo_amd import _lib, generator, velocity
from yelmo_amd.params import SsaParams

pytestmark = pytest.mark.gpu


def _config0(dtype=np.float32):
    return generator.make_problem(81, 16e3, seed=0, dtype=dtype)


def test_velocity_ssa_drop_in_matches_fp32_oracle():
    print("case ygpu_velocity_ssa 81x81 ref", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="ref")
    p = _config0()
    ref_st, ref = oracle.calc_velocity_ssa(p, par, dp=False)
    assert ref["rc"] == 0
    g = p.copy()
    info = velocity.ygpu_velocity_ssa(g, par)
    err = vel_rel_l2(g.state, ref_st)
    print("iters gpu", info.ssa_iter_now, "oracle", ref["ssa_iter_now"], "lim_n",
          info.ssa_lim_n, ref["ssa_lim_n"], "rel L2", err, flush=True)
    assert info.status == 0
    assert g.state["ux_b"].dtype == np.float32
    assert abs(info.ssa_iter_now - ref["ssa_iter_now"]) <= 1
    assert err <= 1e-5
    assert rel_l2(g.state["visc_eff_int"], ref_st["visc_eff_int"]) <= 1e-4
    assert abs(info.ssa_lim_n - ref["ssa_lim_n"]) <= max(2, ref["ssa_lim_n"] // 50)


def test_velocity_diva_drop_in_matches_fp32_oracle():
    print("case ygpu_velocity_diva 81x81 ref", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="ref")
    p = _config0()
    ref_st, ref = oracle.calc_velocity_diva(p, par, dp=False)
    assert ref["rc"] == 0
    g = p.copy()
    info = velocity.ygpu_velocity_diva(g, par)
    eb = vel_rel_l2(g.state, ref_st, "ux_bar", "uy_bar")
    e3 = vel_rel_l2(g.state, ref_st, "ux", "uy")
    ei = vel_rel_l2(g.state, ref_st, "ux_i", "uy_i")
    print("iters gpu", info.ssa_iter_now, "oracle", ref["ssa_iter_now"], "ubar", eb, "3D", e3,
          "shear part", ei, flush=True)
    assert info.status == 0
    assert abs(info.ssa_iter_now - ref["ssa_iter_now"]) <= 1
    assert eb <= 1e-5 and e3 <= 1e-5
    assert ei <= 1e-4


def _speed_hook(c0, calls):
    """neff_hook: c_bed = c0 / (1 + |u_b| / 1000 m/yr), computed in the
    caller's real type (an effective pressure that falls with sliding)."""
    base = {4: np.asfortranarray(c0, dtype=np.float32), 8: np.asfortranarray(c0, dtype=np.float64)}

    def hook(ctx, nx, ny, rb, c_bed, ux_b, uy_b):
        ct = C.c_float if rb == 4 else C.c_double
        n = nx * ny
        cb = np.ctypeslib.as_array(C.cast(c_bed, C.POINTER(ct)), shape=(n,))
        ux = np.ctypeslib.as_array(C.cast(ux_b, C.POINTER(ct)), shape=(n,))
        uy = np.ctypeslib.as_array(C.cast(uy_b, C.POINTER(ct)), shape=(n,))
        dt = cb.dtype.type
        s = np.sqrt(ux * ux + uy * uy)
        cb[:] = base[rb].ravel(order="F") / (dt(1) + s / dt(1000))
        calls.append(1)

    return _lib.NEFF_HOOK(hook)


def test_diva_neff_hook_fp64_matches_oracle():
    print("case diva neff_hook 41x41 fp64", flush=True)
    par = SsaParams(visc_method=2, beta_method=4, precision="fp64")
    p = generator.make_problem(41, 32e3, seed=0, dtype=np.float64)
    c0 = p.fields["c_bed"].copy()
    ocalls, gcalls = [], []
    ohook = _speed_hook(c0, ocalls)
    ref_st, ref = oracle.calc_velocity_diva(p, par, dp=True, neff_hook=ohook)
    assert ref["rc"] == 0
    g = oracle.diva_state(p.copy())
    ghook = _speed_hook(c0, gcalls)
    info = velocity.calc_velocity_diva(g, par, neff_hook=ghook)
    eb = vel_rel_l2(g.state, ref_st, "ux_bar", "uy_bar")
    ec = rel_l2(g.fields["c_bed"], ref_st["c_bed"])
    print("iters gpu", info.ssa_iter_now, "oracle", ref["ssa_iter_now"], "hook calls",
          len(gcalls), len(ocalls), "ubar", eb, "c_bed", ec, flush=True)
    assert info.status == 0
    assert len(gcalls) == info.ssa_iter_now and len(ocalls) == ref["ssa_iter_now"]
    assert abs(info.ssa_iter_now - ref["ssa_iter_now"]) <= 1
    assert eb <= 1e-8 and ec <= 1e-8
    # the hook's last update stays in c_bed, and it did change it
    assert rel_l2(g.fields["c_bed"], c0) > 1e-3
    # without the hook the same domain converges in different iterations
    plain = oracle.diva_state(p.copy())
    plain_info = velocity.calc_velocity_diva(plain, par)
    assert vel_rel_l2(plain.state, g.state, "ux_bar", "uy_bar") > 1e-3
    print("without hook: iters", plain_info.ssa_iter_now, flush=True)


def test_diva_neff_hook_drop_in_fp32():
    print("case ygpu_velocity_diva neff_hook 41x41 ref", flush=True)
    par = SsaParams(visc_method=1, beta_method=2, precision="ref")
    p = generator.make_problem(41, 32e3, seed=0, dtype=np.float32)
    c0 = p.fields["c_bed"].copy()
    ocalls, gcalls = [], []
    ref_st, ref = oracle.calc_velocity_diva(p, par, dp=False, neff_hook=_speed_hook(c0, ocalls))
    g = p.copy()
    info = velocity.ygpu_velocity_diva(g, par, neff_hook=_speed_hook(c0, gcalls))
    eb = vel_rel_l2(g.state, ref_st, "ux_bar", "uy_bar")
    ec = rel_l2(g.fields["c_bed"], ref_st["c_bed"])
    print("iters gpu", info.ssa_iter_now, "oracle", ref["ssa_iter_now"], "ubar", eb, "c_bed", ec,
          flush=True)
    assert info.status == 0
    assert abs(info.ssa_iter_now - ref["ssa_iter_now"]) <= 1
    assert eb <= 1e-5 and ec <= 1e-5


def test_hook_rejected_outside_diva_one_shot():
    p = generator.make_problem(41, 32e3, seed=0, dtype=np.float64)
    calls = []
    with pytest.raises(_lib.YgpuError):
        velocity._run_batch([p], SsaParams(), False, hooks=[_speed_hook(p.fields["c_bed"], calls)])
    assert not calls


def test_substituted_solver_leaves_a_note():
    h = _lib.default_handle()
    p = generator.make_problem(41, 32e3, seed=0, dtype=np.float64)
    for name in ("minres", "bicgsafe"):
        par = SsaParams(ssa_lis_opt=f"-i {name} -p ilu -maxiter 1000 -tol 1.0e-8")
        info = velocity.calc_velocity_ssa(p.copy(), par, handle=h)
        assert info.status == 0
        assert h.note() == f"-i {name} solved with BiCGSTAB"
    velocity.calc_velocity_ssa(p.copy(), SsaParams(), handle=h)
    assert h.note() == ""
