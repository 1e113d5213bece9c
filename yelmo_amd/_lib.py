"""ctypes binding of libygpu.so (the C-ABI declared in include/ygpu.h).

The HIP library is the only implementation of the solver: if it is missing
or the GPU cannot be opened, every entry point raises.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libygpu.so"

YGPU_OK = 0
YGPU_ERR_ARG = -1
YGPU_ERR_DEVICE = -2
YGPU_ERR_UNSUPPORTED = -3
YGPU_ERR_NONFINITE = -6

PREC = {"fp64": 0, "ref": 1, "mixed": 2}
BND = {"zeros": 0, "infinite": 1, "mask": 2, "MISMIP3D": 3, "TROUGH": 4, "periodic": 5,
       "periodic-x": 6, "periodic-y": 7}


class SsaParamsC(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "ssa_solver", "boundaries", "visc_method", "beta_method", "beta_gl_scale",
        "beta_gl_stag", "ssa_vel_lim_method", "ssa_iter_max", "ssa_write_log", "precision")]
    _fields_ += [(n, C.c_double) for n in (
        "visc_const", "beta_const", "beta_q", "beta_u0", "beta_gl_f", "H_grnd_lim", "beta_min",
        "eps_0", "ssa_vel_lim_tau", "ssa_vel_max", "ssa_iter_rel", "ssa_iter_conv", "rho_ice",
        "rho_sw", "g")]
    _fields_ += [("no_slip", C.c_int32), ("block", C.c_int32), ("ssa_lis_opt", C.c_char * 256)]


_P = C.c_void_p
PROBLEM_PTRS_INOUT = ("ux_b", "uy_b", "taub_acx", "taub_acy", "visc_eff", "visc_eff_int")
PROBLEM_PTRS_ERR = ("ssa_err_acx", "ssa_err_acy", "beta", "beta_acx", "beta_acy")
PROBLEM_PTRS_IN = ("c_bed", "f_slide", "taud_acx", "taud_acy", "taul_int_acx", "taul_int_acy",
                   "H_ice", "f_ice", "H_grnd", "f_grnd", "f_grnd_acx", "f_grnd_acy", "ATT",
                   "zeta_aa", "z_sl", "z_bed", "z_srf")
PROBLEM_PTRS_DIVA = ("ux", "uy", "ux_bar", "uy_bar", "ux_i", "uy_i", "beta_eff", "duxdz", "duydz")


class SsaProblemC(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("real_bytes", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("n_glen", C.c_double)]
    _fields_ += [(n, _P) for n in PROBLEM_PTRS_INOUT]
    _fields_ += [("ssa_mask_acx", _P), ("ssa_mask_acy", _P)]
    _fields_ += [(n, _P) for n in PROBLEM_PTRS_ERR]
    _fields_ += [(n, _P) for n in PROBLEM_PTRS_IN]
    _fields_ += [(n, _P) for n in PROBLEM_PTRS_DIVA]
    _fields_ += [(n, C.c_int32) for n in ("ssa_iter_now", "ssa_lin_iter", "ssa_lin_fail",
                                          "ssa_lim_n", "status", "pad")]
    _fields_ += [("picard_resid", C.c_double), ("solve_seconds", C.c_double)]
    _fields_ += [("neff_hook", C.c_void_p), ("hook_ctx", C.c_void_p)]


# neff_hook of calc_velocity_diva: (ctx, nx, ny, real_bytes, c_bed, ux_b, uy_b)
NEFF_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                        C.c_void_p, C.c_void_p)


# every symbol include/ygpu.h declares
EXPORTS = ("ygpu_create", "ygpu_destroy", "ygpu_last_error", "ygpu_last_note",
           "ygpu_default_params", "ygpu_csr_solve", "ygpu_csr_solve_noalloc", "ygpu_dyn_inputs",
           "ygpu_velocity_ssa", "ygpu_velocity_diva",
           "ygpu_velocity_ssa_batch",
           "ygpu_velocity_diva_batch", "ygpu_batch_create", "ygpu_batch_run",
           "ygpu_batch_download", "ygpu_batch_destroy", "ygpu_batch_spmv", "ygpu_csr_spmv", "ygpu_session_create", "ygpu_session_solve",
           "ygpu_session_destroy", "ygpu_stats_enable", "ygpu_stats_get",
           "ygpu_server_create", "ygpu_submit", "ygpu_server_step", "ygpu_poll", "ygpu_fetch",
           "ygpu_server_destroy", "ygpu_autotune", "ygpu_set_schedule", "ygpu_footprint_bytes")

YGPU_SEQ_CONTINUE = 1
YGPU_SEQ_LAST = 2


class StatsC(C.Structure):
    _fields_ = [("krylov_launches", C.c_int64), ("krylov_ms", C.c_double),
                ("krylov_bytes", C.c_double), ("krylov_iters", C.c_int64),
                ("unknown_iters", C.c_double), ("picard_iters", C.c_int64),
                ("krylov_group", C.c_int32), ("krylov_block", C.c_int32),
                ("refine_sweeps", C.c_int64)]

_lib = None


def load() -> C.CDLL:
    """Load libygpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m yelmo_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    I, D, PP = C.c_int32, C.c_double, C.c_void_p
    lib.ygpu_create.argtypes = [I, C.POINTER(C.c_void_p)]
    lib.ygpu_create.restype = C.c_int
    lib.ygpu_destroy.argtypes = [PP]
    lib.ygpu_destroy.restype = C.c_int
    lib.ygpu_last_error.argtypes = [PP]
    lib.ygpu_last_error.restype = C.c_char_p
    lib.ygpu_last_note.argtypes = [PP]
    lib.ygpu_last_note.restype = C.c_char_p
    # the outer drop-ins take calc_velocity_ssa's / calc_velocity_diva's
    # arguments one by one: pointers for the arrays, then dx, dy, n_glen
    lib.ygpu_velocity_ssa.argtypes = [PP, C.POINTER(SsaParamsC), I, I, I] + [PP] * 34 + \
        [C.c_float] * 3
    lib.ygpu_velocity_ssa.restype = C.c_int
    lib.ygpu_velocity_diva.argtypes = [PP, C.POINTER(SsaParamsC), I, I, I] + [PP] * 43 + \
        [C.c_float] * 3 + [NEFF_HOOK, PP]
    lib.ygpu_velocity_diva.restype = C.c_int
    lib.ygpu_default_params.argtypes = [C.POINTER(SsaParamsC)]
    lib.ygpu_default_params.restype = None
    lib.ygpu_csr_solve.argtypes = [PP, I, I, PP, PP, PP, PP, PP, C.c_char_p, C.POINTER(I),
                                   C.POINTER(I), C.POINTER(D), C.POINTER(D)]
    lib.ygpu_csr_solve.restype = C.c_int
    lib.ygpu_csr_solve_noalloc.argtypes = lib.ygpu_csr_solve.argtypes
    lib.ygpu_csr_solve_noalloc.restype = C.c_int
    lib.ygpu_dyn_inputs.argtypes = ([C.c_void_p, C.POINTER(SsaParamsC), C.c_int32, C.c_int32]
                                    + [C.c_float] * 6 + [C.c_int32] + [C.c_void_p] * 15)
    lib.ygpu_dyn_inputs.restype = C.c_int
    lib.ygpu_velocity_ssa_batch.argtypes = [PP, C.POINTER(SsaParamsC), I, C.POINTER(SsaProblemC)]
    lib.ygpu_velocity_ssa_batch.restype = C.c_int
    lib.ygpu_velocity_diva_batch.argtypes = [PP, C.POINTER(SsaParamsC), I, C.POINTER(SsaProblemC)]
    lib.ygpu_velocity_diva_batch.restype = C.c_int
    lib.ygpu_session_create.argtypes = [PP, C.POINTER(SsaParamsC), C.POINTER(SsaProblemC),
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]
    lib.ygpu_session_create.restype = C.c_int
    lib.ygpu_session_solve.argtypes = [PP, C.POINTER(SsaProblemC)]
    lib.ygpu_session_solve.restype = C.c_int
    lib.ygpu_session_destroy.argtypes = [PP]
    lib.ygpu_session_destroy.restype = C.c_int
    lib.ygpu_batch_create.argtypes = [PP, C.POINTER(SsaParamsC), I, C.POINTER(SsaProblemC), I,
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]
    lib.ygpu_batch_create.restype = C.c_int
    lib.ygpu_batch_run.argtypes = [PP, I, C.POINTER(D)]
    lib.ygpu_batch_run.restype = C.c_int
    lib.ygpu_batch_download.argtypes = [PP, C.POINTER(SsaProblemC)]
    lib.ygpu_batch_download.restype = C.c_int
    lib.ygpu_batch_spmv.argtypes = [PP, I, I, PP, C.POINTER(D), C.POINTER(D)]
    lib.ygpu_batch_spmv.restype = C.c_int
    lib.ygpu_csr_spmv.argtypes = [PP, I, I, PP, PP, PP, PP, PP, C.c_char_p]
    lib.ygpu_csr_spmv.restype = C.c_int
    lib.ygpu_batch_destroy.argtypes = [PP]
    lib.ygpu_batch_destroy.restype = C.c_int
    lib.ygpu_stats_enable.argtypes = [PP, I]
    lib.ygpu_stats_enable.restype = C.c_int
    lib.ygpu_stats_get.argtypes = [PP, C.POINTER(StatsC)]
    lib.ygpu_stats_get.restype = C.c_int
    lib.ygpu_server_create.argtypes = [PP, C.POINTER(SsaParamsC), I, I, PP, PP, I, C.c_int64,
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]
    lib.ygpu_server_create.restype = C.c_int
    lib.ygpu_submit.argtypes = [PP, C.POINTER(SsaProblemC), C.c_int64, I, C.POINTER(C.c_int64)]
    lib.ygpu_submit.restype = C.c_int
    lib.ygpu_server_step.argtypes = [PP, C.POINTER(I), C.POINTER(I)]
    lib.ygpu_server_step.restype = C.c_int
    lib.ygpu_poll.argtypes = [PP, C.POINTER(C.c_int64)]
    lib.ygpu_poll.restype = C.c_int
    lib.ygpu_fetch.argtypes = [PP, C.c_int64, C.POINTER(SsaProblemC)]
    lib.ygpu_fetch.restype = C.c_int
    lib.ygpu_server_destroy.argtypes = [PP]
    lib.ygpu_server_destroy.restype = C.c_int
    lib.ygpu_autotune.argtypes = [PP, C.POINTER(SsaParamsC), C.POINTER(SsaProblemC), I, I,
                                  C.POINTER(I), C.POINTER(I), C.POINTER(D)]
    lib.ygpu_autotune.restype = C.c_int
    lib.ygpu_set_schedule.argtypes = [PP, I, I, I, I]
    lib.ygpu_set_schedule.restype = C.c_int
    lib.ygpu_footprint_bytes.argtypes = [C.POINTER(SsaParamsC), I, I, I]
    lib.ygpu_footprint_bytes.restype = C.c_int64
    _lib = lib
    return lib


class YgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ygpu error {code}: {msg}")
        self.code = code


class Handle:
    """One device handle (its own HIP stream)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.ygpu_create(device, C.byref(h))
        if rc != YGPU_OK:
            raise YgpuError(rc, f"ygpu_create(device={device}) failed (no GPU visible?)")
        self.h = h

    def check(self, rc):
        if rc != YGPU_OK:
            raise YgpuError(rc, self.lib.ygpu_last_error(self.h).decode())

    def note(self) -> str:
        """What the last call that read options did differently ("" if nothing)."""
        return self.lib.ygpu_last_note(self.h).decode()

    def stats_enable(self, on: bool = True):
        self.check(self.lib.ygpu_stats_enable(self.h, int(on)))

    def stats(self) -> dict:
        st = StatsC()
        self.check(self.lib.ygpu_stats_get(self.h, C.byref(st)))
        return {n: getattr(st, n) for n, _ in StatsC._fields_}

    def close(self):
        if self.h:
            self.lib.ygpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_handle = None


def default_handle() -> Handle:
    global _default_handle
    if _default_handle is None:
        _default_handle = Handle(int(os.environ.get("YGPU_DEVICE", "0")))
    return _default_handle


def ptr(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)
