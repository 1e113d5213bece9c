"""Host-side mirror of Yelmo's velocity-solver interface.

``calc_velocity_ssa`` / ``calc_velocity_diva`` take the same fields as
``calc_velocity_ssa`` (src/physics/velocity_ssa.f90:60-63) and
``calc_velocity_diva`` (src/physics/velocity_diva.f90:59-64), held in a
``generator.SsaProblem``, and run the whole Picard loop on the GPU through
libygpu.so.  ``linear_solver_matrix_solve`` mirrors the reference's linear
solver boundary (src/physics/solver_linear.F90:151) on an ``Lgs`` holder of
the 1-based CSR arrays (``linear_solver_class``, solver_linear.F90:11-34).
Every call goes to the HIP library; there is no CPU path here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .generator import SsaProblem
from .params import SsaParams

LGS_SUCCESS, LGS_BREAKDOWN, LGS_MAXITER = 0, 2, 4


@dataclass
class SolveInfo:
    ssa_iter_now: int
    ssa_lin_iter: int
    ssa_lin_fail: int
    ssa_lim_n: int
    status: int
    picard_resid: float
    seconds: float


def _real_dtype(prob: SsaProblem):
    dt = prob.fields["H_ice"].dtype
    if dt not in (np.float32, np.float64):
        raise TypeError("fields must be float32 or float64")
    return dt


def _fill_struct(prob: SsaProblem, diva: bool) -> tuple[_lib.SsaProblemC, list]:
    dt = _real_dtype(prob)
    keep = []

    def arr(a, dtype=dt):
        a = np.asfortranarray(a, dtype=dtype)
        keep.append(a)
        return a

    s = _lib.SsaProblemC()
    s.nx, s.ny, s.nz = prob.nx, prob.ny, prob.nz
    s.real_bytes = np.dtype(dt).itemsize
    s.dx, s.dy, s.n_glen = prob.dx, prob.dy, prob.n_glen
    st, fl = prob.state, prob.fields
    for n in _lib.PROBLEM_PTRS_INOUT + _lib.PROBLEM_PTRS_ERR:
        st[n] = arr(st[n])
        setattr(s, n, st[n].ctypes.data)
    # fields and masks set to None pass NULL (what the library derives
    # itself, e.g. under -ygpu_inputs device)
    for n in ("ssa_mask_acx", "ssa_mask_acy"):
        if st.get(n) is None:
            setattr(s, n, None)
            continue
        st[n] = arr(st[n], np.int32)
        setattr(s, n, st[n].ctypes.data)
    for n in _lib.PROBLEM_PTRS_IN:
        if fl.get(n) is None:
            setattr(s, n, None)
            continue
        a = arr(fl[n])
        setattr(s, n, a.ctypes.data)
    if diva:
        nx, ny, nz = prob.nx, prob.ny, prob.nz
        for n, shape in (("ux", (nx, ny, nz)), ("uy", (nx, ny, nz)), ("ux_bar", (nx, ny)),
                         ("uy_bar", (nx, ny)), ("ux_i", (nx, ny, nz)), ("uy_i", (nx, ny, nz)),
                         ("beta_eff", (nx, ny)), ("duxdz", (nx, ny, nz)), ("duydz", (nx, ny, nz))):
            if n not in st:
                st[n] = np.zeros(shape, dtype=dt, order="F")
            st[n] = arr(st[n])
            setattr(s, n, st[n].ctypes.data)
    return s, keep


def _info(s: _lib.SsaProblemC) -> SolveInfo:
    return SolveInfo(s.ssa_iter_now, s.ssa_lin_iter, s.ssa_lin_fail, s.ssa_lim_n, s.status,
                     s.picard_resid, s.solve_seconds)


def _run_batch(problems, par: SsaParams, diva: bool, handle=None, hooks=None):
    h = handle or _lib.default_handle()
    arr = (_lib.SsaProblemC * len(problems))()
    keep = []
    for k, p in enumerate(problems):
        hook = hooks[k] if hooks is not None else None
        if hook is not None:
            # c_bed is INOUT under a hook: the caller's array receives its updates
            p.fields["c_bed"] = np.asfortranarray(p.fields["c_bed"], dtype=_real_dtype(p))
        s, kp = _fill_struct(p, diva)
        if hook is not None:
            s.neff_hook = C.cast(hook, C.c_void_p)
            kp.append(hook)
        arr[k] = s
        keep.append(kp)
    pc = par.to_c()
    fn = h.lib.ygpu_velocity_diva_batch if diva else h.lib.ygpu_velocity_ssa_batch
    h.check(fn(h.h, C.byref(pc), len(problems), arr))
    return [_info(arr[k]) for k in range(len(problems))]


def calc_velocity_ssa(prob: SsaProblem, par: SsaParams, handle=None) -> SolveInfo:
    """One SSA solve (calc_velocity_ssa); updates prob.state in place."""
    return _run_batch([prob], par, False, handle)[0]


def calc_velocity_ssa_batch(problems, par: SsaParams, handle=None) -> list[SolveInfo]:
    """Independent SSA solves of many domains (uneven sizes allowed) in one
    device call; each member runs its own Picard loop."""
    return _run_batch(problems, par, False, handle)


def calc_velocity_diva(prob: SsaProblem, par: SsaParams, handle=None,
                       neff_hook=None) -> SolveInfo:
    """One DIVA solve (calc_velocity_diva); updates prob.state in place.
    neff_hook (an _lib.NEFF_HOOK) is the reference's optional hook: it
    updates c_bed (prob.fields["c_bed"], in place) from u_b after every
    Picard iteration."""
    return _run_batch([prob], par, True, handle, hooks=[neff_hook])[0]


# ---- the outer drop-ins, argument by argument ---------------------------------

_SSA_ARGS = ("ux_b", "uy_b", "taub_acx", "taub_acy", "visc_eff", "visc_eff_int", "ssa_mask_acx",
             "ssa_mask_acy", "ssa_err_acx", "ssa_err_acy", "#ssa_iter_now", "#ssa_lin_iter",
             "#ssa_lin_fail", "#ssa_lim_n", "beta", "beta_acx", "beta_acy", "c_bed", "f_slide",
             "taud_acx", "taud_acy", "taul_int_acx", "taul_int_acy", "H_ice", "f_ice", "H_grnd",
             "f_grnd", "f_grnd_acx", "f_grnd_acy", "ATT", "zeta_aa", "z_sl", "z_bed", "z_srf")
_DIVA_ARGS = ("ux", "uy", "ux_bar", "uy_bar", "ux_b", "uy_b", "ux_i", "uy_i", "taub_acx",
              "taub_acy", "beta", "beta_acx", "beta_acy", "beta_eff", "visc_eff", "visc_eff_int",
              "duxdz", "duydz", "ssa_mask_acx", "ssa_mask_acy", "ssa_err_acx", "ssa_err_acy",
              "#ssa_iter_now", "#ssa_lin_iter", "#ssa_lin_fail", "#ssa_lim_n", "c_bed", "f_slide",
              "taud_acx", "taud_acy", "taul_int_acx", "taul_int_acy", "H_ice", "f_ice", "H_grnd",
              "f_grnd", "f_grnd_acx", "f_grnd_acy", "ATT", "zeta_aa", "z_sl", "z_bed", "z_srf")


def _drop_in_args(prob: SsaProblem, names):
    """float32 / int32 Fortran arrays of prob in the order of `names`
    (the prob's own arrays are replaced by them, so results land there);
    '#name' is an integer scalar output."""
    st, fl = prob.state, prob.fields
    outs, ptrs = {}, []
    for n in names:
        if n.startswith("#"):
            outs[n[1:]] = C.c_int32()
            ptrs.append(C.cast(C.byref(outs[n[1:]]), C.c_void_p))
            continue
        d = st if n in st else fl
        dt = np.int32 if n.startswith("ssa_mask") else np.float32
        d[n] = np.asfortranarray(d[n], dtype=dt)
        ptrs.append(C.c_void_p(d[n].ctypes.data))
    return outs, ptrs


def ygpu_velocity_ssa(prob: SsaProblem, par: SsaParams, handle=None) -> SolveInfo:
    """calc_velocity_ssa through the argument-order drop-in ygpu_velocity_ssa
    (wp = sp arrays: prob's arrays become float32)."""
    h = handle or _lib.default_handle()
    outs, ptrs = _drop_in_args(prob, _SSA_ARGS)
    pc = par.to_c()
    rc = h.lib.ygpu_velocity_ssa(h.h, C.byref(pc), prob.nx, prob.ny, prob.nz, *ptrs,
                                 prob.dx, prob.dy, prob.n_glen)
    if rc not in (_lib.YGPU_OK, _lib.YGPU_ERR_NONFINITE):
        h.check(rc)
    return SolveInfo(outs["ssa_iter_now"].value, outs["ssa_lin_iter"].value,
                     outs["ssa_lin_fail"].value, outs["ssa_lim_n"].value, rc, float("nan"), 0.0)


def ygpu_velocity_diva(prob: SsaProblem, par: SsaParams, handle=None,
                       neff_hook=None) -> SolveInfo:
    """calc_velocity_diva through the argument-order drop-in
    ygpu_velocity_diva (wp = sp arrays); neff_hook as in calc_velocity_diva."""
    from .generator import add_diva_state
    h = handle or _lib.default_handle()
    add_diva_state(prob)
    outs, ptrs = _drop_in_args(prob, _DIVA_ARGS)
    pc = par.to_c()
    hook = neff_hook if neff_hook is not None else _lib.NEFF_HOOK()
    rc = h.lib.ygpu_velocity_diva(h.h, C.byref(pc), prob.nx, prob.ny, prob.nz, *ptrs,
                                  prob.dx, prob.dy, prob.n_glen, hook, None)
    if rc not in (_lib.YGPU_OK, _lib.YGPU_ERR_NONFINITE):
        h.check(rc)
    return SolveInfo(outs["ssa_iter_now"].value, outs["ssa_lin_iter"].value,
                     outs["ssa_lin_fail"].value, outs["ssa_lim_n"].value, rc, float("nan"), 0.0)


def calc_velocity_diva_batch(problems, par: SsaParams, handle=None) -> list[SolveInfo]:
    return _run_batch(problems, par, True, handle)


# ---- the inner boundary ------------------------------------------------------

@dataclass
class Lgs:
    """linear_solver_class (solver_linear.F90:11-34): 1-based CSR."""
    a_ptr: np.ndarray
    a_index: np.ndarray
    a_value: np.ndarray
    b_value: np.ndarray
    x_value: np.ndarray
    lin_iter: int = 0
    status: int = 0
    L2_rel_norm: float = 0.0
    solver_time: float = 0.0

    @property
    def nmax(self) -> int:
        return len(self.b_value)


def csr_spmv(a_ptr, a_index, a_value, x, opts: str = "", handle=None) -> np.ndarray:
    """y = A x on the device for 1-based CSR arrays of the C-grid stencil
    (lis_matvec at the linear-solver boundary)."""
    h = handle or _lib.default_handle()
    ap = np.ascontiguousarray(a_ptr, dtype=np.int32)
    ai = np.ascontiguousarray(a_index, dtype=np.int32)
    av = np.ascontiguousarray(a_value, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(xx)
    n = len(ap) - 1
    h.check(h.lib.ygpu_csr_spmv(h.h, n, int(ap[-1] - 1), ap.ctypes.data, ai.ctypes.data,
                                av.ctypes.data, xx.ctypes.data, y.ctypes.data, opts.encode()))
    return y


def linear_solver_matrix_solve(lgs: Lgs, lis_settings: str, handle=None) -> Lgs:
    """linear_solver_matrix_solve (solver_linear.F90:151): solves in place on
    the GPU, LIS option grammar, status 0/2/4."""
    h = handle or _lib.default_handle()
    n = lgs.nmax
    nnz = int(lgs.a_ptr[n]) - 1
    ap = np.ascontiguousarray(lgs.a_ptr[: n + 1], dtype=np.int32)
    ai = np.ascontiguousarray(lgs.a_index[:nnz], dtype=np.int32)
    av = np.ascontiguousarray(lgs.a_value[:nnz], dtype=np.float64)
    b = np.ascontiguousarray(lgs.b_value, dtype=np.float64)
    x = np.ascontiguousarray(lgs.x_value, dtype=np.float64).copy()
    it, st = C.c_int32(), C.c_int32()
    res, secs = C.c_double(), C.c_double()
    rc = h.lib.ygpu_csr_solve(h.h, n, nnz, ap.ctypes.data, ai.ctypes.data, av.ctypes.data,
                              b.ctypes.data, x.ctypes.data, lis_settings.encode(), C.byref(it),
                              C.byref(st), C.byref(res), C.byref(secs))
    h.check(rc)
    lgs.x_value = x
    lgs.lin_iter, lgs.status = it.value, st.value
    lgs.L2_rel_norm, lgs.solver_time = res.value, secs.value
    return lgs


class ResidentBatch:
    """A batch of domains kept on the device (ygpu_batch_*): inputs are
    uploaded once; ``run()`` solves all domains from the resident initial
    state and returns the device seconds; ``download()`` writes the results
    back into the problems' state arrays."""

    def __init__(self, problems, par: SsaParams, diva: bool = False, handle=None):
        self.h = handle or _lib.default_handle()
        self.problems = problems
        self.arr = (_lib.SsaProblemC * len(problems))()
        self._keep = []
        for k, p in enumerate(problems):
            s, kp = _fill_struct(p, diva)
            self.arr[k] = s
            self._keep.append(kp)
        self.pc = par.to_c()
        b = C.c_void_p()
        nbytes = C.c_int64()
        self.h.check(self.h.lib.ygpu_batch_create(self.h.h, C.byref(self.pc), len(problems),
                                                  self.arr, int(diva), C.byref(b),
                                                  C.byref(nbytes)))
        self.b = b
        self.device_bytes = nbytes.value

    def run(self, reset: bool = True) -> float:
        secs = C.c_double()
        self.h.check(self.h.lib.ygpu_batch_run(self.b, int(reset), C.byref(secs)))
        return secs.value

    def download(self) -> list[SolveInfo]:
        self.h.check(self.h.lib.ygpu_batch_download(self.b, self.arr))
        return [_info(self.arr[k]) for k in range(len(self.problems))]

    def spmv(self, fp32: bool = False, reps: int = 10, want_y: bool = False):
        """y = A x of every domain's last assembled system (x: its last
        solution) in one launch; returns (y or None, ms per launch,
        algorithmic bytes per launch)."""
        n = 2 * sum(p.nx * p.ny for p in self.problems)
        y = np.zeros(n) if want_y else None
        ms, nbytes = C.c_double(), C.c_double()
        self.h.check(self.h.lib.ygpu_batch_spmv(self.b, int(fp32), int(reps),
                                                 None if y is None else y.ctypes.data,
                                                 C.byref(ms), C.byref(nbytes)))
        return y, ms.value, nbytes.value

    def close(self):
        if self.b:
            self.h.lib.ygpu_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Session:
    """One domain kept on the device between solves (ygpu_session_*): the
    velocity, viscosity and friction of a solve are the warm start of the
    next; ``solve(prob)`` uploads prob's IN fields (its INOUT state only on
    the first solve) and writes the results into prob.state."""

    def __init__(self, prob: SsaProblem, par: SsaParams, handle=None):
        self.h = handle or _lib.default_handle()
        self.pc = par.to_c()
        s, keep = _fill_struct(prob, False)
        ss = C.c_void_p()
        nbytes = C.c_int64()
        self.h.check(self.h.lib.ygpu_session_create(self.h.h, C.byref(self.pc), C.byref(s),
                                                    C.byref(ss), C.byref(nbytes)))
        self.s = ss
        self.device_bytes = nbytes.value

    def solve(self, prob: SsaProblem) -> SolveInfo:
        s, keep = _fill_struct(prob, False)
        self.h.check(self.h.lib.ygpu_session_solve(self.s, C.byref(s)))
        return _info(s)

    def close(self):
        if self.s:
            self.h.lib.ygpu_session_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


LATERAL_BC = {"none": 0, "floating": 1, "float": 1, "marine": 2, "all": 3}


def calc_dyn_inputs(fields: dict, par: SsaParams, dx: float, dy: float, taud_lim: float = 5e5,
                    grad_lim: float = 0.1, lateral_bc: str = "all", slope_bg=(0.0, 0.0),
                    handle=None) -> dict:
    """The per-solve inputs of the velocity solve on the device
    (``ygpu_dyn_inputs``): what calc_ydyn computes before the call
    (yelmo_dynamics.f90:105-122,464) from the geometry in ``fields`` (H_ice,
    f_ice, f_grnd, z_bed, z_sl, z_srf; float32, Fortran order).  Returns
    mask_frnt, dzsdx, dzsdy, taud_acx/acy, taul_int_acx/acy and
    ssa_mask_acx/acy."""
    h = handle or _lib.default_handle()
    nx, ny = fields["H_ice"].shape
    ins = [np.asfortranarray(fields[k], dtype=np.float32)
           for k in ("H_ice", "f_ice", "f_grnd", "z_bed", "z_sl", "z_srf")]
    f = lambda: np.zeros((nx, ny), dtype=np.float32, order="F")
    i = lambda: np.zeros((nx, ny), dtype=np.int32, order="F")
    out = dict(mask_frnt=i(), dzsdx=f(), dzsdy=f(), taud_acx=f(), taud_acy=f(),
               taul_int_acx=f(), taul_int_acy=f(), ssa_mask_acx=i(), ssa_mask_acy=i())
    pc = par.to_c()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    h.check(h.lib.ygpu_dyn_inputs(h.h, C.byref(pc), nx, ny, dx, dy, taud_lim, grad_lim,
                                  slope_bg[0], slope_bg[1], LATERAL_BC[lateral_bc],
                                  *[ptr(a) for a in ins], *[ptr(a) for a in out.values()]))
    return out
