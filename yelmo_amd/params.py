"""Solver parameters: ssa_param_class / diva_param_class as a dataclass.

Field names follow ssa_param_class (src/physics/velocity_ssa.f90:22-52) and
the &ydyn namelist (input/yelmo_defaults.nml:140-173); strings that the
reference switches on become the enum codes of include/ygpu.h.
"""
from __future__ import annotations

from dataclasses import dataclass, fields, replace

from . import _lib

DEFAULT_LIS_OPT = "-i bicgstab -p ilu -maxiter 1000 -tol 1.0e-8 -initx_zeros false"


ENERGY_LIS_OPT = "-i cg -p jacobi -maxiter 200 -tol 1.0e-2 -initx_zeros false"


@dataclass
class SsaParams:
    """Defaults are the benchmark workloads' settings (residual assembly,
    BiCGSTAB + ILU(0) at rtol 1e-8, Picard tolerance 1e-4, power-law
    friction q = 0.2); ``SsaParams.yelmo_defaults()`` gives the &ydyn
    defaults of the reference."""
    ssa_solver: str = "residual"
    ssa_lis_opt: str = DEFAULT_LIS_OPT
    boundaries: str = "zeros"
    no_slip: bool = False
    visc_method: int = 1
    visc_const: float = 1e7
    beta_method: int = 2
    beta_const: float = 1e3
    beta_q: float = 0.2
    beta_u0: float = 1.0
    beta_gl_scale: int = 0
    beta_gl_stag: int = 1
    beta_gl_f: float = 1.0
    H_grnd_lim: float = 500.0
    beta_min: float = 100.0
    eps_0: float = 1e-6
    ssa_vel_lim_method: str = "drag"
    ssa_vel_lim_tau: float = 1e5
    ssa_vel_max: float = 1e4
    ssa_iter_max: int = 50
    ssa_iter_rel: float = 0.7
    ssa_iter_conv: float = 1e-4
    ssa_write_log: bool = False
    rho_ice: float = 910.0
    rho_sw: float = 1028.0
    g: float = 9.81
    # arithmetic of the GPU solve: "fp64", "ref" (FP32 fields + FP64 solve), "mixed"
    precision: str = "fp64"
    block: int = 0

    def with_(self, **kw) -> "SsaParams":
        return replace(self, **kw)

    @classmethod
    def yelmo_defaults(cls, **kw) -> "SsaParams":
        """The &ydyn defaults (input/yelmo_defaults.nml:140-173): energy
        assembly with CG + Jacobi at rtol 1e-2, linear friction
        (beta_method 1, q 1, u0 100), Picard tolerance 1e-2 within 20
        iterations."""
        base = cls(ssa_solver="energy", ssa_lis_opt=ENERGY_LIS_OPT, beta_method=1, beta_q=1.0,
                   beta_u0=100.0, ssa_iter_max=20, ssa_iter_conv=1e-2)
        return replace(base, **kw)

    def to_c(self) -> _lib.SsaParamsC:
        c = _lib.SsaParamsC()
        c.ssa_solver = {"residual": 0, "energy": 1}[self.ssa_solver]
        c.boundaries = _lib.BND[self.boundaries]
        c.visc_method = self.visc_method
        c.beta_method = self.beta_method
        c.beta_gl_scale = self.beta_gl_scale
        c.beta_gl_stag = self.beta_gl_stag
        c.ssa_vel_lim_method = {"clip": 0, "drag": 1}[self.ssa_vel_lim_method]
        c.ssa_iter_max = self.ssa_iter_max
        c.ssa_write_log = int(self.ssa_write_log)
        c.precision = _lib.PREC[self.precision]
        for n in ("visc_const", "beta_const", "beta_q", "beta_u0", "beta_gl_f", "H_grnd_lim",
                  "beta_min", "eps_0", "ssa_vel_lim_tau", "ssa_vel_max", "ssa_iter_rel",
                  "ssa_iter_conv", "rho_ice", "rho_sw", "g"):
            setattr(c, n, float(getattr(self, n)))
        c.no_slip = int(self.no_slip)
        c.block = int(self.block)
        c.ssa_lis_opt = self.ssa_lis_opt.encode()
        return c


def benchmark_params(**kw) -> SsaParams:
    """Solver settings of the benchmark workloads: residual assembly,
    BiCGSTAB + ILU(0) at rtol 1e-8, Picard relaxation 0.7, tolerance 1e-4, at
    most 50 iterations, power-law friction q = 0.2, u0 = 1 m/yr, limit drag."""
    return SsaParams(**kw)
