"""Autotuner of the mixed-precision Krylov schedule (ygpu_autotune).

``autotune(prob, par)`` times the Picard solve (its first ten iterations) of
a device-filling batch of copies of ``prob`` for each candidate schedule
(ILU(0) line blocks, threads per member block; for one workgroup per member
and for each group size a batch may pick) and records the fastest in the
handle per grid size and group size; every batch, session and server created
on the handle later uses it for members of that size.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

from . import _lib
from .generator import SsaProblem
from .params import SsaParams
from .velocity import _fill_struct


@dataclass
class Schedule:
    ilu_blocks: int
    block: int
    solve_ms: float


def autotune(prob: SsaProblem, par: SsaParams, *, diva: bool = False, reps: int = 2,
             handle=None) -> Schedule:
    h = handle or _lib.default_handle()
    s, keep = _fill_struct(prob.copy(), diva)
    pc = par.to_c()
    nb, blk, ms = C.c_int32(), C.c_int32(), C.c_double()
    h.check(h.lib.ygpu_autotune(h.h, C.byref(pc), C.byref(s), int(diva), int(reps),
                                C.byref(nb), C.byref(blk), C.byref(ms)))
    return Schedule(nb.value, blk.value, ms.value)


def autotune_sizes(problems, par: SsaParams, *, diva: bool = False, reps: int = 2,
                   handle=None, progress=None) -> dict:
    """Tunes once per distinct grid size among ``problems``; ``progress``
    (a file, e.g. sys.stderr) gets one line per tuned size."""
    out = {}
    for p in problems:
        if (p.nx, p.ny) not in out:
            t0 = time.perf_counter()
            out[(p.nx, p.ny)] = autotune(p, par, diva=diva, reps=reps, handle=handle)
            if progress is not None:
                print(f"# tuned {p.nx}x{p.ny} in {time.perf_counter() - t0:.1f}s", file=progress,
                      flush=True)
    return out


def set_schedule(nx: int, ny: int, ilu_blocks: int, block: int = 0, handle=None):
    h = handle or _lib.default_handle()
    h.check(h.lib.ygpu_set_schedule(h.h, nx, ny, ilu_blocks, block))
