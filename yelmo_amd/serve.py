"""Served solve sequences with continuous batching (ygpu_server_*).

A ``Server`` takes solves as they arrive (``submit``), admits them into
device slots as slots and HBM free up, and advances every admitted solve by
one Picard iteration per ``step()``; finished solves are reported by
``poll()`` and their results copied out by ``fetch()``.  Each request is a
sequence of solves: the first carries the INOUT state, later ones
(``cont=True``) start from the previous solve's device-resident velocity,
viscosity and friction, and ``last=True`` frees the sequence's slot when
that solve finishes.  The reference solves one domain per call
(calc_velocity_ssa, src/physics/velocity_ssa.f90:60-357); a sequence here
gives the same results as those calls made one after another with each
call's outputs as the next call's INOUT arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .generator import SsaProblem
from .params import SsaParams
from .velocity import SolveInfo, _fill_struct, _info


class Server:
    """``lanes``: (max side, slots) per size class, sides increasing; each
    lane steps on its own stream.  A single (max_side, max_slots) lane is
    the default."""

    def __init__(self, par: SsaParams, *, diva: bool = False, max_slots: int = 64,
                 max_side: int = 641, nz_max: int = 11, hbm_budget: int = 8 << 30,
                 lanes=None, handle=None):
        self.h = handle or _lib.default_handle()
        self.diva = diva
        self.pc = par.to_c()
        lanes = list(lanes) if lanes else [(max_side, max_slots)]
        side = np.array([int(a) for a, _ in lanes], dtype=np.int32)
        slots = np.array([int(b) for _, b in lanes], dtype=np.int32)
        srv = C.c_void_p()
        nbytes = C.c_int64()
        self.h.check(self.h.lib.ygpu_server_create(self.h.h, C.byref(self.pc), int(diva),
                                                   len(lanes), side.ctypes.data,
                                                   slots.ctypes.data, nz_max, int(hbm_budget),
                                                   C.byref(srv), C.byref(nbytes)))
        self.s = srv
        self.device_bytes = nbytes.value

    def submit(self, prob: SsaProblem, seq: int, *, cont: bool = False,
               last: bool = False) -> int:
        """Queues one solve of sequence ``seq``; returns its ticket.  The
        arrays are copied before this returns."""
        s, keep = _fill_struct(prob, self.diva)
        flags = (_lib.YGPU_SEQ_CONTINUE if cont else 0) | (_lib.YGPU_SEQ_LAST if last else 0)
        t = C.c_int64()
        self.h.check(self.h.lib.ygpu_submit(self.s, C.byref(s), int(seq), flags, C.byref(t)))
        return t.value

    def step(self) -> tuple[int, int]:
        """Admits what fits and runs one Picard iteration of every admitted
        solve; returns (solves running, solves queued)."""
        nr, nq = C.c_int32(), C.c_int32()
        self.h.check(self.h.lib.ygpu_server_step(self.s, C.byref(nr), C.byref(nq)))
        return nr.value, nq.value

    def poll(self) -> list[int]:
        """Tickets finished since the last poll, in completion order."""
        out = []
        t = C.c_int64()
        while True:
            rc = self.h.lib.ygpu_poll(self.s, C.byref(t))
            if rc == 0:
                return out
            if rc != 1:
                self.h.check(rc)
            out.append(t.value)

    def fetch(self, ticket: int, prob: SsaProblem) -> SolveInfo:
        """Copies a finished solve's results into prob.state."""
        s, keep = _fill_struct(prob, self.diva)
        self.h.check(self.h.lib.ygpu_fetch(self.s, int(ticket), C.byref(s)))
        return _info(s)

    def close(self):
        if self.s:
            self.h.lib.ygpu_server_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
