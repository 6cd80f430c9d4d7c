"""Many independent 3-D acoustic problems (uneven mesh sizes and orders)
advanced together through ``mjw_batch_*``: one stiffness launch per colour
over every problem's elements, one tail block per problem, the step loop
captured in a HIP graph."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MjwError, dptr
from .problem import WaveProblem


class Batch:
    def __init__(self, problems: list[WaveProblem]):
        self.problems = list(problems)
        arr = (C.c_void_p * len(problems))(*[p.handle.value for p in problems])
        self._h = C.c_void_p()
        rc = _lib.lib().mjw_batch_create(C.byref(self._h), len(problems), arr)
        if rc != 0:
            raise MjwError(f"mjw_batch_create failed ({rc}): "
                           f"{_lib.lib().mjw_last_error(None).decode()}")
        self.last_nsteps = 0

    def _chk(self, rc, what):
        if rc != 0:
            raise MjwError(f"{what} failed ({rc}): "
                           f"{_lib.lib().mjw_batch_last_error(self._h).decode()}")

    def config(self, order_mode: int = 0):
        self._chk(_lib.lib().mjw_batch_config(self._h, 8, order_mode), "batch_config")

    def advance(self, nsteps: int, graph: bool = True):
        self._chk(_lib.lib().mjw_batch_advance(self._h, nsteps, int(bool(graph))), "batch_advance")
        self.last_nsteps = nsteps

    def traces(self, i: int) -> np.ndarray:
        P = self.problems[i]
        out = np.zeros((self.last_nsteps, P.nrec))
        self._chk(_lib.lib().mjw_batch_traces(self._h, i, dptr(out)), "batch_traces")
        return out

    def ms_per_step(self) -> float:
        v = C.c_double()
        self._chk(_lib.lib().mjw_batch_timing(self._h, C.byref(v)), "batch_timing")
        return v.value

    def close(self):
        if self._h:
            _lib.lib().mjw_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
