"""Batched persistent form (cosmos_batch_* in include/cosmos_fv.h): many
independent Newtonian / special-relativistic problems, 1D and 2D, of uneven
sizes, advanced together in one captured hipGraph launch.

A problem is described by a dict with keys physics, dim, nx, ny (2D),
cell_spacing, cfl, final_time (default inf: run max_steps steps), max_steps and
adiabatic_index; its state is a numpy array of primitives shaped as for
Solver.upload ((ny, nx, 5) in 2D, (nx, 5) in 1D).
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Sequence

import numpy as np

from . import native
from .solver import VANLEER


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_schedule(physics: int) -> Dict[str, int]:
    s = native.BatchSchedule()
    native.check(native.load().cosmos_batch_default_schedule(physics, ctypes.byref(s)))
    return {k: getattr(s, k) for k, _ in native.BatchSchedule._fields_}


class Batch:
    def __init__(self, problems: Sequence[dict], order: int = 2, bias: float = 0.0, slope_limiter: int = VANLEER):
        self._lib = native.load()
        self.problems = [dict(p) for p in problems]
        arr = (native.BatchProblem * len(problems))()
        for k, p in enumerate(self.problems):
            p.setdefault("ny", 1)
            p.setdefault("final_time", math.inf)
            arr[k].physics = p["physics"]
            arr[k].dim = p["dim"]
            arr[k].nx = p["nx"]
            arr[k].ny = p["ny"] if p["dim"] == 2 else 1
            arr[k].cell_spacing = p["cell_spacing"]
            arr[k].cfl = p["cfl"]
            arr[k].final_time = p["final_time"]
            arr[k].max_steps = p["max_steps"]
            arr[k].adiabatic_index = p["adiabatic_index"]
        h = ctypes.c_void_p()
        native.check(self._lib.cosmos_batch_create(ctypes.byref(h), arr, len(problems), order, bias, slope_limiter),
                     None, batch=True)
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.cosmos_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self.problems)

    def _check(self, rc: int) -> None:
        native.check(rc, self._h, batch=True)

    def shape(self, k: int):
        p = self.problems[k]
        return (p["ny"], p["nx"], 5) if p["dim"] == 2 else (p["nx"], 5)

    def set_con2prim(self, tolerance: float, max_iterations: int) -> None:
        self._check(self._lib.cosmos_batch_set_con2prim(self._h, tolerance, max_iterations))

    def set_schedule(self, physics: int, **fields) -> None:
        cur = self.schedule(physics)
        cur.update(fields)
        s = native.BatchSchedule(**cur)
        self._check(self._lib.cosmos_batch_set_schedule(self._h, physics, ctypes.byref(s)))

    def schedule(self, physics: int) -> Dict[str, int]:
        s = native.BatchSchedule()
        self._check(self._lib.cosmos_batch_get_schedule(self._h, physics, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in native.BatchSchedule._fields_}

    def upload(self, k: int, prim: np.ndarray) -> None:
        a = np.ascontiguousarray(prim, dtype=np.float64)
        if a.shape != self.shape(k):
            raise ValueError(f"problem {k}: state shape {a.shape} != {self.shape(k)}")
        self._check(self._lib.cosmos_batch_upload(self._h, k, _dp(a)))

    def download(self, k: int) -> np.ndarray:
        out = np.empty(self.shape(k))
        self._check(self._lib.cosmos_batch_download(self._h, k, _dp(out)))
        return out

    def run(self) -> None:
        self._check(self._lib.cosmos_batch_run(self._h))

    def run_timed(self) -> float:
        ms = ctypes.c_double()
        self._check(self._lib.cosmos_batch_run_timed(self._h, ctypes.byref(ms)))
        return ms.value

    def stats(self):
        """(time reached, steps taken, completion time in us after the run started) per problem."""
        n = len(self.problems)
        t, fin = np.empty(n), np.empty(n)
        steps = np.empty(n, dtype=np.int32)
        self._check(self._lib.cosmos_batch_stats(self._h, _dp(t), steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                 _dp(fin)))
        return t, steps, fin

    def autotune(self, trial_steps: int = 10) -> float:
        ms = ctypes.c_double()
        self._check(self._lib.cosmos_batch_autotune(self._h, trial_steps, ctypes.byref(ms)))
        return ms.value


def cell_updates(problems: List[dict], steps: np.ndarray) -> float:
    """Cells times steps taken, summed over the problems."""
    return float(sum(p["nx"] * (p["ny"] if p["dim"] == 2 else 1) * int(s) for p, s in zip(problems, steps)))
