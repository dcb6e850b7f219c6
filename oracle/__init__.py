"""CPU checkers for the finite-volume path (TEST INFRASTRUCTURE ONLY).

    load_oracle()  the C restatement, oracle/liboracle.so (orc_* functions)
    load_ref()     the reference's own sources built by oracle/Makefile,
                   oracle/_ref/libcosmos_ref.so (ref_* functions)
    load_ref_omp() the same sources with OpenMP on, _ref/libcosmos_ref_omp.so
                   (timing only: its parallel loops race)

Both expose the same signatures (see cosmos_oracle.h); wrap() gives them a
common numpy-facing interface.  Only tests, the golden-fixture script,
__graft_entry__.smoke() and the CPU-baseline leg of bench.py use this module,
as the checker; the product path never does.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcosmos_ref.so")
REF_OMP_SO = os.path.join(HERE, "_ref", "libcosmos_ref_omp.so")
REFERENCE_ROOT = "/root/reference"

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "sweep": (_I, [_I, _I, _I, _I, _P, _I, _I, _D, _D, _D, _I, _D]),
    "max_wavespeed": (_D, [_I, _I, _P, _I, _I, _D]),
    "stable_dt": (_D, [_I, _I, _P, _I, _I, _D, _D, _D, _D, _I, _D]),
    "stable_dt_scalar": (_D, [_D, _D, _D, _I]),
    "run": (_I, [_I, _I, _I, _P, _I, _I, _D, _D, _D, _I, _D, _I, _D, _P]),
    "slic_flux": (None, [_I, _I, _P, _P, _P, _P, _D, _D, _D, _I, _D, _P]),
    "force_flux": (None, [_I, _I, _P, _P, _D, _D, _D, _P]),
    "half_step": (None, [_I, _I, _P, _P, _P, _D, _D, _D, _I, _I, _D, _P]),
    "p2c": (None, [_I, _P, _P, _D]),
    "c2p": (None, [_I, _P, _P, _D]),
    "flux": (None, [_I, _I, _P, _P, _D]),
    "limiter": (_D, [_D, _D, _I]),
    "slope": (None, [_P, _P, _P, _D, _I, _P]),
    # general relativistic (black hole M, a at (bx, by, bz))
    "gr_sweep": (_I, [_I, _I, _I, _P, _I, _I, _D, _D, _D, _I, _D, _D, _D, _D, _D, _D]),
    "gr_forcing": (_I, [_I, _P, _I, _I, _D, _D, _D, _I, _D, _D, _D, _D, _D, _D]),
    "gr_max_wavespeed": (_D, [_I, _P, _I, _I, _D, _D, _D, _D, _D, _D]),
    "gr_run": (_I, [_I, _I, _P, _I, _I, _D, _D, _D, _I, _D, _I, _I, _D, _D, _D, _D, _D, _D, _P]),
    "gr_point": (None, [_I, _P, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _P]),
}

#: what-codes of gr_point
GR_P2C, GR_C2P, GR_XFLUX, GR_YFLUX, GR_SOURCE, GR_SOUND, GR_METRIC = range(7)
# point-level statics (ORC_PT_* of cosmos_oracle.h)
(PT_P2C, PT_C2P, PT_FLUX, PT_LF, PT_RICHTMYER, PT_FORCE, PT_SLIC, PT_HALF, PT_SLOPE, PT_FRACTIONAL, PT_FORCING,
 PT_VEC_LF, PT_VEC_RICHTMYER, PT_VEC_FORCE, PT_VEC_UPDATE, PT_VEC_EVOLUTION, PT_VEC_SLOPE, PT_LIMITER) = range(18)


def build(ref: bool = True) -> None:
    """Build liboracle.so, and _ref/ when the reference sources are present."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir(REFERENCE_ROOT):
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)


def _dp(a):
    return a.ctypes.data_as(_P)


class CpuPath:
    """numpy-facing wrapper of liboracle.so (prefix orc_) or the reference build (ref_)."""

    def __init__(self, path: str, prefix: str):
        self.path, self.prefix = path, prefix
        self.lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(self.lib, prefix + name)
            fn.restype = res
            fn.argtypes = args
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        if prefix == "ref_" and hasattr(self.lib, "ref_amr_level2"):
            self.lib.ref_amr_level2.restype = _I
            self.lib.ref_amr_level2.argtypes = [_I, _I, _I, _P, _I, _I, _D, _D, _D, _D, _D, _P, _P, _P, u8, u8, u8]
        if prefix == "orc_":
            self.lib.orc_amr_level2.restype = _I
            self.lib.orc_amr_level2.argtypes = [_I, _I, _I, _P, _I, _I, _D, _D, _D, _D, _D, _I, _P, _P, _P, u8, u8,
                                                u8, ctypes.POINTER(_D)]
        if prefix == "ref_" and hasattr(self.lib, "ref_gr_amr_level1"):
            self.lib.ref_gr_amr_level1.restype = _I
            self.lib.ref_gr_amr_level1.argtypes = [_I, _I, _P, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _P,
                                                   _P, u8, u8]
            self.lib.ref_gr_amr_level2.restype = _I
            self.lib.ref_gr_amr_level2.argtypes = [_I, _P, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _P, _P, _P,
                                                   u8, u8, u8]
        if prefix == "ref_" and hasattr(self.lib, "ref_set_threads"):
            self.lib.ref_set_threads.restype = _I
            self.lib.ref_set_threads.argtypes = [_I]
        if prefix == "ref_" and hasattr(self.lib, "ref_amr_level1"):
            self.lib.ref_amr_level1.restype = _I
            self.lib.ref_amr_level1.argtypes = [_I, _I, _I, _P, _I, _I, _D, _D, _D, _D, _D, _P, _P, u8, u8]
        if prefix == "orc_":
            self.lib.orc_amr_refine_flags.restype = None
            self.lib.orc_amr_refine_flags.argtypes = [_I, _I, _P, _I, _I, _D, _D, _D, u8, u8]
            self.lib.orc_amr_level1.restype = _I
            self.lib.orc_amr_level1.argtypes = [_I, _I, _I, _P, _I, _I, _D, _D, _D, _D, _D, _I, _P, _P, u8, u8,
                                                ctypes.POINTER(_D)]
            self.lib.orc_set_c2p.argtypes = [_D, _I]
            self.lib.orc_gr_set_c2p.argtypes = [_D, _I]
            self.lib.orc_gr_run_dts.restype = None
            self.lib.orc_gr_run_dts.argtypes = [_I, _I, _P, _I, _I, _D, _P, _I, _D, _I, _I, _D, _D, _D, _D, _D, _D]
            self.lib.orc_trace_begin.argtypes = [ctypes.POINTER(ctypes.c_ubyte), ctypes.c_long]
            self.lib.orc_trace_end.restype = ctypes.c_long
            self.lib.orc_trace_replay.argtypes = [ctypes.POINTER(ctypes.c_ubyte), ctypes.c_long]
            self.lib.orc_events_read.argtypes = [ctypes.POINTER(ctypes.c_longlong)]
            self.lib.orc_set_threads.argtypes = [_I]
            self.lib.orc_gr_set_excision.argtypes = [_D]
            self.lib.orc_gr_amr_level1.restype = _I
            self.lib.orc_gr_amr_level1.argtypes = [_I, _I, _P, _I, _I, _D, _D, _D, _D, _I, _D, _D, _D, _D, _D, _D,
                                                   _P, _P, u8, u8, ctypes.POINTER(_D)]
            self.lib.orc_gr_amr_level2.restype = _I
            self.lib.orc_gr_amr_level2.argtypes = [_I, _I, _P, _I, _I, _D, _D, _D, _D, _I, _D, _D, _D, _D, _D, _D,
                                                   _P, _P, _P, u8, u8, u8, ctypes.POINTER(_D)]

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    # ---- point-level statics ----
    def point(self, fn, physics, axis, states, dx=1.0, dt=1.0, bias=0.0, frac=0.0, gamma=1.4, limiter=1, side=0,
              hole=None, pos=None):
        """One point-level static (PT_*): states = up to five 5-vectors (primitive
        states, or conserved / flux vectors where the static takes those)."""
        s = np.zeros(25)
        flat = np.ascontiguousarray(states, dtype=np.float64).ravel()
        s[:flat.size] = flat
        par = np.array([dx, dt, bias, frac, gamma], dtype=np.float64)
        st = np.array(hole if hole is not None else (0.0,) * 5, dtype=np.float64)
        p = np.array(pos if pos is not None else (0.0, 0.0, 0.0), dtype=np.float64)
        out = np.full(5, np.nan)
        f = self._f("point")
        f.restype = None
        f.argtypes = [_I, _I, _I, _P, _P, _I, _I, _P, _P, _P]
        f(fn, physics, axis, _dp(s), _dp(par), limiter, side, _dp(st), _dp(p), _dp(out))
        return out[:1] if fn == PT_LIMITER else out

    def insert_boundary(self, W, physics, size):
        """insertBoundaryCells{,2D}(cells, size) of the reference build (ref_ only)."""
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        shape = (nx + 2 * size, 5) if dim == 1 else (ny + 2 * size, nx + 2 * size, 5)
        out = np.full(shape, np.nan)
        f = self.lib.ref_insert_boundary
        f.restype = None
        f.argtypes = [_I, _I, _I, _P, _I, _I, _P]
        f(physics, dim, size, _dp(a), nx, ny, _dp(out))
        return out

    @staticmethod
    def _dims(W):
        if W.ndim == 2:
            return 1, W.shape[0], 1
        return 2, W.shape[1], W.shape[0]

    def sweep(self, W, physics, order, axis, dx, dt, bias=0.0, limiter=1, gamma=1.4):
        a = np.ascontiguousarray(W, dtype=np.float64).copy()
        dim, nx, ny = self._dims(a)
        rc = self._f("sweep")(physics, order, axis, dim, _dp(a), nx, ny, dx, dt, bias, limiter, gamma)
        if rc:
            raise ValueError("sweep rejected its arguments")
        return a

    def max_wavespeed(self, W, physics, gamma=1.4):
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        return self._f("max_wavespeed")(physics, dim, _dp(a), nx, ny, gamma)

    def stable_dt(self, W, physics, dx, cfl, t, tf, iteration, gamma=1.4):
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        return self._f("stable_dt")(physics, dim, _dp(a), nx, ny, dx, cfl, t, tf, iteration, gamma)

    def stable_dt_scalar(self, dt, t, tf, iteration):
        return self._f("stable_dt_scalar")(dt, t, tf, iteration)

    def run(self, W, physics, order, dx, cfl, tf=math.inf, max_steps=2 ** 31 - 1, bias=0.0, limiter=1,
            gamma=1.4):
        a = np.ascontiguousarray(W, dtype=np.float64).copy()
        dim, nx, ny = self._dims(a)
        t = _D()
        n = self._f("run")(physics, order, dim, _dp(a), nx, ny, dx, cfl, tf, max_steps, bias, limiter, gamma,
                           ctypes.byref(t))
        return a, t.value, n

    def slic_flux(self, physics, axis, ll, l, r, rr, dx, dt, bias=0.0, limiter=1, gamma=1.4):
        out = np.empty(5)
        args = [np.ascontiguousarray(x, dtype=np.float64) for x in (ll, l, r, rr)]
        self._f("slic_flux")(physics, axis, *[_dp(x) for x in args], dx, dt, bias, limiter, gamma, _dp(out))
        return out

    def force_flux(self, physics, axis, l, r, dx, dt, gamma=1.4):
        out = np.empty(5)
        a, b = (np.ascontiguousarray(x, dtype=np.float64) for x in (l, r))
        self._f("force_flux")(physics, axis, _dp(a), _dp(b), dx, dt, gamma, _dp(out))
        return out

    def half_step(self, physics, axis, l, m, r, dx, dt, bias, limiter, side, gamma=1.4):
        out = np.empty(5)
        a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (l, m, r))
        self._f("half_step")(physics, axis, _dp(a), _dp(b), _dp(c), dx, dt, bias, limiter, side, gamma, _dp(out))
        return out

    def p2c(self, physics, w, gamma=1.4):
        out = np.empty(5)
        a = np.ascontiguousarray(w, dtype=np.float64)
        self._f("p2c")(physics, _dp(a), _dp(out), gamma)
        return out

    def c2p(self, physics, u, gamma=1.4):
        out = np.empty(5)
        a = np.ascontiguousarray(u, dtype=np.float64)
        self._f("c2p")(physics, _dp(a), _dp(out), gamma)
        return out

    def flux(self, physics, axis, u, gamma=1.4):
        out = np.empty(5)
        a = np.ascontiguousarray(u, dtype=np.float64)
        self._f("flux")(physics, axis, _dp(a), _dp(out), gamma)
        return out

    def limiter(self, r, bias, limiter):
        return self._f("limiter")(r, bias, limiter)

    def slope(self, ul, um, ur, bias, limiter):
        out = np.empty(5)
        a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (ul, um, ur))
        self._f("slope")(_dp(a), _dp(b), _dp(c), bias, limiter, _dp(out))
        return out

    # ---- general relativistic; hole = (M, a, bx, by, bz) ----
    def gr_sweep(self, W, order, axis, h, dt, hole, bias=0.0, limiter=1, gamma=1.4):
        a = np.ascontiguousarray(W, dtype=np.float64).copy()
        dim, nx, ny = self._dims(a)
        self._f("gr_sweep")(order, axis, dim, _dp(a), nx, ny, h, dt, bias, limiter, gamma, *hole)
        return a

    def gr_forcing(self, W, h, dt, hole, bias=0.0, limiter=1, gamma=1.4):
        a = np.ascontiguousarray(W, dtype=np.float64).copy()
        dim, nx, ny = self._dims(a)
        self._f("gr_forcing")(dim, _dp(a), nx, ny, h, dt, bias, limiter, gamma, *hole)
        return a

    def gr_max_wavespeed(self, W, hole, gamma=1.4):
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        return self._f("gr_max_wavespeed")(dim, _dp(a), nx, ny, gamma, *hole)

    def gr_run(self, W, order, h, cfl, hole, tf=math.inf, max_steps=2 ** 31 - 1, bias=0.0, limiter=1,
               subcycles=1, gamma=1.4):
        a = np.ascontiguousarray(W, dtype=np.float64).copy()
        dim, nx, ny = self._dims(a)
        t = _D()
        n = self._f("gr_run")(order, dim, _dp(a), nx, ny, h, cfl, tf, max_steps, bias, limiter, subcycles, gamma,
                              *hole, ctypes.byref(t))
        return a, t.value, n

    def gr_run_dts(self, W, order, h, dts, hole, bias=0.0, limiter=1, subcycles=1, gamma=1.4):
        """The black-hole solve loop's steps with the time steps dts held fixed."""
        a = np.ascontiguousarray(W, dtype=np.float64).copy()
        dim, nx, ny = self._dims(a)
        d = np.ascontiguousarray(dts, dtype=np.float64)
        self.lib.orc_gr_run_dts(order, dim, _dp(a), nx, ny, h, _dp(d), len(d), bias, limiter, subcycles, gamma, *hole)
        return a

    def traced(self, fn, capacity=1 << 26):
        """(fn(), trace bytes): the slope limiter's branch trace of fn, run on one thread."""
        buf = np.zeros(capacity, dtype=np.uint8)
        self.set_threads(1)
        self.lib.orc_trace_begin(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), capacity)
        try:
            res = fn()
        finally:
            n = self.lib.orc_trace_end()
            self.set_threads(0)
        if n > capacity:
            raise ValueError(f"trace of {n} bytes exceeds the {capacity}-byte buffer")
        return res, buf[:n]

    def on_pieces(self, trace, fn):
        """fn() run on one thread with every slope component held to the
        clamps and limiter piece `trace` (from traced()) recorded for it."""
        buf = np.ascontiguousarray(trace, dtype=np.uint8)
        self.set_threads(1)
        self.lib.orc_trace_replay(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), len(buf))
        try:
            res = fn()
        finally:
            n = self.lib.orc_trace_end()
            self.set_threads(0)
        if n != len(buf):
            raise ValueError(f"replayed run evaluated {n} slope components, the trace holds {len(buf)}")
        return res

    def gr_point(self, what, v, x, y, z, h, hole, gamma=1.4):
        out = np.full(44, np.nan)
        a = np.ascontiguousarray(v if v is not None else np.zeros(5), dtype=np.float64)
        self._f("gr_point")(what, _dp(a), x, y, z, h, gamma, *hole, _dp(out))
        n = {GR_SOUND: 1, GR_METRIC: 44}.get(what, 5)
        return out[:n]

    # ---- two-level AMR ----
    def amr_level1(self, W, physics, order, dx, cfl, tf, tol, gamma=1.4, max_steps=-1):
        """(coarse, fine, coarse_flags, fine_flags[, t, steps]) of solve{,2D}Level1AMR."""
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        fshape = (2 * nx, 5) if dim == 1 else (2 * ny, 2 * nx, 5)
        coarse = np.empty_like(a)
        fine = np.empty(fshape)
        cf = np.empty(a.shape[:-1], dtype=np.uint8)
        ff = np.empty(fshape[:-1], dtype=np.uint8)
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        args = [physics, dim, order, _dp(a), nx, ny, dx, cfl, tf, tol, gamma]
        if self.prefix == "ref_":
            self.lib.ref_amr_level1(*args, _dp(coarse), _dp(fine), cf.ctypes.data_as(u8), ff.ctypes.data_as(u8))
            return coarse, fine, cf, ff
        t = _D()
        n = self.lib.orc_amr_level1(*args, max_steps, _dp(coarse), _dp(fine), cf.ctypes.data_as(u8),
                                    ff.ctypes.data_as(u8), ctypes.byref(t))
        return coarse, fine, cf, ff, t.value, n

    def amr_level2(self, W, physics, order, dx, cfl, tf, tol, gamma=1.4, max_steps=-1):
        """(coarse, inter, fine, coarse_flags, inter_flags, fine_flags[, t, steps]) of solve{,2D}Level2AMR."""
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        shape = (lambda m: (m * nx, 5)) if dim == 1 else (lambda m: (m * ny, m * nx, 5))
        coarse, inter, fine = np.empty_like(a), np.empty(shape(2)), np.empty(shape(4))
        cf, jf, ff = (np.empty(shape(m)[:-1], dtype=np.uint8) for m in (1, 2, 4))
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        args = [physics, dim, order, _dp(a), nx, ny, dx, cfl, tf, tol, gamma]
        outs = [_dp(coarse), _dp(inter), _dp(fine), cf.ctypes.data_as(u8), jf.ctypes.data_as(u8),
                ff.ctypes.data_as(u8)]
        if self.prefix == "ref_":
            self.lib.ref_amr_level2(*args, *outs)
            return coarse, inter, fine, cf, jf, ff
        t = _D()
        k = self.lib.orc_amr_level2(*args, max_steps, *outs, ctypes.byref(t))
        return coarse, inter, fine, cf, jf, ff, t.value, k

    def gr_amr_level1(self, W, order, dx, cfl, tf, tol, hole, gamma=4.0 / 3.0, max_steps=-1):
        """(coarse, fine, coarse_flags, fine_flags[, t, steps]) of BlackHoleEulerAMR::solve{,2D}Level1AMR."""
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        fshape = (2 * nx, 5) if dim == 1 else (2 * ny, 2 * nx, 5)
        coarse, fine = np.empty_like(a), np.empty(fshape)
        cf = np.empty(a.shape[:-1], dtype=np.uint8)
        ff = np.empty(fshape[:-1], dtype=np.uint8)
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        if self.prefix == "ref_":
            self.lib.ref_gr_amr_level1(order, dim, _dp(a), nx, ny, dx, cfl, tf, tol, gamma, *hole, _dp(coarse),
                                       _dp(fine), cf.ctypes.data_as(u8), ff.ctypes.data_as(u8))
            return coarse, fine, cf, ff
        t = _D()
        n = self.lib.orc_gr_amr_level1(order, dim, _dp(a), nx, ny, dx, cfl, tf, tol, max_steps, gamma, *hole,
                                       _dp(coarse), _dp(fine), cf.ctypes.data_as(u8), ff.ctypes.data_as(u8),
                                       ctypes.byref(t))
        return coarse, fine, cf, ff, t.value, n

    def gr_amr_level2(self, W, order, dx, cfl, tf, tol, hole, gamma=4.0 / 3.0, max_steps=-1):
        """(coarse, inter, fine, flags x3[, t, steps]) of BlackHoleEulerAMR::solve{,2D}Level2AMR
        (the reference build: 1D only, its 2D driver being truncated)."""
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        shape = (lambda m: (m * nx, 5)) if dim == 1 else (lambda m: (m * ny, m * nx, 5))
        coarse, inter, fine = np.empty_like(a), np.empty(shape(2)), np.empty(shape(4))
        cf, jf, ff = (np.empty(shape(m)[:-1], dtype=np.uint8) for m in (1, 2, 4))
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        outs = [_dp(coarse), _dp(inter), _dp(fine), cf.ctypes.data_as(u8), jf.ctypes.data_as(u8),
                ff.ctypes.data_as(u8)]
        if self.prefix == "ref_":
            if dim != 1:
                raise ValueError("the reference's 2D black-hole level-2 driver does not compile")
            self.lib.ref_gr_amr_level2(order, _dp(a), nx, dx, cfl, tf, tol, gamma, *hole, *outs)
            return coarse, inter, fine, cf, jf, ff
        t = _D()
        k = self.lib.orc_gr_amr_level2(order, dim, _dp(a), nx, ny, dx, cfl, tf, tol, max_steps, gamma, *hole, *outs,
                                       ctypes.byref(t))
        return coarse, inter, fine, cf, jf, ff, t.value, k

    def amr_refine_flags(self, W, physics, dx, tol, cflags, fflags, gamma=1.4):
        """One sticky flag pass; returns updated copies of the flag arrays."""
        a = np.ascontiguousarray(W, dtype=np.float64)
        dim, nx, ny = self._dims(a)
        cf = np.array(cflags, dtype=np.uint8, copy=True)
        ff = np.array(fflags, dtype=np.uint8, copy=True)
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        self.lib.orc_amr_refine_flags(physics, dim, _dp(a), nx, ny, dx, tol, gamma, cf.ctypes.data_as(u8),
                                      ff.ctypes.data_as(u8))
        return cf, ff

    # oracle-only controls
    def set_con2prim(self, tol, max_iterations):
        self.lib.orc_set_c2p(tol, max_iterations)

    def gr_set_con2prim(self, tol, max_iterations=1000):
        """Black-hole con2prim: tol > 0 replaces every per-call tolerance; 0 restores the reference's."""
        self.lib.orc_gr_set_c2p(tol, max_iterations)

    def events_reset(self):
        self.lib.orc_events_reset()

    def events(self):
        """Special-relativistic con2prim events since the last reset:
        (cap hits, q floors, density floors, pressure floors)."""
        out = (ctypes.c_longlong * 4)()
        self.lib.orc_events_read(out)
        return tuple(int(v) for v in out)

    def gr_run_nan_counts(self, W, order, h, cfl, hole, steps, subcycles=1, gamma=1.4):
        """(state, per-step counts of cells holding a NaN) of a black-hole solve of `steps` steps."""
        buf = np.full(steps, -1, dtype=np.intc)
        fn = self._f("set_step_nan_counts")
        fn.argtypes = [ctypes.POINTER(ctypes.c_int), _I]
        fn(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), steps)
        try:
            out = self.gr_run(W, order, h, cfl, hole, max_steps=steps, subcycles=subcycles, gamma=gamma)
        finally:
            fn(None, 0)
        return out[0], buf

    def set_threads(self, n):
        """Threads of the OpenMP loops; returns the count they will use."""
        if self.prefix == "ref_":
            return self.lib.ref_set_threads(n)
        self.lib.orc_set_threads(n)
        return n

    def gr_set_excision(self, radius):
        """Excision radius override (<= 0 restores the reference's horizon radius)."""
        self.lib.orc_gr_set_excision(radius)


def load_oracle() -> CpuPath:
    if not os.path.exists(ORACLE_SO):
        raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
    return CpuPath(ORACLE_SO, "orc_")


def load_ref() -> CpuPath | None:
    """The reference build, or None where it has not been built."""
    if not os.path.exists(REF_SO):
        return None
    return CpuPath(REF_SO, "ref_")


def load_ref_omp() -> CpuPath | None:
    """The reference's sources built with their OpenMP loops enabled, for
    timing only (those loops race on shared maxima and flag vectors, so its
    results are never compared), or None where it has not been built."""
    if not os.path.exists(REF_OMP_SO):
        return None
    return CpuPath(REF_OMP_SO, "ref_")
