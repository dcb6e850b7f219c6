"""Python handle over one device grid (a thin wrapper of include/cosmos_fv.h).

Method names follow the reference's static solver methods they stand for:
    sweep             compute{X,Y}SLICTimeStep2D / computeSLICTimeStep and the FORCE ones
    max_wave_speed    computeMaximumWaveSpeed{,2D}
    stable_time_step  computeStableTimeStep{,2D}
    solve             solve / solve2D (to a final time, or a fixed step count)
    forcing           BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep{,2D}
    solve_forced      BlackHoleEuler{First,Second}OrderSolver::solve{,2D}
States are numpy float64 arrays of primitives, shape (ny, nx, 5) in 2D or
(nx, 5) in 1D, k in (rho, u, v, w, p): vector<vector<StateVector>> flattened.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import native

NEWTONIAN = 0
RELATIVISTIC = 1
BLACKHOLE = 2
SUPERBEE, VANLEER, MINBEE = 0, 1, 2
EXACT, FAST = 0, 1
TAPE_NONE, TAPE_RERUN, TAPE_SEGMENT = 0, 1, 2


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _HostBlock:
    """Page-locked host memory from cosmos_host_alloc, freed with the last array over it."""

    def __init__(self, nbytes: int):
        self._lib = native.load()
        p = ctypes.c_void_p()
        if self._lib.cosmos_host_alloc(nbytes, ctypes.byref(p)) != 0 or not p.value:
            raise native.CosmosError(f"cosmos_host_alloc({nbytes}) failed")
        self.ptr = p.value
        self.nbytes = nbytes

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.cosmos_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = None


class HostArray(np.ndarray):
    """A numpy array in page-locked host memory: Solver.upload / download copy it
    by DMA directly instead of staging it."""

    def __array_finalize__(self, obj):
        self._block = getattr(obj, "_block", None)


def host_empty(shape, dtype=np.float64) -> HostArray:
    """An uninitialised HostArray of the given shape."""
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    block = _HostBlock(max(nbytes, 1))
    raw = (ctypes.c_char * block.nbytes).from_address(block.ptr)
    a = np.frombuffer(raw, dtype=np.uint8)[:nbytes].view(dtype).reshape(shape).view(HostArray)
    a._block = block
    return a


class Solver:
    def __init__(self, physics: int, dim: int, nx: int, ny: int = 1, adiabatic_index: float = 1.4):
        self._lib = native.load()
        self.physics, self.dim = physics, dim
        self.nx, self.ny = nx, (ny if dim == 2 else 1)
        h = ctypes.c_void_p()
        rc = self._lib.cosmos_create(ctypes.byref(h), physics, dim, nx, self.ny, adiabatic_index)
        native.check(rc, None)
        self._h = h

    # -- lifetime -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.cosmos_destroy(self._h)
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

    def _check(self, rc: int) -> None:
        native.check(rc, self._h)

    @property
    def shape(self):
        return (self.ny, self.nx, 5) if self.dim == 2 else (self.nx, 5)

    # -- state ---------------------------------------------------------------
    def upload(self, prim: np.ndarray) -> None:
        a = np.ascontiguousarray(prim, dtype=np.float64)
        if a.size != self.nx * self.ny * 5:
            raise ValueError(f"state has {a.size} values, grid needs {self.nx * self.ny * 5}")
        self._check(self._lib.cosmos_upload(self._h, _dp(a)))

    def download(self, out: np.ndarray | None = None) -> np.ndarray:
        """The state (into ``out`` when given: a C-contiguous float64 array of the
        grid's size, e.g. a HostArray, which is copied by DMA directly)."""
        if out is None:
            out = np.empty(self.shape, dtype=np.float64)
        elif out.dtype != np.float64 or not out.flags.c_contiguous or out.size != self.nx * self.ny * 5:
            raise ValueError("out must be a C-contiguous float64 array of the grid's size")
        self._check(self._lib.cosmos_download(self._h, _dp(out)))
        return out

    def upload_device(self, ptr: int) -> None:
        self._check(self._lib.cosmos_upload_device(self._h, ctypes.c_void_p(ptr)))

    def download_device(self, ptr: int) -> None:
        self._check(self._lib.cosmos_download_device(self._h, ctypes.c_void_p(ptr)))

    def set_mode(self, mode: int) -> None:
        """EXACT (bitwise reference arithmetic, default) or FAST (Newtonian SLIC/van Leer)."""
        self._check(self._lib.cosmos_set_mode(self._h, mode))

    def set_con2prim(self, tolerance: float, max_iterations: int) -> None:
        self._check(self._lib.cosmos_set_con2prim(self._h, tolerance, max_iterations))

    def set_launch(self, y_rows: int = 0, ws_mode: int = 0) -> None:
        """Launch shape: rows per y-sweep thread (16, 32, 64; 0 = default for the
        physics) and the wave-speed hand-off (1 per-block partials, 2 per-wave
        slot atomics, 0 = default)."""
        self._check(self._lib.cosmos_set_launch(self._h, y_rows, ws_mode))

    def set_y_tile(self, columns: int = -1) -> None:
        """y-sweep layout: 0 column march, 8 or 16 LDS tile columns, -1 default by physics."""
        self._check(self._lib.cosmos_set_y_tile(self._h, columns))

    def set_step_fusion(self, enable: bool = True, rows_per_block: int = 0) -> None:
        """Fast mode, 2D Newtonian SLIC / van Leer / bias 0 on >= 8 x 8 cells: run
        each queued step as one fused X Y X kernel (default) or as three sweeps;
        rows_per_block is the fused kernel's launch shape (0: automatic)."""
        self._check(self._lib.cosmos_set_step_fusion(self._h, 1 if enable else 0, rows_per_block))

    def set_spacetime(self, mass: float = 1.0, spin: float = 0.0, x: float = 0.5, y: float = 0.5,
                      z: float = 0.0) -> None:
        """BlackHoleSpacetime(mass, spin, x, y, z) of a BLACKHOLE solver."""
        self._check(self._lib.cosmos_set_spacetime(self._h, mass, spin, x, y, z))

    # -- numerics ---------------------------------------------------------------
    def sweep(self, order: int, axis: int, cell_spacing: float, time_step: float, bias: float = 0.0,
              slope_limiter: int = VANLEER) -> None:
        self._check(self._lib.cosmos_sweep(self._h, order, axis, cell_spacing, time_step, bias, slope_limiter))

    def max_wave_speed(self) -> float:
        v = ctypes.c_double()
        self._check(self._lib.cosmos_max_wave_speed(self._h, ctypes.byref(v)))
        return v.value

    def stable_time_step(self, cell_spacing: float, cfl: float, current_time: float, final_time: float,
                         current_iteration: int) -> float:
        v = ctypes.c_double()
        self._check(self._lib.cosmos_stable_time_step(self._h, cell_spacing, cfl, current_time, final_time,
                                                      current_iteration, ctypes.byref(v)))
        return v.value

    def begin(self, order: int, cell_spacing: float, cfl: float, final_time: float = math.inf,
              max_steps: int = -1, bias: float = 0.0, slope_limiter: int = VANLEER, start_time: float = 0.0,
              start_iteration: int = 0) -> None:
        """cosmos_begin, or cosmos_begin_at for a run resumed at (start_time, start_iteration)."""
        if start_time == 0.0 and start_iteration == 0:
            self._check(self._lib.cosmos_begin(self._h, order, cell_spacing, cfl, final_time, max_steps, bias,
                                               slope_limiter))
        else:
            self._check(self._lib.cosmos_begin_at(self._h, order, cell_spacing, cfl, final_time, max_steps, bias,
                                                  slope_limiter, start_time, start_iteration))

    def advance(self, n_steps: int) -> None:
        self._check(self._lib.cosmos_advance(self._h, n_steps))

    def progress(self):
        t, n, stopped = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.cosmos_progress(self._h, ctypes.byref(t), ctypes.byref(n), ctypes.byref(stopped)))
        return t.value, n.value, bool(stopped.value)

    def solve(self, order: int, cell_spacing: float, cfl: float, final_time: float = math.inf,
              max_steps: int = -1, bias: float = 0.0, slope_limiter: int = VANLEER):
        """Returns (time reached, steps taken)."""
        t, n = ctypes.c_double(), ctypes.c_int()
        self._check(self._lib.cosmos_solve(self._h, order, cell_spacing, cfl, final_time, max_steps, bias,
                                           slope_limiter, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def status(self) -> dict:
        """cosmos_status: the numerical-fault counters and flags."""
        st = native.Status()
        self._check(self._lib.cosmos_status(self._h, ctypes.byref(st)))
        return {name: int(getattr(st, name)) for name, _ in native.Status._fields_ if name != "reserved"}

    def status_reset(self) -> None:
        self._check(self._lib.cosmos_status_reset(self._h))

    def set_geometry_cache(self, enable: bool) -> None:
        """Black hole: tabulate the position-only metric values (default on)."""
        self._check(self._lib.cosmos_set_geometry_cache(self._h, 1 if enable else 0))

    def set_excision(self, radius: float) -> None:
        """Excision radius override of a BLACKHOLE solver (<= 0: the reference's horizon)."""
        self._check(self._lib.cosmos_set_excision(self._h, radius))

    def forcing(self, cell_spacing: float, time_step: float, bias: float = 0.0,
                slope_limiter: int = VANLEER) -> None:
        self._check(self._lib.cosmos_forcing(self._h, cell_spacing, time_step, bias, slope_limiter))

    def solve_forced(self, order: int, cell_spacing: float, cfl: float, final_time: float = math.inf,
                     max_steps: int = -1, bias: float = 0.0, slope_limiter: int = VANLEER, subcycles: int = 1,
                     start_time: float = 0.0, start_iteration: int = 0):
        """Returns (time reached, steps taken); resumes at (start_time, start_iteration) when given."""
        t, n = ctypes.c_double(), ctypes.c_int()
        self._check(self._lib.cosmos_solve_forced_at(self._h, order, cell_spacing, cfl, final_time, max_steps, bias,
                                                     slope_limiter, subcycles, start_time, start_iteration,
                                                     ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    # -- checkpoint / restart and text output --------------------------------
    def checkpoint_save(self, path: str, time: float, iteration: int) -> None:
        self._check(self._lib.cosmos_checkpoint_save(self._h, str(path).encode(), time, iteration))

    def checkpoint_load(self, path: str):
        """Restores the state (and a black hole's spacetime); returns (time, iteration)."""
        t, it = ctypes.c_double(), ctypes.c_int()
        self._check(self._lib.cosmos_checkpoint_load(self._h, str(path).encode(), ctypes.byref(t), ctypes.byref(it)))
        return t.value, it.value

    def write_dat(self, directory: str) -> None:
        """density.dat ... pressure.dat in the reference's text layout."""
        self._check(self._lib.cosmos_write_dat(self._h, str(directory).encode()))

    def synchronize(self) -> None:
        self._check(self._lib.cosmos_synchronize(self._h))

    @property
    def stream(self) -> int:
        return self._lib.cosmos_stream(self._h) or 0

    def state_planes(self):
        arr = (ctypes.c_void_p * 5)()
        pitch = ctypes.c_long()
        self._check(self._lib.cosmos_state_planes(self._h, arr, ctypes.byref(pitch)))
        return [arr[k] for k in range(5)], pitch.value


class Amr:
    """AMR context (cosmos_amr_*): a coarse grid, a fine grid of twice the
    resolution (and, after solve_level2, one of four times) and one refinement
    flag per cell of each; Newtonian, special-relativistic or black-hole."""

    def __init__(self, physics: int, dim: int, nx: int, ny: int = 1, adiabatic_index: float = 1.4):
        self._lib = native.load()
        self.physics, self.dim = physics, dim
        self.nx, self.ny = nx, (ny if dim == 2 else 1)
        h = ctypes.c_void_p()
        native.check(self._lib.cosmos_amr_create(ctypes.byref(h), physics, dim, nx, self.ny, adiabatic_index), None,
                     amr=True)
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.cosmos_amr_destroy(self._h)
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

    def _check(self, rc: int) -> None:
        native.check(rc, self._h, amr=True)

    def set_con2prim(self, tolerance: float, max_iterations: int) -> None:
        self._check(self._lib.cosmos_amr_set_con2prim(self._h, tolerance, max_iterations))

    def set_spacetime(self, mass: float, spin: float, x: float, y: float, z: float = 0.0) -> None:
        """Black-hole contexts: the Kerr-Schild hole of every level."""
        self._check(self._lib.cosmos_amr_set_spacetime(self._h, mass, spin, x, y, z))

    def set_excision(self, radius: float) -> None:
        self._check(self._lib.cosmos_amr_set_excision(self._h, radius))

    AMR_KERNEL_KINDS = ("refine", "exchange", "derefine", "merge")

    def profile(self, on: bool = True) -> None:
        """Bracket every AMR pass kernel with HIP events while on."""
        self._check(self._lib.cosmos_amr_profile_enable(self._h, 1 if on else 0))

    def profile_read(self) -> dict:
        """{kind: (launches, total_ms)} of the timed AMR kernels since the last read."""
        tot = (ctypes.c_double * 4)()
        cnt = (ctypes.c_int * 4)()
        self._check(self._lib.cosmos_amr_profile_read(self._h, tot, cnt))
        return {name: (cnt[k], tot[k]) for k, name in enumerate(self.AMR_KERNEL_KINDS)}

    def level_stream(self, level: int = 0) -> int:
        """The HIP stream of a level's solver handle (0 coarse, 1 fine, 2 finest of a three-level run)."""
        s = self._lib.cosmos_amr_level_solver(self._h, level)
        if not s:
            raise native.CosmosError(f"AMR level {level} has no solver handle")
        return self._lib.cosmos_stream(s)

    def upload(self, prim: np.ndarray) -> None:
        a = np.ascontiguousarray(prim, dtype=np.float64)
        if a.size != self.nx * self.ny * 5:
            raise ValueError("state does not match the coarse grid")
        self._check(self._lib.cosmos_amr_upload(self._h, _dp(a)))

    def download(self):
        """(coarse, fine, coarse_flags, fine_flags)"""
        if self.dim == 1:
            cs, fs = (self.nx, 5), (2 * self.nx, 5)
        else:
            cs, fs = (self.ny, self.nx, 5), (2 * self.ny, 2 * self.nx, 5)
        c, f = np.empty(cs), np.empty(fs)
        cf, ff = np.empty(cs[:-1], dtype=np.uint8), np.empty(fs[:-1], dtype=np.uint8)
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        self._check(self._lib.cosmos_amr_download(self._h, _dp(c), _dp(f), cf.ctypes.data_as(u8),
                                                  ff.ctypes.data_as(u8)))
        return c, f, cf, ff

    def refine_flags(self, cell_spacing: float, tolerance: float) -> None:
        self._check(self._lib.cosmos_amr_refine_flags(self._h, cell_spacing, tolerance))

    def sweep(self, level: int, order: int, axis: int, cell_spacing: float, time_step: float, bias: float = 0.0,
              slope_limiter: int = VANLEER) -> None:
        self._check(self._lib.cosmos_amr_sweep(self._h, level, order, axis, cell_spacing, time_step, bias,
                                               slope_limiter))

    def restrict(self) -> None:
        self._check(self._lib.cosmos_amr_restrict(self._h))

    def derefine(self, cell_spacing: float, tolerance: float) -> None:
        self._check(self._lib.cosmos_amr_derefine(self._h, cell_spacing, tolerance))

    def solve_level1(self, order: int, cell_spacing: float, cfl: float, final_time: float, tolerance: float,
                     max_coarse_steps: int = -1):
        """Returns (time reached, coarse steps taken)."""
        t, n = ctypes.c_double(), ctypes.c_int()
        self._check(self._lib.cosmos_amr_solve_level1(self._h, order, cell_spacing, cfl, final_time, tolerance,
                                                      max_coarse_steps, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def solve_level2(self, order: int, cell_spacing: float, cfl: float, final_time: float, tolerance: float,
                     max_coarse_steps: int = -1):
        """Three-level run from the uploaded coarse state; returns (time reached, coarse steps taken)."""
        t, n = ctypes.c_double(), ctypes.c_int()
        self._check(self._lib.cosmos_amr_solve_level2(self._h, order, cell_spacing, cfl, final_time, tolerance,
                                                      max_coarse_steps, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def download_level2(self):
        """(coarse, intermediate, fine, coarse_flags, intermediate_flags, fine_flags) of a three-level run."""
        shape = (lambda k: (k * self.nx, 5)) if self.dim == 1 else (lambda k: (k * self.ny, k * self.nx, 5))
        c, m, f = np.empty(shape(1)), np.empty(shape(2)), np.empty(shape(4))
        cf, mf, ff = (np.empty(shape(k)[:-1], dtype=np.uint8) for k in (1, 2, 4))
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        self._check(self._lib.cosmos_amr_download_level2(self._h, _dp(c), _dp(m), _dp(f), cf.ctypes.data_as(u8),
                                                         mf.ctypes.data_as(u8), ff.ctypes.data_as(u8)))
        return c, m, f, cf, mf, ff


class Adjoint:
    """Reverse mode of the Newtonian or black-hole update (cosmos_adjoint_*):
    a solver plus the recorded trajectory (dt schedule and checkpoints) of a
    forward run."""

    def __init__(self, dim: int, nx: int, ny: int = 1, adiabatic_index: float = 1.4, physics: int = NEWTONIAN):
        self._lib = native.load()
        self.dim = dim
        self.physics = physics
        self.nx, self.ny = nx, (ny if dim == 2 else 1)
        h = ctypes.c_void_p()
        native.check(self._lib.cosmos_adjoint_create(ctypes.byref(h), physics, dim, nx, self.ny, adiabatic_index),
                     None, adjoint=True)
        self._h = h

    def _solver(self):
        return self._lib.cosmos_adjoint_solver(self._h)

    def set_spacetime(self, mass: float = 1.0, spin: float = 0.0, x: float = 0.5, y: float = 0.5,
                      z: float = 0.0) -> None:
        s = self._solver()
        native.check(self._lib.cosmos_set_spacetime(s, mass, spin, x, y, z), s)

    def set_excision(self, radius: float) -> None:
        s = self._solver()
        native.check(self._lib.cosmos_set_excision(s, radius), s)

    def set_con2prim(self, tolerance: float, max_iterations: int) -> None:
        s = self._solver()
        native.check(self._lib.cosmos_set_con2prim(s, tolerance, max_iterations), s)

    def set_subcycles(self, subcycles: int) -> None:
        self._check(self._lib.cosmos_adjoint_set_subcycles(self._h, subcycles))

    def set_tape(self, mode: int, reserve_bytes: int = 8 << 30) -> None:
        """Black hole: TAPE_NONE, TAPE_RERUN or TAPE_SEGMENT (cosmos_adjoint_set_tape)."""
        self._check(self._lib.cosmos_adjoint_set_tape(self._h, mode, reserve_bytes))

    def release(self) -> None:
        """Free the segment buffers and tapes (cosmos_adjoint_release)."""
        self._check(self._lib.cosmos_adjoint_release(self._h))

    def forcing(self, cell_spacing: float, time_step: float, lam_out: np.ndarray) -> np.ndarray:
        """Advance the state by one RK4 forcing step; return (d out / d in)^T lam_out."""
        lo = self._arr(lam_out)
        li = np.empty(self.shape)
        self._check(self._lib.cosmos_adjoint_forcing(self._h, cell_spacing, time_step, _dp(lo), _dp(li)))
        return li

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.cosmos_adjoint_destroy(self._h)
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

    @property
    def shape(self):
        return (self.ny, self.nx, 5) if self.dim == 2 else (self.nx, 5)

    def _check(self, rc: int) -> None:
        native.check(rc, self._h, adjoint=True)

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.nx * self.ny * 5:
            raise ValueError("array does not match the grid")
        return a

    def upload(self, prim: np.ndarray) -> None:
        s = self._lib.cosmos_adjoint_solver(self._h)
        a = self._arr(prim)
        native.check(self._lib.cosmos_upload(s, _dp(a)), s)

    def download(self) -> np.ndarray:
        s = self._lib.cosmos_adjoint_solver(self._h)
        out = np.empty(self.shape)
        native.check(self._lib.cosmos_download(s, _dp(out)), s)
        return out

    def sweep(self, order: int, axis: int, cell_spacing: float, time_step: float, lam_out: np.ndarray,
              bias: float = 0.0, slope_limiter: int = VANLEER) -> np.ndarray:
        """Advance the state by one sweep; return (d out / d in)^T lam_out."""
        lo = self._arr(lam_out)
        li = np.empty(self.shape)
        self._check(self._lib.cosmos_adjoint_sweep(self._h, order, axis, cell_spacing, time_step, bias,
                                                   slope_limiter, _dp(lo), _dp(li)))
        return li

    def forward(self, order: int, cell_spacing: float, cfl: float, n_steps: int, bias: float = 0.0,
                slope_limiter: int = VANLEER, checkpoint_every: int = 20) -> None:
        self._check(self._lib.cosmos_adjoint_forward(self._h, order, cell_spacing, cfl, n_steps, bias,
                                                     slope_limiter, checkpoint_every))

    def backward(self, grad_out: np.ndarray) -> np.ndarray:
        go = self._arr(grad_out)
        gi = np.empty(self.shape)
        self._check(self._lib.cosmos_adjoint_backward(self._h, _dp(go), _dp(gi)))
        return gi

    def time_steps(self) -> np.ndarray:
        """The dt schedule the last forward pass recorded."""
        n = ctypes.c_int()
        self._check(self._lib.cosmos_adjoint_time_steps(self._h, None, 0, ctypes.byref(n)))
        out = np.empty(n.value)
        self._check(self._lib.cosmos_adjoint_time_steps(self._h, _dp(out), n.value, ctypes.byref(n)))
        return out

    def replay(self, prim_in: np.ndarray) -> np.ndarray:
        a = self._arr(prim_in)
        out = np.empty(self.shape)
        self._check(self._lib.cosmos_adjoint_replay(self._h, _dp(a), _dp(out)))
        return out
