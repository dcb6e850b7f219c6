"""The numerical-fault status word (cosmos_status): special-relativistic
con2prim events of the sweeps counted exactly as the oracle counts them (every
distinct solve of a line once), NaN cells of the state, a NaN time step that
ends a solve loop, and a black-hole grid excised from the wave speed
everywhere.  The reference reports none of these
(RelativisticEuler/RelativisticEulerStateVector.cpp:45-50, 59-80, 93-100;
Solvers/EulerSecondOrderSolver.cpp:109; Solvers/BlackHoleEulerSolvers.cpp:88-140)."""
import ctypes
import math

import numpy as np
import pytest

from conftest import same, sr_extreme_2d
from cosmos_amd import BLACKHOLE, NEWTONIAN, RELATIVISTIC, Solver, native, workloads

pytestmark = pytest.mark.gpu

G_SR = 5.0 / 3.0
KEYS = ("c2p_cap", "c2p_q_floor", "c2p_density_floor", "c2p_pressure_floor")


def events(st):
    return tuple(st[k] for k in KEYS)


@pytest.mark.parametrize("nx,ny", [(17, 9), (40, 33), (5, 4), (300, 70)])
def test_sr_con2prim_events_match_oracle(orc, nx, ny):
    """Every sweep of the extreme-state grids, each y layout and both orders:
    the device counters equal the oracle's counts of the same sweep, and the
    regime is exercised (cap hits and every floor occur)."""
    W = sr_extreme_2d(nx, ny, seed=3)
    dx, dt = 1.0 / 40, 0.1 / 40
    total = np.zeros(4, dtype=np.int64)
    for order in (1, 2):
        for axis in (0, 1):
            orc.events_reset()
            orc.sweep(W, 1, order, axis, dx, dt, 0.0, 1, G_SR)
            exp = orc.events()
            total += np.array(exp)
            layouts = [(0, 16), (0, 64), (8, 0), (16, 0)] if axis == 1 else [(-1, 0)]
            for columns, rows in layouts:
                with Solver(RELATIVISTIC, 2, nx, ny, G_SR) as s:
                    if columns >= 0:
                        s.set_y_tile(columns)
                    if rows:
                        s.set_launch(rows)
                    s.upload(W)
                    s.sweep(order, axis, dx, dt, 0.0, 1)
                    got = events(s.status())
                assert got == exp, (order, axis, columns, rows, got, exp)
    row = W[ny // 2].copy()
    for order in (1, 2):
        orc.events_reset()
        orc.sweep(row, 1, order, 0, dx, dt, 0.0, 1, G_SR)
        with Solver(RELATIVISTIC, 1, nx, 1, G_SR) as s:
            s.upload(row)
            s.sweep(order, 0, dx, dt, 0.0, 1)
            assert events(s.status()) == orc.events(), order
    if nx * ny > 100:
        assert (total > 0).all(), total


def test_sr_events_accumulate_over_a_solve_and_reset(orc):
    """The counters accumulate over every sweep of a solve loop (equal to the
    oracle's run) and cosmos_status_reset clears them."""
    W = sr_extreme_2d(40, 33, seed=5)
    orc.events_reset()
    exp, te, ke = orc.run(W, 1, 2, 1.0 / 40, 0.5, max_steps=4, gamma=G_SR)
    with Solver(RELATIVISTIC, 2, 40, 33, G_SR) as s:
        s.upload(W)
        t, k = s.solve(2, 1.0 / 40, 0.5, math.inf, 4)
        st = s.status()
        assert same(s.download(), exp) and (t, k) == (te, ke)
        assert events(st) == orc.events()
        assert st["nan_cells"] == int(np.isnan(exp).any(axis=-1).sum())
        s.status_reset()
        assert all(v == 0 for v in events(s.status()))


def test_nan_cells_counted(orc):
    W = workloads.quadrants_2d(50, seed=4, ny=30)
    W[3, 7, 0] = np.nan
    W[20, 40, 4] = np.nan
    W[20, 41, 1:3] = np.nan
    with Solver(NEWTONIAN, 2, 50, 30, 1.4) as s:
        s.upload(W)
        assert s.status()["nan_cells"] == 3
        s.sweep(2, 0, 1 / 50, 0.002)
        out = s.download()
        assert s.status()["nan_cells"] == int(np.isnan(out).any(axis=-1).sum()) > 3


def test_nan_time_step_flag(orc):
    """A medium at rest with zero pressure has wave speed 0; with CFL 0 the time
    step 0 * (dx / 0) is NaN: the loop takes that step and stops, as the
    reference's does, and the flag is set."""
    W = np.zeros((12, 10, 5))
    W[..., 0] = 1.0
    exp, te, ke = orc.run(W, 0, 2, 0.1, 0.0, max_steps=5)
    with Solver(NEWTONIAN, 2, 10, 12, 1.4) as s:
        s.upload(W)
        t, k = s.solve(2, 0.1, 0.0, math.inf, 5)
        st = s.status()
        assert k == ke == 1 and math.isnan(t) and math.isnan(te)
        assert same(s.download(), exp)
        assert st["nan_time_step"] == 1 and st["nan_cells"] == 120
    with Solver(NEWTONIAN, 2, 10, 12, 1.4) as s:
        s.upload(workloads.quadrants_2d(10, seed=1, ny=12))
        s.solve(2, 0.1, 0.9, math.inf, 3)
        assert s.status()["nan_time_step"] == 0


def test_excised_grid_flag():
    """Every cell excised at the wave speed's unit-square positions: maximum 0,
    flag set; a grid with cells outside the hole leaves it clear."""
    W = np.zeros((8, 8, 5))
    W[..., 0] = 1.0
    W[..., 4] = 0.1
    with Solver(BLACKHOLE, 2, 8, 8, G_SR) as s:
        s.set_spacetime(2.0, 0.0, 0.5, 0.5, 0.0)
        s.upload(W)
        assert s.max_wave_speed() == 0.0
        assert s.status()["excised_grid"] == 1
    with Solver(BLACKHOLE, 2, 8, 8, G_SR) as s:
        s.set_spacetime(0.05, 0.0, 0.5, 0.5, 0.0)
        s.upload(W)
        assert s.max_wave_speed() > 0.0
        assert s.status()["excised_grid"] == 0


def test_status_async_copy():
    """cosmos_status_async queues the copy; after a synchronize the host struct
    holds the same values as cosmos_status."""
    lib = native.load()
    W = sr_extreme_2d(40, 33, seed=3)
    with Solver(RELATIVISTIC, 2, 40, 33, G_SR) as s:
        s.upload(W)
        s.sweep(2, 0, 1 / 40, 0.1 / 40)
        st = native.Status()
        assert lib.cosmos_status_async(s._h, ctypes.byref(st)) == 0
        s.synchronize()
        sync = s.status()
        assert {k: int(getattr(st, k)) for k in sync} == sync
