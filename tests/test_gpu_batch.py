"""GPU parity of the batched persistent form (cosmos_batch_*): every problem of
a mixed batch (Newtonian and special relativistic, 1D and 2D, uneven and
ragged sizes, whole-block and tiled execution, packed short rows and rows
longer than one segment) against the CPU oracle's solve loop, bitwise, for
several task schedules; the full 1024-problem workload against the per-problem
solver on a sample."""
import math

import numpy as np
import pytest

from conftest import rel_linf, same
from cosmos_amd import Solver, workloads
from cosmos_amd.batch import Batch, cell_updates

pytestmark = pytest.mark.gpu

G_N, G_SR = 1.4, 5.0 / 3.0


def _state(shape, rng, physics):
    W = np.empty(tuple(shape) + (5,))
    W[..., 0] = rng.uniform(0.2, 4.0, shape)
    vmax = 0.5 if physics else 1.0
    W[..., 1] = rng.uniform(-vmax, vmax, shape)
    W[..., 2] = rng.uniform(-vmax, vmax, shape) if len(shape) == 2 else 0.0
    W[..., 3] = 0.0
    W[..., 4] = rng.uniform(0.2, 4.0, shape)
    if physics:
        v2 = W[..., 1] ** 2 + W[..., 2] ** 2
        W[..., 1:3] *= np.where(v2 > 0.7, np.sqrt(0.7 / np.maximum(v2, 1e-300)), 1.0)[..., None]
    return W


def mixed_batch(seed=5, steps=9):
    """Small problems covering the execution paths: (physics, dim, nx, ny, final_time)."""
    shapes = [
        (0, 1, 16, 1, math.inf), (0, 1, 37, 1, math.inf), (0, 1, 251, 1, math.inf), (0, 1, 253, 1, math.inf),
        (0, 1, 600, 1, math.inf), (1, 1, 40, 1, math.inf), (1, 1, 300, 1, math.inf), (0, 1, 2, 1, math.inf),
        (0, 2, 6, 5, math.inf), (0, 2, 13, 29, math.inf), (0, 2, 32, 32, math.inf), (0, 2, 48, 40, math.inf),
        (0, 2, 260, 6, math.inf), (0, 2, 5, 70, math.inf), (1, 2, 9, 11, math.inf), (1, 2, 24, 20, math.inf),
        (1, 2, 40, 36, math.inf), (1, 2, 255, 4, math.inf), (0, 2, 2, 2, math.inf), (0, 1, 90, 1, 0.004),
        (1, 2, 20, 20, 0.003), (0, 2, 64, 64, math.inf),
    ]
    rng = np.random.default_rng(seed)
    problems, states = [], []
    for k, (phys, dim, nx, ny, tf) in enumerate(shapes):
        shape = (ny, nx) if dim == 2 else (nx,)
        states.append(_state(shape, rng, phys))
        problems.append({"physics": phys, "dim": dim, "nx": nx, "ny": ny, "cell_spacing": 1.0 / max(nx, ny),
                         "cfl": 0.5 if phys else 0.9, "final_time": tf, "max_steps": steps + (k % 3),
                         "adiabatic_index": G_SR if phys else G_N})
    return problems, states


def oracle_runs(orc, problems, states, order, bias, limiter, runs=1):
    out = []
    for p, W in zip(problems, states):
        for _ in range(runs):
            W, t, n = orc.run(W, p["physics"], order, p["cell_spacing"], p["cfl"], p["final_time"], p["max_steps"],
                              bias, limiter, p["adiabatic_index"])
        out.append((W, t, n))
    return out


def run_batch(problems, states, order=2, bias=0.0, limiter=1, schedules=None, runs=1, c2p=None):
    with Batch(problems, order, bias, limiter) as b:
        if c2p:
            b.set_con2prim(*c2p)
        for phys, fields in (schedules or {}).items():
            b.set_schedule(phys, **fields)
        for k, W in enumerate(states):
            b.upload(k, W)
        for _ in range(runs):
            b.run()
        t, steps, fin = b.stats()
        return [b.download(k) for k in range(len(problems))], t, steps, fin


SCHEDULES = {
    "default": None,
    "all_whole": {0: {"whole_max_cells": 1 << 20, "y_tile_cols": 0}, 1: {"whole_max_cells": 1 << 20, "y_tile_cols": 0}},
    "all_whole_tiles": {0: {"whole_max_cells": 1 << 20, "y_tile_cols": 16},
                        1: {"whole_max_cells": 1 << 20, "y_tile_cols": 8}},
    "all_tiled": {0: {"whole_max_cells": 0, "x_segments_per_task": 1, "y_rows": 4, "y_tile_cols": 0},
                  1: {"whole_max_cells": 0, "x_segments_per_task": 1, "y_rows": 4, "y_tile_cols": 0}},
    "tiled_odd": {0: {"whole_max_cells": 0, "x_segments_per_task": 3, "y_rows": 7, "blocks_per_cu": 1,
                      "y_tile_cols": 0},
                  1: {"whole_max_cells": 100, "x_segments_per_task": 2, "y_rows": 5, "blocks_per_cu": 1,
                      "y_tile_cols": 0}},
    "tiles": {0: {"whole_max_cells": 0, "x_segments_per_task": 1, "y_tile_cols": 8},
              1: {"whole_max_cells": 300, "x_segments_per_task": 3, "y_tile_cols": 16}},
}


@pytest.fixture(scope="module")
def mixed(orc):
    problems, states = mixed_batch()
    return problems, states, oracle_runs(orc, problems, states, 2, 0.0, 1)


@pytest.mark.parametrize("name", sorted(SCHEDULES))
def test_batch_matches_oracle(mixed, name):
    problems, states, exp = mixed
    got, t, steps, fin = run_batch(problems, states, schedules=SCHEDULES[name])
    for k, (W, (E, te, ne)) in enumerate(zip(got, exp)):
        assert same(W, E), (name, k, problems[k], rel_linf(W, E))
        assert t[k] == te and steps[k] == ne, (name, k, t[k], te, steps[k], ne)
    assert np.all(fin > 0) and np.all(np.isfinite(fin))


@pytest.mark.parametrize("order,bias,limiter", [(1, 0.0, 1), (2, 0.3, 0), (2, -0.2, 2)])
def test_batch_other_schemes(orc, order, bias, limiter):
    problems, states = mixed_batch(seed=6, steps=6)
    exp = oracle_runs(orc, problems, states, order, bias, limiter)
    for name in ("tiled_odd", "tiles"):
        got, t, steps, _ = run_batch(problems, states, order, bias, limiter, schedules=SCHEDULES[name])
        for k, (W, (E, te, ne)) in enumerate(zip(got, exp)):
            assert same(W, E), (name, k, problems[k], rel_linf(W, E))
            assert t[k] == te and steps[k] == ne
    got, t, steps, _ = run_batch(problems, states, order, bias, limiter, schedules=SCHEDULES["all_whole_tiles"])
    for k, (W, (E, te, ne)) in enumerate(zip(got, exp)):
        assert same(W, E), (k, problems[k], rel_linf(W, E))
        assert t[k] == te and steps[k] == ne


def test_batch_repeated_runs_and_tight_con2prim(orc):
    problems, states = mixed_batch(seed=7, steps=5)
    orc.set_con2prim(1e-12, 50)
    try:
        exp = oracle_runs(orc, problems, states, 2, 0.0, 1, runs=2)
    finally:
        orc.set_con2prim(1e-8, 1000)
    got, t, steps, _ = run_batch(problems, states, schedules=SCHEDULES["all_tiled"], runs=2, c2p=(1e-12, 50))
    for k, (W, (E, te, ne)) in enumerate(zip(got, exp)):
        assert same(W, E), (k, problems[k], rel_linf(W, E))
        assert t[k] == te and steps[k] == ne


def test_batch_full_workload_against_solver():
    """The 1024-problem workload (200 steps each) in one launch; a sample of
    problems of every kind compared with the per-problem solver, bitwise."""
    problems, states = workloads.riemann_batch(1024, seed=0, steps=200)
    with Batch(problems) as b:
        for k, W in enumerate(states):
            b.upload(k, W)
        ms = b.run_timed()
        t, steps, fin = b.stats()
        assert np.all(steps == 200)
        assert cell_updates(problems, steps) > 4.8e9
        assert 0 < fin.max() <= ms * 1e3 * 1.01
        cells = np.array([p["nx"] * p["ny"] for p in problems])
        picks = set()
        for phys in (0, 1):
            for dim in (1, 2):
                idx = [k for k, p in enumerate(problems) if p["physics"] == phys and p["dim"] == dim]
                idx.sort(key=lambda k: cells[k])
                picks.update([idx[0], idx[len(idx) // 2], idx[-1]])
        for k in sorted(picks):
            p = problems[k]
            with Solver(p["physics"], p["dim"], p["nx"], p["ny"], p["adiabatic_index"]) as s:
                s.upload(states[k])
                te, ne = s.solve(2, p["cell_spacing"], p["cfl"], math.inf, 200)
                E = s.download()
            W = b.download(k)
            assert same(W, E), (k, p, rel_linf(W, E))
            assert t[k] == te and steps[k] == ne


def test_batch_relativistic_extreme_states_bitwise(orc):
    """Special-relativistic problems holding rootless, clamped and floored
    states (1D and 2D, whole-block and tiled) in one batch with regular
    Newtonian ones, each bitwise the oracle's solve loop."""
    from conftest import sr_extreme_2d

    shapes = [(17, 9), (40, 33), (300, 70), (5, 4)]
    problems, states = [], []
    for k, (nx, ny) in enumerate(shapes):
        W = sr_extreme_2d(nx, ny, seed=3 + k)
        for dim, st in ((2, W), (1, W[ny // 2].copy())):
            states.append(st)
            problems.append({"physics": 1, "dim": dim, "nx": nx, "ny": ny if dim == 2 else 1, "cell_spacing": 1 / 40,
                             "cfl": 0.5, "final_time": math.inf, "max_steps": 4 + k, "adiabatic_index": G_SR})
    rng = np.random.default_rng(9)
    states.append(_state((24, 20), rng, 0))
    problems.append({"physics": 0, "dim": 2, "nx": 20, "ny": 24, "cell_spacing": 1 / 24, "cfl": 0.9,
                     "final_time": math.inf, "max_steps": 6, "adiabatic_index": G_N})
    exp = oracle_runs(orc, problems, states, 2, 0.0, 1)
    for sched in (None, {1: {"whole_max_cells": 0}}):
        outs, t, steps, _ = run_batch(problems, states, schedules=sched)
        for k, (W, te, ne) in enumerate(exp):
            assert steps[k] == ne and t[k] == te, (k, sched)
            assert same(outs[k], W), (k, sched, rel_linf(outs[k], W))

