"""Host-side checks of the batched form (no GPU): the seeded 1024-problem
workload and the argument checks of cosmos_batch_create."""
import ctypes
import math

import numpy as np

from cosmos_amd import native, workloads


def test_riemann_batch_workload():
    problems, states = workloads.riemann_batch(1024, seed=0, steps=200)
    again, _ = workloads.riemann_batch(1024, seed=0, steps=200)
    assert problems == again
    assert len(problems) == len(states) == 1024
    cells = np.array([p["nx"] * p["ny"] for p in problems])
    # about 24.6 M cells and 4.9e9 cell-updates expected for this mix
    assert 2.0e7 < cells.sum() < 3.0e7
    sr = sum(p["physics"] for p in problems) / 1024
    two_d = sum(p["dim"] == 2 for p in problems) / 1024
    assert abs(sr - 0.3) < 0.05 and abs(two_d - 0.5) < 0.05
    for p, W in zip(problems, states):
        n = p["nx"]
        if p["dim"] == 1:
            assert 64 <= n <= 4096 and W.shape == (n, 5) and p["ny"] == 1
            left, right = W[0], W[-1]
            assert np.all(W[: n // 2] == left) and np.all(W[n // 2:] == right)
        else:
            assert 32 <= n <= 512 and W.shape == (n, n, 5) and p["ny"] == n
            left, right = W[0, 0], W[0, -1]
            assert np.all(W[:, : n // 2] == left) and np.all(W[:, n // 2:] == right)
        for s in (left, right):
            assert 0.1 <= s[0] <= 10.0 and 0.1 <= s[4] <= 10.0 and abs(s[1]) <= 0.8
            assert s[2] == 0.0 and s[3] == 0.0
        assert p["cell_spacing"] == 1.0 / n and p["max_steps"] == 200 and math.isinf(p["final_time"])
        assert (p["cfl"], p["adiabatic_index"]) == ((0.5, 5.0 / 3.0) if p["physics"] else (0.9, 1.4))


def _problem(**kw):
    p = native.BatchProblem(physics=0, dim=1, nx=16, ny=1, cell_spacing=1 / 16, cfl=0.9, final_time=math.inf,
                            max_steps=10, adiabatic_index=1.4)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_batch_create_rejects_bad_problems():
    lib = native.load()
    h = ctypes.c_void_p()
    bad = [dict(physics=2), dict(dim=3), dict(nx=1), dict(dim=2, ny=1), dict(cfl=0.0), dict(cell_spacing=-1.0),
           dict(max_steps=0), dict(max_steps=2_000_000), dict(adiabatic_index=1.0)]
    for kw in bad:
        arr = (native.BatchProblem * 2)(_problem(), _problem(**kw))
        assert lib.cosmos_batch_create(ctypes.byref(h), arr, 2, 2, 0.0, 1) == -1, kw
        assert b"problem 1" in lib.cosmos_batch_last_error(None), kw
    arr = (native.BatchProblem * 1)(_problem())
    assert lib.cosmos_batch_create(ctypes.byref(h), arr, 1, 3, 0.0, 1) == -1
    assert lib.cosmos_batch_create(ctypes.byref(h), arr, 0, 2, 0.0, 1) == -1
    assert lib.cosmos_batch_create(None, arr, 1, 2, 0.0, 1) == -1
    assert lib.cosmos_batch_run(None) == -1
    s = native.BatchSchedule()
    assert lib.cosmos_batch_default_schedule(0, ctypes.byref(s)) == 0
    assert s.whole_max_cells > 0 and s.x_segments_per_task >= 1 and s.y_rows >= 1
