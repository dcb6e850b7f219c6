"""cosmos_step_noalloc (the step of the solve loop on caller-owned device
planes, with its working memory in a fixed device array instead of
allocations) against the existing entry point cosmos_solve over one step:
bitwise equal states and time steps, on two calls in a row, the second with
a smaller input, for Newtonian and special-relativistic, 1D and 2D, FORCE and
SLIC, damped and undamped steps and a final-time clamp."""
import ctypes
import math

import numpy as np
import pytest

from conftest import same
from cosmos_amd import NEWTONIAN, RELATIVISTIC, Solver, native, workloads

pytestmark = pytest.mark.gpu


def _holder(W, physics, gamma):
    """A solver whose device state buffer serves as plain device memory: its
    five planes are contiguous (plane stride pitch * ny), the layout the entry
    point takes."""
    dim = 1 if W.ndim == 2 else 2
    nx = W.shape[-2]
    ny = W.shape[0] if dim == 2 else 1
    s = Solver(physics, dim, nx, ny, gamma)
    s.upload(W)
    s.synchronize()
    return s


def noalloc_step(W, physics, order, dx, cfl, t, tf, it, gamma, bias=0.0, limiter=1):
    lib = native.load()
    dim = 1 if W.ndim == 2 else 2
    nx = W.shape[-2]
    ny = W.shape[0] if dim == 2 else 1
    with _holder(W, physics, gamma) as src, _holder(np.zeros_like(W), physics, gamma) as dst, \
            _holder(np.full((2, 5), np.nan), NEWTONIAN, 1.4) as dtbuf:
        (sp, pitch), (dp, dpitch), (tp, _) = src.state_planes(), dst.state_planes(), dtbuf.state_planes()
        assert pitch == dpitch
        rc = lib.cosmos_step_noalloc(physics, dim, nx, ny, pitch, sp[0], dp[0], order, dx, cfl, t, tf, it, bias,
                                     limiter, gamma, tp[0], dst.stream)
        assert rc == 0, lib.cosmos_noalloc_last_error()
        dst.synchronize()
        out = dst.download()
        assert same(src.download(), W)  # the input is not written
        dt = float(dtbuf.download()[0, 0])
    return out, dt


def solver_step(W, physics, order, dx, cfl, t, tf, it, gamma, bias=0.0, limiter=1):
    dim = 1 if W.ndim == 2 else 2
    nx = W.shape[-2]
    ny = W.shape[0] if dim == 2 else 1
    with Solver(physics, dim, nx, ny, gamma) as s:
        s.upload(W)
        s.begin(order, dx, cfl, tf, 1, bias, limiter, start_time=t, start_iteration=it)
        s.advance(1)
        t1, n, _ = s.progress()
        return s.download(), t1 - t if t == 0.0 else None, n


CASES = [
    # (state, physics, order, cfl, t, tf, iteration, gamma, limiter, bias)
    ("quad", NEWTONIAN, 2, 0.9, 0.0, math.inf, 0, 1.4, 1, 0.0),
    ("sr", RELATIVISTIC, 1, 0.5, 0.0, math.inf, 0, 5.0 / 3.0, 1, 0.0),
    ("quad", NEWTONIAN, 2, 0.9, 0.0, 1e-5, 7, 1.4, 0, 0.3),
    ("sr", RELATIVISTIC, 2, 0.5, 0.0, math.inf, 9, 5.0 / 3.0, 2, 0.0),
]


def _state(kind, n, ny, seed, dim):
    W = workloads.quadrants_2d(n, seed=seed, ny=ny) if kind == "quad" else workloads.sr_blast_2d(max(n, ny), seed)[:ny, :n]
    W = np.ascontiguousarray(W)
    return W if dim == 2 else np.ascontiguousarray(W[ny // 2])


@pytest.mark.parametrize("dim", [2, 1])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_step_matches_cosmos_solve_twice(case, dim):
    kind, physics, order, cfl, t, tf, it, gamma, lim, bias = CASES[case]
    for n, ny, seed in ((300, 170, 1), (97, 40, 2)):  # the second call on a smaller grid
        W = _state(kind, n, ny, seed, dim)
        dx = 1.0 / n
        out, dt = noalloc_step(W, physics, order, dx, cfl, t, tf, it, gamma, bias, lim)
        exp, dt_exp, steps = solver_step(W, physics, order, dx, cfl, t, tf, it, gamma, bias, lim)
        assert steps == 1
        assert same(out, exp), (n, ny, float(np.nanmax(np.abs(out - exp))))
        if dt_exp is not None:
            assert dt == dt_exp
        if tf != math.inf:
            assert dt == tf - t


def test_noalloc_rejects_oversized_and_bad_arguments():
    lib = native.load()
    limit = lib.cosmos_noalloc_grid_limit()
    ny = 1000
    pitch = limit // (5 * ny) + 1
    rc = lib.cosmos_step_noalloc(NEWTONIAN, 2, pitch, ny, pitch, 16, 32, 2, 0.1, 0.9, 0.0, 1.0, 0, 0.0, 1, 1.4,
                                 None, None)
    assert rc == -1  # COSMOS_ERR_ARGUMENT, before any device work
    assert b"exceeds the no-allocation limit" in lib.cosmos_noalloc_last_error()
    rc = lib.cosmos_step_noalloc(NEWTONIAN, 2, 8, 8, 8, 16, 32, 2, 0.1, 0.9, 1.0, 1.0, 0, 0.0, 1, 1.4, None, None)
    assert rc == -1 and b"final_time" in lib.cosmos_noalloc_last_error()
