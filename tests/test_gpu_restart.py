"""Checkpoint / restart and the reference's .dat text output, through the C-ABI:
  - a checkpoint round trip restores the state bitwise with its time and
    iteration counter, and a run resumed from it (cosmos_begin_at /
    cosmos_solve_forced_at) is bitwise the uninterrupted run (the first six
    iterations' dt damping keys on the restored counter, so a resume inside
    them is covered too);
  - a mismatched or corrupted file is refused;
  - cosmos_write_dat writes the layout of EulerTests::outputSolution{,2D} and
    BlackHoleEulerTests::outputSolution{,2D}: "x value" / "y x value" rows with the
    C++ stream's default formatting (printf %g), a blank line after each grid
    row, coordinates on the unit square, excised cells as 0."""
import math

import numpy as np
import pytest

from cosmos_amd import BLACKHOLE, NEWTONIAN, RELATIVISTIC, CosmosError, Solver, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("physics,gamma,split", [(NEWTONIAN, 1.4, 3), (NEWTONIAN, 1.4, 9),
                                                 (RELATIVISTIC, 5.0 / 3.0, 4)])
def test_resume_is_bitwise_the_uninterrupted_run(tmp_path, physics, gamma, split):
    n, steps = 48, 12
    W = workloads.quadrants_2d(n, seed=5)
    if physics == RELATIVISTIC:
        W[..., 1:3] *= 0.5
    with Solver(physics, 2, n, n, gamma) as s:
        s.upload(W)
        t_all, k_all = s.solve(2, 1.0 / n, 0.5, math.inf, steps)
        full = s.download()
    path = tmp_path / "run.ck"
    with Solver(physics, 2, n, n, gamma) as s:
        s.upload(W)
        t1, k1 = s.solve(2, 1.0 / n, 0.5, math.inf, split)
        s.checkpoint_save(path, t1, k1)
    with Solver(physics, 2, n, n, gamma) as s:
        t0, it0 = s.checkpoint_load(path)
        assert (t0, it0) == (t1, split)
        s.begin(2, 1.0 / n, 0.5, math.inf, steps - split, start_time=t0, start_iteration=it0)
        s.advance(steps)
        t2, k2, stopped = s.progress()
        assert stopped and k2 == steps - split and t2 == t_all
        assert np.array_equal(s.download(), full)


def test_blackhole_resume_and_spacetime_restore(tmp_path):
    n, steps = 32, 6
    W, h, hole = workloads.gr_random_field_2d(n, seed=3)
    with Solver(BLACKHOLE, 2, n, n, 4.0 / 3.0) as s:
        s.set_spacetime(*hole)
        s.set_excision(2.5)
        s.upload(W)
        t_all, _ = s.solve_forced(2, h, 0.4, math.inf, steps)
        full = s.download()
    path = tmp_path / "bh.ck"
    with Solver(BLACKHOLE, 2, n, n, 4.0 / 3.0) as s:
        s.set_spacetime(*hole)
        s.set_excision(2.5)
        s.upload(W)
        t1, k1 = s.solve_forced(2, h, 0.4, math.inf, 2)
        s.checkpoint_save(path, t1, k1)
    with Solver(BLACKHOLE, 2, n, n, 4.0 / 3.0) as s:  # spacetime and excision come from the file
        t0, it0 = s.checkpoint_load(path)
        t2, k2 = s.solve_forced(2, h, 0.4, math.inf, steps - it0, start_time=t0, start_iteration=it0)
        assert k2 == steps - 2 and t2 == t_all
        assert np.array_equal(s.download(), full)


def test_checkpoint_round_trip_and_refusals(tmp_path):
    n = 20
    rng = np.random.default_rng(1)
    W = rng.uniform(0.1, 2.0, (n, n + 3, 5))
    W[..., 3] = rng.standard_normal((n, n + 3))  # a nonzero z velocity survives too
    path = tmp_path / "a.ck"
    with Solver(NEWTONIAN, 2, n + 3, n, 1.4) as s:
        s.upload(W)
        s.checkpoint_save(path, 0.125, 7)
    with Solver(NEWTONIAN, 2, n + 3, n, 1.4) as s:
        assert s.checkpoint_load(path) == (0.125, 7)
        assert np.array_equal(s.download(), W)
    with Solver(NEWTONIAN, 2, n, n + 3, 1.4) as s:  # transposed grid
        with pytest.raises(CosmosError, match="grid or physics"):
            s.checkpoint_load(path)
    with Solver(NEWTONIAN, 2, n + 3, n, 5.0 / 3.0) as s:
        with pytest.raises(CosmosError, match="adiabatic index"):
            s.checkpoint_load(path)
    raw = bytearray(path.read_bytes())
    raw[200] ^= 0x01
    bad = tmp_path / "bad.ck"
    bad.write_bytes(bytes(raw))
    with Solver(NEWTONIAN, 2, n + 3, n, 1.4) as s:
        with pytest.raises(CosmosError, match="checksum"):
            s.checkpoint_load(bad)
        (tmp_path / "short.ck").write_bytes(bytes(raw[:300]))
        with pytest.raises(CosmosError, match="truncated"):
            s.checkpoint_load(tmp_path / "short.ck")


def expected_dat_2d(W, k, mask=None):
    ny, nx = W.shape[:2]
    hr, hc = 1.0 / ny, 1.0 / nx
    lines = []
    for i in range(ny):
        for j in range(nx):
            v = 0.0 if mask is not None and mask[i, j] else W[i, j, k]
            lines.append("%g %g %g" % (hr * i, hc * j, v))
        lines.append("")
    return "\n".join(lines) + "\n"


NAMES = ["density", "xVelocity", "yVelocity", "zVelocity", "pressure"]


def test_write_dat_layout(tmp_path):
    rng = np.random.default_rng(2)
    W = rng.uniform(-3.0, 3.0, (6, 7, 5)) * 10.0 ** rng.integers(-8, 8, (6, 7, 5))
    with Solver(NEWTONIAN, 2, 7, 6, 1.4) as s:
        s.upload(W)
        s.write_dat(tmp_path)
    for k, name in enumerate(NAMES):
        assert (tmp_path / f"{name}.dat").read_text() == expected_dat_2d(W, k)
    V = rng.uniform(0.1, 2.0, (11, 5))
    d1 = tmp_path / "one"
    d1.mkdir()
    with Solver(NEWTONIAN, 1, 11, 1, 1.4) as s:
        s.upload(V)
        s.write_dat(d1)
    for k, name in enumerate(NAMES):
        want = "".join("%g %g\n" % ((1.0 / 11) * i, V[i, k]) for i in range(11))
        assert (d1 / f"{name}.dat").read_text() == want


def test_write_dat_blackhole_excision(tmp_path):
    n = 16
    rng = np.random.default_rng(3)
    W = rng.uniform(0.5, 1.5, (n, n, 5))
    with Solver(BLACKHOLE, 2, n, n, 4.0 / 3.0) as s:
        s.set_spacetime(0.05, 0.0, 0.5, 0.5, 0.0)  # horizon radius 0.1 on the unit square
        s.upload(W)
        s.write_dat(tmp_path)
    i, j = np.mgrid[0:n, 0:n]
    mask = np.sqrt((j * (1.0 / n) - 0.5) ** 2 + (i * (1.0 / n) - 0.5) ** 2) <= 0.1
    assert 0 < mask.sum() < n * n
    for k, name in enumerate(NAMES):
        assert (tmp_path / f"{name}.dat").read_text() == expected_dat_2d(W, k, mask)
