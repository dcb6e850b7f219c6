"""The reference's scenario drivers (Tests/EulerTests.cpp, Tests/RelativisticEulerTests.cpp,
Tests/BlackHoleEulerTests.cpp) as initial data plus run parameters, for
regression runs through the library and comparison with the oracle.

Each scenario restates one driver's set-up: the cell-index conditions and
states (the drivers compare i (and j) against fractions of cellCount), the
adiabatic index, the CFL number and the final time, with cell spacing
1 / cellCount; the drivers then call solve / solve2D with bias 0 and van Leer
(second order) or the FORCE solver (first order), and the black-hole ones with
`subcyclingIterations` forcing sub-steps.  Transcendental initial data
(sin, atan2, sqrt) use Python's math module (the C library's functions).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NEWTONIAN, RELATIVISTIC, BLACKHOLE = 0, 1, 2


@dataclass
class Scenario:
    name: str
    physics: int
    dim: int
    gamma: float
    cfl: float
    final_time: float
    W: np.ndarray
    hole: tuple = field(default=(1.0, 0.0, 0.5, 0.5, 0.0))

    @property
    def cell_spacing(self) -> float:
        return 1.0 / self.W.shape[-2]


def _line(n, inside, outside, cond):
    W = np.empty((n, 5))
    for i in range(n):
        W[i] = inside(i) if cond(i) else outside(i)
    return W


def _disc(n, inside, outside, cond):
    W = np.empty((n, n, 5))
    for i in range(n):
        for j in range(n):
            W[i, j] = inside(i, j) if cond(i, j) else outside(i, j)
    return W


def _left(n):
    return lambda i: i <= (0.5 * n)


def _centre(n, radius_frac):
    return lambda i, j: math.sqrt(((i - (0.5 * n)) * (i - (0.5 * n))) + ((j - (0.5 * n)) * (j - (0.5 * n)))) <= (
        radius_frac * n)


def _const(*s):
    return lambda *ij: s


# Tests/EulerTests.cpp:3-187 (gamma 1.4, CFL 0.95)
def toro1(n: int) -> Scenario:  # :3-30
    return Scenario("toro1", NEWTONIAN, 1, 1.4, 0.95, 0.25,
                    _line(n, _const(1.0, 0.0, 0.0, 0.0, 1.0), _const(0.125, 0.0, 0.0, 0.0, 0.1), _left(n)))


def toro2(n: int) -> Scenario:  # :32-59
    return Scenario("toro2", NEWTONIAN, 1, 1.4, 0.95, 0.15,
                    _line(n, _const(1.0, -2.0, 0.0, 0.0, 0.4), _const(1.0, 2.0, 0.0, 0.0, 0.4), _left(n)))


def toro3(n: int) -> Scenario:  # :61-88
    return Scenario("toro3", NEWTONIAN, 1, 1.4, 0.95, 0.012,
                    _line(n, _const(1.0, 0.0, 0.0, 0.0, 1000.0), _const(1.0, 0.0, 0.0, 0.0, 0.01), _left(n)))


def toro1_2d(n: int) -> Scenario:  # :90-120
    return Scenario("toro1_2d", NEWTONIAN, 2, 1.4, 0.95, 0.125,
                    _disc(n, _const(1.0, 0.0, 0.0, 0.0, 1.0), _const(0.125, 0.0, 0.0, 0.0, 0.1), _centre(n, 0.2)))


def toro2_2d(n: int) -> Scenario:  # :122-154
    def ang(i, j):
        return math.atan2(j - (0.5 * n), i - (0.5 * n))

    inside = lambda i, j: (1.0, -2.0 * math.sin(ang(i, j)), -2.0 * math.cos(ang(i, j)), 0.0, 0.4)  # noqa: E731
    outside = lambda i, j: (1.0, 2.0 * math.sin(ang(i, j)), 2.0 * math.cos(ang(i, j)), 0.0, 0.4)  # noqa: E731
    return Scenario("toro2_2d", NEWTONIAN, 2, 1.4, 0.95, 0.075, _disc(n, inside, outside, _centre(n, 0.2)))


def toro3_2d(n: int) -> Scenario:  # :156-187
    return Scenario("toro3_2d", NEWTONIAN, 2, 1.4, 0.95, 0.006,
                    _disc(n, _const(1.0, 0.0, 0.0, 0.0, 1000.0), _const(1.0, 0.0, 0.0, 0.0, 0.01), _centre(n, 0.2)))


# Tests/RelativisticEulerTests.cpp:3-187 (gamma 5/3, CFL 0.95)
def sr_shock(n: int, order: int = 2) -> Scenario:  # :3-30 (first order runs to 0.4, second to 0.01)
    return Scenario("sr_shock", RELATIVISTIC, 1, 5.0 / 3.0, 0.95, 0.4 if order == 1 else 0.01,
                    _line(n, _const(10.0, 0.0, 0.0, 0.0, 13.3), _const(1.0, 0.0, 0.0, 0.0, 0.0), _left(n)))


def sr_blast(n: int) -> Scenario:  # :32-59
    return Scenario("sr_blast", RELATIVISTIC, 1, 5.0 / 3.0, 0.95, 0.4,
                    _line(n, _const(1.0, 0.0, 0.0, 0.0, 1000.0), _const(1.0, 0.0, 0.0, 0.0, 0.01), _left(n)))


def sr_perturbed(n: int) -> Scenario:  # :61-88
    h = 1.0 / n
    return Scenario("sr_perturbed", RELATIVISTIC, 1, 5.0 / 3.0, 0.95, 0.35,
                    _line(n, _const(5.0, 0.0, 0.0, 0.0, 50.0),
                          lambda i: (2.0 + 0.3 * math.sin(50.0 * (h * i)), 0.0, 0.0, 0.0, 5.0), _left(n)))


def sr_shock_2d(n: int) -> Scenario:  # :90-120
    return Scenario("sr_shock_2d", RELATIVISTIC, 2, 5.0 / 3.0, 0.95, 0.2,
                    _disc(n, _const(10.0, 0.0, 0.0, 0.0, 13.3), _const(1.0, 0.0, 0.0, 0.0, 0.0), _centre(n, 0.2)))


def sr_blast_2d(n: int) -> Scenario:  # :122-152
    return Scenario("sr_blast_2d", RELATIVISTIC, 2, 5.0 / 3.0, 0.95, 0.15,
                    _disc(n, _const(1.0, 0.0, 0.0, 0.0, 1000.0), _const(1.0, 0.0, 0.0, 0.0, 0.01), _centre(n, 0.2)))


def sr_perturbed_2d(n: int) -> Scenario:  # :154-187
    h = 1.0 / n

    def outside(i, j):
        a, b = h * (i - (0.5 * n)), h * (j - (0.5 * n))
        return (2.0 + 0.3 * math.sin(50.0 * math.sqrt((a * a) + (b * b))), 0.0, 0.0, 0.0, 5.0)

    return Scenario("sr_perturbed_2d", RELATIVISTIC, 2, 5.0 / 3.0, 0.95, 0.175,
                    _disc(n, _const(5.0, 0.0, 0.0, 0.0, 50.0), outside, _centre(n, 0.2)))


# Tests/BlackHoleEulerTests.cpp:3-199 (gamma 5/3, CFL 0.95, t = 0.45); holes as
# (M, a, x, y, z): BlackHoleSpacetime(M, x, y, z) has a = 0
def _shell_1d(n):
    return _line(n, _const(10.0, 0.0, 0.0, 0.0, 0.1), _const(0.1, 0.0, 0.0, 0.0, 0.1),
                 lambda i: i >= (0.1 * n) and i <= (0.9 * n))


def bh_spherical_1d(n: int) -> Scenario:  # :3-32
    return Scenario("bh_spherical_1d", BLACKHOLE, 1, 5.0 / 3.0, 0.95, 0.45, _shell_1d(n), (0.05, 0.0, 0.5, 0.0, 0.0))


def bh_spinning_1d(n: int) -> Scenario:  # :34-63
    return Scenario("bh_spinning_1d", BLACKHOLE, 1, 5.0 / 3.0, 0.95, 0.45, _shell_1d(n), (0.1, 0.99, 0.5, 0.0, 0.0))


def _ball(n, ax, ay):
    def cond(i, j):
        return math.sqrt(((i - (0.5 * n)) * (i - (0.5 * n)) * ax) + ((j - (0.5 * n)) * (j - (0.5 * n)) * ay)) <= (
            0.4 * n)

    return _disc(n, _const(10.0, 0.0, 0.0, 0.0, 0.1), _const(0.1, 0.0, 0.0, 0.0, 0.1), cond)


def _ball_round(n):
    return _disc(n, _const(10.0, 0.0, 0.0, 0.0, 0.1), _const(0.1, 0.0, 0.0, 0.0, 0.1), _centre(n, 0.4))


def bh_spherical_2d(n: int) -> Scenario:  # :65-97
    return Scenario("bh_spherical_2d", BLACKHOLE, 2, 5.0 / 3.0, 0.95, 0.45, _ball_round(n), (0.05, 0.0, 0.5, 0.5, 0.0))


def bh_spinning_2d(n: int) -> Scenario:  # :99-131
    return Scenario("bh_spinning_2d", BLACKHOLE, 2, 5.0 / 3.0, 0.95, 0.45, _ball_round(n), (0.1, 0.99, 0.5, 0.5, 0.0))


def bh_spheroidal_2d(n: int) -> Scenario:  # :133-165
    return Scenario("bh_spheroidal_2d", BLACKHOLE, 2, 5.0 / 3.0, 0.95, 0.45, _ball(n, 0.8, 1.2),
                    (0.05, 0.0, 0.5, 0.5, 0.0))


def bh_spinning_spheroidal_2d(n: int) -> Scenario:  # :167-199
    return Scenario("bh_spinning_spheroidal_2d", BLACKHOLE, 2, 5.0 / 3.0, 0.95, 0.45, _ball(n, 0.8, 1.2),
                    (0.1, 0.99, 0.5, 0.5, 0.0))


ALL = [toro1, toro2, toro3, toro1_2d, toro2_2d, toro3_2d, sr_shock, sr_blast, sr_perturbed, sr_shock_2d,
       sr_blast_2d, sr_perturbed_2d, bh_spherical_1d, bh_spinning_1d, bh_spherical_2d, bh_spinning_2d,
       bh_spheroidal_2d, bh_spinning_spheroidal_2d]


# ---------------------------------------------------------------------------
# The AMR drivers of Tests/EulerAMRTests.cpp, Tests/RelativisticEulerAMRTests.cpp
# and Tests/BlackHoleEulerAMRTests.cpp: their set-ups (CFL 0.95 throughout),
# with the tolerance, order and level main.cpp:20-46 passes to each.
# ---------------------------------------------------------------------------


@dataclass
class AmrScenario:
    scenario: Scenario
    tolerance: float
    order: int
    levels: int


def _amr(name, physics, dim, gamma, final_time, W, tol, order, levels, hole=(1.0, 0.0, 0.5, 0.5, 0.0)):
    return AmrScenario(Scenario(name, physics, dim, gamma, 0.95, final_time, W, hole), tol, order, levels)


def euler_toro1_amr(n: int) -> AmrScenario:  # Tests/EulerAMRTests.cpp:3-30
    W = _line(n, _const(1.0, 0.0, 0.0, 0.0, 1.0), _const(0.125, 0.0, 0.0, 0.0, 0.1), _left(n))
    return _amr("euler_toro1_amr", NEWTONIAN, 1, 1.4, 0.25, W, 1.0, 2, 2)


def euler_toro1_2d_amr(n: int) -> AmrScenario:  # :32-62
    W = _disc(n, _const(1.0, 0.0, 0.0, 0.0, 1.0), _const(0.125, 0.0, 0.0, 0.0, 0.1), _centre(n, 0.2))
    return _amr("euler_toro1_2d_amr", NEWTONIAN, 2, 1.4, 0.125, W, 1.0, 2, 2)


def sr_mild_shock_amr(n: int) -> AmrScenario:  # Tests/RelativisticEulerAMRTests.cpp:3-30
    W = _line(n, _const(10.0, 0.0, 0.0, 0.0, 13.3), _const(1.0, 0.0, 0.0, 0.0, 0.0), _left(n))
    return _amr("sr_mild_shock_amr", RELATIVISTIC, 1, 5.0 / 3.0, 0.4, W, 5.0, 2, 2)


def sr_strong_blast_amr(n: int) -> AmrScenario:  # :32-59
    W = _line(n, _const(1.0, 0.0, 0.0, 0.0, 1000.0), _const(1.0, 0.0, 0.0, 0.0, 0.01), _left(n))
    return _amr("sr_strong_blast_amr", RELATIVISTIC, 1, 5.0 / 3.0, 0.4, W, 100.0, 2, 2)


def sr_perturbed_amr(n: int) -> AmrScenario:  # :61-88
    h = 1.0 / n
    W = _line(n, _const(5.0, 0.0, 0.0, 0.0, 50.0),
              lambda i: (2.0 + 0.3 * math.sin(50 * (h * i)), 0.0, 0.0, 0.0, 5.0), _left(n))
    return _amr("sr_perturbed_amr", RELATIVISTIC, 1, 5.0 / 3.0, 0.35, W, 10.0, 2, 2)


def sr_mild_shock_2d_amr(n: int) -> AmrScenario:  # :90-120
    W = _disc(n, _const(10.0, 0.0, 0.0, 0.0, 13.3), _const(1.0, 0.0, 0.0, 0.0, 0.0), _centre(n, 0.2))
    return _amr("sr_mild_shock_2d_amr", RELATIVISTIC, 2, 5.0 / 3.0, 0.2, W, 10.0, 2, 2)


def sr_strong_blast_2d_amr(n: int) -> AmrScenario:  # :122-152
    W = _disc(n, _const(1.0, 0.0, 0.0, 0.0, 1000.0), _const(1.0, 0.0, 0.0, 0.0, 0.01), _centre(n, 0.2))
    return _amr("sr_strong_blast_2d_amr", RELATIVISTIC, 2, 5.0 / 3.0, 0.2, W, 1000.0, 1, 2)


def sr_perturbed_2d_amr(n: int) -> AmrScenario:  # :154-187
    h = 1.0 / n

    def outside(i, j):
        a, b = h * (i - (0.5 * n)), h * (j - (0.5 * n))
        return (2.0 + 0.3 * math.sin(50.0 * math.sqrt((a * a) + (b * b))), 0.0, 0.0, 0.0, 5.0)

    W = _disc(n, _const(5.0, 0.0, 0.0, 0.0, 50.0), outside, _centre(n, 0.2))
    return _amr("sr_perturbed_2d_amr", RELATIVISTIC, 2, 5.0 / 3.0, 0.175, W, 20.0, 2, 2)


def _bh_amr(make, levels):
    def f(n: int) -> AmrScenario:
        S = make(n)
        return AmrScenario(S, 5.0, 2, levels)

    f.__name__ = make.__name__ + "_amr"
    return f


# Tests/BlackHoleEulerAMRTests.cpp:3-131: the black-hole scenario states; main.cpp runs the 1D ones at
# level 2 and the 2D ones at level 1 (its 2D level-2 driver is truncated)
bh_spherical_1d_amr = _bh_amr(bh_spherical_1d, 2)
bh_spinning_1d_amr = _bh_amr(bh_spinning_1d, 2)
bh_spherical_2d_amr = _bh_amr(bh_spherical_2d, 1)
bh_spinning_2d_amr = _bh_amr(bh_spinning_2d, 1)
bh_spinning_spheroidal_2d_amr = _bh_amr(bh_spinning_spheroidal_2d, 1)

AMR_ALL = [euler_toro1_amr, euler_toro1_2d_amr, sr_mild_shock_amr, sr_strong_blast_amr, sr_perturbed_amr,
           sr_mild_shock_2d_amr, sr_strong_blast_2d_amr, sr_perturbed_2d_amr, bh_spherical_1d_amr,
           bh_spinning_1d_amr, bh_spherical_2d_amr, bh_spinning_2d_amr, bh_spinning_spheroidal_2d_amr]

