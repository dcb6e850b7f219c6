"""Regression runs of the reference's scenario drivers (Tests/EulerTests.cpp,
Tests/RelativisticEulerTests.cpp, Tests/BlackHoleEulerTests.cpp) through the
C-ABI, each bitwise against the oracle's solve loop on the same initial data:
Toro 1-3 in 1D and 2D, the special-relativistic shock, blast and perturbed
density in 1D and 2D, and black-hole accretion (spherical, spinning,
spheroidal) in 1D and 2D, at both orders.  Grids are reduced so that the
oracle finishes in seconds; the Newtonian and special-relativistic runs go to
the drivers' own final times, the black-hole ones stop after a few steps (or
at the final time, whichever comes first)."""
import math

import numpy as np
import pytest

from conftest import same
from cosmos_amd import BLACKHOLE, Solver, scenarios

pytestmark = pytest.mark.gpu

SIZES = {1: 160, 2: 40}
BH_SIZES = {1: 64, 2: 24}
BH_STEPS = {1: 40, 2: 6}


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("make", [f for f in scenarios.ALL if f(8).physics != scenarios.BLACKHOLE],
                         ids=lambda f: f.__name__)
def test_scenario_matches_oracle(orc, make, order):
    sc = make(SIZES[make(8).dim], order) if make is scenarios.sr_shock else make(SIZES[make(8).dim])
    nx = sc.W.shape[-2]
    ny = sc.W.shape[0] if sc.dim == 2 else 1
    with Solver(sc.physics, sc.dim, nx, ny, sc.gamma) as s:
        s.upload(sc.W)
        t, k = s.solve(order, sc.cell_spacing, sc.cfl, sc.final_time)
        out = s.download()
    exp, te, ke = orc.run(sc.W, sc.physics, order, sc.cell_spacing, sc.cfl, tf=sc.final_time, gamma=sc.gamma)
    assert t == sc.final_time or t == te
    assert (k, t) == (ke, te) and k > 3
    assert same(out, exp), sc.name


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("make", [f for f in scenarios.ALL if f(8).physics == scenarios.BLACKHOLE],
                         ids=lambda f: f.__name__)
def test_blackhole_scenario_matches_oracle(orc, make, order):
    probe = make(8)
    sc = make(BH_SIZES[probe.dim])
    steps = BH_STEPS[sc.dim]
    nx = sc.W.shape[-2]
    ny = sc.W.shape[0] if sc.dim == 2 else 1
    with Solver(BLACKHOLE, sc.dim, nx, ny, sc.gamma) as s:
        s.set_spacetime(*sc.hole)
        s.upload(sc.W)
        t, k = s.solve_forced(order, sc.cell_spacing, sc.cfl, sc.final_time, steps)
        out = s.download()
    exp, te, ke = orc.gr_run(sc.W, order, sc.cell_spacing, sc.cfl, sc.hole, tf=sc.final_time, max_steps=steps,
                             subcycles=1, gamma=sc.gamma)
    assert (k, t) == (ke, te) and (k == steps or t == sc.final_time) and k > 3
    assert same(out, exp), sc.name
