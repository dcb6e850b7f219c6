"""The scenario set-ups restate the reference's drivers: checks of the cell
conditions at their edges and of the states (Tests/EulerTests.cpp,
Tests/RelativisticEulerTests.cpp, Tests/BlackHoleEulerTests.cpp)."""
import math

import numpy as np

from cosmos_amd import scenarios


def test_riemann_split_includes_the_midpoint_cell():
    # i <= 0.5 cellCount: cell n/2 takes the left state (unlike the x < 0.5 split)
    sc = scenarios.toro1(10)
    assert list(sc.W[:6, 0]) == [1.0] * 6 and list(sc.W[6:, 0]) == [0.125] * 4
    assert sc.cell_spacing == 0.1 and sc.final_time == 0.25 and sc.gamma == 1.4
    sc = scenarios.toro2(9)  # 0.5 * 9 = 4.5: cells 0..4 left
    assert list(sc.W[:, 1]) == [-2.0] * 5 + [2.0] * 4
    assert scenarios.sr_shock(10, 1).final_time == 0.4 and scenarios.sr_shock(10, 2).final_time == 0.01


def test_disc_radius_is_inclusive_and_centred_on_the_cell_index():
    n = 10
    sc = scenarios.toro1_2d(n)
    inside = sc.W[..., 0] == 1.0
    i, j = np.mgrid[0:n, 0:n]
    assert np.array_equal(inside, np.sqrt((i - 5.0) ** 2 + (j - 5.0) ** 2) <= 2.0)
    assert inside[5, 7] and inside[3, 5] and not inside[4, 3]  # radius exactly 2 is inside


def test_toro2_2d_velocity_points_along_atan2_of_j_then_i():
    n = 20
    sc = scenarios.toro2_2d(n)
    i, j = 3, 11  # outside the disc
    a = math.atan2(j - 10.0, i - 10.0)
    assert tuple(sc.W[i, j]) == (1.0, 2.0 * math.sin(a), 2.0 * math.cos(a), 0.0, 0.4)
    i, j = 12, 9  # inside
    a = math.atan2(j - 10.0, i - 10.0)
    assert tuple(sc.W[i, j]) == (1.0, -2.0 * math.sin(a), -2.0 * math.cos(a), 0.0, 0.4)


def test_perturbed_density_and_blackhole_set_ups():
    n = 40
    sc = scenarios.sr_perturbed(n)
    assert sc.W[30, 0] == 2.0 + 0.3 * math.sin(50.0 * ((1.0 / n) * 30)) and sc.W[20, 4] == 50.0
    sc = scenarios.sr_perturbed_2d(n)
    a, b = (1.0 / n) * (0 - 20.0), (1.0 / n) * (3 - 20.0)
    assert sc.W[0, 3, 0] == 2.0 + 0.3 * math.sin(50.0 * math.sqrt(a * a + b * b))
    sc = scenarios.bh_spherical_1d(n)
    assert list(sc.W[[3, 4, 36, 37], 0]) == [0.1, 10.0, 10.0, 0.1]  # 0.1 n <= i <= 0.9 n
    assert sc.hole == (0.05, 0.0, 0.5, 0.0, 0.0) and scenarios.bh_spinning_2d(n).hole == (0.1, 0.99, 0.5, 0.5, 0.0)
    sp = scenarios.bh_spheroidal_2d(n).W[..., 0] == 10.0
    i, j = np.mgrid[0:n, 0:n]
    assert np.array_equal(sp, np.sqrt((i - 20.0) ** 2 * 0.8 + (j - 20.0) ** 2 * 1.2) <= 16.0)
    assert all(f(8).final_time > 0 for f in scenarios.ALL) and len(scenarios.ALL) == 18
