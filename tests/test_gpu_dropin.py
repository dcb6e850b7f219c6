"""The C++ drop-in classes (cosmos_amd/dropin, same class and method names as
the reference's solvers) called the way the reference's scenario drivers
call them, compared bitwise against the oracle.  The check program is built
by cosmos_amd.build; this test runs it on the GPU and reads its output."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, same, golden
from cosmos_amd import workloads

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "dropin", "dropin_check")


def test_dropin_classes_match_oracle(orc, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/dropin/dropin_check is not built (python -m cosmos_amd.build)")
    out = tmp_path / "dropin.bin"
    subprocess.run([EXE, str(out)], check=True, timeout=120)
    v = np.fromfile(out, dtype=np.float64)
    pos = 0

    def take(shape):
        nonlocal pos
        n = int(np.prod(shape))
        a = v[pos:pos + n].reshape(shape)
        pos += n
        return a

    g = golden("sod_1d.npz")
    assert same(take((400, 5)), g["order2_output"])
    assert same(take((400, 5)), g["order1_output"])

    q = [(1.2, 0.3, -0.4, 0.0, 1.5), (0.4, -0.2, 0.1, 0.0, 0.6), (0.9, 0.5, 0.5, 0.0, 0.3), (1.7, -0.6, -0.1, 0.0, 1.1)]
    W = np.empty((20, 24, 5))
    for i in range(20):
        for j in range(24):
            W[i, j] = q[2 * (i >= 10) + (j >= 12)]
    exp, _, _ = orc.run(W, 0, 2, 1 / 24, 0.9, tf=0.02)
    assert same(take((20, 24, 5)), exp)
    dt = orc.stable_dt(W, 0, 1 / 24, 0.9, 0.0, 1.0, 0)
    e = orc.sweep(W, 0, 2, 0, 1 / 24, 0.5 * dt)
    e = orc.sweep(e, 0, 2, 1, 1 / 24, dt)
    e = orc.sweep(e, 0, 2, 0, 1 / 24, 0.5 * dt)
    assert same(take((20, 24, 5)), e)

    S = np.empty((200, 5))
    S[:100] = (10.0, 0.0, 0.0, 0.0, 13.33)
    S[100:] = (1.0, 0.0, 0.0, 0.0, 1e-6)
    exp, _, _ = orc.run(S, 1, 2, 1 / 200, 0.5, tf=0.05, gamma=5 / 3)
    assert same(take((200, 5)), exp)

    ws, dt2 = take((2,))
    assert ws == orc.max_wavespeed(W, 0) and dt2 == dt

    lim = take((3, 41))
    for k in range(3):
        assert same(lim[k], [orc.limiter(-1.0 + 0.1 * i, 0.0, k) for i in range(41)])

    g = 5.0 / 3.0
    W2, h2, hole2 = workloads.gr_accretion_2d(16, spin=True)
    exp, _, _ = orc.gr_run(W2, 2, h2, 0.95, hole2, tf=0.01, subcycles=1, gamma=g)
    assert same(take((16, 16, 5)), exp)
    W1, h1, hole1 = workloads.gr_accretion_1d(60)
    exp, _, _ = orc.gr_run(W1, 1, h1, 0.95, hole1, tf=0.02, subcycles=1, gamma=g)
    assert same(take((60, 5)), exp)
    e = orc.gr_forcing(W2, h2, 0.001, hole2, 0.0, 1, g)
    e = orc.gr_sweep(e, 2, 0, h2, 0.001, hole2, 0.0, 1, g)
    e = orc.gr_sweep(e, 2, 1, h2, 0.002, hole2, 0.0, 1, g)
    e = orc.gr_sweep(e, 2, 0, h2, 0.001, hole2, 0.0, 1, g)
    assert same(take((16, 16, 5)), e)
    assert take((1,))[0] == orc.gr_max_wavespeed(W2, hole2, g)

    # AMR drop-in classes
    S = np.zeros((40, 5))
    S[:20] = (1.0, 0.0, 0.0, 0.0, 1.0)
    S[20:] = (0.125, 0.0, 0.0, 0.0, 0.1)
    c, fn, cf, ff, _, _ = orc.amr_level1(S, 0, 2, 1 / 40, 0.8, 0.02, 5.0, 1.4)
    assert same(take((40, 5)), c) and same(take((80, 5)), fn)
    assert same(take((40,)), cf.astype(float)) and same(take((80,)), ff.astype(float))
    assert cf.sum() < 40  # the shock was refined
    Q = np.empty((8, 10, 5))
    for i in range(8):
        for j in range(10):
            Q[i, j] = q[2 * (i >= 4) + (j >= 5)]
    exp = orc.amr_level2(Q, 1, 2, 1 / 10, 0.8, 0.02, 3.0, 5 / 3)
    for k, shape in enumerate(((8, 10, 5), (16, 20, 5), (32, 40, 5))):
        assert same(take(shape), exp[k]), ("level", k)
    for k, shape in enumerate(((8, 10), (16, 20), (32, 40))):
        assert same(take(shape), exp[3 + k].astype(float)), ("flags", k)
    assert exp[5].sum() > 0
    mask = (np.add.outer(np.arange(20), np.arange(24)) % 2) == 0
    e = orc.sweep(W, 0, 2, 0, 1 / 24, 0.001)
    assert same(take((20, 24, 5)), np.where(mask[..., None], e, W))

    # black-hole AMR drop-in class
    from cosmos_amd import scenarios

    S = scenarios.bh_spherical_1d(40)
    c, fn, cf, ff, _, _ = orc.gr_amr_level1(S.W, 2, 1 / 40, 0.95, 0.05, 0.5, S.hole, g)
    assert same(take((40, 5)), c) and same(take((80, 5)), fn)
    assert same(take((40,)), cf.astype(float)) and same(take((80,)), ff.astype(float))
    S = scenarios.bh_spinning_spheroidal_2d(12)
    c, fn, cf, ff, _, _ = orc.gr_amr_level1(S.W, 1, 1 / 12, 0.95, 0.02, 0.5, S.hole, g)
    assert same(take((12, 12, 5)), c) and same(take((24, 24, 5)), fn)
    assert same(take((12, 12)), cf.astype(float)) and same(take((24, 24)), ff.astype(float))
    S = scenarios.bh_spinning_1d(32)
    exp = orc.gr_amr_level2(S.W, 1, 1 / 32, 0.95, 0.03, 0.5, S.hole, g)
    for k, n in enumerate((32, 64, 128)):
        assert same(take((n, 5)), exp[k]), ("level", k)
    for k, n in enumerate((32, 64, 128)):
        assert same(take((n,)), exp[3 + k].astype(float)), ("flags", k)
    S = scenarios.bh_spinning_spheroidal_2d(12)
    exp = orc.gr_amr_level2(S.W, 1, 1 / 12, 0.95, 0.02, 0.5, S.hole, g)
    for k, n in enumerate((12, 24, 48)):
        assert same(take((n, n, 5)), exp[k]), ("2D level", k)
    for k, n in enumerate((12, 24, 48)):
        assert same(take((n, n)), exp[3 + k].astype(float)), ("2D flags", k)
    # per-sweep calls through the device-resident copy, with an edit by the caller between them
    e = orc.sweep(W, 0, 2, 0, 1 / 24, 0.001)
    e[5, 7, 0] = 1.5 * e[5, 7, 0]
    e = orc.sweep(e, 0, 2, 0, 1 / 24, 0.001)
    e = orc.sweep(e, 0, 2, 1, 1 / 24, 0.001)
    e = orc.sweep(e, 0, 1, 0, 1 / 24, 0.001)
    assert same(take((20, 24, 5)), e)
    # the point-level statics (sections 20-22), after the device work of the same process
    from test_point import bits, expected_points
    assert bits(take((1005,)), expected_points(orc))
    assert pos == v.size
