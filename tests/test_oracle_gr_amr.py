"""The black-hole AMR part of the CPU oracle (orc_gr_amr_level1 / level2 in
oracle/cosmos_oracle_gr.c) pinned three ways:
  1. bitwise against the reference's own BlackHoleEulerAMR::solveLevel1AMR,
     solve2DLevel1AMR and 1D solveLevel2AMR (oracle/_ref, compiled from the
     complete functions of AMR/BlackHoleEulerAMR.cpp): every level and flag
     grid, FORCE and SLIC, static and spinning holes (NaN states inside the
     horizon at the same places);
  2. against the committed golden fixtures written from that build;
  3. against what the loops fix without the reference: a tolerance no jump
     reaches is the plain solve without forcing (both drivers, every level
     but the coarse a copy); a zero tolerance (hole off the line, so no
     conversion is singular) refines every cell, so after one coarse step
     the fine level is the plain solve of the doubled grid up to the first
     step that comes within 1e-4 of the coarse time, and the coarse level the
     restriction of its children, each converted at its own coordinate
     (p2c / c2p of the single-point oracle calls).
"""
import os
import sys

import numpy as np
import pytest

from conftest import golden, same
from oracle import GR_C2P, GR_P2C

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import GR_AMR_NAMES, gr_amr_cases  # noqa: E402


def run(o, levels, order, S, tol, tf, **kw):
    f = o.gr_amr_level1 if levels == 1 else o.gr_amr_level2
    return f(S.W, order, S.cell_spacing, S.cfl, tf, tol, S.hole, S.gamma, **kw)


@pytest.mark.parametrize("case", range(len(gr_amr_cases())))
def test_gr_amr_oracle_matches_reference(orc, ref, case):
    tag, levels, order, S, tol, tf = gr_amr_cases()[case]
    exp = run(ref, levels, order, S, tol, tf)
    got = run(orc, levels, order, S, tol, tf)
    for k, name in enumerate(GR_AMR_NAMES[levels]):
        assert same(got[k], exp[k]), (tag, name)
    assert got[len(GR_AMR_NAMES[levels]) - 1].sum() > 0  # the finest level was used
    assert got[-2] == pytest.approx(tf) and got[-1] >= 2, tag


def test_gr_amr_oracle_matches_golden(orc):
    g = golden("gr_amr.npz")
    cases = gr_amr_cases()
    for tag, levels, *_ in cases:
        lv, order, tol, tf, cfl, gamma, *hole = g[f"{tag}_meta"]
        W = g[f"{tag}_input"]
        f = orc.gr_amr_level1 if int(lv) == 1 else orc.gr_amr_level2
        got = f(W, int(order), 1.0 / W.shape[-2], cfl, tf, tol, tuple(hole), gamma)
        for k, name in enumerate(GR_AMR_NAMES[int(lv)]):
            assert same(got[k], g[f"{tag}_{name}"]), (tag, name)
    assert any(np.isnan(g[f"{t}_coarse"]).any() for t, *_ in cases)  # spinning holes: NaN inside the horizon


def line(n=40):
    from cosmos_amd import scenarios

    return scenarios.bh_spherical_1d(n)


@pytest.mark.parametrize("levels", [1, 2])
def test_gr_amr_without_refinement_is_the_plain_solve(orc, levels):
    S = line()
    dx, G = S.cell_spacing, S.gamma
    out = run(orc, levels, 2, S, 1e30, 0.05)
    exp, te, ke = orc.gr_run(S.W, 2, dx, S.cfl, S.hole, tf=0.05, subcycles=0, gamma=G)
    assert same(out[0], exp) and (out[-2], out[-1]) == (te, ke)
    if levels == 1:
        c, f, cf, ff = out[:4]
        assert cf.all() and not ff.any() and same(f, np.repeat(c, 2, axis=0))
    else:
        c, m, f, cf, mf, ff = out[:6]
        assert cf.all() and not mf.any() and not ff.any()
        assert same(m, np.repeat(c, 2, axis=0)) and same(f, np.repeat(S.W, 4, axis=0))


def test_gr_amr_full_refinement_one_step(orc):
    S = line()
    hole = (0.05, 0.3, -0.3, 0.0, 0.0)  # off the line: every conversion is regular
    dx, G, n = S.cell_spacing, S.gamma, S.W.shape[0]
    c, f, cf, ff, t, k = orc.gr_amr_level1(S.W, 2, dx, S.cfl, 0.05, 0.0, hole, G, max_steps=1)
    assert k == 1 and not cf.any() and ff.all()
    fine0 = np.repeat(S.W, 2, axis=0)
    for cap in range(1, 50):
        e, te, ke = orc.gr_run(fine0, 2, dx * 0.5, S.cfl, hole, tf=t, max_steps=cap, subcycles=0, gamma=G)
        if te >= t - 1e-4:
            break
    assert ke >= 2 and same(f, e)
    R = np.empty_like(c)
    for i in range(n):
        a = orc.gr_point(GR_P2C, f[2 * i], (i * 2) * (dx * 0.5), 0.0, 0.0, 0.0, hole, G)
        b = orc.gr_point(GR_P2C, f[2 * i + 1], ((i * 2) + 1) * (dx * 0.5), 0.0, 0.0, 0.0, hole, G)
        R[i] = orc.gr_point(GR_C2P, 0.5 * (a + b), i * dx, 0.0, 0.0, 0.0, hole, G)
    assert same(R, c)
    # the same restriction with the parent converted one cell over differs: the coordinate matters
    a = orc.gr_point(GR_P2C, f[20], 20 * (dx * 0.5), 0.0, 0.0, 0.0, hole, G)
    b = orc.gr_point(GR_P2C, f[21], 21 * (dx * 0.5), 0.0, 0.0, 0.0, hole, G)
    assert not same(orc.gr_point(GR_C2P, 0.5 * (a + b), 11 * dx, 0.0, 0.0, 0.0, hole, G), c[10])


def test_gr_amr_level2_2d_without_refinement_is_the_plain_solve(orc):
    """The 2D three-level driver (completed past the reference's truncation):
    with no flag below the coarse level it is the plain 2D solve without
    forcing, the intermediate level its copy and the fine level the initial
    state."""
    from cosmos_amd import scenarios

    S = scenarios.bh_spherical_2d(12)
    c, m, f, cf, mf, ff, t, k = orc.gr_amr_level2(S.W, 2, S.cell_spacing, S.cfl, 0.05, 1e30, S.hole, S.gamma)
    exp, te, ke = orc.gr_run(S.W, 2, S.cell_spacing, S.cfl, S.hole, tf=0.05, subcycles=0, gamma=S.gamma)
    assert same(c, exp) and (t, k) == (te, ke)
    assert cf.all() and not mf.any() and not ff.any()
    assert same(m, np.kron(c, np.ones((2, 2, 1)))) and same(f, np.kron(S.W, np.ones((4, 4, 1))))


def test_gr_amr_level2_2d_uses_every_level(orc):
    from cosmos_amd import scenarios

    S = scenarios.bh_spherical_2d(12)
    c, m, f, cf, mf, ff, t, k = orc.gr_amr_level2(S.W, 2, S.cell_spacing, S.cfl, 0.03, 0.5, S.hole, S.gamma)
    assert k >= 2 and not cf.all() and mf.any() and ff.any()
    assert t == pytest.approx(0.03) and np.isfinite(f).all()
