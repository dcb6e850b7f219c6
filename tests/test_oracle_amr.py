"""The two-level AMR part of the CPU oracle (oracle/cosmos_oracle_amr.c)
pinned three ways:
  1. bitwise against the reference's own solve{,2D}Level1AMR (oracle/_ref):
     coarse and fine states and both flag grids, Newtonian and special
     relativistic, 1D and 2D, FORCE and SLIC;
  2. against the committed golden fixtures written from that build;
  3. against an independent brute-force restatement of the flag rule
     (numpy), the stickiness of the flags, a uniform state (no flags), and
     the special-relativistic quirks: the diagonal "bottom" neighbour and the
     default-state corner of the boundary-padded grid.
"""
import numpy as np
import pytest

from conftest import golden, same
from cosmos_amd import workloads

G_N, G_SR = 1.4, 5.0 / 3.0


def cases():
    sp = workloads.quadrants_2d(10, seed=6, ny=8)
    sp[..., 1:3] *= 0.5
    sp[3, 4] = (1.9, 0.2, -0.3, 0.0, 1.7)
    out = []
    for physics, g in ((0, G_N), (1, G_SR)):
        for order in (1, 2):
            W1 = workloads.sod_1d(40) if physics == 0 else workloads.sr_riemann_1d(40, seed=3)
            W2 = workloads.quadrants_2d(12, seed=1, ny=10)
            if physics == 1:
                W2[..., 1:3] *= 0.5
            tol = 5.0 if physics == 0 else 20.0
            out.append((physics, order, W1, tol, g))
            out.append((physics, order, W2, tol, g))
    out.append((1, 2, sp, 3.0, G_SR))
    out.append((0, 2, sp, 3.0, G_N))
    return out


@pytest.mark.parametrize("case", range(10))
def test_amr_level1_matches_reference(ref, orc, case):
    physics, order, W, tol, g = cases()[case]
    nx = W.shape[-2]
    a = orc.amr_level1(W, physics, order, 1.0 / nx, 0.8, 0.02, tol, g)
    b = ref.amr_level1(W, physics, order, 1.0 / nx, 0.8, 0.02, tol, g)
    for x, y in zip(a[:4], b):
        assert same(x, y)
    assert a[4] == 0.02


def test_amr_golden(orc):
    g = golden("amr.npz")
    tags = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_meta")})
    assert len(tags) == 5
    for tag in tags:
        physics, order, tol, gamma = g[f"{tag}_meta"]
        W = g[f"{tag}_input"]
        c, f, cf, ff, t, n = orc.amr_level1(W, int(physics), int(order), 1.0 / W.shape[-2], 0.8, 0.02, tol, gamma)
        assert same(c, g[f"{tag}_coarse"]) and same(f, g[f"{tag}_fine"]), tag
        assert same(cf, g[f"{tag}_cflags"]) and same(ff, g[f"{tag}_fflags"]), tag


def brute_flags(U, dx, tol, diag, corner):
    """Independent restatement of the coarse refinement test on conserved U
    (ny, nx, 5): padded with a copy boundary, corners set to `corner`."""
    ny, nx, _ = U.shape
    P = np.empty((ny + 2, nx + 2, 5))
    P[1:-1, 1:-1] = U
    P[0, 1:-1], P[-1, 1:-1] = U[0], U[-1]
    P[1:-1, 0], P[1:-1, -1] = U[:, 0], U[:, -1]
    for r, c in ((0, 0), (0, -1), (-1, 0), (-1, -1)):
        P[r, c] = corner
    hit = np.zeros((ny, nx), dtype=bool)
    for i in range(ny):
        for j in range(nx):
            m = P[i + 1, j + 1]
            bottom = P[i + 2, j + 2] if diag else P[i + 2, j + 1]
            for nb, sgn in ((P[i, j + 1], 1), (bottom, -1), (P[i + 1, j], 1), (P[i + 1, j + 2], -1)):
                d = (m - nb) if sgn > 0 else (nb - m)
                if np.any(np.abs(d / dx) >= tol):
                    hit[i, j] = True
    return hit


@pytest.mark.parametrize("physics,gamma", [(0, G_N), (1, G_SR)])
def test_refine_flags_brute_force_and_sticky(orc, physics, gamma):
    rng = np.random.default_rng(4)
    W = np.empty((7, 9, 5))
    W[..., 0] = 1.0
    W[..., 1:4] = 0.0
    W[..., 4] = 1.0
    for (i, j) in ((2, 3), (5, 7), (6, 8), (0, 0)):
        W[i, j] = (rng.uniform(1.5, 2.0), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 0.0, rng.uniform(1.5, 2.0))
    U = np.array([orc.p2c(physics, w, gamma) for w in W.reshape(-1, 5)]).reshape(W.shape)
    corner = orc.p2c(physics, np.array([1.0, 0.0, 0.0, 0.0, 1.0]), gamma)
    dx, tol = 0.1, 2.0
    cf = np.ones((7, 9), dtype=np.uint8)
    ff = np.zeros((14, 18), dtype=np.uint8)
    cf1, ff1 = orc.amr_refine_flags(W, physics, dx, tol, cf, ff, gamma)
    hit = brute_flags(U, dx, tol, diag=(physics == 1), corner=corner)
    assert np.array_equal(cf1 == 0, hit)
    assert np.array_equal(ff1, np.kron(hit, np.ones((2, 2), dtype=bool)).astype(np.uint8))
    # sticky: a uniform state afterwards clears nothing
    Wu = np.tile(np.array([1.0, 0.0, 0.0, 0.0, 1.0]), (7, 9, 1))
    cf2, ff2 = orc.amr_refine_flags(Wu, physics, dx, tol, cf1, ff1, gamma)
    assert np.array_equal(cf2, cf1) and np.array_equal(ff2, ff1)


def test_special_relativistic_quirks(orc):
    """A uniform state differing from the default (1, 0, 0, 0, 1): Newtonian
    flags nothing; the special-relativistic test flags exactly the
    bottom-right coarse cell, whose diagonal "bottom" is the padded grid's
    default-constructed corner."""
    W = np.tile(np.array([2.0, 0.1, 0.0, 0.0, 3.0]), (6, 5, 1))
    cf = np.ones((6, 5), dtype=np.uint8)
    ff = np.zeros((12, 10), dtype=np.uint8)
    cfn, ffn = orc.amr_refine_flags(W, 0, 0.1, 1.0, cf, ff, G_N)
    assert cfn.all() and not ffn.any()
    cfs, ffs = orc.amr_refine_flags(W, 1, 0.1, 1.0, cf, ff, G_SR)
    expect = np.ones((6, 5), dtype=np.uint8)
    expect[5, 4] = 0
    assert np.array_equal(cfs, expect)
    assert ffs.sum() == 4 and ffs[10:, 8:].all()


def test_uniform_state_never_refines(orc):
    W = np.tile(np.array([1.3, 0.2, -0.1, 0.0, 0.7]), (8, 8, 1))
    c, f, cf, ff, t, n = orc.amr_level1(W, 0, 2, 1.0 / 8, 0.8, 0.05, 0.5, G_N)
    assert cf.all() and not ff.any()
    assert same(c, W) and same(f, np.kron(W, np.ones((2, 2, 1))))


# ---- three-level AMR (EulerAMR::solve{,2D}Level2AMR and the special-relativistic copies) ----

def level2_cases():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import amr_level2_cases

    return amr_level2_cases()


LEVEL2_NAMES = ("coarse", "inter", "fine", "cflags", "iflags", "fflags")


@pytest.mark.parametrize("case", range(7))
def test_level2_oracle_matches_reference(orc, ref, case):
    tag, physics, order, W, tol, g, tf = level2_cases()[case]
    n = W.shape[-2]
    exp = ref.amr_level2(W, physics, order, 1.0 / n, 0.8, tf, tol, g)
    got = orc.amr_level2(W, physics, order, 1.0 / n, 0.8, tf, tol, g)
    for k, name in enumerate(LEVEL2_NAMES):
        assert same(got[k], exp[k]), (tag, name)
    assert got[5].sum() > 0 and got[4].sum() > 0  # both finer levels were used


def test_level2_oracle_matches_golden(orc):
    g = golden("amr_level2.npz")
    for tag, *_ in level2_cases():
        physics, order, tol, gamma, tf = g[f"{tag}_meta"]
        W = g[f"{tag}_input"]
        got = orc.amr_level2(W, int(physics), int(order), 1.0 / W.shape[-2], 0.8, tf, tol, gamma)
        for k, name in enumerate(LEVEL2_NAMES):
            assert same(got[k], g[f"{tag}_{name}"]), (tag, name)


@pytest.mark.parametrize("physics,g", [(0, G_N), (1, G_SR)])
def test_level2_without_refinement_is_the_plain_solve(orc, physics, g):
    """A tolerance no jump reaches: no flag is ever set below the coarse level,
    every coarse step advances the step counter once, so the coarse level is the
    plain solve loop, the intermediate level its copy (injected after every
    coarse step) and the fine level the initial state (it is only ever written
    in intermediate sub-steps)."""
    W = workloads.sod_1d(50) if physics == 0 else workloads.sr_riemann_1d(50, seed=4)
    c, m, f, cf, mf, ff, t, k = orc.amr_level2(W, physics, 2, 1.0 / 50, 0.8, 0.05, 1e30, g)
    exp, te, ke = orc.run(W, physics, 2, 1.0 / 50, 0.8, tf=0.05, gamma=g)
    assert same(c, exp) and (t, k) == (te, ke)
    assert cf.all() and not mf.any() and not ff.any()
    assert same(m, np.repeat(c, 2, axis=0)) and same(f, np.repeat(W, 4, axis=0))


def test_level2_uniform_state_never_refines(orc):
    W = np.tile([0.7, 0.1, 0.0, 0.0, 0.9], (30, 1))
    c, m, f, cf, mf, ff, t, k = orc.amr_level2(W, 0, 2, 1.0 / 30, 0.8, 0.1, 1e-9, G_N)
    assert k > 3 and cf.all() and not mf.any() and not ff.any()
