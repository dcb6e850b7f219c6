"""GPU parity of the two-level AMR path (cosmos_amr_*): refinement flags,
masked sweeps and the level-1 driver against the CPU oracle, bitwise for
states and flags."""
import numpy as np
import pytest

from conftest import conserved, golden, rel_linf, same
from cosmos_amd import Amr, workloads
from test_oracle_amr import cases

pytestmark = pytest.mark.gpu

G_N, G_SR = 1.4, 5.0 / 3.0


def amr_for(W, physics, gamma):
    dim = 1 if W.ndim == 2 else 2
    nx = W.shape[-2]
    ny = W.shape[0] if dim == 2 else 1
    a = Amr(physics, dim, nx, ny, gamma)
    a.upload(W)
    return a


@pytest.mark.parametrize("case", range(10))
def test_amr_level1_bitwise(orc, case):
    physics, order, W, tol, g = cases()[case]
    nx = W.shape[-2]
    exp = orc.amr_level1(W, physics, order, 1.0 / nx, 0.8, 0.02, tol, g)
    with amr_for(W, physics, g) as a:
        t, n = a.solve_level1(order, 1.0 / nx, 0.8, 0.02, tol)
        got = a.download()
    for k, name in enumerate(("coarse", "fine", "coarse flags", "fine flags")):
        assert same(got[k], exp[k]), (name, rel_linf(got[k], exp[k]) if k < 2 else None)
    assert (t, n) == (exp[4], exp[5])


def test_amr_matches_golden():
    g = golden("amr.npz")
    for tag in sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_meta")}):
        physics, order, tol, gamma = g[f"{tag}_meta"]
        W = g[f"{tag}_input"]
        with amr_for(W, int(physics), gamma) as a:
            a.solve_level1(int(order), 1.0 / W.shape[-2], 0.8, 0.02, tol)
            c, f, cf, ff = a.download()
        assert same(c, g[f"{tag}_coarse"]) and same(f, g[f"{tag}_fine"]), tag
        assert same(cf, g[f"{tag}_cflags"]) and same(ff, g[f"{tag}_fflags"]), tag


@pytest.mark.parametrize("physics,gamma", [(0, G_N), (1, G_SR)])
@pytest.mark.parametrize("shape", [(37,), (23, 30), (64, 48)])
def test_amr_flags_and_masked_sweep(orc, physics, gamma, shape):
    rng = np.random.default_rng(len(shape) * 7 + physics)
    W = np.empty(tuple(shape) + (5,))
    W[..., 0] = rng.uniform(0.5, 2.0, shape)
    W[..., 1:3] = rng.uniform(-0.4, 0.4, tuple(shape) + (2,))
    W[..., 3] = 0.0
    W[..., 4] = rng.uniform(0.5, 2.0, shape)
    nx = shape[-1]
    dx, tol = 1.0 / nx, 60.0
    cshape = shape
    fshape = (2 * shape[0],) if len(shape) == 1 else (2 * shape[0], 2 * shape[1])
    cf0 = np.ones(cshape, dtype=np.uint8)
    ff0 = np.zeros(fshape, dtype=np.uint8)
    ecf, eff = orc.amr_refine_flags(W, physics, dx, tol, cf0, ff0, gamma)
    assert 0 < (ecf == 0).sum() < ecf.size
    with amr_for(W, physics, gamma) as a:
        a.refine_flags(dx, tol)
        c, f, cf, ff = a.download()
        assert same(cf, ecf) and same(ff, eff)
        for axis in ((0,) if len(shape) == 1 else (0, 1)):
            a.sweep(0, 2, axis, dx, 0.1 * dx)
        c2, _, _, _ = a.download()
    exp = W
    for axis in ((0,) if len(shape) == 1 else (0, 1)):
        full = orc.sweep(exp, physics, 2, axis, dx, 0.1 * dx, 0.0, 1, gamma)
        exp = np.where(ecf[..., None] == 1, full, exp)
    assert same(c2, exp), rel_linf(c2, exp)


@pytest.mark.parametrize("physics,gamma", [(0, G_N), (1, G_SR)])
@pytest.mark.parametrize("shape,dx", [((300,), 1.0 / 256), ((300,), 1.0 / 300), ((9, 70), 1.0 / 64),
                                      ((9, 70), 0.013), ((130, 67), 2.0 ** -7), ((130, 67), 1.0 / 67)])
def test_refine_flags_tiles_and_exact_reciprocal(orc, physics, gamma, shape, dx):
    """The tiled flag kernel (partial tiles at the right and bottom edges, the
    copy boundary from the tile ring, the SR diagonal "bottom" and corner) and
    its multiply-by-1/dx form for power-of-two spacings against the oracle's
    divide; a second pass over already-set flags changes nothing (sticky)."""
    rng = np.random.default_rng(sum(shape) + physics)
    W = np.empty(tuple(shape) + (5,))
    W[..., 0] = rng.uniform(0.5, 2.0, shape)
    W[..., 1:3] = rng.uniform(-0.4, 0.4, tuple(shape) + (2,))
    W[..., 3] = 0.0
    W[..., 4] = rng.uniform(0.5, 2.0, shape)
    U = conserved(W, gamma, physics).reshape(W.shape)
    tol = float(np.quantile(np.abs(np.diff(U, axis=-2)).max(axis=-1), 0.8)) / dx
    fshape = (2 * shape[0],) if len(shape) == 1 else (2 * shape[0], 2 * shape[1])
    ecf, eff = orc.amr_refine_flags(W, physics, dx, tol, np.ones(shape, np.uint8), np.zeros(fshape, np.uint8),
                                    gamma)
    assert 0 < (ecf == 0).sum() < ecf.size
    with amr_for(W, physics, gamma) as a:
        a.refine_flags(dx, tol)
        _, _, cf, ff = a.download()
        a.refine_flags(dx, tol)
        _, _, cf2, ff2 = a.download()
    assert same(cf, ecf) and same(ff, eff)
    assert same(cf2, ecf) and same(ff2, eff)


def test_amr_2d_larger_newtonian(orc):
    W = workloads.quadrants_2d(48, seed=9, ny=40)
    exp = orc.amr_level1(W, 0, 2, 1.0 / 48, 0.9, 1.0, 8.0, G_N, max_steps=6)
    with amr_for(W, 0, G_N) as a:
        t, n = a.solve_level1(2, 1.0 / 48, 0.9, 1.0, 8.0, 6)
        got = a.download()
    for k in range(4):
        assert same(got[k], exp[k]), k
    assert (t, n) == (exp[4], exp[5]) and n == 6


# ---- three-level AMR ----

from test_oracle_amr import LEVEL2_NAMES, level2_cases  # noqa: E402


@pytest.mark.parametrize("physics", [0, 1])
def test_amr_drivers_on_levels_wider_than_a_segment(orc, physics):
    """Finer levels wider than one x-sweep row segment (252 cells) and taller
    than one y-sweep tile, a few percent of their cells flagged: the masked
    sweeps then launch only the listed blocks, and every fine sub-step after
    the first re-reads only the flagged cells' wave speeds.  Both drivers,
    six coarse steps with sub-steps, every level and flag bitwise the
    oracle's."""
    g = G_N if physics == 0 else G_SR
    W = workloads.quadrants_2d(140, seed=21, ny=72)
    if physics:
        W[..., 1:3] *= 0.5
    exp = orc.amr_level1(W, physics, 2, 1.0 / 140, 0.8, 1.0, 8.0, g, max_steps=6)
    assert 0.0 < (exp[3] != 0).mean() < 0.2
    with amr_for(W, physics, g) as a:
        t, n = a.solve_level1(2, 1.0 / 140, 0.8, 1.0, 8.0, 6)
        got = a.download()
    for k in range(4):
        assert same(got[k], exp[k]), k
    assert (t, n) == (exp[4], exp[5])
    W = workloads.quadrants_2d(72, seed=22, ny=40)
    if physics:
        W[..., 1:3] *= 0.5
    exp = orc.amr_level2(W, physics, 2, 1.0 / 72, 0.8, 1.0, 8.0, g, max_steps=6)
    assert 0.0 < (exp[5] != 0).mean() < 0.2
    with Amr(physics, 2, 72, 40, g) as a:
        a.upload(W)
        t, k = a.solve_level2(2, 1.0 / 72, 0.8, 1.0, 8.0, 6)
        got = a.download_level2()
    for k2, name in enumerate(LEVEL2_NAMES):
        assert same(got[k2], exp[k2]), name
    assert (t, k) == (exp[6], exp[7])


@pytest.mark.parametrize("case", range(7))
def test_amr_level2_bitwise(orc, case):
    tag, physics, order, W, tol, g, tf = level2_cases()[case]
    n = W.shape[-2]
    exp = orc.amr_level2(W, physics, order, 1.0 / n, 0.8, tf, tol, g)
    dim = 1 if W.ndim == 2 else 2
    with Amr(physics, dim, n, W.shape[0] if dim == 2 else 1, g) as a:
        a.upload(W)
        t, k = a.solve_level2(order, 1.0 / n, 0.8, tf, tol)
        got = a.download_level2()
        # a second run from the coarse state it left restarts every level from it
        a.upload(W)
        a.solve_level2(order, 1.0 / n, 0.8, tf, tol)
        again = a.download_level2()
    for k2, name in enumerate(LEVEL2_NAMES):
        assert same(got[k2], exp[k2]), (tag, name)
        assert same(again[k2], exp[k2]), (tag, name)
    assert (t, k) == (exp[6], exp[7])


def test_amr_level2_matches_golden():
    g = golden("amr_level2.npz")
    for tag, *_ in level2_cases():
        physics, order, tol, gamma, tf = g[f"{tag}_meta"]
        W = g[f"{tag}_input"]
        dim = 1 if W.ndim == 2 else 2
        with Amr(int(physics), dim, W.shape[-2], W.shape[0] if dim == 2 else 1, gamma) as a:
            a.upload(W)
            a.solve_level2(int(order), 1.0 / W.shape[-2], 0.8, tf, tol)
            got = a.download_level2()
        for k, name in enumerate(LEVEL2_NAMES):
            assert same(got[k], g[f"{tag}_{name}"]), (tag, name)


# ---- black-hole AMR (BlackHoleEulerAMR::solve{,2D}Level1AMR, 1D solveLevel2AMR) ----

from cosmos_amd import BLACKHOLE  # noqa: E402
from test_oracle_gr_amr import GR_AMR_NAMES, gr_amr_cases  # noqa: E402


def gr_amr_gpu(W, levels, order, dx, cfl, tf, tol, hole, gamma, excision=0.0):
    dim = 1 if W.ndim == 2 else 2
    with Amr(BLACKHOLE, dim, W.shape[-2], W.shape[0] if dim == 2 else 1, gamma) as a:
        a.set_spacetime(*hole)
        if excision:
            a.set_excision(excision)
        a.upload(W)
        if levels == 1:
            t, k = a.solve_level1(order, dx, cfl, tf, tol)
            return a.download(), t, k
        t, k = a.solve_level2(order, dx, cfl, tf, tol)
        return a.download_level2(), t, k


@pytest.mark.parametrize("case", range(len(gr_amr_cases())))
def test_gr_amr_bitwise(orc, case):
    tag, levels, order, S, tol, tf = gr_amr_cases()[case]
    f = orc.gr_amr_level1 if levels == 1 else orc.gr_amr_level2
    exp = f(S.W, order, S.cell_spacing, S.cfl, tf, tol, S.hole, S.gamma)
    print(f"black-hole AMR {tag}: levels {levels} order {order} shape {S.W.shape}", flush=True)
    got, t, k = gr_amr_gpu(S.W, levels, order, S.cell_spacing, S.cfl, tf, tol, S.hole, S.gamma)
    for i, name in enumerate(GR_AMR_NAMES[levels]):
        assert same(got[i], exp[i]), (tag, name)
    assert (t, k) == (exp[-2], exp[-1])


def test_gr_amr_matches_golden():
    g = golden("gr_amr.npz")
    for tag, *_ in gr_amr_cases():
        lv, order, tol, tf, cfl, gamma, *hole = g[f"{tag}_meta"]
        W = g[f"{tag}_input"]
        got, _, _ = gr_amr_gpu(W, int(lv), int(order), 1.0 / W.shape[-2], cfl, tf, tol, tuple(hole), gamma)
        for i, name in enumerate(GR_AMR_NAMES[int(lv)]):
            assert same(got[i], g[f"{tag}_{name}"]), (tag, name)


@pytest.mark.parametrize("levels", [1, 2])
def test_gr_amr_excision_radius(orc, levels):
    """A set excision radius reaches every level's sweeps (the oracle's global
    radius, reset afterwards)."""
    from cosmos_amd import scenarios

    S = scenarios.bh_spherical_1d(48)
    f = orc.gr_amr_level1 if levels == 1 else orc.gr_amr_level2
    try:
        orc.gr_set_excision(0.12)
        exp = f(S.W, 2, S.cell_spacing, S.cfl, 0.05, 0.3, S.hole, S.gamma)
    finally:
        orc.gr_set_excision(0.0)
    plain = f(S.W, 2, S.cell_spacing, S.cfl, 0.05, 0.3, S.hole, S.gamma)
    got, t, k = gr_amr_gpu(S.W, levels, 2, S.cell_spacing, S.cfl, 0.05, 0.3, S.hole, S.gamma, excision=0.12)
    for i, name in enumerate(GR_AMR_NAMES[levels]):
        assert same(got[i], exp[i]), name
    assert not same(plain[0], exp[0])  # the radius changed the run


@pytest.mark.parametrize("order,spin", [(2, False), (1, True)])
def test_gr_amr_level2_2d_completion(orc, order, spin):
    """The 2D three-level black-hole driver (the reference's is truncated;
    both sides complete it the same way) against the oracle."""
    from cosmos_amd import scenarios

    S = (scenarios.bh_spinning_spheroidal_2d if spin else scenarios.bh_spherical_2d)(12)
    exp = orc.gr_amr_level2(S.W, order, S.cell_spacing, S.cfl, 0.03, 0.5, S.hole, S.gamma)
    print(f"black-hole 2D level 2: order {order} spin {spin}", flush=True)
    got, t, k = gr_amr_gpu(S.W, 2, order, S.cell_spacing, S.cfl, 0.03, 0.5, S.hole, S.gamma)
    for i, name in enumerate(GR_AMR_NAMES[2]):
        assert same(got[i], exp[i]), name
    assert (t, k) == (exp[-2], exp[-1]) and exp[5].any()


def test_gr_amr_argument_errors():
    with Amr(BLACKHOLE, 1, 16, 1, 5.0 / 3.0) as a:
        with pytest.raises(Exception, match="cell spacing"):
            a.restrict()
    with Amr(0, 1, 16, 1, 1.4) as a:
        with pytest.raises(Exception, match="COSMOS_BLACKHOLE"):
            a.set_spacetime(1.0, 0.0, 0.5, 0.0)


def test_amr_profile_counts_each_pass():
    """cosmos_amr_profile_*: one timed launch per pass kind, positive times,
    reset by the read; the level solver handles expose their streams."""
    W = workloads.quadrants_2d(40, seed=3, ny=24)
    with amr_for(W, 0, G_N) as a:
        assert a.level_stream(0) and a.level_stream(1)
        with pytest.raises(Exception):
            a.level_stream(2)  # no three-level run yet
        a.profile(True)
        a.refine_flags(1.0 / 40, 5.0)
        a.sweep(0, 2, 0, 1.0 / 40, 1e-4)
        a.restrict()
        a.derefine(1.0 / 40, 5.0)
        got = a.profile_read()
        assert {k: n for k, (n, _) in got.items()} == {"refine": 1, "exchange": 1, "derefine": 1, "merge": 1}
        assert all(ms > 0.0 for _, ms in got.values())
        a.profile(False)
        a.refine_flags(1.0 / 40, 5.0)
        assert all(n == 0 for n, _ in a.profile_read().values())
