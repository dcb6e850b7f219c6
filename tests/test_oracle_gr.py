"""The general-relativistic part of the CPU oracle (oracle/cosmos_oracle_gr.c)
pinned three ways:
  1. bitwise against the reference's own sources (oracle/_ref) on the same
     inputs: point functions, sweeps, forcing, wave speed and solve loops;
  2. bitwise against the committed golden fixtures written from that build;
  3. against what the mathematics fixes: alpha sqrt(det gamma) = 1 for any
     Kerr-Schild metric, gamma^{ij} gamma_jk = delta, g^{mu nu} g_{nu rho} =
     delta for the reference's 4-vector (with its (r + (y - y_BH)) term), the
     Schwarzschild closed forms, the excision radius, the flat-space limit
     (GR c2p and y-flux equal the special-relativistic ones bitwise, tau with
     W in p2c, zero source), and the unit-square wave-speed placement.
"""
import math

import numpy as np
import pytest

from conftest import golden, same
from cosmos_amd import workloads
from oracle import GR_C2P, GR_METRIC, GR_P2C, GR_SOUND, GR_SOURCE, GR_XFLUX, GR_YFLUX

G = 5.0 / 3.0
HOLES = [(0.05, 0.0, 0.5, 0.5, 0.0), (0.1, 0.9, 0.5, 0.4, 0.0), (0.3, 0.7, 0.5, 0.4, 0.05)]


def states(n, seed):
    rng = np.random.default_rng(seed)
    W = np.empty((n, 5))
    W[:, 0] = rng.uniform(0.2, 3.0, n)
    W[:, 1:4] = rng.uniform(-0.3, 0.3, (n, 3))
    W[:, 4] = rng.uniform(0.05, 2.0, n)
    return W


def grid(shape, seed):
    W = states(int(np.prod(shape)), seed).reshape(tuple(shape) + (5,))
    return W


# ---------------------------------------------------------------------------
# 1. oracle == reference build
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("hole", HOLES)
def test_gr_points_match_reference(ref, orc, hole):
    rng = np.random.default_rng(7)
    Ws = states(40, 8)
    for k in range(40):
        x, y, z = rng.uniform(-0.5, 1.5, 3)
        if k % 3:
            z = 0.0
        h = rng.uniform(0.002, 0.1)
        for what in range(7):
            v = Ws[k] if what in (GR_P2C, GR_SOUND, GR_METRIC) else ref.gr_point(GR_P2C, Ws[k], x, y, z, h, hole, G)
            assert same(orc.gr_point(what, v, x, y, z, h, hole, G), ref.gr_point(what, v, x, y, z, h, hole, G)), what


@pytest.mark.parametrize("hole", HOLES)
@pytest.mark.parametrize("shape", [(30,), (10, 16)])
def test_gr_sweeps_forcing_wavespeed_match_reference(ref, orc, hole, shape):
    W = grid(shape, 11)
    h = 1.0 / max(shape)
    for order in (1, 2):
        for axis in ((0,) if len(shape) == 1 else (0, 1)):
            for lim, bias in ((1, 0.0), (0, 0.3)):
                a = orc.gr_sweep(W, order, axis, h, 0.2 * h, hole, bias, lim, G)
                b = ref.gr_sweep(W, order, axis, h, 0.2 * h, hole, bias, lim, G)
                assert same(a, b), (order, axis, lim)
    assert same(orc.gr_forcing(W, h, 0.1 * h, hole, 0.0, 1, G), ref.gr_forcing(W, h, 0.1 * h, hole, 0.0, 1, G))
    assert orc.gr_max_wavespeed(W, hole, G) == ref.gr_max_wavespeed(W, hole, G)


@pytest.mark.parametrize("order", [1, 2])
def test_gr_runs_match_reference(ref, orc, order):
    W, h, hole = workloads.gr_accretion_2d(14, spin=order == 2)
    a = orc.gr_run(W, order, h, 0.95, hole, tf=0.45, max_steps=2, subcycles=1, gamma=G)
    b = ref.gr_run(W, order, h, 0.95, hole, tf=0.45, max_steps=2, subcycles=1, gamma=G)
    assert same(a[0], b[0]) and a[1:] == b[1:]
    W, h, hole = workloads.gr_accretion_1d(40, spin=order == 1)
    a = orc.gr_run(W, order, h, 0.95, hole, tf=0.45, max_steps=4, subcycles=2, gamma=G)
    b = ref.gr_run(W, order, h, 0.95, hole, tf=0.45, max_steps=4, subcycles=2, gamma=G)
    assert same(a[0], b[0]) and a[1:] == b[1:]


# ---------------------------------------------------------------------------
# 2. oracle == golden fixtures (written from the reference build)
# ---------------------------------------------------------------------------
def test_gr_golden(orc):
    g = golden("blackhole.npz")
    a, t, _ = orc.gr_run(g["accretion_2d_input"], 2, float(g["accretion_2d_h"]), 0.95, tuple(g["accretion_2d_hole"]),
                         tf=0.45, max_steps=2, subcycles=1, gamma=G)
    assert same(a, g["accretion_2d_output"]) and t == float(g["accretion_2d_time"])
    a, t, _ = orc.gr_run(g["accretion_1d_input"], 1, float(g["accretion_1d_h"]), 0.95, tuple(g["accretion_1d_hole"]),
                         tf=0.45, max_steps=5, subcycles=1, gamma=G)
    assert same(a, g["accretion_1d_output"]) and t == float(g["accretion_1d_time"])
    hole = tuple(g["point_hole"])
    for k, (w, xyz) in enumerate(zip(g["point_states"], g["point_xyz"])):
        assert same(orc.gr_point(GR_P2C, w, *xyz, 0.01, hole, G), g["point_p2c"][k])
        assert same(orc.gr_point(GR_SOUND, w, *xyz, 0.01, hole, G), g["point_sound"][k])
        assert same(orc.gr_point(GR_METRIC, w, *xyz, 0.01, hole, G), g["point_metric"][k])
        u = g["point_p2c"][k]
        for what, tag in ((GR_C2P, "c2p"), (GR_XFLUX, "xflux"), (GR_YFLUX, "yflux"), (GR_SOURCE, "source")):
            assert same(orc.gr_point(what, u, *xyz, 0.01, hole, G), g[f"point_{tag}"][k]), tag


# ---------------------------------------------------------------------------
# 3. closed forms and invariants
# ---------------------------------------------------------------------------
def metric(orc, x, y, z, hole, h=0.01):
    m = orc.gr_point(GR_METRIC, None, x, y, z, h, hole, G)
    return dict(H=m[0], l=m[1:4], alpha=m[4], beta=m[5:8], det=m[8], ginv=m[9:18].reshape(3, 3),
                g4inv=m[18:34].reshape(4, 4), K=m[34:43].reshape(3, 3), excised=m[43])


@pytest.mark.parametrize("hole", HOLES)
def test_kerr_schild_invariants(orc, hole):
    M, a, bx, by, bz = hole
    rng = np.random.default_rng(3)
    for _ in range(50):
        x, y, z = rng.uniform(-1.0, 2.0, 3)
        q = metric(orc, x, y, z, hole)
        if q["excised"]:
            continue
        H, l = q["H"], q["l"]
        # l is unit in the flat metric; alpha sqrt(det gamma) = 1
        assert abs(l @ l - 1.0) < 1e-13
        assert abs(q["alpha"] * math.sqrt(q["det"]) - 1.0) < 1e-13
        gam = np.eye(3) - 2.0 * H * np.outer(l, l)
        assert np.abs(q["ginv"] @ gam - np.eye(3)).max() < 1e-12
        assert np.allclose(q["beta"], (2.0 * H / (1.0 - 2.0 * H)) * l, rtol=1e-14, atol=1e-16)
        # the 4-metric is built from (-1, l_x, (r + dy - a M dx)/(r^2 + a^2 M^2) negated, l_z)
        dx, dy, dz = x - bx, y - by, z - bz
        ns = dx * dx + dy * dy + dz * dz
        ma = M * M * a * a
        r = math.sqrt(0.5 * ((ns - ma) + math.sqrt((ns - ma) ** 2 + 4.0 * ma * dz * dz)))
        l4 = np.array([-1.0, l[0], -((r + dy) - a * M * dx) / (r * r + ma), l[2]])
        g4 = np.diag([-1.0, 1.0, 1.0, 1.0]) - 2.0 * H * np.outer(l4, l4)
        assert np.abs(q["g4inv"] @ g4 - np.eye(4)).max() < 1e-11
        assert abs(H + M * r ** 3 / (r ** 4 + ma * dz * dz)) <= 1e-14 * abs(H)


def test_schwarzschild_closed_forms(orc):
    M = 0.2
    hole = (M, 0.0, 0.5, 0.5, 0.0)
    for x, y in ((1.3, 0.2), (0.1, 0.9), (0.8, 0.8)):
        q = metric(orc, x, y, 0.0, hole)
        r = math.hypot(x - 0.5, y - 0.5)
        assert abs(q["H"] + M / r) < 1e-14 * (M / r)
        assert np.allclose(q["l"], -np.array([x - 0.5, y - 0.5, 0.0]) / r, rtol=1e-14, atol=1e-17)
        assert abs(q["det"] - (1.0 + 2.0 * M / r)) < 1e-14
        assert abs(q["alpha"] - 1.0 / math.sqrt(1.0 + 2.0 * M / r)) < 1e-14


def test_excision_radius(orc):
    for M, a in ((0.1, 0.0), (0.1, 0.99), (0.3, 0.5)):
        hole = (M, a, 0.5, 0.5, 0.0)
        rh = M * (1.0 + math.sqrt(1.0 - a * a))
        assert metric(orc, 0.5 + rh * (1.0 - 1e-9), 0.5, 0.0, hole)["excised"] == 1.0
        assert metric(orc, 0.5, 0.5 + rh * (1.0 + 1e-9), 0.0, hole)["excised"] == 0.0


def test_flat_space_limit(orc):
    flat = (0.0, 0.0, 0.5, 0.5, 0.0)
    Ws = states(30, 21)
    for w in Ws:
        U = orc.gr_point(GR_P2C, w, 0.3, 0.7, 0.0, 0.01, flat, G)
        Usr = orc.p2c(1, w, G)
        assert same(U[:4], Usr[:4])
        # tau with W, not W^2 (the reference's p2c)
        rho, p = w[0], w[4]
        vsq = w[1] * w[1] + w[2] * w[2] + w[3] * w[3]
        lf = 1.0 / math.sqrt(1.0 - vsq)
        hh = 1.0 + (p / rho) * (G / (G - 1.0))
        assert abs(U[4] - (rho * hh * lf - p - rho * lf)) < 1e-13 * abs(rho * hh * lf)
        # con2prim and the y flux reduce to the special-relativistic ones (same tolerance)
        assert same(orc.gr_point(GR_C2P, Usr, 0.3, 0.7, 0.0, 0.01, flat, G), orc.c2p(1, Usr, G))
        assert same(orc.gr_point(GR_YFLUX, Usr, 0.3, 0.7, 0.0, 0.01, flat, G), orc.flux(1, 1, Usr, G))
        S = orc.gr_point(GR_SOURCE, Usr, 0.3, 0.7, 0.0, 0.01, flat, G)
        assert np.all(S == 0.0)


def test_wavespeed_unit_square_placement(orc):
    """The reference evaluates the 2D wave speed on the unit square (1/N per
    cell) whatever the spacing, testing excision at (j/nx, i/ny) and the
    eigenvalues at (i/ny, j/nx)."""
    hole = (0.15, 0.5, 0.2, 0.7, 0.0)
    W = grid((9, 13), 4)
    best = 0.0
    for i in range(9):
        for j in range(13):
            if metric(orc, j / 13, i / 9, 0.0, hole)["excised"]:
                continue
            ws = abs(orc.gr_point(GR_SOUND, W[i, j], i / 9, j / 13, 0.0, 0.01, hole, G)[0])
            best = max(best, ws)
    assert orc.gr_max_wavespeed(W, hole, G) == best


def test_excision_radius_override(orc):
    hole = (0.1, 0.0, 0.5, 0.5, 0.0)
    try:
        orc.gr_set_excision(0.35)
        assert metric(orc, 0.5 + 0.35 * (1.0 - 1e-9), 0.5, 0.0, hole)["excised"] == 1.0
        assert metric(orc, 0.5, 0.5 + 0.35 * (1.0 + 1e-9), 0.0, hole)["excised"] == 0.0
        # a sweep leaves every cell whose +-2 stencil meets the region as it was
        W = grid((20, 20), 9)
        h = 1.0 / 20
        out = orc.gr_sweep(W, 2, 0, h, 0.01, hole, 0.0, 1, G)

        def inside(ii, jj):  # the cell centre (jj h, ii h), formed as the sweep forms it
            dx, dy = jj * h - 0.5, ii * h - 0.5
            return math.sqrt((dx * dx + dy * dy) + 0.0) <= 0.35

        for i in range(20):
            for j in range(20):
                near = any(inside(ii, jj) for ii, jj in
                           [(i, j)] + [(i, j + d) for d in (-2, -1, 1, 2)] + [(i + d, j) for d in (-2, -1, 1, 2)])
                assert same(out[i, j], W[i, j]) == near, (i, j)
    finally:
        orc.gr_set_excision(0.0)
    assert metric(orc, 0.5 + 0.2 * (1.0 - 1e-9), 0.5, 0.0, hole)["excised"] == 1.0
    assert metric(orc, 0.5 + 0.3, 0.5, 0.0, hole)["excised"] == 0.0


def _sr_cons(W, sd, g):
    """Flat special-relativistic p2c scaled by sqrt(det gamma): the conserved
    state whose con2prim root is W (the GR con2prim inverts this map, with the
    flat |S|^2; BlackHoleEuler/BlackHoleEulerStateVector.cpp:30-103)."""
    rho, v, p = W[0], W[1:4], W[4]
    lf = 1.0 / math.sqrt(1.0 - float(np.dot(v, v)))
    h = 1.0 + (p / rho) * (g / (g - 1.0))
    return sd * np.array([rho * lf, *(rho * h * lf * lf * v), rho * h * lf * lf - p - rho * lf])


def test_gr_tight_con2prim_reaches_the_root(orc):
    """With the tolerance control set (1e-13, the gradient-check mode), c2p
    returns primitives whose conserved state reproduces the input to rounding;
    at the reference's tolerance the kept Newton iterate leaves residuals
    orders of magnitude larger (measured: 1.5e-13 against 1.9e-8).  A control
    that is not applied fails the first bound.  The other call sites (fluxes,
    source) are compared bitwise with the GPU's tight mode in the GPU tests."""
    hole = (1.0, 0.0, 20.0, 20.0, 0.0)
    rng = np.random.default_rng(21)
    Ws = states(200, 22)
    worst = {}
    for tol in (1e-13, 0.0):
        orc.gr_set_con2prim(tol, 200 if tol else 1000)
        try:
            res = 0.0
            for W in Ws:
                x, y = rng.uniform(0.0, 40.0, 2)
                if math.hypot(x - 20.0, y - 20.0) < 3.0:
                    continue
                sd = math.sqrt(orc.gr_point(GR_METRIC, None, x, y, 0.0, 0.04, hole, G)[8])
                U = _sr_cons(W, sd, G)
                R = orc.gr_point(GR_C2P, U, x, y, 0.0, 0.04, hole, G)
                res = max(res, float(np.max(np.abs(_sr_cons(R, sd, G) - U) / np.max(np.abs(U)))))
            worst[tol] = res
        finally:
            orc.gr_set_con2prim(0.0, 1000)
    assert worst[1e-13] <= 1e-12, worst
    assert worst[0.0] > 1e3 * worst[1e-13], worst


def test_gr_nan_onset_matches_reference_fixture(orc):
    """The perturbed black-hole medium at 64^2 (gamma 4/3, p = 0.01, CFL 0.4,
    SLIC + RK4 forcing, the reference's excision) goes NaN in the reference's
    own arithmetic: no NaN for 33 steps, 24 cells after step 34, then spreading.
    The oracle reproduces the per-step counts and the state after 40 steps
    (tests/golden/gr_nan_onset.npz, written from the reference build).  This is
    why the black-hole workloads run 20 steps, not 200."""
    g = golden("gr_nan_onset.npz")
    W, h, hole = workloads.gr_random_field_2d(64, seed=1)
    out, counts = orc.gr_run_nan_counts(W, 2, h, 0.4, hole, len(g["nan_counts"]), gamma=4.0 / 3.0)
    assert np.array_equal(counts, g["nan_counts"])
    assert same(out, g["output"])
    first = int(np.argmax(counts > 0))
    assert first == 33 and counts[first] == 24 and (counts[:first] == 0).all()
