"""Point-level statics of the solver classes (one state, interface or cell per
call: the state conversions, the Lax-Friedrichs / Richtmyer / FORCE / SLIC
interface fluxes, the half-step and fractional-step evolutions, the slope
vectors, the black-hole RK4 forcing of one cell and the vector arithmetic of
the flux schemes), for the Newtonian, special-relativistic and black-hole
physics, along x and y.

  1. The oracle's restatement (orc_point) is pinned bitwise against the
     reference's own statics compiled from its sources (oracle/_ref,
     ref_point) on seeded inputs, including z != 0 positions and the
     special-relativistic con2prim extreme states.
  2. The library's host-evaluated entry points (cosmos_point_*,
     cosmos_vector_*, through cosmos_amd.point) equal the oracle bitwise on
     the same inputs; they need no GPU.
  3. Identities the mathematics fixes: uniform neighbours give a zero slope
     and a half step that returns the state itself (equal face fluxes), so
     the SLIC flux of four equal states is their FORCE flux, which is their
     physical flux; FORCE is the mean of Lax-Friedrichs and Richtmyer; the
     fractional step by 0 is the state at its own position.
"""
import numpy as np
import pytest

from conftest import SR_EXTREME_STATES
from cosmos_amd import point as pt
from oracle import (PT_C2P, PT_FLUX, PT_FORCE, PT_FORCING, PT_FRACTIONAL, PT_HALF, PT_LF, PT_LIMITER, PT_P2C,
                    PT_RICHTMYER, PT_SLIC, PT_SLOPE, PT_VEC_EVOLUTION, PT_VEC_FORCE, PT_VEC_LF, PT_VEC_RICHTMYER,
                    PT_VEC_SLOPE, PT_VEC_UPDATE)

HOLES = [(0.05, 0.0, 0.5, 0.5, 0.0), (0.1, 0.9, 0.5, 0.4, 0.0), (0.3, 0.7, 0.5, 0.4, 0.05)]
GAMMA = {0: 1.4, 1: 5.0 / 3.0, 2: 4.0 / 3.0}
STATE_FNS = [PT_P2C, PT_C2P, PT_FLUX, PT_LF, PT_RICHTMYER, PT_FORCE, PT_SLIC, PT_HALF, PT_SLOPE]
GR_FNS = [PT_FRACTIONAL, PT_FORCING]
VEC_FNS = [PT_VEC_LF, PT_VEC_RICHTMYER, PT_VEC_FORCE, PT_VEC_UPDATE, PT_VEC_EVOLUTION, PT_VEC_SLOPE, PT_LIMITER]


def bits(a, b):
    """Bitwise equality (signed zeros distinguished; NaN matches NaN)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.uint64) == b.view(np.uint64))))


def prim_states(physics, n, rng):
    W = np.empty((n, 5))
    W[:, 0] = rng.uniform(0.2, 3.0, n)
    vmax = {0: 2.0, 1: 0.5, 2: 0.3}[physics]
    W[:, 1:4] = rng.uniform(-vmax, vmax, (n, 3))
    W[:, 4] = rng.uniform(0.05, 2.0, n)
    return W


class Case:
    """One seeded call: states, parameters and (black hole) spacetime and position."""

    def __init__(self, fn, physics, axis, rng, orc, hole=None):
        self.fn, self.physics, self.axis = fn, physics, axis
        self.gamma = GAMMA[physics]
        self.dx = float(rng.uniform(0.002, 0.1))
        self.dt = float(self.dx * rng.uniform(0.05, 0.5))
        self.bias = float(rng.choice([0.0, 0.3, -0.5]))
        self.limiter = int(rng.integers(0, 3))
        self.side = int(rng.integers(0, 2))
        self.frac = float(rng.choice([0.25, 0.5, 1.0]))
        self.hole = hole
        self.pos = None
        if physics == 2:
            z = 0.0 if rng.uniform() < 0.6 else float(rng.uniform(-0.2, 0.2))
            self.pos = (float(rng.uniform(-0.5, 1.5)), float(rng.uniform(-0.5, 1.5)), z)
        S = prim_states(physics, 5, rng)
        if fn in (PT_C2P, PT_FLUX, PT_FORCING):
            # conserved vectors (of the seeded states at the evaluation point)
            S = np.array([orc.point(PT_P2C, physics, 0, s, gamma=self.gamma, hole=hole, pos=self.pos) for s in S])
        if fn >= PT_VEC_LF:
            S = rng.uniform(-2.0, 2.0, (5, 5))
            if fn == PT_LIMITER:
                S[0, 0] = rng.choice([rng.uniform(-1.0, 3.0), 0.0, 0.5, 1.0, 2.0])
        self.S = S

    def kw(self):
        return dict(dx=self.dx, dt=self.dt, bias=self.bias, frac=self.frac, gamma=self.gamma, limiter=self.limiter,
                    side=self.side, hole=self.hole, pos=self.pos)

    def product(self):
        """The same call through the library's C-ABI."""
        fn, ph, ax, S = self.fn, self.physics, self.axis, self.S
        g, h, p = self.gamma, self.hole, self.pos
        dx, dt, b, lim = self.dx, self.dt, self.bias, self.limiter
        if fn == PT_P2C:
            return pt.conserved(ph, S[0], g, h, p)
        if fn == PT_C2P:
            return pt.primitive(ph, S[0], g, h, p)
        if fn == PT_FLUX:
            return pt.flux(ph, ax, S[0], g, h, p)
        if fn in (PT_LF, PT_RICHTMYER, PT_FORCE):
            kind = {PT_LF: pt.LAX_FRIEDRICHS, PT_RICHTMYER: pt.RICHTMYER, PT_FORCE: pt.FORCE}[fn]
            return pt.interface_flux(ph, kind, ax, S[0], S[1], dx, dt, g, h, p)
        if fn == PT_SLIC:
            return pt.slic_flux(ph, ax, S[0], S[1], S[2], S[3], dx, dt, b, lim, g, h, p)
        if fn == PT_HALF:
            return pt.half_step(ph, ax, S[0], S[1], S[2], dx, dt, b, lim, self.side, g, h, p)
        if fn == PT_SLOPE:
            return pt.slope_vector(ph, ax, S[0], S[1], S[2], dx, b, lim, g, h, p)
        if fn == PT_FRACTIONAL:
            return pt.fractional_step(ax, self.frac, S[0], S[1], S[2], dx, dt, b, lim, g, h, p)
        if fn == PT_FORCING:
            top, bottom = (S[3], S[4]) if ax == 1 else (None, None)
            return pt.forcing(ax + 1, S[0], S[1], S[2], top, bottom, dx, dt, b, lim, g, h, p)
        if fn == PT_VEC_LF:
            return pt.vector_lax_friedrichs(S[0], S[1], S[2], S[3], dx, dt)
        if fn == PT_VEC_RICHTMYER:
            return pt.vector_richtmyer(S[0], S[1], S[2], S[3], dx, dt)
        if fn == PT_VEC_FORCE:
            return pt.vector_force(S[0], S[1])
        if fn == PT_VEC_UPDATE:
            return pt.vector_force_update(S[0], S[1], S[2], dx, dt)
        if fn == PT_VEC_EVOLUTION:
            return pt.vector_evolution(self.frac, S[0], S[1], dx, dt)
        if fn == PT_VEC_SLOPE:
            return pt.vector_slope(S[0], S[1], S[2], b, lim)
        if fn == PT_LIMITER:
            return np.array([pt.slope_limiter(S[0, 0], b, lim)])
        raise AssertionError(fn)


def cases(orc, physics, fns, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for fn in fns:
        for axis in (0, 1):
            for k in range(n):
                hole = HOLES[k % len(HOLES)] if physics == 2 else None
                out.append(Case(fn, physics, axis, rng, orc, hole))
    return out


# ---------------------------------------------------------------------------
# 1. oracle == reference build
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("physics", [0, 1, 2])
def test_point_oracle_matches_reference(ref, orc, physics):
    fns = STATE_FNS + (GR_FNS if physics == 2 else [])
    for c in cases(orc, physics, fns, 12, 100 + physics):
        e = ref.point(c.fn, physics, c.axis, c.S, **c.kw())
        assert bits(orc.point(c.fn, physics, c.axis, c.S, **c.kw()), e), (c.fn, physics, c.axis)


def test_point_vectors_oracle_matches_reference(ref, orc):
    for c in cases(orc, 0, VEC_FNS, 20, 7):
        e = ref.point(c.fn, 0, c.axis, c.S, **c.kw())
        assert bits(orc.point(c.fn, 0, c.axis, c.S, **c.kw()), e), c.fn


def test_point_sr_extreme_states_match_reference(ref, orc):
    """con2prim at its cap and floors, in every static that runs it."""
    ex = np.array(SR_EXTREME_STATES)
    for fn in STATE_FNS:
        for axis in (0, 1):
            for k in range(len(ex)):
                S = np.roll(ex, k, axis=0)[:5]
                if fn in (PT_C2P, PT_FLUX):
                    S = np.array([orc.point(PT_P2C, 1, 0, s, gamma=5 / 3) for s in S])
                kw = dict(dx=0.01, dt=0.004, bias=0.0, limiter=k % 3, side=k % 2, gamma=5 / 3)
                assert bits(orc.point(fn, 1, axis, S, **kw), ref.point(fn, 1, axis, S, **kw)), (fn, axis, k)


# ---------------------------------------------------------------------------
# 2. library (host-evaluated point entry points) == oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("physics", [0, 1, 2])
def test_point_library_matches_oracle(orc, physics):
    fns = STATE_FNS + (GR_FNS if physics == 2 else [])
    for c in cases(orc, physics, fns, 25, 200 + physics):
        e = orc.point(c.fn, physics, c.axis, c.S, **c.kw())
        assert bits(c.product(), e), (c.fn, physics, c.axis, c.S.tolist(), c.kw())


def test_point_vectors_library_matches_oracle(orc):
    for c in cases(orc, 0, VEC_FNS, 40, 9):
        assert bits(c.product(), orc.point(c.fn, 0, c.axis, c.S, **c.kw())), c.fn


def test_point_library_sr_extreme_states(orc):
    ex = np.array(SR_EXTREME_STATES)
    g = 5 / 3
    for axis in (0, 1):
        for k in range(len(ex)):
            S = np.roll(ex, k, axis=0)[:5]
            U = np.array([orc.point(PT_P2C, 1, 0, s, gamma=g) for s in S])
            lim, side = k % 3, k % 2
            kw = dict(dx=0.01, dt=0.004, bias=0.0, limiter=lim, side=side, gamma=g)
            assert bits(pt.conserved(1, S[0], g), orc.point(PT_P2C, 1, axis, S, **kw))
            assert bits(pt.primitive(1, U[0], g), orc.point(PT_C2P, 1, axis, U, **kw))
            assert bits(pt.flux(1, axis, U[0], g), orc.point(PT_FLUX, 1, axis, U, **kw))
            assert bits(pt.interface_flux(1, pt.FORCE, axis, S[0], S[1], 0.01, 0.004, g),
                        orc.point(PT_FORCE, 1, axis, S, **kw))
            assert bits(pt.slic_flux(1, axis, S[0], S[1], S[2], S[3], 0.01, 0.004, 0.0, lim, g),
                        orc.point(PT_SLIC, 1, axis, S, **kw))
            assert bits(pt.half_step(1, axis, S[0], S[1], S[2], 0.01, 0.004, 0.0, lim, side, g),
                        orc.point(PT_HALF, 1, axis, S, **kw))


def test_point_argument_errors():
    with pytest.raises(Exception, match="axis"):
        pt.flux(0, 2, np.ones(5), 1.4)
    with pytest.raises(Exception, match="spacetime"):
        pt.conserved(2, np.ones(5), 1.4)
    with pytest.raises(Exception, match="kind"):
        pt.interface_flux(0, 3, 0, np.ones(5), np.ones(5), 0.1, 0.01, 1.4)
    with pytest.raises(Exception, match="side"):
        pt.half_step(0, 0, np.ones(5), np.ones(5), np.ones(5), 0.1, 0.01, 0.0, 1, 2, 1.4)
    with pytest.raises(Exception, match="physics"):
        pt.slic_flux(5, 0, *(np.ones(5),) * 4, 0.1, 0.01, 0.0, 1, 1.4)


# ---------------------------------------------------------------------------
# 3. identities
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("physics", [0, 1, 2])
def test_point_uniform_identities(orc, physics):
    rng = np.random.default_rng(31 + physics)
    g = GAMMA[physics]
    for k in range(10):
        w = prim_states(physics, 1, rng)[0]
        hole = HOLES[k % 3] if physics == 2 else None
        for axis in (0, 1):
            pos = (0.9 + 0.1 * k, 1.3, 0.0) if physics == 2 else None
            if physics == 2:
                # a uniform state is not uniform in conserved variables on a curved
                # background; check the flat-space hole (mass 0) there
                hole = (0.0, 0.0, 0.5, 0.5, 0.0)
            dx, dt = 0.01, 0.004
            s = pt.slope_vector(physics, axis, w, w, w, dx, 0.0, k % 3, g, hole, pos)
            assert np.all(s == 0.0)
            for side in (0, 1):
                hs = pt.half_step(physics, axis, w, w, w, dx, dt, 0.0, k % 3, side, g, hole, pos)
                # the evolved face state round-trips through c2p(p2c(w))
                ww = pt.primitive(physics, pt.conserved(physics, w, g, hole, pos), g, hole, pos)
                if physics != 2:
                    assert bits(hs, ww)
            lf = pt.interface_flux(physics, pt.LAX_FRIEDRICHS, axis, w, w, dx, dt, g, hole, pos)
            ri = pt.interface_flux(physics, pt.RICHTMYER, axis, w, w, dx, dt, g, hole, pos)
            fo = pt.interface_flux(physics, pt.FORCE, axis, w, w, dx, dt, g, hole, pos)
            assert bits(fo, pt.vector_force(lf, ri))
            if physics != 2:
                u = pt.conserved(physics, w, g)
                f = pt.flux(physics, axis, u, g)
                assert bits(lf, f)  # 0.5 (2 F) exactly, the dx/dt term vanishes
                sl = pt.slic_flux(physics, axis, w, w, w, w, dx, dt, 0.0, k % 3, g)
                hw = pt.half_step(physics, axis, w, w, w, dx, dt, 0.0, k % 3, 0, g)
                assert bits(sl, pt.interface_flux(physics, pt.FORCE, axis, hw, hw, dx, dt, g))
    if physics == 2:
        w = prim_states(2, 1, rng)[0]
        hole, pos = HOLES[1], (1.2, 0.3, 0.0)
        e = pt.fractional_step(0, 0.0, w, w, w, 0.01, 0.004, 0.0, 1, g, hole, pos)
        assert bits(e, pt.primitive(2, pt.conserved(2, w, g, hole, pos), g, hole, pos))


# ---------------------------------------------------------------------------
# The drop-in classes' point-level statics (tests/dropin/dropin_check.cpp,
# sections 20-22), through the C++ drop-in library against the oracle
# ---------------------------------------------------------------------------
def point_states(a):
    return np.array([(1.0 + 0.1 * k, a * (0.2 - 0.05 * k), a * (-0.1 + 0.03 * k), a * (0.04 * k), 1.0 + 0.2 * k)
                     for k in range(5)])


def point_vectors():
    return np.array([[(-1.0 if c % 2 else 1.0) * (0.3 + 0.1 * k + 0.05 * c) for c in range(5)] for k in range(5)])


def padded_1d(s, size):
    """insertBoundaryCells (Solvers/EulerSolvers.cpp:3-28)."""
    n = len(s)
    out = np.tile(np.array([1.0, 0.0, 0.0, 0.0, 1.0]), (n + 2 * size, 1))
    out[size:size + n] = s
    if size == 1:
        out[0], out[n + 1] = s[0], s[n - 1]
    else:
        out[0], out[1], out[n + 2], out[n + 3] = s[1], s[0], s[n - 1], s[n - 2]
    return out


def padded_2d(g, size):
    """insertBoundaryCells2D (Solvers/EulerSolvers.cpp:30-85); corners default (1, 0, 0, 0, 1)."""
    R, C = g.shape[:2]
    out = np.tile(np.array([1.0, 0.0, 0.0, 0.0, 1.0]), (R + 2 * size, C + 2 * size, 1))
    out[size:size + R, size:size + C] = g
    if size == 1:
        out[1:R + 1, 0], out[1:R + 1, C + 1] = g[:, 0], g[:, C - 1]
        out[0, 1:C + 1], out[R + 1, 1:C + 1] = g[0], g[R - 1]
    else:
        out[2:R + 2, 0], out[2:R + 2, 1] = g[:, 1], g[:, 0]
        out[2:R + 2, C + 2], out[2:R + 2, C + 3] = g[:, C - 1], g[:, C - 2]
        out[0, 2:C + 2], out[1, 2:C + 2] = g[1], g[0]
        out[R + 2, 2:C + 2], out[R + 3, 2:C + 2] = g[R - 1], g[R - 2]
    return out


def expected_points(orc, ref=None):
    """The values sections 20-22 of the drop-in check write, from the oracle
    (and the boundary layouts from the reference build where it is present)."""
    dx, dt, bias, lim = 0.01, 0.004, 0.3, 0
    v = point_vectors()
    xl = np.array([1.0, 0.1, -0.2, 0.05, 2.5])
    xr = np.array([1.1, 0.2, -0.1, 0.0, 2.7])
    e = 0.01 * v[0]
    out = []
    grid = lambda s: np.array([[s[(i + j) % 5] for j in range(4)] for i in range(3)])

    def pad1(ph, s, size):
        return ref.insert_boundary(s, ph, size) if ref is not None else padded_1d(s, size)

    def pad2(ph, g, size):
        return ref.insert_boundary(g, ph, size) if ref is not None else padded_2d(g, size)

    for ph, a, g in ((0, 1.0, 1.4), (1, 0.5, 5 / 3)):
        s = point_states(a)
        kw = dict(dx=dx, dt=dt, bias=bias, limiter=lim, gamma=g)
        for axis in (0, 1):
            for fn in (PT_LF, PT_RICHTMYER, PT_FORCE):
                out.append(orc.point(fn, ph, axis, s[:2], **kw))
        for axis in (0, 1):
            out.append(orc.point(PT_SLIC, ph, axis, s[:4], **kw))
        for axis in (0, 1):
            for side in (0, 1):
                out.append(orc.point(PT_HALF, ph, axis, s[:3], side=side, **kw))
        out.append(orc.point(PT_SLOPE, ph, 0, s[:3], **kw))
        if ph == 0:
            out.append(orc.point(PT_VEC_LF, 0, 0, v[:4], **kw))
            out.append(orc.point(PT_VEC_RICHTMYER, 0, 0, v[:4], **kw))
            out.append(orc.point(PT_VEC_FORCE, 0, 0, v[:2], **kw))
            out.append(orc.point(PT_VEC_UPDATE, 0, 0, v[:3], **kw))
            out.append(orc.point(PT_VEC_EVOLUTION, 0, 0, v[:2], frac=0.25, **kw))
            out.append(orc.point(PT_VEC_EVOLUTION, 0, 0, v[:2], frac=0.5, **kw))
            out.append(orc.point(PT_VEC_SLOPE, 0, 0, v[:3], **kw))
        for side in (0, 1):
            out.append(orc.point(PT_C2P, ph, 0, (xl if side == 0 else xr) + e, gamma=g))
        out.append(pad1(ph, s, 2))
        if ph == 0:
            out.append(pad1(0, s, 1))
            out.append(pad2(0, grid(s), 2))
            out.append(pad2(0, grid(s), 1))
    hole, pos = (0.1, 0.9, 0.5, 0.4, 0.0), (1.1, 0.7, 0.05)
    s = point_states(0.3)
    kw = dict(dx=dx, dt=dt, bias=bias, limiter=lim, gamma=4 / 3, hole=hole, pos=pos)
    for axis in (0, 1):
        for fn in (PT_LF, PT_RICHTMYER, PT_FORCE):
            out.append(orc.point(fn, 2, axis, s[:2], **kw))
    for axis in (0, 1):
        out.append(orc.point(PT_SLIC, 2, axis, s[:4], **kw))
    for axis in (0, 1):
        for side in (0, 1):
            out.append(orc.point(PT_HALF, 2, axis, s[:3], side=side, **kw))
    for axis in (0, 1):
        out.append(orc.point(PT_SLOPE, 2, axis, s[:3], **kw))
    for axis in (0, 1):
        out.append(orc.point(PT_FRACTIONAL, 2, axis, s[:3], frac=0.5, **kw))
    x, y, z = pos
    for axis in (0, 1):
        for side in (0, 1):
            off = -(dx * 0.5) if side == 0 else dx * 0.5
            p = (x + off, y, z) if axis == 0 else (x, y + off, z)
            out.append(orc.point(PT_C2P, 2, 0, (xl if side == 0 else xr) + e, gamma=4 / 3, hole=hole, pos=p))
    out.append(orc.point(PT_C2P, 2, 0, xl + e, gamma=4 / 3, hole=hole, pos=pos))
    c = np.array([orc.point(PT_P2C, 2, 0, w, gamma=4 / 3, hole=hole, pos=pos) for w in s])
    out.append(orc.point(PT_FORCING, 2, 0, c, **kw))
    out.append(orc.point(PT_FORCING, 2, 1, c, **kw))
    out.append(pad2(2, grid(s), 1))
    return np.concatenate([np.asarray(o).ravel() for o in out])


def test_boundary_layout_matches_reference(ref):
    """The boundary layouts the expectation uses, against the reference's own insertBoundaryCells{,2D}."""
    s = point_states(1.0)
    g = np.array([[s[(i + j) % 5] for j in range(4)] for i in range(3)])
    for ph in (0, 1, 2):
        for size in (1, 2):
            assert bits(ref.insert_boundary(s, ph, size), padded_1d(s, size))
            assert bits(ref.insert_boundary(g, ph, size), padded_2d(g, size))


def test_dropin_point_statics_match_oracle(orc, tmp_path):
    """The drop-in classes' point-level statics (host-evaluated: no GPU)."""
    import os
    import subprocess

    from conftest import ROOT

    exe = os.path.join(ROOT, "tests", "dropin", "dropin_check")
    if not os.path.exists(exe):
        pytest.fail("tests/dropin/dropin_check is not built (python -m cosmos_amd.build)")
    out = tmp_path / "points.bin"
    subprocess.run([exe, str(out), "points"], check=True, timeout=60)
    got = np.fromfile(out, dtype=np.float64)
    exp = expected_points(orc)
    assert got.size == exp.size == 1005
    assert bits(got, exp), np.flatnonzero(~((got == exp) | (np.isnan(got) & np.isnan(exp))))[:10]
