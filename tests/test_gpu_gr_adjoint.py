"""The adjoint of the black-hole (Kerr-Schild) update, checked against central
finite differences of the forward map on the GPU:
  - per operation, the dot-product test <lam, (S(W + h e) - S(W - h e)) / 2h>
    = <S'^T lam, e> for an x and a y SLIC / FORCE sweep with excision and for
    the RK4 forcing step;
  - per trajectory (forcing, X Y X sweeps, forcing per step), the gradient of
    L = sum (rho_N - rho*)^2 along seeded directions with the recorded dt
    schedule held fixed (cosmos_adjoint_replay).
Con2prim runs with a tight tolerance (1e-13) so that the Newton iteration
counts, and so the forward map, do not jump inside the finite-difference
stencil.  States are cell-wise random, which keeps the limiters' kinks away
from a single operation's stencil: h = 1e-6 relative, bar 1e-6.  Trajectories
use h = 1e-7 with the 1e-6 bar in at least three of four directions and 1e-3
in all (a wrong adjoint is off by O(1))."""
import numpy as np
import pytest

from cosmos_amd import BLACKHOLE, TAPE_NONE, TAPE_RERUN, TAPE_SEGMENT, Adjoint

pytestmark = pytest.mark.gpu

G = 4.0 / 3.0
TOL = 1e-6


def bh_state(n, seed, ny=None):
    """A perturbed, slowly moving medium around a hole at the centre of
    [0, 40]^2 in cell coordinates (cell (i, j) at (j h, i h))."""
    ny = ny or n
    rng = np.random.default_rng(seed)
    W = np.empty((ny, n, 5))
    W[..., 0] = rng.uniform(0.8, 1.2, (ny, n))
    W[..., 1] = rng.uniform(-0.1, 0.1, (ny, n))
    W[..., 2] = rng.uniform(-0.1, 0.1, (ny, n))
    W[..., 3] = 0.0
    W[..., 4] = rng.uniform(0.05, 0.15, (ny, n))
    return W, 40.0 / n


def make(n, ny=None, spin=0.0, excision=0.0):
    ny = ny or n
    a = Adjoint(2, n, ny, G, physics=BLACKHOLE)
    a.set_spacetime(1.0, spin, 20.0, 20.0, 0.0)
    if excision:
        a.set_excision(excision)
    a.set_con2prim(1e-13, 200)
    return a


def rel(x, y):
    return abs(x - y) / max(abs(x), abs(y), 1e-300)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("axis", [0, 1])
def test_gr_sweep_adjoint_dot_product(order, axis):
    W, h = bh_state(24, 3, ny=20)
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(W.shape)
    dt = 0.1 * h
    H = 1e-6
    with make(24, 20) as a:
        a.upload(W)
        g = a.sweep(order, axis, h, dt, lam)
        for trial in range(3):
            e = rng.standard_normal(W.shape) * np.abs(W).clip(0.05)
            exact = float(np.sum(g * e))
            a.upload(W + H * e)
            a.sweep(order, axis, h, dt, lam)
            p = a.download()
            a.upload(W - H * e)
            a.sweep(order, axis, h, dt, lam)
            m = a.download()
            fd = float(np.sum(lam * (p - m))) / (2.0 * H)
            assert abs(exact) > 1e-6 and rel(fd, exact) < TOL, (trial, fd, exact)


@pytest.mark.parametrize("spin,excision", [(0.0, 0.0), (0.5, 0.0), (0.0, 2.5)])
def test_gr_forcing_adjoint_dot_product(spin, excision):
    W, h = bh_state(20, 5, ny=22)
    rng = np.random.default_rng(6)
    lam = rng.standard_normal(W.shape)
    dt = 0.2 * h
    H = 1e-6
    with make(20, 22, spin, excision) as a:
        a.upload(W)
        g = a.forcing(h, dt, lam)
        for trial in range(3):
            e = rng.standard_normal(W.shape) * np.abs(W).clip(0.05)
            exact = float(np.sum(g * e))
            a.upload(W + H * e)
            a.forcing(h, dt, lam)
            p = a.download()
            a.upload(W - H * e)
            a.forcing(h, dt, lam)
            m = a.download()
            fd = float(np.sum(lam * (p - m))) / (2.0 * H)
            assert abs(exact) > 1e-6 and rel(fd, exact) < TOL, (trial, fd, exact)


def loss_and_grad(WN, target):
    r = WN[..., 0] - target
    g = np.zeros_like(WN)
    g[..., 0] = 2.0 * r
    return float(np.sum(r * r)), g


@pytest.mark.parametrize("order,every", [(2, 2), (1, 3)])
def test_gr_trajectory_gradient(order, every):
    W0, h = bh_state(24, 11)
    target = bh_state(24, 12)[0][..., 0]
    with make(24) as a:
        a.upload(W0)
        a.forward(order, h, 0.4, 4, 0.0, 1, every)
        WN = a.download()
        assert np.all(np.isfinite(WN))
        _, gout = loss_and_grad(WN, target)
        grad = a.backward(gout)
        assert np.array_equal(a.download(), WN)
        assert np.array_equal(a.replay(W0), WN)
        rng = np.random.default_rng(13)
        hh = 1e-7
        errs = []
        for _ in range(4):
            e = rng.standard_normal(W0.shape) * np.abs(W0).clip(0.05)
            exact = float(np.sum(grad * e))
            lp = loss_and_grad(a.replay(W0 + hh * e), target)[0]
            lm = loss_and_grad(a.replay(W0 - hh * e), target)[0]
            errs.append(rel((lp - lm) / (2 * hh), exact))
        errs = np.array(errs)
        assert np.abs(grad).max() > 1e-6
        assert (errs < TOL).sum() >= 3 and errs.max() < 1e-3, errs


def close(a, b, tol=1e-11):
    """Agreement to rounding: the adjoint kernels accumulate with atomic adds."""
    scale = max(np.abs(a).max(), np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


@pytest.mark.parametrize("spin,excision", [(0.0, 2.5), (0.5, 0.0)])
def test_gr_forcing_tape_modes(spin, excision):
    """The taped forcing adjoint (the chain written by a forward kernel, then
    walked back) against the adjoint kernel that re-runs the chain in
    registers; the tape kernel's rewrite of the step's output is bitwise the
    forward forcing."""
    W, h = bh_state(20, 21, ny=22)
    lam = np.random.default_rng(22).standard_normal(W.shape)
    dt = 0.2 * h
    out = {}
    for mode in (TAPE_NONE, TAPE_RERUN):
        with make(20, 22, spin, excision) as a:
            a.set_tape(mode)
            a.upload(W)
            out[mode] = (a.forcing(h, dt, lam), a.download())
    assert np.array_equal(out[TAPE_NONE][1], out[TAPE_RERUN][1])
    assert np.abs(out[TAPE_NONE][0]).max() > 1e-3
    assert close(out[TAPE_NONE][0], out[TAPE_RERUN][0])


def test_gr_trajectory_tape_modes():
    """The trajectory gradient is the same whether the forcing chains come from
    per-segment tapes, a re-run tape or the in-register re-run."""
    W0, h = bh_state(24, 31)
    target = bh_state(24, 32)[0][..., 0]
    grads = {}
    for mode in (TAPE_NONE, TAPE_RERUN, TAPE_SEGMENT):
        with make(24, excision=2.5) as a:
            a.set_tape(mode, 0)
            a.upload(W0)
            a.forward(2, h, 0.4, 5, 0.0, 1, 2)
            WN = a.download()
            grads[mode] = a.backward(loss_and_grad(WN, target)[1])
            assert np.array_equal(a.download(), WN)
    assert np.abs(grads[TAPE_NONE]).max() > 1e-6
    assert close(grads[TAPE_NONE], grads[TAPE_RERUN])
    assert close(grads[TAPE_NONE], grads[TAPE_SEGMENT])


@pytest.mark.parametrize("steps,every", [(6, 2), (5, 5)])
def test_gr_backward_repeats_and_recorded_segment(steps, every):
    """The forward pass leaves the last segment (states and tapes) in the
    segment buffers and the first backward uses it without recomputing it; a
    second backward (which recomputes every segment the buffers no longer
    hold) and a context that recomputes everything (TAPE_RERUN: the forward
    recorded no tapes) give the same gradient."""
    W0, h = bh_state(20, 41)
    target = bh_state(20, 42)[0][..., 0]
    grads = []
    for mode in (TAPE_SEGMENT, TAPE_RERUN):
        with make(20, excision=2.5) as a:
            a.set_tape(mode, 0)
            a.upload(W0)
            a.forward(2, h, 0.4, steps, 0.0, 1, every)
            WN = a.download()
            assert np.array_equal(a.replay(W0), WN)
            a.upload(WN)
            gout = loss_and_grad(WN, target)[1]
            grads.append(a.backward(gout))
            grads.append(a.backward(gout))
    assert np.abs(grads[0]).max() > 1e-6
    for g in grads[1:]:
        assert close(grads[0], g)


# ---------------------------------------------------------------------------
# The config-5 gradient against the oracle's finite differences.  The
# perturbed medium at 64^2 (seed 7, target seed 8, excision 2.5M, tight
# con2prim), 4 steps, L = sum (rho_N - rho*)^2, 32 seeded directions of the
# initial density field.  L(W + he) - L(W - he) is summed cell by cell as
# (rho+ - rho-)(rho+ + rho- - 2 rho*), which keeps the rounding of L itself out
# of the quotient.
# ---------------------------------------------------------------------------
C5_N, C5_STEPS, C5_DIRS = 64, 4, 32


@pytest.fixture(scope="module")
def c5(orc):
    from cosmos_amd import workloads

    W, h, hole = workloads.gr_random_field_2d(C5_N, seed=7)
    target = workloads.gr_random_field_2d(C5_N, seed=8)[0][..., 0]
    with Adjoint(2, C5_N, C5_N, G, physics=BLACKHOLE) as a:
        a.set_spacetime(*hole)
        a.set_excision(2.5)
        a.set_con2prim(1e-13, 200)
        a.upload(W)
        a.forward(2, h, 0.4, C5_STEPS, 0.0, 1, 20)
        WN = a.download()
        dts = a.time_steps()
        g = np.zeros_like(WN)
        g[..., 0] = 2.0 * (WN[..., 0] - target)
        grad = a.backward(g)
    rng = np.random.default_rng(9)
    dirs = []
    for _ in range(C5_DIRS):
        e = np.zeros_like(W)
        e[..., 0] = rng.standard_normal(W.shape[:-1])
        dirs.append(e)
    return dict(W=W, h=h, hole=hole, target=target, WN=WN, dts=dts, grad=grad, dirs=dirs)


class _oracle_c5_mode:
    """The oracle in the workload's mode: excision 2.5M, con2prim 1e-13 / 200."""

    def __init__(self, orc):
        self.orc = orc

    def __enter__(self):
        self.orc.gr_set_excision(2.5)
        self.orc.gr_set_con2prim(1e-13, 200)
        return self.orc

    def __exit__(self, *exc):
        self.orc.gr_set_excision(0.0)
        self.orc.gr_set_con2prim(0.0, 1000)


def _fd(p, q, target, hh):
    return float(np.sum((p[..., 0] - q[..., 0]) * (p[..., 0] + q[..., 0] - 2.0 * target))) / (2.0 * hh)


def test_c5_tight_forward_is_the_oracle(orc, c5):
    """The adjoint's recorded forward pass (GPU, tight con2prim) is bitwise the
    oracle's solve loop: the same dt schedule and the same final state."""
    with _oracle_c5_mode(orc) as o:
        exp, te, ne = o.gr_run(c5["W"], 2, c5["h"], 0.4, c5["hole"], max_steps=C5_STEPS, subcycles=1, gamma=G)
        t = 0.0
        for dt in c5["dts"]:
            t += float(dt)
        assert ne == C5_STEPS and te == t
        assert np.array_equal(exp, c5["WN"])
        assert np.array_equal(o.gr_run_dts(c5["W"], 2, c5["h"], c5["dts"], c5["hole"], gamma=G), c5["WN"])


@pytest.mark.parametrize("half", [0, 1])
def test_c5_gradient_matches_oracle_fd_on_the_forward_pieces(orc, c5, half):
    """h = 1e-6 central differences of the oracle's forward map held to the
    branches the unperturbed run took (each slope component keeps its recorded
    clamps and limiter piece: the map whose derivative the adjoint's
    subgradients are) agree with the GPU gradient to 1e-6 in every direction."""
    hh = 1e-6
    with _oracle_c5_mode(orc) as o:
        run = lambda W0: o.gr_run_dts(W0, 2, c5["h"], c5["dts"], c5["hole"], gamma=G)  # noqa: E731
        base, trace = o.traced(lambda: run(c5["W"]))
        assert np.array_equal(base, c5["WN"])
        errs = []
        for e in c5["dirs"][16 * half:16 * half + 16]:
            p = o.on_pieces(trace, lambda: run(c5["W"] + hh * e))
            q = o.on_pieces(trace, lambda: run(c5["W"] - hh * e))
            fd, exact = _fd(p, q, c5["target"], hh), float(np.sum(c5["grad"] * e))
            errs.append(abs(fd - exact) / max(abs(fd), abs(exact)))
    print("piece-held FD h=1e-6 rel errors:", " ".join(f"{x:.1e}" for x in errs))
    assert max(errs) <= TOL, errs


def test_c5_gradient_matches_oracle_fd_and_branch_flips(orc, c5):
    """Free central differences of the oracle's forward map: at h = 1e-8 no
    slope component changes its clamps or limiter piece between W + he and
    W - he, and the differences agree with the GPU gradient to 1e-6 in all but
    one direction and to 1e-5 in all (at h = 1e-8 the rounding of the states
    themselves, ~1e-15 per cell over 4096 cells, is ~1e-7 of the quotient and
    more where the directional derivative is small); at h = 1e-6 every
    direction flips tens of them (|delta| < 1e-5 -> +1e-5,
    Solvers/SlopeLimiters.cpp:71-101, sits where the p = 0.01 medium's tau
    differences do), which is what moves h = 1e-6 differences off the
    gradient (held to the unperturbed pieces they agree, test above)."""
    with _oracle_c5_mode(orc) as o:
        run = lambda W0: o.gr_run_dts(W0, 2, c5["h"], c5["dts"], c5["hole"], gamma=G)  # noqa: E731
        errs = []
        for e in c5["dirs"]:
            p, q = run(c5["W"] + 1e-8 * e), run(c5["W"] - 1e-8 * e)
            fd, exact = _fd(p, q, c5["target"], 1e-8), float(np.sum(c5["grad"] * e))
            errs.append(abs(fd - exact) / max(abs(fd), abs(exact)))
        flips = {}
        for hh in (1e-8, 1e-6):
            flips[hh] = []
            for e in c5["dirs"][:4]:
                _, tp = o.traced(lambda: run(c5["W"] + hh * e))
                _, tq = o.traced(lambda: run(c5["W"] - hh * e))
                flips[hh].append(int(np.count_nonzero(tp != tq)))
    print("free FD h=1e-8 rel errors:", " ".join(f"{x:.1e}" for x in errs))
    print("branch flips per direction:", flips)
    assert sum(e <= TOL for e in errs) >= C5_DIRS - 1 and max(errs) <= 10 * TOL, errs
    assert all(f == 0 for f in flips[1e-8]), flips
    assert all(f > 0 for f in flips[1e-6]), flips
