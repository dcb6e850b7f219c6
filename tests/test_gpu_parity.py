"""GPU parity: the HIP library (through the C-ABI) against the CPU oracle on
the same seeded inputs.  The bar is bitwise equality of every primitive
value (the kernels keep the reference's IEEE operation order), which is
stronger than the 1e-11 (Newtonian) / 1e-9 (SR) relative bound."""
import math

import numpy as np
import pytest

from conftest import golden, rel_linf, same
from cosmos_amd import NEWTONIAN, RELATIVISTIC, Solver, workloads

pytestmark = pytest.mark.gpu

G_N, G_SR = 1.4, 5.0 / 3.0


def gpu_run(W, physics, order, dx, cfl, tf=math.inf, max_steps=-1, bias=0.0, limiter=1, gamma=G_N):
    dim = 1 if W.ndim == 2 else 2
    nx = W.shape[-2]
    ny = W.shape[0] if dim == 2 else 1
    with Solver(physics, dim, nx, ny, gamma) as s:
        s.upload(W)
        t, n = s.solve(order, dx, cfl, tf, max_steps, bias, limiter)
        return s.download(), t, n


def test_sod_config1_matches_golden():
    g = golden("sod_1d.npz")
    for order in (1, 2):
        out, t, n = gpu_run(g["input"], NEWTONIAN, order, 1.0 / 400, 0.8, tf=0.25)
        assert same(out, g[f"order{order}_output"]), rel_linf(out, g[f"order{order}_output"])
        assert t == float(g[f"order{order}_time"]) and n == int(g[f"order{order}_steps"])


def test_newtonian_2d_matches_golden():
    g = golden("newtonian_2d.npz")
    for order in (1, 2):
        for lim in ((0, 1, 2) if order == 2 else (1,)):
            out, t, n = gpu_run(g["input"], NEWTONIAN, order, 1.0 / 24, 0.9, max_steps=3, limiter=lim)
            assert same(out, g[f"o{order}_l{lim}_output"]), (order, lim)
            assert t == float(g[f"o{order}_l{lim}_time"])
    out, _, _ = gpu_run(g["input"], NEWTONIAN, 2, 1.0 / 24, 0.9, max_steps=3, bias=0.3, limiter=1)
    assert same(out, g["o2_l1_bias03_output"])


def test_relativistic_matches_golden():
    g = golden("relativistic.npz")
    for order in (1, 2):
        out, t, _ = gpu_run(g["riemann_input"], RELATIVISTIC, order, 1.0 / 200, 0.5, max_steps=20, gamma=G_SR)
        assert same(out, g[f"riemann_o{order}_output"]), order
        assert t == float(g[f"riemann_o{order}_time"])
    out, t, _ = gpu_run(g["blast_input"], RELATIVISTIC, 2, 1.0 / 20, 0.5, max_steps=3, gamma=G_SR)
    assert same(out, g["blast_o2_output"])


# shapes: minimum SLIC grid, ragged sizes that do not divide the x tile (252) or the
# y march (64 rows), a single column of x tiles, and a wide strip.
SHAPES = [(2, 2), (3, 5), (17, 9), (300, 257), (253, 65), (700, 3)]


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("order", [1, 2])
def test_newtonian_sweeps_bitwise(orc, nx, ny, order):
    W = workloads.quadrants_2d(nx, seed=11, ny=ny)
    dx = 1.0 / max(nx, ny)
    for axis in (0, 1):
        for lim, bias in ((1, 0.0), (0, 0.0), (2, 0.0), (1, 0.3)):
            if order == 1 and (lim, bias) != (1, 0.0):
                continue
            with Solver(NEWTONIAN, 2, nx, ny, G_N) as s:
                s.upload(W)
                s.sweep(order, axis, dx, 0.2 * dx, bias, lim)
                out = s.download()
            exp = orc.sweep(W, 0, order, axis, dx, 0.2 * dx, bias, lim, G_N)
            assert same(out, exp), (axis, lim, bias, rel_linf(out, exp))


@pytest.mark.parametrize("nx,ny", [(5, 4), (300, 257), (64, 130)])
def test_relativistic_sweeps_bitwise(orc, nx, ny):
    W = workloads.sr_blast_2d(max(nx, ny), seed=12)[:ny, :nx].copy()
    dx = 1.0 / max(nx, ny)
    for order in (1, 2):
        for axis in (0, 1):
            with Solver(RELATIVISTIC, 2, nx, ny, G_SR) as s:
                s.upload(W)
                s.sweep(order, axis, dx, 0.1 * dx, 0.0, 1)
                out = s.download()
            exp = orc.sweep(W, 1, order, axis, dx, 0.1 * dx, 0.0, 1, G_SR)
            assert same(out, exp), (order, axis, rel_linf(out, exp))


def test_wavespeed_and_stable_dt_exact(orc):
    for physics, gamma, W in ((0, G_N, workloads.quadrants_2d(97, seed=3, ny=61)),
                              (1, G_SR, workloads.sr_blast_2d(80, seed=5))):
        with Solver(physics, 2, W.shape[1], W.shape[0], gamma) as s:
            s.upload(W)
            assert s.max_wave_speed() == orc.max_wavespeed(W, physics, gamma)
            for it, tf in ((0, 1.0), (6, 1.0), (7, 1e-6), (2, math.inf)):
                assert s.stable_time_step(0.01, 0.9, 0.0, tf, it) == orc.stable_dt(W, physics, 0.01, 0.9, 0.0, tf, it,
                                                                                  gamma)
        W1 = W[W.shape[0] // 2]
        with Solver(physics, 1, W1.shape[0], 1, gamma) as s:
            s.upload(W1)
            assert s.max_wave_speed() == orc.max_wavespeed(W1, physics, gamma)


@pytest.mark.parametrize("physics,gamma,n,steps", [(0, G_N, 256, 40), (1, G_SR, 128, 12)])
@pytest.mark.parametrize("order", [1, 2])
def test_solve_fixed_steps_bitwise(orc, physics, gamma, n, steps, order):
    W = workloads.quadrants_2d(n, seed=1) if physics == 0 else workloads.sr_blast_2d(n, seed=2)
    cfl = 0.9 if physics == 0 else 0.5
    out, t, k = gpu_run(W, physics, order, 1.0 / n, cfl, max_steps=steps, gamma=gamma)
    exp, te, ke = orc.run(W, physics, order, 1.0 / n, cfl, max_steps=steps, gamma=gamma)
    assert k == ke == steps and t == te
    assert same(out, exp), rel_linf(out, exp)


def test_solve_to_time_and_resume(orc):
    """Run to a final time (the loop stops and clamps dt), then a second solve
    restarts t and the damping counter like the reference's solve()."""
    W = workloads.quadrants_2d(96, seed=8, ny=80)
    with Solver(NEWTONIAN, 2, 96, 80, G_N) as s:
        s.upload(W)
        t, k = s.solve(2, 1 / 96, 0.9, 0.02)
        out1 = s.download()
        t2, k2 = s.solve(2, 1 / 96, 0.9, 0.01)
        out2 = s.download()
    e1, te, ke = orc.run(W, 0, 2, 1 / 96, 0.9, tf=0.02)
    assert same(out1, e1) and t == te == 0.02 and k == ke
    e2, te2, ke2 = orc.run(e1, 0, 2, 1 / 96, 0.9, tf=0.01)
    assert same(out2, e2) and t2 == te2 and k2 == ke2


def test_advance_in_chunks_matches_one_run(orc):
    W = workloads.quadrants_2d(128, seed=21)
    with Solver(NEWTONIAN, 2, 128, 128, G_N) as s:
        s.upload(W)
        s.begin(2, 1 / 128, 0.9, math.inf, 10)
        s.advance(3)
        s.advance(4)
        s.advance(9)  # the loop stops at 10; the rest are no-ops
        t, n, stopped = s.progress()
        out = s.download()
    exp, te, ne = orc.run(W, 0, 2, 1 / 128, 0.9, max_steps=10)
    assert n == 10 and stopped and t == te and same(out, exp)


def test_config2_1024_100_steps_bitwise(orc):
    """The 4-quadrant workload at 1024^2 for the full 100 steps, bitwise."""
    W = workloads.quadrants_2d(1024, seed=0)
    out, t, k = gpu_run(W, NEWTONIAN, 2, 1.0 / 1024, 0.9, max_steps=100)
    exp, te, ke = orc.run(W, 0, 2, 1.0 / 1024, 0.9, max_steps=100)
    assert k == ke == 100 and t == te
    assert same(out, exp), rel_linf(out, exp)


def test_config2_full_size_bitwise(orc):
    """4096^2, three full Strang steps, bitwise against the oracle."""
    W = workloads.quadrants_2d(4096, seed=1)
    out, t, k = gpu_run(W, NEWTONIAN, 2, 1.0 / 4096, 0.9, max_steps=3)
    exp, te, ke = orc.run(W, 0, 2, 1.0 / 4096, 0.9, max_steps=3)
    assert t == te
    assert same(out, exp), rel_linf(out, exp)


def test_config3_relativistic_reduced_bitwise(orc):
    """The seeded SR blast at 512^2 for 20 steps (reference con2prim constants)."""
    W = workloads.sr_blast_2d(512, seed=0)
    out, t, k = gpu_run(W, RELATIVISTIC, 2, 1.0 / 512, 0.5, max_steps=20, gamma=G_SR)
    exp, te, ke = orc.run(W, 1, 2, 1.0 / 512, 0.5, max_steps=20, gamma=G_SR)
    assert t == te and same(out, exp), rel_linf(out, exp)


@pytest.mark.parametrize("nx,ny", [(17, 9), (40, 33), (5, 4), (300, 70)])
def test_relativistic_extreme_states_bitwise(orc, nx, ny):
    """Rootless (|v| > 1: Newton to the 1000-iteration cap), W-clamped,
    density- and pressure-floored and negative-pressure states planted among
    the blast's cells: every sweep (orders 1 and 2, both axes, every y-sweep
    layout), the 1D sweep of a row and a short solve, bitwise against the
    oracle (RelativisticEuler/RelativisticEulerStateVector.cpp:30-136)."""
    from conftest import sr_extreme_2d

    W = sr_extreme_2d(nx, ny, seed=3)
    dx, dt = 1.0 / 40, 0.1 / 40
    for order in (1, 2):
        for axis in (0, 1):
            exp = orc.sweep(W, 1, order, axis, dx, dt, 0.0, 1, G_SR)
            for columns in ((0, 8, 16) if axis == 1 else (-1,)):
                with Solver(RELATIVISTIC, 2, nx, ny, G_SR) as s:
                    if columns >= 0:
                        s.set_y_tile(columns)
                    s.upload(W)
                    s.sweep(order, axis, dx, dt, 0.0, 1)
                    out = s.download()
                assert same(out, exp), (order, axis, columns, rel_linf(out, exp))
    row = W[ny // 2].copy()
    for order in (1, 2):
        with Solver(RELATIVISTIC, 1, nx, 1, G_SR) as s:
            s.upload(row)
            s.sweep(order, 0, dx, dt, 0.0, 1)
            assert same(s.download(), orc.sweep(row, 1, order, 0, dx, dt, 0.0, 1, G_SR)), order
    out, t, k = gpu_run(W, RELATIVISTIC, 2, dx, 0.5, max_steps=4, gamma=G_SR)
    exp, te, ke = orc.run(W, 1, 2, dx, 0.5, max_steps=4, gamma=G_SR)
    assert (t, k) == (te, ke) and same(out, exp), rel_linf(out, exp)


def test_config3_2048_past_the_cap_hitting_steps_bitwise(orc):
    """The seeded SR blast at its full 2048^2 for 20 steps: from step 13 on,
    interface states with no con2prim root appear and their Newton solves run
    to the reference's 1000-iteration cap; the GPU stays bitwise the oracle."""
    W = workloads.sr_blast_2d(2048, seed=0)
    with Solver(RELATIVISTIC, 2, 2048, 2048, G_SR) as s:
        s.upload(W)
        t, k = s.solve(2, 1.0 / 2048, 0.5, math.inf, 20)
        out = s.download()
        st = s.status()
    orc.events_reset()
    exp, te, ke = orc.run(W, 1, 2, 1.0 / 2048, 0.5, max_steps=20, gamma=G_SR)
    ev = orc.events()
    print(f"SR blast 2048^2 x 20 steps: oracle con2prim events (cap, q, rho, p floors) {ev}, device {st}")
    assert ev[0] > 0
    assert k == ke == 20 and t == te
    assert same(out, exp), rel_linf(out, exp)
    assert (st["c2p_cap"], st["c2p_q_floor"], st["c2p_density_floor"], st["c2p_pressure_floor"]) == ev


def test_relativistic_config_mode_con2prim(orc):
    """Tighter con2prim (tol 1e-12, <= 50 iterations) on both sides."""
    W = workloads.sr_blast_2d(96, seed=6)
    orc.set_con2prim(1e-12, 50)
    try:
        exp, te, _ = orc.run(W, 1, 2, 1.0 / 96, 0.5, max_steps=5, gamma=G_SR)
    finally:
        orc.set_con2prim(1e-8, 1000)
    with Solver(RELATIVISTIC, 2, 96, 96, G_SR) as s:
        s.set_con2prim(1e-12, 50)
        s.upload(W)
        t, _ = s.solve(2, 1.0 / 96, 0.5, math.inf, 5)
        out = s.download()
    assert t == te and same(out, exp)


def test_nan_propagates_like_reference(orc):
    """A NaN cell is carried as a value (the reference has no checks), and the
    wave-speed maximum skips it."""
    W = workloads.quadrants_2d(40, seed=2)
    W[10, 10, 4] = np.nan
    with Solver(NEWTONIAN, 2, 40, 40, G_N) as s:
        s.upload(W)
        assert s.max_wave_speed() == orc.max_wavespeed(W, 0, G_N)
        s.sweep(2, 0, 1 / 40, 0.002)
        out = s.download()
    exp = orc.sweep(W, 0, 2, 0, 1 / 40, 0.002)
    assert same(out, exp) and np.isnan(out).any()


# ---------------------------------------------------------------------------
# Fast mode: Newtonian SLIC / van Leer re-associated for throughput.  The bar
# is the relative L-infinity difference per conserved variable against the
# reference arithmetic after the full step count: <= 1e-11.
# ---------------------------------------------------------------------------
FAST_TOL = 1e-11


def fast_run(W, n, steps):
    from cosmos_amd import FAST

    with Solver(NEWTONIAN, 2, W.shape[1], W.shape[0], G_N) as s:
        s.set_mode(FAST)
        s.upload(W)
        t, k = s.solve(2, 1.0 / n, 0.9, math.inf, steps)
        return s.download(), t, k


@pytest.mark.parametrize("n,seed", [(256, 0), (1024, 0), (1024, 2)])
def test_fast_mode_100_steps_within_tolerance(orc, n, seed):
    from conftest import cons_rel_linf

    W = workloads.quadrants_2d(n, seed=seed)
    out, t, k = fast_run(W, n, 100)
    exp, te, ke = orc.run(W, 0, 2, 1.0 / n, 0.9, max_steps=100)
    err = cons_rel_linf(out, exp)
    assert k == 100 and abs(t - te) <= FAST_TOL * te
    assert err <= FAST_TOL, err


def test_fast_mode_with_nonzero_z_velocity(orc):
    """The z-velocity plane is carried when it is not identically zero."""
    from conftest import cons_rel_linf

    W = workloads.quadrants_2d(300, seed=4, ny=200)
    W[..., 3] = 0.1 * W[..., 1]
    out, t, k = fast_run(W, 300, 30)
    exp, te, ke = orc.run(W, 0, 2, 1.0 / 300, 0.9, max_steps=30)
    assert cons_rel_linf(out, exp) <= FAST_TOL


def test_fast_mode_config2_full_size_100_steps(orc):
    """The benchmark workload itself: 4096^2, 100 steps, fast mode against the
    reference arithmetic (the oracle, multi-threaded) within 1e-11."""
    from conftest import cons_rel_linf

    W = workloads.quadrants_2d(4096, seed=0)
    out, t, k = fast_run(W, 4096, 100)
    exp, te, ke = orc.run(W, 0, 2, 1.0 / 4096, 0.9, max_steps=100)
    err = cons_rel_linf(out, exp)
    print(f"fast mode 4096^2 x 100 steps: rel L-inf per conserved variable {err:.3e}")
    assert k == 100 and err <= FAST_TOL, err


def fast_solve(W, steps, fuse, tf=math.inf, rows=0):
    """A fast-mode solve with the fused step kernel (rows per block: rows, 0
    automatic) or with three sweeps: (state, time, steps)."""
    from cosmos_amd import FAST

    ny, nx = W.shape[:2]
    with Solver(NEWTONIAN, 2, nx, ny, G_N) as s:
        s.set_mode(FAST)
        s.set_step_fusion(fuse, rows)
        s.upload(W)
        t, k = s.solve(2, 1.0 / max(nx, ny), 0.9, tf, steps)
        return s.download(), t, k


def fast_pair(W, steps, tf=math.inf, rows=0):
    """The same fast-mode solve with the fused step kernel and with three sweeps."""
    return [fast_solve(W, steps, fuse, tf, rows) for fuse in (True, False)]


# The fused step hands each sweep's updated conserved vector straight to the
# next sweep, where the three sweeps store primitives between them and convert
# back: the two differ by rounding only (each against the oracle: FAST_TOL).
FUSED_TOL = 1e-12


@pytest.mark.parametrize("zvel", [False, True])
def test_fused_step_matches_the_three_sweeps(zvel):
    """The fused X(dt/2) Y(dt) X(dt/2) kernel against the three fast sweeps:
    equal step counts, states and times within rounding, on shapes at and
    around its 248-column strips and 64-row blocks (edge strips that need the
    mirrored ghost columns on both sides, a single strip narrower than 248, a
    last strip of one column), with and without the z-velocity plane; a wrong
    halo or ghost column would differ by the size of the state's jumps."""
    from conftest import cons_rel_linf

    shapes = ((8, 8), (9, 13), (247, 70), (248, 65), (249, 130), (250, 64), (253, 9), (497, 129), (744, 200),
              (1000, 67))
    worst = 0.0
    for idx, (nx, ny) in enumerate(shapes):
        W = workloads.quadrants_2d(nx, seed=20 + idx, ny=ny)
        if zvel:
            W[..., 3] = 0.1 * W[..., 2]
        (a, ta, ka), (b, tb, kb) = fast_pair(W, 9)
        err = cons_rel_linf(a, b)
        worst = max(worst, err)
        assert ka == kb == 9 and abs(ta - tb) <= FUSED_TOL * tb, (nx, ny, ta, tb)
        assert err <= FUSED_TOL, (nx, ny, err)
    print(f"fused against three sweeps, 9 steps: worst rel L-inf {worst:.3e}")
    # final-time clamp and early stop inside the queued steps
    W = workloads.quadrants_2d(300, seed=31, ny=140)
    (a, ta, ka), (b, tb, kb) = fast_pair(W, 1000, tf=0.01)
    assert ka == kb and 0 < ka < 1000 and abs(ta - tb) <= FUSED_TOL * tb
    assert cons_rel_linf(a, b) <= FUSED_TOL


def test_fused_step_rows_per_block_is_a_launch_shape():
    """Rows per block of the fused kernel change only the launch: bitwise
    equal states, times and step counts."""
    for rows in (4, 13, 200):
        W = workloads.quadrants_2d(253, seed=40 + rows, ny=150)
        a, ta, ka = fast_solve(W, 5, True, rows=rows)
        b, tb, kb = fast_solve(W, 5, True)
        assert ka == kb and ta == tb and same(a, b), rows


def test_consecutive_solves_continue_from_the_last_state(orc):
    """A second cosmos_solve starts from the state the first one left (its
    queued steps settled first), in exact mode bitwise the oracle's two runs,
    and in fast mode the same with and without the fused step."""
    from cosmos_amd import FAST

    W = workloads.quadrants_2d(96, seed=33, ny=80)
    e1, _, _ = orc.run(W, 0, 2, 1.0 / 96, 0.9, max_steps=5)
    e2, te, ke = orc.run(e1, 0, 2, 1.0 / 96, 0.9, max_steps=7)
    with Solver(NEWTONIAN, 2, 96, 80, G_N) as s:
        s.upload(W)
        s.solve(2, 1.0 / 96, 0.9, math.inf, 5)
        t, k = s.solve(2, 1.0 / 96, 0.9, math.inf, 7)
        assert same(s.download(), e2) and (t, k) == (te, ke)
    outs = []
    for fuse in (True, False):
        with Solver(NEWTONIAN, 2, 96, 80, G_N) as s:
            s.set_mode(FAST)
            s.set_step_fusion(fuse)
            s.upload(W)
            s.solve(2, 1.0 / 96, 0.9, math.inf, 5)
            outs.append((s.solve(2, 1.0 / 96, 0.9, math.inf, 7), s.download()))
    from conftest import cons_rel_linf

    (t0, k0), (t1, k1) = outs[0][0], outs[1][0]
    assert k0 == k1 and abs(t0 - t1) <= FUSED_TOL * t1
    assert cons_rel_linf(outs[0][1], outs[1][1]) <= FUSED_TOL
    assert cons_rel_linf(outs[0][1], e2) <= FAST_TOL


def test_fused_step_full_size_against_oracle(orc):
    """Fused fast steps at 1024^2 against the reference arithmetic within 1e-11
    (and against the three fast sweeps within rounding)."""
    from conftest import cons_rel_linf

    W = workloads.quadrants_2d(1024, seed=5)
    (a, ta, ka), (b, tb, kb) = fast_pair(W, 40)
    exp, te, ke = orc.run(W, 0, 2, 1.0 / 1024, 0.9, max_steps=40)
    assert ka == kb == ke == 40 and abs(ta - te) <= FAST_TOL * te
    assert cons_rel_linf(a, b) <= FUSED_TOL
    assert cons_rel_linf(a, exp) <= FAST_TOL


@pytest.mark.parametrize("columns", [0, 8, 16])
@pytest.mark.parametrize("physics", [NEWTONIAN, RELATIVISTIC])
def test_y_sweep_layout_does_not_change_results(orc, columns, physics):
    """Column march or LDS tiles of 8 / 16 columns for the y sweep: bitwise
    the oracle's sweep on shapes that do not divide the tiles, for every
    order and limiter, and through the fused-wave-speed solve loop."""
    gamma = G_N if physics == NEWTONIAN else G_SR
    for nx, ny in ((37, 70), (130, 33), (5, 3), (17, 2)):
        W = workloads.quadrants_2d(nx, seed=14, ny=ny)
        if physics == RELATIVISTIC:
            W[..., 1:3] *= 0.5
        dx = 1.0 / max(nx, ny)
        for order, limiter, bias in ((1, 1, 0.0), (2, 1, 0.0), (2, 0, 0.4), (2, 2, -0.3)):
            with Solver(physics, 2, nx, ny, gamma) as s:
                s.set_y_tile(columns)
                s.upload(W)
                s.sweep(order, 1, dx, 0.2 * dx, bias, limiter)
                out = s.download()
            exp = orc.sweep(W, physics, order, 1, dx, 0.2 * dx, bias, limiter, gamma)
            assert same(out, exp), (nx, ny, order, limiter, rel_linf(out, exp))
    W = workloads.quadrants_2d(48, seed=15, ny=40)
    if physics == RELATIVISTIC:
        W[..., 1:3] *= 0.5
    exp, te, ne = orc.run(W, physics, 2, 1.0 / 48, 0.5, math.inf, 7, 0.0, 1, gamma)
    with Solver(physics, 2, 48, 40, gamma) as s:
        s.set_y_tile(columns)
        s.upload(W)
        t, n = s.solve(2, 1.0 / 48, 0.5, math.inf, 7)
        assert same(s.download(), exp) and (t, n) == (te, ne)


@pytest.mark.parametrize("y_rows", [16, 32, 64])
@pytest.mark.parametrize("physics", [NEWTONIAN, RELATIVISTIC])
def test_y_sweep_tile_does_not_change_results(orc, y_rows, physics):
    """The y-sweep march length is a launch shape only: bitwise the same on
    shapes that do not divide 16, 32 or 64."""
    gamma = G_N if physics == NEWTONIAN else G_SR
    for nx, ny in ((37, 70), (130, 33)):
        W = workloads.quadrants_2d(nx, seed=13, ny=ny)
        if physics == RELATIVISTIC:
            W[..., 1:3] *= 0.5
        dx = 1.0 / max(nx, ny)
        for order in (1, 2):
            with Solver(physics, 2, nx, ny, gamma) as s:
                s.set_y_tile(0)
                s.set_launch(y_rows)
                s.upload(W)
                s.sweep(order, 1, dx, 0.2 * dx)
                out = s.download()
            assert same(out, orc.sweep(W, physics, order, 1, dx, 0.2 * dx, 0.0, 1, gamma)), (nx, ny, order)
