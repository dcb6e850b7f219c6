"""Writes the golden fixtures in tests/golden/*.npz from the reference's own
sources, compiled by oracle/Makefile into oracle/_ref/libcosmos_ref.so.

Run in a container that has the reference checked out (it is not needed
afterwards; the fixtures are committed):
    make -C oracle ref && python tests/golden/make_golden.py
    (python tests/golden/make_golden.py amr_level2 writes only amr_level2.npz,
     ... gr_amr only gr_amr.npz, ... gr_nan_onset only gr_nan_onset.npz)

Every fixture stores inputs and expected outputs as raw float64 arrays.
"""
from __future__ import annotations

import ctypes
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cosmos_amd import workloads  # noqa: E402
from oracle import load_ref  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def amr_cases():
    """(tag, physics, order, input, tolerance, gamma) of the AMR fixtures."""
    spiky = workloads.quadrants_2d(10, seed=6, ny=8)
    spiky[..., 1:3] *= 0.5
    spiky[3, 4] = (1.9, 0.2, -0.3, 0.0, 1.7)
    spiky[6, 1] = (0.3, -0.4, 0.1, 0.0, 0.4)
    return [("n1_slic", 0, 2, workloads.sod_1d(40), 5.0, 1.4),
            ("n2_slic", 0, 2, workloads.quadrants_2d(12, seed=1, ny=10), 5.0, 1.4),
            ("n2_force", 0, 1, workloads.quadrants_2d(9, seed=2, ny=11), 4.0, 1.4),
            ("sr1_force", 1, 1, workloads.sr_riemann_1d(40, seed=3), 20.0, 5.0 / 3.0),
            ("sr2_slic", 1, 2, spiky, 3.0, 5.0 / 3.0)]


def amr_level2_cases():
    """(tag, physics, order, input, tolerance, gamma, final time) of the three-level AMR fixtures."""
    from cosmos_amd import scenarios

    spiky = workloads.quadrants_2d(10, seed=6, ny=8)
    spiky[..., 1:3] *= 0.5
    return [("n_slic", 0, 2, scenarios.toro1(64).W, 1.0, 1.4, 0.05),
            ("n_force", 0, 1, workloads.sod_1d(48), 0.5, 1.4, 0.08),
            ("sr_slic", 1, 2, scenarios.sr_blast(64).W, 1.0, 5.0 / 3.0, 0.1),
            ("sr_force", 1, 1, workloads.sr_riemann_1d(40, seed=3), 2.0, 5.0 / 3.0, 0.1),
            ("n2_slic", 0, 2, workloads.quadrants_2d(12, seed=1, ny=10), 3.0, 1.4, 0.02),
            ("n2_force", 0, 1, scenarios.toro1_2d(16).W, 1.0, 1.4, 0.02),
            ("sr2_slic", 1, 2, spiky, 3.0, 5.0 / 3.0, 0.02)]


def amr_level2_fixtures(ref) -> None:
    """Three-level AMR: the reference's solve{,2D}Level2AMR (Newtonian and special
    relativistic, FORCE and SLIC, 1D and 2D): all three levels and their flags."""
    fx = {}
    for tag, physics, order, W, tol, gamma, tf in amr_level2_cases():
        n = W.shape[-2]
        out = ref.amr_level2(W, physics, order, 1.0 / n, 0.8, tf, tol, gamma)
        fx[f"{tag}_input"] = W
        fx[f"{tag}_meta"] = np.array([physics, order, tol, gamma, tf])
        for name, v in zip(("coarse", "inter", "fine", "cflags", "iflags", "fflags"), out):
            fx[f"{tag}_{name}"] = v
    np.savez_compressed(os.path.join(OUT, "amr_level2.npz"), **fx)


def gr_amr_cases():
    """(tag, levels, order, scenario, tolerance, final time) of the black-hole AMR
    fixtures: the set-ups of Tests/BlackHoleEulerAMRTests.cpp (the black-hole
    scenario states, gamma 5/3, CFL 0.95) at small sizes and short times."""
    from cosmos_amd import scenarios as sc

    return [("l1_1d_slic", 1, 2, sc.bh_spherical_1d(40), 0.5, 0.05),
            ("l1_1d_force_spin", 1, 1, sc.bh_spinning_1d(40), 0.5, 0.05),
            ("l1_2d_slic", 1, 2, sc.bh_spherical_2d(16), 0.5, 0.03),
            ("l1_2d_force_spin", 1, 1, sc.bh_spinning_spheroidal_2d(16), 0.5, 0.03),
            ("l2_1d_slic", 2, 2, sc.bh_spherical_1d(64), 0.2, 0.08),
            ("l2_1d_force_spin", 2, 1, sc.bh_spinning_1d(40), 0.5, 0.05)]


GR_AMR_NAMES = {1: ("coarse", "fine", "cflags", "fflags"),
                2: ("coarse", "inter", "fine", "cflags", "iflags", "fflags")}


def gr_amr_fixtures(ref) -> None:
    """Black-hole AMR: the reference's BlackHoleEulerAMR::solveLevel1AMR,
    solve2DLevel1AMR and 1D solveLevel2AMR, every level and its flags."""
    fx = {}
    for tag, levels, order, S, tol, tf in gr_amr_cases():
        run = ref.gr_amr_level1 if levels == 1 else ref.gr_amr_level2
        out = run(S.W, order, S.cell_spacing, S.cfl, tf, tol, S.hole, S.gamma)
        fx[f"{tag}_input"] = S.W
        fx[f"{tag}_meta"] = np.array([levels, order, tol, tf, S.cfl, S.gamma, *S.hole])
        for name, v in zip(GR_AMR_NAMES[levels], out):
            fx[f"{tag}_{name}"] = v
    np.savez_compressed(os.path.join(OUT, "gr_amr.npz"), **fx)


GR_NAN_ONSET_STEPS = 40


def gr_nan_onset_fixture(ref) -> None:
    """The perturbed black-hole medium (2D, 64^2 on [-20M, 20M]^2, seed 1,
    gamma 4/3, SLIC + RK4 forcing, CFL 0.4, the reference's horizon excision)
    through the reference's solve loop for 40 steps: the count of cells holding
    a NaN after each step, and the final state (its NaN cells included)."""
    W, h, hole = workloads.gr_random_field_2d(64, seed=1)
    out, counts = ref.gr_run_nan_counts(W, 2, h, 0.4, hole, GR_NAN_ONSET_STEPS, gamma=4.0 / 3.0)
    np.savez_compressed(os.path.join(OUT, "gr_nan_onset.npz"), nan_counts=counts.astype(np.int32), output=out,
                        meta=np.array([64, 1, h, 0.4, 4.0 / 3.0, GR_NAN_ONSET_STEPS]))


def main() -> None:
    ref = load_ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libcosmos_ref.so is not built (make -C oracle ref)")
    if sys.argv[1:] == ["amr_level2"]:
        amr_level2_fixtures(ref)
        return
    if sys.argv[1:] == ["gr_amr"]:
        gr_amr_fixtures(ref)
        return
    if sys.argv[1:] == ["gr_nan_onset"]:
        gr_nan_onset_fixture(ref)
        return

    # 1. Toro test 1 (Sod), 400 cells, SLIC + van Leer, CFL 0.8, to t = 0.25 through the
    #    reference's own EulerSecondOrderSolver::solve, and the same with FORCE.
    W0 = workloads.sod_1d(400)
    fx = {"input": W0}
    D, I, P = ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    ref.lib.ref_solve.argtypes = [I, I, I, P, I, I, D, D, D, D, I, D]
    for order in (1, 2):
        a, t, n = ref.run(W0, 0, order, 1.0 / 400, 0.8, tf=0.25)
        # the driver loop must agree with the reference's own solve()
        s = W0.copy()
        ref.lib.ref_solve(0, order, 1, s.ctypes.data_as(P), 400, 1, 1.0 / 400, 0.8, 0.25, 0.0, 1, 1.4)
        assert np.array_equal(s, a), "driver loop differs from EulerSecondOrderSolver::solve"
        fx[f"order{order}_output"] = a
        fx[f"order{order}_time"] = np.array(t)
        fx[f"order{order}_steps"] = np.array(n)
    np.savez_compressed(os.path.join(OUT, "sod_1d.npz"), **fx)

    # 2. 2D Newtonian four-quadrant problem, 24x20 cells (ragged), 3 Strang steps, every
    #    limiter and both orders, CFL 0.9.
    fx = {}
    W0 = workloads.quadrants_2d(24, seed=0, ny=20)
    fx["input"] = W0
    for order in (1, 2):
        for lim in ((0, 1, 2) if order == 2 else (1,)):
            a, t, n = ref.run(W0, 0, order, 1.0 / 24, 0.9, max_steps=3, limiter=lim)
            fx[f"o{order}_l{lim}_output"] = a
            fx[f"o{order}_l{lim}_time"] = np.array(t)
    # biased SLIC (bias 0.3, van Leer) to cover the general limiter path
    a, t, n = ref.run(W0, 0, 2, 1.0 / 24, 0.9, max_steps=3, bias=0.3, limiter=1)
    fx["o2_l1_bias03_output"] = a
    np.savez_compressed(os.path.join(OUT, "newtonian_2d.npz"), **fx)

    # 3. Special relativistic: a 1D Riemann problem (200 cells, 20 steps, CFL 0.5) and a
    #    2D blast (20x20, 3 steps), gamma 5/3.
    fx = {}
    W1 = workloads.sr_riemann_1d(200, seed=3)
    fx["riemann_input"] = W1
    for order in (1, 2):
        a, t, n = ref.run(W1, 1, order, 1.0 / 200, 0.5, max_steps=20, gamma=5.0 / 3.0)
        fx[f"riemann_o{order}_output"] = a
        fx[f"riemann_o{order}_time"] = np.array(t)
    W2 = workloads.sr_blast_2d(20, seed=4, vmax=0.9)
    fx["blast_input"] = W2
    a, t, n = ref.run(W2, 1, 2, 1.0 / 20, 0.5, max_steps=3, gamma=5.0 / 3.0)
    fx["blast_o2_output"] = a
    fx["blast_o2_time"] = np.array(t)
    np.savez_compressed(os.path.join(OUT, "relativistic.npz"), **fx)

    # 4. Known answers of single calls.
    fx = {}
    r = np.concatenate([np.linspace(-2.0, 4.0, 121), [0.0, 0.5, 1.0, 2.0, 1e300, -1e-300, 1e-300]])
    fx["limiter_r"] = r
    for lim in (0, 1, 2):
        for bias, tag in ((0.0, "0"), (0.3, "p3"), (-0.5, "m5")):
            fx[f"limiter_{lim}_{tag}"] = np.array([ref.limiter(x, bias, lim) for x in r])
    for phys, g in ((0, 1.4), (1, 5.0 / 3.0)):
        Ws = workloads.random_states(64, seed=10 + phys, physics=phys)
        fx[f"p{phys}_states"] = Ws
        fx[f"p{phys}_p2c"] = np.array([ref.p2c(phys, w, g) for w in Ws])
        Us = fx[f"p{phys}_p2c"]
        fx[f"p{phys}_c2p"] = np.array([ref.c2p(phys, u, g) for u in Us])
        fx[f"p{phys}_fluxx"] = np.array([ref.flux(phys, 0, u, g) for u in Us])
        fx[f"p{phys}_fluxy"] = np.array([ref.flux(phys, 1, u, g) for u in Us])
        sl = []
        hs = []
        for k in range(0, 60, 4):
            for axis in (0, 1):
                sl.append(ref.slic_flux(phys, axis, Ws[k], Ws[k + 1], Ws[k + 2], Ws[k + 3], 0.01, 0.004, 0.0, 1, g))
                hs.append(ref.half_step(phys, axis, Ws[k], Ws[k + 1], Ws[k + 2], 0.01, 0.004, 0.0, 1, k % 2, g))
        fx[f"p{phys}_slic_flux"] = np.array(sl)
        fx[f"p{phys}_half_step"] = np.array(hs)
        fx[f"p{phys}_force_flux"] = np.array(
            [ref.force_flux(phys, k % 2, Ws[k], Ws[k + 1], 0.01, 0.004, g) for k in range(0, 60, 2)])
    U = workloads.random_states(30, seed=20, physics=0)
    fx["slope_inputs"] = U
    fx["slope_out"] = np.array([ref.slope(U[k], U[k + 1], U[k + 2], b, lim)
                                for k in range(0, 27, 3) for lim in (0, 1, 2) for b in (0.0, 0.3)])
    fx["dt_scalar"] = np.array([ref.stable_dt_scalar(0.01, t, tf, it)
                                for (t, tf, it) in ((0.0, 1.0, 0), (0.5, 1.0, 5), (0.5, 1.0, 6), (0.995, 1.0, 7),
                                                    (0.0, math.inf, 3))])
    np.savez_compressed(os.path.join(OUT, "single_calls.npz"), **fx)

    # 5. General relativistic (black hole): the reference's 2D spinning accretion set-up
    #    at 16^2 (2 steps of solve2D's loop, one forcing subcycle), its 1D accretion set-up
    #    at 50 cells (FORCE, 5 steps), and known answers of the point functions.
    fx = {}
    W, h, hole = workloads.gr_accretion_2d(16, spin=True)
    a, t, n = ref.gr_run(W, 2, h, 0.95, hole, tf=0.45, max_steps=2, subcycles=1, gamma=5.0 / 3.0)
    fx.update(accretion_2d_input=W, accretion_2d_h=np.array(h), accretion_2d_hole=np.array(hole),
              accretion_2d_output=a, accretion_2d_time=np.array(t))
    W, h, hole = workloads.gr_accretion_1d(50)
    a, t, n = ref.gr_run(W, 1, h, 0.95, hole, tf=0.45, max_steps=5, subcycles=1, gamma=5.0 / 3.0)
    fx.update(accretion_1d_input=W, accretion_1d_h=np.array(h), accretion_1d_hole=np.array(hole),
              accretion_1d_output=a, accretion_1d_time=np.array(t))
    rng = np.random.default_rng(31)
    pts = np.column_stack([rng.uniform(-0.5, 1.5, 24), rng.uniform(-0.5, 1.5, 24), np.zeros(24)])
    pts[::3, 2] = rng.uniform(-0.2, 0.2, 8)
    states = workloads.random_states(24, seed=32, physics=1)
    fx["point_xyz"] = pts
    fx["point_states"] = states
    kat_hole = (0.3, 0.7, 0.5, 0.4, 0.05)
    fx["point_hole"] = np.array(kat_hole)
    for what, tag in ((0, "p2c"), (5, "sound"), (6, "metric")):
        fx[f"point_{tag}"] = np.array([ref.gr_point(what, w, *xyz, 0.01, kat_hole, 5.0 / 3.0)
                                         for w, xyz in zip(states, pts)])
    U = fx["point_p2c"]
    for what, tag in ((1, "c2p"), (2, "xflux"), (3, "yflux"), (4, "source")):
        fx[f"point_{tag}"] = np.array([ref.gr_point(what, u, *xyz, 0.01, kat_hole, 5.0 / 3.0)
                                         for u, xyz in zip(U, pts)])
    np.savez_compressed(os.path.join(OUT, "blackhole.npz"), **fx)

    # 6. Two-level AMR: the reference's solve{,2D}Level1AMR on small Newtonian and
    #    special-relativistic problems, coarse and fine states and both flag grids.
    fx = {}
    for tag, physics, order, W, tol, gamma in amr_cases():
        nx = W.shape[-2]
        c, f, cf, ff = ref.amr_level1(W, physics, order, 1.0 / nx, 0.8, 0.02, tol, gamma)
        fx.update({f"{tag}_input": W, f"{tag}_coarse": c, f"{tag}_fine": f, f"{tag}_cflags": cf,
                   f"{tag}_fflags": ff, f"{tag}_meta": np.array([physics, order, tol, gamma])})
    np.savez_compressed(os.path.join(OUT, "amr.npz"), **fx)
    amr_level2_fixtures(ref)
    gr_amr_fixtures(ref)
    gr_nan_onset_fixture(ref)
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
