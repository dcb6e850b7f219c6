#!/usr/bin/env python3
"""Benchmark of the finite-volume hot path on one MI355X.

Workload: 2D Newtonian Euler, 4096 x 4096 FP64 cells on [0, 1]^2, gamma 1.4,
SLIC + van Leer, dimensionally split X(dt/2) Y(dt) X(dt/2), CFL 0.9,
transmissive boundaries, seeded four-quadrant initial data.  A step is one
full time step: the CFL time step (max wave speed, fused into the step's
last pass) and the three sweeps.  The state is resident in HBM before timing.
Fast mode (the headline) runs each step as one fused X Y X kernel; the same
mode with the three sweeps launched separately is reported beside it.

    python bench.py [--steps 100] [--warmup 5] [--size 4096] [--no-cpu]

Prints one JSON line.  value = cell-updates/s.  hbm_fraction = cell-updates/s
x 240 B (the three unfused sweeps' compulsory traffic: five FP64 primitives
read and five written per sweep) / 8 TB/s; hbm_fraction_measured = the step
kernels' HBM bytes from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
/ step time / 8 TB/s.  roofline = the dominant kernel's algorithmic bytes per
launch (80 B per cell: the state read once and written once) over its average
launch duration, timed with HIP events on the solver's stream; cpu_baseline =
the reference's own sources (oracle/_ref, built with their OpenMP loops on)
timed on every host thread on a bounded sample, cpu_baseline_serial = the same
sources' serial build on one core.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s & %HBM peak (2D SLIC FP64 4096²); adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_SWEEP = 80.0      # 5 FP64 read + 5 FP64 written
SWEEPS_PER_STEP = 3
KIND_NAMES = ["step_head", "sweep_x", "sweep_y", "sweep_x_wavespeed", "step_fused"]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sweeps_4096_traffic.json")
VALU_FILE = os.path.join(ROOT, "profiles", "step_kernels_4096_valu.json")
# MI355X at its 2.4 GHz peak engine clock: 256 CUs x 4 SIMDs, one 64-lane VALU
# instruction (FP64 included) per SIMD every 4 cycles; FP64 FMA peak 78.6 TFLOP/s
VALU_PEAK_INSTR_PER_S = 256 * 4 * 2.4e9 / 4
FP64_PEAK_TFLOPS = 78.6


def measured_traffic(size, mode, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by the copy-kernel calibration), or None when the
    profiled workload is not this one."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
        if d["grid"] != size:
            return None
        k = d["kernels"][mode][kernel]
        return float(k["fetch_bytes"] + k["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def measured_valu(size, mode, kernel):
    """VALU wave instructions and FP64 FLOPs per launch of `kernel` from the
    committed rocprofv3 SQ counter passes, or None when the profiled workload
    is not this one."""
    try:
        with open(VALU_FILE) as f:
            d = json.load(f)
        if d["grid"] != size:
            return None
        k = d["kernels"][mode][kernel]
        return float(k["valu_wave_instructions"]), float(k["fp64_flop"])
    except (OSError, KeyError, ValueError):
        return None


def issue_roofline(size, mode, kernel, avg_ms):
    """The dominant kernel against the VALU issue rate and the FP64 peak."""
    v = measured_valu(size, mode, kernel)
    if v is None:
        return None
    instr, flop = v
    rate = instr / (avg_ms * 1e-3)
    tflops = flop / (avg_ms * 1e-3) / 1e12
    return {"valu_instr_per_s": rate, "valu_peak_instr_per_s": VALU_PEAK_INSTR_PER_S,
            "valu_frac": rate / VALU_PEAK_INSTR_PER_S, "fp64_tflops": tflops, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
            "fp64_frac": tflops / FP64_PEAK_TFLOPS,
            "source": "rocprofv3 SQ_INSTS_VALU and SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 per launch (committed), over "
                      "the live average launch time"}


LEG_VALU_FILE = os.path.join(ROOT, "profiles", "relativistic_blackhole_valu.json")


def leg_valu(leg, size, seed, steps=None, mode=None):
    """(VALU wave instructions, FP64 FLOPs) of a special-relativistic run (all
    of its steps, for `mode`) or of one black-hole step, from the committed
    rocprofv3 SQ counter pass, or None when the profiled workload differs."""
    try:
        with open(LEG_VALU_FILE) as f:
            d = json.load(f)[leg]
        if d["grid"] != size or d["seed"] != seed:
            return None
        if leg == "relativistic":
            if d["steps"] != steps:
                return None
            m = d["modes"][mode]
            return float(m["valu_wave_instructions"]), float(m["fp64_flop"])
        return float(d["per_step"]["valu_wave_instructions"]), float(d["per_step"]["fp64_flop"])
    except (OSError, KeyError, ValueError):
        return None


def leg_issue(valu, flop, ms):
    """VALU issue and FP64 rates of a whole leg over its live time."""
    rate = valu / (ms * 1e-3)
    tflops = flop / (ms * 1e-3) / 1e12
    return {"valu_instr_per_s": rate, "valu_frac": rate / VALU_PEAK_INSTR_PER_S, "fp64_tflops": tflops,
            "fp64_frac": tflops / FP64_PEAK_TFLOPS,
            "source": "rocprofv3 SQ_INSTS_VALU and SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 of the same run (committed), "
                      "over the live time"}


def headline_roofline(size, mode, g):
    """The dominant kernel's roofline.  A sweep kernel is charged its
    algorithmic 80 B per cell (5 FP64 read + 5 written).  The fused step kernel
    performs a whole step (three sweeps) reading and writing the state once:
    its achieved bandwidth is the rocprofv3-measured HBM traffic over its time,
    and the 240 B-per-cell-update basis of the three unfused sweeps is reported
    beside it (that figure exceeds 1 because of the fusion, not because more
    bytes move than the peak allows).  The bound is whichever of HBM and VALU
    issue the kernel is closer to."""
    kern, avg_ms = g["dominant"], g["dominant_avg_ms"]
    cells = float(size) * size
    fused = kern == "step_fused"
    alg = BYTES_PER_CELL_SWEEP * cells * (SWEEPS_PER_STEP if fused else 1)
    alg_gbs = alg / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(size, mode, kern)
    meas_gbs = None if traffic is None else traffic / (avg_ms * 1e-3) / 1e9
    issue = issue_roofline(size, mode, kern, avg_ms)
    achieved = meas_gbs if (fused and meas_gbs is not None) else alg_gbs
    hbm_frac = achieved / HBM_PEAK_GBS
    bound = "hbm"
    if issue is not None and issue["valu_frac"] > hbm_frac:
        bound = "valu_issue"
    r = {"bound": bound, "kernel": kern, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": hbm_frac, "traffic": traffic,
         "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
         "achieved_basis": "measured HBM traffic" if achieved is meas_gbs else "algorithmic bytes (80 B per cell)",
         "algorithmic_bytes": alg, "issue": issue, "avg_launch_ms": avg_ms}
    if fused:
        r["algorithmic_frac_240B_basis"] = alg_gbs / HBM_PEAK_GBS
        r["algorithmic_note"] = ("240 B per cell-update (three sweeps x 5 FP64 read + 5 written) over the fused "
                                 "kernel's time; above 1 because one launch does the three sweeps' work with a "
                                 "single read and write of the state")
    return r


def measured_step_traffic(size, mode, kernels, steps):
    """HBM bytes per step of the step's sweep kernels (the dt head's few bytes
    aside) from the committed counter passes, or None if one is missing."""
    total = 0.0
    for name, k in kernels.items():
        if name == "step_head":
            continue
        b = measured_traffic(size, mode, name)
        if b is None:
            return None
        total += b * k["launches"] / steps
    return total


class HipEvents:
    """hipEvent_t timing through the HIP runtime (events on a given stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def new(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return float(ms.value)


def gpu_leg(args, mode):
    """mode: "exact", "fast" (fused step kernel) or "fast_sweeps" (fast mode,
    three sweep launches per step)."""
    from cosmos_amd import EXACT, FAST, NEWTONIAN, VANLEER, Solver, host_empty, native, workloads

    n = args.size
    W = workloads.quadrants_2d(n, seed=args.seed)
    cells = n * n
    dx = 1.0 / n
    ev = HipEvents()
    with Solver(NEWTONIAN, 2, n, n, 1.4) as s:
        s.set_mode(EXACT if mode == "exact" else FAST)
        s.set_step_fusion(mode == "fast")
        # host transfers, each path timed on its second use (the first allocates
        # the device and page-locked staging buffers): pageable numpy arrays
        # (staged through page-locked chunks) and page-locked arrays (host_empty,
        # what the drop-in classes pack into), copied by DMA directly
        H = host_empty(W.shape)
        H[...] = W
        Hout = host_empty(W.shape)
        pageable_out = np.empty_like(W)
        xfer = {}
        for name, src, dst in (("pageable", W, pageable_out), ("page_locked", H, Hout)):
            for rep in range(2):
                t0 = time.perf_counter()
                s.upload(src)
                t1 = time.perf_counter()
                s.download(dst)
                t2 = time.perf_counter()
            xfer[name] = {"upload_s": t1 - t0, "download_s": t2 - t1,
                          "upload_gbs": W.nbytes / (t1 - t0) / 1e9, "download_gbs": W.nbytes / (t2 - t1) / 1e9}
        assert np.array_equal(Hout, W) and np.array_equal(pageable_out, W)
        transfer_s = xfer["page_locked"]["upload_s"] + xfer["page_locked"]["download_s"]
        transfer_pageable_s = xfer["pageable"]["upload_s"] + xfer["pageable"]["download_s"]
        s.upload(H)
        del W, H, Hout, pageable_out
        stream = s.stream

        # timed region: K full steps after W warm-up steps (device-resident loop)
        s.begin(2, dx, 0.9, math.inf, args.warmup + args.steps, 0.0, VANLEER)
        s.advance(args.warmup)
        s.synchronize()
        e0, e1 = ev.new(), ev.new()
        t0 = time.perf_counter()
        ev.record(e0, stream)
        s.advance(args.steps)
        ev.record(e1, stream)
        s.synchronize()
        wall = time.perf_counter() - t0
        ms_total = ev.elapsed_ms(e0, e1)
        t_reached, steps_done, _ = s.progress()
        assert steps_done == args.warmup + args.steps, steps_done
        out = s.download()
        state_ok = bool(np.isfinite(out[:: max(1, n // 64), :: max(1, n // 64)]).all())
        del out

        # the same steps again with every kernel bracketed by HIP events on the stream
        s.begin(2, dx, 0.9, math.inf, args.warmup + args.steps, 0.0, VANLEER)
        s.advance(args.warmup)
        s.synchronize()
        lib = native.load()
        lib.cosmos_profile_enable(s._h, 1)
        s.advance(args.steps)
        tot = (ctypes.c_double * 5)()
        cnt = (ctypes.c_int * 5)()
        native.check(lib.cosmos_profile_read(s._h, tot, cnt), s._h)
        lib.cosmos_profile_enable(s._h, 0)
        kernels = {}
        for k, name in enumerate(KIND_NAMES):
            if cnt[k]:
                kernels[name] = {"launches": cnt[k], "avg_ms": tot[k] / cnt[k], "total_ms": tot[k]}

        # achievable bandwidth: a plain copy kernel moving the first half of the
        # (now finished) state buffer onto its second half
        planes, pitch = s.state_planes()
        nbytes = 5 * n * pitch * 8
        src = planes[0]
        half = (nbytes // 2) // 16 * 16
        c0, c1 = ev.new(), ev.new()
        for _ in range(2):
            native.check(lib.cosmos_device_copy(ctypes.c_void_p(src + half), ctypes.c_void_p(src), half,
                                                ctypes.c_void_p(stream)), None)
        ev.record(c0, stream)
        reps = 5
        for _ in range(reps):
            native.check(lib.cosmos_device_copy(ctypes.c_void_p(src + half), ctypes.c_void_p(src), half,
                                                ctypes.c_void_p(stream)), None)
        ev.record(c1, stream)
        copy_ms = ev.elapsed_ms(c0, c1) / reps
        copy_gbs = 2.0 * half / (copy_ms * 1e-3) / 1e9

    ms_per_step = ms_total / args.steps
    cell_updates_per_s = cells / (ms_per_step * 1e-3)
    dominant = max(kernels.items(), key=lambda kv: kv[1]["total_ms"])
    dom_name, dom = dominant
    alg_bytes = BYTES_PER_CELL_SWEEP * cells
    achieved = alg_bytes / (dom["avg_ms"] * 1e-3) / 1e9
    return {
        "ms_per_step": ms_per_step,
        "wall_s": wall,
        "cell_updates_per_s": cell_updates_per_s,
        "step_hbm_gbs": cell_updates_per_s * BYTES_PER_CELL_SWEEP * SWEEPS_PER_STEP / 1e9,
        "kernels": kernels,
        "dominant": dom_name,
        "dominant_alg_bytes": alg_bytes,
        "dominant_avg_ms": dom["avg_ms"],
        "dominant_achieved_gbs": achieved,
        "copy_gbs": copy_gbs,
        "transfer_s": transfer_s,
        "transfer_pageable_s": transfer_pageable_s,
        "transfers": xfer,
        "t_reached": t_reached,
        "state_finite": state_ok,
    }


def gr_leg(args):
    """The black-hole workload: 2D accretion of a perturbed medium on [-20M, 20M]^2
    (M = 1, spin 0, gamma 4/3, excision r <= 2.5M, SLIC + van Leer with one RK4
    forcing subcycle before and after the Strang flux step, CFL 0.4), through the
    device loop of cosmos_solve_forced.  The reference's own dynamics on this
    workload reach NaN after 30-40 steps at every resolution tried (its wave
    speed is evaluated on the unit square, far from the hole, so the time step
    ignores the fast near-horizon characteristics); the measured steps stay
    before that."""
    from cosmos_amd import BLACKHOLE, Solver, native, workloads

    n = args.gr_size
    W, h, hole = workloads.gr_random_field_2d(n, seed=args.seed)
    ev = HipEvents()
    lib = native.load()
    with Solver(BLACKHOLE, 2, n, n, 4.0 / 3.0) as s:
        s.set_spacetime(*hole)
        s.set_excision(2.5)
        s.upload(W)
        stream = s.stream
        s.solve_forced(2, h, 0.4, max_steps=1, subcycles=1)   # warm-up (not timed)
        s.upload(W)
        s.synchronize()
        e0, e1 = ev.new(), ev.new()
        t0 = time.perf_counter()
        ev.record(e0, stream)
        t_reached, steps = s.solve_forced(2, h, 0.4, max_steps=args.gr_steps, subcycles=1)
        ev.record(e1, stream)
        ms_total = ev.elapsed_ms(e0, e1)
        wall = time.perf_counter() - t0
        out = s.download()
        finite = bool(np.isfinite(out).all())
        # per-kernel split: one more step with events around every launch
        lib.cosmos_profile_enable(s._h, 1)
        s.solve_forced(2, h, 0.4, max_steps=1, subcycles=1)
        tot = (ctypes.c_double * 5)()
        cnt = (ctypes.c_int * 5)()
        native.check(lib.cosmos_profile_read(s._h, tot, cnt), s._h)
        lib.cosmos_profile_enable(s._h, 0)
    names = {0: "forcing", 1: "sweep_x", 2: "sweep_y", 3: "wavespeed_and_step_head"}
    kernels = {names[k]: {"launches": cnt[k], "avg_ms": tot[k] / cnt[k]} for k in names if cnt[k]}
    ms_per_step = ms_total / steps
    v = leg_valu("blackhole", n, args.seed)
    issue = leg_issue(v[0] * steps, v[1] * steps, ms_total) if v else None
    return {
        "workload": f"2D black-hole accretion, {n}x{n} FP64 on [-20M,20M]^2, M=1, a=0, gamma 4/3, "
                    f"excision r<=2.5M, SLIC+van Leer, RK4 forcing (1 subcycle), CFL 0.4, {steps} steps, "
                    f"perturbed uniform medium (seed {args.seed})",
        "cell_updates_per_s": n * n / (ms_per_step * 1e-3),
        "ms_per_step": ms_per_step,
        "wall_s": wall,
        "t_reached": t_reached,
        "state_finite": finite,
        "kernels": kernels,
        "bound": "fp64",
        "issue": issue,
    }


def sr_leg(args):
    """The special-relativistic blast: 2048^2 FP64, gamma 5/3, SLIC + van Leer,
    CFL 0.5, seeded cylindrical blast with |v| <= 0.9, in the reference's
    con2prim mode (tolerance 1e-8, <= 1000 iterations) and in the tighter
    mode (1e-12, <= 50 iterations)."""
    from cosmos_amd import RELATIVISTIC, VANLEER, Solver, workloads

    n = args.sr_size
    W = workloads.sr_blast_2d(n, seed=args.seed, vmax=0.9)
    ev = HipEvents()
    res = {}
    for tag, (tol, maxit) in (("reference_c2p", (1e-8, 1000)), ("tight_c2p", (1e-12, 50))):
        with Solver(RELATIVISTIC, 2, n, n, 5.0 / 3.0) as s:
            s.set_con2prim(tol, maxit)
            s.upload(W)
            stream = s.stream
            s.begin(2, 1.0 / n, 0.5, math.inf, args.sr_steps, 0.0, VANLEER)
            s.synchronize()
            e0, e1 = ev.new(), ev.new()
            ev.record(e0, stream)
            s.advance(args.sr_steps)
            ev.record(e1, stream)
            ms = ev.elapsed_ms(e0, e1)
            _, steps, _ = s.progress()
        res[tag] = {"cell_updates_per_s": n * n * steps / (ms * 1e-3), "ms_per_step": ms / steps, "steps": steps}
        v = leg_valu("relativistic", n, args.seed, steps, tag)
        res[tag]["issue"] = leg_issue(v[0], v[1], ms) if v else None
    return {"workload": f"2D special-relativistic blast, {n}x{n} FP64, gamma 5/3, SLIC+van Leer, CFL 0.5, "
                        f"{args.sr_steps} steps (seed {args.seed})", "bound": "fp64", **res}


AMR_FLAG_BYTES_PER_CELL = 41.0   # 5 FP64 read + 1 flag byte per coarse cell


def amr_leg(args):
    """The AMR refinement-flag pass (EulerAMR's level-1 flag loop) on the
    headline state, Newtonian 2D, coarse args.amr_size^2 (seeded 4-quadrant,
    tolerance 5 over dx), timed launch by launch with HIP events, against
    41 B of algorithmic traffic per coarse cell; then the level-1 driver for a
    few coarse steps (SLIC, CFL 0.9) with its kernel split."""
    from cosmos_amd import NEWTONIAN, Amr, native, workloads

    n = args.amr_size
    W = workloads.quadrants_2d(n, seed=args.seed)
    dx, tol, reps = 1.0 / n, 5.0, 20
    names = ("refine", "exchange", "derefine", "merge")
    lib = native.load()

    def read(a):
        tot = (ctypes.c_double * 4)()
        cnt = (ctypes.c_int * 4)()
        native.check(lib.cosmos_amr_profile_read(a._h, tot, cnt), a._h, amr=True)
        return {names[k]: {"launches": cnt[k], "avg_ms": tot[k] / cnt[k], "total_ms": tot[k]} for k in range(4)
                if cnt[k]}

    with Amr(NEWTONIAN, 2, n, n, 1.4) as a:
        a.upload(W)
        a.refine_flags(dx, tol)   # warm-up; the flags are sticky, so every pass reads and tests the same state
        native.check(lib.cosmos_amr_profile_enable(a._h, 1), a._h, amr=True)
        for _ in range(reps):
            a.refine_flags(dx, tol)
        flags = read(a)["refine"]
        refined = int((a.download()[2] == 0).sum())
        a.upload(W)
        t0 = time.perf_counter()
        t, k = a.solve_level1(2, dx, 0.9, math.inf, tol, args.amr_steps)
        wall = time.perf_counter() - t0
        driver = read(a)
        native.check(lib.cosmos_amr_profile_enable(a._h, 0), a._h, amr=True)
    level2 = amr_level2_leg(args, lib, read, tol)
    alg = AMR_FLAG_BYTES_PER_CELL * n * n
    achieved = alg / (flags["avg_ms"] * 1e-3) / 1e9
    return {
        "workload": f"AMR level-1 refinement flags, Newtonian 2D, coarse {n}x{n} (fine {2 * n}x{2 * n}) FP64, "
                    f"seeded 4-quadrant (seed {args.seed}), tolerance {tol:g} over dx; level-1 driver "
                    f"{args.amr_steps} coarse steps, SLIC, CFL 0.9",
        "flag_pass": {"launches": flags["launches"], "avg_ms": flags["avg_ms"], "refined_coarse_cells": refined,
                      "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": alg}},
        "level1_driver": {"coarse_steps": k, "t_reached": t, "wall_s": wall, "kernels": driver,
                          "note": "merge: the masked sweeps' write-back of flagged cells; their sweep kernels are "
                                  "the level solvers' (level2_driver splits them per level)"},
        "level2_driver": level2,
    }


def amr_level2_leg(args, lib, read, tol):
    """The three-level driver (EulerAMR::solve2DLevel2AMR) on the seeded
    4-quadrant state at coarse args.amr_l2_size^2 (finest level 4x that per
    axis), SLIC, CFL 0.9, args.amr_l2_steps coarse steps, with the AMR pass
    kernels and each level's masked sweeps timed by HIP events."""
    from cosmos_amd import NEWTONIAN, Amr, native, workloads

    m = args.amr_l2_size
    W = workloads.quadrants_2d(m, seed=args.seed)
    dx = args.amr_l2_extent / m
    names = ("coarse", "intermediate", "fine")
    with Amr(NEWTONIAN, 2, m, m, 1.4) as a:
        a.upload(W)
        a.solve_level2(2, dx, 0.9, math.inf, tol, 0)  # creates the finest level; no step
        a.upload(W)
        t0 = time.perf_counter()
        t, k = a.solve_level2(2, dx, 0.9, math.inf, tol, args.amr_l2_steps)
        wall = time.perf_counter() - t0
        # the same solve again with every kernel bracketed by events for the split
        a.upload(W)
        levels = [lib.cosmos_amr_level_solver(a._h, k) for k in range(3)]
        for h in levels:
            native.check(lib.cosmos_profile_enable(h, 1), h)
        native.check(lib.cosmos_amr_profile_enable(a._h, 1), a._h, amr=True)
        read(a)
        t0 = time.perf_counter()
        a.solve_level2(2, dx, 0.9, math.inf, tol, args.amr_l2_steps)
        wall_profiled = time.perf_counter() - t0
        passes = read(a)
        sweeps = {}
        for name, h in zip(names, levels):
            tot = (ctypes.c_double * 5)()
            cnt = (ctypes.c_int * 5)()
            native.check(lib.cosmos_profile_read(h, tot, cnt), h)
            native.check(lib.cosmos_profile_enable(h, 0), h)
            sweeps[name] = {kind: {"launches": cnt[q], "avg_ms": tot[q] / cnt[q], "total_ms": tot[q]}
                            for kind, q in (("x", 1), ("y", 2)) if cnt[q]}
        native.check(lib.cosmos_amr_profile_enable(a._h, 0), a._h, amr=True)
        flags = a.download_level2()[3:]
    kernel_ms = sum(v["total_ms"] for v in passes.values()) + sum(
        v["total_ms"] for lv in sweeps.values() for v in lv.values())
    return {"workload": f"AMR three-level driver, Newtonian 2D, coarse {m}x{m} (intermediate {2 * m}x{2 * m}, "
                        f"finest {4 * m}x{4 * m}) FP64 on [0,{args.amr_l2_extent:g}]^2, seeded 4-quadrant "
                        f"(seed {args.seed}), tolerance {tol:g}, "
                        f"SLIC, CFL 0.9, {args.amr_l2_steps} coarse steps",
            "coarse_steps": k, "t_reached": t, "wall_s": wall, "wall_profiled_s": wall_profiled,
            "kernel_ms": kernel_ms,
            "flagged_cells": {name: int((f != 0).sum()) for name, f in zip(names, flags)},
            "cells": {name: int(f.size) for name, f in zip(names, flags)},
            "pass_kernels": passes, "masked_sweeps": sweeps,
            "note": "finer levels' masked sweeps sweep only the tiles that hold a flagged cell and write only "
                    "flagged cells (merge: their write-back); the coarse level's are full sweeps with the unflagged "
                    "cells restored (merge); wall_s (an unprofiled solve) includes the host round trips for each time "
                    "step and refinement test; kernel_ms and the splits come from a second, event-bracketed solve "
                    "(wall_profiled_s).  The reference's drivers sub-step a level only while its parent's time step "
                    "exceeds 1e-4 (t < t_parent + dt_parent - 1e-4); on [0,1]^2 at this resolution the coarse time "
                    "step stays below 1e-4 and no level sub-steps, so the domain is [0,4]^2 (4x the time step)"}


def batch_leg(args):
    """The batched form: args.batch_count independent Riemann problems
    (70 % Newtonian / 30 % special relativistic, half 1D in [64, 4096] cells,
    half 2D in [32, 512]^2, log-uniform sizes; SLIC + van Leer; CFL 0.9 / 0.5;
    args.batch_steps steps each) advanced in one hipGraph launch after the
    schedule autotuner's trials, against the same problems solved one by one
    through the per-problem API (cosmos_solve per problem, in turn) and with
    every problem's solve loop queued on its own stream before one wait.
    Completion times per problem come from the device wall clock."""
    from cosmos_amd import Solver, VANLEER, workloads
    from cosmos_amd.batch import Batch, cell_updates

    problems, states = workloads.riemann_batch(args.batch_count, seed=args.seed, steps=args.batch_steps)
    total = cell_updates(problems, [p["max_steps"] for p in problems])
    res = {"workload": f"{args.batch_count} Riemann problems (seed {args.seed}): 70% Newtonian (gamma 1.4, CFL 0.9) / "
                       "30% special relativistic (gamma 5/3, CFL 0.5, reference con2prim), 50% 1D [64, 4096] cells / "
                       f"50% 2D [32, 512]^2, log-uniform; SLIC + van Leer; {args.batch_steps} steps each",
           "problems": len(problems), "cells": int(sum(p["nx"] * p["ny"] for p in problems)),
           "cell_updates": total}
    with Batch(problems) as b:
        for k, W in enumerate(states):
            b.upload(k, W)
        t0 = time.perf_counter()
        res["tune_trial_ms"] = b.autotune(args.batch_tune_steps)
        res["tune_s"] = time.perf_counter() - t0
        res["schedule"] = {"newtonian": b.schedule(0), "relativistic": b.schedule(1)}
        b.run()  # builds the graph for the tuned schedule
        for k, W in enumerate(states):
            b.upload(k, W)
        t0 = time.perf_counter()
        ms = b.run_timed()
        wall = time.perf_counter() - t0
        _, steps, fin = b.stats()
    done = cell_updates(problems, steps)
    res.update({"ms": ms, "wall_s": wall, "cell_updates_per_s": done / (ms * 1e-3),
                "p50_ms": float(np.percentile(fin, 50)) * 1e-3, "p99_ms": float(np.percentile(fin, 99)) * 1e-3,
                "max_ms": float(fin.max()) * 1e-3, "all_steps_taken": bool(np.all(steps == args.batch_steps))})
    if not args.no_batch_baseline:
        solvers = []
        try:
            for p, W in zip(problems, states):
                s = Solver(p["physics"], p["dim"], p["nx"], p["ny"], p["adiabatic_index"])
                solvers.append(s)
                s.upload(W)
            t0 = time.perf_counter()
            for p, s in zip(problems, solvers):
                s.solve(2, p["cell_spacing"], p["cfl"], math.inf, p["max_steps"], 0.0, VANLEER)
            seq = time.perf_counter() - t0
            for s, W in zip(solvers, states):
                s.upload(W)
            t0 = time.perf_counter()
            for p, s in zip(problems, solvers):
                s.begin(2, p["cell_spacing"], p["cfl"], math.inf, p["max_steps"], 0.0, VANLEER)
            for p, s in zip(problems, solvers):
                s.advance(p["max_steps"])
            for s in solvers:
                s.synchronize()
            streams = time.perf_counter() - t0
        finally:
            for s in solvers:
                s.close()
        res["per_problem_sequential"] = {"s": seq, "cell_updates_per_s": total / seq}
        res["per_problem_streams"] = {"s": streams, "cell_updates_per_s": total / streams}
        res["speedup_vs_sequential"] = seq / (ms * 1e-3)
        res["speedup_vs_streams"] = streams / (ms * 1e-3)
    res["parity"] = "each problem bitwise equal to the CPU oracle's solve loop (tests/test_gpu_batch.py)"
    return res


def gr_adjoint_leg(args):
    """Adjoint of the black-hole update on the differentiable-GR workload:
    n x n on [-20M, 20M]^2 (hole at the centre, M = 1, a = 0), gamma 4/3,
    excision r < 2.5M, rho = 1 + 0.05 GRF (correlation length 2M, seed),
    p = 0.01, v = 0; SLIC + van Leer with one RK4 forcing sub-step on each side
    of the sweeps, CFL 0.4, args.gra_steps steps, a checkpoint every
    args.gra_every.  forward = the solve loop (cosmos_solve_forced) over the
    same steps; adjoint = cosmos_adjoint_backward (the reverse over the segment
    the forward pass taped; then again with every segment recomputed) for
    L = sum (rho_N - rho*)^2, rho* from seed + 1.  The finite-difference
    check runs the same workload at args.gra_fd_size with a tight con2prim
    (1e-13), central differences with h = 1e-6 and 1e-8 along 32 seeded
    directions of the initial density field."""
    from cosmos_amd import BLACKHOLE, VANLEER, Adjoint, Solver, workloads

    n, steps = args.gra_size, args.gra_steps
    W, h, hole = workloads.gr_random_field_2d(n, seed=args.seed)
    target = workloads.gr_random_field_2d(n, seed=args.seed + 1)[0][..., 0]
    with Solver(BLACKHOLE, 2, n, n, 4.0 / 3.0) as s:
        s.set_spacetime(*hole)
        s.set_excision(2.5)
        s.upload(W)
        s.synchronize()
        t0 = time.perf_counter()
        s.solve_forced(2, h, 0.4, math.inf, steps, 0.0, VANLEER, 1)
        fwd = time.perf_counter() - t0
    with Adjoint(2, n, n, 4.0 / 3.0, physics=BLACKHOLE) as a:
        a.set_spacetime(*hole)
        a.set_excision(2.5)
        a.upload(W)
        t0 = time.perf_counter()
        a.forward(2, h, 0.4, steps, 0.0, VANLEER, args.gra_every)
        rec = time.perf_counter() - t0
        WN = a.download()
        g = np.zeros_like(WN)
        g[..., 0] = 2.0 * (WN[..., 0] - target)
        t0 = time.perf_counter()
        grad = a.backward(g)
        adj = time.perf_counter() - t0
        # the same with nothing left by the forward pass: every segment recomputed
        a.release()
        t0 = time.perf_counter()
        grad2 = a.backward(g)
        adj_rc = time.perf_counter() - t0
        finite = bool(np.all(np.isfinite(WN)) and np.all(np.isfinite(grad)))
        same_grad = float(np.abs(grad2 - grad).max() / max(np.abs(grad).max(), 1e-300))

    m, fsteps = args.gra_fd_size, args.gra_fd_steps
    W, h, hole = workloads.gr_random_field_2d(m, seed=args.seed + 7)
    target = workloads.gr_random_field_2d(m, seed=args.seed + 8)[0][..., 0]
    errs = []
    with Adjoint(2, m, m, 4.0 / 3.0, physics=BLACKHOLE) as a:
        a.set_spacetime(*hole)
        a.set_excision(2.5)
        a.set_con2prim(1e-13, 200)
        a.upload(W)
        a.forward(2, h, 0.4, fsteps, 0.0, VANLEER, args.gra_every)
        WN = a.download()
        g = np.zeros_like(WN)
        g[..., 0] = 2.0 * (WN[..., 0] - target)
        gr = a.backward(g)
        rng = np.random.default_rng(args.seed + 9)
        dirs = []
        for _ in range(32):
            e = np.zeros_like(W)
            e[..., 0] = rng.standard_normal(W.shape[:-1])  # the initial density field
            dirs.append(e)
        fd_by_h = {}
        for hh in (1e-6, 1e-8):
            errs = []
            for e in dirs:
                exact = float(np.sum(gr * e))
                lp = float(np.sum((a.replay(W + hh * e)[..., 0] - target) ** 2))
                lm = float(np.sum((a.replay(W - hh * e)[..., 0] - target) ** 2))
                fd = (lp - lm) / (2 * hh)
                errs.append(abs(fd - exact) / max(abs(fd), abs(exact), 1e-300))
            errs = np.array(errs)
            fd_by_h[f"h={hh:g}"] = {"median_rel_err": float(np.median(errs)), "max_rel_err": float(errs.max()),
                                    "within_1e-6": int((errs < 1e-6).sum())}
    return {"workload": f"2D black-hole accretion, {n}x{n} FP64 on [-20M,20M]^2, M=1, a=0, gamma 4/3, excision "
                        f"r<2.5M, rho = 1 + 0.05 GRF (seed {args.seed}), p = 0.01, SLIC+van Leer, RK4 forcing "
                        f"(1 subcycle), CFL 0.4, {steps} steps, checkpoint every {args.gra_every}, dt schedule frozen",
            "forward_s": fwd, "forward_recording_s": rec, "adjoint_recomputed_s": adj_rc,
            "adjoint_forward_ratio": adj_rc / fwd,
            "adjoint_forward_ratio_definition": "reverse pass including the recomputation of its segment from the "
                                                "checkpoint, over the forward solve of the same steps (the per-step "
                                                "cost of every segment but the last of a checkpointed run)",
            "adjoint_taped_segment_s": adj, "adjoint_taped_segment_forward_ratio": adj / fwd,
            "recording_plus_adjoint_forward_ratio": (rec + adj) / fwd,
            "recomputed_vs_first_gradient_max_rel_diff": same_grad,
            "note": "adjoint_taped_segment_s: the reverse pass right after the recording forward pass, which "
                    "leaves its last segment (here the only one) taped in HBM; adjoint_recomputed_s: the reverse "
                    "pass that recomputes every segment from its checkpoint (cosmos_adjoint_release first), as "
                    "every segment but the last does in a longer run",
            "state_and_gradient_finite": finite,
            "fd_check": {"grid": [m, m], "steps": fsteps, "con2prim_tolerance": 1e-13, "directions": 32,
                         "direction": "seeded normal perturbations of the initial density field",
                         **fd_by_h,
                         "note": "differences shrink with h: the reference's slope clamp (|delta| < 1e-5 -> +1e-5, "
                                 "sign dropped) and van Leer kink make the forward map piecewise; central "
                                 "differences straddle those pieces at h = 1e-6"}}


def adjoint_leg(args):
    """Adjoint of the Newtonian SLIC update: the 4-quadrant workload at
    n x n (CFL 0.9, van Leer), args.adj_steps steps with a checkpoint every
    20.  forward = the device solve loop (exact arithmetic) over the same
    steps; adjoint = cosmos_adjoint_backward (segment recompute, except the
    segment the forward pass left in the buffers, + reverse sweeps), and the
    same with every segment recomputed.  Gradient agreement of L = sum (rho_N - rho*)^2 against central
    differences along 32 seeded directions, dt schedule held fixed."""
    from cosmos_amd import NEWTONIAN, VANLEER, Adjoint, Solver, workloads

    n, steps = args.adj_size, args.adj_steps
    W = workloads.quadrants_2d(n, seed=args.seed)
    ev = HipEvents()
    with Solver(NEWTONIAN, 2, n, n, 1.4) as s:
        s.upload(W)
        s.begin(2, 1.0 / n, 0.9, math.inf, steps, 0.0, VANLEER)
        s.synchronize()
        e0, e1 = ev.new(), ev.new()
        ev.record(e0, s.stream)
        s.advance(steps)
        ev.record(e1, s.stream)
        fwd_ms = ev.elapsed_ms(e0, e1)
    with Adjoint(2, n, n, 1.4) as a:
        a.upload(W)
        t0 = time.perf_counter()
        a.forward(2, 1.0 / n, 0.9, steps, 0.0, VANLEER, 20)
        rec_s = time.perf_counter() - t0
        WN = a.download()
        target = workloads.quadrants_2d(n, seed=args.seed + 1)[..., 0]
        g = np.zeros_like(WN)
        g[..., 0] = 2.0 * (WN[..., 0] - target)
        t0 = time.perf_counter()
        a.backward(g)
        bwd_s = time.perf_counter() - t0
        a.release()
        t0 = time.perf_counter()
        a.backward(g)
        bwd_rc_s = time.perf_counter() - t0
    # Gradient agreement: central differences with the dt schedule frozen, along
    # 32 seeded directions.  The SLIC limiters are piecewise smooth (kinks at
    # r = 0 and r = 1 for van Leer); a finite-difference stencil that straddles
    # a kink differs from the one-sided derivative the adjoint takes, so the
    # SLIC check uses a small cell-wise random problem over few steps (kinks
    # rarely within reach of the stencil) and the kink-free FORCE scheme is
    # checked at scale.
    def fd_check(shape, order, n_steps, h, seed):
        rng = np.random.default_rng(seed)
        W0 = np.empty(shape + (5,))
        W0[..., 0] = rng.uniform(0.5, 2.0, shape)
        W0[..., 1:3] = rng.uniform(-0.5, 0.5, shape + (2,))
        W0[..., 3] = 0.0
        W0[..., 4] = rng.uniform(0.5, 2.0, shape)
        tgt = rng.uniform(0.5, 2.0, shape)
        errs = []
        with Adjoint(2, shape[1], shape[0], 1.4) as a:
            a.upload(W0)
            a.forward(order, 1.0 / shape[1], 0.5, n_steps, 0.0, VANLEER, 5)
            WN = a.download()
            g = np.zeros_like(WN)
            g[..., 0] = 2.0 * (WN[..., 0] - tgt)
            grad = a.backward(g)

            def loss(X):
                r = a.replay(X)[..., 0] - tgt
                return float(np.sum(r * r))

            for _ in range(32):
                e = rng.standard_normal(W0.shape)
                exact = float(np.sum(grad * e))
                fd = (loss(W0 + h * e) - loss(W0 - h * e)) / (2 * h)
                errs.append(abs(fd - exact) / max(abs(exact), 1e-300))
        errs = np.array(errs)
        return {"grid": list(shape), "order": order, "steps": n_steps, "h": h, "directions": 32,
                "median_rel_err": float(np.median(errs)), "max_rel_err": float(errs.max()),
                "within_1e-6": int((errs < 1e-6).sum())}

    fd = {"slic_random_24x32_5_steps": fd_check((24, 32), 2, 5, 1e-7, args.seed + 7),
          "force_random_256x256_20_steps": fd_check((256, 256), 1, 20, 1e-6, args.seed + 8)}
    return {
        "workload": f"2D Newtonian SLIC, van Leer, {n}x{n} FP64, CFL 0.9, 4-quadrant (seed {args.seed}), "
                    f"{steps} steps, checkpoint every 20, dt schedule frozen",
        "forward_ms": fwd_ms,
        "forward_recording_s": rec_s,
        "adjoint_recomputed_s": bwd_rc_s,
        "adjoint_forward_ratio": bwd_rc_s * 1e3 / fwd_ms,
        "adjoint_forward_ratio_definition": "reverse pass with every segment recomputed from its checkpoint, over "
                                            "the forward solve of the same steps",
        "adjoint_after_recording_s": bwd_s,
        "adjoint_after_recording_forward_ratio": bwd_s * 1e3 / fwd_ms,
        "note": "adjoint_after_recording_s: the reverse pass after the recording forward pass (which leaves its "
                "last segment in the segment buffers); adjoint_recomputed_s: every segment recomputed "
                "(cosmos_adjoint_release first)",
        "fd_check": fd,
    }


def cpu_leg(args):
    """The reference's own sources on a bounded sample of the same workload:
    built with their OpenMP loops on, on every host thread (cpu_baseline), and
    serial on one core (cpu_baseline_serial); the C restatement
    (oracle/liboracle.so, OpenMP) on one full-size step for comparison."""
    from cosmos_amd import workloads
    import oracle

    out = {}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    ref = oracle.load_ref()
    ref_omp = oracle.load_ref_omp()
    n = args.cpu_size
    W = workloads.quadrants_2d(n, seed=args.seed)
    if ref_omp is not None:
        used = ref_omp.set_threads(threads)
        t0 = time.perf_counter()
        _, _, steps = ref_omp.run(W, 0, 2, 1.0 / n, 0.9, max_steps=args.cpu_steps)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": n * n * steps / dt, "unit": "cell-updates/s", "cores": used, "kind": "reference",
            "sample": f"reference sources built with their own OpenMP loops (-fopenmp, {used} threads) on the "
                      f"same 4-quadrant SLIC workload at {n}x{n}, {steps} full steps, {dt:.1f} s",
        }
    if ref is not None:
        t0 = time.perf_counter()
        _, _, steps = ref.run(W, 0, 2, 1.0 / n, 0.9, max_steps=args.cpu_steps)
        dt = time.perf_counter() - t0
        out["cpu_baseline" if "cpu_baseline" not in out else "cpu_baseline_serial"] = {
            "value": n * n * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": f"reference sources (serial build) on the same 4-quadrant SLIC workload at {n}x{n}, "
                      f"{steps} full steps, {dt:.1f} s",
        }
        if not args.no_extra:
            # the same reference sources on bounded samples of the SR and black-hole workloads
            m = 128
            Wsr = workloads.sr_blast_2d(m, seed=args.seed, vmax=0.9)
            t0 = time.perf_counter()
            _, _, steps = ref.run(Wsr, 1, 2, 1.0 / m, 0.5, max_steps=2, gamma=5.0 / 3.0)
            dt = time.perf_counter() - t0
            out["cpu_baseline_relativistic"] = {
                "value": m * m * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                "sample": f"reference sources (serial) on the SR blast at {m}x{m}, {steps} steps, {dt:.1f} s"}
            m = 64
            Wgr, h, hole = workloads.gr_random_field_2d(m, seed=args.seed)
            t0 = time.perf_counter()
            _, _, steps = ref.gr_run(Wgr, 2, h, 0.4, hole, max_steps=2, subcycles=1, gamma=4.0 / 3.0)
            dt = time.perf_counter() - t0
            out["cpu_baseline_blackhole"] = {
                "value": m * m * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                "sample": f"reference sources (serial) on the black-hole workload at {m}x{m} (its own "
                          f"excision radius 2M), {steps} steps, {dt:.1f} s"}
    orc = oracle.load_oracle()
    orc.set_threads(threads)
    N = args.size
    WF = workloads.quadrants_2d(N, seed=args.seed)
    t0 = time.perf_counter()
    _, _, steps = orc.run(WF, 0, 2, 1.0 / N, 0.9, max_steps=1)
    dt = time.perf_counter() - t0
    port = {"value": N * N * steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"C restatement (OpenMP over rows) on the full {N}x{N} workload, {steps} step, {dt:.1f} s"}
    if "cpu_baseline" not in out:
        out["cpu_baseline"] = port
    else:
        out["cpu_port"] = port
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        model = platform.processor()
    out["cpu_host"] = {"model": model, "nproc": os.cpu_count(), "omp_threads": threads}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mode", choices=["fast", "exact", "fast_sweeps"], default="fast",
                    help="headline mode (the others are measured and reported beside it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-size", type=int, default=768)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--gr-size", type=int, default=1024)
    ap.add_argument("--gr-steps", type=int, default=20)
    ap.add_argument("--sr-size", type=int, default=2048)
    ap.add_argument("--sr-steps", type=int, default=200)
    ap.add_argument("--adj-size", type=int, default=1024)
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--gra-size", type=int, default=1024)
    ap.add_argument("--gra-steps", type=int, default=20)
    ap.add_argument("--gra-every", type=int, default=20)
    ap.add_argument("--gra-fd-size", type=int, default=64)
    ap.add_argument("--gra-fd-steps", type=int, default=4)
    ap.add_argument("--batch-count", type=int, default=1024)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-tune-steps", type=int, default=10)
    ap.add_argument("--no-batch-baseline", action="store_true", help="skip the per-problem launch baselines")
    ap.add_argument("--no-extra", action="store_true", help="skip the SR, black-hole, adjoint and batched legs")
    ap.add_argument("--amr-size", type=int, default=4096)
    ap.add_argument("--amr-steps", type=int, default=3)
    ap.add_argument("--amr-l2-size", type=int, default=1024)
    ap.add_argument("--amr-l2-steps", type=int, default=8)
    ap.add_argument("--amr-l2-extent", type=float, default=4.0)
    ap.add_argument("--legs", default="relativistic,blackhole,adjoint,adjoint_blackhole,batched,amr",
                    help="comma-separated further legs to run (default: all)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    modes = {m: gpu_leg(args, m) for m in ("exact", "fast", "fast_sweeps")}
    g = modes[args.mode]
    step_traffic = measured_step_traffic(args.size, args.mode, g["kernels"], args.steps)
    line = {
        "metric": METRIC,
        "value": g["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D Newtonian SLIC, van Leer, {args.size}x{args.size} FP64, gamma 1.4, CFL 0.9, "
                               f"transmissive, seeded 4-quadrant (seed {args.seed}), {args.steps} steps",
                   "grid": [args.size, args.size], "order": 2, "limiter": "van Leer", "cfl": 0.9,
                   "mode": args.mode, "parallelism": "single GPU"},
        "hbm_fraction_measured": None if step_traffic is None
        else step_traffic / (g["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "step_measured_bytes_per_cell_update": None if step_traffic is None else step_traffic / args.size ** 2,
        "algorithmic_hbm_fraction_240B_basis": g["step_hbm_gbs"] / HBM_PEAK_GBS,
        "algorithmic_note": "cell-updates/s x 240 B (three unfused sweeps' compulsory bytes) / 8 TB/s; exceeds 1 "
                            "when the step is fused, which moves fewer bytes (hbm_fraction_measured is the bandwidth "
                            "fraction)",
        "copy_kernel_gbs": g["copy_gbs"],
        "copy_kernel_frac_of_peak": g["copy_gbs"] / HBM_PEAK_GBS,
        "pcie_inclusive_cell_updates_per_s": float(args.size) ** 2 * args.steps
        / (g["ms_per_step"] * 1e-3 * args.steps + g["transfer_s"]),
        "pcie_inclusive_pageable_cell_updates_per_s": float(args.size) ** 2 * args.steps
        / (g["ms_per_step"] * 1e-3 * args.steps + g["transfer_pageable_s"]),
        "transfers": g["transfers"],
        "pcie_note": "the run plus one upload of the state from host memory and one download back (what the "
                     "drop-in classes do around a solve): page-locked host arrays (the drop-in's pack buffer) "
                     "and, _pageable, numpy arrays staged through page-locked chunks; value excludes both "
                     "(state resident in HBM)",
        "roofline": headline_roofline(args.size, args.mode, g),
        "kernels": g["kernels"],
        "modes": {m: {"cell_updates_per_s": r["cell_updates_per_s"], "ms_per_step": r["ms_per_step"],
                      "algorithmic_hbm_fraction_240B_basis": r["step_hbm_gbs"] / HBM_PEAK_GBS,
                      "roofline": headline_roofline(args.size, m, r), "kernels": r["kernels"]}
                  for m, r in modes.items()},
        "parity": {"exact": "bitwise equal to the reference arithmetic (tests/test_gpu_parity.py)",
                   "fast": "relative L-inf per conserved variable <= 1e-11 after 100 steps at 4096^2 "
                           "(tests/test_gpu_parity.py::test_fast_mode_config2_full_size_100_steps); the fused "
                           "step kernel hands conserved vectors between its sweeps and agrees with the three "
                           "fast sweeps to 1e-12 (tests/test_gpu_parity.py::"
                           "test_fused_step_matches_the_three_sweeps)"},
        "timing": "HIP events on the solver stream; per-kernel averages from a second identical region with "
                  "every launch bracketed by events",
    }
    if not args.no_extra:
        legs = {"relativistic": sr_leg, "blackhole": gr_leg, "adjoint": adjoint_leg,
                "adjoint_blackhole": gr_adjoint_leg, "batched": batch_leg, "amr": amr_leg}
        for name in args.legs.split(","):
            if name not in legs:
                ap.error(f"unknown leg {name!r} (choose from {', '.join(legs)})")
            line[name] = legs[name](args)
    if not args.no_cpu:
        line.update(cpu_leg(args))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
