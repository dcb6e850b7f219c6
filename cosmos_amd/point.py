"""Point-level statics of the reference's solver classes (one state, one
interface or one cell per call), through the C-ABI's cosmos_point_* and
cosmos_vector_* entry points of the HIP library, which evaluate them on the
host with the same physics code the kernels run.

Physics: 0 Newtonian, 1 special relativistic, 2 black hole.  Primitive states
are (rho, vx, vy, vz, p); for the black hole, ``hole`` = (mass, spin, x, y, z)
of BlackHoleSpacetime and ``position`` = the (x, y, z) the reference's method
takes.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native

LAX_FRIEDRICHS, RICHTMYER, FORCE = 0, 1, 2

_P = ctypes.POINTER(ctypes.c_double)


def _v(a, n=5):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if a.size != n:
        raise ValueError(f"expected {n} components, got {a.size}")
    return a


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def _opt(a, n):
    return None if a is None else _v(a, n)


def _check(rc):
    if rc != 0:
        msg = native.load().cosmos_point_last_error()
        raise native.CosmosError(f"cosmos error {rc}: {msg.decode() if msg else ''}")


def _call(name, *args):
    _check(getattr(native.load(), name)(*args))


# ---- vector arithmetic (Solvers/EulerFirstOrderSolver.cpp:3-24, 96-101; Solvers/EulerSolvers.cpp:163-172) ----
def vector_lax_friedrichs(ul, ur, fl, fr, dx, dt):
    ul, ur, fl, fr = (np.ascontiguousarray(a, dtype=np.float64) for a in (ul, ur, fl, fr))
    out = np.empty_like(ul)
    _call("cosmos_vector_lax_friedrichs", _p(ul), _p(ur), _p(fl), _p(fr), ul.size, dx, dt, _p(out))
    return out


def vector_richtmyer(ul, ur, fl, fr, dx, dt):
    ul, ur, fl, fr = (np.ascontiguousarray(a, dtype=np.float64) for a in (ul, ur, fl, fr))
    out = np.empty_like(ul)
    _call("cosmos_vector_richtmyer", _p(ul), _p(ur), _p(fl), _p(fr), ul.size, dx, dt, _p(out))
    return out


def vector_force(lf, ri):
    lf, ri = (np.ascontiguousarray(a, dtype=np.float64) for a in (lf, ri))
    out = np.empty_like(lf)
    _call("cosmos_vector_force", _p(lf), _p(ri), lf.size, _p(out))
    return out


def vector_force_update(u, fl, fr, dx, dt):
    u, fl, fr = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, fl, fr))
    out = np.empty_like(u)
    _call("cosmos_vector_force_update", _p(u), _p(fl), _p(fr), u.size, dx, dt, _p(out))
    return out


def vector_evolution(fraction, fl, fr, dx, dt):
    fl, fr = (np.ascontiguousarray(a, dtype=np.float64) for a in (fl, fr))
    out = np.empty_like(fl)
    _call("cosmos_vector_evolution", fraction, _p(fl), _p(fr), fl.size, dx, dt, _p(out))
    return out


def vector_slope(ul, um, ur, bias, limiter):
    ul, um, ur = (np.ascontiguousarray(a, dtype=np.float64) for a in (ul, um, ur))
    out = np.empty_like(ul)
    _call("cosmos_vector_slope", _p(ul), _p(um), _p(ur), ul.size, bias, limiter, _p(out))
    return out


def slope_limiter(steepness, bias, limiter):
    out = ctypes.c_double()
    _call("cosmos_point_slope_limiter", steepness, bias, limiter, ctypes.byref(out))
    return out.value


# ---- states ----
def conserved(physics, prim, gamma, hole=None, position=None):
    out = np.empty(5)
    _call("cosmos_point_conserved", physics, _p(_v(prim)), gamma, _p(_opt(hole, 5)), _p(_opt(position, 3)), _p(out))
    return out


def primitive(physics, cons, gamma, hole=None, position=None):
    out = np.empty(5)
    _call("cosmos_point_primitive", physics, _p(_v(cons)), gamma, _p(_opt(hole, 5)), _p(_opt(position, 3)), _p(out))
    return out


def flux(physics, axis, cons, gamma, hole=None, position=None):
    out = np.empty(5)
    _call("cosmos_point_flux", physics, axis, _p(_v(cons)), gamma, _p(_opt(hole, 5)), _p(_opt(position, 3)),
          _p(out))
    return out


# ---- solver statics ----
def interface_flux(physics, kind, axis, left, right, dx, dt, gamma, hole=None, position=None):
    out = np.empty(5)
    _call("cosmos_point_interface_flux", physics, kind, axis, _p(_v(left)), _p(_v(right)), dx, dt, gamma,
          _p(_opt(hole, 5)), _p(_opt(position, 3)), _p(out))
    return out


def slic_flux(physics, axis, ll, l, r, rr, dx, dt, bias, limiter, gamma, hole=None, position=None):
    out = np.empty(5)
    _call("cosmos_point_slic_flux", physics, axis, _p(_v(ll)), _p(_v(l)), _p(_v(r)), _p(_v(rr)), dx, dt, bias,
          limiter, gamma, _p(_opt(hole, 5)), _p(_opt(position, 3)), _p(out))
    return out


def half_step(physics, axis, l, m, r, dx, dt, bias, limiter, side, gamma, hole=None, position=None):
    out = np.empty(5)
    _call("cosmos_point_half_step", physics, axis, _p(_v(l)), _p(_v(m)), _p(_v(r)), dx, dt, bias, limiter, side,
          gamma, _p(_opt(hole, 5)), _p(_opt(position, 3)), _p(out))
    return out


def fractional_step(axis, fraction, l, m, r, dx, dt, bias, limiter, gamma, hole, position):
    out = np.empty(5)
    _call("cosmos_point_fractional_step", axis, fraction, _p(_v(l)), _p(_v(m)), _p(_v(r)), dx, dt, bias, limiter,
          gamma, _p(_v(hole)), _p(_v(position, 3)), _p(out))
    return out


def slope_vector(physics, axis, l, m, r, dx, bias, limiter, gamma, hole=None, position=None):
    out = np.empty(5)
    _call("cosmos_point_slope_vector", physics, axis, _p(_v(l)), _p(_v(m)), _p(_v(r)), dx, bias, limiter, gamma,
          _p(_opt(hole, 5)), _p(_opt(position, 3)), _p(out))
    return out


def forcing(dim, left, middle, right, top, bottom, dx, dt, bias, limiter, gamma, hole, position):
    out = np.empty(5)
    _call("cosmos_point_forcing", dim, _p(_v(left)), _p(_v(middle)), _p(_v(right)), _p(_opt(top, 5)),
          _p(_opt(bottom, 5)), dx, dt, bias, limiter, gamma, _p(_v(hole)), _p(_v(position, 3)), _p(out))
    return out
