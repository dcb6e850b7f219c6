// Point-level functions of the reference's solver classes (one state, one
// interface or one cell at a time), evaluated on the host by the same physics
// code the kernels run: fv_physics.hpp, fv_sweep_core.hpp, fv_gr.hpp and
// fv_gr_ops.hpp are __host__ __device__, and this file instantiates them for
// the host.  A caller that evaluates single states gets the kernels' bits
// without a device round trip per call (a launch and two copies cost tens of
// microseconds; one GR interface flux costs a few).  The grid-level work stays
// on the device (cosmos_sweep, cosmos_forcing, the solve loops).
//
// Every entry point takes primitive states as (rho, vx, vy, vz, p) and
// conserved vectors as the reference orders them, five doubles each.  The
// black-hole functions take the spacetime as (mass, spin, x, y, z) of
// BlackHoleSpacetime(mass, spin, x, y, z) and the evaluation point
// (x, y, z) exactly as the reference's statics take their coordinates.
#include <string>
#include <type_traits>

#include "../../include/cosmos_fv.h"
#include "fv_gr_ops.hpp"
#include "fv_sweep_core.hpp"

namespace {

using namespace cfv;
using namespace cfv::core;

thread_local std::string g_error;

int bad(const char* msg) {
    g_error = msg;
    return COSMOS_ERR_ARGUMENT;
}

// The ratios the reference forms at each use (Solvers/EulerFirstOrderSolver.cpp:3-19,
// Solvers/EulerSolvers.cpp:163-172).
SweepCoef coef_of(double dx, double dt) {
    SweepCoef c{};
    c.dt_dx = dt / dx;
    c.dx_dt = dx / dt;
    c.hdt_dx = 0.5 * (dt / dx);
    return c;
}

// The GR launch record of one point evaluation: evaluated at the call site
// (GEO 2) with z from the position, no excision (the reference's point
// functions do not excise; only its grid loops do).
GRLaunch gr_point(double gamma, const double* st, const double* pos, double h, double dt, double bias, int lim) {
    GRLaunch L{};
    L.P = make_params(gamma);
    L.hole = make_hole(st[0], st[1], st[2], st[3], st[4], 0.0);
    L.h = h;
    L.dt = dt;
    L.bias = bias;
    L.limiter = lim;
    L.coef = coef_of(h, dt);
    L.z = pos[2];
    L.dim = 2;
    return L;
}

using Newt = NewtonianT<false>;
using Rel = RelativisticT<false>;

// Dispatch on the Newtonian / special-relativistic physics and the axis.
template <class F>
int by_physics_axis(int physics, int axis, F&& f) {
    if (axis != 0 && axis != 1) return bad("axis must be 0 (x) or 1 (y)");
    if (physics == COSMOS_NEWTONIAN) {
        if (axis == 0) f(Newt{}, std::integral_constant<int, 0>{});
        else f(Newt{}, std::integral_constant<int, 1>{});
    } else if (physics == COSMOS_RELATIVISTIC) {
        if (axis == 0) f(Rel{}, std::integral_constant<int, 0>{});
        else f(Rel{}, std::integral_constant<int, 1>{});
    } else {
        return bad("physics must be COSMOS_NEWTONIAN or COSMOS_RELATIVISTIC here");
    }
    return COSMOS_OK;
}

int check_gr(const double* spacetime, const double* position) {
    if (!spacetime || !position) return bad("COSMOS_BLACKHOLE needs the spacetime and the position");
    return COSMOS_OK;
}

// p2c and flux of a primitive state (compute{X,Y}FluxVector(mp) = the flux of
// computeConservedVariableVector(mp), Euler/EulerStateVector.cpp:90-93).
template <class Phys, int AXIS>
void cons_and_flux(const double* W, const PhysParams& P, double* U, double* F) {
    Phys::p2c(W, U, P);
    Phys::template flux<AXIS>(U, F, Phys::inv_density(U), P);
}

// compute{X,Y}{LaxFriedrichs,Richtmyer,FORCE}Flux(state, state)
// (Solvers/EulerFirstOrderSolver.cpp:27-94 and the relativistic copies).
template <class Phys, int AXIS>
void interface_flux(int kind, const double* WL, const double* WR, const SweepCoef& cf, const PhysParams& P,
                    double* out) {
    double UL[5], UR[5], FL[5], FR[5];
    cons_and_flux<Phys, AXIS>(WL, P, UL, FL);
    cons_and_flux<Phys, AXIS>(WR, P, UR, FR);
    if (kind == COSMOS_FLUX_LAX_FRIEDRICHS) {
        for (int k = 0; k < 5; ++k) out[k] = lax_friedrichs_k(UL[k], UR[k], FL[k], FR[k], cf.dx_dt);
    } else if (kind == COSMOS_FLUX_RICHTMYER) {
        double us[5];
        for (int k = 0; k < 5; ++k) us[k] = richtmyer_k(UL[k], UR[k], FL[k], FR[k], cf.dt_dx);
        Phys::template flux<AXIS>(us, out, Phys::inv_density(us), P);
    } else {
        force_flux<Phys, AXIS, false>(UL, FL, UR, FR, cf, P, out);
    }
}

// The evolved face state of evolveStateByHalf{X,Y}TimeStep (state form).
template <class Phys, int AXIS>
void half_state(const double* WL, const double* WM, const double* WR, const SweepCoef& cf, double bias, int lim,
                int side, const PhysParams& P, double* out) {
    double ul[5], um[5], ur[5], a[5], b[5];
    Phys::p2c(WL, ul, P);
    Phys::p2c(WM, um, P);
    Phys::p2c(WR, ur, P);
    half_faces<Phys, AXIS, 0, false>(ul, um, ur, cf, bias, lim, P, a, b);
    const double* e = side == 0 ? a : b;
    Phys::c2p(e, out, Phys::inv_density(e), P);
}

template <int AXIS>
void gr_interface_flux(int kind, const double* WL, const double* WR, double x, double y, const GRLaunch& L,
                       double* out) {
    if (kind == COSMOS_FLUX_LAX_FRIEDRICHS) gr_force<AXIS, 2, 0>(WL, WR, x, y, 0, 0, L, L.coef, out);
    else if (kind == COSMOS_FLUX_RICHTMYER) gr_force<AXIS, 2, 1>(WL, WR, x, y, 0, 0, L, L.coef, out);
    else gr_force<AXIS, 2, 2>(WL, WR, x, y, 0, 0, L, L.coef, out);
}

// compute{X,Y}SLICFlux (Solvers/BlackHoleEulerSecondOrderSolver.cpp:3-35) at the
// interface (x, y): the half steps of the two neighbours at their centres
// x -/+ h*0.5, then the FORCE flux of the evolved states at the interface.
template <int AXIS>
void gr_slic(const double* WLL, const double* WL, const double* WR, const double* WRR, double x, double y,
             const GRLaunch& L, double* out) {
    const double hh = L.h * 0.5;
    double pr[5], pl[5];
    if (AXIS == 0) {
        gr_half_step<0, 2>(WLL, WL, WR, x - hh, y, 0, 0, 1, L, L.coef, pr);
        gr_half_step<0, 2>(WL, WR, WRR, x + hh, y, 0, 0, 0, L, L.coef, pl);
    } else {
        gr_half_step<1, 2>(WLL, WL, WR, x, y - hh, 0, 0, 1, L, L.coef, pr);
        gr_half_step<1, 2>(WL, WR, WRR, x, y + hh, 0, 0, 0, L, L.coef, pl);
    }
    gr_force<AXIS, 2>(pr, pl, x, y, 0, 0, L, L.coef, out);
}

}  // namespace

extern "C" {

const char* cosmos_point_last_error(void) { return g_error.c_str(); }

// ---- vector arithmetic -------------------------------------------------------

int cosmos_vector_lax_friedrichs(const double* left_cons, const double* right_cons, const double* left_flux,
                                 const double* right_flux, long n, double cell_spacing, double time_step,
                                 double* out) {
    if (!left_cons || !right_cons || !left_flux || !right_flux || !out || n < 0) return bad("null vector or n < 0");
    const double dx_dt = cell_spacing / time_step;
    for (long k = 0; k < n; ++k) out[k] = lax_friedrichs_k(left_cons[k], right_cons[k], left_flux[k], right_flux[k], dx_dt);
    return COSMOS_OK;
}

int cosmos_vector_richtmyer(const double* left_cons, const double* right_cons, const double* left_flux,
                            const double* right_flux, long n, double cell_spacing, double time_step, double* out) {
    if (!left_cons || !right_cons || !left_flux || !right_flux || !out || n < 0) return bad("null vector or n < 0");
    const double dt_dx = time_step / cell_spacing;
    for (long k = 0; k < n; ++k) out[k] = richtmyer_k(left_cons[k], right_cons[k], left_flux[k], right_flux[k], dt_dx);
    return COSMOS_OK;
}

int cosmos_vector_force(const double* lax_friedrichs, const double* richtmyer, long n, double* out) {
    if (!lax_friedrichs || !richtmyer || !out || n < 0) return bad("null vector or n < 0");
    for (long k = 0; k < n; ++k) out[k] = force_k(lax_friedrichs[k], richtmyer[k]);
    return COSMOS_OK;
}

int cosmos_vector_force_update(const double* cons, const double* left_flux, const double* right_flux, long n,
                               double cell_spacing, double time_step, double* out) {
    if (!cons || !left_flux || !right_flux || !out || n < 0) return bad("null vector or n < 0");
    const double dt_dx = time_step / cell_spacing;
    for (long k = 0; k < n; ++k) out[k] = force_update_k(cons[k], left_flux[k], right_flux[k], dt_dx);
    return COSMOS_OK;
}

int cosmos_vector_evolution(double step_fraction, const double* left_flux, const double* right_flux, long n,
                            double cell_spacing, double time_step, double* out) {
    if (!left_flux || !right_flux || !out || n < 0) return bad("null vector or n < 0");
    const double f = step_fraction * (time_step / cell_spacing);
    for (long k = 0; k < n; ++k) out[k] = evolution_k(f, left_flux[k], right_flux[k]);
    return COSMOS_OK;
}

int cosmos_vector_slope(const double* left, const double* middle, const double* right, long n, double bias,
                        int slope_limiter, double* out) {
    if (!left || !middle || !right || !out || n < 0) return bad("null vector or n < 0");
    for (long k = 0; k < n; ++k) out[k] = slope_component<0>(left[k], middle[k], right[k], bias, slope_limiter);
    return COSMOS_OK;
}

int cosmos_point_slope_limiter(double steepness, double bias, int slope_limiter, double* out) {
    if (!out) return bad("null output");
    *out = limiter_general(steepness, bias, slope_limiter);
    return COSMOS_OK;
}

// ---- states --------------------------------------------------------------------

int cosmos_point_conserved(int physics, const double* prim, double adiabatic_index, const double* spacetime,
                           const double* position, double* cons) {
    if (!prim || !cons) return bad("null state");
    const PhysParams P = make_params(adiabatic_index);
    if (physics == COSMOS_BLACKHOLE) {
        if (int rc = check_gr(spacetime, position)) return rc;
        const Hole b = make_hole(spacetime[0], spacetime[1], spacetime[2], spacetime[3], spacetime[4], 0.0);
        gr_p2c(prim, geo_at(b, position[0], position[1], position[2]), cons, P);
        return COSMOS_OK;
    }
    return by_physics_axis(physics, 0, [&](auto phys, auto) { decltype(phys)::p2c(prim, cons, P); });
}

int cosmos_point_primitive(int physics, const double* cons, double adiabatic_index, const double* spacetime,
                           const double* position, double* prim) {
    if (!prim || !cons) return bad("null state");
    const PhysParams P = make_params(adiabatic_index);
    if (physics == COSMOS_BLACKHOLE) {
        if (int rc = check_gr(spacetime, position)) return rc;
        const Hole b = make_hole(spacetime[0], spacetime[1], spacetime[2], spacetime[3], spacetime[4], 0.0);
        gr_c2p(cons, geo_at(b, position[0], position[1], position[2]).sdet, P, prim);
        return COSMOS_OK;
    }
    return by_physics_axis(physics, 0, [&](auto phys, auto) {
        using Phys = decltype(phys);
        Phys::c2p(cons, prim, Phys::inv_density(cons), P);
    });
}

int cosmos_point_flux(int physics, int axis, const double* cons, double adiabatic_index, const double* spacetime,
                      const double* position, double* flux) {
    if (!cons || !flux) return bad("null state");
    const PhysParams P = make_params(adiabatic_index);
    if (physics == COSMOS_BLACKHOLE) {
        if (int rc = check_gr(spacetime, position)) return rc;
        if (axis != 0 && axis != 1) return bad("axis must be 0 (x) or 1 (y)");
        const Hole b = make_hole(spacetime[0], spacetime[1], spacetime[2], spacetime[3], spacetime[4], 0.0);
        const Geo g = geo_at(b, position[0], position[1], position[2]);
        if (axis == 0) gr_flux<0>(cons, g, P, flux);
        else gr_flux<1>(cons, g, P, flux);
        return COSMOS_OK;
    }
    return by_physics_axis(physics, axis, [&](auto phys, auto ax) {
        using Phys = decltype(phys);
        Phys::template flux<decltype(ax)::value>(cons, flux, Phys::inv_density(cons), P);
    });
}

// ---- solver statics --------------------------------------------------------------

int cosmos_point_interface_flux(int physics, int kind, int axis, const double* left, const double* right,
                                double cell_spacing, double time_step, double adiabatic_index,
                                const double* spacetime, const double* position, double* out) {
    if (!left || !right || !out) return bad("null state");
    if (kind != COSMOS_FLUX_LAX_FRIEDRICHS && kind != COSMOS_FLUX_RICHTMYER && kind != COSMOS_FLUX_FORCE)
        return bad("kind must be COSMOS_FLUX_LAX_FRIEDRICHS, COSMOS_FLUX_RICHTMYER or COSMOS_FLUX_FORCE");
    if (physics == COSMOS_BLACKHOLE) {
        if (int rc = check_gr(spacetime, position)) return rc;
        if (axis != 0 && axis != 1) return bad("axis must be 0 (x) or 1 (y)");
        const GRLaunch L = gr_point(adiabatic_index, spacetime, position, cell_spacing, time_step, 0.0, 0);
        if (axis == 0) gr_interface_flux<0>(kind, left, right, position[0], position[1], L, out);
        else gr_interface_flux<1>(kind, left, right, position[0], position[1], L, out);
        return COSMOS_OK;
    }
    const PhysParams P = make_params(adiabatic_index);
    const SweepCoef cf = coef_of(cell_spacing, time_step);
    return by_physics_axis(physics, axis, [&](auto phys, auto ax) {
        interface_flux<decltype(phys), decltype(ax)::value>(kind, left, right, cf, P, out);
    });
}

int cosmos_point_slic_flux(int physics, int axis, const double* left_left, const double* left, const double* right,
                           const double* right_right, double cell_spacing, double time_step, double bias,
                           int slope_limiter, double adiabatic_index, const double* spacetime,
                           const double* position, double* out) {
    if (!left_left || !left || !right || !right_right || !out) return bad("null state");
    if (physics == COSMOS_BLACKHOLE) {
        if (int rc = check_gr(spacetime, position)) return rc;
        if (axis != 0 && axis != 1) return bad("axis must be 0 (x) or 1 (y)");
        const GRLaunch L =
            gr_point(adiabatic_index, spacetime, position, cell_spacing, time_step, bias, slope_limiter);
        if (axis == 0) gr_slic<0>(left_left, left, right, right_right, position[0], position[1], L, out);
        else gr_slic<1>(left_left, left, right, right_right, position[0], position[1], L, out);
        return COSMOS_OK;
    }
    const PhysParams P = make_params(adiabatic_index);
    const SweepCoef cf = coef_of(cell_spacing, time_step);
    // computeXSLICFlux (Solvers/EulerSecondOrderSolver.cpp:3-14): the right face
    // of the left cell and the left face of the right cell, half-step evolved,
    // then their FORCE flux
    return by_physics_axis(physics, axis, [&](auto phys, auto ax) {
        using Phys = decltype(phys);
        constexpr int A = decltype(ax)::value;
        double u0[5], u1[5], u2[5], u3[5], Q0[5], G0[5], Q1[5], G1[5], Qx[5], Gx[5];
        Phys::p2c(left_left, u0, P);
        Phys::p2c(left, u1, P);
        Phys::p2c(right, u2, P);
        Phys::p2c(right_right, u3, P);
        half_step<Phys, A, 0, false>(u0, u1, u2, cf, bias, slope_limiter, P, Qx, Gx, Q1, G1);
        half_step<Phys, A, 0, false>(u1, u2, u3, cf, bias, slope_limiter, P, Q0, G0, Qx, Gx);
        force_flux<Phys, A, false>(Q1, G1, Q0, G0, cf, P, out);
    });
}

int cosmos_point_half_step(int physics, int axis, const double* left, const double* middle, const double* right,
                           double cell_spacing, double time_step, double bias, int slope_limiter, int side,
                           double adiabatic_index, const double* spacetime, const double* position,
                           double* prim_out) {
    if (!left || !middle || !right || !prim_out) return bad("null state");
    if (side != 0 && side != 1) return bad("side must be 0 (left / top face) or 1 (right / bottom face)");
    if (physics == COSMOS_BLACKHOLE) {
        if (int rc = check_gr(spacetime, position)) return rc;
        if (axis != 0 && axis != 1) return bad("axis must be 0 (x) or 1 (y)");
        const GRLaunch L =
            gr_point(adiabatic_index, spacetime, position, cell_spacing, time_step, bias, slope_limiter);
        if (axis == 0) gr_half_step<0, 2>(left, middle, right, position[0], position[1], 0, 0, side, L, L.coef, prim_out);
        else gr_half_step<1, 2>(left, middle, right, position[0], position[1], 0, 0, side, L, L.coef, prim_out);
        return COSMOS_OK;
    }
    const PhysParams P = make_params(adiabatic_index);
    const SweepCoef cf = coef_of(cell_spacing, time_step);
    return by_physics_axis(physics, axis, [&](auto phys, auto ax) {
        half_state<decltype(phys), decltype(ax)::value>(left, middle, right, cf, bias, slope_limiter, side, P,
                                                        prim_out);
    });
}

int cosmos_point_fractional_step(int axis, double step_fraction, const double* left, const double* middle,
                                 const double* right, double cell_spacing, double time_step, double bias,
                                 int slope_limiter, double adiabatic_index, const double* spacetime,
                                 const double* position, double* prim_out) {
    if (!left || !middle || !right || !prim_out) return bad("null state");
    if (int rc = check_gr(spacetime, position)) return rc;
    if (axis != 0 && axis != 1) return bad("axis must be 0 (x) or 1 (y)");
    const GRLaunch L = gr_point(adiabatic_index, spacetime, position, cell_spacing, time_step, bias, slope_limiter);
    const double f = step_fraction * (time_step / cell_spacing);
    if (axis == 0) gr_fractional<0, 2>(f, left, middle, right, position[0], position[1], L, prim_out);
    else gr_fractional<1, 2>(f, left, middle, right, position[0], position[1], L, prim_out);
    return COSMOS_OK;
}

int cosmos_point_slope_vector(int physics, int axis, const double* left, const double* middle, const double* right,
                              double cell_spacing, double bias, int slope_limiter, double adiabatic_index,
                              const double* spacetime, const double* position, double* out) {
    if (!left || !middle || !right || !out) return bad("null state");
    if (physics == COSMOS_BLACKHOLE) {
        if (int rc = check_gr(spacetime, position)) return rc;
        if (axis != 0 && axis != 1) return bad("axis must be 0 (x) or 1 (y)");
        const GRLaunch L = gr_point(adiabatic_index, spacetime, position, cell_spacing, 1.0, bias, slope_limiter);
        double um[5];
        if (axis == 0) gr_slope_at<0, 2>(left, middle, right, position[0], position[1], 0, 0, L, um, out);
        else gr_slope_at<1, 2>(left, middle, right, position[0], position[1], 0, 0, L, um, out);
        return COSMOS_OK;
    }
    const PhysParams P = make_params(adiabatic_index);
    return by_physics_axis(physics, 0, [&](auto phys, auto) {
        using Phys = decltype(phys);
        double ul[5], um[5], ur[5];
        Phys::p2c(left, ul, P);
        Phys::p2c(middle, um, P);
        Phys::p2c(right, ur, P);
        for (int k = 0; k < 5; ++k) out[k] = slope_component<0>(ul[k], um[k], ur[k], bias, slope_limiter);
    });
}

int cosmos_point_forcing(int dim, const double* left, const double* middle, const double* right, const double* top,
                         const double* bottom, double cell_spacing, double time_step, double bias, int slope_limiter,
                         double adiabatic_index, const double* spacetime, const double* position,
                         double* cons_out) {
    if (dim != 1 && dim != 2) return bad("dim must be 1 or 2");
    if (!left || !middle || !right || !cons_out || (dim == 2 && (!top || !bottom))) return bad("null state");
    if (int rc = check_gr(spacetime, position)) return rc;
    GRLaunch L = gr_point(adiabatic_index, spacetime, position, cell_spacing, time_step, bias, slope_limiter);
    // the stage fractions of evolveConservedVariableVector{,2D}
    // (Solvers/BlackHoleEulerForcingSolver.cpp:21-53, 98-167) times dt / h
    static const double f1[3] = {0.5, 0.5, 1.0}, fx[3] = {0.25, 0.25, 0.5}, fy[3] = {0.5, 0.5, 1.0};
    for (int k = 0; k < 3; ++k) {
        L.frac_dt_h[k] = (dim == 1 ? f1[k] : fx[k]) * (time_step / cell_spacing);
        L.frac_dt_h_y[k] = fy[k] * (time_step / cell_spacing);
    }
    const double x = position[0], y = position[1], z = position[2];
    const Geo gc = geo_at(L.hole, x, y, z);
    SourceGeo sg;
    source_geo(L.hole, x, y, z, cell_spacing, sg);
    const GRStep S = gr_step_of(L);
    if (dim == 1) gr_rk4_cons<1, 2>(left, middle, right, nullptr, nullptr, x, y, gc, sg, L, S, cons_out);
    else gr_rk4_cons<2, 2>(left, middle, right, top, bottom, x, y, gc, sg, L, S, cons_out);
    return COSMOS_OK;
}

}  // extern "C"
