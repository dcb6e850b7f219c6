// Building blocks of the general-relativistic solver shared by the gfx950
// kernels (fv_gr.hip) and the point-level entry points evaluated on the host
// (fv_point.hip): the limited slope, the half-step and fractional-step
// evolutions, the FORCE flux of one interface and the RK4 source-term chain
// of one cell (Solvers/BlackHoleEulerSolvers.cpp, BlackHoleEulerFirstOrderSolver.cpp,
// BlackHoleEulerForcingSolver.cpp).
//
// Where the position-only values come from is a template parameter GEO:
//   0  evaluated at the call site's position in the z = 0 plane (the sweeps
//      and the forcing, whose grids lie in that plane)
//   1  read from the geometry cache (L.tab) at the half-lattice point (ax, ay)
//   2  evaluated at the call site's position with z = L.z (the point-level
//      functions, which take any (x, y, z) as the reference's statics do)
#pragma once

#include "fv_kernels.hpp"

namespace cfv {

template <int GEO>
CFV_HD double z_of(const GRLaunch& L) {
    if constexpr (GEO == 2) return L.z;
    else return 0.0;
}

// The Geo at (x, y), half-lattice point (ax, ay).
template <int GEO>
CFV_HD Geo geo_pick(const GRLaunch& L, double x, double y, int ax, int ay) {
    if constexpr (GEO == 1) return geo_from_table(L.tab, L.dim, ax, ay);
    else return geo_at(L.hole, x, y, z_of<GEO>(L));
}

// position (x, y) moved by off (d half-lattice units) along the sweep axis
template <int AXIS, int GEO>
CFV_HD Geo geo_off(const GRLaunch& L, double x, double y, double off, int ax, int ay, int d) {
    return AXIS == 0 ? geo_pick<GEO>(L, x + off, y, ax + d, ay) : geo_pick<GEO>(L, x, y + off, ax, ay + d);
}

// Limited slope of the conserved vectors (SlopeLimiters::computeSlopeVector{X,Y},
// Solvers/SlopeLimiters.cpp:123-147, on top of computeSlopeVector :71-101).
CFV_HD void slope5(const double* ul, const double* um, const double* ur, double bias, int limiter, double* s) {
    if (bias == 0.0 && limiter == 1) {
#pragma unroll
        for (int k = 0; k < 5; ++k) s[k] = slope_component<1>(ul[k], um[k], ur[k], bias, limiter);
    } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) s[k] = slope_component<0>(ul[k], um[k], ur[k], bias, limiter);
    }
}

// SlopeLimiters::computeSlopeVector{X,Y} of the middle cell centred at (x, y):
// the three states' p2c at x - h, x, x + h.
template <int AXIS, int GEO>
CFV_HD void gr_slope_at(const double* WL, const double* WM, const double* WR, double x, double y, int ax, int ay,
                        const GRLaunch& L, double* um, double* s) {
    const double h = L.h;
    double ul[5], ur[5];
    {
        Geo g = geo_off<AXIS, GEO>(L, x, y, -h, ax, ay, -2);
        gr_p2c(WL, g, ul, L.P);
    }
    {
        Geo g = geo_pick<GEO>(L, x, y, ax, ay);
        gr_p2c(WM, g, um, L.P);
    }
    {
        Geo g = geo_off<AXIS, GEO>(L, x, y, h, ax, ay, 2);
        gr_p2c(WR, g, ur, L.P);
    }
    slope5(ul, um, ur, L.bias, L.limiter, s);
}

// evolveStateByHalf{X,Y}TimeStep (Solvers/BlackHoleEulerSolvers.cpp:219-307) of the
// middle cell centred at (x, y): slope from p2c at x - h, x, x + h; face fluxes
// at x -/+ h/2; the selected face evolved by 0.5 (dt/h) (F_L - F_R) and
// converted back at that face.  gr_half_evolve stops before that conversion:
// e, the evolved conserved face state, and sdet, sqrt(det gamma) at the face.
template <int AXIS, int GEO>
__host__ __device__ void gr_half_evolve(const double* WL, const double* WM, const double* WR, double x, double y,
                                        int ax, int ay, int side, const GRLaunch& L, const SweepCoef& cf, double* e,
                                        double& sdet) {
    const double h = L.h;
    double um[5], s[5];
    gr_slope_at<AXIS, GEO>(WL, WM, WR, x, y, ax, ay, L, um, s);
    double xl[5], xr[5], fl[5], fr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    const double hh = h * 0.5;
    Geo gl = geo_off<AXIS, GEO>(L, x, y, -hh, ax, ay, -1);
    Geo gr = geo_off<AXIS, GEO>(L, x, y, hh, ax, ay, 1);
    gr_flux2<AXIS>(xl, gl, xr, gr, L.P, fl, fr);
#pragma unroll
    for (int k = 0; k < 5; ++k) e[k] = (side == 0 ? xl[k] : xr[k]) + evolution_k(cf.hdt_dx, fl[k], fr[k]);
    sdet = side == 0 ? gl.sdet : gr.sdet;
}

template <int AXIS, int GEO>
__host__ __device__ void gr_half_step(const double* WL, const double* WM, const double* WR, double x, double y, int ax,
                                      int ay, int side, const GRLaunch& L, const SweepCoef& cf, double* out) {
    double e[5], sdet;
    gr_half_evolve<AXIS, GEO>(WL, WM, WR, x, y, ax, ay, side, L, cf, e, sdet);
    gr_c2p(e, sdet, L.P, out);
}

// The Lax-Friedrichs flux (KIND 0), the Richtmyer flux (1) or the FORCE flux
// (2) at the interface (x, y) (Solvers/BlackHoleEulerFirstOrderSolver.cpp:3-99):
// the states' p2c and fluxes at the interface -/+ h/2, the Richtmyer state's
// flux at the interface.
template <int AXIS, int GEO, int KIND = 2>
__host__ __device__ void gr_force(const double* WL, const double* WR, double x, double y, int ax, int ay,
                                  const GRLaunch& L, const SweepCoef& cf, double* F) {
    const double hh = L.h * 0.5;
    double ul[5], ur[5], fl[5], fr[5], us[5], fs[5];
    {
        Geo gl = geo_off<AXIS, GEO>(L, x, y, -hh, ax, ay, -1);
        Geo gr = geo_off<AXIS, GEO>(L, x, y, hh, ax, ay, 1);
        gr_p2c(WL, gl, ul, L.P);
        gr_p2c(WR, gr, ur, L.P);
        gr_flux2<AXIS>(ul, gl, ur, gr, L.P, fl, fr);
    }
    if constexpr (KIND == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) F[k] = lax_friedrichs_k(ul[k], ur[k], fl[k], fr[k], cf.dx_dt);
        return;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) us[k] = richtmyer_k(ul[k], ur[k], fl[k], fr[k], cf.dt_dx);
    {
        Geo g = geo_pick<GEO>(L, x, y, ax, ay);
        gr_flux<AXIS>(us, g, L.P, KIND == 1 ? F : fs);
    }
    if constexpr (KIND == 2) {
#pragma unroll
        for (int k = 0; k < 5; ++k) F[k] = force_k(lax_friedrichs_k(ul[k], ur[k], fl[k], fr[k], cf.dx_dt), fs[k]);
    }
}

// evolveStateByFractional{X,Y}TimeStep (Solvers/BlackHoleEulerSolvers.cpp:171-217):
// slope from p2c at x - h, x, x + h; face fluxes at x -/+ h (a whole cell,
// as in the reference); U_M + frac (dt/h) (F_L - F_R) converted back at x.
// ul, ur: the neighbours' conserved vectors at x -/+ h (formed by the caller).
template <int AXIS>
CFV_HD void gr_fractional_core(double fdt_h, const double* ul, const double* WM, const double* ur, const Geo& gl,
                               const Geo& gc, const Geo& gr, const GRLaunch& L, double* out) {
    double um[5], s[5];
    gr_p2c(WM, gc, um, L.P);
    slope5(ul, um, ur, L.bias, L.limiter, s);
    double xl[5], xr[5], fl[5], fr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    gr_flux2<AXIS>(xl, gl, xr, gr, L.P, fl, fr);
    double e[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) e[k] = um[k] + evolution_k(fdt_h, fl[k], fr[k]);
    gr_c2p(e, gc.sdet, L.P, out);
}

template <int AXIS, int GEO>
__host__ __device__ void gr_fractional(double fdt_h, const double* WL, const double* WM, const double* WR, double x,
                                       double y, const GRLaunch& L, double* out) {
    const Hole& b = L.hole;
    const double h = L.h;
    const double z = z_of<GEO>(L);
    double ul[5], ur[5];
    // the neighbours' p2c and the face fluxes share their positions x -/+ h
    const Geo gl = AXIS == 0 ? geo_at(b, x - h, y, z) : geo_at(b, x, y - h, z);
    const Geo gc = geo_at(b, x, y, z);
    const Geo gr = AXIS == 0 ? geo_at(b, x + h, y, z) : geo_at(b, x, y + h, z);
    gr_p2c(WL, gl, ul, L.P);
    gr_p2c(WR, gr, ur, L.P);
    gr_fractional_core<AXIS>(fdt_h, ul, WM, ur, gl, gc, gr, L, out);
}

// The RK4 bookkeeping of the forcing: k1 .. k4 as named registers and the
// current stage's k and fractions carried through the stage loop (a
// dynamically indexed k[4][5] or fraction table would live in scratch memory).
// Stage st's middle state is U + k/2 (st = 0, 1) or U + k (st = 2) with k the
// previous stage's; the result is U + (k1 + 2 k2 + 2 k3 + k4) / 6 associated
// as the reference writes it.
struct RK4 {
    double k1[5], k2[5], k3[5], k4[5], kc[5];
    double fx, fy, fx1, fx2, fy1, fy2;  // the current stage's fractions and the later ones
    CFV_HD void start(const double* src, const GRStep& S) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            k1[k] = S.dt * src[k];
            kc[k] = k1[k];
            k2[k] = k3[k] = k4[k] = 0.0;
        }
        fx = S.fx0;
        fy = S.fy0;
        fx1 = S.fx1;
        fx2 = S.fx2;
        fy1 = S.fy1;
        fy2 = S.fy2;
    }
    CFV_HD void middle(int st, const double* uM, double* tmp) const {
#pragma unroll
        for (int k = 0; k < 5; ++k) tmp[k] = st < 2 ? uM[k] + (0.5 * kc[k]) : uM[k] + kc[k];
    }
    // value selects on the uniform stage index (no store through a selected
    // address, which would keep these arrays in memory)
    CFV_HD void next(int st, const double* src, const GRStep& S) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            kc[k] = S.dt * src[k];
            k2[k] = st == 0 ? kc[k] : k2[k];
            k3[k] = st == 1 ? kc[k] : k3[k];
            k4[k] = st == 2 ? kc[k] : k4[k];
        }
        fx = st == 0 ? fx1 : fx2;
        fy = st == 0 ? fy1 : fy2;
    }
    CFV_HD void finish(const double* uM, double* un) const {
#pragma unroll
        for (int k = 0; k < 5; ++k) un[k] = uM[k] + ((1.0 / 6.0) * (k1[k] + ((2.0 * k2[k]) + ((2.0 * k3[k]) + k4[k]))));
    }
};

// evolveConservedVariableVector{,2D} (Solvers/BlackHoleEulerForcingSolver.cpp:3-185) at
// (x, y) evaluated at the call sites (GEO 0 or 2): the neighbours' conserved
// vectors uL, uR (uT, uB) back to primitives at x -/+ h (y -/+ h), then RK4 on
// the source term with each stage's middle state evolved by fractional sweeps
// against them (1D: x by 1/2, 1/2, 1; 2D: x 1/4 y 1/2 x 1/4, twice, then
// x 1/2 y 1 x 1/2: S.fx, S.fy), U + (k1 + 2 k2 + 2 k3 + k4) / 6 with k = S.dt times
// the source.  uM: the cell's conserved vector at (x, y); gc: the Geo there;
// sg: its SourceGeo.  un: the evolved conserved vector.
template <int DIM, int GEO>
CFV_HD void gr_rk4_cons(const double* uL, const double* uM, const double* uR, const double* uT, const double* uB,
                        double x, double y, const Geo& gc, const SourceGeo& sg, const GRLaunch& L, const GRStep& S,
                        double* un) {
    const Hole& b = L.hole;
    const double h = L.h;
    const double z = z_of<GEO>(L);
    double WLn[5], WRn[5], WTn[5], WBn[5];
    gr_c2p(uL, geo_at(b, x - h, y, z).sdet, L.P, WLn);
    gr_c2p(uR, geo_at(b, x + h, y, z).sdet, L.P, WRn);
    if (DIM == 2) {
        gr_c2p(uT, geo_at(b, x, y - h, z).sdet, L.P, WTn);
        gr_c2p(uB, geo_at(b, x, y + h, z).sdet, L.P, WBn);
    }
    double src[5];
    gr_source_of(uM, SrcRef{sg}, L.P, src);
    RK4 rk;
    rk.start(src, S);
    for (int st = 0; st < 3; ++st) {
        double tmp[5], Wm[5], e1[5], um[5];
        rk.middle(st, uM, tmp);
        gr_c2p(tmp, gc.sdet, L.P, Wm);
        if (DIM == 1) {
            gr_fractional<0, GEO>(rk.fx, WLn, Wm, WRn, x, y, L, e1);
        } else {
            double e2[5], e3[5];
            gr_fractional<0, GEO>(rk.fx, WLn, Wm, WRn, x, y, L, e2);
            gr_fractional<1, GEO>(rk.fy, WTn, e2, WBn, x, y, L, e3);
            gr_fractional<0, GEO>(rk.fx, WLn, e3, WRn, x, y, L, e1);
        }
        gr_p2c(e1, gc, um, L.P);
        gr_source_of(um, SrcRef{sg}, L.P, src);
        rk.next(st, src, S);
    }
    rk.finish(uM, un);
}

}  // namespace cfv
