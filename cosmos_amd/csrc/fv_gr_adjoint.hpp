// Reverse-mode (adjoint) building blocks of the general-relativistic update:
// vector-Jacobian products of the Kerr-Schild p2c, the Newton con2prim, the
// x / y fluxes and the source term (BlackHoleEuler/BlackHoleEulerStateVector.cpp:30-517),
// and of the per-interface and per-cell stencils built from them: the SLIC
// half step and FORCE flux about an interface
// (Solvers/BlackHoleEulerSolvers.cpp:219-307, Solvers/BlackHoleEulerFirstOrderSolver.cpp:3-99),
// the fractional sweep of the forcing stages (Solvers/BlackHoleEulerSolvers.cpp:171-217).
//
// The metric is a function of position only, so it is a constant here; the
// con2prim is differentiated through its root by the implicit-function
// theorem (f(x; c) = 0 gives dx = -(df/dc) dc / f'(x)), evaluated at the
// iterate the forward pass kept.  That is the derivative of the exact root;
// it differs from the derivative of the discrete map by terms of the size of
// the stopping tolerance.  Limiters take the branch the forward evaluation
// took (one-sided derivatives at kinks); floors (rho, p >= 1e-8) and the
// Lorentz-factor clamp pass no gradient where they are active.
//
// Everything is __host__ __device__ so that a host build can check each
// product against finite differences of the forward functions.  FP64.
#pragma once

#include "fv_adjoint_physics.hpp"
#include "fv_kernels.hpp"

namespace cfv {
namespace adjgr {

// gW += (dU/dW)^T gU for U = gr_p2c(W, G)
CFV_HD void vjp_p2c(const double* W, const Geo& G, const PhysParams& P, const double* gU, double* gW) {
    double vsq = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) vsq += (G.g[i * 3 + j] * W[1 + i]) * W[1 + j];
    const bool clamp = vsq > 1.0 - 1e-8;
    const double lf = clamp ? 1.0 / __builtin_sqrt(1.0 - 1e-8) : 1.0 / __builtin_sqrt(1.0 - vsq);
    const double rho = W[0], p = W[4];
    const double h = 1.0 + ((p / rho) * P.sr_gg1);
    const double sr = G.sdet * rho;
    const double m = (sr * h) * (lf * lf);
    // U0 = sr lf, U_k = m v_k, U4 = sdet (rho h lf - p - rho lf)
    const double gm = (gU[1] * W[1] + gU[2] * W[2]) + gU[3] * W[3];
    const double gsr = gU[0] * lf + gm * h * lf * lf;
    const double glf = gU[0] * sr + gm * 2.0 * sr * h * lf + gU[4] * G.sdet * (rho * h - rho);
    const double gh = gm * sr * lf * lf + gU[4] * G.sdet * rho * lf;
    gW[0] += gsr * G.sdet + gU[4] * G.sdet * (h * lf - lf) - gh * (p / (rho * rho)) * P.sr_gg1;
    gW[4] += -gU[4] * G.sdet + gh * P.sr_gg1 / rho;
    double gv[3] = {gU[1] * m, gU[2] * m, gU[3] * m};
    if (!clamp) {
        const double gvsq = glf * 0.5 * lf * lf * lf;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) a += (G.g[i * 3 + j] + G.g[j * 3 + i]) * W[1 + j];
            gv[i] += gvsq * a;
        }
    }
    gW[1] += gv[0];
    gW[2] += gv[1];
    gW[3] += gv[2];
}

// gU += the adjoint of con2prim at the recorded point c, from the adjoints of
// R (floored rho, v, floored p), the Lorentz factor and the enthalpy.
CFV_HD void vjp_con2prim(const C2PRec& c, double sdet, const PhysParams& P, const double* gR, double glf, double gh,
                         double* gU) {
    double grho = c.rho < 1e-8 ? 0.0 : gR[0];
    const double gp = c.p < 1e-8 ? 0.0 : gR[4];
    double gS[3], grhw = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gS[k] = gR[1 + k] / c.rhw;
        grhw -= gR[1 + k] * c.R[1 + k] / c.rhw;
    }
    // p = (rhw - D) - tau
    grhw += gp;
    double gD = -gp, gtau = -gp;
    // rhw = (rho h) lf^2
    grho += grhw * c.h * c.lf * c.lf;
    gh += grhw * c.rho * c.lf * c.lf;
    glf += grhw * 2.0 * c.rho * c.h * c.lf;
    // rho = D / lf
    gD += grho / c.lf;
    glf -= grho * c.rho / c.lf;
    // h = 1 / (dc x)
    double gdc = -gh * c.h / c.dc;
    double gx = -gh * c.h / c.x;
    // lf = 0.5 ec x (1 + sqrt(A)), A = 1 + k (1 - dc x) / (ec^2 x^2)
    const double k = P.sr_4gm1g;
    const double M = (c.ec * c.ec) * (c.x * c.x);
    const double N = 1.0 - c.dc * c.x;
    const double B = __builtin_sqrt(1.0 + k * N / M);
    const double gA = glf * 0.5 * c.ec * c.x * 0.5 / B;
    double gec = glf * 0.5 * c.x * (1.0 + B);
    gx += glf * 0.5 * c.ec * (1.0 + B);
    gdc += gA * (-k * c.x / M);
    gx += gA * k * (-c.dc / M - 2.0 * N / (M * c.x));
    gec += gA * (-2.0 * k * N / (M * c.ec));
    // x: root of f(x) = c4 x^3 (x - eta) + c2 x^2 + c1 x + c0
    const double w = -gx / c.fd;
    const double x2 = c.x * c.x, x3 = x2 * c.x;
    const double gc1 = w * c.x, gc2 = w * x2, gc4 = w * (x3 * (c.x - c.eta)), geta = w * (-c.c4 * x3);
    // c1 = -2 dc c1f, c2 = gm2g (ec^2 - 1) + 1 - dc^2 gm1g^2, c4 = ec^2 - 1, eta = 2 dc gm1g
    gdc += gc1 * (-2.0 * P.sr_c1f) + gc2 * (-2.0 * c.dc * P.sr_gm1g * P.sr_gm1g) + geta * (2.0 * P.sr_gm1g);
    gec += gc2 * (2.0 * P.sr_gm2g * c.ec) + gc4 * (2.0 * c.ec);
    // dc = D / sq, ec = (D + tau) / sq
    gD += (gdc + gec) / c.sq;
    gtau += gec / c.sq;
    const double gsq = -(gdc * c.dc + gec * c.ec) / c.sq;
    // sq = sqrt(q) (unless floored), q = (tau + D)^2 - |S|^2
    const double gq = c.q < 1e-8 ? 0.0 : gsq * 0.5 / c.sq;
    const double td = c.tau + c.D;
    gtau += gq * 2.0 * td;
    gD += gq * 2.0 * td;
#pragma unroll
    for (int q = 0; q < 3; ++q) gS[q] -= gq * 2.0 * c.S[q];
    // D, S, tau = U / sqrt(gamma)
    gU[0] += gD / sdet;
    gU[1] += gS[0] / sdet;
    gU[2] += gS[1] / sdet;
    gU[3] += gS[2] / sdet;
    gU[4] += gtau / sdet;
}

// gU += (dW/dU)^T gW for W = gr_c2p(U) (tolerance tol)
CFV_HD void vjp_c2p(const double* U, double sdet, double tol, const PhysParams& P, const double* gW, double* gU) {
    C2PRec c;
    gr_con2prim_rec(U, sdet, tol, P, c);
    vjp_con2prim(c, sdet, P, gW, 0.0, 0.0, gU);
}

// gU += (dF/dU)^T gF for F = gr_flux<AXIS>(U, G), from the con2prim record c of U
template <int AXIS>
CFV_HD void vjp_flux_rec(const C2PRec& c, const Geo& G, const PhysParams& P, const double* gF, double* gU) {
    const double rho = c.R[0], p = c.R[4], lf = c.lf, h = c.h;
    const double as = G.alpha * G.sdet;
    const double vn = c.R[1 + AXIS];
    const double vt = vn - (G.beta[AXIS] / G.alpha);
    const double rhw = (rho * h) * (lf * lf);
    // F0 = as rho lf vt; Fk = as (rhw v_k vt + [k == AXIS] p);
    // F4 = as ((rhw - p - rho lf) vt + p vn)
    double gvt = as * (gF[0] * rho * lf + gF[4] * ((rhw - p) - rho * lf));
    double grhw = as * gF[4] * vt;
    double gR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gvt += as * gF[1 + k] * rhw * c.R[1 + k];
        grhw += as * gF[1 + k] * c.R[1 + k] * vt;
        gR[1 + k] = as * gF[1 + k] * rhw * vt;
    }
    gR[1 + AXIS] += gvt + as * gF[4] * p;
    gR[4] = as * gF[1 + AXIS] + as * gF[4] * (vn - vt);
    gR[0] = as * (gF[0] - gF[4]) * lf * vt + grhw * h * lf * lf;
    const double glf = as * (gF[0] - gF[4]) * rho * vt + grhw * 2.0 * rho * h * lf;
    const double gh = grhw * rho * lf * lf;
    vjp_con2prim(c, G.sdet, P, gR, glf, gh, gU);
}

template <int AXIS>
CFV_HD void vjp_flux(const double* U, const Geo& G, const PhysParams& P, const double* gF, double* gU) {
    C2PRec c;
    gr_con2prim_rec(U, G.sdet, gr_flux_tol(P, AXIS), P, c);
    vjp_flux_rec<AXIS>(c, G, P, gF, gU);
}

// The flux of U with its con2prim record kept for the reverse pass.
template <int AXIS>
CFV_HD void flux_keep(const double* U, const Geo& G, const PhysParams& P, C2PRec& c, double* F) {
    gr_con2prim_rec(U, G.sdet, gr_flux_tol(P, AXIS), P, c);
    gr_flux_of<AXIS>(c, G, F);
}

// Two independent fluxes with their records, the Newton solves interleaved
// (gr_con2prim_rec2): bitwise flux_keep of each.
template <int AXIS>
CFV_HD void flux_keep2(const double* Ua, const Geo& Ga, C2PRec& ca, double* Fa, const double* Ub, const Geo& Gb,
                       C2PRec& cb, double* Fb, const PhysParams& P) {
    const double tol = gr_flux_tol(P, AXIS);
    gr_con2prim_rec2(Ua, Ga.sdet, tol, Ub, Gb.sdet, tol, P, ca, cb);
    gr_flux_of<AXIS>(ca, Ga, Fa);
    gr_flux_of<AXIS>(cb, Gb, Fb);
}

// gU += (dS/dU)^T gS for S = gr_source(U, SG)
CFV_HD void vjp_source(const double* U, const SourceGeo& S, const PhysParams& P, const double* gout, double* gU) {
    C2PRec c;
    gr_con2prim_rec(U, S.sdet, gr_tol(P, 1e-8), P, c);
    const double rho = c.R[0], p = c.R[4], lf = c.lf, hh = c.h;
    const double al = S.alpha;
    double u[4];
    u[0] = lf / al;
#pragma unroll
    for (int k = 0; k < 3; ++k) u[1 + k] = ((al * u[0]) * c.R[1 + k]) - (S.beta[k] * u[0]);
    double vc[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        vc[i] = ((S.gm[i * 3 + 0] * c.R[1]) + (S.gm[i * 3 + 1] * c.R[2])) + (S.gm[i * 3 + 2] * c.R[3]);
    const double rhw = (rho * hh) * (lf * lf);
    const double f = al * S.sdet;
    double gs[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) gs[k] = f * gout[k];
    // adjoints of T (row-major 4x4) and of rhw vc_j
    double gT[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) gT[k] = 0.0;
    double grv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double g = gs[1 + i];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double d = S.dg[i * 9 + j * 3 + k];
                gT[0] += g * ((0.5 * S.beta[j]) * S.beta[k]) * d;
                gT[j + 1] += g * S.beta[k] * d;
                gT[(j + 1) * 4 + (k + 1)] += g * 0.5 * d;
            }
            grv[j] += g * S.db[i * 3 + j] / al;
        }
        gT[0] -= g * al * S.da[i];
    }
    {
        const double g = gs[4];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double K = S.K[i * 3 + j];
                gT[0] += g * (S.beta[i] * S.beta[j]) * K;
                gT[i + 1] += g * (2.0 * S.beta[j]) * K;
                gT[(i + 1) * 4 + (j + 1)] += g * K;
            }
            gT[0] -= g * S.beta[i] * S.da[i];
            gT[i + 1] -= g * S.da[i];
        }
    }
    // T = rho hh u u^T + p g^{-1}
    double grhh = 0.0, gp = 0.0, gu[4] = {0.0, 0.0, 0.0, 0.0};
    const double rhh = rho * hh;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double g = gT[i * 4 + j];
            grhh += g * u[i] * u[j];
            gp += g * S.gi4[i * 4 + j];
            gu[i] += g * rhh * u[j];
            gu[j] += g * rhh * u[i];
        }
    // rhw vc_j
    double grhw = 0.0, gR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        grhw += grv[j] * vc[j];
        const double gvc = grv[j] * rhw;
#pragma unroll
        for (int q = 0; q < 3; ++q) gR[1 + q] += gvc * S.gm[j * 3 + q];
    }
    // u_k = al u0 v_k - beta_k u0, u0 = lf / al
    double gu0 = gu[0];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gu0 += gu[1 + k] * (al * c.R[1 + k] - S.beta[k]);
        gR[1 + k] += gu[1 + k] * al * u[0];
    }
    double glf = gu0 / al;
    // rhh = rho hh, rhw = rho hh lf^2
    gR[0] = grhh * hh + grhw * hh * lf * lf;
    double gh = grhh * rho + grhw * rho * lf * lf;
    glf += grhw * 2.0 * rho * hh * lf;
    gR[4] = gp;
    vjp_con2prim(c, S.sdet, P, gR, glf, gh, gU);
}

// The metric at (x, y) moved by off along the sweep axis.
template <int AXIS>
CFV_HD Geo geo_shift(const Hole& b, double x, double y, double off) {
    return AXIS == 0 ? geo_at(b, x + off, y, 0.0) : geo_at(b, x, y + off, 0.0);
}

// The limited slopes of three conserved vectors and their partials by
// dl = m - l and dr = r - m (computeSlopeVector on conserved vectors).
CFV_HD void slopes(const double* ul, const double* um, const double* ur, double bias, int limiter, double* s, double* d0,
                   double* d1) {
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] = adj::slope_d(ul[k], um[k], ur[k], bias, limiter, d0[k], d1[k]);
}

// The metric of a FORCE flux at an interface: at the interface -/+ h/2 and at it.
struct ForceGeo {
    Geo l, r, c;
};
// The metric of a half step about a centre: the centre -/+ h, the centre, and
// its faces -/+ h/2.
struct HalfGeo {
    Geo m0, m, m1, fl, fr;
};

template <int AXIS>
CFV_HD ForceGeo force_geo(const Hole& b, double x, double y, double h) {
    const double hh = h * 0.5;
    return ForceGeo{geo_shift<AXIS>(b, x, y, -hh), geo_shift<AXIS>(b, x, y, hh), geo_at(b, x, y, 0.0)};
}

template <int AXIS>
CFV_HD HalfGeo half_geo(const Hole& b, double x, double y, double h) {
    return HalfGeo{geo_shift<AXIS>(b, x, y, -h), geo_at(b, x, y, 0.0), geo_shift<AXIS>(b, x, y, h),
                   geo_shift<AXIS>(b, x, y, -h * 0.5), geo_shift<AXIS>(b, x, y, h * 0.5)};
}

// FORCE flux at an interface from the states WL (at -h/2) and WR (at +h/2)
// (gr_force): the forward and its adjoint.
template <int AXIS>
CFV_HD void force_fwd_g(const double* WL, const double* WR, const ForceGeo& G, const SweepCoef& cf, const PhysParams& P,
                        double* F) {
    double ul[5], ur[5], fl[5], fr[5], us[5], fs[5];
    gr_p2c(WL, G.l, ul, P);
    gr_p2c(WR, G.r, ur, P);
    gr_flux2<AXIS>(ul, G.l, ur, G.r, P, fl, fr);
#pragma unroll
    for (int k = 0; k < 5; ++k) us[k] = 0.5 * ((ul[k] + ur[k]) + (cf.dt_dx * (fl[k] - fr[k])));
    gr_flux<AXIS>(us, G.c, P, fs);
#pragma unroll
    for (int k = 0; k < 5; ++k) F[k] = 0.5 * ((0.5 * ((fl[k] + fr[k]) + (cf.dx_dt * (ul[k] - ur[k])))) + fs[k]);
}

template <int AXIS>
CFV_HD void vjp_force_g(const double* WL, const double* WR, const ForceGeo& G, const SweepCoef& cf,
                        const PhysParams& P, const double* gF, double* gWL, double* gWR) {
    double ul[5], ur[5], fl[5], fr[5], us[5];
    C2PRec cl, cr;
    gr_p2c(WL, G.l, ul, P);
    gr_p2c(WR, G.r, ur, P);
    flux_keep2<AXIS>(ul, G.l, cl, fl, ur, G.r, cr, fr, P);
#pragma unroll
    for (int k = 0; k < 5; ++k) us[k] = 0.5 * ((ul[k] + ur[k]) + (cf.dt_dx * (fl[k] - fr[k])));
    double gfs[5], gfl[5], gfr[5], gul[5], gur[5], gus[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        gfs[k] = 0.5 * gF[k];
        const double glf = 0.25 * gF[k];
        gfl[k] = glf;
        gfr[k] = glf;
        gul[k] = cf.dx_dt * glf;
        gur[k] = -cf.dx_dt * glf;
    }
    vjp_flux<AXIS>(us, G.c, P, gfs, gus);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        gul[k] += 0.5 * gus[k];
        gur[k] += 0.5 * gus[k];
        gfl[k] += 0.5 * cf.dt_dx * gus[k];
        gfr[k] -= 0.5 * cf.dt_dx * gus[k];
    }
    vjp_flux_rec<AXIS>(cl, G.l, P, gfl, gul);
    vjp_flux_rec<AXIS>(cr, G.r, P, gfr, gur);
    vjp_p2c(WL, G.l, P, gul, gWL);
    vjp_p2c(WR, G.r, P, gur, gWR);
}

template <int AXIS>
CFV_HD void force_fwd(const double* WL, const double* WR, double x, double y, const Hole& b, double h,
                      const SweepCoef& cf, const PhysParams& P, double* F) {
    force_fwd_g<AXIS>(WL, WR, force_geo<AXIS>(b, x, y, h), cf, P, F);
}

template <int AXIS>
CFV_HD void vjp_force(const double* WL, const double* WR, double x, double y, const Hole& b, double h,
                      const SweepCoef& cf, const PhysParams& P, const double* gF, double* gWL, double* gWR) {
    vjp_force_g<AXIS>(WL, WR, force_geo<AXIS>(b, x, y, h), cf, P, gF, gWL, gWR);
}

// SLIC half step of a middle cell (gr_half_step): the forward and its adjoint.
template <int AXIS>
CFV_HD void half_fwd_g(const double* WL, const double* WM, const double* WR, int side, const HalfGeo& G,
                       const SweepCoef& cf, double bias, int limiter, const PhysParams& P, double* out) {
    double ul[5], um[5], ur[5], s[5], d0[5], d1[5];
    gr_p2c(WL, G.m0, ul, P);
    gr_p2c(WM, G.m, um, P);
    gr_p2c(WR, G.m1, ur, P);
    slopes(ul, um, ur, bias, limiter, s, d0, d1);
    double xl[5], xr[5], fl[5], fr[5], e[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    gr_flux2<AXIS>(xl, G.fl, xr, G.fr, P, fl, fr);
#pragma unroll
    for (int k = 0; k < 5; ++k) e[k] = (side == 0 ? xl[k] : xr[k]) + (cf.hdt_dx * (fl[k] - fr[k]));
    gr_c2p(e, side == 0 ? G.fl.sdet : G.fr.sdet, P, out);
}

template <int AXIS>
CFV_HD void vjp_half_g(const double* WL, const double* WM, const double* WR, int side, const HalfGeo& G,
                       const SweepCoef& cf, double bias, int limiter, const PhysParams& P, const double* gout,
                       double* gWL, double* gWM, double* gWR) {
    double ul[5], um[5], ur[5], s[5], d0[5], d1[5];
    gr_p2c(WL, G.m0, ul, P);
    gr_p2c(WM, G.m, um, P);
    gr_p2c(WR, G.m1, ur, P);
    slopes(ul, um, ur, bias, limiter, s, d0, d1);
    double xl[5], xr[5], fl[5], fr[5], e[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    C2PRec cl, cr;
    flux_keep2<AXIS>(xl, G.fl, cl, fl, xr, G.fr, cr, fr, P);
#pragma unroll
    for (int k = 0; k < 5; ++k) e[k] = (side == 0 ? xl[k] : xr[k]) + (cf.hdt_dx * (fl[k] - fr[k]));
    double ge[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    vjp_c2p(e, side == 0 ? G.fl.sdet : G.fr.sdet, gr_tol(P, 1e-8), P, gout, ge);
    double gxl[5], gxr[5], gfl[5], gfr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        gxl[k] = side == 0 ? ge[k] : 0.0;
        gxr[k] = side == 0 ? 0.0 : ge[k];
        gfl[k] = cf.hdt_dx * ge[k];
        gfr[k] = -cf.hdt_dx * ge[k];
    }
    vjp_flux_rec<AXIS>(cl, G.fl, P, gfl, gxl);
    vjp_flux_rec<AXIS>(cr, G.fr, P, gfr, gxr);
    double gul[5], gum[5], gur[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double gs = 0.5 * (gxr[k] - gxl[k]);
        gum[k] = gxl[k] + gxr[k] + gs * (d0[k] - d1[k]);
        gul[k] = -gs * d0[k];
        gur[k] = gs * d1[k];
    }
    vjp_p2c(WL, G.m0, P, gul, gWL);
    vjp_p2c(WM, G.m, P, gum, gWM);
    vjp_p2c(WR, G.m1, P, gur, gWR);
}

template <int AXIS>
CFV_HD void half_fwd(const double* WL, const double* WM, const double* WR, double x, double y, int side, const Hole& b,
                     double h, const SweepCoef& cf, double bias, int limiter, const PhysParams& P, double* out) {
    half_fwd_g<AXIS>(WL, WM, WR, side, half_geo<AXIS>(b, x, y, h), cf, bias, limiter, P, out);
}

template <int AXIS>
CFV_HD void vjp_half(const double* WL, const double* WM, const double* WR, double x, double y, int side, const Hole& b,
                     double h, const SweepCoef& cf, double bias, int limiter, const PhysParams& P, const double* gout,
                     double* gWL, double* gWM, double* gWR) {
    vjp_half_g<AXIS>(WL, WM, WR, side, half_geo<AXIS>(b, x, y, h), cf, bias, limiter, P, gout, gWL, gWM, gWR);
}

// Fractional sweep of a forcing stage about a cell centre with
// fdt_h = fraction * dt / h (gr_fractional), the metric given at the centre
// (gc) and one cell either side along the sweep (gl, gr): the forward and its
// adjoint.
template <int AXIS>
CFV_HD void frac_fwd_g(double fdt_h, const double* WL, const double* WM, const double* WR, const Geo& gl, const Geo& gc,
                       const Geo& gr, double bias, int limiter, const PhysParams& P, double* out) {
    double ul[5], um[5], ur[5], s[5], d0[5], d1[5];
    gr_p2c(WL, gl, ul, P);
    gr_p2c(WM, gc, um, P);
    gr_p2c(WR, gr, ur, P);
    slopes(ul, um, ur, bias, limiter, s, d0, d1);
    double xl[5], xr[5], fl[5], fr[5], e[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    gr_flux2<AXIS>(xl, gl, xr, gr, P, fl, fr);
#pragma unroll
    for (int k = 0; k < 5; ++k) e[k] = um[k] + (fdt_h * (fl[k] - fr[k]));
    gr_c2p(e, gc.sdet, P, out);
}

template <int AXIS>
CFV_HD void vjp_frac_g(double fdt_h, const double* WL, const double* WM, const double* WR, const Geo& gl, const Geo& gc,
                       const Geo& gr, double bias, int limiter, const PhysParams& P, const double* gout, double* gWL,
                       double* gWM, double* gWR) {
    double ul[5], um[5], ur[5], s[5], d0[5], d1[5];
    gr_p2c(WL, gl, ul, P);
    gr_p2c(WM, gc, um, P);
    gr_p2c(WR, gr, ur, P);
    slopes(ul, um, ur, bias, limiter, s, d0, d1);
    double xl[5], xr[5], fl[5], fr[5], e[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    C2PRec cl, cr;
    flux_keep2<AXIS>(xl, gl, cl, fl, xr, gr, cr, fr, P);
#pragma unroll
    for (int k = 0; k < 5; ++k) e[k] = um[k] + (fdt_h * (fl[k] - fr[k]));
    double ge[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    vjp_c2p(e, gc.sdet, gr_tol(P, 1e-8), P, gout, ge);
    double gxl[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gxr[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gfl[5], gfr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        gfl[k] = fdt_h * ge[k];
        gfr[k] = -fdt_h * ge[k];
    }
    vjp_flux_rec<AXIS>(cl, gl, P, gfl, gxl);
    vjp_flux_rec<AXIS>(cr, gr, P, gfr, gxr);
    double gul[5], gum[5], gur[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double gs = 0.5 * (gxr[k] - gxl[k]);
        gum[k] = ge[k] + gxl[k] + gxr[k] + gs * (d0[k] - d1[k]);
        gul[k] = -gs * d0[k];
        gur[k] = gs * d1[k];
    }
    vjp_p2c(WL, gl, P, gul, gWL);
    vjp_p2c(WM, gc, P, gum, gWM);
    vjp_p2c(WR, gr, P, gur, gWR);
}

// The same about (x, y), the metric formed here.
template <int AXIS>
CFV_HD void frac_fwd(double fdt_h, const double* WL, const double* WM, const double* WR, double x, double y,
                     const Hole& b, double h, double bias, int limiter, const PhysParams& P, double* out) {
    frac_fwd_g<AXIS>(fdt_h, WL, WM, WR, geo_shift<AXIS>(b, x, y, -h), geo_at(b, x, y, 0.0), geo_shift<AXIS>(b, x, y, h),
                     bias, limiter, P, out);
}

template <int AXIS>
CFV_HD void vjp_frac(double fdt_h, const double* WL, const double* WM, const double* WR, double x, double y,
                     const Hole& b, double h, double bias, int limiter, const PhysParams& P, const double* gout,
                     double* gWL, double* gWM, double* gWR) {
    vjp_frac_g<AXIS>(fdt_h, WL, WM, WR, geo_shift<AXIS>(b, x, y, -h), geo_at(b, x, y, 0.0),
                     geo_shift<AXIS>(b, x, y, h), bias, limiter, P, gout, gWL, gWM, gWR);
}

}  // namespace adjgr
}  // namespace cfv
