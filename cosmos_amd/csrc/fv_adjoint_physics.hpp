// Reverse-mode (adjoint) building blocks of the Newtonian finite-volume
// update: vector-Jacobian products of prim <-> cons, the Euler flux, the
// slope limiters (one-sided derivatives at their kinks, as the forward
// evaluation selects the branch), the SLIC half step and the FORCE flux.
//
// The adjoint differentiates the mathematical update the reference
// computes (Solvers/EulerSolvers.cpp:163-223, Solvers/EulerFirstOrderSolver.cpp:3-24,
// Solvers/SlopeLimiters.cpp:3-101, Euler/EulerStateVector.cpp:30-134).  The
// reference's prim -> cons -> prim round trips of the evolved half-step
// states are the identity in exact arithmetic and are not differentiated.
// All arithmetic is FP64.
#pragma once

#include "fv_physics.hpp"

namespace cfv {

// 1 / x for the Newtonian reverse sweep: on the GPU the reciprocal estimate
// with one Newton step (within an ulp of the quotient, fast_rcp), on the host
// (the finite-difference checks of these functions) the quotient.  That
// kernel's own re-evaluations of forward values need no bitwise agreement
// with the forward pass, only its branches up to the kinks.
CFV_HD double adj_rcp(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    return fast_rcp(x);
#else
    return 1.0 / x;
#endif
}
namespace adj {

struct Coef {
    double k;     // dt / dx
    double kinv;  // dx / dt
    double c;     // 0.5 * dt / dx (half step)
    double gm1;   // gamma - 1
    double igm1;  // 1 / (gamma - 1)
    double bias;
    int limiter;
};

// primitives (rho, u, v, w, p) -> conserved (rho, m, E)
CFV_HD void p2c(const double* W, double* U, const Coef& C) {
    const double r = W[0];
    U[0] = r;
    U[1] = r * W[1];
    U[2] = r * W[2];
    U[3] = r * W[3];
    U[4] = 0.5 * r * ((W[1] * W[1] + W[2] * W[2]) + W[3] * W[3]) + W[4] * C.igm1;
}

// gW = (dU/dW)^T gU at W
CFV_HD void vjp_p2c(const double* W, const double* gU, double* gW, const Coef& C) {
    const double r = W[0];
    const double q = 0.5 * ((W[1] * W[1] + W[2] * W[2]) + W[3] * W[3]);
    gW[0] = gU[0] + gU[1] * W[1] + gU[2] * W[2] + gU[3] * W[3] + gU[4] * q;
    gW[1] = r * (gU[1] + gU[4] * W[1]);
    gW[2] = r * (gU[2] + gU[4] * W[2]);
    gW[3] = r * (gU[3] + gU[4] * W[3]);
    gW[4] = gU[4] * C.igm1;
}

// primitives of U (rho, velocity, p) with 1/rho
CFV_HD void c2p(const double* U, double* W, double& irho, const Coef& C) {
    irho = adj_rcp(U[0]);
    W[0] = U[0];
    W[1] = U[1] * irho;
    W[2] = U[2] * irho;
    W[3] = U[3] * irho;
    W[4] = C.gm1 * (U[4] - 0.5 * U[0] * ((W[1] * W[1] + W[2] * W[2]) + W[3] * W[3]));
}

// gU = (dW/dU)^T gW, the Jacobian of c2p taken at the primitives W = c2p(U)
CFV_HD void vjp_c2p(const double* W, const double* gW, double* gU, const Coef& C) {
    const double ir = adj_rcp(W[0]);
    const double q = 0.5 * ((W[1] * W[1] + W[2] * W[2]) + W[3] * W[3]);
    const double gp = gW[4] * C.gm1;
    gU[0] = gW[0] - ir * ((gW[1] * W[1] + gW[2] * W[2]) + gW[3] * W[3]) + gp * q;
    gU[1] = gW[1] * ir - gp * W[1];
    gU[2] = gW[2] * ir - gp * W[2];
    gU[3] = gW[3] * ir - gp * W[3];
    gU[4] = gp;
}

// Euler flux along AXIS (0: x, 1: y) of conserved U
template <int AXIS>
CFV_HD void flux(const double* U, double* F, const Coef& C) {
    double W[5], ir;
    c2p(U, W, ir, C);
    const double q = W[1 + AXIS];
    F[0] = U[0] * q;
    F[1] = U[1] * q;
    F[2] = U[2] * q;
    F[3] = U[3] * q;
    F[1 + AXIS] += W[4];
    F[4] = (U[4] + W[4]) * q;
}

// gU += (dF/dU)^T gF at U
template <int AXIS>
CFV_HD void vjp_flux_acc(const double* U, const double* gF, double* gU, const Coef& C) {
    double W[5], ir;
    c2p(U, W, ir, C);
    const double r = W[0], q = W[1 + AXIS], p = W[4], E = U[4];
    // F = (r q, r q v_k + p delta, (E + p) q) over (r, v, p, E)
    double gW[5];
    gW[0] = gF[0] * q + q * ((gF[1] * W[1] + gF[2] * W[2]) + gF[3] * W[3]);
    gW[1] = gF[1] * r * q;
    gW[2] = gF[2] * r * q;
    gW[3] = gF[3] * r * q;
    gW[1 + AXIS] = gF[0] * r + r * ((gF[1] * W[1] + gF[2] * W[2]) + gF[3] * W[3]) + gF[1 + AXIS] * r * q +
                   gF[4] * (E + p);
    gW[4] = gF[1 + AXIS] + gF[4] * q;
    double g[5];
    vjp_c2p(W, gW, g, C);
    gU[0] += g[0];
    gU[1] += g[1];
    gU[2] += g[2];
    gU[3] += g[3];
    gU[4] += g[4] + gF[4] * q;  // E enters F4 directly
}

// Limiter value phi(r) and its derivative on the branch the forward selects
// (SlopeLimiters::computeSlopeLimiter, Solvers/SlopeLimiters.cpp:3-69).
// FASTDIV: quotients by the reciprocal estimate (adj_rcp), for the Newtonian
// reverse sweep; the black-hole adjoint, whose taping forward passes
// re-evaluate the forward slopes bitwise, keeps the quotients.
template <bool FASTDIV = false>
CFV_HD double limiter_d(double r, double bias, int limiter, double& dphi) {
    const double den = (1.0 - bias) + (1.0 + bias) * r;
    const double rden = FASTDIV ? adj_rcp(den) : 1.0;
    const double xi = FASTDIV ? 2.0 * rden : 2.0 / den;
    const double dxi = FASTDIV ? (-2.0 * (1.0 + bias)) * (rden * rden) : -2.0 * (1.0 + bias) / (den * den);
    dphi = 0.0;
    if (r < 0.0) return 0.0;
    if (limiter == 0) {  // SuperBee
        if (r < 0.5) {
            dphi = 2.0;
            return 2.0 * r;
        }
        if (r < 1.0) return 1.0;
        double m = r, dm = 1.0;
        if (xi < m) {
            m = xi;
            dm = dxi;
        }
        if (2.0 < m) return 2.0;
        dphi = dm;
        return m;
    }
    if (limiter == 1) {  // van Leer
        const double r1 = FASTDIV ? adj_rcp(1.0 + r) : 1.0;
        const double a = FASTDIV ? (2.0 * r) * r1 : (2.0 * r) / (1.0 + r);
        if (xi < a) {
            dphi = dxi;
            return xi;
        }
        dphi = FASTDIV ? 2.0 * (r1 * r1) : 2.0 / ((1.0 + r) * (1.0 + r));
        return a;
    }
    if (r < 1.0) {  // MinBee
        dphi = 1.0;
        return r;
    }
    if (xi < 1.0) {
        dphi = dxi;
        return xi;
    }
    return 1.0;
}

// One component of the limited slope Delta = s phi(r) (SlopeLimiters::computeSlopeVector,
// Solvers/SlopeLimiters.cpp:71-101) and its partial derivatives by dl = m - l and dr = r - m.
template <bool FASTDIV = false>
CFV_HD double slope_d(double ul, double um, double ur, double bias, int limiter, double& d_dl,
                                          double& d_dr) {
    const double dl = um - ul, dr = ur - um;
    const double s = 0.5 * (1.0 + bias) * dl + 0.5 * (1.0 - bias) * dr;
    const bool cl = __builtin_fabs(dl) < 1e-5, cr = __builtin_fabs(dr) < 1e-5;
    const double num = cl ? 1e-5 : dl, den = cr ? 1e-5 : dr;
    const double rden = FASTDIV ? adj_rcp(den) : 1.0;
    const double r = FASTDIV ? num * rden : num / den;
    double dphi;
    const double phi = limiter_d<FASTDIV>(r, bias, limiter, dphi);
    const double drdl = cl ? 0.0 : (FASTDIV ? rden : 1.0 / den);
    const double drdr = cr ? 0.0 : (FASTDIV ? -r * rden : -r / den);
    d_dl = 0.5 * (1.0 + bias) * phi + s * dphi * drdl;
    d_dr = 0.5 * (1.0 - bias) * phi + s * dphi * drdr;
    return s * phi;
}

// Half-step evolution of a cell (EulerSolvers::evolveStateByHalf{X,Y}TimeStep,
// Solvers/EulerSolvers.cpp:174-223): the evolved left (side 0) and right
// (side 1) face states, conserved.
template <int AXIS>
CFV_HD void half_step(const double* UL, const double* UM, const double* UR, const Coef& C,
                                          double* a, double* b) {
    double xl[5], xr[5], fl[5], fr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double d0, d1;
        const double D = slope_d<true>(UL[k], UM[k], UR[k], C.bias, C.limiter, d0, d1);
        xl[k] = UM[k] - 0.5 * D;
        xr[k] = UM[k] + 0.5 * D;
    }
    flux<AXIS>(xl, fl, C);
    flux<AXIS>(xr, fr, C);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double e = C.c * (fl[k] - fr[k]);
        a[k] = xl[k] + e;
        b[k] = xr[k] + e;
    }
}

// Adjoint of the half step: from the adjoints ga, gb of its two outputs,
// the contributions to the adjoints of UL, UM, UR.
template <int AXIS>
CFV_HD void vjp_half_step(const double* UL, const double* UM, const double* UR, const double* ga,
                                              const double* gb, const Coef& C, double* gL, double* gM, double* gR) {
    double xl[5], xr[5], dd0[5], dd1[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double D = slope_d<true>(UL[k], UM[k], UR[k], C.bias, C.limiter, dd0[k], dd1[k]);
        xl[k] = UM[k] - 0.5 * D;
        xr[k] = UM[k] + 0.5 * D;
    }
    double ge[5], gxl[5], gxr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        ge[k] = ga[k] + gb[k];
        gxl[k] = ga[k];
        gxr[k] = gb[k];
    }
    // e = c (F(xl) - F(xr))
    double t[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) t[k] = C.c * ge[k];
    vjp_flux_acc<AXIS>(xl, t, gxl, C);
#pragma unroll
    for (int k = 0; k < 5; ++k) t[k] = -t[k];
    vjp_flux_acc<AXIS>(xr, t, gxr, C);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double gD = 0.5 * (gxr[k] - gxl[k]);
        gM[k] = gxl[k] + gxr[k] + gD * (dd0[k] - dd1[k]);
        gL[k] = -gD * dd0[k];
        gR[k] = gD * dd1[k];
    }
}

// FORCE flux of an interface (computeFORCEFlux, Solvers/EulerFirstOrderSolver.cpp:3-24)
// is not needed forward; its adjoint from the flux adjoint phi gives the
// adjoints of the left and right states.
template <int AXIS>
CFV_HD void vjp_force(const double* UL, const double* UR, const double* phi, const Coef& C,
                                          double* gUL, double* gUR) {
    double FL[5], FR[5], us[5];
    flux<AXIS>(UL, FL, C);
    flux<AXIS>(UR, FR, C);
#pragma unroll
    for (int k = 0; k < 5; ++k) us[k] = 0.5 * ((UL[k] + UR[k]) + C.k * (FL[k] - FR[k]));
    double hphi[5], gus[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 5; ++k) hphi[k] = 0.5 * phi[k];
    vjp_flux_acc<AXIS>(us, hphi, gus, C);
    double gFL[5], gFR[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        gUL[k] = 0.5 * gus[k] + 0.5 * C.kinv * hphi[k];
        gUR[k] = 0.5 * gus[k] - 0.5 * C.kinv * hphi[k];
        gFL[k] = 0.5 * C.k * gus[k] + 0.5 * hphi[k];
        gFR[k] = -0.5 * C.k * gus[k] + 0.5 * hphi[k];
    }
    vjp_flux_acc<AXIS>(UL, gFL, gUL, C);
    vjp_flux_acc<AXIS>(UR, gFR, gUR, C);
}

}  // namespace adj
}  // namespace cfv
