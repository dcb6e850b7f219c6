// Device-side general-relativistic physics: an ideal gas on a fixed
// Kerr-Schild black-hole spacetime in 3+1 form (gfx950, FP64).
//
// Every function reproduces the reference's IEEE operation sequence
// (BlackHoleEuler/BlackHoleSpacetime.cpp, BlackHoleEuler/BlackHoleEulerStateVector.cpp,
// BlackHoleEuler/BlackHoleEulerEigenvalues.cpp, Mathematics/MatrixAlgebra.cpp), so that
// results are bitwise equal to the reference CPU path.  The reference
// re-derives the Kerr-Schild radius, scalar and null vector from scratch in
// every metric accessor (a spatial-metric derivative alone evaluates the
// radius 12 times); here each position's radius, H and l are evaluated once
// and every quantity at that position is formed from them.  That is exact:
// the same operations on the same inputs give the same bits.
//
// Reference quirks kept: the (r + (y - y_BH)) term of the Kerr-Schild
// 4-vector (BlackHoleSpacetime.cpp:108), tau formed with W rather than W^2
// in p2c (BlackHoleEulerStateVector.cpp:154), the flat |S|^2 of con2prim, the
// x-flux con2prim tolerance of 1e-10 against 1e-8 elsewhere, and the
// previous-iterate return of the Newton loop.
#pragma once

#include "fv_physics.hpp"

namespace cfv {

// BlackHoleSpacetime(M, a, x, y, z) (BlackHoleEuler/BlackHoleSpacetime.cpp:43-52)
// with the constants its accessors form, computed once on the host with the
// same expressions.
struct Hole {
    double M, a, x, y, z;
    double ma;     // (M * M) * (a * a)
    double aM;     // a * M
    double r_exc;  // M * (1.0 + sqrt(1.0 - (a * a))), the excision radius (:372-386)
};

// BlackHoleSpacetime(M, a, x, y, z) and the constants its accessors form
// (BlackHoleEuler/BlackHoleSpacetime.cpp:54-113, 372-386); a positive
// excision radius replaces the horizon radius.
inline Hole make_hole(double M, double a, double x, double y, double z, double excision) {
    Hole b{};
    b.M = M;
    b.a = a;
    b.x = x;
    b.y = y;
    b.z = z;
    b.ma = (M * M) * (a * a);
    b.aM = a * M;
    b.r_exc = excision > 0.0 ? excision : M * (1.0 + __builtin_sqrt(1.0 - (a * a)));
    return b;
}

// Kerr-Schild radius, scalar H and spatial null vector l at one point
// (computeKerrSchildScalar :54-68, computeKerrSchildVector :70-90).
struct KSPoint {
    double r, dx, dy, dz, H, l0, l1, l2;
};

CFV_HD KSPoint ks_point(const Hole& b, double x, double y, double z) {
    KSPoint k;
    k.dx = x - b.x;
    k.dy = y - b.y;
    k.dz = z - b.z;
    double ns = ((k.dx * k.dx) + (k.dy * k.dy)) + (k.dz * k.dz);
    double nm = ns - b.ma;
    double dz2 = k.dz * k.dz;
    double r = __builtin_sqrt(0.5 * (nm + __builtin_sqrt((nm * nm) + (4.0 * (b.ma * dz2)))));
    double r2 = r * r;
    double r3 = r2 * r;
    k.r = r;
    k.H = (-(b.M * r3)) / ((r3 * r) + (b.ma * dz2));
    double den = r2 + b.ma;
    double iden = 1.0 / den;
    k.l0 = div_by(-((r * k.dx) + (b.aM * k.dy)), den, iden);
    k.l1 = div_by(-((r * k.dy) - (b.aM * k.dx)), den, iden);
    k.l2 = (-k.dz) / r;
    return k;
}

// The spacetime 4-vector's second spatial component with the reference's
// (r + (y - y_BH)) term (computeKerrSchildSpacetimeVector :92-113).
CFV_HD double ks_l4y(const Hole& b, const KSPoint& k) {
    return (-((k.r + k.dy) - (b.aM * k.dx))) / ((k.r * k.r) + b.ma);
}

// Determinant of a 3x3 matrix by cofactor expansion along row 0
// (MatrixAlgebra::computeMatrixDeterminant, Mathematics/MatrixAlgebra.cpp:106-143).
CFV_HD double det3(const double* m) {
    double d = 0.0;
    d += m[0] * ((m[4] * m[8]) - (m[7] * m[5]));
    d -= m[1] * ((m[3] * m[8]) - (m[6] * m[5]));
    d += m[2] * ((m[3] * m[7]) - (m[6] * m[4]));
    return d;
}

// gamma_ij = delta_ij - 2 H l_i l_j (computeSpatialMetricTensor :153-186).
CFV_HD void spatial_metric(const KSPoint& k, double* g) {
    const double l[3] = {k.l0, k.l1, k.l2};
    double H2 = 2.0 * k.H;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) g[i * 3 + j] = (i == j ? 1.0 : 0.0) - ((H2 * l[i]) * l[j]);
}

// lapse 1 / sqrt(1 - 2H) (:278-281)
CFV_HD double lapse_of(const KSPoint& k) { return 1.0 / __builtin_sqrt(1.0 - (2.0 * k.H)); }

// shift (2H / (1 - 2H)) l (:283-296)
CFV_HD void shift_of(const KSPoint& k, double* be) {
    double f = (2.0 * k.H) / (1.0 - (2.0 * k.H));
    be[0] = f * k.l0;
    be[1] = f * k.l1;
    be[2] = f * k.l2;
}

// What p2c, c2p and the fluxes need at one position.
struct Geo {
    double g[9];     // spatial metric
    double sdet;     // sqrt(det gamma)
    double alpha;    // lapse
    double beta[3];  // shift
};

CFV_HD Geo geo_at(const Hole& b, double x, double y, double z) {
    Geo G;
    KSPoint k = ks_point(b, x, y, z);
    spatial_metric(k, G.g);
    G.sdet = __builtin_sqrt(det3(G.g));
    G.alpha = lapse_of(k);
    shift_of(k, G.beta);
    return G;
}

// inExcisionRegion (:372-386)
CFV_HD bool excised(const Hole& b, double x, double y, double z) {
    double dx = x - b.x, dy = y - b.y, dz = z - b.z;
    return __builtin_sqrt(((dx * dx) + (dy * dy)) + (dz * dz)) <= b.r_exc;
}

// ---------------------------------------------------------------------------
// State functions (BlackHoleEuler/BlackHoleEulerStateVector.cpp)
// ---------------------------------------------------------------------------

// computeConservedVariableVector (:118-157): densitised by sqrt(det gamma),
// v^2 = gamma_ij v^i v^j, W clamped, tau with W (not W^2).
CFV_HD void gr_p2c(const double* W, const Geo& G, double* U, const PhysParams& P) {
    double vsq = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) vsq += (G.g[i * 3 + j] * W[1 + i]) * W[1 + j];
    double lf = 1.0 / __builtin_sqrt(1.0 - vsq);
    if (vsq > 1.0 - 1e-8) lf = 1.0 / __builtin_sqrt(1.0 - 1e-8);
    double h = 1.0 + ((W[4] / W[0]) * P.sr_gg1);
    double sr = G.sdet * W[0];
    double m = (sr * h) * (lf * lf);
    U[0] = sr * lf;
    U[1] = m * W[1];
    U[2] = m * W[2];
    U[3] = m * W[3];
    U[4] = G.sdet * ((((W[0] * h) * lf) - W[4]) - (W[0] * lf));
}

// Primitive recovery shared by setConservedVariableVector (:30-103), the
// fluxes (:159-357) and the source term (:359-517) (iteration cap: the
// reference's 1000 unless configured): divide by sqrt(det gamma), then the
// special-relativistic quartic Newton iteration from 1.0 with the flat |S|^2,
// stopping when |x - x_new| < tol and keeping the previous iterate; rho and p
// floored at 1e-8 after the velocities are formed.  C2PRec keeps the
// intermediate values the reverse-mode code differentiates through.
struct C2PRec {
    double D, S[3], tau;   // undensitised conserved state
    double q, sq;          // (tau + D)^2 - |S|^2 and the root used (floored q: sqrt(1e-8))
    double dc, ec;         // D / sq, (D + tau) / sq
    double c2, c4, eta;    // quartic coefficients (c1 follows from dc)
    double x, fd;          // the kept iterate and f'(x) there
    double lf, h, rho, rhw, p;
    double R[5];           // (rho, v^x, v^y, v^z, p), rho and p floored
};

// The three parts of the solve: the undensitised state and the quartic's
// coefficients (setup), the Newton iteration, and the primitives from its
// kept iterate (finish).
struct C2PCoef {
    double c1, c2x2, c4x4, e3c4;
};

CFV_HD void c2p_setup(const double* U, double sdet, const PhysParams& P, C2PRec& c, C2PCoef& k) {
    double is = 1.0 / sdet;
    c.D = div_by(U[0], sdet, is);
    c.S[0] = div_by(U[1], sdet, is);
    c.S[1] = div_by(U[2], sdet, is);
    c.S[2] = div_by(U[3], sdet, is);
    c.tau = div_by(U[4], sdet, is);
    const double D = c.D, Sx = c.S[0], Sy = c.S[1], Sz = c.S[2], tau = c.tau;
    c.q = ((tau + D) * (tau + D)) - (((Sx * Sx) + (Sy * Sy)) + (Sz * Sz));
    c.sq = (c.q < 1e-8) ? __builtin_sqrt(1e-8) : __builtin_sqrt(c.q);
    double isq = 1.0 / c.sq;
    c.dc = div_by(D, c.sq, isq);
    c.ec = div_by(D + tau, c.sq, isq);
    const double dc = c.dc, ec = c.ec;
    k.c1 = (-2.0 * dc) * P.sr_c1f;
    c.c2 = ((P.sr_gm2g * ((ec * ec) - 1.0)) + 1.0) - (((dc * dc) * P.sr_gm1g) * P.sr_gm1g);
    c.c4 = (ec * ec) - 1.0;
    c.eta = (2.0 * dc) * P.sr_gm1g;
    k.c2x2 = 2.0 * c.c2;
    k.c4x4 = 4.0 * c.c4;
    k.e3c4 = (3.0 * c.eta) * c.c4;
}

// One Newton step of the quartic at x: returns x_new, fd = f'(x).
CFV_HD double c2p_newton(const C2PRec& c, const C2PCoef& k, double c0, double x, double& fd) {
    double x2 = x * x;
    double x3 = x2 * x;
    double f = ((((c.c4 * x3) * (x - c.eta)) + (c.c2 * x2)) + (k.c1 * x)) + c0;
    fd = ((k.c1 + (k.c2x2 * x)) + (k.c4x4 * x3)) - (k.e3c4 * x2);
    return x - (f / fd);
}

CFV_HD void c2p_finish(C2PRec& c, double x, double fd, const PhysParams& P) {
    const double dc = c.dc, ec = c.ec, D = c.D, tau = c.tau;
    double lf = ((0.5 * ec) * x) * (1.0 + __builtin_sqrt(1.0 + (P.sr_4gm1g * ((1.0 - (dc * x)) / ((ec * ec) * (x * x))))));
    double h = 1.0 / (dc * x);
    double rho = D / lf;
    double rhw = (rho * h) * (lf * lf);
    double irhw = 1.0 / rhw;
    c.R[1] = div_by(c.S[0], rhw, irhw);
    c.R[2] = div_by(c.S[1], rhw, irhw);
    c.R[3] = div_by(c.S[2], rhw, irhw);
    double p = (rhw - D) - tau;
    c.R[0] = (rho < 1e-8) ? 1e-8 : rho;
    c.R[4] = (p < 1e-8) ? 1e-8 : p;
    c.x = x;
    c.fd = fd;
    c.lf = lf;
    c.h = h;
    c.rho = rho;
    c.rhw = rhw;
    c.p = p;
}

CFV_HD void gr_con2prim_rec(const double* U, double sdet, double tol, const PhysParams& P, C2PRec& c) {
    C2PCoef k;
    c2p_setup(U, sdet, P, c, k);
    double x = 1.0, fd = 1.0;
    for (int it = 0; it < P.c2p_maxit; ++it) {
        double xn = c2p_newton(c, k, P.sr_c0, x, fd);
        if (__builtin_fabs(x - xn) < tol) break;
        x = xn;
    }
    c2p_finish(c, x, fd, P);
}

// Two independent primitive recoveries side by side.  Each chain takes exactly
// the steps of gr_con2prim_rec (the same iterates, the same kept x and f'(x),
// the same finish), so each record is bitwise what gr_con2prim_rec gives; the
// Newton steps of the two are interleaved so that one chain's dependent FP64
// latency hides the other's.  A chain that has stopped is held by selects
// while the other continues (its extra steps are computed and discarded).
CFV_HD void gr_con2prim_rec2(const double* Ua, double sdeta, double tola, const double* Ub, double sdetb,
                             double tolb, const PhysParams& P, C2PRec& ca, C2PRec& cb) {
    C2PCoef ka, kb;
    c2p_setup(Ua, sdeta, P, ca, ka);
    c2p_setup(Ub, sdetb, P, cb, kb);
    double xa = 1.0, fda = 1.0, xb = 1.0, fdb = 1.0;
    bool ra = true, rb = true;
    for (int it = 0; it < P.c2p_maxit; ++it) {
        double ga, gb;
        const double xna = c2p_newton(ca, ka, P.sr_c0, xa, ga);
        const double xnb = c2p_newton(cb, kb, P.sr_c0, xb, gb);
        const bool sa = __builtin_fabs(xa - xna) < tola, sb = __builtin_fabs(xb - xnb) < tolb;
        fda = ra ? ga : fda;
        fdb = rb ? gb : fdb;
        xa = (ra && !sa) ? xna : xa;
        xb = (rb && !sb) ? xnb : xb;
        ra = ra && !sa;
        rb = rb && !sb;
        if (!ra && !rb) break;
    }
    c2p_finish(ca, xa, fda, P);
    c2p_finish(cb, xb, fdb, P);
}

// R = (rho, v^x, v^y, v^z, p); lf = W; h = specific enthalpy.
CFV_HD void gr_con2prim(const double* U, double sdet, double tol, const PhysParams& P, double* R, double& lf_out,
                        double& h_out) {
    C2PRec c;
    gr_con2prim_rec(U, sdet, tol, P, c);
#pragma unroll
    for (int k = 0; k < 5; ++k) R[k] = c.R[k];
    lf_out = c.lf;
    h_out = c.h;
}

CFV_HD double gr_tol(const PhysParams& P, double reference) { return P.gr_tol > 0.0 ? P.gr_tol : reference; }

CFV_HD void gr_c2p(const double* U, double sdet, const PhysParams& P, double* W) {
    double lf, h;
    gr_con2prim(U, sdet, gr_tol(P, 1e-8), P, W, lf, h);
}

// Two independent gr_c2p side by side (bitwise each gr_c2p).
CFV_HD void gr_c2p2(const double* Ua, double sdeta, const double* Ub, double sdetb, const PhysParams& P, double* Wa,
                    double* Wb) {
    C2PRec ca, cb;
    const double tol = gr_tol(P, 1e-8);
    gr_con2prim_rec2(Ua, sdeta, tol, Ub, sdetb, tol, P, ca, cb);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        Wa[k] = ca.R[k];
        Wb[k] = cb.R[k];
    }
}

// compute{X,Y}FluxVector(cons) (:159-257 x, tolerance 1e-10; :259-357 y, 1e-8):
// alpha sqrt(gamma) (rho W vt, rho h W^2 v_i vt + p delta_i, (rho h W^2 - p - rho W) vt + p v)
// with vt = v^axis - beta^axis / alpha, from the floored density and pressure.
// gr_flux_of forms the flux from a con2prim record of U (so that the reverse
// pass can reuse the record of its forward evaluation).
CFV_HD double gr_flux_tol(const PhysParams& P, int axis) { return gr_tol(P, axis == 0 ? 1e-10 : 1e-8); }

template <int AXIS>
CFV_HD void gr_flux_of(const C2PRec& c, const Geo& G, double* F) {
    const double* R = c.R;
    const double lf = c.lf, h = c.h;
    double rho = R[0], p = R[4];
    double as = G.alpha * G.sdet;
    double vn = R[1 + AXIS];
    double vt = vn - (G.beta[AXIS] / G.alpha);
    double rhw = (rho * h) * (lf * lf);
    F[0] = as * ((rho * lf) * vt);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double t = rhw * (R[1 + k] * vt);
        if (k == AXIS) t = t + p;
        F[1 + k] = as * t;
    }
    F[4] = as * ((((rhw - p) - (rho * lf)) * vt) + (p * vn));
}

template <int AXIS>
CFV_HD void gr_flux(const double* U, const Geo& G, const PhysParams& P, double* F) {
    C2PRec c;
    gr_con2prim_rec(U, G.sdet, gr_flux_tol(P, AXIS), P, c);
    gr_flux_of<AXIS>(c, G, F);
}

// Two independent fluxes along one axis (their Newton solves interleaved):
// bitwise gr_flux of each.
template <int AXIS>
CFV_HD void gr_flux2(const double* Ua, const Geo& Ga, const double* Ub, const Geo& Gb, const PhysParams& P,
                     double* Fa, double* Fb) {
    C2PRec ca, cb;
    const double tol = gr_flux_tol(P, AXIS);
    gr_con2prim_rec2(Ua, Ga.sdet, tol, Ub, Gb.sdet, tol, P, ca, cb);
    gr_flux_of<AXIS>(ca, Ga, Fa);
    gr_flux_of<AXIS>(cb, Gb, Fb);
}

// ---------------------------------------------------------------------------
// Source term (computeSourceTermVector, :359-517) at one cell centre
// ---------------------------------------------------------------------------

// 4x4 inverse by Cayley-Hamilton (MatrixAlgebra::computeMatrixInverse,
// Mathematics/MatrixAlgebra.cpp:145-194) with the cofactor determinant;
// matrix products accumulate from 0.0 over k (:56-82).
CFV_HD double det4(const double* m) {
    double d = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double c[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) c[r * 3 + q] = m[(r + 1) * 4 + (q < i ? q : q + 1)];
        if (i % 2 == 0) d += m[i] * det3(c);
        else d -= m[i] * det3(c);
    }
    return d;
}

CFV_HD void inverse4(const double* m, double* out) {
    double det = det4(m);
    double tr = 0.0;
    tr += m[0];
    tr += m[5];
    tr += m[10];
    tr += m[15];
    double m2[16], m3[16];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) s += m[i * 4 + k] * m[k * 4 + j];
            m2[i * 4 + j] = s;
        }
    double tr2 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) tr2 += m2[i * 5];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) s += m2[i * 4 + k] * m[k * 4 + j];
            m3[i * 4 + j] = s;
        }
    double tr3 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) tr3 += m3[i * 5];
    double c3 = (1.0 / 6.0) * ((((tr * tr) * tr) - ((3.0 * tr) * tr2)) + (2.0 * tr3));
    double c2 = 0.5 * ((tr * tr) - tr2);
    double id = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double e = c3 * (i == j ? 1.0 : 0.0);
            e = e - (c2 * m[i * 4 + j]);
            e = e + (tr * m2[i * 4 + j]);
            e = e - m3[i * 4 + j];
            out[i * 4 + j] = id * e;
        }
}

// Position-only part of the source term at (x, y, z) with finite-difference
// spacing h: everything but the fluid state.
struct SourceGeo {
    double alpha, sdet;
    double beta[3];
    double gm[9];     // spatial metric
    double gi4[16];   // inverse spacetime metric (from the 4-vector with the reference's term)
    double da[3];     // lapse derivative
    double db[9];     // shift derivative [direction][component]
    double dg[27];    // spatial-metric derivative [direction][i][j]
    double K[9];      // extrinsic curvature
};

// The reference's central differences (f(x + h/2) - f(x - h/2)) / h
// (computeLapseFunctionDerivative :298-310, computeKerrSchildScalarDerivative :115-127)
// and (1/h) (v(x + h/2) - v(x - h/2)) for vectors and tensors
// (:129-151, :250-276, :312-334), at offsets x + (0.5 * h).
CFV_HD void source_geo(const Hole& b, double x, double y, double z, double h, SourceGeo& S) {
    KSPoint c = ks_point(b, x, y, z);
    S.alpha = lapse_of(c);
    shift_of(c, S.beta);
    spatial_metric(c, S.gm);
    S.sdet = __builtin_sqrt(det3(S.gm));
    // g_{mu nu} = eta - 2 H l_mu l_nu with l = (-1, l_x, l4y, l_z) (:188-227)
    {
        double l4[4] = {-1.0, c.l0, ks_l4y(b, c), c.l2};
        double g4[16];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                double eta = (i == j) ? (i == 0 ? -1.0 : 1.0) : 0.0;
                g4[i * 4 + j] = eta - (((2.0 * c.H) * l4[i]) * l4[j]);
            }
        inverse4(g4, S.gi4);
    }
    double ih = 1.0 / h;
    double dH[3], dl[9];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double o = 0.5 * h;
        KSPoint p = ks_point(b, d == 0 ? x + o : x, d == 1 ? y + o : y, d == 2 ? z + o : z);
        KSPoint m = ks_point(b, d == 0 ? x - o : x, d == 1 ? y - o : y, d == 2 ? z - o : z);
        S.da[d] = (lapse_of(p) - lapse_of(m)) / h;
        dH[d] = (p.H - m.H) / h;
        double bp[3], bm[3], gp[9], gmm[9];
        shift_of(p, bp);
        shift_of(m, bm);
        spatial_metric(p, gp);
        spatial_metric(m, gmm);
#pragma unroll
        for (int i = 0; i < 3; ++i) S.db[d * 3 + i] = ih * (bp[i] - bm[i]);
#pragma unroll
        for (int i = 0; i < 9; ++i) S.dg[d * 9 + i] = ih * (gp[i] - gmm[i]);
        dl[d * 3 + 0] = ih * (p.l0 - m.l0);
        dl[d * 3 + 1] = ih * (p.l1 - m.l1);
        dl[d * 3 + 2] = ih * (p.l2 - m.l2);
    }
    // computeExtrinsicCurvatureTensor (:336-370)
    const double l[3] = {c.l0, c.l1, c.l2};
    double al = S.alpha, H = c.H;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double k = al * ((((-(l[i] * dH[j])) - (l[j] * dH[i])) - (H * dl[j * 3 + i])) - (H * dl[i * 3 + j]));
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                k += ((al * 2.0) * (H * H)) * (((l[i] * l[q]) * dl[q * 3 + j]) + ((l[j] * l[q]) * dl[q * 3 + i]));
                k += (((((al * 2.0) * H) * l[i]) * l[j]) * l[q]) * dH[q];
            }
            S.K[i * 3 + j] = k;
        }
}

// ---------------------------------------------------------------------------
// Geometry cache.  Every metric quantity depends on position only, and the
// positions the kernels evaluate form the half lattice (ax h / 2, ay h / 2):
// cell centres (ax, ay even), x interfaces (ax odd) and y interfaces (ay odd).
// Where h's significand is short enough that every such position and every
// sum or difference of two of them is exact (h = 40 / 1024 = 5 * 2^-7, for
// one), the expressions the reference forms at its call sites ((c - 0.5) * h,
// x - h, (j - 1) * h, x + 0.5 * h ...) all land on the exact lattice value,
// so a table of geo_at over the lattice, built with the same function, gives
// the same bits as evaluating at the call site.  The source term's position
// part (SourceGeo) is formed at x = j h, y = i h with the step's h, the
// expressions of its only call site, so its table is exact for any h.
// ---------------------------------------------------------------------------
constexpr int kGeoVals = 14;  // g[9], sdet, alpha, beta[3]
constexpr int kSrcVals = 78;  // alpha, sdet, beta[3], gm[9], gi4[16], da[3], db[9], dg[27], K[9]

struct GeoTab {
    const double* geo;  // three classes (centres, x faces, y faces) of kGeoVals planes, rows x pitch each
    long pitch;         // doubles per row: nx + 4 points (x index (ax + 4) / 2, ax from -4 / -3)
    long plane;         // rows * pitch; rows = ny + 4 (2D, y index (ay + 4) / 2) or 1 (1D)
    const double* src;  // kSrcVals planes of ny x spitch: SourceGeo at the cell centres
    long spitch, splane;
    const double* q;    // roundtrip planes (5) over cells [-1, nx] x [-1, ny], pitch qpitch
    long qpitch, qplane;
};

CFV_HD long geo_index(const GeoTab& T, int dim, int ax, int ay) {
    const int cls = (ax & 1) | ((ay & 1) << 1);
    const long row = dim == 2 ? (long)((ay + 4) >> 1) : 0;
    return cls * kGeoVals * T.plane + row * T.pitch + ((ax + 4) >> 1);
}

CFV_HD Geo geo_from_table(const GeoTab& T, int dim, int ax, int ay) {
    const double* p = T.geo + geo_index(T, dim, ax, ay);
    Geo G;
#pragma unroll
    for (int k = 0; k < 9; ++k) G.g[k] = p[k * T.plane];
    G.sdet = p[9 * T.plane];
    G.alpha = p[10 * T.plane];
#pragma unroll
    for (int k = 0; k < 3; ++k) G.beta[k] = p[(11 + k) * T.plane];
    return G;
}

CFV_HD void geo_to_table(const Geo& G, double* p, long plane) {
#pragma unroll
    for (int k = 0; k < 9; ++k) p[k * plane] = G.g[k];
    p[9 * plane] = G.sdet;
    p[10 * plane] = G.alpha;
#pragma unroll
    for (int k = 0; k < 3; ++k) p[(11 + k) * plane] = G.beta[k];
}

CFV_HD void source_to_table(const SourceGeo& S, double* p, long plane) {
    int k = 0;
    p[(k++) * plane] = S.alpha;
    p[(k++) * plane] = S.sdet;
    for (int i = 0; i < 3; ++i) p[(k++) * plane] = S.beta[i];
    for (int i = 0; i < 9; ++i) p[(k++) * plane] = S.gm[i];
    for (int i = 0; i < 16; ++i) p[(k++) * plane] = S.gi4[i];
    for (int i = 0; i < 3; ++i) p[(k++) * plane] = S.da[i];
    for (int i = 0; i < 9; ++i) p[(k++) * plane] = S.db[i];
    for (int i = 0; i < 27; ++i) p[(k++) * plane] = S.dg[i];
    for (int i = 0; i < 9; ++i) p[(k++) * plane] = S.K[i];
}

CFV_HD void source_from_table(const double* p, long plane, SourceGeo& S) {
    int k = 0;
    S.alpha = p[(k++) * plane];
    S.sdet = p[(k++) * plane];
#pragma unroll
    for (int i = 0; i < 3; ++i) S.beta[i] = p[(k++) * plane];
#pragma unroll
    for (int i = 0; i < 9; ++i) S.gm[i] = p[(k++) * plane];
#pragma unroll
    for (int i = 0; i < 16; ++i) S.gi4[i] = p[(k++) * plane];
#pragma unroll
    for (int i = 0; i < 3; ++i) S.da[i] = p[(k++) * plane];
#pragma unroll
    for (int i = 0; i < 9; ++i) S.db[i] = p[(k++) * plane];
#pragma unroll
    for (int i = 0; i < 27; ++i) S.dg[i] = p[(k++) * plane];
#pragma unroll
    for (int i = 0; i < 9; ++i) S.K[i] = p[(k++) * plane];
}

// Read access to a position's SourceGeo (the source term is written against
// this accessor).
struct SrcRef {
    const SourceGeo& s;
    CFV_HD double alpha() const { return s.alpha; }
    CFV_HD double sdet() const { return s.sdet; }
    CFV_HD double beta(int k) const { return s.beta[k]; }
    CFV_HD double gm(int k) const { return s.gm[k]; }
    CFV_HD double gi4(int k) const { return s.gi4[k]; }
    CFV_HD double da(int k) const { return s.da[k]; }
    CFV_HD double db(int k) const { return s.db[k]; }
    CFV_HD double dg(int k) const { return s.dg[k]; }
    CFV_HD double K(int k) const { return s.K[k]; }
};

// computeSourceTermVector from the fluid state and the position's SourceGeo:
// u^mu (:438-442), T^{mu nu} = rho h u^mu u^nu + p g^{mu nu} (:444-453), the
// momentum source from d gamma, beta, T, d beta and d alpha (:478-494), the
// energy source from K and d alpha (:496-507), all times alpha sqrt(gamma) (:509).
template <class S>
CFV_HD void gr_source_of(const double* U, const S& sg, const PhysParams& P, double* out) {
    double R[5], lf, hh;
    gr_con2prim(U, sg.sdet(), gr_tol(P, 1e-8), P, R, lf, hh);
    double rho = R[0], p = R[4];
    double al = sg.alpha();
    double u[4];
    u[0] = lf / al;
#pragma unroll
    for (int k = 0; k < 3; ++k) u[1 + k] = ((al * u[0]) * R[1 + k]) - (sg.beta(k) * u[0]);
    double T[16];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) T[i * 4 + j] = (((rho * hh) * u[i]) * u[j]) + (p * sg.gi4(i * 4 + j));
    double vc[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        vc[i] = ((sg.gm(i * 3 + 0) * R[1]) + (sg.gm(i * 3 + 1) * R[2])) + (sg.gm(i * 3 + 2) * R[3]);
    double rhw = (rho * hh) * (lf * lf);
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double d = sg.dg(i * 9 + j * 3 + k);
                s[i + 1] += T[0] * (((0.5 * sg.beta(j)) * sg.beta(k)) * d);
                s[i + 1] += (T[j + 1] * sg.beta(k)) * d;
                s[i + 1] += 0.5 * (T[(j + 1) * 4 + (k + 1)] * d);
            }
            s[i + 1] += ((rhw * vc[j]) / al) * sg.db(i * 3 + j);
        }
        s[i + 1] -= (T[0] * al) * sg.da(i);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double K = sg.K(i * 3 + j);
            s[4] += T[0] * ((sg.beta(i) * sg.beta(j)) * K);
            s[4] += T[i + 1] * ((2.0 * sg.beta(j)) * K);
            s[4] += T[(i + 1) * 4 + (j + 1)] * K;
        }
        s[4] -= T[0] * (sg.beta(i) * sg.da(i));
        s[4] -= T[i + 1] * sg.da(i);
    }
    double f = al * sg.sdet();
#pragma unroll
    for (int k = 0; k < 5; ++k) out[k] = f * s[k];
}

CFV_HD void gr_source(const double* U, const SourceGeo& S, const PhysParams& P, double* out) {
    gr_source_of(U, SrcRef{S}, P, out);
}

// computeSoundSpeed (:519-549) with BlackHoleEulerEigenvalues (:3-96): the
// largest signed material / fast / slow eigenvalue over the three
// directions, floored at 0.  Only the diagonal of the Cayley-Hamilton
// inverse spatial metric is used.
CFV_HD double gr_sound_speed(const double* W, const Hole& b, double x, double y, double z,
                                                 const PhysParams& P) {
    KSPoint k = ks_point(b, x, y, z);
    double al = lapse_of(k);
    double be[3], gm[9];
    shift_of(k, be);
    spatial_metric(k, gm);
    // inverse3 diagonal: (1/det) (c I - tr m + m^2), c = (tr^2 - tr(m^2)) / 2
    double det = det3(gm);
    double tr = 0.0;
    tr += gm[0];
    tr += gm[4];
    tr += gm[8];
    double m2d[3], tr2 = 0.0;  // diagonal of m^2, accumulated from 0.0 over q
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) s += gm[i * 3 + q] * gm[q * 3 + i];
        m2d[i] = s;
        tr2 += s;
    }
    double c = 0.5 * ((tr * tr) - tr2);
    double id = 1.0 / det;
    double num = (P.gamma * W[4]) / W[0];
    double den = 1.0 + ((W[4] / W[0]) * P.sr_gg1);
    double cs = __builtin_sqrt(num / den);
    double vsq = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) vsq += (gm[i * 3 + j] * W[1 + i]) * W[1 + j];
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double gii = id * (((c * 1.0) - (tr * gm[i * 4])) + m2d[i]);
        double v = W[1 + i];
        double mat = (al * v) - be[i];
        double pre = al / (1.0 - (vsq * (cs * cs)));
        double rt = cs * __builtin_sqrt((1.0 - vsq) * ((gii * (1.0 - (vsq * (cs * cs)))) - ((v * v) * (1.0 - (cs * cs)))));
        double fast = (pre * ((v * (1.0 - (cs * cs))) + rt)) - be[i];
        double slow = (pre * ((v * (1.0 - (cs * cs))) - rt)) - be[i];
        if (mat > s) s = mat;
        if (fast > s) s = fast;
        if (slow > s) s = slow;
    }
    return s;
}

}  // namespace cfv
