/* CPU restatement of the Cosmos finite-volume arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: tests, the golden-fixture
 * scripts and the CPU-baseline leg of bench.py use it as the checker.  The
 * product path (the HIP library) never links, loads or calls it.
 *
 * Every expression keeps the reference's evaluation order, written out with
 * explicit parentheses, so that results are bitwise equal to the reference
 * built serially (x86-64, SSE2, no FMA contraction).  Each function cites
 * the reference lines it restates.  VectorAlgebra::subtractVectors(a, b) is
 * a + (-1 * b) (Mathematics/VectorAlgebra.cpp:17-20), which is bitwise a - b,
 * so it is written as a - b.
 *
 * Parity pins (tests/test_oracle.py): bitwise against oracle/_ref (the
 * reference's own sources) on every entry point, against the committed
 * golden fixtures, and against closed forms (limiter tables, uniform-state
 * invariance, conservation, the SR con2prim round trip).
 */
#include "cosmos_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double g_c2p_tol = 1e-8;
static int g_c2p_maxit = 1000;
static int g_threads = 0;

void orc_set_c2p(double tol, int max_iterations) {
    g_c2p_tol = tol;
    g_c2p_maxit = max_iterations;
}

void orc_set_threads(int n) { g_threads = n; }

int orc_thread_count(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* Con2prim event counters (a diagnostic with no counterpart in the reference,
 * which accepts these states silently): [0] solves that ran to the iteration
 * cap, [1] solves whose (tau + D)^2 - |S|^2 fell below 1e-8 (root taken at
 * sqrt(1e-8)), [2] solves whose density and [3] whose pressure were raised to
 * 1e-8 (RelativisticEuler/RelativisticEulerStateVector.cpp:45-50, 59-80,
 * 93-100).  Each distinct solve of a sweep is counted once: every half step,
 * interface flux and cell update of a line is evaluated once here (the
 * reference repeats some with identical inputs). */
static long long g_events[4];

void orc_events_reset(void) { memset(g_events, 0, sizeof g_events); }

void orc_events_read(long long* out) { memcpy(out, g_events, sizeof g_events); }

static void count_event(int k) {
#pragma omp atomic
    g_events[k] += 1;
}

/* ------------------------------------------------------------------------ */
/* Newtonian physics                                                        */
/* ------------------------------------------------------------------------ */

/* EulerStateVector::computeConservedVariableVector (Euler/EulerStateVector.cpp:55-66),
 * computeTotalEnergy (:129-134), EulerEquationOfState::computeSpecificInternalEnergy
 * (Euler/EulerEquationOfState.cpp:3-8). */
static void nw_p2c(const double* W, double* U, double g) {
    double rho = W[0], u = W[1], v = W[2], w = W[3], p = W[4];
    double vsq = ((u * u) + (v * v)) + (w * w);
    double e = p / ((g - 1.0) * rho);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * w;
    U[4] = rho * ((0.5 * vsq) + e);
}

/* EulerStateVector::setConservedVariableVector (Euler/EulerStateVector.cpp:30-40),
 * EulerEquationOfState::computePressure (Euler/EulerEquationOfState.cpp:10-15). */
static void nw_c2p(const double* U, double* W, double g) {
    double rho = U[0];
    double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
    double vsq = ((u * u) + (v * v)) + (w * w);
    double e = (U[4] / rho) - (0.5 * vsq);
    W[0] = rho;
    W[1] = u;
    W[2] = v;
    W[3] = w;
    W[4] = (e * (g - 1.0)) * rho;
}

/* EulerStateVector::compute{X,Y}FluxVector(cons) (Euler/EulerStateVector.cpp:68-89, 96-117). */
static void nw_flux(int axis, const double* U, double* F, double g) {
    double rho = U[0];
    double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
    double vsq = ((u * u) + (v * v)) + (w * w);
    double E = U[4] / rho;
    double e = E - (0.5 * vsq);
    double p = (e * (g - 1.0)) * rho;
    if (axis == 0) {
        F[0] = rho * u;
        F[1] = (rho * (u * u)) + p;
        F[2] = rho * (u * v);
        F[3] = rho * (u * w);
        F[4] = (rho * (u * E)) + (u * p);
    } else {
        F[0] = rho * v;
        F[1] = rho * (v * u);
        F[2] = (rho * (v * v)) + p;
        F[3] = rho * (v * w);
        F[4] = (rho * (v * E)) + (v * p);
    }
}

/* EulerEquationOfState::computeSoundSpeed (Euler/EulerEquationOfState.cpp:17-22). */
static double nw_sound_speed(const double* W, double g) { return sqrt((g * W[4]) / W[0]); }

/* ------------------------------------------------------------------------ */
/* Special-relativistic physics                                             */
/* ------------------------------------------------------------------------ */

/* RelativisticEulerStateVector::computeConservedVariableVector
 * (RelativisticEuler/RelativisticEulerStateVector.cpp:116-136) and
 * RelativisticEulerEquationOfState::computeSpecificEnthalpy
 * (RelativisticEuler/RelativisticEulerEquationOfState.cpp:3-8). */
static void sr_p2c(const double* W, double* U, double g) {
    double rho = W[0], u = W[1], v = W[2], w = W[3], p = W[4];
    double vsq = ((u * u) + (v * v)) + (w * w);
    double lf = 1.0 / sqrt(1.0 - vsq);
    if (vsq > 1.0 - 1e-8) lf = 1.0 / sqrt(1.0 - 1e-8);
    double h = 1.0 + ((p / rho) * (g / (g - 1.0)));
    double rhw = (rho * h) * (lf * lf);
    U[0] = rho * lf;
    U[1] = rhw * u;
    U[2] = rhw * v;
    U[3] = rhw * w;
    U[4] = (rhw - p) - (rho * lf);
}

/* The con2prim shared by setConservedVariableVector
 * (RelativisticEuler/RelativisticEulerStateVector.cpp:30-101) and the two
 * flux functions (:138-220, :227-309): quartic Newton from 1.0, stopping when
 * |x - x_new| < tol and keeping the old iterate, floors 1e-8 on rho and p.
 * Outputs rho, v, p (floored), W and h. */
static void sr_con2prim(const double* U, double g, double* out_rho, double* out_v /*3*/,
                        double* out_p, double* out_lf, double* out_h) {
    double D = U[0], Sx = U[1], Sy = U[2], Sz = U[3], tau = U[4];
    double q = ((tau + D) * (tau + D)) - (((Sx * Sx) + (Sy * Sy)) + (Sz * Sz));
    double dc = D / sqrt(q);
    double ec = (D + tau) / sqrt(q);
    if (q < 1e-8) {
        dc = D / sqrt(1e-8);
        ec = (D + tau) / sqrt(1e-8);
        count_event(1);
    }
    double gm1g = (g - 1.0) / g;
    double c0 = -1.0 / (g * g);
    double c1 = (-2.0 * dc) * ((g - 1.0) / (g * g));
    double c2 = ((((g - 2.0) / g) * ((ec * ec) - 1.0)) + 1.0) - (((dc * dc) * gm1g) * gm1g);
    double c4 = (ec * ec) - 1.0;
    double eta = (2.0 * dc) * gm1g;
    double x = 1.0;
    int it = 0, converged = 0;
    while (it < g_c2p_maxit) {
        double x2 = x * x;
        double x3 = x2 * x;
        double f = (((c4 * x3) * (x - eta)) + (c2 * x2)) + (c1 * x) + c0;
        double fd = ((c1 + ((2.0 * c2) * x)) + ((4.0 * c4) * x3)) - (((3.0 * eta) * c4) * x2);
        double xn = x - (f / fd);
        if (fabs(x - xn) < g_c2p_tol) {
            it = g_c2p_maxit;
            converged = 1;
        } else {
            it += 1;
            x = xn;
        }
    }
    if (!converged) count_event(0);
    double lf = ((0.5 * ec) * x) * (1.0 + sqrt(1.0 + ((4.0 * gm1g) * ((1.0 - (dc * x)) / ((ec * ec) * (x * x))))));
    double h = 1.0 / (dc * x);
    double rho = D / lf;
    double rhw = (rho * h) * (lf * lf);
    out_v[0] = Sx / rhw;
    out_v[1] = Sy / rhw;
    out_v[2] = Sz / rhw;
    double p = (rhw - D) - tau;
    if (rho < 1e-8) {
        rho = 1e-8;
        count_event(2);
    }
    if (p < 1e-8) {
        p = 1e-8;
        count_event(3);
    }
    *out_rho = rho;
    *out_p = p;
    *out_lf = lf;
    *out_h = h;
}

static void sr_c2p(const double* U, double* W, double g) {
    double rho, v[3], p, lf, h;
    sr_con2prim(U, g, &rho, v, &p, &lf, &h);
    W[0] = rho;
    W[1] = v[0];
    W[2] = v[1];
    W[3] = v[2];
    W[4] = p;
}

/* RelativisticEulerStateVector::compute{X,Y}FluxVector(cons)
 * (RelativisticEuler/RelativisticEulerStateVector.cpp:138-220, 227-309). */
static void sr_flux(int axis, const double* U, double* F, double g) {
    double rho, v[3], p, lf, h;
    sr_con2prim(U, g, &rho, v, &p, &lf, &h);
    double rhw = (rho * h) * (lf * lf);
    double vn = v[axis];
    F[0] = (rho * lf) * vn;
    F[1] = rhw * (v[0] * vn);
    F[2] = rhw * (v[1] * vn);
    F[3] = rhw * (v[2] * vn);
    F[1 + axis] = F[1 + axis] + p;
    F[4] = (((rhw - p) - (rho * lf)) * vn) + (p * vn);
}

/* RelativisticEulerEquationOfState::computeSoundSpeed
 * (RelativisticEuler/RelativisticEulerEquationOfState.cpp:10-18). */
static double sr_sound_speed(const double* W, double g) {
    double num = (g * W[4]) / W[0];
    double den = 1.0 + ((W[4] / W[0]) * (g / (g - 1.0)));
    return sqrt(num / den);
}

/* ------------------------------------------------------------------------ */
/* Physics dispatch                                                         */
/* ------------------------------------------------------------------------ */

static void p2c(int phys, const double* W, double* U, double g) {
    if (phys == 0) nw_p2c(W, U, g); else sr_p2c(W, U, g);
}
static void c2p(int phys, const double* U, double* W, double g) {
    if (phys == 0) nw_c2p(U, W, g); else sr_c2p(U, W, g);
}
static void flux(int phys, int axis, const double* U, double* F, double g) {
    if (phys == 0) nw_flux(axis, U, F, g); else sr_flux(axis, U, F, g);
}

/* ------------------------------------------------------------------------ */
/* Slope limiters (Solvers/SlopeLimiters.cpp)                               */
/* ------------------------------------------------------------------------ */

/* std::min / std::max as libstdc++ defines them: min(a,b) = (b < a) ? b : a. */
static double std_min(double a, double b) { return (b < a) ? b : a; }
static double std_max(double a, double b) { return (a < b) ? b : a; }

/* computeGradientRatio (:3-6) */
static double gradient_ratio(double r, double bias) { return 2.0 / ((1.0 - bias) + ((1.0 + bias) * r)); }

/* computeSuperBeeSlopeLimiter (:8-25), computeVanLeerSlopeLimiter (:27-37),
 * computeMinBeeSlopeLimiter (:39-53), computeSlopeLimiter (:55-69). */
double orc_limiter(double r, double bias, int limiter) {
    if (limiter == 0) {
        if (r < 0.0) return 0.0;
        if (r >= 0.0 && r < 0.5) return 2.0 * r;
        if (r >= 0.5 && r < 1.0) return 1.0;
        return std_min(std_min(r, gradient_ratio(r, bias)), 2.0);
    }
    if (limiter == 1) {
        if (r < 0.0) return 0.0;
        return std_min((2.0 * r) / (1.0 + r), gradient_ratio(r, bias));
    }
    if (r < 0.0) return 0.0;
    if (r >= 0.0 && r < 1.0) return r;
    return std_min(1.0, gradient_ratio(r, bias));
}

/* Which piece of computeSlopeLimiter's definition (:8-69) gives phi(r): 0 for
 * r < 0 (or NaN), then the pieces in the order the reference tests them, the
 * std::min candidates last (1 the first, 2 the gradient ratio, 3 the constant). */
static int limiter_piece(double r, double bias, int limiter) {
    if (!(r >= 0.0)) return 0;
    double xi = gradient_ratio(r, bias);
    if (limiter == 0) {
        if (r < 0.5) return 4;
        if (r < 1.0) return 5;
        double m = std_min(r, xi);
        return (2.0 < m) ? 3 : ((xi < r) ? 2 : 1);
    }
    if (limiter == 1) return (xi < ((2.0 * r) / (1.0 + r))) ? 2 : 1;
    if (r < 1.0) return 4;
    return (xi < 1.0) ? 2 : 3;
}

/* Branch trace (a diagnostic for finite-difference checks): while a buffer is
 * set, each slope component appends one byte, bit 0 = |dl| clamped, bit 1 =
 * |dr| clamped, bits 2-4 = limiter_piece.  The byte order is the call order,
 * so it is only meaningful with one thread (orc_set_threads(1)). */
static unsigned char* g_trace;
static long g_trace_cap, g_trace_len;
static int g_trace_replay;

void orc_trace_begin(unsigned char* buf, long capacity) {
    g_trace = buf;
    g_trace_cap = capacity;
    g_trace_len = 0;
    g_trace_replay = 0;
}

/* Replay a recorded trace instead: each slope component takes the clamps and
 * the limiter piece the trace recorded for it, whatever its own differences
 * are.  The forward map then follows the pieces of the recorded run: the map
 * a reverse pass that differentiates the branches taken (subgradients at the
 * kinks) is the derivative of. */
void orc_trace_replay(unsigned char* buf, long length) {
    g_trace = buf;
    g_trace_cap = length;
    g_trace_len = 0;
    g_trace_replay = 1;
}

long orc_trace_end(void) {
    g_trace = NULL;
    g_trace_replay = 0;
    return g_trace_len;
}

/* phi(r) on a given piece of limiter_piece's numbering. */
static double limiter_on_piece(double r, double bias, int limiter, int piece) {
    switch (piece) {
        case 0: return 0.0;
        case 1: return limiter == 1 ? (2.0 * r) / (1.0 + r) : r;
        case 2: return gradient_ratio(r, bias);
        case 3: return limiter == 0 ? 2.0 : 1.0;
        case 4: return limiter == 0 ? 2.0 * r : r;
        default: return 1.0;
    }
}

/* computeSlopeVector(cons L, M, R) (:71-101): the ratio clamps |d| < 1e-5 to
 * +1e-5, discarding the sign. */
void orc_slope(const double* ul, const double* um, const double* ur, double bias, int limiter, double* out) {
    for (int k = 0; k < 5; ++k) {
        double dl = um[k] - ul[k];
        double dr = ur[k] - um[k];
        double s = ((0.5 * (1.0 + bias)) * dl) + ((0.5 * (1.0 - bias)) * dr);
        double num = dl, den = dr;
        int bits = 0;
        if (g_trace && g_trace_replay) {
            int rec = g_trace_len < g_trace_cap ? g_trace[g_trace_len] : 0;
            g_trace_len += 1;
            if (rec & 1) num = 1e-5;
            if (rec & 2) den = 1e-5;
            out[k] = s * limiter_on_piece(num / den, bias, limiter, rec >> 2);
            continue;
        }
        if (fabs(num) < 1e-5) num = 1e-5, bits |= 1;
        if (fabs(den) < 1e-5) den = 1e-5, bits |= 2;
        out[k] = s * orc_limiter(num / den, bias, limiter);
        if (g_trace) {
            if (g_trace_len < g_trace_cap)
                g_trace[g_trace_len] = (unsigned char)(bits | (limiter_piece(num / den, bias, limiter) << 2));
            g_trace_len += 1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Per-interface building blocks                                            */
/* ------------------------------------------------------------------------ */

/* EulerSolvers::evolveStateByHalf{X,Y}TimeStep (Solvers/EulerSolvers.cpp:191-223),
 * computeEvolutionVector (:163-172), evolveStateByHalfTimeStep (:174-189), on
 * conserved neighbours.  Writes both sides: p0 = side 0 (left face),
 * p1 = side 1 (right face), as primitives. */
static void half_step_cons(int phys, int axis, const double* ul, const double* um, const double* ur,
                           double dt_dx, double bias, int limiter, double g, double* p0, double* p1) {
    double s[5], xl[5], xr[5], fl[5], fr[5], a[5], b[5];
    orc_slope(ul, um, ur, bias, limiter, s);
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    flux(phys, axis, xl, fl, g);
    flux(phys, axis, xr, fr, g);
    double f = 0.5 * dt_dx;
    for (int k = 0; k < 5; ++k) {
        double ev = f * (fl[k] - fr[k]);
        a[k] = xl[k] + ev;
        b[k] = xr[k] + ev;
    }
    c2p(phys, a, p0, g);
    c2p(phys, b, p1, g);
}

/* EulerFirstOrderSolver::compute{X,Y}FORCEFlux(states) (Solvers/EulerFirstOrderSolver.cpp:52-94)
 * with computeLaxFriedrichsFlux (:3-11), computeRichtmyerFlux (:13-19) and
 * computeFORCEFlux (:21-24).  LF and Richtmyer both evaluate p2c and the
 * instance flux (p2c, then the cons flux) of each state; the values are the
 * same, so they are evaluated once. */
static void force_flux_cons(int phys, int axis, const double* ul, const double* fl, const double* ur,
                            const double* fr, double dt_dx, double dx_dt, double g, double* out) {
    double us[5], fs[5];
    for (int k = 0; k < 5; ++k) us[k] = 0.5 * ((ul[k] + ur[k]) + (dt_dx * (fl[k] - fr[k])));
    flux(phys, axis, us, fs, g);
    for (int k = 0; k < 5; ++k) {
        double lf = 0.5 * ((fl[k] + fr[k]) + (dx_dt * (ul[k] - ur[k])));
        out[k] = 0.5 * (lf + fs[k]);
    }
}

static void force_flux_states(int phys, int axis, const double* wl, const double* wr, double dt_dx,
                              double dx_dt, double g, double* out) {
    double ul[5], ur[5], fl[5], fr[5];
    p2c(phys, wl, ul, g);
    p2c(phys, wr, ur, g);
    flux(phys, axis, ul, fl, g);
    flux(phys, axis, ur, fr, g);
    force_flux_cons(phys, axis, ul, fl, ur, fr, dt_dx, dx_dt, g, out);
}

void orc_force_flux(int physics, int axis, const double* l, const double* r, double dx, double dt,
                    double gamma, double* out) {
    force_flux_states(physics, axis, l, r, dt / dx, dx / dt, gamma, out);
}

/* EulerFirstOrderSolver::compute{X,Y}LaxFriedrichsFlux(states) (Solvers/EulerFirstOrderSolver.cpp:27-37,
 * 62-72): computeLaxFriedrichsFlux (:3-11) of the states' p2c and instance fluxes. */
void orc_lf_flux(int physics, int axis, const double* l, const double* r, double dx, double dt, double gamma,
                 double* out) {
    double ul[5], ur[5], fl[5], fr[5];
    p2c(physics, l, ul, gamma);
    p2c(physics, r, ur, gamma);
    flux(physics, axis, ul, fl, gamma);
    flux(physics, axis, ur, fr, gamma);
    for (int k = 0; k < 5; ++k) out[k] = 0.5 * ((fl[k] + fr[k]) + ((dx / dt) * (ul[k] - ur[k])));
}

/* EulerFirstOrderSolver::compute{X,Y}RichtmyerFlux(states) (:39-50, 74-85): the cons flux of
 * computeRichtmyerFlux's intermediate state (:13-19). */
void orc_richtmyer_flux(int physics, int axis, const double* l, const double* r, double dx, double dt,
                        double gamma, double* out) {
    double ul[5], ur[5], fl[5], fr[5], us[5];
    p2c(physics, l, ul, gamma);
    p2c(physics, r, ur, gamma);
    flux(physics, axis, ul, fl, gamma);
    flux(physics, axis, ur, fr, gamma);
    for (int k = 0; k < 5; ++k) us[k] = 0.5 * ((ul[k] + ur[k]) + ((dt / dx) * (fl[k] - fr[k])));
    flux(physics, axis, us, out, gamma);
}

void orc_half_step(int physics, int axis, const double* l, const double* m, const double* r, double dx,
                   double dt, double bias, int limiter, int side, double gamma, double* out) {
    double ul[5], um[5], ur[5], p0[5], p1[5];
    p2c(physics, l, ul, gamma);
    p2c(physics, m, um, gamma);
    p2c(physics, r, ur, gamma);
    half_step_cons(physics, axis, ul, um, ur, dt / dx, bias, limiter, gamma, p0, p1);
    memcpy(out, side == 0 ? p0 : p1, sizeof p0);
}

/* EulerSecondOrderSolver::compute{X,Y}SLICFlux (Solvers/EulerSecondOrderSolver.cpp:3-25). */
void orc_slic_flux(int physics, int axis, const double* ll, const double* l, const double* r,
                   const double* rr, double dx, double dt, double bias, int limiter, double gamma,
                   double* out) {
    double pr[5], pl[5];
    orc_half_step(physics, axis, ll, l, r, dx, dt, bias, limiter, 1, gamma, pr);
    orc_half_step(physics, axis, l, r, rr, dx, dt, bias, limiter, 0, gamma, pl);
    force_flux_states(physics, axis, pr, pl, dt / dx, dx / dt, gamma, out);
}

void orc_p2c(int physics, const double* w, double* u, double gamma) { p2c(physics, w, u, gamma); }
void orc_c2p(int physics, const double* u, double* w, double gamma) { c2p(physics, u, w, gamma); }
void orc_flux(int physics, int axis, const double* u, double* f, double gamma) { flux(physics, axis, u, f, gamma); }

/* ------------------------------------------------------------------------ */
/* One line (a row for x sweeps, a column for y sweeps)                     */
/* ------------------------------------------------------------------------ */

/* Updates n cells of one line in place.  line[j * stride] is cell j.
 * order 2: computeSLICTimeStep / compute{X,Y}SLICTimeStep2D
 *   (Solvers/EulerSecondOrderSolver.cpp:27-131) with the two-ghost mirrored
 *   boundary of insertBoundaryCells{,2D}(.., 2) (Solvers/EulerSolvers.cpp:3-85).
 * order 1: computeFORCETimeStep / compute{X,Y}FORCETimeStep2D
 *   (Solvers/EulerFirstOrderSolver.cpp:103-174) with the one-ghost copy boundary.
 * Each interface flux is evaluated once; the reference evaluates it from both
 * neighbouring cells with identical inputs and arithmetic. */
static void update_line(int phys, int order, int axis, double* line, long stride, int n, double dt_dx,
                        double dx_dt, double bias, int limiter, double g, double* work) {
    int ng = order == 2 ? 2 : 1;
    int m = n + 2 * ng;
    double* ext = work;            /* m primitive states */
    double* cons = ext + 5 * m;    /* m conserved states */
    double* p0 = cons + 5 * m;     /* m evolved left-face states */
    double* p1 = p0 + 5 * m;       /* m evolved right-face states */
    double* fx = p1 + 5 * m;       /* m - 1 interface fluxes */
    for (int k = 0; k < m; ++k) {
        int c = k - ng;
        if (order == 2) {
            if (c == -2) c = 1;
            else if (c == -1) c = 0;
            else if (c == n) c = n - 1;
            else if (c == n + 1) c = n - 2;
        } else {
            if (c < 0) c = 0;
            else if (c >= n) c = n - 1;
        }
        memcpy(ext + 5 * k, line + (long)c * stride, 5 * sizeof(double));
        p2c(phys, ext + 5 * k, cons + 5 * k, g);
    }
    if (order == 2) {
        for (int k = 1; k < m - 1; ++k)
            half_step_cons(phys, axis, cons + 5 * (k - 1), cons + 5 * k, cons + 5 * (k + 1), dt_dx, bias, limiter, g,
                           p0 + 5 * k, p1 + 5 * k);
        /* interface between ext k and k + 1, k = 1 .. m - 3 */
        for (int k = 1; k < m - 2; ++k)
            force_flux_states(phys, axis, p1 + 5 * k, p0 + 5 * (k + 1), dt_dx, dx_dt, g, fx + 5 * k);
    } else {
        /* the flux of each cell's conserved state once (p1 is free in FORCE) */
        for (int k = 0; k < m; ++k) flux(phys, axis, cons + 5 * k, p1 + 5 * k, g);
        for (int k = 0; k < m - 1; ++k)
            force_flux_cons(phys, axis, cons + 5 * k, p1 + 5 * k, cons + 5 * (k + 1), p1 + 5 * (k + 1), dt_dx,
                            dx_dt, g, fx + 5 * k);
    }
    for (int j = 0; j < n; ++j) {
        int k = j + ng; /* ext index of cell j; left interface k - 1, right interface k */
        const double* u = cons + 5 * k;
        const double* fl = fx + 5 * (k - 1);
        const double* fr = fx + 5 * k;
        double un[5];
        for (int c = 0; c < 5; ++c) un[c] = u[c] + (dt_dx * (fl[c] - fr[c]));
        c2p(phys, un, line + (long)j * stride, g);
    }
}

int orc_sweep(int physics, int order, int axis, int dim, double* prim, int nx, int ny, double dx, double dt,
              double bias, int limiter, double gamma) {
    if (dim == 1) ny = 1, axis = 0;
    int n = axis == 0 ? nx : ny;
    int lines = axis == 0 ? ny : nx;
    if (n < 2) return -1;
    long stride = axis == 0 ? 5 : 5L * nx;
    long line_step = axis == 0 ? 5L * nx : 5;
    double dt_dx = dt / dx, dx_dt = dx / dt;
#pragma omp parallel num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    {
        double* work = (double*)malloc(sizeof(double) * 25 * (size_t)(n + 4));
#pragma omp for schedule(static)
        for (int l = 0; l < lines; ++l) {
            if (axis == 0) {
                update_line(physics, order, axis, prim + (long)l * line_step, stride, n, dt_dx, dx_dt, bias, limiter,
                            gamma, work);
            } else {
                /* gather the column so that the line kernel works on contiguous memory */
                double* col = (double*)malloc(sizeof(double) * 5 * (size_t)n);
                for (int i = 0; i < n; ++i) memcpy(col + 5 * i, prim + (long)l * line_step + (long)i * stride, 40);
                update_line(physics, order, axis, col, 5, n, dt_dx, dx_dt, bias, limiter, gamma, work);
                for (int i = 0; i < n; ++i) memcpy(prim + (long)l * line_step + (long)i * stride, col + 5 * i, 40);
                free(col);
            }
        }
        free(work);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Wave speed and time step                                                 */
/* ------------------------------------------------------------------------ */

/* EulerSolvers::computeMaximumWaveSpeed{,2D} (Solvers/EulerSolvers.cpp:87-128):
 * 1D |u| + |c|; 2D max(|u|, |v| + |c|) (the reference's grouping).
 * RelativisticEulerSolvers::computeMaximumWaveSpeed{,2D}
 * (Solvers/RelativisticEulerSolvers.cpp:89-130): 2D max(|u|, |v|) + |c|.
 * The running maximum starts at 0 and skips NaN (waveSpeed > max is false). */
static double cell_wavespeed(int phys, int dim, const double* W, double g) {
    double c = phys == 0 ? nw_sound_speed(W, g) : sr_sound_speed(W, g);
    if (dim == 1) return fabs(W[1]) + fabs(c);
    if (phys == 0) return std_max(fabs(W[1]), fabs(W[2]) + fabs(c));
    return std_max(fabs(W[1]), fabs(W[2])) + fabs(c);
}

double orc_max_wavespeed(int physics, int dim, const double* prim, int nx, int ny, double gamma) {
    long cells = (long)nx * (dim == 1 ? 1 : ny);
    double best = 0.0;
#pragma omp parallel num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    {
        double local = 0.0;
#pragma omp for schedule(static)
        for (long c = 0; c < cells; ++c) {
            double ws = cell_wavespeed(physics, dim, prim + 5 * c, gamma);
            if (ws > local) local = ws;
        }
#pragma omp critical
        {
            if (local > best) best = local;
        }
    }
    return best;
}

/* EulerSolvers::computeStableTimeStep(dt, t, T, n) (Solvers/EulerSolvers.cpp:130-145). */
double orc_stable_dt_scalar(double dt, double t, double tf, int iteration) {
    double d = dt;
    if (iteration <= 5) d *= 0.2;
    if ((t + d) > tf) d = tf - t;
    return d;
}

/* computeStableTimeStep{,2D} (Solvers/EulerSolvers.cpp:147-161). */
double orc_stable_dt(int physics, int dim, const double* prim, int nx, int ny, double dx, double cfl, double t,
                     double tf, int iteration, double gamma) {
    double dt = cfl * (dx / orc_max_wavespeed(physics, dim, prim, nx, ny, gamma));
    return orc_stable_dt_scalar(dt, t, tf, iteration);
}

/* solve / solve2D loops (Solvers/EulerSecondOrderSolver.cpp:133-178,
 * Solvers/EulerFirstOrderSolver.cpp:220-265): Strang X(dt/2) Y(dt) X(dt/2)
 * in 2D, capped at max_steps iterations. */
int orc_run(int physics, int order, int dim, double* prim, int nx, int ny, double dx, double cfl, double tf,
            int max_steps, double bias, int limiter, double gamma, double* t_out) {
    double t = 0.0;
    int it = 0;
    if (dim == 1) ny = 1;
    while (t < tf && it < max_steps) {
        double dt = orc_stable_dt(physics, dim, prim, nx, ny, dx, cfl, t, tf, it, gamma);
        if (dim == 1) {
            orc_sweep(physics, order, 0, 1, prim, nx, 1, dx, dt, bias, limiter, gamma);
        } else {
            orc_sweep(physics, order, 0, 2, prim, nx, ny, dx, 0.5 * dt, bias, limiter, gamma);
            orc_sweep(physics, order, 1, 2, prim, nx, ny, dx, dt, bias, limiter, gamma);
            orc_sweep(physics, order, 0, 2, prim, nx, ny, dx, 0.5 * dt, bias, limiter, gamma);
        }
        t += dt;
        it += 1;
    }
    if (t_out) *t_out = t;
    return it;
}
