// Flat-array driver around the reference's own solver classes.
//
// This file is test infrastructure: it is compiled together with the
// reference's unmodified C++ sources (see oracle/Makefile) into
// oracle/_ref/libcosmos_ref.so, and only tests, golden-fixture scripts and the
// CPU-baseline leg of bench.py load it.  It never ships in the product path.
//
// Every entry point takes primitive states as a flat row-major AoS array
// prim[(i * nx + j) * 5 + k], i = row (y), j = column (x), k in
// (rho, u, v, w, p), converts them to the reference's state-vector objects,
// calls the reference's static method, and converts back.  Physics codes:
// 0 = Newtonian (Euler/), 1 = special relativistic (RelativisticEuler/).

#include <cstring>
#include <iostream>
#include <sstream>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "Solvers/EulerFirstOrderSolver.hpp"
#include "Solvers/EulerSecondOrderSolver.hpp"
#include "Solvers/RelativisticEulerFirstOrderSolver.hpp"
#include "Solvers/RelativisticEulerSecondOrderSolver.hpp"

namespace {

// The reference's solve/solve2D print one status line per iteration
// (Solvers/EulerSolvers.cpp:266-269); silence it while a driver call runs.
struct QuietCout {
    std::ostringstream sink;
    std::streambuf* saved;
    QuietCout() : saved(std::cout.rdbuf(sink.rdbuf())) {}
    ~QuietCout() { std::cout.rdbuf(saved); }
};

struct NewtonianTraits {
    typedef EulerStateVector State;
    typedef EulerSolvers Helpers;
    typedef EulerFirstOrderSolver First;
    typedef EulerSecondOrderSolver Second;
};

struct RelativisticTraits {
    typedef RelativisticEulerStateVector State;
    typedef RelativisticEulerSolvers Helpers;
    typedef RelativisticEulerFirstOrderSolver First;
    typedef RelativisticEulerSecondOrderSolver Second;
};

template <class S>
S make_state(const double* w) { return S(w[0], w[1], w[2], w[3], w[4]); }

template <class S>
void store_state(S& s, double* w) {
    w[0] = s.getDensity(); w[1] = s.getXVelocity(); w[2] = s.getYVelocity();
    w[3] = s.getZVelocity(); w[4] = s.getPressure();
}

template <class S>
std::vector<S> to_1d(const double* prim, int n) {
    std::vector<S> cells;
    cells.reserve(n);
    for (int j = 0; j < n; ++j) cells.push_back(make_state<S>(prim + 5 * j));
    return cells;
}

template <class S>
void from_1d(std::vector<S>& cells, double* prim) {
    for (size_t j = 0; j < cells.size(); ++j) store_state(cells[j], prim + 5 * j);
}

template <class S>
std::vector<std::vector<S> > to_2d(const double* prim, int nx, int ny) {
    std::vector<std::vector<S> > cells(ny);
    for (int i = 0; i < ny; ++i) cells[i] = to_1d<S>(prim + 5 * (size_t)i * nx, nx);
    return cells;
}

template <class S>
void from_2d(std::vector<std::vector<S> >& cells, double* prim) {
    size_t nx = cells[0].size();
    for (size_t i = 0; i < cells.size(); ++i) from_1d(cells[i], prim + 5 * i * nx);
}

template <class T>
void sweep(int order, int axis, int dim, double* prim, int nx, int ny, double dx, double dt,
           double bias, int limiter, EulerMaterialParameters mp) {
    typedef typename T::State S;
    if (dim == 1) {
        std::vector<S> cells = to_1d<S>(prim, nx);
        if (order == 1) T::First::computeFORCETimeStep(cells, dx, dt, mp);
        else T::Second::computeSLICTimeStep(cells, dx, dt, bias, limiter, mp);
        from_1d(cells, prim);
        return;
    }
    std::vector<std::vector<S> > cells = to_2d<S>(prim, nx, ny);
    if (order == 1) {
        if (axis == 0) T::First::computeXFORCETimeStep2D(cells, dx, dt, mp);
        else T::First::computeYFORCETimeStep2D(cells, dx, dt, mp);
    } else {
        if (axis == 0) T::Second::computeXSLICTimeStep2D(cells, dx, dt, bias, limiter, mp);
        else T::Second::computeYSLICTimeStep2D(cells, dx, dt, bias, limiter, mp);
    }
    from_2d(cells, prim);
}

template <class T>
double max_wavespeed(int dim, const double* prim, int nx, int ny, EulerMaterialParameters mp) {
    typedef typename T::State S;
    if (dim == 1) {
        std::vector<S> cells = to_1d<S>(prim, nx);
        return T::Helpers::computeMaximumWaveSpeed(cells, mp);
    }
    std::vector<std::vector<S> > cells = to_2d<S>(prim, nx, ny);
    return T::Helpers::computeMaximumWaveSpeed2D(cells, mp);
}

template <class T>
double stable_dt(int dim, const double* prim, int nx, int ny, double dx, double cfl, double t,
                 double tf, int iteration, EulerMaterialParameters mp) {
    typedef typename T::State S;
    if (dim == 1) {
        std::vector<S> cells = to_1d<S>(prim, nx);
        return T::Helpers::computeStableTimeStep(cells, dx, cfl, t, tf, iteration, mp);
    }
    std::vector<std::vector<S> > cells = to_2d<S>(prim, nx, ny);
    return T::Helpers::computeStableTimeStep2D(cells, dx, cfl, t, tf, iteration, mp);
}

// The loop of solve/solve2D (Solvers/EulerSecondOrderSolver.cpp:133-178 and
// Solvers/EulerFirstOrderSolver.cpp:220-265), capped at max_steps iterations
// so that a fixed step count can be run (pass finalTime = +inf for no cap).
template <class T>
int run(int order, int dim, double* prim, int nx, int ny, double dx, double cfl, double tf,
        int max_steps, double bias, int limiter, EulerMaterialParameters mp, double* t_out) {
    typedef typename T::State S;
    double t = 0.0;
    int it = 0;
    if (dim == 1) {
        std::vector<S> cells = to_1d<S>(prim, nx);
        while (t < tf && it < max_steps) {
            double dt = T::Helpers::computeStableTimeStep(cells, dx, cfl, t, tf, it, mp);
            if (order == 1) T::First::computeFORCETimeStep(cells, dx, dt, mp);
            else T::Second::computeSLICTimeStep(cells, dx, dt, bias, limiter, mp);
            t += dt;
            it += 1;
        }
        from_1d(cells, prim);
    } else {
        std::vector<std::vector<S> > cells = to_2d<S>(prim, nx, ny);
        while (t < tf && it < max_steps) {
            double dt = T::Helpers::computeStableTimeStep2D(cells, dx, cfl, t, tf, it, mp);
            if (order == 1) {
                T::First::computeXFORCETimeStep2D(cells, dx, dt * 0.5, mp);
                T::First::computeYFORCETimeStep2D(cells, dx, dt, mp);
                T::First::computeXFORCETimeStep2D(cells, dx, dt * 0.5, mp);
            } else {
                T::Second::computeXSLICTimeStep2D(cells, dx, 0.5 * dt, bias, limiter, mp);
                T::Second::computeYSLICTimeStep2D(cells, dx, dt, bias, limiter, mp);
                T::Second::computeXSLICTimeStep2D(cells, dx, 0.5 * dt, bias, limiter, mp);
            }
            t += dt;
            it += 1;
        }
        from_2d(cells, prim);
    }
    if (t_out) *t_out = t;
    return it;
}

// The reference's own solve/solve2D, run to finalTime.
template <class T>
void solve(int order, int dim, double* prim, int nx, int ny, double dx, double cfl, double tf,
           double bias, int limiter, EulerMaterialParameters mp) {
    typedef typename T::State S;
    QuietCout quiet;
    if (dim == 1) {
        std::vector<S> cells = to_1d<S>(prim, nx);
        std::vector<S> out = (order == 1) ? T::First::solve(cells, dx, cfl, tf, mp)
                                          : T::Second::solve(cells, dx, cfl, tf, bias, limiter, mp);
        from_1d(out, prim);
    } else {
        std::vector<std::vector<S> > cells = to_2d<S>(prim, nx, ny);
        std::vector<std::vector<S> > out =
            (order == 1) ? T::First::solve2D(cells, dx, cfl, tf, mp)
                         : T::Second::solve2D(cells, dx, cfl, tf, bias, limiter, mp);
        from_2d(out, prim);
    }
}

template <class T>
void slic_flux(int axis, const double* ll, const double* l, const double* r, const double* rr,
               double dx, double dt, double bias, int limiter, EulerMaterialParameters mp,
               double* out) {
    typedef typename T::State S;
    std::vector<double> f =
        axis == 0 ? T::Second::computeXSLICFlux(make_state<S>(ll), make_state<S>(l), make_state<S>(r),
                                                make_state<S>(rr), dx, dt, bias, limiter, mp)
                  : T::Second::computeYSLICFlux(make_state<S>(ll), make_state<S>(l), make_state<S>(r),
                                                make_state<S>(rr), dx, dt, bias, limiter, mp);
    for (int k = 0; k < 5; ++k) out[k] = f[k];
}

template <class T>
void force_flux(int axis, const double* l, const double* r, double dx, double dt,
                EulerMaterialParameters mp, double* out) {
    typedef typename T::State S;
    std::vector<double> f = axis == 0 ? T::First::computeXFORCEFlux(make_state<S>(l), make_state<S>(r), dx, dt, mp)
                                      : T::First::computeYFORCEFlux(make_state<S>(l), make_state<S>(r), dx, dt, mp);
    for (int k = 0; k < 5; ++k) out[k] = f[k];
}

template <class T>
void half_step(int axis, const double* l, const double* m, const double* r, double dx, double dt,
               double bias, int limiter, int side, EulerMaterialParameters mp, double* out) {
    typedef typename T::State S;
    S s = axis == 0 ? T::Helpers::evolveStateByHalfXTimeStep(make_state<S>(l), make_state<S>(m), make_state<S>(r),
                                                             dx, dt, bias, limiter, side, mp)
                    : T::Helpers::evolveStateByHalfYTimeStep(make_state<S>(l), make_state<S>(m), make_state<S>(r),
                                                             dx, dt, bias, limiter, side, mp);
    store_state(s, out);
}

template <class S>
void p2c(const double* w, double* u, EulerMaterialParameters mp) {
    S s = make_state<S>(w);
    std::vector<double> c = s.computeConservedVariableVector(mp);
    for (int k = 0; k < 5; ++k) u[k] = c[k];
}

template <class S>
void c2p(const double* u, double* w, EulerMaterialParameters mp) {
    S s;
    s.setConservedVariableVector(std::vector<double>(u, u + 5), mp);
    store_state(s, w);
}

template <class S>
void flux(int axis, const double* u, double* f, EulerMaterialParameters mp) {
    std::vector<double> c(u, u + 5);
    std::vector<double> r = axis == 0 ? S::computeXFluxVector(c, mp) : S::computeYFluxVector(c, mp);
    for (int k = 0; k < 5; ++k) f[k] = r[k];
}

}  // namespace

extern "C" {

int ref_sweep(int physics, int order, int axis, int dim, double* prim, int nx, int ny, double dx,
              double dt, double bias, int limiter, double gamma) {
    EulerMaterialParameters mp(gamma);
    if (physics == 0) sweep<NewtonianTraits>(order, axis, dim, prim, nx, ny, dx, dt, bias, limiter, mp);
    else sweep<RelativisticTraits>(order, axis, dim, prim, nx, ny, dx, dt, bias, limiter, mp);
    return 0;
}

double ref_max_wavespeed(int physics, int dim, const double* prim, int nx, int ny, double gamma) {
    EulerMaterialParameters mp(gamma);
    return physics == 0 ? max_wavespeed<NewtonianTraits>(dim, prim, nx, ny, mp)
                        : max_wavespeed<RelativisticTraits>(dim, prim, nx, ny, mp);
}

double ref_stable_dt(int physics, int dim, const double* prim, int nx, int ny, double dx, double cfl,
                     double t, double tf, int iteration, double gamma) {
    EulerMaterialParameters mp(gamma);
    return physics == 0 ? stable_dt<NewtonianTraits>(dim, prim, nx, ny, dx, cfl, t, tf, iteration, mp)
                        : stable_dt<RelativisticTraits>(dim, prim, nx, ny, dx, cfl, t, tf, iteration, mp);
}

double ref_stable_dt_scalar(double dt, double t, double tf, int iteration) {
    return EulerSolvers::computeStableTimeStep(dt, t, tf, iteration);
}

int ref_run(int physics, int order, int dim, double* prim, int nx, int ny, double dx, double cfl,
            double tf, int max_steps, double bias, int limiter, double gamma, double* t_out) {
    EulerMaterialParameters mp(gamma);
    return physics == 0
               ? run<NewtonianTraits>(order, dim, prim, nx, ny, dx, cfl, tf, max_steps, bias, limiter, mp, t_out)
               : run<RelativisticTraits>(order, dim, prim, nx, ny, dx, cfl, tf, max_steps, bias, limiter, mp, t_out);
}

int ref_solve(int physics, int order, int dim, double* prim, int nx, int ny, double dx, double cfl,
              double tf, double bias, int limiter, double gamma) {
    EulerMaterialParameters mp(gamma);
    if (physics == 0) solve<NewtonianTraits>(order, dim, prim, nx, ny, dx, cfl, tf, bias, limiter, mp);
    else solve<RelativisticTraits>(order, dim, prim, nx, ny, dx, cfl, tf, bias, limiter, mp);
    return 0;
}

void ref_slic_flux(int physics, int axis, const double* ll, const double* l, const double* r,
                   const double* rr, double dx, double dt, double bias, int limiter, double gamma,
                   double* out) {
    EulerMaterialParameters mp(gamma);
    if (physics == 0) slic_flux<NewtonianTraits>(axis, ll, l, r, rr, dx, dt, bias, limiter, mp, out);
    else slic_flux<RelativisticTraits>(axis, ll, l, r, rr, dx, dt, bias, limiter, mp, out);
}

void ref_force_flux(int physics, int axis, const double* l, const double* r, double dx, double dt,
                    double gamma, double* out) {
    EulerMaterialParameters mp(gamma);
    if (physics == 0) force_flux<NewtonianTraits>(axis, l, r, dx, dt, mp, out);
    else force_flux<RelativisticTraits>(axis, l, r, dx, dt, mp, out);
}

void ref_half_step(int physics, int axis, const double* l, const double* m, const double* r, double dx,
                   double dt, double bias, int limiter, int side, double gamma, double* out) {
    EulerMaterialParameters mp(gamma);
    if (physics == 0) half_step<NewtonianTraits>(axis, l, m, r, dx, dt, bias, limiter, side, mp, out);
    else half_step<RelativisticTraits>(axis, l, m, r, dx, dt, bias, limiter, side, mp, out);
}

void ref_p2c(int physics, const double* w, double* u, double gamma) {
    EulerMaterialParameters mp(gamma);
    if (physics == 0) p2c<EulerStateVector>(w, u, mp);
    else p2c<RelativisticEulerStateVector>(w, u, mp);
}

void ref_c2p(int physics, const double* u, double* w, double gamma) {
    EulerMaterialParameters mp(gamma);
    if (physics == 0) c2p<EulerStateVector>(u, w, mp);
    else c2p<RelativisticEulerStateVector>(u, w, mp);
}

void ref_flux(int physics, int axis, const double* u, double* f, double gamma) {
    EulerMaterialParameters mp(gamma);
    if (physics == 0) flux<EulerStateVector>(axis, u, f, mp);
    else flux<RelativisticEulerStateVector>(axis, u, f, mp);
}

double ref_limiter(double r, double bias, int limiter) {
    return SlopeLimiters::computeSlopeLimiter(r, bias, limiter);
}

void ref_slope(const double* ul, const double* um, const double* ur, double bias, int limiter, double* out) {
    std::vector<double> s = SlopeLimiters::computeSlopeVector(std::vector<double>(ul, ul + 5),
                                                              std::vector<double>(um, um + 5),
                                                              std::vector<double>(ur, ur + 5), bias, limiter);
    for (int k = 0; k < 5; ++k) out[k] = s[k];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// General relativistic (BlackHoleEuler/): the same flat layout; the black hole
// is BlackHoleSpacetime(M, a, x, y, z) (BlackHoleEuler/BlackHoleSpacetime.cpp:43-52).
// ---------------------------------------------------------------------------
#include "Solvers/BlackHoleEulerFirstOrderSolver.hpp"
#include "Solvers/BlackHoleEulerForcingSolver.hpp"
#include "Solvers/BlackHoleEulerSecondOrderSolver.hpp"

// Per-step NaN cell counts of ref_gr_run (a diagnostic of this driver): while a
// buffer is set, step k's count of cells holding a NaN goes to buf[k].
static int* g_step_nans = nullptr;
static int g_step_nans_cap = 0;

template <class S>
static int count_nan_cells(std::vector<S>& cells) {
    int n = 0;
    for (size_t j = 0; j < cells.size(); ++j) {
        double w[5];
        store_state(cells[j], w);
        bool bad = false;
        for (int k = 0; k < 5; ++k) bad = bad || (w[k] != w[k]);
        n += bad ? 1 : 0;
    }
    return n;
}

template <class S>
static void record_step_nans(std::vector<S>& cells, int step) {
    if (g_step_nans && step < g_step_nans_cap) g_step_nans[step] = count_nan_cells(cells);
}

template <class S>
static void record_step_nans(std::vector<std::vector<S> >& cells, int step) {
    if (!g_step_nans || step >= g_step_nans_cap) return;
    int n = 0;
    for (size_t i = 0; i < cells.size(); ++i) n += count_nan_cells(cells[i]);
    g_step_nans[step] = n;
}

extern "C" {

void ref_set_step_nan_counts(int* buf, int capacity) {
    g_step_nans = buf;
    g_step_nans_cap = buf ? capacity : 0;
}

int ref_gr_sweep(int order, int axis, int dim, double* prim, int nx, int ny, double h, double dt, double bias,
                 int lim, double gamma, double M, double a, double bx, double by, double bz) {
    EulerMaterialParameters mp(gamma);
    BlackHoleSpacetime bh(M, a, bx, by, bz);
    typedef BlackHoleEulerStateVector S;
    if (dim == 1) {
        std::vector<S> c = to_1d<S>(prim, nx);
        if (order == 1) BlackHoleEulerFirstOrderSolver::computeFORCETimeStep(c, h, dt, mp, bh);
        else BlackHoleEulerSecondOrderSolver::computeSLICTimeStep(c, h, dt, bias, lim, mp, bh);
        from_1d(c, prim);
        return 0;
    }
    std::vector<std::vector<S> > c = to_2d<S>(prim, nx, ny);
    if (order == 1) {
        if (axis == 0) BlackHoleEulerFirstOrderSolver::computeXFORCETimeStep2D(c, h, dt, mp, bh);
        else BlackHoleEulerFirstOrderSolver::computeYFORCETimeStep2D(c, h, dt, mp, bh);
    } else {
        if (axis == 0) BlackHoleEulerSecondOrderSolver::computeXSLICTimeStep2D(c, h, dt, bias, lim, mp, bh);
        else BlackHoleEulerSecondOrderSolver::computeYSLICTimeStep2D(c, h, dt, bias, lim, mp, bh);
    }
    from_2d(c, prim);
    return 0;
}

int ref_gr_forcing(int dim, double* prim, int nx, int ny, double h, double dt, double bias, int lim, double gamma,
                   double M, double a, double bx, double by, double bz) {
    EulerMaterialParameters mp(gamma);
    BlackHoleSpacetime bh(M, a, bx, by, bz);
    typedef BlackHoleEulerStateVector S;
    if (dim == 1) {
        std::vector<S> c = to_1d<S>(prim, nx);
        BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep(c, h, dt, bias, lim, mp, bh);
        from_1d(c, prim);
        return 0;
    }
    std::vector<std::vector<S> > c = to_2d<S>(prim, nx, ny);
    BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep2D(c, h, dt, bias, lim, mp, bh);
    from_2d(c, prim);
    return 0;
}

double ref_gr_max_wavespeed(int dim, const double* prim, int nx, int ny, double gamma, double M, double a, double bx,
                            double by, double bz) {
    EulerMaterialParameters mp(gamma);
    BlackHoleSpacetime bh(M, a, bx, by, bz);
    typedef BlackHoleEulerStateVector S;
    if (dim == 1) {
        std::vector<S> c = to_1d<S>(prim, nx);
        return BlackHoleEulerSolvers::computeMaximumWaveSpeed(c, mp, bh);
    }
    std::vector<std::vector<S> > c = to_2d<S>(prim, nx, ny);
    return BlackHoleEulerSolvers::computeMaximumWaveSpeed2D(c, mp, bh);
}

// The solve / solve2D loops (Solvers/BlackHoleEulerSecondOrderSolver.cpp:145-218,
// Solvers/BlackHoleEulerFirstOrderSolver.cpp:183-253), capped at max_steps.
int ref_gr_run(int order, int dim, double* prim, int nx, int ny, double h, double cfl, double tf, int max_steps,
               double bias, int lim, int subcycles, double gamma, double M, double a, double bx, double by, double bz,
               double* t_out) {
    EulerMaterialParameters mp(gamma);
    BlackHoleSpacetime bh(M, a, bx, by, bz);
    typedef BlackHoleEulerStateVector S;
    double t = 0.0;
    int it = 0;
    if (dim == 1) {
        std::vector<S> c = to_1d<S>(prim, nx);
        while (t < tf && it < max_steps) {
            double dt = BlackHoleEulerSolvers::computeStableTimeStep(c, h, cfl, t, tf, it, mp, bh);
            for (int s = 0; s < subcycles; ++s)
                BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep(c, h, 0.5 * (dt / subcycles), 0.0, 1, mp, bh);
            if (order == 1) BlackHoleEulerFirstOrderSolver::computeFORCETimeStep(c, h, dt, mp, bh);
            else BlackHoleEulerSecondOrderSolver::computeSLICTimeStep(c, h, dt, bias, lim, mp, bh);
            for (int s = 0; s < subcycles; ++s)
                BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep(c, h, 0.5 * (dt / subcycles), 0.0, 1, mp, bh);
            t += dt;
            record_step_nans(c, it);
            it += 1;
        }
        from_1d(c, prim);
    } else {
        std::vector<std::vector<S> > c = to_2d<S>(prim, nx, ny);
        while (t < tf && it < max_steps) {
            double dt = BlackHoleEulerSolvers::computeStableTimeStep2D(c, h, cfl, t, tf, it, mp, bh);
            for (int s = 0; s < subcycles; ++s)
                BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep2D(c, h, 0.5 * (dt / subcycles), 0.0, 1, mp, bh);
            if (order == 1) {
                BlackHoleEulerFirstOrderSolver::computeXFORCETimeStep2D(c, h, dt * 0.5, mp, bh);
                BlackHoleEulerFirstOrderSolver::computeYFORCETimeStep2D(c, h, dt, mp, bh);
                BlackHoleEulerFirstOrderSolver::computeXFORCETimeStep2D(c, h, dt * 0.5, mp, bh);
            } else {
                BlackHoleEulerSecondOrderSolver::computeXSLICTimeStep2D(c, h, dt * 0.5, bias, lim, mp, bh);
                BlackHoleEulerSecondOrderSolver::computeYSLICTimeStep2D(c, h, dt, bias, lim, mp, bh);
                BlackHoleEulerSecondOrderSolver::computeXSLICTimeStep2D(c, h, dt * 0.5, bias, lim, mp, bh);
            }
            for (int s = 0; s < subcycles; ++s)
                BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep2D(c, h, 0.5 * (dt / subcycles), 0.0, 1, mp, bh);
            t += dt;
            record_step_nans(c, it);
            it += 1;
        }
        from_2d(c, prim);
    }
    if (t_out) *t_out = t;
    return it;
}

// what: 0 p2c, 1 c2p, 2 x flux, 3 y flux, 4 source term, 5 sound speed,
// 6 metric pieces (H, l[3], alpha, beta[3], det gamma, gamma^{ij}[9],
// g^{mu nu}[16], K[9], excised) -- 44 values.
void ref_gr_point(int what, const double* in, double x, double y, double z, double h, double gamma, double M,
                  double a, double bx, double by, double bz, double* out) {
    EulerMaterialParameters mp(gamma);
    BlackHoleSpacetime bh(M, a, bx, by, bz);
    std::vector<double> v(in, in + 5), r;
    BlackHoleEulerStateVector s(in[0], in[1], in[2], in[3], in[4]);
    switch (what) {
        case 0: r = s.computeConservedVariableVector(x, y, z, mp, bh); break;
        case 1: {
            BlackHoleEulerStateVector t;
            t.setConservedVariableVector(v, x, y, z, mp, bh);
            store_state(t, out);
            return;
        }
        case 2: r = BlackHoleEulerStateVector::computeXFluxVector(v, x, y, z, mp, bh); break;
        case 3: r = BlackHoleEulerStateVector::computeYFluxVector(v, x, y, z, mp, bh); break;
        case 4: r = BlackHoleEulerStateVector::computeSourceTermVector(v, x, y, z, h, mp, bh); break;
        case 5: out[0] = s.computeSoundSpeed(x, y, z, mp, bh); return;
        case 6: {
            out[0] = bh.computeKerrSchildScalar(x, y, z);
            std::vector<double> l = bh.computeKerrSchildVector(x, y, z);
            out[1] = l[0]; out[2] = l[1]; out[3] = l[2];
            out[4] = bh.computeLapseFunction(x, y, z);
            std::vector<double> be = bh.computeShiftVector(x, y, z);
            out[5] = be[0]; out[6] = be[1]; out[7] = be[2];
            out[8] = bh.computeSpatialMetricDeterminant(x, y, z);
            std::vector<std::vector<double> > gi = bh.computeInverseSpatialMetricTensor(x, y, z);
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[9 + 3 * i + j] = gi[i][j];
            std::vector<std::vector<double> > g4 = bh.computeInverseSpacetimeMetricTensor(x, y, z);
            for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) out[18 + 4 * i + j] = g4[i][j];
            std::vector<std::vector<double> > K = bh.computeExtrinsicCurvatureTensor(x, y, z, h);
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[34 + 3 * i + j] = K[i][j];
            out[43] = bh.inExcisionRegion(x, y, z) ? 1.0 : 0.0;
            return;
        }
        default: return;
    }
    for (size_t k = 0; k < r.size(); ++k) out[k] = r[k];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// AMR level 1 (AMR/EulerAMR.cpp:169-590, AMR/RelativisticEulerAMR.cpp): the reference's
// own solve{,2D}Level1AMR, run to final_time; coarse [ny][nx], fine [2ny][2nx],
// flags as 0/1 bytes.
// ---------------------------------------------------------------------------
#include "AMR/EulerAMR.hpp"
#include "AMR/RelativisticEulerAMR.hpp"

namespace {
template <class AMR, class S>
void amr_level1(int dim, int order, const double* prim, int nx, int ny, double dx, double cfl, double tf, double tol,
                double gamma, double* coarse, double* fine, unsigned char* cflags, unsigned char* fflags) {
    QuietCout quiet;
    EulerMaterialParameters mp(gamma);
    if (dim == 1) {
        std::vector<S> c = to_1d<S>(prim, nx);
        auto r = AMR::solveLevel1AMR(c, dx, cfl, tf, tol, order, mp);
        from_1d(std::get<0>(r), coarse);
        from_1d(std::get<1>(r), fine);
        for (int i = 0; i < nx; ++i) cflags[i] = std::get<2>(r)[i] ? 1 : 0;
        for (int i = 0; i < 2 * nx; ++i) fflags[i] = std::get<3>(r)[i] ? 1 : 0;
        return;
    }
    std::vector<std::vector<S> > c = to_2d<S>(prim, nx, ny);
    auto r = AMR::solve2DLevel1AMR(c, dx, cfl, tf, tol, order, mp);
    from_2d(std::get<0>(r), coarse);
    from_2d(std::get<1>(r), fine);
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) cflags[(size_t)i * nx + j] = std::get<2>(r)[i][j] ? 1 : 0;
    for (int i = 0; i < 2 * ny; ++i)
        for (int j = 0; j < 2 * nx; ++j) fflags[(size_t)i * 2 * nx + j] = std::get<3>(r)[i][j] ? 1 : 0;
}
}  // namespace

extern "C" int ref_amr_level1(int physics, int dim, int order, const double* prim, int nx, int ny, double dx,
                              double cfl, double tf, double tol, double gamma, double* coarse, double* fine,
                              unsigned char* cflags, unsigned char* fflags) {
    if (physics == 0)
        amr_level1<EulerAMR, EulerStateVector>(dim, order, prim, nx, ny, dx, cfl, tf, tol, gamma, coarse, fine,
                                               cflags, fflags);
    else
        amr_level1<RelativisticEulerAMR, RelativisticEulerStateVector>(dim, order, prim, nx, ny, dx, cfl, tf, tol,
                                                                       gamma, coarse, fine, cflags, fflags);
    return 0;
}

// AMR level 2 (AMR/EulerAMR.cpp:593-1381, AMR/RelativisticEulerAMR.cpp:605-1394): the
// reference's own solve{,2D}Level2AMR; coarse [ny][nx], intermediate [2ny][2nx],
// fine [4ny][4nx] (1D: n, 2n, 4n cells).
namespace {
template <class V>
void flags_out(const std::vector<V>& f, unsigned char* out) {
    for (size_t i = 0; i < f.size(); ++i) out[i] = f[i] ? 1 : 0;
}
template <class V>
void flags_out(const std::vector<std::vector<V> >& f, unsigned char* out) {
    size_t k = 0;
    for (const auto& row : f)
        for (bool v : row) out[k++] = v ? 1 : 0;
}

template <class AMR, class S>
void amr_level2(int dim, int order, const double* prim, int nx, int ny, double dx, double cfl, double tf, double tol,
                double gamma, double* coarse, double* inter, double* fine, unsigned char* cflags,
                unsigned char* iflags, unsigned char* fflags) {
    QuietCout quiet;
    EulerMaterialParameters mp(gamma);
    if (dim == 1) {
        std::vector<S> c = to_1d<S>(prim, nx);
        auto r = AMR::solveLevel2AMR(c, dx, cfl, tf, tol, order, mp);
        from_1d(std::get<0>(r), coarse);
        from_1d(std::get<1>(r), inter);
        from_1d(std::get<2>(r), fine);
        flags_out(std::get<3>(r), cflags);
        flags_out(std::get<4>(r), iflags);
        flags_out(std::get<5>(r), fflags);
        return;
    }
    std::vector<std::vector<S> > c = to_2d<S>(prim, nx, ny);
    auto r = AMR::solve2DLevel2AMR(c, dx, cfl, tf, tol, order, mp);
    from_2d(std::get<0>(r), coarse);
    from_2d(std::get<1>(r), inter);
    from_2d(std::get<2>(r), fine);
    flags_out(std::get<3>(r), cflags);
    flags_out(std::get<4>(r), iflags);
    flags_out(std::get<5>(r), fflags);
}
}  // namespace

extern "C" int ref_amr_level2(int physics, int dim, int order, const double* prim, int nx, int ny, double dx,
                              double cfl, double tf, double tol, double gamma, double* coarse, double* inter,
                              double* fine, unsigned char* cflags, unsigned char* iflags, unsigned char* fflags) {
    if (physics == 0)
        amr_level2<EulerAMR, EulerStateVector>(dim, order, prim, nx, ny, dx, cfl, tf, tol, gamma, coarse, inter, fine,
                                               cflags, iflags, fflags);
    else
        amr_level2<RelativisticEulerAMR, RelativisticEulerStateVector>(dim, order, prim, nx, ny, dx, cfl, tf, tol,
                                                                       gamma, coarse, inter, fine, cflags, iflags,
                                                                       fflags);
    return 0;
}

// ---------------------------------------------------------------------------
// Black-hole AMR (AMR/BlackHoleEulerAMR.cpp:238-1055): the reference's own
// solveLevel1AMR, solve2DLevel1AMR and 1D solveLevel2AMR, run to final_time,
// on BlackHoleSpacetime(M, a, x, y, z); layouts as above.
// ---------------------------------------------------------------------------
#include "AMR/BlackHoleEulerAMR.hpp"

extern "C" int ref_gr_amr_level1(int order, int dim, const double* prim, int nx, int ny, double dx, double cfl,
                                 double tf, double tol, double gamma, double M, double a, double bx, double by,
                                 double bz, double* coarse, double* fine, unsigned char* cflags,
                                 unsigned char* fflags) {
    QuietCout quiet;
    EulerMaterialParameters mp(gamma);
    BlackHoleSpacetime bh(M, a, bx, by, bz);
    typedef BlackHoleEulerStateVector S;
    if (dim == 1) {
        std::vector<S> c = to_1d<S>(prim, nx);
        auto r = BlackHoleEulerAMR::solveLevel1AMR(c, dx, cfl, tf, tol, order, mp, bh);
        from_1d(std::get<0>(r), coarse);
        from_1d(std::get<1>(r), fine);
        flags_out(std::get<2>(r), cflags);
        flags_out(std::get<3>(r), fflags);
        return 0;
    }
    std::vector<std::vector<S> > c = to_2d<S>(prim, nx, ny);
    auto r = BlackHoleEulerAMR::solve2DLevel1AMR(c, dx, cfl, tf, tol, order, mp, bh);
    from_2d(std::get<0>(r), coarse);
    from_2d(std::get<1>(r), fine);
    flags_out(std::get<2>(r), cflags);
    flags_out(std::get<3>(r), fflags);
    return 0;
}

extern "C" int ref_gr_amr_level2(int order, const double* prim, int nx, double dx, double cfl, double tf, double tol,
                                 double gamma, double M, double a, double bx, double by, double bz, double* coarse,
                                 double* inter, double* fine, unsigned char* cflags, unsigned char* iflags,
                                 unsigned char* fflags) {
    QuietCout quiet;
    EulerMaterialParameters mp(gamma);
    BlackHoleSpacetime bh(M, a, bx, by, bz);
    std::vector<BlackHoleEulerStateVector> c = to_1d<BlackHoleEulerStateVector>(prim, nx);
    auto r = BlackHoleEulerAMR::solveLevel2AMR(c, dx, cfl, tf, tol, order, mp, bh);
    from_1d(std::get<0>(r), coarse);
    from_1d(std::get<1>(r), inter);
    from_1d(std::get<2>(r), fine);
    flags_out(std::get<3>(r), cflags);
    flags_out(std::get<4>(r), iflags);
    flags_out(std::get<5>(r), fflags);
    return 0;
}

// ---------------------------------------------------------------------------
// Point-level statics of the solver classes, one call each, numbered as
// ORC_PT_* in cosmos_oracle.h: s = up to five 5-vectors, par = {dx, dt, bias,
// step fraction, gamma}, black hole: spacetime = {M, a, bx, by, bz},
// pos = {x, y, z}.
// ---------------------------------------------------------------------------
namespace {

enum {
    PT_P2C, PT_C2P, PT_FLUX, PT_LF, PT_RICHTMYER, PT_FORCE, PT_SLIC, PT_HALF, PT_SLOPE, PT_FRACTIONAL, PT_FORCING,
    PT_VEC_LF, PT_VEC_RICHTMYER, PT_VEC_FORCE, PT_VEC_UPDATE, PT_VEC_EVOLUTION, PT_VEC_SLOPE, PT_LIMITER
};

std::vector<double> vec5(const double* v) { return std::vector<double>(v, v + 5); }

template <class T>
void point_sr(int fn, int axis, const double* s, const double* par, int lim, int side, double* out) {
    typedef typename T::State S;
    EulerMaterialParameters mp(par[4]);
    const double dx = par[0], dt = par[1], bias = par[2];
    S a = make_state<S>(s), b = make_state<S>(s + 5), c = make_state<S>(s + 10), d = make_state<S>(s + 15);
    std::vector<double> r;
    switch (fn) {
        case PT_P2C: p2c<S>(s, out, mp); return;
        case PT_C2P: c2p<S>(s, out, mp); return;
        case PT_FLUX: flux<S>(axis, s, out, mp); return;
        case PT_LF:
            r = axis == 0 ? T::First::computeXLaxFriedrichsFlux(a, b, dx, dt, mp)
                          : T::First::computeYLaxFriedrichsFlux(a, b, dx, dt, mp);
            break;
        case PT_RICHTMYER:
            r = axis == 0 ? T::First::computeXRichtmyerFlux(a, b, dx, dt, mp)
                          : T::First::computeYRichtmyerFlux(a, b, dx, dt, mp);
            break;
        case PT_FORCE:
            r = axis == 0 ? T::First::computeXFORCEFlux(a, b, dx, dt, mp) : T::First::computeYFORCEFlux(a, b, dx, dt, mp);
            break;
        case PT_SLIC:
            r = axis == 0 ? T::Second::computeXSLICFlux(a, b, c, d, dx, dt, bias, lim, mp)
                          : T::Second::computeYSLICFlux(a, b, c, d, dx, dt, bias, lim, mp);
            break;
        case PT_HALF: {
            S e = axis == 0 ? T::Helpers::evolveStateByHalfXTimeStep(a, b, c, dx, dt, bias, lim, side, mp)
                            : T::Helpers::evolveStateByHalfYTimeStep(a, b, c, dx, dt, bias, lim, side, mp);
            store_state(e, out);
            return;
        }
        case PT_SLOPE: r = SlopeLimiters::computeSlopeVector(a, b, c, bias, lim, mp); break;
        default: return;
    }
    for (int k = 0; k < 5; ++k) out[k] = r[k];
}

void point_bh(int fn, int axis, const double* s, const double* par, int lim, int side, const double* st,
              const double* pos, double* out) {
    typedef BlackHoleEulerStateVector S;
    EulerMaterialParameters mp(par[4]);
    BlackHoleSpacetime bh(st[0], st[1], st[2], st[3], st[4]);
    const double dx = par[0], dt = par[1], bias = par[2], frac = par[3];
    const double x = pos[0], y = pos[1], z = pos[2];
    S a = make_state<S>(s), b = make_state<S>(s + 5), c = make_state<S>(s + 10), d = make_state<S>(s + 15);
    std::vector<double> r;
    switch (fn) {
        case PT_P2C: r = a.computeConservedVariableVector(x, y, z, mp, bh); break;
        case PT_C2P: {
            S t;
            t.setConservedVariableVector(vec5(s), x, y, z, mp, bh);
            store_state(t, out);
            return;
        }
        case PT_FLUX:
            r = axis == 0 ? S::computeXFluxVector(vec5(s), x, y, z, mp, bh) : S::computeYFluxVector(vec5(s), x, y, z, mp, bh);
            break;
        case PT_LF:
            r = axis == 0 ? BlackHoleEulerFirstOrderSolver::computeXLaxFriedrichsFlux(a, b, x, y, z, dx, dt, mp, bh)
                          : BlackHoleEulerFirstOrderSolver::computeYLaxFriedrichsFlux(a, b, x, y, z, dx, dt, mp, bh);
            break;
        case PT_RICHTMYER:
            r = axis == 0 ? BlackHoleEulerFirstOrderSolver::computeXRichtmyerFlux(a, b, x, y, z, dx, dt, mp, bh)
                          : BlackHoleEulerFirstOrderSolver::computeYRichtmyerFlux(a, b, x, y, z, dx, dt, mp, bh);
            break;
        case PT_FORCE:
            r = axis == 0 ? BlackHoleEulerFirstOrderSolver::computeXFORCEFlux(a, b, x, y, z, dx, dt, mp, bh)
                          : BlackHoleEulerFirstOrderSolver::computeYFORCEFlux(a, b, x, y, z, dx, dt, mp, bh);
            break;
        case PT_SLIC:
            r = axis == 0
                    ? BlackHoleEulerSecondOrderSolver::computeXSLICFlux(a, b, c, d, dx, dt, bias, lim, x, y, z, mp, bh)
                    : BlackHoleEulerSecondOrderSolver::computeYSLICFlux(a, b, c, d, dx, dt, bias, lim, x, y, z, mp, bh);
            break;
        case PT_HALF: {
            S e = axis == 0 ? BlackHoleEulerSolvers::evolveStateByHalfXTimeStep(a, b, c, dx, dt, bias, lim, side, x, y, z,
                                                                               mp, bh)
                            : BlackHoleEulerSolvers::evolveStateByHalfYTimeStep(a, b, c, dx, dt, bias, lim, side, x, y, z,
                                                                               mp, bh);
            store_state(e, out);
            return;
        }
        case PT_SLOPE:
            r = axis == 0 ? SlopeLimiters::computeSlopeVectorX(a, b, c, dx, bias, lim, x, y, z, mp, bh)
                          : SlopeLimiters::computeSlopeVectorY(a, b, c, dx, bias, lim, x, y, z, mp, bh);
            break;
        case PT_FRACTIONAL: {
            S e = axis == 0 ? BlackHoleEulerSolvers::evolveStateByFractionalXTimeStep(frac, a, b, c, dx, dt, bias, lim, x,
                                                                                     y, z, mp, bh)
                            : BlackHoleEulerSolvers::evolveStateByFractionalYTimeStep(frac, a, b, c, dx, dt, bias, lim, x,
                                                                                     y, z, mp, bh);
            store_state(e, out);
            return;
        }
        case PT_FORCING:
            r = axis == 0 ? BlackHoleEulerForcingSolver::evolveConservedVariableVector(vec5(s), vec5(s + 5), vec5(s + 10),
                                                                                      x, y, z, dx, dt, bias, lim, mp, bh)
                          : BlackHoleEulerForcingSolver::evolveConservedVariableVector2D(
                                vec5(s), vec5(s + 5), vec5(s + 10), vec5(s + 15), vec5(s + 20), x, y, z, dx, dt, bias,
                                lim, mp, bh);
            break;
        default: return;
    }
    for (int k = 0; k < 5; ++k) out[k] = r[k];
}

}  // namespace

extern "C" void ref_point(int fn, int physics, int axis, const double* s, const double* par, int lim, int side,
                          const double* spacetime, const double* pos, double* out) {
    const double dx = par[0], dt = par[1], bias = par[2], frac = par[3];
    std::vector<double> r;
    switch (fn) {
        case PT_VEC_LF:
            r = EulerFirstOrderSolver::computeLaxFriedrichsFlux(vec5(s), vec5(s + 5), vec5(s + 10), vec5(s + 15), dx, dt);
            break;
        case PT_VEC_RICHTMYER:
            r = EulerFirstOrderSolver::computeRichtmyerFlux(vec5(s), vec5(s + 5), vec5(s + 10), vec5(s + 15), dx, dt);
            break;
        case PT_VEC_FORCE: r = EulerFirstOrderSolver::computeFORCEFlux(vec5(s), vec5(s + 5)); break;
        case PT_VEC_UPDATE: r = EulerFirstOrderSolver::computeFORCEUpdate(vec5(s), vec5(s + 5), vec5(s + 10), dx, dt); break;
        case PT_VEC_EVOLUTION: r = EulerSolvers::computeFractionalEvolutionVector(frac, vec5(s), vec5(s + 5), dx, dt); break;
        case PT_VEC_SLOPE: r = SlopeLimiters::computeSlopeVector(vec5(s), vec5(s + 5), vec5(s + 10), bias, lim); break;
        case PT_LIMITER: out[0] = SlopeLimiters::computeSlopeLimiter(s[0], bias, lim); return;
        default:
            if (physics == 0) point_sr<NewtonianTraits>(fn, axis, s, par, lim, side, out);
            else if (physics == 1) point_sr<RelativisticTraits>(fn, axis, s, par, lim, side, out);
            else point_bh(fn, axis, s, par, lim, side, spacetime, pos, out);
            return;
    }
    for (int k = 0; k < 5; ++k) out[k] = r[k];
}

// EulerSolvers::insertBoundaryCells{,2D} (Solvers/EulerSolvers.cpp:3-85) and the
// relativistic and black-hole copies: the padded grid, (nx + 2 size) x
// (ny + 2 size) in 2D (its corners are default-constructed states).
extern "C" void ref_insert_boundary(int physics, int dim, int size, const double* prim, int nx, int ny, double* out) {
    if (dim == 1) {
        if (physics == 0) {
            std::vector<EulerStateVector> c = to_1d<EulerStateVector>(prim, nx);
            std::vector<EulerStateVector> r = EulerSolvers::insertBoundaryCells(c, size);
            from_1d(r, out);
        } else if (physics == 1) {
            std::vector<RelativisticEulerStateVector> c = to_1d<RelativisticEulerStateVector>(prim, nx);
            std::vector<RelativisticEulerStateVector> r = RelativisticEulerSolvers::insertBoundaryCells(c, size);
            from_1d(r, out);
        } else {
            std::vector<BlackHoleEulerStateVector> c = to_1d<BlackHoleEulerStateVector>(prim, nx);
            std::vector<BlackHoleEulerStateVector> r = BlackHoleEulerSolvers::insertBoundaryCells(c, size);
            from_1d(r, out);
        }
        return;
    }
    if (physics == 0) {
        std::vector<std::vector<EulerStateVector> > c = to_2d<EulerStateVector>(prim, nx, ny);
        std::vector<std::vector<EulerStateVector> > r = EulerSolvers::insertBoundaryCells2D(c, size);
        from_2d(r, out);
    } else if (physics == 1) {
        std::vector<std::vector<RelativisticEulerStateVector> > c = to_2d<RelativisticEulerStateVector>(prim, nx, ny);
        std::vector<std::vector<RelativisticEulerStateVector> > r = RelativisticEulerSolvers::insertBoundaryCells2D(c, size);
        from_2d(r, out);
    } else {
        std::vector<std::vector<BlackHoleEulerStateVector> > c = to_2d<BlackHoleEulerStateVector>(prim, nx, ny);
        std::vector<std::vector<BlackHoleEulerStateVector> > r = BlackHoleEulerSolvers::insertBoundaryCells2D(c, size);
        from_2d(r, out);
    }
}

// Threads of the reference's own OpenMP loops (the -fopenmp build,
// libcosmos_ref_omp.so; the serial build has one).  n <= 0 leaves the
// runtime's default.  Returns the thread count the loops will use.
extern "C" int ref_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
