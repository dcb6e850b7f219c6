// Drop-in solver classes over the C-ABI (see cosmos_dropin.hpp).
//
// Each call converts the caller's vector<...StateVector> grid to the flat AoS
// primitive layout, runs the sweep / loop on the device through a cached
// solver handle (one per thread, re-created when the grid shape, physics or
// adiabatic index changes), and writes the result back.  solve/solve2D do a
// single upload and download around the device-resident loop.
#include "cosmos_dropin.hpp"

#include <cmath>
#include <cstring>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>

#include "cosmos_fv.h"

namespace {

// The thread's cached solver handle and a page-locked AoS copy of the state
// its device buffer holds.  `held` marks that copy valid: a call whose grid is
// bitwise the copy (a caller passing back the grid it received, the usual
// pattern of per-sweep loops) skips the upload; any call that changes the
// device state clears `held` until its result is downloaded into the copy.
struct CachedHandle {
    cosmos_solver* s = nullptr;
    int physics = -1, dim = 0, nx = 0, ny = 0;
    double gamma = 0.0;
    double* host = nullptr;
    bool held = false;
    ~CachedHandle() {
        if (s) cosmos_destroy(s);
        if (host) cosmos_host_free(host);
    }
};

thread_local CachedHandle g_handle;

void check(int rc, cosmos_solver* s) {
    if (rc != COSMOS_OK) throw std::runtime_error(std::string("cosmos: ") + cosmos_last_error(s));
}

cosmos_solver* handle(int physics, int dim, int nx, int ny, double gamma) {
    CachedHandle& h = g_handle;
    if (h.s && h.physics == physics && h.dim == dim && h.nx == nx && h.ny == ny && h.gamma == gamma) return h.s;
    if (h.s) cosmos_destroy(h.s);
    if (h.host) cosmos_host_free(h.host);
    h.s = nullptr;
    h.host = nullptr;
    h.held = false;
    check(cosmos_create(&h.s, physics, dim, nx, ny, gamma), nullptr);
    void* p = nullptr;
    if (cosmos_host_alloc(sizeof(double) * 5 * (size_t)nx * ny, &p) != COSMOS_OK)
        throw std::runtime_error("cosmos: page-locked host allocation failed");
    h.host = static_cast<double*>(p);
    h.physics = physics;
    h.dim = dim;
    h.nx = nx;
    h.ny = ny;
    h.gamma = gamma;
    return h.s;
}

template <class S>
void pack(vector<S>& cells, double* out) {
    for (size_t j = 0; j < cells.size(); ++j) {
        out[5 * j + 0] = cells[j].getDensity();
        out[5 * j + 1] = cells[j].getXVelocity();
        out[5 * j + 2] = cells[j].getYVelocity();
        out[5 * j + 3] = cells[j].getZVelocity();
        out[5 * j + 4] = cells[j].getPressure();
    }
}

template <class S>
void unpack(const double* in, vector<S>& cells) {
    for (size_t j = 0; j < cells.size(); ++j) {
        cells[j].setDensity(in[5 * j + 0]);
        cells[j].setXVelocity(in[5 * j + 1]);
        cells[j].setYVelocity(in[5 * j + 2]);
        cells[j].setZVelocity(in[5 * j + 3]);
        cells[j].setPressure(in[5 * j + 4]);
    }
}

// Are the cells bitwise the AoS copy h?
template <class S>
bool holds(vector<S>& cells, const double* h) {
    for (size_t j = 0; j < cells.size(); ++j) {
        const double w[5] = {cells[j].getDensity(), cells[j].getXVelocity(), cells[j].getYVelocity(),
                             cells[j].getZVelocity(), cells[j].getPressure()};
        if (std::memcmp(w, h + 5 * j, sizeof w) != 0) return false;
    }
    return true;
}

// The device holds `cells` on return; modifies: the caller is about to change
// the device state (so the host copy stops describing it).
template <class S>
cosmos_solver* load_1d(int physics, vector<S>& cells, double gamma, bool modifies = true) {
    int nx = (int)cells.size();
    cosmos_solver* s = handle(physics, 1, nx, 1, gamma);
    CachedHandle& h = g_handle;
    if (!(h.held && holds(cells, h.host))) {
        h.held = false;
        pack(cells, h.host);
        check(cosmos_upload(s, h.host), s);
        h.held = true;
    }
    if (modifies) h.held = false;
    return s;
}

template <class S>
cosmos_solver* load_2d(int physics, vector<vector<S> >& cells, double gamma, bool modifies = true) {
    int ny = (int)cells.size();
    int nx = ny ? (int)cells[0].size() : 0;
    for (auto& row : cells)
        if ((int)row.size() != nx) throw std::invalid_argument("cosmos: ragged grid rows");
    cosmos_solver* s = handle(physics, 2, nx, ny, gamma);
    CachedHandle& h = g_handle;
    bool same = h.held;
    for (int i = 0; same && i < ny; ++i) same = holds(cells[i], h.host + 5 * (size_t)i * nx);
    if (!same) {
        h.held = false;
        for (int i = 0; i < ny; ++i) pack(cells[i], h.host + 5 * (size_t)i * nx);
        check(cosmos_upload(s, h.host), s);
        h.held = true;
    }
    if (modifies) h.held = false;
    return s;
}

template <class S>
void store_1d(cosmos_solver* s, vector<S>& cells) {
    CachedHandle& h = g_handle;
    check(cosmos_download(s, h.host), s);
    h.held = true;
    unpack(h.host, cells);
}

template <class S>
void store_2d(cosmos_solver* s, vector<vector<S> >& cells) {
    CachedHandle& h = g_handle;
    check(cosmos_download(s, h.host), s);
    h.held = true;
    size_t nx = cells.empty() ? 0 : cells[0].size();
    for (size_t i = 0; i < cells.size(); ++i) unpack(h.host + 5 * i * nx, cells[i]);
}

template <class S>
void sweep_1d(int physics, int order, vector<S>& cells, double dx, double dt, double bias, int lim, double gamma) {
    cosmos_solver* s = load_1d(physics, cells, gamma);
    check(cosmos_sweep(s, order, 0, dx, dt, bias, lim), s);
    store_1d(s, cells);
}

template <class S>
void sweep_2d(int physics, int order, int axis, vector<vector<S> >& cells, double dx, double dt, double bias, int lim,
              double gamma) {
    cosmos_solver* s = load_2d(physics, cells, gamma);
    check(cosmos_sweep(s, order, axis, dx, dt, bias, lim), s);
    store_2d(s, cells);
}

template <class S>
vector<S> solve_1d(int physics, int order, vector<S> cells, double dx, double cfl, double tf, double bias, int lim,
                   double gamma) {
    cosmos_solver* s = load_1d(physics, cells, gamma);
    check(cosmos_solve(s, order, dx, cfl, tf, -1, bias, lim, nullptr, nullptr), s);
    store_1d(s, cells);
    return cells;
}

template <class S>
vector<vector<S> > solve_2d(int physics, int order, vector<vector<S> > cells, double dx, double cfl, double tf,
                            double bias, int lim, double gamma) {
    cosmos_solver* s = load_2d(physics, cells, gamma);
    check(cosmos_solve(s, order, dx, cfl, tf, -1, bias, lim, nullptr, nullptr), s);
    store_2d(s, cells);
    return cells;
}

template <class S>
double wave_speed_1d(int physics, vector<S>& cells, double gamma) {
    cosmos_solver* s = load_1d(physics, cells, gamma, false);
    double v = 0.0;
    check(cosmos_max_wave_speed(s, &v), s);
    return v;
}

template <class S>
double wave_speed_2d(int physics, vector<vector<S> >& cells, double gamma) {
    cosmos_solver* s = load_2d(physics, cells, gamma, false);
    double v = 0.0;
    check(cosmos_max_wave_speed(s, &v), s);
    return v;
}

}  // namespace

// --- SlopeLimiters (computeSlopeLimiter and the slope vectors: cosmos_dropin_point.cpp) ---
double SlopeLimiters::computeSuperBeeSlopeLimiter(double steepness, double bias) {
    return computeSlopeLimiter(steepness, bias, 0);
}
double SlopeLimiters::computeVanLeerSlopeLimiter(double steepness, double bias) {
    return computeSlopeLimiter(steepness, bias, 1);
}
double SlopeLimiters::computeMinBeeSlopeLimiter(double steepness, double bias) {
    return computeSlopeLimiter(steepness, bias, 2);
}
double SlopeLimiters::computeGradientRatio(double steepness, double bias) {
    // xi_R(r) = 2 / ((1 - bias) + (1 + bias) r) (Solvers/SlopeLimiters.cpp:3-6): a scalar helper
    // the limiters use; the device evaluates the same expression inside the sweeps.
    return 2.0 / ((1.0 - bias) + ((1.0 + bias) * steepness));
}

// --- Newtonian ---------------------------------------------------------------
double EulerSolvers::computeMaximumWaveSpeed(vector<EulerStateVector>& c, EulerMaterialParameters mp) {
    return wave_speed_1d(COSMOS_NEWTONIAN, c, mp.getAdiabaticIndex());
}
double EulerSolvers::computeMaximumWaveSpeed2D(vector<vector<EulerStateVector> >& c, EulerMaterialParameters mp) {
    return wave_speed_2d(COSMOS_NEWTONIAN, c, mp.getAdiabaticIndex());
}
double EulerSolvers::computeStableTimeStep(double timeStep, double currentTime, double finalTime, int currentIteration) {
    // Solvers/EulerSolvers.cpp:130-145 (scalar bookkeeping on the caller's values)
    double dt = timeStep;
    if (currentIteration <= 5) dt *= 0.2;
    if ((currentTime + dt) > finalTime) dt = finalTime - currentTime;
    return dt;
}
double EulerSolvers::computeStableTimeStep(vector<EulerStateVector> c, double dx, double cfl, double t, double tf,
                                           int it, EulerMaterialParameters mp) {
    cosmos_solver* s = load_1d(COSMOS_NEWTONIAN, c, mp.getAdiabaticIndex(), false);
    double dt = 0.0;
    check(cosmos_stable_time_step(s, dx, cfl, t, tf, it, &dt), s);
    return dt;
}
double EulerSolvers::computeStableTimeStep2D(vector<vector<EulerStateVector> > c, double dx, double cfl, double t,
                                             double tf, int it, EulerMaterialParameters mp) {
    cosmos_solver* s = load_2d(COSMOS_NEWTONIAN, c, mp.getAdiabaticIndex(), false);
    double dt = 0.0;
    check(cosmos_stable_time_step(s, dx, cfl, t, tf, it, &dt), s);
    return dt;
}
void EulerSolvers::outputStatus(int currentIteration, double currentTime, double timeStep) {
    std::cout << "Iteration = " << currentIteration << "; Time = " << currentTime << "; Timestep = " << timeStep
              << std::endl;
}

void EulerFirstOrderSolver::computeFORCETimeStep(vector<EulerStateVector>& c, double dx, double dt,
                                                 EulerMaterialParameters mp) {
    sweep_1d(COSMOS_NEWTONIAN, 1, c, dx, dt, 0.0, 1, mp.getAdiabaticIndex());
}
void EulerFirstOrderSolver::computeXFORCETimeStep2D(vector<vector<EulerStateVector> >& c, double dx, double dt,
                                                    EulerMaterialParameters mp) {
    sweep_2d(COSMOS_NEWTONIAN, 1, 0, c, dx, dt, 0.0, 1, mp.getAdiabaticIndex());
}
void EulerFirstOrderSolver::computeYFORCETimeStep2D(vector<vector<EulerStateVector> >& c, double dx, double dt,
                                                    EulerMaterialParameters mp) {
    sweep_2d(COSMOS_NEWTONIAN, 1, 1, c, dx, dt, 0.0, 1, mp.getAdiabaticIndex());
}
vector<EulerStateVector> EulerFirstOrderSolver::solve(vector<EulerStateVector> c, double dx, double cfl, double tf,
                                                      EulerMaterialParameters mp) {
    return solve_1d(COSMOS_NEWTONIAN, 1, c, dx, cfl, tf, 0.0, 1, mp.getAdiabaticIndex());
}
vector<vector<EulerStateVector> > EulerFirstOrderSolver::solve2D(vector<vector<EulerStateVector> > c, double dx,
                                                                 double cfl, double tf, EulerMaterialParameters mp) {
    return solve_2d(COSMOS_NEWTONIAN, 1, c, dx, cfl, tf, 0.0, 1, mp.getAdiabaticIndex());
}

void EulerSecondOrderSolver::computeSLICTimeStep(vector<EulerStateVector>& c, double dx, double dt, double bias,
                                                 int lim, EulerMaterialParameters mp) {
    sweep_1d(COSMOS_NEWTONIAN, 2, c, dx, dt, bias, lim, mp.getAdiabaticIndex());
}
void EulerSecondOrderSolver::computeXSLICTimeStep2D(vector<vector<EulerStateVector> >& c, double dx, double dt,
                                                    double bias, int lim, EulerMaterialParameters mp) {
    sweep_2d(COSMOS_NEWTONIAN, 2, 0, c, dx, dt, bias, lim, mp.getAdiabaticIndex());
}
void EulerSecondOrderSolver::computeYSLICTimeStep2D(vector<vector<EulerStateVector> >& c, double dx, double dt,
                                                    double bias, int lim, EulerMaterialParameters mp) {
    sweep_2d(COSMOS_NEWTONIAN, 2, 1, c, dx, dt, bias, lim, mp.getAdiabaticIndex());
}
vector<EulerStateVector> EulerSecondOrderSolver::solve(vector<EulerStateVector> c, double dx, double cfl, double tf,
                                                       double bias, int lim, EulerMaterialParameters mp) {
    return solve_1d(COSMOS_NEWTONIAN, 2, c, dx, cfl, tf, bias, lim, mp.getAdiabaticIndex());
}
vector<vector<EulerStateVector> > EulerSecondOrderSolver::solve2D(vector<vector<EulerStateVector> > c, double dx,
                                                                  double cfl, double tf, double bias, int lim,
                                                                  EulerMaterialParameters mp) {
    return solve_2d(COSMOS_NEWTONIAN, 2, c, dx, cfl, tf, bias, lim, mp.getAdiabaticIndex());
}

// --- Special relativistic ----------------------------------------------------
double RelativisticEulerSolvers::computeMaximumWaveSpeed(vector<RelativisticEulerStateVector>& c,
                                                         EulerMaterialParameters mp) {
    return wave_speed_1d(COSMOS_RELATIVISTIC, c, mp.getAdiabaticIndex());
}
double RelativisticEulerSolvers::computeMaximumWaveSpeed2D(vector<vector<RelativisticEulerStateVector> >& c,
                                                           EulerMaterialParameters mp) {
    return wave_speed_2d(COSMOS_RELATIVISTIC, c, mp.getAdiabaticIndex());
}
double RelativisticEulerSolvers::computeStableTimeStep(vector<RelativisticEulerStateVector> c, double dx, double cfl,
                                                       double t, double tf, int it, EulerMaterialParameters mp) {
    cosmos_solver* s = load_1d(COSMOS_RELATIVISTIC, c, mp.getAdiabaticIndex(), false);
    double dt = 0.0;
    check(cosmos_stable_time_step(s, dx, cfl, t, tf, it, &dt), s);
    return dt;
}
double RelativisticEulerSolvers::computeStableTimeStep2D(vector<vector<RelativisticEulerStateVector> > c, double dx,
                                                         double cfl, double t, double tf, int it,
                                                         EulerMaterialParameters mp) {
    cosmos_solver* s = load_2d(COSMOS_RELATIVISTIC, c, mp.getAdiabaticIndex(), false);
    double dt = 0.0;
    check(cosmos_stable_time_step(s, dx, cfl, t, tf, it, &dt), s);
    return dt;
}

void RelativisticEulerFirstOrderSolver::computeFORCETimeStep(vector<RelativisticEulerStateVector>& c, double dx,
                                                             double dt, EulerMaterialParameters mp) {
    sweep_1d(COSMOS_RELATIVISTIC, 1, c, dx, dt, 0.0, 1, mp.getAdiabaticIndex());
}
void RelativisticEulerFirstOrderSolver::computeXFORCETimeStep2D(vector<vector<RelativisticEulerStateVector> >& c,
                                                                double dx, double dt, EulerMaterialParameters mp) {
    sweep_2d(COSMOS_RELATIVISTIC, 1, 0, c, dx, dt, 0.0, 1, mp.getAdiabaticIndex());
}
void RelativisticEulerFirstOrderSolver::computeYFORCETimeStep2D(vector<vector<RelativisticEulerStateVector> >& c,
                                                                double dx, double dt, EulerMaterialParameters mp) {
    sweep_2d(COSMOS_RELATIVISTIC, 1, 1, c, dx, dt, 0.0, 1, mp.getAdiabaticIndex());
}
vector<RelativisticEulerStateVector> RelativisticEulerFirstOrderSolver::solve(vector<RelativisticEulerStateVector> c,
                                                                              double dx, double cfl, double tf,
                                                                              EulerMaterialParameters mp) {
    return solve_1d(COSMOS_RELATIVISTIC, 1, c, dx, cfl, tf, 0.0, 1, mp.getAdiabaticIndex());
}
vector<vector<RelativisticEulerStateVector> > RelativisticEulerFirstOrderSolver::solve2D(
    vector<vector<RelativisticEulerStateVector> > c, double dx, double cfl, double tf, EulerMaterialParameters mp) {
    return solve_2d(COSMOS_RELATIVISTIC, 1, c, dx, cfl, tf, 0.0, 1, mp.getAdiabaticIndex());
}

void RelativisticEulerSecondOrderSolver::computeSLICTimeStep(vector<RelativisticEulerStateVector>& c, double dx,
                                                             double dt, double bias, int lim,
                                                             EulerMaterialParameters mp) {
    sweep_1d(COSMOS_RELATIVISTIC, 2, c, dx, dt, bias, lim, mp.getAdiabaticIndex());
}
void RelativisticEulerSecondOrderSolver::computeXSLICTimeStep2D(vector<vector<RelativisticEulerStateVector> >& c,
                                                                double dx, double dt, double bias, int lim,
                                                                EulerMaterialParameters mp) {
    sweep_2d(COSMOS_RELATIVISTIC, 2, 0, c, dx, dt, bias, lim, mp.getAdiabaticIndex());
}
void RelativisticEulerSecondOrderSolver::computeYSLICTimeStep2D(vector<vector<RelativisticEulerStateVector> >& c,
                                                                double dx, double dt, double bias, int lim,
                                                                EulerMaterialParameters mp) {
    sweep_2d(COSMOS_RELATIVISTIC, 2, 1, c, dx, dt, bias, lim, mp.getAdiabaticIndex());
}
vector<RelativisticEulerStateVector> RelativisticEulerSecondOrderSolver::solve(vector<RelativisticEulerStateVector> c,
                                                                               double dx, double cfl, double tf,
                                                                               double bias, int lim,
                                                                               EulerMaterialParameters mp) {
    return solve_1d(COSMOS_RELATIVISTIC, 2, c, dx, cfl, tf, bias, lim, mp.getAdiabaticIndex());
}
vector<vector<RelativisticEulerStateVector> > RelativisticEulerSecondOrderSolver::solve2D(
    vector<vector<RelativisticEulerStateVector> > c, double dx, double cfl, double tf, double bias, int lim,
    EulerMaterialParameters mp) {
    return solve_2d(COSMOS_RELATIVISTIC, 2, c, dx, cfl, tf, bias, lim, mp.getAdiabaticIndex());
}

// --- General relativistic (black hole) -------------------------------------------
namespace {

// The cached handle with the caller's black hole set (a host-side parameter).
template <class Load, class Cells>
cosmos_solver* load_gr(Load load, Cells& cells, EulerMaterialParameters& mp, BlackHoleSpacetime& bh, bool modifies) {
    cosmos_solver* s = load(COSMOS_BLACKHOLE, cells, mp.getAdiabaticIndex(), modifies);
    check(cosmos_set_spacetime(s, bh.getBlackHoleMass(), bh.getBlackHoleSpin(), bh.getBlackHoleXPosition(),
                               bh.getBlackHoleYPosition(), bh.getBlackHoleZPosition()),
          s);
    return s;
}

typedef vector<BlackHoleEulerStateVector> GRLine;
typedef vector<vector<BlackHoleEulerStateVector> > GRGrid;

cosmos_solver* load_gr_1d(GRLine& c, EulerMaterialParameters& mp, BlackHoleSpacetime& bh, bool modifies = true) {
    return load_gr(load_1d<BlackHoleEulerStateVector>, c, mp, bh, modifies);
}
cosmos_solver* load_gr_2d(GRGrid& c, EulerMaterialParameters& mp, BlackHoleSpacetime& bh, bool modifies = true) {
    return load_gr(load_2d<BlackHoleEulerStateVector>, c, mp, bh, modifies);
}

void gr_sweep_1d(int order, GRLine& c, double dx, double dt, double bias, int lim, EulerMaterialParameters& mp,
                 BlackHoleSpacetime& bh) {
    cosmos_solver* s = load_gr_1d(c, mp, bh);
    check(cosmos_sweep(s, order, 0, dx, dt, bias, lim), s);
    store_1d(s, c);
}

void gr_sweep_2d(int order, int axis, GRGrid& c, double dx, double dt, double bias, int lim,
                 EulerMaterialParameters& mp, BlackHoleSpacetime& bh) {
    cosmos_solver* s = load_gr_2d(c, mp, bh);
    check(cosmos_sweep(s, order, axis, dx, dt, bias, lim), s);
    store_2d(s, c);
}

double gr_stable_dt(cosmos_solver* s, double dx, double cfl, double t, double tf, int it) {
    double dt = 0.0;
    check(cosmos_stable_time_step(s, dx, cfl, t, tf, it, &dt), s);
    return dt;
}

}  // namespace

double BlackHoleEulerSolvers::computeMaximumWaveSpeed(GRLine& c, EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    cosmos_solver* s = load_gr_1d(c, mp, bh, false);
    double v = 0.0;
    check(cosmos_max_wave_speed(s, &v), s);
    return v;
}
double BlackHoleEulerSolvers::computeMaximumWaveSpeed2D(GRGrid& c, EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    cosmos_solver* s = load_gr_2d(c, mp, bh, false);
    double v = 0.0;
    check(cosmos_max_wave_speed(s, &v), s);
    return v;
}
double BlackHoleEulerSolvers::computeStableTimeStep(GRLine c, double dx, double cfl, double t, double tf, int it,
                                                    EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    return gr_stable_dt(load_gr_1d(c, mp, bh, false), dx, cfl, t, tf, it);
}
double BlackHoleEulerSolvers::computeStableTimeStep2D(GRGrid c, double dx, double cfl, double t, double tf, int it,
                                                      EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    return gr_stable_dt(load_gr_2d(c, mp, bh, false), dx, cfl, t, tf, it);
}

void BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep(GRLine& c, double dx, double dt, double bias, int lim,
                                                            EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    cosmos_solver* s = load_gr_1d(c, mp, bh);
    check(cosmos_forcing(s, dx, dt, bias, lim), s);
    store_1d(s, c);
}
void BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep2D(GRGrid& c, double dx, double dt, double bias, int lim,
                                                              EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    cosmos_solver* s = load_gr_2d(c, mp, bh);
    check(cosmos_forcing(s, dx, dt, bias, lim), s);
    store_2d(s, c);
}

void BlackHoleEulerFirstOrderSolver::computeFORCETimeStep(GRLine& c, double dx, double dt, EulerMaterialParameters mp,
                                                          BlackHoleSpacetime bh) {
    gr_sweep_1d(1, c, dx, dt, 0.0, 1, mp, bh);
}
void BlackHoleEulerFirstOrderSolver::computeXFORCETimeStep2D(GRGrid& c, double dx, double dt,
                                                             EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    gr_sweep_2d(1, 0, c, dx, dt, 0.0, 1, mp, bh);
}
void BlackHoleEulerFirstOrderSolver::computeYFORCETimeStep2D(GRGrid& c, double dx, double dt,
                                                             EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    gr_sweep_2d(1, 1, c, dx, dt, 0.0, 1, mp, bh);
}
GRLine BlackHoleEulerFirstOrderSolver::solve(GRLine c, double dx, double cfl, double tf, int sub,
                                             EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    cosmos_solver* s = load_gr_1d(c, mp, bh);
    check(cosmos_solve_forced(s, 1, dx, cfl, tf, -1, 0.0, 1, sub, nullptr, nullptr), s);
    store_1d(s, c);
    return c;
}
GRGrid BlackHoleEulerFirstOrderSolver::solve2D(GRGrid c, double dx, double cfl, double tf, int sub,
                                               EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    cosmos_solver* s = load_gr_2d(c, mp, bh);
    check(cosmos_solve_forced(s, 1, dx, cfl, tf, -1, 0.0, 1, sub, nullptr, nullptr), s);
    store_2d(s, c);
    return c;
}

void BlackHoleEulerSecondOrderSolver::computeSLICTimeStep(GRLine& c, double dx, double dt, double bias, int lim,
                                                          EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    gr_sweep_1d(2, c, dx, dt, bias, lim, mp, bh);
}
void BlackHoleEulerSecondOrderSolver::computeXSLICTimeStep2D(GRGrid& c, double dx, double dt, double bias, int lim,
                                                             EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    gr_sweep_2d(2, 0, c, dx, dt, bias, lim, mp, bh);
}
void BlackHoleEulerSecondOrderSolver::computeYSLICTimeStep2D(GRGrid& c, double dx, double dt, double bias, int lim,
                                                             EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    gr_sweep_2d(2, 1, c, dx, dt, bias, lim, mp, bh);
}
GRLine BlackHoleEulerSecondOrderSolver::solve(GRLine c, double dx, double cfl, double tf, double bias, int lim,
                                              int sub, EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    cosmos_solver* s = load_gr_1d(c, mp, bh);
    check(cosmos_solve_forced(s, 2, dx, cfl, tf, -1, bias, lim, sub, nullptr, nullptr), s);
    store_1d(s, c);
    return c;
}
GRGrid BlackHoleEulerSecondOrderSolver::solve2D(GRGrid c, double dx, double cfl, double tf, double bias, int lim,
                                                int sub, EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    cosmos_solver* s = load_gr_2d(c, mp, bh);
    check(cosmos_solve_forced(s, 2, dx, cfl, tf, -1, bias, lim, sub, nullptr, nullptr), s);
    store_2d(s, c);
    return c;
}

// --- AMR (EulerAMR, RelativisticEulerAMR, BlackHoleEulerAMR) --------------------
namespace {

struct AmrHandle {
    cosmos_amr* a = nullptr;
    // bh: the black hole of a COSMOS_BLACKHOLE context
    AmrHandle(int physics, int dim, int nx, int ny, double gamma, BlackHoleSpacetime* bh = nullptr) {
        if (cosmos_amr_create(&a, physics, dim, nx, ny, gamma) != COSMOS_OK)
            throw std::runtime_error(std::string("cosmos: ") + cosmos_amr_last_error(nullptr));
        if (bh)
            check(cosmos_amr_set_spacetime(a, bh->getBlackHoleMass(), bh->getBlackHoleSpin(),
                                           bh->getBlackHoleXPosition(), bh->getBlackHoleYPosition(),
                                           bh->getBlackHoleZPosition()));
    }
    ~AmrHandle() { cosmos_amr_destroy(a); }
    void check(int rc) const {
        if (rc != COSMOS_OK) throw std::runtime_error(std::string("cosmos: ") + cosmos_amr_last_error(a));
    }
};

template <class S>
void pack_grid(vector<vector<S> >& cells, int& nx, int& ny, vector<double>& buf) {
    ny = (int)cells.size();
    nx = ny ? (int)cells[0].size() : 0;
    for (auto& row : cells)
        if ((int)row.size() != nx) throw std::invalid_argument("cosmos: ragged grid rows");
    buf.resize(5 * (size_t)nx * ny);
    for (int i = 0; i < ny; ++i) pack(cells[i], buf.data() + 5 * (size_t)i * nx);
}

template <class S>
vector<vector<S> > unpack_grid(const vector<double>& buf, int nx, int ny) {
    vector<vector<S> > out(ny, vector<S>(nx));
    for (int i = 0; i < ny; ++i) unpack(buf.data() + 5 * (size_t)i * nx, out[i]);
    return out;
}

vector<bool> bools(const vector<unsigned char>& f) { return vector<bool>(f.begin(), f.end()); }

vector<vector<bool> > bools(const vector<unsigned char>& f, int nx, int ny) {
    vector<vector<bool> > out(ny, vector<bool>(nx));
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) out[i][j] = f[(size_t)i * nx + j] != 0;
    return out;
}

// A masked sweep: the cells as the coarse level of an AMR context, the mask
// as its flags (AMR/EulerAMR.cpp:3-167).
template <class S>
void masked_sweep_1d(int physics, int order, vector<S>& cells, double dx, double dt, double bias, int lim,
                     const vector<bool>& mask, double gamma, BlackHoleSpacetime* bh = nullptr) {
    const int n = (int)cells.size();
    if ((int)mask.size() != n) throw std::invalid_argument("cosmos: AMRStructure size differs from the cells'");
    AmrHandle h(physics, 1, n, 1, gamma, bh);
    vector<double> buf(5 * (size_t)n);
    pack(cells, buf.data());
    h.check(cosmos_amr_upload(h.a, buf.data()));
    vector<unsigned char> f(mask.begin(), mask.end());
    h.check(cosmos_amr_set_flags(h.a, f.data(), nullptr));
    h.check(cosmos_amr_sweep(h.a, 0, order, 0, dx, dt, bias, lim));
    h.check(cosmos_amr_download(h.a, buf.data(), nullptr, nullptr, nullptr));
    unpack(buf.data(), cells);
}

template <class S>
void masked_sweep_2d(int physics, int order, int axis, vector<vector<S> >& cells, double dx, double dt, double bias,
                     int lim, const vector<vector<bool> >& mask, double gamma, BlackHoleSpacetime* bh = nullptr) {
    int nx, ny;
    vector<double> buf;
    pack_grid(cells, nx, ny, buf);
    if ((int)mask.size() != ny) throw std::invalid_argument("cosmos: AMRStructure size differs from the cells'");
    vector<unsigned char> f((size_t)nx * ny);
    for (int i = 0; i < ny; ++i) {
        if ((int)mask[i].size() != nx) throw std::invalid_argument("cosmos: AMRStructure size differs from the cells'");
        for (int j = 0; j < nx; ++j) f[(size_t)i * nx + j] = mask[i][j] ? 1 : 0;
    }
    AmrHandle h(physics, 2, nx, ny, gamma, bh);
    h.check(cosmos_amr_upload(h.a, buf.data()));
    h.check(cosmos_amr_set_flags(h.a, f.data(), nullptr));
    h.check(cosmos_amr_sweep(h.a, 0, order, axis, dx, dt, bias, lim));
    h.check(cosmos_amr_download(h.a, buf.data(), nullptr, nullptr, nullptr));
    cells = unpack_grid<S>(buf, nx, ny);
}

template <class S>
tuple<vector<S>, vector<S>, vector<bool>, vector<bool> > level1_1d(int physics, vector<S>& cells, double dx,
                                                                    double cfl, double tf, double tol, int order,
                                                                    double gamma, BlackHoleSpacetime* bh = nullptr) {
    const int n = (int)cells.size();
    AmrHandle h(physics, 1, n, 1, gamma, bh);
    vector<double> c(5 * (size_t)n), f(10 * (size_t)n);
    pack(cells, c.data());
    h.check(cosmos_amr_upload(h.a, c.data()));
    h.check(cosmos_amr_solve_level1(h.a, order, dx, cfl, tf, tol, -1, nullptr, nullptr));
    vector<unsigned char> cf(n), ff(2 * (size_t)n);
    h.check(cosmos_amr_download(h.a, c.data(), f.data(), cf.data(), ff.data()));
    vector<S> coarse(n), fine(2 * (size_t)n);
    unpack(c.data(), coarse);
    unpack(f.data(), fine);
    return tuple<vector<S>, vector<S>, vector<bool>, vector<bool> >(coarse, fine, bools(cf), bools(ff));
}

template <class S>
tuple<vector<vector<S> >, vector<vector<S> >, vector<vector<bool> >, vector<vector<bool> > > level1_2d(
    int physics, vector<vector<S> >& cells, double dx, double cfl, double tf, double tol, int order, double gamma,
    BlackHoleSpacetime* bh = nullptr) {
    int nx, ny;
    vector<double> c;
    pack_grid(cells, nx, ny, c);
    AmrHandle h(physics, 2, nx, ny, gamma, bh);
    h.check(cosmos_amr_upload(h.a, c.data()));
    h.check(cosmos_amr_solve_level1(h.a, order, dx, cfl, tf, tol, -1, nullptr, nullptr));
    vector<double> f(5 * (size_t)(4 * nx) * ny);
    vector<unsigned char> cf((size_t)nx * ny), ff((size_t)4 * nx * ny);
    h.check(cosmos_amr_download(h.a, c.data(), f.data(), cf.data(), ff.data()));
    return tuple<vector<vector<S> >, vector<vector<S> >, vector<vector<bool> >, vector<vector<bool> > >(
        unpack_grid<S>(c, nx, ny), unpack_grid<S>(f, 2 * nx, 2 * ny), bools(cf, nx, ny), bools(ff, 2 * nx, 2 * ny));
}

template <class S>
tuple<vector<S>, vector<S>, vector<S>, vector<bool>, vector<bool>, vector<bool> > level2_1d(
    int physics, vector<S>& cells, double dx, double cfl, double tf, double tol, int order, double gamma,
    BlackHoleSpacetime* bh = nullptr) {
    const int n = (int)cells.size();
    AmrHandle h(physics, 1, n, 1, gamma, bh);
    vector<double> c(5 * (size_t)n), m(10 * (size_t)n), f(20 * (size_t)n);
    pack(cells, c.data());
    h.check(cosmos_amr_upload(h.a, c.data()));
    h.check(cosmos_amr_solve_level2(h.a, order, dx, cfl, tf, tol, -1, nullptr, nullptr));
    vector<unsigned char> cf(n), mf(2 * (size_t)n), ff(4 * (size_t)n);
    h.check(cosmos_amr_download_level2(h.a, c.data(), m.data(), f.data(), cf.data(), mf.data(), ff.data()));
    vector<S> coarse(n), mid(2 * (size_t)n), fine(4 * (size_t)n);
    unpack(c.data(), coarse);
    unpack(m.data(), mid);
    unpack(f.data(), fine);
    return tuple<vector<S>, vector<S>, vector<S>, vector<bool>, vector<bool>, vector<bool> >(
        coarse, mid, fine, bools(cf), bools(mf), bools(ff));
}

template <class S>
tuple<vector<vector<S> >, vector<vector<S> >, vector<vector<S> >, vector<vector<bool> >, vector<vector<bool> >,
      vector<vector<bool> > >
level2_2d(int physics, vector<vector<S> >& cells, double dx, double cfl, double tf, double tol, int order,
          double gamma, BlackHoleSpacetime* bh = nullptr) {
    int nx, ny;
    vector<double> c;
    pack_grid(cells, nx, ny, c);
    AmrHandle h(physics, 2, nx, ny, gamma, bh);
    h.check(cosmos_amr_upload(h.a, c.data()));
    h.check(cosmos_amr_solve_level2(h.a, order, dx, cfl, tf, tol, -1, nullptr, nullptr));
    const size_t n = (size_t)nx * ny;
    vector<double> m(5 * 4 * n), f(5 * 16 * n);
    vector<unsigned char> cf(n), mf(4 * n), ff(16 * n);
    h.check(cosmos_amr_download_level2(h.a, c.data(), m.data(), f.data(), cf.data(), mf.data(), ff.data()));
    return tuple<vector<vector<S> >, vector<vector<S> >, vector<vector<S> >, vector<vector<bool> >,
                 vector<vector<bool> >, vector<vector<bool> > >(
        unpack_grid<S>(c, nx, ny), unpack_grid<S>(m, 2 * nx, 2 * ny), unpack_grid<S>(f, 4 * nx, 4 * ny),
        bools(cf, nx, ny), bools(mf, 2 * nx, 2 * ny), bools(ff, 4 * nx, 4 * ny));
}

typedef EulerStateVector NS;
typedef RelativisticEulerStateVector RS;
typedef BlackHoleEulerStateVector BS;

}  // namespace

void EulerAMR::computeFORCETimeStepAMR(vector<NS>& c, double dx, double dt, vector<bool> m, EulerMaterialParameters mp) {
    masked_sweep_1d(COSMOS_NEWTONIAN, 1, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex());
}
void EulerAMR::computeXFORCETimeStep2DAMR(vector<vector<NS> >& c, double dx, double dt, vector<vector<bool> > m,
                                          EulerMaterialParameters mp) {
    masked_sweep_2d(COSMOS_NEWTONIAN, 1, 0, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex());
}
void EulerAMR::computeYFORCETimeStep2DAMR(vector<vector<NS> >& c, double dx, double dt, vector<vector<bool> > m,
                                          EulerMaterialParameters mp) {
    masked_sweep_2d(COSMOS_NEWTONIAN, 1, 1, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex());
}
void EulerAMR::computeSLICTimeStepAMR(vector<NS>& c, double dx, double dt, double bias, int lim, vector<bool> m,
                                      EulerMaterialParameters mp) {
    masked_sweep_1d(COSMOS_NEWTONIAN, 2, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex());
}
void EulerAMR::computeXSLICTimeStep2DAMR(vector<vector<NS> >& c, double dx, double dt, double bias, int lim,
                                         vector<vector<bool> > m, EulerMaterialParameters mp) {
    masked_sweep_2d(COSMOS_NEWTONIAN, 2, 0, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex());
}
void EulerAMR::computeYSLICTimeStep2DAMR(vector<vector<NS> >& c, double dx, double dt, double bias, int lim,
                                         vector<vector<bool> > m, EulerMaterialParameters mp) {
    masked_sweep_2d(COSMOS_NEWTONIAN, 2, 1, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex());
}
tuple<vector<NS>, vector<NS>, vector<bool>, vector<bool> > EulerAMR::solveLevel1AMR(
    vector<NS> c, double dx, double cfl, double tf, double tol, int order, EulerMaterialParameters mp) {
    return level1_1d(COSMOS_NEWTONIAN, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex());
}
tuple<vector<vector<NS> >, vector<vector<NS> >, vector<vector<bool> >, vector<vector<bool> > >
EulerAMR::solve2DLevel1AMR(vector<vector<NS> > c, double dx, double cfl, double tf, double tol, int order,
                           EulerMaterialParameters mp) {
    return level1_2d(COSMOS_NEWTONIAN, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex());
}
tuple<vector<NS>, vector<NS>, vector<NS>, vector<bool>, vector<bool>, vector<bool> > EulerAMR::solveLevel2AMR(
    vector<NS> c, double dx, double cfl, double tf, double tol, int order, EulerMaterialParameters mp) {
    return level2_1d(COSMOS_NEWTONIAN, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex());
}
tuple<vector<vector<NS> >, vector<vector<NS> >, vector<vector<NS> >, vector<vector<bool> >, vector<vector<bool> >,
      vector<vector<bool> > >
EulerAMR::solve2DLevel2AMR(vector<vector<NS> > c, double dx, double cfl, double tf, double tol, int order,
                           EulerMaterialParameters mp) {
    return level2_2d(COSMOS_NEWTONIAN, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex());
}

void RelativisticEulerAMR::computeFORCETimeStepAMR(vector<RS>& c, double dx, double dt, vector<bool> m,
                                                   EulerMaterialParameters mp) {
    masked_sweep_1d(COSMOS_RELATIVISTIC, 1, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex());
}
void RelativisticEulerAMR::computeXFORCETimeStep2DAMR(vector<vector<RS> >& c, double dx, double dt,
                                                      vector<vector<bool> > m, EulerMaterialParameters mp) {
    masked_sweep_2d(COSMOS_RELATIVISTIC, 1, 0, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex());
}
void RelativisticEulerAMR::computeYFORCETimeStep2DAMR(vector<vector<RS> >& c, double dx, double dt,
                                                      vector<vector<bool> > m, EulerMaterialParameters mp) {
    masked_sweep_2d(COSMOS_RELATIVISTIC, 1, 1, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex());
}
void RelativisticEulerAMR::computeSLICTimeStepAMR(vector<RS>& c, double dx, double dt, double bias, int lim,
                                                  vector<bool> m, EulerMaterialParameters mp) {
    masked_sweep_1d(COSMOS_RELATIVISTIC, 2, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex());
}
void RelativisticEulerAMR::computeXSLICTimeStep2DAMR(vector<vector<RS> >& c, double dx, double dt, double bias,
                                                     int lim, vector<vector<bool> > m, EulerMaterialParameters mp) {
    masked_sweep_2d(COSMOS_RELATIVISTIC, 2, 0, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex());
}
void RelativisticEulerAMR::computeYSLICTimeStep2DAMR(vector<vector<RS> >& c, double dx, double dt, double bias,
                                                     int lim, vector<vector<bool> > m, EulerMaterialParameters mp) {
    masked_sweep_2d(COSMOS_RELATIVISTIC, 2, 1, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex());
}
tuple<vector<RS>, vector<RS>, vector<bool>, vector<bool> > RelativisticEulerAMR::solveLevel1AMR(
    vector<RS> c, double dx, double cfl, double tf, double tol, int order, EulerMaterialParameters mp) {
    return level1_1d(COSMOS_RELATIVISTIC, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex());
}
tuple<vector<vector<RS> >, vector<vector<RS> >, vector<vector<bool> >, vector<vector<bool> > >
RelativisticEulerAMR::solve2DLevel1AMR(vector<vector<RS> > c, double dx, double cfl, double tf, double tol, int order,
                                       EulerMaterialParameters mp) {
    return level1_2d(COSMOS_RELATIVISTIC, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex());
}
tuple<vector<RS>, vector<RS>, vector<RS>, vector<bool>, vector<bool>, vector<bool> >
RelativisticEulerAMR::solveLevel2AMR(vector<RS> c, double dx, double cfl, double tf, double tol, int order,
                                     EulerMaterialParameters mp) {
    return level2_1d(COSMOS_RELATIVISTIC, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex());
}
tuple<vector<vector<RS> >, vector<vector<RS> >, vector<vector<RS> >, vector<vector<bool> >, vector<vector<bool> >,
      vector<vector<bool> > >
RelativisticEulerAMR::solve2DLevel2AMR(vector<vector<RS> > c, double dx, double cfl, double tf, double tol, int order,
                                       EulerMaterialParameters mp) {
    return level2_2d(COSMOS_RELATIVISTIC, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex());
}

void BlackHoleEulerAMR::computeFORCETimeStepAMR(vector<BS>& c, double dx, double dt, vector<bool> m,
                                               EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    masked_sweep_1d(COSMOS_BLACKHOLE, 1, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex(), &bh);
}
void BlackHoleEulerAMR::computeXFORCETimeStep2DAMR(vector<vector<BS> >& c, double dx, double dt,
                                                  vector<vector<bool> > m, EulerMaterialParameters mp,
                                                  BlackHoleSpacetime bh) {
    masked_sweep_2d(COSMOS_BLACKHOLE, 1, 0, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex(), &bh);
}
void BlackHoleEulerAMR::computeYFORCETimeStep2DAMR(vector<vector<BS> >& c, double dx, double dt,
                                                  vector<vector<bool> > m, EulerMaterialParameters mp,
                                                  BlackHoleSpacetime bh) {
    masked_sweep_2d(COSMOS_BLACKHOLE, 1, 1, c, dx, dt, 0.0, 1, m, mp.getAdiabaticIndex(), &bh);
}
void BlackHoleEulerAMR::computeSLICTimeStepAMR(vector<BS>& c, double dx, double dt, double bias, int lim,
                                              vector<bool> m, EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    masked_sweep_1d(COSMOS_BLACKHOLE, 2, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex(), &bh);
}
void BlackHoleEulerAMR::computeXSLICTimeStep2DAMR(vector<vector<BS> >& c, double dx, double dt, double bias, int lim,
                                                 vector<vector<bool> > m, EulerMaterialParameters mp,
                                                 BlackHoleSpacetime bh) {
    masked_sweep_2d(COSMOS_BLACKHOLE, 2, 0, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex(), &bh);
}
void BlackHoleEulerAMR::computeYSLICTimeStep2DAMR(vector<vector<BS> >& c, double dx, double dt, double bias, int lim,
                                                 vector<vector<bool> > m, EulerMaterialParameters mp,
                                                 BlackHoleSpacetime bh) {
    masked_sweep_2d(COSMOS_BLACKHOLE, 2, 1, c, dx, dt, bias, lim, m, mp.getAdiabaticIndex(), &bh);
}
tuple<vector<BS>, vector<BS>, vector<bool>, vector<bool> > BlackHoleEulerAMR::solveLevel1AMR(
    vector<BS> c, double dx, double cfl, double tf, double tol, int order, EulerMaterialParameters mp,
    BlackHoleSpacetime bh) {
    return level1_1d(COSMOS_BLACKHOLE, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex(), &bh);
}
tuple<vector<vector<BS> >, vector<vector<BS> >, vector<vector<bool> >, vector<vector<bool> > >
BlackHoleEulerAMR::solve2DLevel1AMR(vector<vector<BS> > c, double dx, double cfl, double tf, double tol, int order,
                                    EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    return level1_2d(COSMOS_BLACKHOLE, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex(), &bh);
}
tuple<vector<BS>, vector<BS>, vector<BS>, vector<bool>, vector<bool>, vector<bool> > BlackHoleEulerAMR::solveLevel2AMR(
    vector<BS> c, double dx, double cfl, double tf, double tol, int order, EulerMaterialParameters mp,
    BlackHoleSpacetime bh) {
    return level2_1d(COSMOS_BLACKHOLE, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex(), &bh);
}
tuple<vector<vector<BS> >, vector<vector<BS> >, vector<vector<BS> >, vector<vector<bool> >, vector<vector<bool> >,
      vector<vector<bool> > >
BlackHoleEulerAMR::solve2DLevel2AMR(vector<vector<BS> > c, double dx, double cfl, double tf, double tol, int order,
                                    EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    // the reference's definition is truncated (AMR/BlackHoleEulerAMR.cpp:1319-1325);
    // the library completes it as the 1D and special-relativistic drivers continue
    return level2_2d(COSMOS_BLACKHOLE, c, dx, cfl, tf, tol, order, mp.getAdiabaticIndex(), &bh);
}
