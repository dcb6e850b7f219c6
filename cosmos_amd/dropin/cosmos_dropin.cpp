// Drop-in solver classes over the C-ABI (see cosmos_dropin.hpp).
//
// Each call converts the caller's vector<...StateVector> grid to the flat AoS
// primitive layout, runs the sweep / loop on the device through a cached
// solver handle (one per thread, re-created when the grid shape, physics or
// adiabatic index changes), and writes the result back.  solve/solve2D do a
// single upload and download around the device-resident loop.
#include "cosmos_dropin.hpp"

#include <cmath>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>

#include "cosmos_fv.h"

namespace {

struct CachedHandle {
    cosmos_solver* s = nullptr;
    int physics = -1, dim = 0, nx = 0, ny = 0;
    double gamma = 0.0;
    ~CachedHandle() {
        if (s) cosmos_destroy(s);
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
    h.s = nullptr;
    check(cosmos_create(&h.s, physics, dim, nx, ny, gamma), nullptr);
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

template <class S>
cosmos_solver* load_1d(int physics, vector<S>& cells, double gamma, vector<double>& buf) {
    int nx = (int)cells.size();
    buf.resize(5 * (size_t)nx);
    pack(cells, buf.data());
    cosmos_solver* s = handle(physics, 1, nx, 1, gamma);
    check(cosmos_upload(s, buf.data()), s);
    return s;
}

template <class S>
cosmos_solver* load_2d(int physics, vector<vector<S> >& cells, double gamma, vector<double>& buf) {
    int ny = (int)cells.size();
    int nx = ny ? (int)cells[0].size() : 0;
    for (auto& row : cells)
        if ((int)row.size() != nx) throw std::invalid_argument("cosmos: ragged grid rows");
    buf.resize(5 * (size_t)nx * ny);
    for (int i = 0; i < ny; ++i) pack(cells[i], buf.data() + 5 * (size_t)i * nx);
    cosmos_solver* s = handle(physics, 2, nx, ny, gamma);
    check(cosmos_upload(s, buf.data()), s);
    return s;
}

template <class S>
void store_1d(cosmos_solver* s, vector<S>& cells, vector<double>& buf) {
    check(cosmos_download(s, buf.data()), s);
    unpack(buf.data(), cells);
}

template <class S>
void store_2d(cosmos_solver* s, vector<vector<S> >& cells, vector<double>& buf) {
    check(cosmos_download(s, buf.data()), s);
    size_t nx = cells.empty() ? 0 : cells[0].size();
    for (size_t i = 0; i < cells.size(); ++i) unpack(buf.data() + 5 * i * nx, cells[i]);
}

template <class S>
void sweep_1d(int physics, int order, vector<S>& cells, double dx, double dt, double bias, int lim, double gamma) {
    vector<double> buf;
    cosmos_solver* s = load_1d(physics, cells, gamma, buf);
    check(cosmos_sweep(s, order, 0, dx, dt, bias, lim), s);
    store_1d(s, cells, buf);
}

template <class S>
void sweep_2d(int physics, int order, int axis, vector<vector<S> >& cells, double dx, double dt, double bias, int lim,
              double gamma) {
    vector<double> buf;
    cosmos_solver* s = load_2d(physics, cells, gamma, buf);
    check(cosmos_sweep(s, order, axis, dx, dt, bias, lim), s);
    store_2d(s, cells, buf);
}

template <class S>
vector<S> solve_1d(int physics, int order, vector<S> cells, double dx, double cfl, double tf, double bias, int lim,
                   double gamma) {
    vector<double> buf;
    cosmos_solver* s = load_1d(physics, cells, gamma, buf);
    check(cosmos_solve(s, order, dx, cfl, tf, -1, bias, lim, nullptr, nullptr), s);
    store_1d(s, cells, buf);
    return cells;
}

template <class S>
vector<vector<S> > solve_2d(int physics, int order, vector<vector<S> > cells, double dx, double cfl, double tf,
                            double bias, int lim, double gamma) {
    vector<double> buf;
    cosmos_solver* s = load_2d(physics, cells, gamma, buf);
    check(cosmos_solve(s, order, dx, cfl, tf, -1, bias, lim, nullptr, nullptr), s);
    store_2d(s, cells, buf);
    return cells;
}

template <class S>
double wave_speed_1d(int physics, vector<S>& cells, double gamma) {
    vector<double> buf;
    cosmos_solver* s = load_1d(physics, cells, gamma, buf);
    double v = 0.0;
    check(cosmos_max_wave_speed(s, &v), s);
    return v;
}

template <class S>
double wave_speed_2d(int physics, vector<vector<S> >& cells, double gamma) {
    vector<double> buf;
    cosmos_solver* s = load_2d(physics, cells, gamma, buf);
    double v = 0.0;
    check(cosmos_max_wave_speed(s, &v), s);
    return v;
}

}  // namespace

// --- SlopeLimiters ---------------------------------------------------------
double SlopeLimiters::computeSlopeLimiter(double steepness, double bias, int slopeLimiter) {
    double out = 0.0;
    check(cosmos_slope_limiter(&steepness, 1, bias, slopeLimiter, &out), nullptr);
    return out;
}
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
vector<double> SlopeLimiters::computeSlopeVector(vector<double> left, vector<double> middle, vector<double> right,
                                                 double bias, int slopeLimiter) {
    if (left.size() != 5 || middle.size() != 5 || right.size() != 5)
        throw std::invalid_argument("cosmos: slope vectors have 5 components");
    vector<double> out(5);
    check(cosmos_slope_vector(left.data(), middle.data(), right.data(), 1, bias, slopeLimiter, out.data()), nullptr);
    return out;
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
    vector<double> buf;
    cosmos_solver* s = load_1d(COSMOS_NEWTONIAN, c, mp.getAdiabaticIndex(), buf);
    double dt = 0.0;
    check(cosmos_stable_time_step(s, dx, cfl, t, tf, it, &dt), s);
    return dt;
}
double EulerSolvers::computeStableTimeStep2D(vector<vector<EulerStateVector> > c, double dx, double cfl, double t,
                                             double tf, int it, EulerMaterialParameters mp) {
    vector<double> buf;
    cosmos_solver* s = load_2d(COSMOS_NEWTONIAN, c, mp.getAdiabaticIndex(), buf);
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
    vector<double> buf;
    cosmos_solver* s = load_1d(COSMOS_RELATIVISTIC, c, mp.getAdiabaticIndex(), buf);
    double dt = 0.0;
    check(cosmos_stable_time_step(s, dx, cfl, t, tf, it, &dt), s);
    return dt;
}
double RelativisticEulerSolvers::computeStableTimeStep2D(vector<vector<RelativisticEulerStateVector> > c, double dx,
                                                         double cfl, double t, double tf, int it,
                                                         EulerMaterialParameters mp) {
    vector<double> buf;
    cosmos_solver* s = load_2d(COSMOS_RELATIVISTIC, c, mp.getAdiabaticIndex(), buf);
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
cosmos_solver* load_gr(Load load, Cells& cells, EulerMaterialParameters& mp, BlackHoleSpacetime& bh,
                       vector<double>& buf) {
    cosmos_solver* s = load(COSMOS_BLACKHOLE, cells, mp.getAdiabaticIndex(), buf);
    check(cosmos_set_spacetime(s, bh.getBlackHoleMass(), bh.getBlackHoleSpin(), bh.getBlackHoleXPosition(),
                               bh.getBlackHoleYPosition(), bh.getBlackHoleZPosition()),
          s);
    return s;
}

typedef vector<BlackHoleEulerStateVector> GRLine;
typedef vector<vector<BlackHoleEulerStateVector> > GRGrid;

cosmos_solver* load_gr_1d(GRLine& c, EulerMaterialParameters& mp, BlackHoleSpacetime& bh, vector<double>& buf) {
    return load_gr(load_1d<BlackHoleEulerStateVector>, c, mp, bh, buf);
}
cosmos_solver* load_gr_2d(GRGrid& c, EulerMaterialParameters& mp, BlackHoleSpacetime& bh, vector<double>& buf) {
    return load_gr(load_2d<BlackHoleEulerStateVector>, c, mp, bh, buf);
}

void gr_sweep_1d(int order, GRLine& c, double dx, double dt, double bias, int lim, EulerMaterialParameters& mp,
                 BlackHoleSpacetime& bh) {
    vector<double> buf;
    cosmos_solver* s = load_gr_1d(c, mp, bh, buf);
    check(cosmos_sweep(s, order, 0, dx, dt, bias, lim), s);
    store_1d(s, c, buf);
}

void gr_sweep_2d(int order, int axis, GRGrid& c, double dx, double dt, double bias, int lim,
                 EulerMaterialParameters& mp, BlackHoleSpacetime& bh) {
    vector<double> buf;
    cosmos_solver* s = load_gr_2d(c, mp, bh, buf);
    check(cosmos_sweep(s, order, axis, dx, dt, bias, lim), s);
    store_2d(s, c, buf);
}

double gr_stable_dt(cosmos_solver* s, double dx, double cfl, double t, double tf, int it) {
    double dt = 0.0;
    check(cosmos_stable_time_step(s, dx, cfl, t, tf, it, &dt), s);
    return dt;
}

}  // namespace

double BlackHoleEulerSolvers::computeMaximumWaveSpeed(GRLine& c, EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    vector<double> buf;
    cosmos_solver* s = load_gr_1d(c, mp, bh, buf);
    double v = 0.0;
    check(cosmos_max_wave_speed(s, &v), s);
    return v;
}
double BlackHoleEulerSolvers::computeMaximumWaveSpeed2D(GRGrid& c, EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    vector<double> buf;
    cosmos_solver* s = load_gr_2d(c, mp, bh, buf);
    double v = 0.0;
    check(cosmos_max_wave_speed(s, &v), s);
    return v;
}
double BlackHoleEulerSolvers::computeStableTimeStep(GRLine c, double dx, double cfl, double t, double tf, int it,
                                                    EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    vector<double> buf;
    return gr_stable_dt(load_gr_1d(c, mp, bh, buf), dx, cfl, t, tf, it);
}
double BlackHoleEulerSolvers::computeStableTimeStep2D(GRGrid c, double dx, double cfl, double t, double tf, int it,
                                                      EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    vector<double> buf;
    return gr_stable_dt(load_gr_2d(c, mp, bh, buf), dx, cfl, t, tf, it);
}

void BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep(GRLine& c, double dx, double dt, double bias, int lim,
                                                            EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    vector<double> buf;
    cosmos_solver* s = load_gr_1d(c, mp, bh, buf);
    check(cosmos_forcing(s, dx, dt, bias, lim), s);
    store_1d(s, c, buf);
}
void BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep2D(GRGrid& c, double dx, double dt, double bias, int lim,
                                                              EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    vector<double> buf;
    cosmos_solver* s = load_gr_2d(c, mp, bh, buf);
    check(cosmos_forcing(s, dx, dt, bias, lim), s);
    store_2d(s, c, buf);
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
    vector<double> buf;
    cosmos_solver* s = load_gr_1d(c, mp, bh, buf);
    check(cosmos_solve_forced(s, 1, dx, cfl, tf, -1, 0.0, 1, sub, nullptr, nullptr), s);
    store_1d(s, c, buf);
    return c;
}
GRGrid BlackHoleEulerFirstOrderSolver::solve2D(GRGrid c, double dx, double cfl, double tf, int sub,
                                               EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    vector<double> buf;
    cosmos_solver* s = load_gr_2d(c, mp, bh, buf);
    check(cosmos_solve_forced(s, 1, dx, cfl, tf, -1, 0.0, 1, sub, nullptr, nullptr), s);
    store_2d(s, c, buf);
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
    vector<double> buf;
    cosmos_solver* s = load_gr_1d(c, mp, bh, buf);
    check(cosmos_solve_forced(s, 2, dx, cfl, tf, -1, bias, lim, sub, nullptr, nullptr), s);
    store_1d(s, c, buf);
    return c;
}
GRGrid BlackHoleEulerSecondOrderSolver::solve2D(GRGrid c, double dx, double cfl, double tf, double bias, int lim,
                                                int sub, EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    vector<double> buf;
    cosmos_solver* s = load_gr_2d(c, mp, bh, buf);
    check(cosmos_solve_forced(s, 2, dx, cfl, tf, -1, bias, lim, sub, nullptr, nullptr), s);
    store_2d(s, c, buf);
    return c;
}
