// Exercises the drop-in solver classes the way the reference's scenario
// drivers call them (Tests/EulerTests.cpp:3-30, Tests/RelativisticEulerTests.cpp),
// writing every result as raw float64 to the file named by argv[1] for
// tests/test_gpu_dropin.py to compare against the oracle.
//
// Sections, in order:
//   1. EulerSecondOrderSolver::solve, Sod, 400 cells, CFL 0.8, t = 0.25   (400 x 5)
//   2. EulerFirstOrderSolver::solve on the same input                      (400 x 5)
//   3. EulerSecondOrderSolver::solve2D, 24 x 20 quadrant input, t = 0.02   (20 x 24 x 5)
//   4. X(dt/2) Y(dt) X(dt/2) by per-sweep calls, dt from computeStableTimeStep2D (20 x 24 x 5)
//   5. RelativisticEulerSecondOrderSolver::solve, 200 cells, t = 0.05     (200 x 5)
//   6. EulerSolvers::computeMaximumWaveSpeed2D and the stable time step    (2)
//   7. SlopeLimiters::computeSlopeLimiter for r = -1 .. 3, 3 limiters       (3 x 41)
//   8. BlackHoleEulerSecondOrderSolver::solve2D, the spinning 2D accretion set-up
//      of Tests/BlackHoleEulerTests.cpp at 16 x 16, t = 0.01, one subcycle  (16 x 16 x 5)
//   9. BlackHoleEulerFirstOrderSolver::solve, the 1D accretion set-up at 60 cells,
//      t = 0.02, one subcycle                                               (60 x 5)
//  10. one forcing step + X/Y/X SLIC per-call on the 2D input (dt 0.002)   (16 x 16 x 5)
//  11. BlackHoleEulerSolvers::computeMaximumWaveSpeed2D of the 2D input    (1)
//  12. EulerAMR::solveLevel1AMR, Sod at 40 cells, tol 5, CFL 0.8, t = 0.02,
//      SLIC: coarse, fine, coarse flags, fine flags                         (40x5, 80x5, 40, 80)
//  13. RelativisticEulerAMR::solve2DLevel2AMR, the quadrant states of 3 at
//      10 x 8, tol 3, t = 0.02, SLIC: three levels and their flags
//      (8x10x5, 16x20x5, 32x40x5, 80, 320, 1280)
//  14. EulerAMR::computeXSLICTimeStep2DAMR on the input of 3 with a
//      checkerboard AMRStructure, dt 0.001                                   (20 x 24 x 5)
//  15. BlackHoleEulerAMR::solveLevel1AMR, the 1D accretion set-up of
//      Tests/BlackHoleEulerAMRTests.cpp at 40 cells, tol 0.5, CFL 0.95,
//      t = 0.05, SLIC: coarse, fine, both flag grids                        (40x5, 80x5, 40, 80)
//  16. BlackHoleEulerAMR::solve2DLevel1AMR, the spinning spheroidal set-up at
//      12 x 12, tol 0.5, t = 0.02, FORCE                                     (12x12x5, 24x24x5, 144, 576)
//  17. BlackHoleEulerAMR::solveLevel2AMR, the spinning 1D set-up at 32 cells,
//      tol 0.5, t = 0.03, FORCE: three levels and their flags               (32x5, 64x5, 128x5, 32, 64, 128)
//  18. BlackHoleEulerAMR::solve2DLevel2AMR on the input of 16 (the completion
//      of the reference's truncated driver): three levels and their flags
//      (12x12x5, 24x24x5, 48x48x5, 144, 576, 2304)
//  19. per-sweep calls through the drop-in's device-resident copy: on the
//      input of 3, X SLIC (dt 0.001), the caller's edit of cell (5, 7)
//      (density x 1.5), X SLIC, the maximum wave speed, Y SLIC, X FORCE on
//      the grid passed back each time                                      (20 x 24 x 5)
//  20-22. the point-level statics (host-evaluated; also run alone with the
//      argument "points", which needs no GPU): see point_sections below.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "cosmos_dropin.hpp"
#include "cosmos_fv.h"

using std::vector;

static void put(FILE* f, double x) { fwrite(&x, sizeof x, 1, f); }

template <class S>
static void put_cells(FILE* f, vector<S>& c) {
    for (auto& s : c) {
        put(f, s.getDensity());
        put(f, s.getXVelocity());
        put(f, s.getYVelocity());
        put(f, s.getZVelocity());
        put(f, s.getPressure());
    }
}

static void put_vec(FILE* f, const vector<double>& v) {
    for (double x : v) put(f, x);
}

template <class S>
static void put_state(FILE* f, S s) {
    vector<S> one(1, s);
    put_cells(f, one);
}

// The inputs of the point sections: five primitive states k = 0..4
//   (1 + 0.1 k, a (0.2 - 0.05 k), a (-0.1 + 0.03 k), a 0.04 k, 1 + 0.2 k)
// with velocity scale a, and five vectors k = 0..4 with components
//   v_k[c] = 0.3 + 0.1 k + 0.05 c, negated for odd c.
template <class S>
static vector<S> point_states(double a) {
    vector<S> out;
    for (int k = 0; k < 5; ++k)
        out.push_back(S(1.0 + 0.1 * k, a * (0.2 - 0.05 * k), a * (-0.1 + 0.03 * k), a * (0.04 * k), 1.0 + 0.2 * k));
    return out;
}

static vector<vector<double> > point_vectors() {
    vector<vector<double> > out;
    for (int k = 0; k < 5; ++k) {
        vector<double> v(5);
        for (int c = 0; c < 5; ++c) v[c] = (c % 2 ? -1.0 : 1.0) * (0.3 + 0.1 * k + 0.05 * c);
        out.push_back(v);
    }
    return out;
}

// 20. Newtonian (gamma 1.4, a = 1; dx 0.01, dt 0.004, bias 0.3, limiter 0):
//     X LF, X Richtmyer, X FORCE, Y LF, Y Richtmyer, Y FORCE of (s0, s1); X, Y SLIC of s0..s3;
//     half steps X side 0, X side 1, Y side 0, Y side 1 of s0..s2; slope vector of s0..s2;
//     on the vectors v: LF(v0, v1, v2, v3), Richtmyer(v0..v3), FORCE(v0, v1), FORCE update
//     (v0, v1, v2), fractional evolution (0.25, v0, v1), evolution (v0, v1), slope (v0, v1, v2),
//     half-step state of (xl, xr, e) sides 0, 1 with xl = (1, 0.1, -0.2, 0.05, 2.5),
//     xr = (1.1, 0.2, -0.1, 0, 2.7), e = 0.01 v0; insertBoundaryCells(s, 2), (s, 1) and
//     insertBoundaryCells2D of the 3 x 4 grid g[i][j] = s_{(i + j) % 5}, sizes 2 and 1
//     (24 x 5, 9 x 5, 7 x 5, 7 x 8 x 5, 5 x 6 x 5)
// 21. Relativistic (gamma 5/3, a = 0.5): the state functions of 20 and the half-step state
//     of (xl, xr, e); insertBoundaryCells(s, 2)  (15 x 5 + 2 x 5 + 9 x 5)
// 22. Black hole (gamma 4/3, a = 0.3; hole (0.1, 0.9, 0.5, 0.4, 0), position (1.1, 0.7, 0.05)):
//     the LF / Richtmyer / FORCE fluxes, SLIC fluxes, half steps and slope vectors (X, Y) as in
//     20 (slope X, Y); fractional steps X, Y (0.5) of s0..s2; the vector half steps X side 0, 1,
//     Y side 0, 1 of (xl, xr, e); the fractional-step state of (xl, e); the forcing of the
//     conserved vectors c_k of s_k at the position (1D: c0, c1, c2; 2D: c0..c4);
//     insertBoundaryCells2D of g, size 1  (16 x 5 + 2 x 5 + 4 x 5 + 5 + 2 x 5 + 5 x 6 x 5)
static void point_sections(FILE* f) {
    const double dx = 0.01, dt = 0.004, bias = 0.3;
    const int lim = 0;
    const vector<vector<double> > v = point_vectors();
    const vector<double> xl = {1.0, 0.1, -0.2, 0.05, 2.5}, xr = {1.1, 0.2, -0.1, 0.0, 2.7};
    vector<double> e(5);
    for (int c = 0; c < 5; ++c) e[c] = 0.01 * v[0][c];
    {
        EulerMaterialParameters mp(1.4);
        vector<EulerStateVector> s = point_states<EulerStateVector>(1.0);
        put_vec(f, EulerFirstOrderSolver::computeXLaxFriedrichsFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, EulerFirstOrderSolver::computeXRichtmyerFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, EulerFirstOrderSolver::computeXFORCEFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, EulerFirstOrderSolver::computeYLaxFriedrichsFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, EulerFirstOrderSolver::computeYRichtmyerFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, EulerFirstOrderSolver::computeYFORCEFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, EulerSecondOrderSolver::computeXSLICFlux(s[0], s[1], s[2], s[3], dx, dt, bias, lim, mp));
        put_vec(f, EulerSecondOrderSolver::computeYSLICFlux(s[0], s[1], s[2], s[3], dx, dt, bias, lim, mp));
        put_state(f, EulerSolvers::evolveStateByHalfXTimeStep(s[0], s[1], s[2], dx, dt, bias, lim, 0, mp));
        put_state(f, EulerSolvers::evolveStateByHalfXTimeStep(s[0], s[1], s[2], dx, dt, bias, lim, 1, mp));
        put_state(f, EulerSolvers::evolveStateByHalfYTimeStep(s[0], s[1], s[2], dx, dt, bias, lim, 0, mp));
        put_state(f, EulerSolvers::evolveStateByHalfYTimeStep(s[0], s[1], s[2], dx, dt, bias, lim, 1, mp));
        put_vec(f, SlopeLimiters::computeSlopeVector(s[0], s[1], s[2], bias, lim, mp));
        put_vec(f, EulerFirstOrderSolver::computeLaxFriedrichsFlux(v[0], v[1], v[2], v[3], dx, dt));
        put_vec(f, EulerFirstOrderSolver::computeRichtmyerFlux(v[0], v[1], v[2], v[3], dx, dt));
        put_vec(f, EulerFirstOrderSolver::computeFORCEFlux(v[0], v[1]));
        put_vec(f, EulerFirstOrderSolver::computeFORCEUpdate(v[0], v[1], v[2], dx, dt));
        put_vec(f, EulerSolvers::computeFractionalEvolutionVector(0.25, v[0], v[1], dx, dt));
        put_vec(f, EulerSolvers::computeEvolutionVector(v[0], v[1], dx, dt));
        put_vec(f, SlopeLimiters::computeSlopeVector(v[0], v[1], v[2], bias, lim));
        put_state(f, EulerSolvers::evolveStateByHalfTimeStep(xl, xr, e, 0, mp));
        put_state(f, EulerSolvers::evolveStateByHalfTimeStep(xl, xr, e, 1, mp));
        vector<EulerStateVector> p2 = EulerSolvers::insertBoundaryCells(s, 2);
        put_cells(f, p2);
        vector<EulerStateVector> p1 = EulerSolvers::insertBoundaryCells(s, 1);
        put_cells(f, p1);
        vector<vector<EulerStateVector> > g(3, vector<EulerStateVector>(4));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) g[i][j] = s[(i + j) % 5];
        for (auto& row : EulerSolvers::insertBoundaryCells2D(g, 2)) put_cells(f, row);
        for (auto& row : EulerSolvers::insertBoundaryCells2D(g, 1)) put_cells(f, row);
    }
    {
        EulerMaterialParameters mp(5.0 / 3.0);
        typedef RelativisticEulerStateVector R;
        vector<R> s = point_states<R>(0.5);
        put_vec(f, RelativisticEulerFirstOrderSolver::computeXLaxFriedrichsFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, RelativisticEulerFirstOrderSolver::computeXRichtmyerFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, RelativisticEulerFirstOrderSolver::computeXFORCEFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, RelativisticEulerFirstOrderSolver::computeYLaxFriedrichsFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, RelativisticEulerFirstOrderSolver::computeYRichtmyerFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, RelativisticEulerFirstOrderSolver::computeYFORCEFlux(s[0], s[1], dx, dt, mp));
        put_vec(f, RelativisticEulerSecondOrderSolver::computeXSLICFlux(s[0], s[1], s[2], s[3], dx, dt, bias, lim, mp));
        put_vec(f, RelativisticEulerSecondOrderSolver::computeYSLICFlux(s[0], s[1], s[2], s[3], dx, dt, bias, lim, mp));
        put_state(f, RelativisticEulerSolvers::evolveStateByHalfXTimeStep(s[0], s[1], s[2], dx, dt, bias, lim, 0, mp));
        put_state(f, RelativisticEulerSolvers::evolveStateByHalfXTimeStep(s[0], s[1], s[2], dx, dt, bias, lim, 1, mp));
        put_state(f, RelativisticEulerSolvers::evolveStateByHalfYTimeStep(s[0], s[1], s[2], dx, dt, bias, lim, 0, mp));
        put_state(f, RelativisticEulerSolvers::evolveStateByHalfYTimeStep(s[0], s[1], s[2], dx, dt, bias, lim, 1, mp));
        put_vec(f, SlopeLimiters::computeSlopeVector(s[0], s[1], s[2], bias, lim, mp));
        put_state(f, RelativisticEulerSolvers::evolveStateByHalfTimeStep(xl, xr, e, 0, mp));
        put_state(f, RelativisticEulerSolvers::evolveStateByHalfTimeStep(xl, xr, e, 1, mp));
        vector<R> p2 = RelativisticEulerSolvers::insertBoundaryCells(s, 2);
        put_cells(f, p2);
    }
    {
        EulerMaterialParameters mp(4.0 / 3.0);
        BlackHoleSpacetime bh(0.1, 0.9, 0.5, 0.4, 0.0);
        const double x = 1.1, y = 0.7, z = 0.05;
        typedef BlackHoleEulerStateVector B;
        typedef BlackHoleEulerFirstOrderSolver F1;
        vector<B> s = point_states<B>(0.3);
        put_vec(f, F1::computeXLaxFriedrichsFlux(s[0], s[1], x, y, z, dx, dt, mp, bh));
        put_vec(f, F1::computeXRichtmyerFlux(s[0], s[1], x, y, z, dx, dt, mp, bh));
        put_vec(f, F1::computeXFORCEFlux(s[0], s[1], x, y, z, dx, dt, mp, bh));
        put_vec(f, F1::computeYLaxFriedrichsFlux(s[0], s[1], x, y, z, dx, dt, mp, bh));
        put_vec(f, F1::computeYRichtmyerFlux(s[0], s[1], x, y, z, dx, dt, mp, bh));
        put_vec(f, F1::computeYFORCEFlux(s[0], s[1], x, y, z, dx, dt, mp, bh));
        put_vec(f, BlackHoleEulerSecondOrderSolver::computeXSLICFlux(s[0], s[1], s[2], s[3], dx, dt, bias, lim, x, y, z,
                                                                     mp, bh));
        put_vec(f, BlackHoleEulerSecondOrderSolver::computeYSLICFlux(s[0], s[1], s[2], s[3], dx, dt, bias, lim, x, y, z,
                                                                     mp, bh));
        for (int axis = 0; axis < 2; ++axis)
            for (int side = 0; side < 2; ++side)
                put_state(f, axis == 0 ? BlackHoleEulerSolvers::evolveStateByHalfXTimeStep(s[0], s[1], s[2], dx, dt, bias,
                                                                                          lim, side, x, y, z, mp, bh)
                                       : BlackHoleEulerSolvers::evolveStateByHalfYTimeStep(s[0], s[1], s[2], dx, dt, bias,
                                                                                          lim, side, x, y, z, mp, bh));
        put_vec(f, SlopeLimiters::computeSlopeVectorX(s[0], s[1], s[2], dx, bias, lim, x, y, z, mp, bh));
        put_vec(f, SlopeLimiters::computeSlopeVectorY(s[0], s[1], s[2], dx, bias, lim, x, y, z, mp, bh));
        put_state(f, BlackHoleEulerSolvers::evolveStateByFractionalXTimeStep(0.5, s[0], s[1], s[2], dx, dt, bias, lim, x,
                                                                            y, z, mp, bh));
        put_state(f, BlackHoleEulerSolvers::evolveStateByFractionalYTimeStep(0.5, s[0], s[1], s[2], dx, dt, bias, lim, x,
                                                                            y, z, mp, bh));
        for (int side = 0; side < 2; ++side)
            put_state(f, BlackHoleEulerSolvers::evolveStateByHalfXTimeStep(xl, xr, e, side, dx, x, y, z, mp, bh));
        for (int side = 0; side < 2; ++side)
            put_state(f, BlackHoleEulerSolvers::evolveStateByHalfYTimeStep(xl, xr, e, side, dx, x, y, z, mp, bh));
        put_state(f, BlackHoleEulerSolvers::evolveStateByFractionalTimeStep(xl, e, x, y, z, mp, bh));
        vector<vector<double> > c(5, vector<double>(5));
        for (int k = 0; k < 5; ++k) {
            double w[5] = {s[k].getDensity(), s[k].getXVelocity(), s[k].getYVelocity(), s[k].getZVelocity(),
                           s[k].getPressure()};
            const double hole[5] = {0.1, 0.9, 0.5, 0.4, 0.0}, pos[3] = {x, y, z};
            if (cosmos_point_conserved(COSMOS_BLACKHOLE, w, 4.0 / 3.0, hole, pos, c[k].data()) != COSMOS_OK)
                throw std::runtime_error("cosmos_point_conserved");
        }
        put_vec(f, BlackHoleEulerForcingSolver::evolveConservedVariableVector(c[0], c[1], c[2], x, y, z, dx, dt, bias,
                                                                              lim, mp, bh));
        put_vec(f, BlackHoleEulerForcingSolver::evolveConservedVariableVector2D(c[0], c[1], c[2], c[3], c[4], x, y, z,
                                                                                dx, dt, bias, lim, mp, bh));
        vector<vector<B> > g(3, vector<B>(4));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) g[i][j] = s[(i + j) % 5];
        for (auto& row : BlackHoleEulerSolvers::insertBoundaryCells2D(g, 1)) put_cells(f, row);
    }
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    FILE* f = fopen(argv[1], "wb");
    if (!f) return 3;
    if (argc > 2 && std::string(argv[2]) == "points") {
        point_sections(f);
        fclose(f);
        return 0;
    }
    EulerMaterialParameters mp(1.4);

    vector<EulerStateVector> sod;
    for (int i = 0; i < 400; ++i)
        sod.push_back(i < 200 ? EulerStateVector(1.0, 0.0, 0.0, 0.0, 1.0) : EulerStateVector(0.125, 0.0, 0.0, 0.0, 0.1));
    vector<EulerStateVector> out = EulerSecondOrderSolver::solve(sod, 1.0 / 400, 0.8, 0.25, 0.0, 1, mp);
    put_cells(f, out);
    out = EulerFirstOrderSolver::solve(sod, 1.0 / 400, 0.8, 0.25, mp);
    put_cells(f, out);

    // quadrant states as in the Python workload generator, given explicitly
    const double q[4][5] = {{1.2, 0.3, -0.4, 0.0, 1.5}, {0.4, -0.2, 0.1, 0.0, 0.6},
                            {0.9, 0.5, 0.5, 0.0, 0.3}, {1.7, -0.6, -0.1, 0.0, 1.1}};
    vector<vector<EulerStateVector> > grid(20, vector<EulerStateVector>(24));
    for (int i = 0; i < 20; ++i)
        for (int j = 0; j < 24; ++j) {
            const double* s = q[2 * (i >= 10) + (j >= 12)];
            grid[i][j] = EulerStateVector(s[0], s[1], s[2], s[3], s[4]);
        }
    vector<vector<EulerStateVector> > g2 = EulerSecondOrderSolver::solve2D(grid, 1.0 / 24, 0.9, 0.02, 0.0, 1, mp);
    for (auto& row : g2) put_cells(f, row);

    vector<vector<EulerStateVector> > g3 = grid;
    double dt = EulerSolvers::computeStableTimeStep2D(g3, 1.0 / 24, 0.9, 0.0, 1.0, 0, mp);
    EulerSecondOrderSolver::computeXSLICTimeStep2D(g3, 1.0 / 24, 0.5 * dt, 0.0, 1, mp);
    EulerSecondOrderSolver::computeYSLICTimeStep2D(g3, 1.0 / 24, dt, 0.0, 1, mp);
    EulerSecondOrderSolver::computeXSLICTimeStep2D(g3, 1.0 / 24, 0.5 * dt, 0.0, 1, mp);
    for (auto& row : g3) put_cells(f, row);

    EulerMaterialParameters mps(5.0 / 3.0);
    vector<RelativisticEulerStateVector> sr;
    for (int i = 0; i < 200; ++i)
        sr.push_back(i < 100 ? RelativisticEulerStateVector(10.0, 0.0, 0.0, 0.0, 13.33)
                             : RelativisticEulerStateVector(1.0, 0.0, 0.0, 0.0, 1e-6));
    vector<RelativisticEulerStateVector> sro = RelativisticEulerSecondOrderSolver::solve(sr, 1.0 / 200, 0.5, 0.05, 0.0,
                                                                                         1, mps);
    put_cells(f, sro);

    put(f, EulerSolvers::computeMaximumWaveSpeed2D(grid, mp));
    put(f, dt);

    for (int lim = 0; lim < 3; ++lim)
        for (int k = 0; k <= 40; ++k) put(f, SlopeLimiters::computeSlopeLimiter(-1.0 + 0.1 * k, 0.0, lim));

    // black hole (Tests/BlackHoleEulerTests.cpp:3-31, 125-160 set-ups)
    {
        const int n = 16;
        BlackHoleSpacetime bh(0.1, 0.99, 0.5, 0.5, 0.0);
        vector<vector<BlackHoleEulerStateVector> > g(n, vector<BlackHoleEulerStateVector>(n));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double r = std::sqrt(((i - 0.5 * n) * (i - 0.5 * n)) + ((j - 0.5 * n) * (j - 0.5 * n)));
                g[i][j] = r <= 0.4 * n ? BlackHoleEulerStateVector(10.0, 0.0, 0.0, 0.0, 0.1)
                                       : BlackHoleEulerStateVector(0.1, 0.0, 0.0, 0.0, 0.1);
            }
        vector<vector<BlackHoleEulerStateVector> > g2 =
            BlackHoleEulerSecondOrderSolver::solve2D(g, 1.0 / n, 0.95, 0.01, 0.0, 1, 1, mps, bh);
        for (auto& row : g2) put_cells(f, row);

        const int m = 60;
        BlackHoleSpacetime bh1(0.05, 0.5, 0.0, 0.0);
        vector<BlackHoleEulerStateVector> line(m);
        for (int i = 0; i < m; ++i)
            line[i] = (i >= 0.1 * m && i <= 0.9 * m) ? BlackHoleEulerStateVector(10.0, 0.0, 0.0, 0.0, 0.1)
                                                     : BlackHoleEulerStateVector(0.1, 0.0, 0.0, 0.0, 0.1);
        vector<BlackHoleEulerStateVector> l2 = BlackHoleEulerFirstOrderSolver::solve(line, 1.0 / m, 0.95, 0.02, 1, mps, bh1);
        put_cells(f, l2);

        vector<vector<BlackHoleEulerStateVector> > g3 = g;
        BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep2D(g3, 1.0 / n, 0.001, 0.0, 1, mps, bh);
        BlackHoleEulerSecondOrderSolver::computeXSLICTimeStep2D(g3, 1.0 / n, 0.001, 0.0, 1, mps, bh);
        BlackHoleEulerSecondOrderSolver::computeYSLICTimeStep2D(g3, 1.0 / n, 0.002, 0.0, 1, mps, bh);
        BlackHoleEulerSecondOrderSolver::computeXSLICTimeStep2D(g3, 1.0 / n, 0.001, 0.0, 1, mps, bh);
        for (auto& row : g3) put_cells(f, row);
        put(f, BlackHoleEulerSolvers::computeMaximumWaveSpeed2D(g, mps, bh));
    }
    {
        EulerMaterialParameters mp(1.4);
        const int n = 40;
        vector<EulerStateVector> sod(n);
        for (int i = 0; i < n; ++i)
            sod[i] = i < n / 2 ? EulerStateVector(1.0, 0.0, 0.0, 0.0, 1.0) : EulerStateVector(0.125, 0.0, 0.0, 0.0, 0.1);
        auto r = EulerAMR::solveLevel1AMR(sod, 1.0 / n, 0.8, 0.02, 5.0, 2, mp);
        put_cells(f, std::get<0>(r));
        put_cells(f, std::get<1>(r));
        for (bool b : std::get<2>(r)) put(f, b ? 1.0 : 0.0);
        for (bool b : std::get<3>(r)) put(f, b ? 1.0 : 0.0);

        EulerMaterialParameters mps(5.0 / 3.0);
        const double q[4][5] = {{1.2, 0.3, -0.4, 0.0, 1.5}, {0.4, -0.2, 0.1, 0.0, 0.6}, {0.9, 0.5, 0.5, 0.0, 0.3},
                                {1.7, -0.6, -0.1, 0.0, 1.1}};
        const int nx = 10, ny = 8;
        vector<vector<RelativisticEulerStateVector> > sr(ny, vector<RelativisticEulerStateVector>(nx));
        for (int i = 0; i < ny; ++i)
            for (int j = 0; j < nx; ++j) {
                const double* w = q[2 * (i >= ny / 2) + (j >= nx / 2)];
                sr[i][j] = RelativisticEulerStateVector(w[0], w[1], w[2], w[3], w[4]);
            }
        auto r2 = RelativisticEulerAMR::solve2DLevel2AMR(sr, 1.0 / nx, 0.8, 0.02, 3.0, 2, mps);
        for (auto& row : std::get<0>(r2)) put_cells(f, row);
        for (auto& row : std::get<1>(r2)) put_cells(f, row);
        for (auto& row : std::get<2>(r2)) put_cells(f, row);
        for (auto& row : std::get<3>(r2))
            for (bool b : row) put(f, b ? 1.0 : 0.0);
        for (auto& row : std::get<4>(r2))
            for (bool b : row) put(f, b ? 1.0 : 0.0);
        for (auto& row : std::get<5>(r2))
            for (bool b : row) put(f, b ? 1.0 : 0.0);

        const int mx = 24, my = 20;
        vector<vector<EulerStateVector> > w(my, vector<EulerStateVector>(mx));
        vector<vector<bool> > mask(my, vector<bool>(mx));
        for (int i = 0; i < my; ++i)
            for (int j = 0; j < mx; ++j) {
                const double* s4 = q[2 * (i >= 10) + (j >= 12)];
                w[i][j] = EulerStateVector(s4[0], s4[1], s4[2], s4[3], s4[4]);
                mask[i][j] = ((i + j) % 2) == 0;
            }
        EulerAMR::computeXSLICTimeStep2DAMR(w, 1.0 / mx, 0.001, 0.0, 1, mask, mp);
        for (auto& row : w) put_cells(f, row);
    }
    {
        EulerMaterialParameters mpg(5.0 / 3.0);
        const BlackHoleEulerStateVector dense(10.0, 0.0, 0.0, 0.0, 0.1), thin(0.1, 0.0, 0.0, 0.0, 0.1);
        int n = 40;
        vector<BlackHoleEulerStateVector> line(n);
        for (int i = 0; i < n; ++i) line[i] = (i >= (0.1 * n) && i <= (0.9 * n)) ? dense : thin;
        auto r = BlackHoleEulerAMR::solveLevel1AMR(line, 1.0 / n, 0.95, 0.05, 0.5, 2, mpg,
                                                   BlackHoleSpacetime(0.05, 0.5, 0.0, 0.0));
        put_cells(f, std::get<0>(r));
        put_cells(f, std::get<1>(r));
        for (bool b : std::get<2>(r)) put(f, b ? 1.0 : 0.0);
        for (bool b : std::get<3>(r)) put(f, b ? 1.0 : 0.0);

        const int m = 12;
        vector<vector<BlackHoleEulerStateVector> > ball(m, vector<BlackHoleEulerStateVector>(m));
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j)
                ball[i][j] = std::sqrt(((i - (0.5 * m)) * (i - (0.5 * m)) * 0.8) +
                                       ((j - (0.5 * m)) * (j - (0.5 * m)) * 1.2)) <= (0.4 * m)
                                 ? dense
                                 : thin;
        BlackHoleSpacetime spinning(0.1, 0.99, 0.5, 0.5, 0.0);
        auto r2 = BlackHoleEulerAMR::solve2DLevel1AMR(ball, 1.0 / m, 0.95, 0.02, 0.5, 1, mpg, spinning);
        for (auto& row : std::get<0>(r2)) put_cells(f, row);
        for (auto& row : std::get<1>(r2)) put_cells(f, row);
        for (auto& row : std::get<2>(r2))
            for (bool b : row) put(f, b ? 1.0 : 0.0);
        for (auto& row : std::get<3>(r2))
            for (bool b : row) put(f, b ? 1.0 : 0.0);

        n = 32;
        line.assign(n, thin);
        for (int i = 0; i < n; ++i) line[i] = (i >= (0.1 * n) && i <= (0.9 * n)) ? dense : thin;
        auto r3 = BlackHoleEulerAMR::solveLevel2AMR(line, 1.0 / n, 0.95, 0.03, 0.5, 1, mpg,
                                                    BlackHoleSpacetime(0.1, 0.99, 0.5, 0.0, 0.0));
        put_cells(f, std::get<0>(r3));
        put_cells(f, std::get<1>(r3));
        put_cells(f, std::get<2>(r3));
        for (bool b : std::get<3>(r3)) put(f, b ? 1.0 : 0.0);
        for (bool b : std::get<4>(r3)) put(f, b ? 1.0 : 0.0);
        for (bool b : std::get<5>(r3)) put(f, b ? 1.0 : 0.0);

        auto r4 = BlackHoleEulerAMR::solve2DLevel2AMR(ball, 1.0 / m, 0.95, 0.02, 0.5, 1, mpg, spinning);
        for (auto& row : std::get<0>(r4)) put_cells(f, row);
        for (auto& row : std::get<1>(r4)) put_cells(f, row);
        for (auto& row : std::get<2>(r4)) put_cells(f, row);
        for (auto& row : std::get<3>(r4))
            for (bool b : row) put(f, b ? 1.0 : 0.0);
        for (auto& row : std::get<4>(r4))
            for (bool b : row) put(f, b ? 1.0 : 0.0);
        for (auto& row : std::get<5>(r4))
            for (bool b : row) put(f, b ? 1.0 : 0.0);
    }
    {
        vector<vector<EulerStateVector> > g4 = grid;
        EulerSecondOrderSolver::computeXSLICTimeStep2D(g4, 1.0 / 24, 0.001, 0.0, 1, mp);
        g4[5][7].setDensity(1.5 * g4[5][7].getDensity());
        EulerSecondOrderSolver::computeXSLICTimeStep2D(g4, 1.0 / 24, 0.001, 0.0, 1, mp);
        (void)EulerSolvers::computeMaximumWaveSpeed2D(g4, mp);
        EulerSecondOrderSolver::computeYSLICTimeStep2D(g4, 1.0 / 24, 0.001, 0.0, 1, mp);
        EulerFirstOrderSolver::computeXFORCETimeStep2D(g4, 1.0 / 24, 0.001, mp);
        for (auto& row : g4) put_cells(f, row);
    }
    point_sections(f);
    fclose(f);
    return 0;
}
