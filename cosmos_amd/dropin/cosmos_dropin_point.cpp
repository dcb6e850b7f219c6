// Point-level statics of the drop-in solver classes (see cosmos_dropin.hpp):
// the boundary padding of insertBoundaryCells{,2D} (a layout copy done here)
// and every one-state / one-interface / one-cell method, computed by the
// library's host-evaluated cosmos_point_* and cosmos_vector_* entry points.
#include <stdexcept>
#include <string>

#include "cosmos_dropin.hpp"
#include "cosmos_fv.h"

namespace {

void pcheck(int rc) {
    if (rc != COSMOS_OK) throw std::runtime_error(std::string("cosmos: ") + cosmos_point_last_error());
}

template <class S>
void prim_of(S s, double* w) {
    w[0] = s.getDensity();
    w[1] = s.getXVelocity();
    w[2] = s.getYVelocity();
    w[3] = s.getZVelocity();
    w[4] = s.getPressure();
}

template <class S>
S state_of(const double* w) {
    S s;
    s.setDensity(w[0]);
    s.setXVelocity(w[1]);
    s.setYVelocity(w[2]);
    s.setZVelocity(w[3]);
    s.setPressure(w[4]);
    return s;
}

// A state's five primitives, packed for the C-ABI.
struct Prim {
    double w[5];
    template <class S>
    explicit Prim(S s) {
        prim_of(s, w);
    }
};

const double* five(const vector<double>& v) {
    if (v.size() != 5) throw std::invalid_argument("cosmos: conserved and flux vectors have 5 components");
    return v.data();
}

size_t same_size(const vector<double>& a, const vector<double>& b) {
    if (a.size() != b.size()) throw std::invalid_argument("cosmos: vector arguments differ in length");
    return a.size();
}

// BlackHoleSpacetime(mass, spin, x, y, z) for the C-ABI.
struct Hole {
    double v[5];
    explicit Hole(BlackHoleSpacetime bh)
        : v{bh.getBlackHoleMass(), bh.getBlackHoleSpin(), bh.getBlackHoleXPosition(), bh.getBlackHoleYPosition(),
            bh.getBlackHoleZPosition()} {}
};

double gamma_of(EulerMaterialParameters mp) { return mp.getAdiabaticIndex(); }

// insertBoundaryCells (Solvers/EulerSolvers.cpp:3-28): boundarySize ghosts at
// each end, mirrored (size 2: cells 1, 0 | ... | n-1, n-2) or copied (size 1);
// other sizes leave the ghosts default-constructed.
template <class S>
vector<S> pad_1d(vector<S>& c, int size) {
    const long n = (long)c.size();
    if (size < 0 || n < (size == 2 ? 2 : 1)) throw std::invalid_argument("cosmos: too few cells for the boundary");
    vector<S> out(n + 2 * size);
    if (size == 1) {
        out[0] = c[0];
        out[n + 1] = c[n - 1];
    } else if (size == 2) {
        out[0] = c[1];
        out[1] = c[0];
        out[n + 2] = c[n - 1];
        out[n + 3] = c[n - 2];
    }
    for (long i = 0; i < n; ++i) out[i + size] = c[i];
    return out;
}

// insertBoundaryCells2D (Solvers/EulerSolvers.cpp:30-85): the same along each
// row and column; the corner ghosts stay default-constructed.
template <class S>
vector<vector<S> > pad_2d(vector<vector<S> >& c, int size) {
    const long rows = (long)c.size();
    const long cols = rows ? (long)c[0].size() : 0;
    const long need = size == 2 ? 2 : 1;
    if (size < 0 || rows < need || cols < need) throw std::invalid_argument("cosmos: too few cells for the boundary");
    for (auto& r : c)
        if ((long)r.size() != cols) throw std::invalid_argument("cosmos: ragged grid rows");
    vector<vector<S> > out(rows + 2 * size, vector<S>(cols + 2 * size));
    if (size == 1) {
        for (long i = 0; i < rows; ++i) {
            out[i + 1][0] = c[i][0];
            out[i + 1][cols + 1] = c[i][cols - 1];
        }
        for (long j = 0; j < cols; ++j) {
            out[0][j + 1] = c[0][j];
            out[rows + 1][j + 1] = c[rows - 1][j];
        }
    } else if (size == 2) {
        for (long i = 0; i < rows; ++i) {
            out[i + 2][0] = c[i][1];
            out[i + 2][1] = c[i][0];
            out[i + 2][cols + 2] = c[i][cols - 1];
            out[i + 2][cols + 3] = c[i][cols - 2];
        }
        for (long j = 0; j < cols; ++j) {
            out[0][j + 2] = c[1][j];
            out[1][j + 2] = c[0][j];
            out[rows + 2][j + 2] = c[rows - 1][j];
            out[rows + 3][j + 2] = c[rows - 2][j];
        }
    }
    for (long i = 0; i < rows; ++i)
        for (long j = 0; j < cols; ++j) out[i + size][j + size] = c[i][j];
    return out;
}

// compute{X,Y}{LaxFriedrichs,Richtmyer,FORCE}Flux(state, state).
template <class S>
vector<double> iflux(int physics, int kind, int axis, S l, S r, double dx, double dt, EulerMaterialParameters mp,
                     const double* hole = nullptr, const double* pos = nullptr) {
    Prim a(l), b(r);
    vector<double> out(5);
    pcheck(cosmos_point_interface_flux(physics, kind, axis, a.w, b.w, dx, dt, gamma_of(mp), hole, pos, out.data()));
    return out;
}

template <class S>
vector<double> slic(int physics, int axis, S ll, S l, S r, S rr, double dx, double dt, double bias, int lim,
                    EulerMaterialParameters mp, const double* hole = nullptr, const double* pos = nullptr) {
    Prim a(ll), b(l), c(r), d(rr);
    vector<double> out(5);
    pcheck(cosmos_point_slic_flux(physics, axis, a.w, b.w, c.w, d.w, dx, dt, bias, lim, gamma_of(mp), hole, pos,
                                  out.data()));
    return out;
}

template <class S>
S half(int physics, int axis, S l, S m, S r, double dx, double dt, double bias, int lim, int side,
       EulerMaterialParameters mp, const double* hole = nullptr, const double* pos = nullptr) {
    Prim a(l), b(m), c(r);
    double out[5];
    pcheck(cosmos_point_half_step(physics, axis, a.w, b.w, c.w, dx, dt, bias, lim, side, gamma_of(mp), hole, pos,
                                  out));
    return state_of<S>(out);
}

// The state of a conserved vector u + e (VectorAlgebra::addVectors, then
// setConservedVariableVector) at pos.
template <class S>
S state_of_sum(int physics, const vector<double>& u, const vector<double>& e, EulerMaterialParameters mp,
               const double* hole = nullptr, const double* pos = nullptr) {
    const double* a = five(u);
    const double* b = five(e);
    double sum[5], out[5];
    for (int k = 0; k < 5; ++k) sum[k] = a[k] + b[k];
    pcheck(cosmos_point_primitive(physics, sum, gamma_of(mp), hole, pos, out));
    return state_of<S>(out);
}

template <class S>
vector<double> slope_states(int physics, int axis, S l, S m, S r, double dx, double bias, int lim,
                            EulerMaterialParameters mp, const double* hole = nullptr, const double* pos = nullptr) {
    Prim a(l), b(m), c(r);
    vector<double> out(5);
    pcheck(cosmos_point_slope_vector(physics, axis, a.w, b.w, c.w, dx, bias, lim, gamma_of(mp), hole, pos,
                                     out.data()));
    return out;
}

}  // namespace

// --- SlopeLimiters -----------------------------------------------------------------
double SlopeLimiters::computeSlopeLimiter(double steepness, double bias, int slopeLimiter) {
    double out = 0.0;
    pcheck(cosmos_point_slope_limiter(steepness, bias, slopeLimiter, &out));
    return out;
}
vector<double> SlopeLimiters::computeSlopeVector(vector<double> left, vector<double> middle, vector<double> right,
                                                 double bias, int slopeLimiter) {
    size_t n = same_size(left, middle);
    same_size(middle, right);
    vector<double> out(n);
    pcheck(cosmos_vector_slope(left.data(), middle.data(), right.data(), (long)n, bias, slopeLimiter, out.data()));
    return out;
}
vector<double> SlopeLimiters::computeSlopeVector(EulerStateVector l, EulerStateVector m, EulerStateVector r,
                                                 double bias, int lim, EulerMaterialParameters mp) {
    return slope_states(COSMOS_NEWTONIAN, 0, l, m, r, 1.0, bias, lim, mp);
}
vector<double> SlopeLimiters::computeSlopeVector(RelativisticEulerStateVector l, RelativisticEulerStateVector m,
                                                 RelativisticEulerStateVector r, double bias, int lim,
                                                 EulerMaterialParameters mp) {
    return slope_states(COSMOS_RELATIVISTIC, 0, l, m, r, 1.0, bias, lim, mp);
}
vector<double> SlopeLimiters::computeSlopeVectorX(BlackHoleEulerStateVector l, BlackHoleEulerStateVector m,
                                                  BlackHoleEulerStateVector r, double dx, double bias, int lim,
                                                  double x, double y, double z, EulerMaterialParameters mp,
                                                  BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return slope_states(COSMOS_BLACKHOLE, 0, l, m, r, dx, bias, lim, mp, Hole(bh).v, pos);
}
vector<double> SlopeLimiters::computeSlopeVectorY(BlackHoleEulerStateVector t, BlackHoleEulerStateVector m,
                                                  BlackHoleEulerStateVector b, double dx, double bias, int lim,
                                                  double x, double y, double z, EulerMaterialParameters mp,
                                                  BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return slope_states(COSMOS_BLACKHOLE, 1, t, m, b, dx, bias, lim, mp, Hole(bh).v, pos);
}

// --- EulerSolvers ------------------------------------------------------------------
vector<EulerStateVector> EulerSolvers::insertBoundaryCells(vector<EulerStateVector>& c, int size) {
    return pad_1d(c, size);
}
vector<vector<EulerStateVector> > EulerSolvers::insertBoundaryCells2D(vector<vector<EulerStateVector> >& c, int size) {
    return pad_2d(c, size);
}
vector<double> EulerSolvers::computeFractionalEvolutionVector(double fraction, vector<double> fl, vector<double> fr,
                                                              double dx, double dt) {
    vector<double> out(same_size(fl, fr));
    pcheck(cosmos_vector_evolution(fraction, fl.data(), fr.data(), (long)out.size(), dx, dt, out.data()));
    return out;
}
vector<double> EulerSolvers::computeEvolutionVector(vector<double> fl, vector<double> fr, double dx, double dt) {
    return computeFractionalEvolutionVector(0.5, fl, fr, dx, dt);
}
EulerStateVector EulerSolvers::evolveStateByHalfTimeStep(vector<double> xl, vector<double> xr, vector<double> ev,
                                                         int side, EulerMaterialParameters mp) {
    return state_of_sum<EulerStateVector>(COSMOS_NEWTONIAN, side == 0 ? xl : xr, ev, mp);
}
EulerStateVector EulerSolvers::evolveStateByHalfXTimeStep(EulerStateVector l, EulerStateVector m, EulerStateVector r,
                                                          double dx, double dt, double bias, int lim, int side,
                                                          EulerMaterialParameters mp) {
    return half(COSMOS_NEWTONIAN, 0, l, m, r, dx, dt, bias, lim, side, mp);
}
EulerStateVector EulerSolvers::evolveStateByHalfYTimeStep(EulerStateVector t, EulerStateVector m, EulerStateVector b,
                                                          double dx, double dt, double bias, int lim, int side,
                                                          EulerMaterialParameters mp) {
    return half(COSMOS_NEWTONIAN, 1, t, m, b, dx, dt, bias, lim, side, mp);
}

// --- EulerFirstOrderSolver ---------------------------------------------------------
vector<double> EulerFirstOrderSolver::computeLaxFriedrichsFlux(vector<double> ul, vector<double> ur,
                                                               vector<double> fl, vector<double> fr, double dx,
                                                               double dt) {
    vector<double> out(same_size(ul, ur));
    same_size(ul, fl);
    same_size(ul, fr);
    pcheck(cosmos_vector_lax_friedrichs(ul.data(), ur.data(), fl.data(), fr.data(), (long)out.size(), dx, dt,
                                        out.data()));
    return out;
}
vector<double> EulerFirstOrderSolver::computeRichtmyerFlux(vector<double> ul, vector<double> ur, vector<double> fl,
                                                           vector<double> fr, double dx, double dt) {
    vector<double> out(same_size(ul, ur));
    same_size(ul, fl);
    same_size(ul, fr);
    pcheck(cosmos_vector_richtmyer(ul.data(), ur.data(), fl.data(), fr.data(), (long)out.size(), dx, dt, out.data()));
    return out;
}
vector<double> EulerFirstOrderSolver::computeFORCEFlux(vector<double> lf, vector<double> ri) {
    vector<double> out(same_size(lf, ri));
    pcheck(cosmos_vector_force(lf.data(), ri.data(), (long)out.size(), out.data()));
    return out;
}
vector<double> EulerFirstOrderSolver::computeFORCEUpdate(vector<double> u, vector<double> fl, vector<double> fr,
                                                         double dx, double dt) {
    vector<double> out(same_size(u, fl));
    same_size(u, fr);
    pcheck(cosmos_vector_force_update(u.data(), fl.data(), fr.data(), (long)out.size(), dx, dt, out.data()));
    return out;
}
vector<double> EulerFirstOrderSolver::computeXLaxFriedrichsFlux(EulerStateVector l, EulerStateVector r, double dx,
                                                                double dt, EulerMaterialParameters mp) {
    return iflux(COSMOS_NEWTONIAN, COSMOS_FLUX_LAX_FRIEDRICHS, 0, l, r, dx, dt, mp);
}
vector<double> EulerFirstOrderSolver::computeXRichtmyerFlux(EulerStateVector l, EulerStateVector r, double dx,
                                                            double dt, EulerMaterialParameters mp) {
    return iflux(COSMOS_NEWTONIAN, COSMOS_FLUX_RICHTMYER, 0, l, r, dx, dt, mp);
}
vector<double> EulerFirstOrderSolver::computeXFORCEFlux(EulerStateVector l, EulerStateVector r, double dx, double dt,
                                                        EulerMaterialParameters mp) {
    return iflux(COSMOS_NEWTONIAN, COSMOS_FLUX_FORCE, 0, l, r, dx, dt, mp);
}
vector<double> EulerFirstOrderSolver::computeYLaxFriedrichsFlux(EulerStateVector t, EulerStateVector b, double dx,
                                                                double dt, EulerMaterialParameters mp) {
    return iflux(COSMOS_NEWTONIAN, COSMOS_FLUX_LAX_FRIEDRICHS, 1, t, b, dx, dt, mp);
}
vector<double> EulerFirstOrderSolver::computeYRichtmyerFlux(EulerStateVector t, EulerStateVector b, double dx,
                                                            double dt, EulerMaterialParameters mp) {
    return iflux(COSMOS_NEWTONIAN, COSMOS_FLUX_RICHTMYER, 1, t, b, dx, dt, mp);
}
vector<double> EulerFirstOrderSolver::computeYFORCEFlux(EulerStateVector t, EulerStateVector b, double dx, double dt,
                                                        EulerMaterialParameters mp) {
    return iflux(COSMOS_NEWTONIAN, COSMOS_FLUX_FORCE, 1, t, b, dx, dt, mp);
}

// --- EulerSecondOrderSolver --------------------------------------------------------
vector<double> EulerSecondOrderSolver::computeXSLICFlux(EulerStateVector ll, EulerStateVector l, EulerStateVector r,
                                                        EulerStateVector rr, double dx, double dt, double bias,
                                                        int lim, EulerMaterialParameters mp) {
    return slic(COSMOS_NEWTONIAN, 0, ll, l, r, rr, dx, dt, bias, lim, mp);
}
vector<double> EulerSecondOrderSolver::computeYSLICFlux(EulerStateVector tt, EulerStateVector t, EulerStateVector b,
                                                        EulerStateVector bb, double dx, double dt, double bias,
                                                        int lim, EulerMaterialParameters mp) {
    return slic(COSMOS_NEWTONIAN, 1, tt, t, b, bb, dx, dt, bias, lim, mp);
}

// --- Relativistic ------------------------------------------------------------------
typedef RelativisticEulerStateVector RState;

vector<RState> RelativisticEulerSolvers::insertBoundaryCells(vector<RState>& c, int size) { return pad_1d(c, size); }
vector<vector<RState> > RelativisticEulerSolvers::insertBoundaryCells2D(vector<vector<RState> >& c, int size) {
    return pad_2d(c, size);
}
RState RelativisticEulerSolvers::evolveStateByHalfTimeStep(vector<double> xl, vector<double> xr, vector<double> ev,
                                                           int side, EulerMaterialParameters mp) {
    return state_of_sum<RState>(COSMOS_RELATIVISTIC, side == 0 ? xl : xr, ev, mp);
}
RState RelativisticEulerSolvers::evolveStateByHalfXTimeStep(RState l, RState m, RState r, double dx, double dt,
                                                            double bias, int lim, int side,
                                                            EulerMaterialParameters mp) {
    return half(COSMOS_RELATIVISTIC, 0, l, m, r, dx, dt, bias, lim, side, mp);
}
RState RelativisticEulerSolvers::evolveStateByHalfYTimeStep(RState t, RState m, RState b, double dx, double dt,
                                                            double bias, int lim, int side,
                                                            EulerMaterialParameters mp) {
    return half(COSMOS_RELATIVISTIC, 1, t, m, b, dx, dt, bias, lim, side, mp);
}
vector<double> RelativisticEulerFirstOrderSolver::computeXLaxFriedrichsFlux(RState l, RState r, double dx, double dt,
                                                                            EulerMaterialParameters mp) {
    return iflux(COSMOS_RELATIVISTIC, COSMOS_FLUX_LAX_FRIEDRICHS, 0, l, r, dx, dt, mp);
}
vector<double> RelativisticEulerFirstOrderSolver::computeXRichtmyerFlux(RState l, RState r, double dx, double dt,
                                                                        EulerMaterialParameters mp) {
    return iflux(COSMOS_RELATIVISTIC, COSMOS_FLUX_RICHTMYER, 0, l, r, dx, dt, mp);
}
vector<double> RelativisticEulerFirstOrderSolver::computeXFORCEFlux(RState l, RState r, double dx, double dt,
                                                                    EulerMaterialParameters mp) {
    return iflux(COSMOS_RELATIVISTIC, COSMOS_FLUX_FORCE, 0, l, r, dx, dt, mp);
}
vector<double> RelativisticEulerFirstOrderSolver::computeYLaxFriedrichsFlux(RState t, RState b, double dx, double dt,
                                                                            EulerMaterialParameters mp) {
    return iflux(COSMOS_RELATIVISTIC, COSMOS_FLUX_LAX_FRIEDRICHS, 1, t, b, dx, dt, mp);
}
vector<double> RelativisticEulerFirstOrderSolver::computeYRichtmyerFlux(RState t, RState b, double dx, double dt,
                                                                        EulerMaterialParameters mp) {
    return iflux(COSMOS_RELATIVISTIC, COSMOS_FLUX_RICHTMYER, 1, t, b, dx, dt, mp);
}
vector<double> RelativisticEulerFirstOrderSolver::computeYFORCEFlux(RState t, RState b, double dx, double dt,
                                                                    EulerMaterialParameters mp) {
    return iflux(COSMOS_RELATIVISTIC, COSMOS_FLUX_FORCE, 1, t, b, dx, dt, mp);
}
vector<double> RelativisticEulerSecondOrderSolver::computeXSLICFlux(RState ll, RState l, RState r, RState rr,
                                                                    double dx, double dt, double bias, int lim,
                                                                    EulerMaterialParameters mp) {
    return slic(COSMOS_RELATIVISTIC, 0, ll, l, r, rr, dx, dt, bias, lim, mp);
}
vector<double> RelativisticEulerSecondOrderSolver::computeYSLICFlux(RState tt, RState t, RState b, RState bb,
                                                                    double dx, double dt, double bias, int lim,
                                                                    EulerMaterialParameters mp) {
    return slic(COSMOS_RELATIVISTIC, 1, tt, t, b, bb, dx, dt, bias, lim, mp);
}

// --- Black hole --------------------------------------------------------------------
typedef BlackHoleEulerStateVector BState;

vector<BState> BlackHoleEulerSolvers::insertBoundaryCells(vector<BState>& c, int size) { return pad_1d(c, size); }
vector<vector<BState> > BlackHoleEulerSolvers::insertBoundaryCells2D(vector<vector<BState> >& c, int size) {
    return pad_2d(c, size);
}
BState BlackHoleEulerSolvers::evolveStateByFractionalTimeStep(vector<double> um, vector<double> ev, double x, double y,
                                                              double z, EulerMaterialParameters mp,
                                                              BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return state_of_sum<BState>(COSMOS_BLACKHOLE, um, ev, mp, Hole(bh).v, pos);
}
BState BlackHoleEulerSolvers::evolveStateByFractionalXTimeStep(double fraction, BState l, BState m, BState r,
                                                               double dx, double dt, double bias, int lim, double x,
                                                               double y, double z, EulerMaterialParameters mp,
                                                               BlackHoleSpacetime bh) {
    Prim a(l), b(m), c(r);
    const double pos[3] = {x, y, z};
    double out[5];
    pcheck(cosmos_point_fractional_step(0, fraction, a.w, b.w, c.w, dx, dt, bias, lim, gamma_of(mp), Hole(bh).v, pos,
                                        out));
    return state_of<BState>(out);
}
BState BlackHoleEulerSolvers::evolveStateByFractionalYTimeStep(double fraction, BState t, BState m, BState b_,
                                                               double dx, double dt, double bias, int lim, double x,
                                                               double y, double z, EulerMaterialParameters mp,
                                                               BlackHoleSpacetime bh) {
    Prim a(t), b(m), c(b_);
    const double pos[3] = {x, y, z};
    double out[5];
    pcheck(cosmos_point_fractional_step(1, fraction, a.w, b.w, c.w, dx, dt, bias, lim, gamma_of(mp), Hole(bh).v, pos,
                                        out));
    return state_of<BState>(out);
}
// (Solvers/BlackHoleEulerSolvers.cpp:219-236, 266-283): the evolved face
// state converted at the face, x -/+ cellSpacing * 0.5.
BState BlackHoleEulerSolvers::evolveStateByHalfXTimeStep(vector<double> xl, vector<double> xr, vector<double> ev,
                                                         int side, double dx, double x, double y, double z,
                                                         EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    const double pos[3] = {side == 0 ? x - (dx * 0.5) : x + (dx * 0.5), y, z};
    return state_of_sum<BState>(COSMOS_BLACKHOLE, side == 0 ? xl : xr, ev, mp, Hole(bh).v, pos);
}
BState BlackHoleEulerSolvers::evolveStateByHalfXTimeStep(BState l, BState m, BState r, double dx, double dt,
                                                         double bias, int lim, int side, double x, double y, double z,
                                                         EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return half(COSMOS_BLACKHOLE, 0, l, m, r, dx, dt, bias, lim, side, mp, Hole(bh).v, pos);
}
BState BlackHoleEulerSolvers::evolveStateByHalfYTimeStep(vector<double> xt, vector<double> xb, vector<double> ev,
                                                         int side, double dx, double x, double y, double z,
                                                         EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    const double pos[3] = {x, side == 0 ? y - (dx * 0.5) : y + (dx * 0.5), z};
    return state_of_sum<BState>(COSMOS_BLACKHOLE, side == 0 ? xt : xb, ev, mp, Hole(bh).v, pos);
}
BState BlackHoleEulerSolvers::evolveStateByHalfYTimeStep(BState t, BState m, BState b, double dx, double dt,
                                                         double bias, int lim, int side, double x, double y, double z,
                                                         EulerMaterialParameters mp, BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return half(COSMOS_BLACKHOLE, 1, t, m, b, dx, dt, bias, lim, side, mp, Hole(bh).v, pos);
}

vector<double> BlackHoleEulerFirstOrderSolver::computeXLaxFriedrichsFlux(BState l, BState r, double x, double y,
                                                                         double z, double dx, double dt,
                                                                         EulerMaterialParameters mp,
                                                                         BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return iflux(COSMOS_BLACKHOLE, COSMOS_FLUX_LAX_FRIEDRICHS, 0, l, r, dx, dt, mp, Hole(bh).v, pos);
}
vector<double> BlackHoleEulerFirstOrderSolver::computeXRichtmyerFlux(BState l, BState r, double x, double y, double z,
                                                                     double dx, double dt, EulerMaterialParameters mp,
                                                                     BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return iflux(COSMOS_BLACKHOLE, COSMOS_FLUX_RICHTMYER, 0, l, r, dx, dt, mp, Hole(bh).v, pos);
}
vector<double> BlackHoleEulerFirstOrderSolver::computeXFORCEFlux(BState l, BState r, double x, double y, double z,
                                                                 double dx, double dt, EulerMaterialParameters mp,
                                                                 BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return iflux(COSMOS_BLACKHOLE, COSMOS_FLUX_FORCE, 0, l, r, dx, dt, mp, Hole(bh).v, pos);
}
vector<double> BlackHoleEulerFirstOrderSolver::computeYLaxFriedrichsFlux(BState t, BState b, double x, double y,
                                                                         double z, double dx, double dt,
                                                                         EulerMaterialParameters mp,
                                                                         BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return iflux(COSMOS_BLACKHOLE, COSMOS_FLUX_LAX_FRIEDRICHS, 1, t, b, dx, dt, mp, Hole(bh).v, pos);
}
vector<double> BlackHoleEulerFirstOrderSolver::computeYRichtmyerFlux(BState t, BState b, double x, double y, double z,
                                                                     double dx, double dt, EulerMaterialParameters mp,
                                                                     BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return iflux(COSMOS_BLACKHOLE, COSMOS_FLUX_RICHTMYER, 1, t, b, dx, dt, mp, Hole(bh).v, pos);
}
vector<double> BlackHoleEulerFirstOrderSolver::computeYFORCEFlux(BState t, BState b, double x, double y, double z,
                                                                 double dx, double dt, EulerMaterialParameters mp,
                                                                 BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return iflux(COSMOS_BLACKHOLE, COSMOS_FLUX_FORCE, 1, t, b, dx, dt, mp, Hole(bh).v, pos);
}
vector<double> BlackHoleEulerSecondOrderSolver::computeXSLICFlux(BState ll, BState l, BState r, BState rr, double dx,
                                                                 double dt, double bias, int lim, double x, double y,
                                                                 double z, EulerMaterialParameters mp,
                                                                 BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return slic(COSMOS_BLACKHOLE, 0, ll, l, r, rr, dx, dt, bias, lim, mp, Hole(bh).v, pos);
}
vector<double> BlackHoleEulerSecondOrderSolver::computeYSLICFlux(BState tt, BState t, BState b, BState bb, double dx,
                                                                 double dt, double bias, int lim, double x, double y,
                                                                 double z, EulerMaterialParameters mp,
                                                                 BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    return slic(COSMOS_BLACKHOLE, 1, tt, t, b, bb, dx, dt, bias, lim, mp, Hole(bh).v, pos);
}
vector<double> BlackHoleEulerForcingSolver::evolveConservedVariableVector(vector<double> ul, vector<double> um,
                                                                          vector<double> ur, double x, double y,
                                                                          double z, double dx, double dt, double bias,
                                                                          int lim, EulerMaterialParameters mp,
                                                                          BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    vector<double> out(5);
    pcheck(cosmos_point_forcing(1, five(ul), five(um), five(ur), nullptr, nullptr, dx, dt, bias, lim, gamma_of(mp),
                                Hole(bh).v, pos, out.data()));
    return out;
}
vector<double> BlackHoleEulerForcingSolver::evolveConservedVariableVector2D(vector<double> ul, vector<double> um,
                                                                            vector<double> ur, vector<double> ut,
                                                                            vector<double> ub, double x, double y,
                                                                            double z, double dx, double dt,
                                                                            double bias, int lim,
                                                                            EulerMaterialParameters mp,
                                                                            BlackHoleSpacetime bh) {
    const double pos[3] = {x, y, z};
    vector<double> out(5);
    pcheck(cosmos_point_forcing(2, five(ul), five(um), five(ur), five(ut), five(ub), dx, dt, bias, lim, gamma_of(mp),
                                Hole(bh).v, pos, out.data()));
    return out;
}
