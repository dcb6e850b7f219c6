// Device building blocks of the FORCE / SLIC sweeps shared by the per-sweep
// kernels (fv_kernels.hip) and the batched persistent kernel (fv_batch.hip):
// boundary index maps, the per-cell stages of the update, one row segment of
// an x sweep by a block, and one column march of a y sweep by a thread.
//
// Stages (each half step and each interface flux evaluated once):
//   A  U = p2c(W) for every cell in the tile plus halo
//   B  the half-step evolution of each cell -> the evolved face states P1, P0
//      (primitives, as the reference stores them), their p2c Q and flux G
//   C  the FORCE flux of each interface from (Q1, G1) of its left cell and
//      (Q0, G0) of its right cell
//   D  the conservative update U + dt/dx (F_left - F_right) -> c2p
#pragma once

#include "fv_kernels.hpp"

namespace cfv {
namespace core {

__device__ __forceinline__ int mirror2(int c, int n) {
    if (c < 0) c = -c - 1;
    if (c >= n) c = 2 * n - 1 - c;
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}

__device__ __forceinline__ int clamp1(int c, int n) {
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}

template <bool W0>
__device__ __forceinline__ void load_cell(const ConstPlanes& in, long idx, double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) W[k] = (W0 && k == 3) ? 0.0 : in.p[k][idx];
}

template <bool W0>
__device__ __forceinline__ void store_cell(const Planes& out, long idx, const double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (!(W0 && k == 3)) out.p[k][idx] = W[k];
}

// Half-step evolution of one cell from its conserved neighbours up to the
// evolved face states' conserved vectors a (side 0) and b (side 1)
// (EulerSolvers::evolveStateByHalf{X,Y}TimeStep, Solvers/EulerSolvers.cpp:191-223,
// before its final setConservedVariableVector).  Con2prim events of the face
// fluxes are counted into eh when given.
template <class Phys, int AXIS, int LIMMODE, bool W0>
CFV_HD void half_faces(const double* ul, const double* um, const double* ur, const SweepCoef& cf, double bias,
                       int limiter, const PhysParams& P, double* a, double* b, Ev* eh = nullptr) {
    double xl[5], xr[5], fl[5], fr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (W0 && k == 3) {
            xl[k] = xr[k] = 0.0;
            continue;
        }
        double s;
        if constexpr (LIMMODE == 2) s = slope_component_fast_vl(ul[k], um[k], ur[k]);
        else s = slope_component<LIMMODE>(ul[k], um[k], ur[k], bias, limiter);
        if constexpr (Phys::kFast) {
            xl[k] = __builtin_fma(-0.5, s, um[k]);
            xr[k] = __builtin_fma(0.5, s, um[k]);
        } else {
            xl[k] = um[k] - (0.5 * s);
            xr[k] = um[k] + (0.5 * s);
        }
    }
    Phys::template flux<AXIS>(xl, fl, Phys::inv_density(xl), P, eh);
    Phys::template flux<AXIS>(xr, fr, Phys::inv_density(xr), P, eh);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double ev = (W0 && k == 3) ? 0.0 : evolution_k(cf.hdt_dx, fl[k], fr[k]);
        a[k] = (W0 && k == 3) ? 0.0 : xl[k] + ev;
        b[k] = (W0 && k == 3) ? 0.0 : xr[k] + ev;
    }
}

// The half-step evolution of one cell (half_faces), then the p2c and flux of
// both evolved face states as the FORCE flux of the neighbouring interfaces
// will need them (Solvers/EulerFirstOrderSolver.cpp:52-59).
// Con2prim events are counted into eh (the face fluxes and face states), eg0
// and eg1 (the fluxes of the two face states' conserved vectors) when given.
template <class Phys, int AXIS, int LIMMODE, bool W0>
CFV_HD void half_step(const double* ul, const double* um, const double* ur, const SweepCoef& cf, double bias,
                      int limiter, const PhysParams& P, double* Q0, double* G0, double* Q1, double* G1,
                      Ev* eh = nullptr, Ev* eg0 = nullptr, Ev* eg1 = nullptr) {
    double a[5], b[5], p0[5], p1[5];
    half_faces<Phys, AXIS, LIMMODE, W0>(ul, um, ur, cf, bias, limiter, P, a, b, eh);
    // P0 = c2p(a) (side 0), P1 = c2p(b) (side 1); p2c(P) has density P[0] = a[0] (resp.
    // b[0]), so the flux of p2c(P) divides by the same density.
    double ia = Phys::inv_density(a), ib = Phys::inv_density(b);
    if constexpr (Phys::kFast) {
        // fast mode: the face state's conserved vector is a (resp. b) itself
        Phys::template c2p_flux<AXIS>(a, p0, G0, ia, P);
        Phys::template c2p_flux<AXIS>(b, p1, G1, ib, P);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            Q0[k] = a[k];
            Q1[k] = b[k];
        }
    } else {
        Phys::c2p(a, p0, ia, P, eh);
        Phys::c2p(b, p1, ib, P, eh);
        Phys::p2c(p0, Q0, P);
        Phys::p2c(p1, Q1, P);
        Phys::template flux<AXIS>(Q0, G0, ia, P, eg0);
        Phys::template flux<AXIS>(Q1, G1, ib, P, eg1);
    }
}

// FORCE flux of one interface from p2c and flux of its two states
// (computeLaxFriedrichsFlux, computeRichtmyerFlux, computeFORCEFlux:
// Solvers/EulerFirstOrderSolver.cpp:3-24).
template <class Phys, int AXIS, bool W0>
CFV_HD void force_flux(const double* UL, const double* FL, const double* UR, const double* FR, const SweepCoef& cf,
                       const PhysParams& P, double* F, Ev* ei = nullptr) {
    double us[5], fs[5];
    if constexpr (Phys::kFast) {
        // U* = (UL + UR) / 2 + (dt / 2dx)(FL - FR);  F = (FL + FR) / 4 + (dx / 4dt)(UL - UR) + F(U*) / 2
#pragma unroll
        for (int k = 0; k < 5; ++k)
            us[k] = (W0 && k == 3) ? 0.0 : __builtin_fma(cf.hdt_dx, FL[k] - FR[k], 0.5 * (UL[k] + UR[k]));
        Phys::template flux<AXIS>(us, fs, Phys::inv_density(us), P);
        const double q = 0.25 * cf.dx_dt;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (W0 && k == 3) {
                F[k] = 0.0;
                continue;
            }
            F[k] = __builtin_fma(0.5, fs[k], __builtin_fma(q, UL[k] - UR[k], 0.25 * (FL[k] + FR[k])));
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
        us[k] = (W0 && k == 3) ? 0.0 : richtmyer_k(UL[k], UR[k], FL[k], FR[k], cf.dt_dx);
    Phys::template flux<AXIS>(us, fs, Phys::inv_density(us), P, ei);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (W0 && k == 3) {
            F[k] = 0.0;
            continue;
        }
        F[k] = force_k(lax_friedrichs_k(UL[k], UR[k], FL[k], FR[k], cf.dx_dt), fs[k]);
    }
}

// computeFORCEUpdate (Solvers/EulerFirstOrderSolver.cpp:96-101) then c2p.
template <class Phys, bool W0>
CFV_HD void update_cell(const double* U, const double* Fl, const double* Fr, const SweepCoef& cf,
                        const PhysParams& P, double* W, Ev* eu = nullptr) {
    double un[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (W0 && k == 3) un[k] = 0.0;
        else if constexpr (Phys::kFast) un[k] = __builtin_fma(cf.dt_dx, Fl[k] - Fr[k], U[k]);
        else un[k] = force_update_k(U[k], Fl[k], Fr[k], cf.dt_dx);
    }
    Phys::c2p(un, W, Phys::inv_density(un), P, eu);
}

// The conservative update of update_cell without its con2prim (fast mode):
// the updated conserved vector, for a following sweep that starts from it.
template <class Phys, bool W0>
__device__ __forceinline__ void update_cons(const double* U, const double* Fl, const double* Fr,
                                            const SweepCoef& cf, double* un) {
    static_assert(Phys::kFast, "conserved hand-off between sweeps is a fast-mode re-association");
#pragma unroll
    for (int k = 0; k < 5; ++k) un[k] = (W0 && k == 3) ? 0.0 : __builtin_fma(cf.dt_dx, Fl[k] - Fr[k], U[k]);
}

// Which of a thread's con2prim solves it owns for the event counts: every
// distinct solve of a line is counted once, in the sites the oracle
// evaluates (cells -1 .. n of the line for the half steps and, FORCE, the cell
// fluxes; interfaces (-1|0) .. (n-1|n); updates 0 .. n-1), though tiles
// recompute halo sites of their neighbours.  c: the thread's position along
// the line of n cells; valid: it is an output cell of this block; first,
// last: this block holds the line's first / last cells.
struct Own {
    bool half, g0, g1, iface, upd;
};

__device__ __forceinline__ Own own_of(int c, int n, bool valid, bool first, bool last) {
    Own o;
    o.half = valid || (c == -1 && first) || (c == n && last);
    o.g0 = o.half && c != -1;
    o.g1 = o.half && c != n;
    o.iface = valid || (c == -1 && first);
    o.upd = valid;
    return o;
}

// One sweep pass of a block of NT threads over cells laid out in LDS so that
// the neighbours of thread s along the swept axis are threads s - STRIDE and
// s + STRIDE: a run of consecutive cells of a row (x sweep, STRIDE 1; several
// short rows may be packed side by side) or a tile of STRIDE columns (y
// sweep, one row of the tile per STRIDE threads).  Thread s loads the cell at
// ld_idx (already mapped through the boundary) into LDS column s; each line
// needs G = 2 (SLIC) or 1 (FORCE) halo cells at both ends, whose outputs are
// not stored.  valid: the thread's cell is an output, stored at st_idx.
// lds: 15 rows of NT doubles; the caller must barrier before reusing it.
// Returns the wave speed of the output cell (the 1D or 2D form by ws_dim; 0
// when ws_dim is 0 or the cell is not an output).
// The stages of one such pass on values held by the threads: W, the thread's
// input primitives, gives Wn, its updated primitives (meaningful where the
// thread's cell has its full stencil in the block).
template <class Phys, int ORDER, int LIMMODE, bool W0, int NT, int AXIS, int STRIDE>
__device__ __forceinline__ void line_step(const double* W, double* Wn, const SweepCoef& cf, double bias, int limiter,
                                          const PhysParams& P, double (*lds)[NT], const Own* own = nullptr,
                                          Ev* acc = nullptr) {
    // special-relativistic solves are counted into registers (the per-stage
    // counters must not live behind a pointer that may be null: that would put
    // them in scratch memory); acc, when given, receives the owned ones
    constexpr bool kCount = Phys::kPhysics == 1;
    Ev eh{}, eg0{}, eg1{}, ei{}, eu{};
    Ev* const ph = kCount ? &eh : nullptr;
    Ev* const pg0 = kCount ? &eg0 : nullptr;
    Ev* const pg1 = kCount ? &eg1 : nullptr;
    Ev* const pi = kCount ? &ei : nullptr;
    Ev* const pu = kCount ? &eu : nullptr;
    //   SLIC:  sA = rows 0-4   stage A U, then the interface fluxes
    //          sB = rows 5-14  stage B Q0, G0
    //   FORCE: sA = rows 0-9   U and F(U);  sB = rows 10-14 the interface fluxes
    double(*sA)[NT] = lds;
    double(*sB)[NT] = lds + (ORDER == 2 ? 5 : 10);
    const int s = threadIdx.x;

    double U[5], F[5];
    Phys::p2c(W, U, P);

    const int sl = s >= STRIDE ? s - STRIDE : 0;
    const int sr = s < NT - STRIDE ? s + STRIDE : NT - 1;
    if constexpr (ORDER == 2) {
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (!(W0 && k == 3)) sA[k][s] = U[k];
        __syncthreads();
        double ul[5], ur[5], Q0[5], G0[5], Q1[5], G1[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            ul[k] = (W0 && k == 3) ? 0.0 : sA[k][sl];
            ur[k] = (W0 && k == 3) ? 0.0 : sA[k][sr];
        }
        half_step<Phys, AXIS, LIMMODE, W0>(ul, U, ur, cf, bias, limiter, P, Q0, G0, Q1, G1, ph, pg0, pg1);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (W0 && k == 3) continue;
            sB[k][s] = Q0[k];
            sB[5 + k][s] = G0[k];
        }
        __syncthreads();
        double UR[5], FR[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            UR[k] = (W0 && k == 3) ? 0.0 : sB[k][sr];
            FR[k] = (W0 && k == 3) ? 0.0 : sB[5 + k][sr];
        }
        force_flux<Phys, AXIS, W0>(Q1, G1, UR, FR, cf, P, F, pi);
        // interface (s, s+1) -> sA (stage A values are no longer read)
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (!(W0 && k == 3)) sA[k][s] = F[k];
        __syncthreads();
    } else {
        double FU[5];
        Phys::template flux<AXIS>(U, FU, Phys::inv_density(U), P, ph);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (W0 && k == 3) continue;
            sA[k][s] = U[k];
            sA[5 + k][s] = FU[k];
        }
        __syncthreads();
        double UR[5], FR[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            UR[k] = (W0 && k == 3) ? 0.0 : sA[k][sr];
            FR[k] = (W0 && k == 3) ? 0.0 : sA[5 + k][sr];
        }
        force_flux<Phys, AXIS, W0>(U, FU, UR, FR, cf, P, F, pi);
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (!(W0 && k == 3)) sB[k][s] = F[k];
        __syncthreads();
    }
    double Fl[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Fl[k] = (W0 && k == 3) ? 0.0 : (ORDER == 2 ? sA[k][sl] : sB[k][sl]);
    update_cell<Phys, W0>(U, Fl, F, cf, P, Wn, pu);
    if (kCount && acc) {
        acc->hit += ((eh.cap + eg0.cap) + (eg1.cap + ei.cap)) + eu.cap;
        acc->hit += ((eh.slow + eg0.slow) + (eg1.slow + ei.slow)) + eu.slow;
        // FORCE: the cell's own flux counts where SLIC counts its half step
        ev_add(*acc, eh, own->half);
        ev_add(*acc, eg0, own->g0);
        ev_add(*acc, eg1, own->g1);
        ev_add(*acc, ei, own->iface);
        ev_add(*acc, eu, own->upd);
    }
}

// Rows of LDS (of NT doubles) line_step_3r uses: the conserved vectors, the
// left-face states and fluxes, and the interface fluxes, each of the stored
// components (four when the z velocity is identically zero).
template <bool W0>
constexpr int slic3r_lds_rows() {
    return 4 * (W0 ? 4 : 5);
}

// line_step for a SLIC pass along x at bias 0 with van Leer (the fused step's
// sweeps; IN_CONS / OUT_CONS: W is the conserved vector / Wn receives the
// updated conserved vector, skipping the con2prim of the previous sweep and
// the prim2con of this one), its stages in three disjoint LDS regions: R1 the conserved vectors
// (rows [0, NC)), R2 the left-face states and fluxes ([NC, 3 NC)), R3 the
// interface fluxes ([3 NC, 4 NC)).  Each region is written only after a
// barrier that follows the previous pass's last read of it, so consecutive
// passes need no barrier between them; a caller that reads or writes these
// rows itself between passes uses R2 (free from the end of one pass to the
// first barrier of the next) and barriers before and after.  Every value is
// computed by the functions and in the order line_step uses: bitwise equal.
template <class Phys, int LIMMODE, bool W0, int NT, bool IN_CONS = false, bool OUT_CONS = false>
__device__ __forceinline__ void line_step_3r(const double* W, double* Wn, const SweepCoef& cf, const PhysParams& P,
                                             double (*lds)[NT]) {
    constexpr int NC = W0 ? 4 : 5;
    double(*R1)[NT] = lds;
    double(*Rq)[NT] = lds + NC;
    double(*Rg)[NT] = lds + 2 * NC;
    double(*R3)[NT] = lds + 3 * NC;
    auto slot = [](int k) { return (W0 && k > 3) ? k - 1 : k; };
    const int s = threadIdx.x;
    const int sl = s >= 1 ? s - 1 : 0;
    const int sr = s < NT - 1 ? s + 1 : NT - 1;
    double U[5], F[5];
    if constexpr (IN_CONS) {
#pragma unroll
        for (int k = 0; k < 5; ++k) U[k] = W[k];
    } else {
        Phys::p2c(W, U, P);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (!(W0 && k == 3)) R1[slot(k)][s] = U[k];
    __syncthreads();
    double Q1[5], G1[5];
    {
        double ul[5], ur[5], Q0[5], G0[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            ul[k] = (W0 && k == 3) ? 0.0 : R1[slot(k)][sl];
            ur[k] = (W0 && k == 3) ? 0.0 : R1[slot(k)][sr];
        }
        half_step<Phys, 0, LIMMODE, W0>(ul, U, ur, cf, 0.0, 1, P, Q0, G0, Q1, G1);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (W0 && k == 3) continue;
            Rq[slot(k)][s] = Q0[k];
            Rg[slot(k)][s] = G0[k];
        }
    }
    __syncthreads();
    {
        double UR[5], FR[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            UR[k] = (W0 && k == 3) ? 0.0 : Rq[slot(k)][sr];
            FR[k] = (W0 && k == 3) ? 0.0 : Rg[slot(k)][sr];
        }
        force_flux<Phys, 0, W0>(Q1, G1, UR, FR, cf, P, F);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (!(W0 && k == 3)) R3[slot(k)][s] = F[k];
    __syncthreads();
    double Fl[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Fl[k] = (W0 && k == 3) ? 0.0 : R3[slot(k)][sl];
    if constexpr (OUT_CONS) update_cons<Phys, W0>(U, Fl, F, cf, Wn);
    else update_cell<Phys, W0>(U, Fl, F, cf, P, Wn);
}

template <class Phys, int ORDER, int LIMMODE, bool W0, int NT, int AXIS, int STRIDE>
__device__ __forceinline__ double line_cells(const ConstPlanes& in, const Planes& out, long ld_idx, long st_idx,
                                          bool valid, int ws_dim, const SweepCoef& cf, double bias, int limiter,
                                          const PhysParams& P, double (*lds)[NT], const Own* own = nullptr,
                                          Ev* acc = nullptr) {
    double W[5], Wn[5];
    load_cell<W0>(in, ld_idx, W);
    line_step<Phys, ORDER, LIMMODE, W0, NT, AXIS, STRIDE>(W, Wn, cf, bias, limiter, P, lds, own, acc);
    if (valid) store_cell<W0>(out, st_idx, Wn);
    double m = 0.0;
    if (ws_dim != 0 && valid) {
        double w = ws_dim == 1 ? Phys::template wavespeed<1>(Wn, P) : Phys::template wavespeed<2>(Wn, P);
        if (w > m) m = w;
    }
    return m;
}

template <class Phys, int ORDER, int LIMMODE, bool W0, int NT>
__device__ __forceinline__ double x_cells(const ConstPlanes& in, const Planes& out, long ld_idx, long st_idx,
                                          bool valid, int ws_dim, const SweepCoef& cf, double bias, int limiter,
                                          const PhysParams& P, double (*lds)[NT], const Own* own = nullptr,
                                          Ev* acc = nullptr) {
    return line_cells<Phys, ORDER, LIMMODE, W0, NT, 0, 1>(in, out, ld_idx, st_idx, valid, ws_dim, cf, bias, limiter, P,
                                                           lds, own, acc);
}

// One y-sweep tile by a block of NT threads: TC columns by NT / TC rows,
// thread s at tile row s / TC and column s mod TC, the first and last G rows
// halo.  Outputs rows [r0, r0 + NT / TC - 2G) of columns [c0, c0 + TC)
// (clipped to the grid).
template <class Phys, int ORDER, int LIMMODE, bool W0, int NT, int TC>
__device__ __forceinline__ double y_tile_cells(const ConstPlanes& in, const Planes& out, int nx, int ny, long pitch,
                                               int c0, int r0, int ws_dim, const SweepCoef& cf, double bias,
                                               int limiter, const PhysParams& P, double (*lds)[NT], Ev* acc = nullptr,
                                               const unsigned char* mask = nullptr) {
    constexpr int G = ORDER == 2 ? 2 : 1;
    constexpr int TR = NT / TC;
    const int s = threadIdx.x;
    const int r = s / TC;
    const int i = r0 - G + r;
    const int j = c0 + (s - r * TC);
    const int jj = j < nx ? j : nx - 1;
    const int ii = ORDER == 2 ? mirror2(i, ny) : clamp1(i, ny);
    bool valid = (r >= G) && (r < TR - G) && (i < ny) && (j < nx);
    if (mask) {
        // masked: only flagged cells are outputs; a tile without one is done
        valid = valid && mask[(long)i * nx + j] != 0;
        if (!__syncthreads_or(valid ? 1 : 0)) return 0.0;
    }
    Own own = own_of(i, ny, valid, r0 == 0, r0 + (TR - 2 * G) >= ny);
    if (j >= nx) own = Own{false, false, false, false, false};
    return line_cells<Phys, ORDER, LIMMODE, W0, NT, 1, TC>(in, out, (long)ii * pitch + jj, (long)i * pitch + j, valid,
                                                            ws_dim, cf, bias, limiter, P, lds, &own, acc);
}

// One row segment of an x sweep: thread s holds cell tile * (NT - 2G) - G + s
// of the row at rowbase.  mask (masked sweeps, else null): the row's flags;
// only flagged cells are outputs, and a segment without one returns at once.
template <class Phys, int ORDER, int LIMMODE, int DIM, bool WS, bool W0, int NT>
__device__ __forceinline__ double x_segment(const ConstPlanes& in, const Planes& out, int nx, long rowbase, int tile,
                                            const SweepCoef& cf, double bias, int limiter, const PhysParams& P,
                                            double (*lds)[NT], Ev* acc = nullptr,
                                            const unsigned char* mask = nullptr) {
    constexpr int G = ORDER == 2 ? 2 : 1;
    constexpr int TW = NT - 2 * G;
    const int s = threadIdx.x;
    const int c = tile * TW - G + s;
    const int cc = ORDER == 2 ? mirror2(c, nx) : clamp1(c, nx);
    bool valid = (s >= G) && (s < NT - G) && (c < nx);
    if (mask) {
        valid = valid && mask[c] != 0;
        if (!__syncthreads_or(valid ? 1 : 0)) return 0.0;
    }
    const Own own = own_of(c, nx, valid, tile == 0, (tile + 1) * TW >= nx);
    return x_cells<Phys, ORDER, LIMMODE, W0, NT>(in, out, rowbase + cc, rowbase + c, valid, WS ? DIM : 0, cf, bias,
                                                 limiter, P, lds, &own, acc);
}

// One y-sweep column march: thread's column j (jj = j clamped into the grid,
// colok = j < nx) over output rows [i0, i1), reading rows i0 - G .. i1 + G - 1
// through the boundary map and keeping a three-cell register window.  Returns
// the largest wave speed of the written cells when ws.
// acc: con2prim events of the solves this march owns (rows [i0, i1) and, at
// the grid's edges, the ghost sites the oracle evaluates; see Own).
template <class Phys, int ORDER, int LIMMODE, bool W0>
__device__ __forceinline__ double y_march(const ConstPlanes& in, const Planes& out, int ny, long pitch, int j,
                                          bool colok, int jj, int i0, int i1, bool ws, const SweepCoef& cf,
                                          double bias, int limiter, const PhysParams& P, Ev* acc = nullptr) {
    double wsmax = 0.0;
    // special-relativistic solves are counted into registers (the per-stage
    // counters must not live behind a pointer that may be null: that would put
    // them in scratch memory); acc, when given, receives the owned ones
    constexpr bool kCount = Phys::kPhysics == 1;
    Ev eh{}, eg0{}, eg1{}, ei{}, eu{};
    Ev* const ph = kCount ? &eh : nullptr;
    Ev* const pg0 = kCount ? &eg0 : nullptr;
    Ev* const pg1 = kCount ? &eg1 : nullptr;
    Ev* const pi = kCount ? &ei : nullptr;
    Ev* const pu = kCount ? &eu : nullptr;
    const bool first = i0 == 0, last = i1 >= ny;
    auto owned = [&](int q) { return colok && ((q >= i0 && q < i1) || (q == -1 && first) || (q == ny && last)); };
    auto owned_iface = [&](int q) { return colok && ((q >= i0 && q < i1) || (q == -1 && first)); };

    if constexpr (ORDER == 2) {
        double Ua[5], Ub[5], Q1p[5], G1p[5], Fp[5], Wn[5];
        load_cell<W0>(in, (long)mirror2(i0 - 2, ny) * pitch + jj, Wn);
#pragma unroll 1
        for (int r = i0 - 2; r <= i1 + 1; ++r) {
            double W[5], U[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) W[k] = Wn[k];
            if (r < i1 + 1) load_cell<W0>(in, (long)mirror2(r + 1, ny) * pitch + jj, Wn);
            Phys::p2c(W, U, P);
            if (r >= i0) {
                double Q0[5], G0[5], Q1[5], G1[5];
                half_step<Phys, 1, LIMMODE, W0>(Ua, Ub, U, cf, bias, limiter, P, Q0, G0, Q1, G1, ph, pg0, pg1);
                if (kCount && acc) {
                    // the half step of row r - 1
                    const bool o = owned(r - 1);
                    ev_add(*acc, eh, o);
                    ev_add(*acc, eg0, o && r - 1 != -1);
                    ev_add(*acc, eg1, o && r - 1 != ny);
                    eh = eg0 = eg1 = Ev{};
                }
                if (r >= i0 + 1) {
                    double F[5];
                    force_flux<Phys, 1, W0>(Q1p, G1p, Q0, G0, cf, P, F, pi);
                    if (kCount && acc) {
                        // interface (r - 2 | r - 1)
                        ev_add(*acc, ei, owned_iface(r - 2));
                        ei = Ev{};
                    }
                    if (r >= i0 + 2) {
                        double Wo[5];
                        update_cell<Phys, W0>(Ua, Fp, F, cf, P, Wo, pu);
                        if (kCount && acc) {
                            ev_add(*acc, eu, colok);
                            eu = Ev{};
                        }
                        if (colok) store_cell<W0>(out, (long)(r - 2) * pitch + j, Wo);
                        if (ws) {
                            double w = Phys::template wavespeed<2>(Wo, P);
                            if (colok && w > wsmax) wsmax = w;
                        }
                    }
#pragma unroll
                    for (int k = 0; k < 5; ++k) Fp[k] = F[k];
                }
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    Q1p[k] = Q1[k];
                    G1p[k] = G1[k];
                }
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                Ua[k] = Ub[k];
                Ub[k] = U[k];
            }
        }
    } else {
        double Up[5], FUp[5], Fp[5], Wn[5];
        load_cell<W0>(in, (long)clamp1(i0 - 1, ny) * pitch + jj, Wn);
#pragma unroll 1
        for (int r = i0 - 1; r <= i1; ++r) {
            double W[5], U[5], FU[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) W[k] = Wn[k];
            if (r < i1) load_cell<W0>(in, (long)clamp1(r + 1, ny) * pitch + jj, Wn);
            Phys::p2c(W, U, P);
            Phys::template flux<1>(U, FU, Phys::inv_density(U), P, ph);
            if (kCount && acc) {
                // the flux of row r's conserved vector
                ev_add(*acc, eh, owned(r));
                eh = Ev{};
            }
            if (r >= i0) {
                double F[5];
                force_flux<Phys, 1, W0>(Up, FUp, U, FU, cf, P, F, pi);
                if (kCount && acc) {
                    // interface (r - 1 | r)
                    ev_add(*acc, ei, owned_iface(r - 1));
                    ei = Ev{};
                }
                if (r >= i0 + 1) {
                    double Wo[5];
                    update_cell<Phys, W0>(Up, Fp, F, cf, P, Wo, pu);
                    if (kCount && acc) {
                        ev_add(*acc, eu, colok);
                        eu = Ev{};
                    }
                    if (colok) store_cell<W0>(out, (long)(r - 1) * pitch + j, Wo);
                    if (ws) {
                        double w = Phys::template wavespeed<2>(Wo, P);
                        if (colok && w > wsmax) wsmax = w;
                    }
                }
#pragma unroll
                for (int k = 0; k < 5; ++k) Fp[k] = F[k];
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                Up[k] = U[k];
                FUp[k] = FU[k];
            }
        }
    }
    return wsmax;
}

}  // namespace core
}  // namespace cfv
