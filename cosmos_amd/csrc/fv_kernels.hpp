// Launch interface between the C-ABI host layer and the gfx950 sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_gr.hpp"
#include "fv_physics.hpp"

namespace cfv {

// Running maxima (of non-negative doubles, as IEEE bit patterns) are merged
// with integer atomics into kMaxSlots words, one per 128-byte line, picked by
// block index: tens of thousands of blocks then contend on 32 addresses
// instead of one.  Readers take the max over the slots.
constexpr int kMaxSlots = 32;
constexpr int kSlotStride = 16;                 // u64 words per slot (128 B)
constexpr int kSlotWords = kMaxSlots * kSlotStride;

// Block-wide max of non-negative wave speeds, merged into one of the slots
// with an integer atomic max (the IEEE bit patterns of non-negative doubles
// order like the values).  The reference's running max skips NaN
// (ws > max is false), so NaN never enters.
template <int NT>
__device__ __forceinline__ void block_max_to(double m, unsigned long long* slots) {
    __shared__ double red[NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = (o > m) ? o : m;
    }
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (NT > 64) {
        if (lane == 0) red[wave] = m;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double b = m;
        if (NT > 64) {
            b = red[0];
#pragma unroll
            for (int w = 1; w < NT / 64; ++w) b = (red[w] > b) ? red[w] : b;
        }
        if (b > 0.0)
            atomicMax(slots + (blockIdx.x % kMaxSlots) * kSlotStride, (unsigned long long)__double_as_longlong(b));
    }
}

// PhysParams of an adiabatic index, each constant formed with the reference's
// expression (RelativisticEuler/RelativisticEulerStateVector.cpp:52-57).
inline PhysParams make_params(double g) {
    PhysParams P{};
    P.gamma = g;
    P.gm1 = g - 1.0;
    P.c2p_tol = 1e-8;
    P.c2p_maxit = 1000;
    P.sr_c0 = -1.0 / (g * g);
    P.sr_c1f = (g - 1.0) / (g * g);
    P.sr_gm1g = (g - 1.0) / g;
    P.sr_gm2g = (g - 2.0) / g;
    P.sr_gg1 = g / (g - 1.0);
    P.sr_4gm1g = 4.0 * ((g - 1.0) / g);
    P.inv_gm1 = 1.0 / (g - 1.0);
    return P;
}

// Per-sweep coefficients: dt/dx, dx/dt and 0.5*(dt/dx), with dt the sweep's
// step (0.5*dt for the outer Strang x sweeps).  Formed exactly as the
// reference forms them at each use (Solvers/EulerSolvers.cpp:163-172,
// Solvers/EulerFirstOrderSolver.cpp:3-19, 96-101).
struct SweepCoef {
    double dt_dx;
    double dx_dt;
    double hdt_dx;
    double pad;
};

// Device-resident time-stepping state for solve/solve2D-style loops
// (Solvers/EulerSecondOrderSolver.cpp:133-178): t, the iteration counter
// (which drives the 0.2x damping of the first six steps), the final time and
// the step cap.  The sweeps of a step read their coefficients from here, so
// an N-step loop needs no host round trip and can be captured in a graph.
struct StepState {
    double t;
    double tf;
    double dt;
    double cfl;
    double dx;
    unsigned long long* ws_slots;  // max wave speed of the state at cosmos_begin (kSlotWords words)
    double* ws_part;               // per-block maxima written by the last sweep of each step
    int ws_nparts;                 // entries of ws_part to read at the next step head
    int ws_nparts_sweep;           // blocks of the last sweep of a step (0 with slot atomics)
    int ws_slot_atomics;           // the last sweep merges per-wave maxima into ws_slots
    int iter;
    int max_steps;
    int done;
    int active_steps;            // steps actually taken since the loop began
    unsigned long long* status;  // the solver's status block (null: none); [5] set on a NaN time step
    SweepCoef half;              // the outer x sweeps of a 2D step: 0.5 * dt
    SweepCoef full;              // the y sweep of a 2D step and the 1D sweep: dt
    // black-hole loops (gr != 0): the forcing's step 0.5 * (dt / subcycles) and its
    // RK4 stage fractions times that step over dx (Solvers/BlackHoleEulerForcingSolver.cpp)
    int gr;
    int gr_subcycles;
    int gr_dim;
    int pad;
    double gr_dt;
    double gr_frac[3];           // x (1D: 1/2 1/2 1; 2D: 1/4 1/4 1/2)
    double gr_frac_y[3];         // 2D y: 1/2 1/2 1
};

struct Planes {
    double* p[5];
};

struct ConstPlanes {
    const double* p[5];
};

struct SweepLaunch {
    ConstPlanes in;
    Planes out;
    int nx, ny;
    long pitch;                 // doubles per row
    double bias;
    int limiter;
    int order;                  // 1 = FORCE, 2 = SLIC
    int physics;                // 0 Newtonian, 1 SR
    int dim;                    // 1 or 2
    int w_zero;                 // the z-velocity plane is identically zero: skip it
    int fast;                   // fast (re-associated, not bitwise) Newtonian SLIC / van Leer kernels
    PhysParams P;
    SweepCoef coef;             // used when step == nullptr
    StepState* step;            // device state: coefficients and done flag
    int use_full;               // which StepState coefficients (0 half, 1 full)
    int fuse_ws;                // merge the max wave speed of the output into step->ws_slots
    int y_rows;                 // rows per y-sweep thread (16, 32, 64; 0: by physics)
    int fuse_rows;  // fused step: output rows per block (0: chosen from the occupancy)
    int y_tile;                 // y-sweep layout: 0 column march, 8 or 16 columns per LDS tile; -1: by physics
    unsigned long long* status; // special relativity: con2prim event counters (cap, q, rho, p floors), or null
    // Special relativity in a solve loop (else null): the dispatch order of the
    // blocks, those whose solves ran to the iteration cap in the previous step
    // first (their wandering Newton loops then overlap the other blocks instead
    // of trailing the launch), and the marks for the next step's order.  Used
    // when order_n equals the launch's grid; any order gives the same results.
    const int* block_order;
    int order_n;
    unsigned char* hot;
    // AMR masked sweeps (else null): the level's refinement flags ([ny][nx]
    // bytes).  Only flagged cells are written to out; a block whose tile holds
    // no flagged output cell returns at once.  The y sweep then runs as LDS
    // tiles (the column march takes no mask).  With a mask, block_order may
    // list the blocks that hold a flagged output cell (from launch_mask_lists):
    // the launch then runs order_n blocks, block k the listed block_order[k],
    // and none at all when order_n is 0.
    const unsigned char* mask;
};

// All return hipSuccess or the launch error.
hipError_t launch_sweep_x(const SweepLaunch& L, hipStream_t s);
hipError_t launch_sweep_y(const SweepLaunch& L, hipStream_t s);
// dst[c] = src[c] for every cell c whose mask byte is set (the five planes of
// an nx-by-ny grid of row pitch `pitch`): the write-back of a masked sweep.
// segs (else null): the n_segs 64-cell segments (row-major cell order)
// holding a flagged cell (from launch_mask_lists); only those are visited.
hipError_t launch_mask_writeback(const Planes& dst, const ConstPlanes& src, const unsigned char* mask, int nx, int ny,
                                 long pitch, const int* segs, int n_segs, hipStream_t s);
// The blocks of a masked level's launches that hold a flagged cell of mask
// ([ny][nx] bytes), for sweeps of this order (1 or 2) and y-tile width tc
// (8 or 16; 2D only): counts[0] x-sweep row segments into xl, counts[1]
// y-sweep tiles into yl, counts[2] 64-cell write-back segments into wl, each
// list in no particular order (the blocks write disjoint cells).  counts are
// cleared first.  xl, yl and wl need mask_list_sizes() entries; hit is their
// total in bytes of mark maps, all zero on entry and left all zero.
hipError_t launch_mask_lists(const unsigned char* mask, int nx, int ny, int dim, int order, int tc,
                             unsigned char* hit, int* xl, int* yl, int* wl, int* counts, hipStream_t s);
// Candidate blocks of each list: x-sweep segments, y-sweep tiles (0 in 1D),
// 64-cell write-back segments.
void mask_list_sizes(int nx, int ny, int dim, int order, int tc, long sizes[3]);
// Columns of the masked y sweep's tiles for this launch record (8 or 16).
int mask_tile_cols(const SweepLaunch& L);
hipError_t launch_wavespeed(const ConstPlanes& in, int nx, int ny, long pitch, int physics, int dim,
                            const PhysParams& P, unsigned long long* slots, hipStream_t s);
// The wave-speed maximum of a masked AMR level (Newtonian or special
// relativistic), in two forms that give launch_wavespeed's slots: split
// reads every cell and also leaves the maximum over the unflagged cells in
// the slots rest (kSlotWords words, zeroed by the caller); segs reads only
// the flagged cells of the n_segs listed 64-cell segments and merges rest,
// valid while no unflagged cell has changed since the split pass.
hipError_t launch_wavespeed_split(const ConstPlanes& in, int nx, int ny, long pitch, int physics, int dim,
                                  const PhysParams& P, const unsigned char* mask, unsigned long long* slots,
                                  unsigned long long* rest, hipStream_t s);
hipError_t launch_wavespeed_segs(const ConstPlanes& in, int nx, int ny, long pitch, int physics, int dim,
                                 const PhysParams& P, const unsigned char* mask, const int* segs, int n_segs,
                                 const unsigned long long* rest, unsigned long long* slots, hipStream_t s);
hipError_t launch_step_begin(StepState* st, hipStream_t s);
// The dispatch order of an n-block sweep from its blocks' marks: the marked
// blocks first, then the rest in the XCD-contiguous order; clears the marks.
// row > 0: a block also counts as marked when a neighbour in the block grid of
// row width `row` was (see k_hot_order); 0: its own mark only.
hipError_t launch_hot_order(unsigned char* hot, int* order, int n, int row, hipStream_t s);
// Blocks per row of the block grid of an x sweep (axis 0) or of a y-tile sweep
// (axis 1; 0 when the y sweep is the column march).
int sweep_row_blocks(const SweepLaunch& L, int axis);
// Blocks of the sweeps launch_sweep_x / launch_sweep_y run for L (0 when the
// y sweep is the column march, which takes no order).
int sweep_x_grid(const SweepLaunch& L);
int sweep_y_tile_grid(const SweepLaunch& L);
// One whole 2D step X(dt/2) Y(dt) X(dt/2) of the fast Newtonian SLIC / van
// Leer / bias 0 update in one kernel (coefficients from L.step, the wave-speed
// maximum merged into the step's slots).  Needs nx, ny >= 8 and slot atomics.
hipError_t launch_step_fused(const SweepLaunch& L, hipStream_t s);
// Blocks of an x-sweep launch (the number of per-block maxima it writes).
int sweep_x_blocks(int nx, int ny, int order);
hipError_t launch_zero_slots(unsigned long long* slots, hipStream_t s);
// p[0 .. n) = v (a fill kernel, no memset node).
hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s);
// *count += cells of the state holding a NaN (count zeroed by the caller).
hipError_t launch_nan_cells(const ConstPlanes& in, int nx, int ny, long pitch, unsigned long long* count,
                            hipStream_t s);
// flag[0] = 1 if any value of the nx-by-ny plane (row pitch `pitch`) is nonzero, else 0.
hipError_t launch_any_nonzero(const double* plane, int nx, int ny, long pitch, int* flag, hipStream_t s);

// ---------------------------------------------------------------------------
// General relativistic (Kerr-Schild) sweeps, forcing and wave speed.  Each
// sweep is two launches: the FORCE / SLIC flux of every interface into the
// flux planes, then the conservative update of every cell.
// ---------------------------------------------------------------------------
struct GRLaunch {
    ConstPlanes in;
    Planes out;
    Planes flux;                // interface fluxes: x sweep [ny][fpitch], y sweep [ny + 1][fpitch]
    double* half;               // SLIC sweeps, optional: the two evolved face states of each interface
                                // (10 planes of half_plane doubles in the flux layout), kept for the adjoint
    long half_plane;
    int nx, ny;
    long pitch;                 // doubles per state row
    long fpitch;                // doubles per flux row (>= nx + 1)
    int dim;                    // 1 or 2
    int order;                  // 1 = FORCE, 2 = SLIC
    int limiter;
    double bias;
    double h;                   // cell spacing (also the finite-difference spacing of the source)
    SweepCoef coef;             // from the sweep's (or the forcing's) time step
    double dt;                  // the forcing's time step
    double frac_dt_h[3];        // RK4 stage fractions times dt / h (1D: 1/2 1/2 1; 2D x: 1/4 1/4 1/2)
    double frac_dt_h_y[3];      // 2D y: 1/2 1/2 1
    PhysParams P;
    Hole hole;
    GeoTab tab;                 // geometry cache of this h (tab.geo null: evaluate at the call sites)
    double z;                   // z of the evaluation positions of the point-level functions (the
                                // sweeps and the forcing work in the z = 0 plane)
    const StepState* step;      // a device-resident loop: the step's coefficients come from here, and
                                // every launch returns at once after the loop stopped (null: coef, dt,
                                // frac_dt_h and frac_dt_h_y above)
    int use_full;               // with step: the sweep takes step->full (else step->half)
};

// The time-step coefficients of one black-hole launch: the sweep's and the
// forcing's (its step and RK4 stage fractions), from the launch record or,
// in a device-resident loop, from the step state.
struct GRStep {
    SweepCoef cf;
    double dt;
    double fx0, fx1, fx2;
    double fy0, fy1, fy2;
    CFV_HD double fx(int st) const { return st == 0 ? fx0 : (st == 1 ? fx1 : fx2); }
    CFV_HD double fy(int st) const { return st == 0 ? fy0 : (st == 1 ? fy1 : fy2); }
};

CFV_HD GRStep gr_step_of(const GRLaunch& L) {
    GRStep S;
    if (L.step) {
        const StepState& t = *L.step;
        S.cf = L.use_full ? t.full : t.half;
        S.dt = t.gr_dt;
        S.fx0 = t.gr_frac[0];
        S.fx1 = t.gr_frac[1];
        S.fx2 = t.gr_frac[2];
        S.fy0 = t.gr_frac_y[0];
        S.fy1 = t.gr_frac_y[1];
        S.fy2 = t.gr_frac_y[2];
    } else {
        S.cf = L.coef;
        S.dt = L.dt;
        S.fx0 = L.frac_dt_h[0];
        S.fx1 = L.frac_dt_h[1];
        S.fx2 = L.frac_dt_h[2];
        S.fy0 = L.frac_dt_h_y[0];
        S.fy1 = L.frac_dt_h_y[1];
        S.fy2 = L.frac_dt_h_y[2];
    }
    return S;
}

// A launch of a device-resident loop after the loop stopped does nothing.
CFV_HD bool gr_stopped(const GRLaunch& L) { return L.step && L.step->done; }

hipError_t launch_gr_sweep(const GRLaunch& L, int axis, hipStream_t s);
// Fill the geometry cache of L.h, L.hole: geo_at over the half lattice of the
// grid plus halo (tab.geo) and source_geo at the cell centres (tab.src).
hipError_t launch_gr_geo_build(const GRLaunch& L, double* geo, double* src, hipStream_t s);
hipError_t launch_gr_forcing(const GRLaunch& L, hipStream_t s);
hipError_t launch_gr_wavespeed(const GRLaunch& L, unsigned long long* slots, hipStream_t s);

// Max over the slots on the host (after copying the kSlotWords words back).
inline double slots_max_host(const unsigned long long* words) {
    unsigned long long best = 0;
    for (int k = 0; k < kMaxSlots; ++k)
        if (words[k * kSlotStride] > best) best = words[k * kSlotStride];
    double v;
    __builtin_memcpy(&v, &best, sizeof v);
    return v;
}

}  // namespace cfv
