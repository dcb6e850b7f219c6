// gfx950 kernels for the general-relativistic (Kerr-Schild black hole) path:
// FORCE / SLIC sweeps with excision, the RK4 source-term forcing step and the
// maximum wave speed.
//
// Layout: the same SoA primitive planes as the Newtonian and special-
// relativistic sweeps.  Cell (i, j) of a 2D grid sits at (j*h, i*h), cell j of
// a 1D grid at (j*h, 0) (Solvers/BlackHoleEulerSecondOrderSolver.cpp:37-143).
// Interfaces sit at (c -/+ 0.5)*h.  Every coordinate is formed with the
// expression the reference uses at that call site, so position-dependent
// metric values are bitwise those of the reference.
//
// A sweep is two launches.  k_gr_iface computes each interface flux once
// (the reference computes it twice, once from each neighbouring cell, with
// identical inputs) and only where a neighbouring cell is updated;
// k_gr_update applies the conservative update where the cell's stencil is
// clear of the excision region and copies the cell otherwise.  The GR work
// per interface is thousands of FP64 operations (metric evaluations and
// Newton con2prim solves), so the flux round trip through HBM (80 B per
// interface) is negligible and the kernels are FP64-issue-bound.
//
// Geometry cache (GEO = 1, where the host found the half lattice of h exact;
// the position-only values then come from the tables GeoTab holds instead of
// being re-derived at each call site, bitwise the same): every Geo of the
// sweeps and the forcing, each forcing cell's SourceGeo, and the forcing's
// neighbour states.  A forcing cell converts each neighbour to conserved
// variables at the neighbour's centre, back to primitives at its own x -/+ h
// (y -/+ h), and to conserved again there in every fractional sweep; on the
// exact lattice those positions are the neighbour's centre, so that chain is
// a function of the neighbour alone, formed once per cell by k_gr_roundtrip
// (tab.q) instead of four times per neighbour per step.
#include "fv_gr_ops.hpp"

namespace cfv {
namespace {

constexpr int kGRThreads = 256;
// The forcing runs at one wave per SIMD (its RK4 chain holds ~490 registers
// per lane): one-wave blocks let each SIMD start its next wave as soon as its
// own finishes, instead of waiting for the slowest of a 4-wave block.
constexpr int kForceThreads = 64;
// The interface fluxes run at two waves per SIMD; one-wave blocks likewise.
constexpr int kIfaceThreads = 64;

__device__ __forceinline__ int mirror2_gr(int c, int n) {
    if (c < 0) c = -c - 1;
    if (c >= n) c = 2 * n - 1 - c;
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}

__device__ __forceinline__ int clamp1_gr(int c, int n) {
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}

__device__ __forceinline__ void load5(const ConstPlanes& in, long idx, double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) W[k] = in.p[k][idx];
}

__device__ __forceinline__ void store5(const Planes& out, long idx, const double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) out.p[k][idx] = W[k];
}

// A cell is updated only if it and its neighbours within w cells along x
// (and, in 2D, along y, for either sweep direction) lie outside the excision
// region (Solvers/BlackHoleEulerSecondOrderSolver.cpp:50-52, 81-85, 115-119;
// Solvers/BlackHoleEulerFirstOrderSolver.cpp:102-103, 131-133, 160-162;
// Solvers/BlackHoleEulerForcingSolver.cpp:195-197, 228-231).
__device__ __forceinline__ bool stencil_clear(const Hole& b, int dim, int i, int j, int w, double h) {
    double y = dim == 2 ? i * h : 0.0;
    if (excised(b, j * h, y, 0.0)) return false;
    for (int d = 1; d <= w; ++d) {
        if (excised(b, (j - d) * h, y, 0.0) || excised(b, (j + d) * h, y, 0.0)) return false;
        if (dim == 2 && (excised(b, j * h, (i - d) * h, 0.0) || excised(b, j * h, (i + d) * h, 0.0))) return false;
    }
    return true;
}

// One flux per interface.  x sweep: thread -> (row i, interface c), c in [0, nx];
// y sweep: thread -> (interface c, column j), c in [0, ny].  Interface c lies
// between cells c - 1 and c at (c - 0.5) * h.
template <int ORDER, int AXIS, int GEO>
__global__ __launch_bounds__(kIfaceThreads) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_gr_iface(GRLaunch L) {
    if (gr_stopped(L)) return;
    const SweepCoef cf = gr_step_of(L).cf;
    const int n = AXIS == 0 ? L.nx : L.ny;
    const long lines = AXIS == 0 ? L.ny : L.nx;
    const long per = AXIS == 0 ? (long)(n + 1) : lines;
    const long total = (long)(n + 1) * lines;
    long idx = (long)blockIdx.x * kIfaceThreads + threadIdx.x;
    if (idx >= total) return;
    int c, i, j;
    if (AXIS == 0) {
        i = (int)(idx / per);
        c = (int)(idx - (long)i * per);
        j = c;
    } else {
        c = (int)(idx / per);
        j = (int)(idx - (long)c * per);
        i = c;
    }
    const int gh = ORDER == 2 ? 2 : 1;
    // needed when a neighbouring cell is updated
    bool need;
    if (AXIS == 0)
        need = (c > 0 && stencil_clear(L.hole, L.dim, i, c - 1, gh, L.h)) ||
               (c < n && stencil_clear(L.hole, L.dim, i, c, gh, L.h));
    else
        need = (c > 0 && stencil_clear(L.hole, L.dim, c - 1, j, gh, L.h)) ||
               (c < n && stencil_clear(L.hole, L.dim, c, j, gh, L.h));
    if (!need) return;
    auto at = [&](int cc) -> long {
        int m = ORDER == 2 ? mirror2_gr(cc, n) : clamp1_gr(cc, n);
        return AXIS == 0 ? (long)i * L.pitch + m : (long)m * L.pitch + j;
    };
    double x, y;
    int ax, ay;  // half-lattice point of the interface
    if (AXIS == 0) {
        x = (c - 0.5) * L.h;
        y = L.dim == 2 ? i * L.h : 0.0;
        ax = 2 * c - 1;
        ay = L.dim == 2 ? 2 * i : 0;
    } else {
        x = j * L.h;
        y = (c - 0.5) * L.h;
        ax = 2 * j;
        ay = 2 * c - 1;
    }
    double F[5];
    if (ORDER == 1) {
        double WL[5], WR[5];
        load5(L.in, at(c - 1), WL);
        load5(L.in, at(c), WR);
        gr_force<AXIS, GEO>(WL, WR, x, y, ax, ay, L, cf, F);
    } else {
        // computeXSLICFlux (Solvers/BlackHoleEulerSecondOrderSolver.cpp:3-35): the left
        // cell's right-face state evolved about the interface - h/2, the right cell's
        // left-face state about the interface + h/2, then FORCE at the interface.
        const double hh = L.h * 0.5;
        double W0[5], W1[5], W2[5], pr[5], pl[5];
        load5(L.in, at(c - 2), W0);
        load5(L.in, at(c - 1), W1);
        load5(L.in, at(c), W2);
        // the two face states are independent: both evolved, then their
        // primitive recoveries run as one interleaved pair
        double er[5], el[5], sr, sl;
        if (AXIS == 0) gr_half_evolve<AXIS, GEO>(W0, W1, W2, x - hh, y, ax - 1, ay, 1, L, cf, er, sr);
        else gr_half_evolve<AXIS, GEO>(W0, W1, W2, x, y - hh, ax, ay - 1, 1, L, cf, er, sr);
        load5(L.in, at(c + 1), W0);
        if (AXIS == 0) gr_half_evolve<AXIS, GEO>(W1, W2, W0, x + hh, y, ax + 1, ay, 0, L, cf, el, sl);
        else gr_half_evolve<AXIS, GEO>(W1, W2, W0, x, y + hh, ax, ay + 1, 0, L, cf, el, sl);
        gr_c2p2(er, sr, el, sl, L.P, pr, pl);
        gr_force<AXIS, GEO>(pr, pl, x, y, ax, ay, L, cf, F);
        if (L.half) {
            const long f = AXIS == 0 ? (long)i * L.fpitch + c : (long)c * L.fpitch + j;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                L.half[k * L.half_plane + f] = pr[k];
                L.half[(5 + k) * L.half_plane + f] = pl[k];
            }
        }
    }
    long f = AXIS == 0 ? (long)i * L.fpitch + c : (long)c * L.fpitch + j;
#pragma unroll
    for (int k = 0; k < 5; ++k) L.flux.p[k][f] = F[k];
}

// The conservative update of each cell (computeFORCEUpdate with p2c / c2p at
// the cell centre) where its stencil is clear; a copy elsewhere.
template <int ORDER, int AXIS, int GEO>
__global__ __launch_bounds__(kGRThreads) void k_gr_update(GRLaunch L) {
    if (gr_stopped(L)) return;
    const double dt_dx = gr_step_of(L).cf.dt_dx;
    long idx = (long)blockIdx.x * kGRThreads + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return;
    int i = (int)(idx / L.nx), j = (int)(idx - (long)i * L.nx);
    long cell = (long)i * L.pitch + j;
    double W[5];
    load5(L.in, cell, W);
    if (stencil_clear(L.hole, L.dim, i, j, ORDER == 2 ? 2 : 1, L.h)) {
        Geo g = geo_pick<GEO>(L, j * L.h, L.dim == 2 ? i * L.h : 0.0, 2 * j, L.dim == 2 ? 2 * i : 0);
        double U[5], un[5];
        gr_p2c(W, g, U, L.P);
        // x: interfaces j, j + 1 of row i; y: interfaces i, i + 1 of column j
        long fl = (long)i * L.fpitch + j;
        long fr = AXIS == 0 ? fl + 1 : fl + L.fpitch;
#pragma unroll
        for (int k = 0; k < 5; ++k) un[k] = U[k] + (dt_dx * (L.flux.p[k][fl] - L.flux.p[k][fr]));
        gr_c2p(un, g.sdet, L.P, W);
    }
    store5(L.out, cell, W);
}

// The neighbour chain of a forcing cell on the exact lattice (GEO): for every
// cell c of [-1, nx] x [-1, ny] (its data clamped into the grid, its geometry
// at its own centre), p2c(c2p(p2c(W_c))) with the Geo of that centre: what the
// fractional sweeps of a neighbouring cell convert it to
// (computeRungeKuttaTimeStep{,2D} p2c at the neighbour's centre, c2p at x -/+ h
// inside evolveConservedVariableVector{,2D}, p2c at x -/+ h in each sweep).
template <int DIM>
__global__ __launch_bounds__(kGRThreads) void k_gr_roundtrip(GRLaunch L, double* q) {
    if (gr_stopped(L)) return;
    const int qw = L.nx + 2, qh = DIM == 2 ? L.ny + 2 : 1;
    long idx = (long)blockIdx.x * kGRThreads + threadIdx.x;
    if (idx >= (long)qw * qh) return;
    const int r = (int)(idx / qw), cc = (int)(idx - (long)r * qw);
    const int i = DIM == 2 ? r - 1 : 0, j = cc - 1;
    // the forcing reads this chain only from cells whose stencil is clear, so
    // never at an excised position (whose recoveries near the singularity can
    // run to the iteration cap): skipped there
    if (excised(L.hole, j * L.h, DIM == 2 ? i * L.h : 0.0, 0.0)) return;
    double W[5], U[5], R[5];
    load5(L.in, (long)clamp1_gr(i, L.ny) * L.pitch + clamp1_gr(j, L.nx), W);
    const Geo g = geo_from_table(L.tab, DIM, 2 * j, 2 * i);
    gr_p2c(W, g, U, L.P);
    gr_c2p(U, g.sdet, L.P, R);
    gr_p2c(R, g, U, L.P);
    const long o = (long)r * L.tab.qpitch + cc;
#pragma unroll
    for (int k = 0; k < 5; ++k) q[k * L.tab.qplane + o] = U[k];
}

// computeRungeKuttaTimeStep{,2D} (Solvers/BlackHoleEulerForcingSolver.cpp:187-250) with
// evolveConservedVariableVector{,2D} (:3-185): per cell, RK4 on the source term,
// each stage's middle state first evolved by fractional flux sweeps against
// the step-start neighbours (1D: x by 1/2, 1/2, 1; 2D: x 1/4 y 1/2 x 1/4,
// twice, then x 1/2 y 1 x 1/2), U + (k1 + 2 k2 + 2 k3 + k4) / 6.
template <int DIM, int GEO>
__global__ __launch_bounds__(kForceThreads) void k_gr_forcing(GRLaunch L) {
    if (gr_stopped(L)) return;
    const GRStep S = gr_step_of(L);
    long idx = (long)blockIdx.x * kForceThreads + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return;
    int i = (int)(idx / L.nx), j = (int)(idx - (long)i * L.nx);
    long cell = (long)i * L.pitch + j;
    double W[5];
    load5(L.in, cell, W);
    const Hole& b = L.hole;
    const double h = L.h;
    if (!stencil_clear(b, DIM, i, j, 1, h)) {
        store5(L.out, cell, W);
        return;
    }
    const double x = j * h, y = DIM == 2 ? i * h : 0.0;
    double uM[5];
    SourceGeo sg;
    if constexpr (GEO == 1) {
        double src[5];
        // every position-only value and the neighbours' chain read once, at the
        // start, from the tables (tab.geo, tab.src, tab.q); the RK4 chain then
        // runs on registers only
        const int ax = 2 * j, ay = DIM == 2 ? 2 * i : 0;
        const Geo gc = geo_from_table(L.tab, DIM, ax, ay);
        Geo gn[4];  // x-, x+, y-, y+ neighbours' centres (flux geometry)
        gn[0] = geo_from_table(L.tab, DIM, ax - 2, ay);
        gn[1] = geo_from_table(L.tab, DIM, ax + 2, ay);
        if (DIM == 2) {
            gn[2] = geo_from_table(L.tab, DIM, ax, ay - 2);
            gn[3] = geo_from_table(L.tab, DIM, ax, ay + 2);
        }
        double qn[4][5];
        {
            const long qo = (long)(DIM == 2 ? i + 1 : 0) * L.tab.qpitch + (j + 1);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double* q = L.tab.q + k * L.tab.qplane + qo;
                qn[0][k] = q[-1];
                qn[1][k] = q[1];
                if (DIM == 2) {
                    qn[2][k] = q[-L.tab.qpitch];
                    qn[3][k] = q[L.tab.qpitch];
                }
            }
        }
        source_from_table(L.tab.src + (long)i * L.tab.spitch + j, L.tab.splane, sg);
        gr_p2c(W, gc, uM, L.P);
        gr_source(uM, sg, L.P, src);
        RK4 rk;
        rk.start(src, S);
        for (int st = 0; st < 3; ++st) {
            double tmp[5], Wm[5], e1[5], um[5];
            rk.middle(st, uM, tmp);
            gr_c2p(tmp, gc.sdet, L.P, Wm);
            if (DIM == 1) {
                gr_fractional_core<0>(rk.fx, qn[0], Wm, qn[1], gn[0], gc, gn[1], L, e1);
            } else {
                double e2[5], e3[5];
                gr_fractional_core<0>(rk.fx, qn[0], Wm, qn[1], gn[0], gc, gn[1], L, e2);
                gr_fractional_core<1>(rk.fy, qn[2], e2, qn[3], gn[2], gc, gn[3], L, e3);
                gr_fractional_core<0>(rk.fx, qn[0], e3, qn[1], gn[0], gc, gn[1], L, e1);
            }
            gr_p2c(e1, gc, um, L.P);
            gr_source(um, sg, L.P, src);
            rk.next(st, src, S);
        }
        double un[5];
        rk.finish(uM, un);
        gr_c2p(un, gc.sdet, L.P, W);
        store5(L.out, cell, W);
        return;
    }
    // neighbours' conserved vectors, each formed at its own centre
    // (the grid loop of computeRungeKuttaTimeStep{,2D})
    double uL[5], uR[5], uT[5], uB[5];
    {
        double Wn[5];
        load5(L.in, (long)i * L.pitch + clamp1_gr(j - 1, L.nx), Wn);
        gr_p2c(Wn, geo_at(b, (j - 1) * h, y, 0.0), uL, L.P);
        load5(L.in, (long)i * L.pitch + clamp1_gr(j + 1, L.nx), Wn);
        gr_p2c(Wn, geo_at(b, (j + 1) * h, y, 0.0), uR, L.P);
        if (DIM == 2) {
            load5(L.in, (long)clamp1_gr(i - 1, L.ny) * L.pitch + j, Wn);
            gr_p2c(Wn, geo_at(b, x, (i - 1) * h, 0.0), uT, L.P);
            load5(L.in, (long)clamp1_gr(i + 1, L.ny) * L.pitch + j, Wn);
            gr_p2c(Wn, geo_at(b, x, (i + 1) * h, 0.0), uB, L.P);
        }
    }
    const Geo gc = geo_at(b, x, y, 0.0);
    gr_p2c(W, gc, uM, L.P);
    if (L.tab.src) source_from_table(L.tab.src + (long)i * L.tab.spitch + j, L.tab.splane, sg);
    else source_geo(b, x, y, 0.0, h, sg);
    double un[5];
    gr_rk4_cons<DIM, 0>(uL, uM, uR, uT, uB, x, y, gc, sg, L, S, un);
    gr_c2p(un, gc.sdet, L.P, W);
    store5(L.out, cell, W);
}

// The geometry tables: geo_at at every half-lattice point of the three
// classes (centres, x faces, y faces) over the grid plus a two-cell halo, at
// the exact position (ax h / 2, ay h / 2); source_geo at every cell centre
// with the forcing's own expressions x = j h, y = i h.
__global__ __launch_bounds__(kGRThreads) void k_gr_geo_build(GRLaunch L, double* geo, long pitch, long plane,
                                                             int rows) {
    const long per = (long)rows * pitch;
    long idx = (long)blockIdx.x * kGRThreads + threadIdx.x;
    if (idx >= 3 * per) return;
    const int cls = (int)(idx / per);
    const long rem = idx - (long)cls * per;
    const int row = (int)(rem / pitch), col = (int)(rem - (long)row * pitch);
    const int ax = 2 * col - 4 + (cls & 1);
    const int ay = L.dim == 2 ? 2 * row - 4 + (cls >> 1) : 0;
    if (L.dim == 1 && cls == 2) return;
    const double hh = 0.5 * L.h;
    const Geo G = geo_at(L.hole, ax * hh, ay * hh, 0.0);
    geo_to_table(G, geo + (long)cls * kGeoVals * plane + (long)row * pitch + col, plane);
}

__global__ __launch_bounds__(kGRThreads) void k_gr_src_build(GRLaunch L, double* src, long spitch, long splane) {
    long idx = (long)blockIdx.x * kGRThreads + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return;
    const int i = (int)(idx / L.nx), j = (int)(idx - (long)i * L.nx);
    const double x = j * L.h, y = L.dim == 2 ? i * L.h : 0.0;
    SourceGeo S;
    source_geo(L.hole, x, y, 0.0, L.h, S);
    source_to_table(S, src + (long)i * spitch + j, splane);
}

// computeMaximumWaveSpeed{,2D} (Solvers/BlackHoleEulerSolvers.cpp:88-140): the
// reference places cells on the unit square whatever the grid spacing
// (1/N per cell), tests excision at (j/nx, i/ny) but evaluates the sound
// speed at (i/ny, j/nx), and takes max |c| over non-excised cells
// (NaN never enters the running max).
template <int DIM>
__global__ __launch_bounds__(kGRThreads) void k_gr_wavespeed(GRLaunch L, double dr, double dc,
                                                             unsigned long long* slots) {
    if (gr_stopped(L)) return;
    double m = 0.0;
    long total = (long)L.nx * L.ny;
    for (long idx = (long)blockIdx.x * kGRThreads + threadIdx.x; idx < total;
         idx += (long)gridDim.x * kGRThreads) {
        int i = (int)(idx / L.nx), j = (int)(idx - (long)i * L.nx);
        double W[5];
        load5(L.in, (long)i * L.pitch + j, W);
        double ws;
        if (DIM == 1) {
            if (excised(L.hole, j * dc, 0.0, 0.0)) continue;
            ws = __builtin_fabs(gr_sound_speed(W, L.hole, j * dc, 0.0, 0.0, L.P));
        } else {
            if (excised(L.hole, j * dc, i * dr, 0.0)) continue;
            ws = __builtin_fabs(gr_sound_speed(W, L.hole, i * dr, j * dc, 0.0, L.P));
        }
        if (ws > m) m = ws;
    }
    block_max_to<kGRThreads>(m, slots);
}

dim3 blocks_for(long work) { return dim3((unsigned)((work + kGRThreads - 1) / kGRThreads)); }
dim3 force_blocks(long cells) { return dim3((unsigned)((cells + kForceThreads - 1) / kForceThreads)); }

}  // namespace

namespace {

template <int GEO>
hipError_t launch_gr_sweep_t(const GRLaunch& L, int axis, hipStream_t s) {
    const dim3 block(kGRThreads);
    const long ifaces = axis == 0 ? (long)(L.nx + 1) * L.ny : (long)(L.ny + 1) * L.nx;
    const dim3 iblock(kIfaceThreads), igrid((unsigned)((ifaces + kIfaceThreads - 1) / kIfaceThreads));
    if (L.order == 2) {
        if (axis == 0) hipLaunchKernelGGL((k_gr_iface<2, 0, GEO>), igrid, iblock, 0, s, L);
        else hipLaunchKernelGGL((k_gr_iface<2, 1, GEO>), igrid, iblock, 0, s, L);
    } else {
        if (axis == 0) hipLaunchKernelGGL((k_gr_iface<1, 0, GEO>), igrid, iblock, 0, s, L);
        else hipLaunchKernelGGL((k_gr_iface<1, 1, GEO>), igrid, iblock, 0, s, L);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long cells = (long)L.nx * L.ny;
    if (L.order == 2) {
        if (axis == 0) hipLaunchKernelGGL((k_gr_update<2, 0, GEO>), blocks_for(cells), block, 0, s, L);
        else hipLaunchKernelGGL((k_gr_update<2, 1, GEO>), blocks_for(cells), block, 0, s, L);
    } else {
        if (axis == 0) hipLaunchKernelGGL((k_gr_update<1, 0, GEO>), blocks_for(cells), block, 0, s, L);
        else hipLaunchKernelGGL((k_gr_update<1, 1, GEO>), blocks_for(cells), block, 0, s, L);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gr_sweep(const GRLaunch& L, int axis, hipStream_t s) {
    return L.tab.geo ? launch_gr_sweep_t<1>(L, axis, s) : launch_gr_sweep_t<0>(L, axis, s);
}

hipError_t launch_gr_forcing(const GRLaunch& L, hipStream_t s) {
    const long cells = (long)L.nx * L.ny;
    if (L.tab.geo && L.tab.src && L.tab.q) {
        // the neighbour chain once per cell, then the forcing
        const long qn = (long)(L.nx + 2) * (L.dim == 2 ? L.ny + 2 : 1);
        double* q = const_cast<double*>(L.tab.q);
        if (L.dim == 1) hipLaunchKernelGGL((k_gr_roundtrip<1>), blocks_for(qn), dim3(kGRThreads), 0, s, L, q);
        else hipLaunchKernelGGL((k_gr_roundtrip<2>), blocks_for(qn), dim3(kGRThreads), 0, s, L, q);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (L.dim == 1) hipLaunchKernelGGL((k_gr_forcing<1, 1>), force_blocks(cells), dim3(kForceThreads), 0, s, L);
        else hipLaunchKernelGGL((k_gr_forcing<2, 1>), force_blocks(cells), dim3(kForceThreads), 0, s, L);
        return hipGetLastError();
    }
    if (L.dim == 1) hipLaunchKernelGGL((k_gr_forcing<1, 0>), force_blocks(cells), dim3(kForceThreads), 0, s, L);
    else hipLaunchKernelGGL((k_gr_forcing<2, 0>), force_blocks(cells), dim3(kForceThreads), 0, s, L);
    return hipGetLastError();
}

hipError_t launch_gr_geo_build(const GRLaunch& L, double* geo, double* src, hipStream_t s) {
    if (geo) {
        const long rows = L.dim == 2 ? L.ny + 4 : 1;
        hipLaunchKernelGGL(k_gr_geo_build, blocks_for(3 * rows * L.tab.pitch), dim3(kGRThreads), 0, s, L, geo,
                           L.tab.pitch, L.tab.plane, (int)rows);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (src) {
        hipLaunchKernelGGL(k_gr_src_build, blocks_for((long)L.nx * L.ny), dim3(kGRThreads), 0, s, L, src,
                           L.tab.spitch, L.tab.splane);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_gr_wavespeed(const GRLaunch& L, unsigned long long* slots, hipStream_t s) {
    const long cells = (long)L.nx * L.ny;
    long nb = (cells + kGRThreads - 1) / kGRThreads;
    if (nb > 8192) nb = 8192;
    // 1.0 / rowCount and 1.0 / columnCount (1D: 1.0 / cellCount)
    const double dr = 1.0 / (double)L.ny, dc = 1.0 / (double)L.nx;
    if (L.dim == 1) hipLaunchKernelGGL((k_gr_wavespeed<1>), dim3((unsigned)nb), dim3(kGRThreads), 0, s, L, dr, dc, slots);
    else hipLaunchKernelGGL((k_gr_wavespeed<2>), dim3((unsigned)nb), dim3(kGRThreads), 0, s, L, dr, dc, slots);
    return hipGetLastError();
}

}  // namespace cfv
