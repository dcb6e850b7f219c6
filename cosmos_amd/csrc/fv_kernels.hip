// gfx950 kernels for the dimensionally split FORCE / SLIC sweeps.
//
// State of record: primitives (rho, u, v, w, p) in five SoA FP64 planes,
// row-major with a padded pitch, no ghost cells.  The transmissive boundary
// of insertBoundaryCells{,2D} (Solvers/EulerSolvers.cpp:3-85) is an index
// map applied at load time: two-ghost mirror order for SLIC
// (g[-1] = c[0], g[-2] = c[1], g[n] = c[n-1], g[n+1] = c[n-2]) and a one-ghost
// copy for FORCE.
//
// Per cell and sweep the reference evaluates the SLIC half-step of each
// cell twice and each interface flux twice (once from each neighbour), with
// identical inputs; here each is evaluated once:
//   stage A  U = p2c(W) for every cell in the tile plus halo
//   stage B  the half-step evolution of each cell -> the evolved right-face
//            and left-face states P1, P0 (stored as primitives, as the
//            reference does), their p2c Q and flux G
//   stage C  the FORCE flux of each interface from (Q1, G1) of its left cell
//            and (Q0, G0) of its right cell
//   stage D  the conservative update U + dt/dx (F_left - F_right) -> c2p
//
// x sweeps: one thread per cell, a 256-cell row segment per block, LDS for
// neighbour exchange between stages (252 outputs per block for SLIC).
// y sweeps: one thread per column, each marching down a segment of rows with a
// three-cell register window; every load and store is a coalesced row access.
//
// Kernels are templated on the physics (with its W0 flag: the z-velocity
// plane is identically zero and is neither read nor written; fast mode
// only), the order, the limiter mode (1: van Leer at bias 0 with one division
// per component, 2: the fast-mode van Leer, 0: any limiter and bias) and
// whether the output's max wave speed is merged.
#include "fv_sweep_core.hpp"

namespace cfv {
namespace {

using namespace core;

constexpr int kXThreads = 256;
constexpr int kYThreads = 64;

// The hardware hands consecutive workgroups to the 8 XCDs round robin, and
// each XCD has its own L2.  Renumber so that each XCD works through one
// contiguous range of tiles: neighbouring tiles, which share halo cache
// lines, then run on the same XCD at about the same time and the second
// reader hits in L2 instead of fetching the line from HBM again.  A
// bijection on [0, n) for any n.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int xcd = b % kXcds, k = b / kXcds;
    const int base = n / kXcds, extra = n % kXcds;
    return xcd * base + (xcd < extra ? xcd : extra) + k;
}

// Block-wide max of non-negative wave speeds stored to part[blockIdx.x] (no
// atomics; the step head reduces the per-block values).  red: NT / 64 doubles
// of LDS the caller no longer reads.
template <int NT>
__device__ __forceinline__ void block_max_store(double m, double* part, double* red) {
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
        part[blockIdx.x] = b;
    }
}

// Wave-wide max of non-negative wave speeds merged with one integer atomic
// max into slot (wave_id mod kMaxSlots) (no LDS, no barrier).
__device__ __forceinline__ void wave_max_to_slot(double m, unsigned long long* slots, int wave_id) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = (o > m) ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m > 0.0)
        atomicMax(slots + (wave_id % kMaxSlots) * kSlotStride, (unsigned long long)__double_as_longlong(m));
}

// A thread's con2prim event counts into the solver's status counters (events
// are rare: most threads add nothing).
__device__ __forceinline__ void status_add(unsigned long long* st, const Ev& e) {
    if (!st) return;
    if (e.cap) atomicAdd(st + 0, (unsigned long long)e.cap);
    if (e.q) atomicAdd(st + 1, (unsigned long long)e.q);
    if (e.rho) atomicAdd(st + 2, (unsigned long long)e.rho);
    if (e.p) atomicAdd(st + 3, (unsigned long long)e.p);
}

// The block this workgroup runs: from the launch's dispatch order where one
// is given for this grid, else the XCD-contiguous renumbering.
__device__ __forceinline__ int block_of(const SweepLaunch& L) {
    if (L.block_order && L.order_n == (int)gridDim.x) return L.block_order[blockIdx.x];
    return xcd_tile(blockIdx.x, gridDim.x);
}

__device__ __forceinline__ bool resolve_coef(const SweepLaunch& L, SweepCoef& cf) {
    cf = L.coef;
    if (L.step) {
        if (L.step->done) return false;
        cf = L.use_full ? L.step->full : L.step->half;
    }
    return true;
}

// ---------------------------------------------------------------------------
// x sweep: thread per cell, LDS exchange
// ---------------------------------------------------------------------------
template <class Phys, int ORDER, int LIMMODE, int DIM, bool WS, bool W0>
__global__ __launch_bounds__(kXThreads) void k_sweep_x(SweepLaunch L) {
    constexpr int NT = kXThreads;
    constexpr int G = ORDER == 2 ? 2 : 1;
    constexpr int TW = NT - 2 * G;
    // One LDS arena of 15 rows (30 KB: five blocks per CU).
    __shared__ double lds[15][NT];

    SweepCoef cf;
    if (!resolve_coef(L, cf)) return;
    const PhysParams& P = L.P;
    const int nx = L.nx;
    const int tiles = (nx + TW - 1) / TW;
    const int blk = block_of(L);
    const int row = blk / tiles;
    const int tile = blk - row * tiles;
    Ev acc{};
    Ev* const pacc = Phys::kPhysics == 1 ? &acc : nullptr;  // counted in registers, added only with L.status
    const double m = x_segment<Phys, ORDER, LIMMODE, DIM, WS, W0, NT>(
        L.in, L.out, nx, (long)row * L.pitch, tile, cf, L.bias, L.limiter, P, lds, pacc,
        L.mask ? L.mask + (long)row * nx : nullptr);
    if (Phys::kPhysics == 1) {
        status_add(L.status, acc);
        if (L.hot && acc.hit) L.hot[blk] = 1;
    }
    if constexpr (WS) {
        if (L.step->ws_slot_atomics) {
            wave_max_to_slot(m, L.step->ws_slots, blockIdx.x * (NT / 64) + (threadIdx.x >> 6));
        } else {
            // rows no longer read in stage D: sB for SLIC, sA for FORCE
            block_max_store<NT>(m, L.step->ws_part, &lds[ORDER == 2 ? 5 : 0][0]);
        }
    }
}

// ---------------------------------------------------------------------------
// y sweep: thread per column, marching down ROWS rows
// ---------------------------------------------------------------------------
// Minimum waves per SIMD, which caps the VGPRs: 4 (128 VGPRs) for the fast
// Newtonian sweep without the z-velocity plane, which fits without spilling.
// At that cap the fast sweep with the plane would spill 20 B per lane, the
// exact Newtonian sweep 44-68 B and the special-relativistic one 172-200 B
// (its con2prim state), so they may use up to 168 (3 waves), 168 and 256
// (2 waves) VGPRs.
template <class Phys, bool W0>
constexpr int y_min_waves() { return Phys::kPhysics == 1 ? 2 : (Phys::kFast && W0 ? 4 : 3); }

template <class Phys, int ORDER, int LIMMODE, bool WS, bool W0, int ROWS>
__global__ __launch_bounds__(kYThreads) __attribute__((amdgpu_waves_per_eu(y_min_waves<Phys, W0>(), 8))) void k_sweep_y(
    SweepLaunch L) {
    SweepCoef cf;
    if (!resolve_coef(L, cf)) return;
    const PhysParams& P = L.P;
    const int nx = L.nx, ny = L.ny;
    const int tx = (nx + kYThreads - 1) / kYThreads;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int by = blk / tx;
    const int bx = blk - by * tx;
    const int j = bx * kYThreads + threadIdx.x;
    const bool colok = j < nx;
    const int jj = colok ? j : nx - 1;
    const int i0 = by * ROWS;
    const int i1 = min(i0 + ROWS, ny);
    Ev acc{};
    Ev* const pacc = Phys::kPhysics == 1 ? &acc : nullptr;  // counted in registers, added only with L.status
    const double wsmax = y_march<Phys, ORDER, LIMMODE, W0>(L.in, L.out, ny, L.pitch, j, colok, jj, i0, i1, WS, cf,
                                                          L.bias, L.limiter, P, pacc);
    if (Phys::kPhysics == 1) status_add(L.status, acc);
    if constexpr (WS) {
        if (L.step->ws_slot_atomics) wave_max_to_slot(wsmax, L.step->ws_slots, blockIdx.x);
        else block_max_store<kYThreads>(wsmax, L.step->ws_part, nullptr);
    }
}

// ---------------------------------------------------------------------------
// y sweep: thread per cell over TC-column tiles, LDS exchange along columns
// (the x sweep's stages with the neighbours TC threads apart).  Each tile
// re-evaluates 2G halo rows, but a thread holds no marching window, so the
// compute-bound special-relativistic sweep runs at the x sweep's occupancy.
// ---------------------------------------------------------------------------
template <class Phys, int ORDER, int LIMMODE, int DIM, bool WS, bool W0, int TC>
__global__ __launch_bounds__(kXThreads) void k_sweep_y_tile(SweepLaunch L) {
    constexpr int NT = kXThreads;
    constexpr int G = ORDER == 2 ? 2 : 1;
    constexpr int TRV = NT / TC - 2 * G;  // output rows per tile
    __shared__ double lds[15][NT];
    SweepCoef cf;
    if (!resolve_coef(L, cf)) return;
    const int tc = (L.nx + TC - 1) / TC;
    const int blk = block_of(L);
    const int br = blk / tc;
    const int bc = blk - br * tc;
    Ev acc{};
    Ev* const pacc = Phys::kPhysics == 1 ? &acc : nullptr;  // counted in registers, added only with L.status
    const double m = y_tile_cells<Phys, ORDER, LIMMODE, W0, NT, TC>(L.in, L.out, L.nx, L.ny, L.pitch, bc * TC, br * TRV,
                                                                     WS ? DIM : 0, cf, L.bias, L.limiter, L.P, lds,
                                                                     pacc, L.mask);
    if (Phys::kPhysics == 1) {
        status_add(L.status, acc);
        if (L.hot && acc.hit) L.hot[blk] = 1;
    }
    if constexpr (WS) {
        if (L.step->ws_slot_atomics) wave_max_to_slot(m, L.step->ws_slots, blockIdx.x * (NT / 64) + (threadIdx.x >> 6));
        else block_max_store<NT>(m, L.step->ws_part, &lds[ORDER == 2 ? 5 : 0][0]);
    }
}

// ---------------------------------------------------------------------------
// One whole 2D step X(dt/2) Y(dt) X(dt/2) in one pass over HBM (fast mode).
// A block owns a strip of kFuseCols output columns (256 threads: 4 halo
// columns each side) and marches down RB output rows.  For each input row it
// runs the first x sweep on the row (LDS exchange, as k_sweep_x), feeds the
// result into the y sweep's three-row register window (as k_sweep_y), and
// every y output row goes straight through the second x sweep and out to HBM
// with the wave-speed maximum.  Per-cell arithmetic is the three separate
// sweeps', except that each sweep hands its updated conserved vector straight
// to the next (the separate sweeps store primitives and convert back), so the
// two differ by rounding; each cell is read and written once per step instead
// of three times, at the cost of recomputing the first x sweep on 4 halo rows
// per RB and every sweep on 8 halo columns per 256.
// Boundaries: columns through mirror2 at load (the first sweep's ghosts); the
// second sweep's ghost columns are the y outputs of the mirrored columns, which
// the y sweep produces from the first sweep's outputs of those columns: blocks
// at the grid's left or right edge copy them across through LDS after the first
// sweep.  Rows through mirror2 as in the y march.
// ---------------------------------------------------------------------------
constexpr int kFuseCols = kXThreads - 8;

template <class Phys, int LIMMODE, bool W0>
__global__ __launch_bounds__(kXThreads) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_step_fused(SweepLaunch L,
                                                                                                     int RB) {
    constexpr int NT = kXThreads;
    constexpr int NC = W0 ? 4 : 5;  // stored components
    __shared__ double lds[slic3r_lds_rows<W0>()][NT];
    if (L.step->done) return;
    const SweepCoef ch = L.step->half, cfu = L.step->full;
    const PhysParams& P = L.P;
    const int nx = L.nx, ny = L.ny;
    const long pitch = L.pitch;
    const int tx = (nx + kFuseCols - 1) / kFuseCols;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int by = blk / tx;
    const int bx = blk - by * tx;
    const int t = threadIdx.x;
    const int base = bx * kFuseCols - 4;
    const int p = base + t;              // the thread's column (may lie outside the grid)
    const int pc = mirror2(p, nx);       // the column it loads
    const bool edge = base < 0 || base + NT > nx;
    const int i0 = by * RB;
    const int i1 = min(i0 + RB, ny);

    // index of the thread's cell of (mirrored) input row rr
    auto at_row = [&](int rr) { return (long)mirror2(rr, ny) * pitch + pc; };
    // first x sweep of one input row (the thread's primitives Wr) -> Wx, its
    // updated conserved vector; the passes of line_step_3r need no barrier
    // between them
    auto x1_row = [&](const double* Wr, double* Wx) {
        line_step_3r<Phys, LIMMODE, W0, NT, false, true>(Wr, Wx, ch, P, lds);
        if (edge) {
            // through the R2 rows of line_step_3r (free until the next pass's
            // first barrier; its last reads were before this pass's last one)
            double(*X)[NT] = lds + NC;
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (!(W0 && k == 3)) X[(W0 && k > 3) ? k - 1 : k][t] = Wx[k];
            __syncthreads();
            if (p < 0 || p >= nx) {
                // the mirror column's slot: in [4, 8) on the left, >= 1 on the
                // right (>= 3 for the columns an output needs); clamped for the
                // far columns no output reads
                int src = pc - base;
                src = src < 0 ? 0 : (src > NT - 1 ? NT - 1 : src);
#pragma unroll
                for (int k = 0; k < 5; ++k)
                    if (!(W0 && k == 3)) Wx[k] = X[(W0 && k > 3) ? k - 1 : k][src];
            }
        }
    };

    // Input rows i0 - 2 .. i1 + 1; each is loaded one row ahead of its first
    // x sweep (Wpre), so the load has a whole row's work to arrive.
    double wsmax = 0.0;
    double Ua[5], Ub[5], Q1p[5], G1p[5], Fp[5], Wn[5], Wpre[5];
    load_cell<W0>(L.in, at_row(i0 - 2), Wpre);
    x1_row(Wpre, Wn);
    load_cell<W0>(L.in, at_row(i0 - 1), Wpre);
#pragma unroll 1
    for (int r = i0 - 2; r <= i1 + 1; ++r) {
        double U[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) U[k] = Wn[k];
        if (r < i1 + 1) {
            x1_row(Wpre, Wn);  // row r + 1
            if (r < i1) load_cell<W0>(L.in, at_row(r + 2), Wpre);
        }
        if (r >= i0) {
            double Q0[5], G0[5], Q1[5], G1[5];
            half_step<Phys, 1, LIMMODE, W0>(Ua, Ub, U, cfu, 0.0, 1, P, Q0, G0, Q1, G1);
            if (r >= i0 + 1) {
                double F[5];
                force_flux<Phys, 1, W0>(Q1p, G1p, Q0, G0, cfu, P, F);
                if (r >= i0 + 2) {
                    // y output of row r - 2 -> second x sweep -> HBM
                    double Uo[5], Wz[5];
                    update_cons<Phys, W0>(Ua, Fp, F, cfu, Uo);
                    line_step_3r<Phys, LIMMODE, W0, NT, true, false>(Uo, Wz, ch, P, lds);
                    if (t >= 4 && t < NT - 4 && p < nx) {
                        store_cell<W0>(L.out, (long)(r - 2) * pitch + p, Wz);
                        const double w = Phys::template wavespeed<2>(Wz, P);
                        wsmax = (w > wsmax) ? w : wsmax;
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
    wave_max_to_slot(wsmax, L.step->ws_slots, blockIdx.x * (NT / 64) + (t >> 6));
}

// ---------------------------------------------------------------------------
// Stand-alone wave speed (computeMaximumWaveSpeed{,2D}) and the step driver
// ---------------------------------------------------------------------------
// The wave-speed kernels' tiles: 256 columns by kWsRows rows per block, a
// thread per column (no index division per cell; the rows' loads overlap).
constexpr int kWsRows = 8;

struct WsTile {
    int j, i0, rows;  // column, first row, rows in range (0 when the column is outside the grid)
};

__device__ __forceinline__ WsTile ws_tile(int nx, int ny) {
    const int gx = (nx + 255) / 256;
    const int by = blockIdx.x / gx, bx = blockIdx.x - by * gx;
    WsTile t;
    t.j = bx * 256 + threadIdx.x;
    t.i0 = by * kWsRows;
    t.rows = t.j < nx ? min(kWsRows, ny - t.i0) : 0;
    return t;
}

int ws_blocks(int nx, int ny) { return ((nx + 255) / 256) * ((ny + kWsRows - 1) / kWsRows); }

template <class Phys, int DIM>
__global__ __launch_bounds__(256) void k_wavespeed(ConstPlanes in, int nx, int ny, long pitch, PhysParams P,
                                                   unsigned long long* slots) {
    const WsTile t = ws_tile(nx, ny);
    double m = 0.0;
    if (t.rows == kWsRows) {
#pragma unroll
        for (int r = 0; r < kWsRows; ++r) {
            double W[5];
            load_cell<false>(in, (long)(t.i0 + r) * pitch + t.j, W);
            const double w = Phys::template wavespeed<DIM>(W, P);
            if (w > m) m = w;
        }
    } else {
        for (int r = 0; r < t.rows; ++r) {
            double W[5];
            load_cell<false>(in, (long)(t.i0 + r) * pitch + t.j, W);
            const double w = Phys::template wavespeed<DIM>(W, P);
            if (w > m) m = w;
        }
    }
    block_max_to<256>(m, slots);
}

// The maximum wave speed of a masked AMR level in two forms whose results
// combine into k_wavespeed's.  k_wavespeed_split: every cell, as
// k_wavespeed, into slots, and the maximum over the cells whose mask byte is
// clear into the slots rest (same layout).  While only flagged cells change
// (a masked level's sub-steps), k_wavespeed_segs then reads only the flagged
// cells of the listed 64-cell segments (a wave per segment) and merges the
// rest slots in: the same maximum, as max is exact in any order.
template <class Phys, int DIM>
__global__ __launch_bounds__(256) void k_wavespeed_split(ConstPlanes in, int nx, int ny, long pitch, PhysParams P,
                                                         const unsigned char* mask, unsigned long long* slots,
                                                         unsigned long long* rest) {
    __shared__ double red[2][4];
    const WsTile t = ws_tile(nx, ny);
    double m = 0.0, r = 0.0;
    auto row = [&](int k) {
        const int i = t.i0 + k;
        double W[5];
        load_cell<false>(in, (long)i * pitch + t.j, W);
        const bool flagged = mask[(long)i * nx + t.j] != 0;
        const double w = Phys::template wavespeed<DIM>(W, P);
        if (w > m) m = w;
        if (!flagged && w > r) r = w;
    };
    if (t.rows == kWsRows) {
#pragma unroll
        for (int k = 0; k < kWsRows; ++k) row(k);
    } else {
        for (int k = 0; k < t.rows; ++k) row(k);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64), q = __shfl_xor(r, off, 64);
        m = (o > m) ? o : m;
        r = (q > r) ? q : r;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = m;
        red[1][wave] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0][0], c = red[1][0];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            b = (red[0][w] > b) ? red[0][w] : b;
            c = (red[1][w] > c) ? red[1][w] : c;
        }
        if (b > 0.0)
            atomicMax(slots + (blockIdx.x % kMaxSlots) * kSlotStride, (unsigned long long)__double_as_longlong(b));
        if (c > 0.0)
            atomicMax(rest + (blockIdx.x % kMaxSlots) * kSlotStride, (unsigned long long)__double_as_longlong(c));
    }
}

template <class Phys, int DIM>
__global__ __launch_bounds__(256) void k_wavespeed_segs(ConstPlanes in, int nx, int ny, long pitch, PhysParams P,
                                                        const unsigned char* mask, const int* segs, int n_segs,
                                                        const unsigned long long* rest, unsigned long long* slots) {
    const long wv = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    double m = 0.0;
    if (wv < n_segs) {
        const long idx = (long)segs[wv] * 64 + (threadIdx.x & 63);
        if (idx < (long)nx * ny && mask[idx]) {
            const long i = idx / nx, j = idx - i * nx;
            double W[5];
            load_cell<false>(in, i * pitch + j, W);
            const double w = Phys::template wavespeed<DIM>(W, P);
            if (w > m) m = w;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < kMaxSlots) {
        const double q = __longlong_as_double((long long)rest[threadIdx.x * kSlotStride]);
        if (q > m) m = q;
    }
    wave_max_to_slot(m, slots, (int)wv);
}

// One iteration of the solve loop's head (Solvers/EulerSecondOrderSolver.cpp:162-172 with
// EulerSolvers::computeStableTimeStep{,2D}, Solvers/EulerSolvers.cpp:130-161):
//   stop unless t < finalTime (and the step cap is not reached);
//   dt = CFL * (dx / S_max); dt *= 0.2 while iteration <= 5; clamp to finalTime - t;
//   t += dt, iteration += 1 (the sweeps of this step read dt from here).
// The maximum wave speed comes from the slots (written at cosmos_begin and
// cleared here) and the per-block maxima of the previous step's last sweep.
constexpr int kHeadThreads = 1024;
__global__ __launch_bounds__(kHeadThreads) void k_step_begin(StepState* st) {
    __shared__ double red[kHeadThreads / 64];
    const int tid = threadIdx.x;
    if (st->done) return;
    double m = 0.0;
    if (tid < kMaxSlots) {
        unsigned long long b = st->ws_slots[tid * kSlotStride];
        m = __longlong_as_double((long long)b);
    }
    const double* part = st->ws_part;
    const int n = st->ws_nparts;
    // eight independent loads in flight per thread
    constexpr int kU = 8;
    double acc[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) acc[u] = 0.0;
    int k = tid;
    for (; k + (kU - 1) * kHeadThreads < n; k += kU * kHeadThreads) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            double v = part[k + u * kHeadThreads];
            acc[u] = (v > acc[u]) ? v : acc[u];
        }
    }
    for (; k < n; k += kHeadThreads) {
        double v = part[k];
        acc[0] = (v > acc[0]) ? v : acc[0];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) m = (acc[u] > m) ? acc[u] : m;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = (o > m) ? o : m;
    }
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    const bool stop = !(st->t < st->tf) || st->iter >= st->max_steps;
    if (stop) {
        if (tid == 0) st->done = 1;
        return;
    }
    if (tid < kMaxSlots) st->ws_slots[tid * kSlotStride] = 0ull;
    if (tid != 0) return;
    double smax = red[0];
    for (int w = 1; w < kHeadThreads / 64; ++w) smax = (red[w] > smax) ? red[w] : smax;
    st->ws_nparts = st->ws_nparts_sweep;
    double dt = st->cfl * (st->dx / smax);
    if (st->iter <= 5) dt *= 0.2;
    if ((st->t + dt) > st->tf) dt = st->tf - st->t;
    st->dt = dt;
    // a NaN step ends the loop after it (t becomes NaN), as in the reference
    if (dt != dt && st->status) st->status[5] = 1ull;
    double dth = 0.5 * dt;
    st->half.dt_dx = dth / st->dx;
    st->half.dx_dt = st->dx / dth;
    st->half.hdt_dx = 0.5 * (dth / st->dx);
    st->full.dt_dx = dt / st->dx;
    st->full.dx_dt = st->dx / dt;
    st->full.hdt_dx = 0.5 * (dt / st->dx);
    if (st->gr) {
        // the forcing of a black-hole step: dt / 2 over its subcycles, and the
        // stage fractions of evolveConservedVariableVector{,2D} times that over dx,
        // each formed as the host forms them for a single forcing call
        const double fdt = 0.5 * (dt / st->gr_subcycles);
        const double f = fdt / st->dx;
        st->gr_dt = fdt;
        st->gr_frac[0] = (st->gr_dim == 1 ? 0.5 : 0.25) * f;
        st->gr_frac[1] = (st->gr_dim == 1 ? 0.5 : 0.25) * f;
        st->gr_frac[2] = (st->gr_dim == 1 ? 1.0 : 0.5) * f;
        st->gr_frac_y[0] = 0.5 * f;
        st->gr_frac_y[1] = 0.5 * f;
        st->gr_frac_y[2] = 1.0 * f;
    }
    st->t = st->t + dt;
    st->iter = st->iter + 1;
    st->active_steps = st->active_steps + 1;
}

// Cells of an nx-by-ny state holding a NaN in any of the five planes, added to *count.
__global__ __launch_bounds__(256) void k_nan_cells(ConstPlanes in, int nx, int ny, long pitch,
                                                   unsigned long long* count) {
    long total = (long)nx * ny;
    unsigned n = 0;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        long i = idx / nx, j = idx - i * nx;
        bool bad = false;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            double v = in.p[k][i * pitch + j];
            bad = bad || (v != v);
        }
        n += bad ? 1u : 0u;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(count, (unsigned long long)n);
}

__global__ __launch_bounds__(64) void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = v;
}

__global__ __launch_bounds__(256) void k_zero_slots(unsigned long long* slots) {
    for (int k = threadIdx.x; k < kSlotWords; k += blockDim.x) slots[k] = 0ull;
}

// A wave per 1024 consecutive cells: each lane tests 16 of their mask bytes
// with one 16-byte read (mask is a device allocation, 16-byte aligned) and a
// wave with no flagged cell leaves at once; otherwise the wave copies its
// cells 64 consecutive ones at a time, each lane the five values of its cell
// where flagged (coalesced rows).
__global__ __launch_bounds__(256) void k_mask_writeback(Planes dst, ConstPlanes src, const unsigned char* mask,
                                                       int nx, int ny, long pitch) {
    const long n = (long)nx * ny;
    const int lane = threadIdx.x & 63;
    const long wbase = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 1024;
    if (wbase >= n) return;
    const long b = wbase + 16L * lane;
    bool any = false;
    if (b + 16 <= n) {
        const uint4 w = *reinterpret_cast<const uint4*>(mask + b);
        any = (w.x | w.y | w.z | w.w) != 0u;
    } else {
        for (long e = b; e < n; ++e) any = any || mask[e] != 0;
    }
    if (!__any(any)) return;
#pragma unroll 1
    for (int e = 0; e < 16; ++e) {
        const long idx = wbase + 64L * e + lane;
        if (idx >= n || !mask[idx]) continue;
        const long i = idx / nx, j = idx - i * nx;
        const long c = i * pitch + j;
#pragma unroll
        for (int k = 0; k < 5; ++k) dst.p[k][c] = src.p[k][c];
    }
}

// The candidate blocks of a masked level's launches that hold a flagged cell,
// in two passes.  k_mask_mark: a thread per 16 mask bytes (one 16-byte read;
// the mask is a device allocation, 16-byte aligned) marks, for each flagged
// cell, its x-sweep row segment, its y-sweep tile and its 64-cell write-back
// segment in the byte maps hx, hy, hw (benign same-value races), and clears
// the three counts.  Flags are sparse, so most threads read zero and leave.
__global__ __launch_bounds__(256) void k_mask_mark(const unsigned char* mask, long n, int nx, int tw, int tiles_x,
                                                  int tc, int trv, int tcn, unsigned char* hx, unsigned char* hy,
                                                  unsigned char* hw, int* counts) {
    const long e0 = ((long)blockIdx.x * 256 + threadIdx.x) * 16;
    if (blockIdx.x == 0 && threadIdx.x < 3) counts[threadIdx.x] = 0;  // k_mask_compact runs after this kernel
    if (e0 >= n) return;
    unsigned int q[4] = {0u, 0u, 0u, 0u};
    if (e0 + 16 <= n) {
        const uint4 v = *reinterpret_cast<const uint4*>(mask + e0);
        q[0] = v.x;
        q[1] = v.y;
        q[2] = v.z;
        q[3] = v.w;
    } else {
        for (long e = e0; e < n; ++e) q[(e - e0) >> 2] |= (unsigned int)mask[e] << (8 * ((e - e0) & 3));
    }
    if ((q[0] | q[1] | q[2] | q[3]) == 0u) return;
    long row = e0 / nx;
    int col = (int)(e0 - row * nx);
    for (int k = 0; k < 16; ++k, ++col) {
        while (col >= nx) {
            col -= nx;
            ++row;
        }
        if (((q[k >> 2] >> (8 * (k & 3))) & 0xffu) == 0u) continue;
        hx[row * tiles_x + col / tw] = 1;
        if (hy) hy[(row / trv) * tcn + col / tc] = 1;
        hw[(e0 + k) >> 6] = 1;
    }
}

// k_mask_compact: eight consecutive candidates per thread, 2048 per block
// (the x segments' blocks first, then the y tiles', then the write-back
// segments'; a block never mixes kinds).  Each marked candidate is appended
// to its kind's list and its mark cleared (the maps are all zero again for
// the next build); a block takes its slots with one atomic and places its
// candidates by a scan over its threads.
constexpr int kCompactPer = 8;
__global__ __launch_bounds__(256) void k_mask_compact(unsigned char* hx, unsigned char* hy, unsigned char* hw, long nxb,
                                                     long nyb, long nwb, int bx, int by, int* xl, int* yl, int* wl,
                                                     int* counts) {
    __shared__ int wave_sum[4];
    __shared__ int block_base;
    const int blk = blockIdx.x;
    const int kind = blk < bx ? 0 : (blk < bx + by ? 1 : 2);
    const long b0 = ((long)(blk - (kind == 0 ? 0 : (kind == 1 ? bx : bx + by))) * 256 + threadIdx.x) * kCompactPer;
    const long nb = kind == 0 ? nxb : (kind == 1 ? nyb : nwb);
    unsigned char* hit = kind == 0 ? hx : (kind == 1 ? hy : hw);
    int* list = kind == 0 ? xl : (kind == 1 ? yl : wl);
    unsigned int marks = 0u;
#pragma unroll
    for (int u = 0; u < kCompactPer; ++u) {
        const long b = b0 + u;
        if (b < nb && hit[b] != 0) {
            hit[b] = 0;
            marks |= 1u << u;
        }
    }
    const int mine = __popc(marks);
    // inclusive scan of the counts over the wave, then over the block's waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        before += w < wave ? wave_sum[w] : 0;
        total += wave_sum[w];
    }
    if (total == 0) return;  // uniform over the block
    if (threadIdx.x == 0) block_base = atomicAdd(counts + kind, total);
    __syncthreads();
    int pos = block_base + before + incl - mine;
    for (int u = 0; u < kCompactPer; ++u)
        if (marks & (1u << u)) list[pos++] = (int)(b0 + u);
}

// A wave per listed 64-cell segment of the grid (segs[k] covers cells
// [64 segs[k], 64 segs[k] + 64) in row-major order): each lane copies the five
// values of its cell where flagged.
__global__ __launch_bounds__(256) void k_mask_writeback_segs(Planes dst, ConstPlanes src, const unsigned char* mask,
                                                            int nx, int ny, long pitch, const int* segs, int n_segs) {
    const long wv = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= n_segs) return;
    const long idx = (long)segs[wv] * 64 + (threadIdx.x & 63);
    if (idx >= (long)nx * ny || !mask[idx]) return;
    const long i = idx / nx, j = idx - i * nx;
    const long c = i * pitch + j;
#pragma unroll
    for (int k = 0; k < 5; ++k) dst.p[k][c] = src.p[k][c];
}

__global__ __launch_bounds__(256) void k_any_nonzero(const double* plane, int nx, int ny, long pitch, int* flag) {
    long total = (long)nx * ny;
    int found = 0;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        long i = idx / nx, j = idx - i * nx;
        if (plane[i * pitch + j] != 0.0) found = 1;
    }
    if (__any(found) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

template <class Phys, int ORDER, int LIMMODE, bool W0>
hipError_t dispatch_x(const SweepLaunch& L, hipStream_t s) {
    constexpr int G = ORDER == 2 ? 2 : 1;
    const int TW = kXThreads - 2 * G;
    const int tiles = (L.nx + TW - 1) / TW;
    const bool listed = L.mask && L.block_order;
    if (listed && L.order_n <= 0) return hipSuccess;
    dim3 grid((unsigned)(listed ? L.order_n : tiles * L.ny)), block(kXThreads);
    if (L.fuse_ws) {
        if (L.dim == 1) hipLaunchKernelGGL((k_sweep_x<Phys, ORDER, LIMMODE, 1, true, W0>), grid, block, 0, s, L);
        else hipLaunchKernelGGL((k_sweep_x<Phys, ORDER, LIMMODE, 2, true, W0>), grid, block, 0, s, L);
    } else {
        hipLaunchKernelGGL((k_sweep_x<Phys, ORDER, LIMMODE, 2, false, W0>), grid, block, 0, s, L);
    }
    return hipGetLastError();
}

// Rows per y-sweep thread: long marches amortise the 4-row halo for the
// HBM-bound Newtonian sweeps; the compute-bound special-relativistic sweeps
// need more waves in flight, so shorter marches (more blocks).
int y_rows_for(const SweepLaunch& L) {
    if (L.y_rows == 16 || L.y_rows == 32 || L.y_rows == 64) return L.y_rows;
    return L.physics == 1 ? 16 : 64;
}

// Columns per y-sweep LDS tile (0: the column march): the march for the
// Newtonian sweeps (no halo recomputation; measured 3-4 % faster than tiles
// at 4096^2) on grids it fills, 16-column tiles below that, 8-column tiles (28 output rows of 32) for the
// special-relativistic ones, which are FP64-bound: the march's register
// window holds them to 2 waves per SIMD, the tiles run at the x sweep's
// occupancy (2048^2 SR step 2.98 ms with the march, 2.22 ms with 8-column
// and 2.35 ms with 16-column tiles).
int masked_tile_cols(int y_tile) { return y_tile == 8 ? 8 : 16; }  // masked sweeps: tiles only

int y_tile_for(const SweepLaunch& L) {
    if (L.mask) return masked_tile_cols(L.y_tile);
    if (L.y_tile == 0 || L.y_tile == 8 || L.y_tile == 16) return L.y_tile;
    if (L.physics == 1) return 8;
    // Newtonian: the march, unless it would run fewer than four blocks per CU
    // (1024 blocks; 2048^2 and smaller grids), where 16-column tiles fill the chip
    const int rows = L.fuse_ws ? 64 : y_rows_for(L);
    const long march = (long)((L.nx + kYThreads - 1) / kYThreads) * ((L.ny + rows - 1) / rows);
    return march < 1024 ? 16 : 0;
}

template <class Phys, int ORDER, int LIMMODE, bool W0, int TC>
hipError_t dispatch_y_tile(const SweepLaunch& L, hipStream_t s) {
    constexpr int G = ORDER == 2 ? 2 : 1;
    constexpr int TRV = kXThreads / TC - 2 * G;
    const int tiles = (L.nx + TC - 1) / TC * ((L.ny + TRV - 1) / TRV);
    const bool listed = L.mask && L.block_order;
    if (listed && L.order_n <= 0) return hipSuccess;
    dim3 grid((unsigned)(listed ? L.order_n : tiles)), block(kXThreads);
    if (L.fuse_ws) hipLaunchKernelGGL((k_sweep_y_tile<Phys, ORDER, LIMMODE, 2, true, W0, TC>), grid, block, 0, s, L);
    else hipLaunchKernelGGL((k_sweep_y_tile<Phys, ORDER, LIMMODE, 2, false, W0, TC>), grid, block, 0, s, L);
    return hipGetLastError();
}

template <class Phys, int ORDER, int LIMMODE, bool W0>
hipError_t dispatch_y(const SweepLaunch& L, hipStream_t s) {
    const int tc = y_tile_for(L);
    if (tc == 16) return dispatch_y_tile<Phys, ORDER, LIMMODE, W0, 16>(L, s);
    if (tc == 8) return dispatch_y_tile<Phys, ORDER, LIMMODE, W0, 8>(L, s);
    const int rows = L.fuse_ws ? 64 : y_rows_for(L);
    const int tx = (L.nx + kYThreads - 1) / kYThreads;
    const int ty = (L.ny + rows - 1) / rows;
    dim3 grid((unsigned)(tx * ty)), block(kYThreads);
    if (L.fuse_ws) hipLaunchKernelGGL((k_sweep_y<Phys, ORDER, LIMMODE, true, W0, 64>), grid, block, 0, s, L);
    else if (rows == 16) hipLaunchKernelGGL((k_sweep_y<Phys, ORDER, LIMMODE, false, W0, 16>), grid, block, 0, s, L);
    else if (rows == 32) hipLaunchKernelGGL((k_sweep_y<Phys, ORDER, LIMMODE, false, W0, 32>), grid, block, 0, s, L);
    else hipLaunchKernelGGL((k_sweep_y<Phys, ORDER, LIMMODE, false, W0, 64>), grid, block, 0, s, L);
    return hipGetLastError();
}

template <template <bool> class PhysT, int AXIS>
hipError_t dispatch_axis(const SweepLaunch& L, hipStream_t s) {
    const bool vl0 = (L.limiter == 1) && (L.bias == 0.0);
#define COSMOS_DISPATCH(ORDER, LM, W0)                                                          \
    return AXIS == 0 ? dispatch_x<PhysT<W0>, ORDER, LM, W0>(L, s) : dispatch_y<PhysT<W0>, ORDER, LM, W0>(L, s)
    if (L.fast && L.physics == 0 && L.order == 2 && vl0) {
#define COSMOS_DISPATCH_FAST(W0)                                                                \
    return AXIS == 0 ? dispatch_x<NewtonianFastT<W0>, 2, 2, W0>(L, s) : dispatch_y<NewtonianFastT<W0>, 2, 2, W0>(L, s)
        if (L.w_zero) COSMOS_DISPATCH_FAST(true);
        COSMOS_DISPATCH_FAST(false);
#undef COSMOS_DISPATCH_FAST
    }
    // Exact mode always carries the z-velocity: skipping it is exact only while
    // the state stays finite (the reference's rho * (u * w) is NaN once u is).
    if (L.order == 1) COSMOS_DISPATCH(1, 0, false);
    if (vl0) COSMOS_DISPATCH(2, 1, false);
    COSMOS_DISPATCH(2, 0, false);
#undef COSMOS_DISPATCH
}

}  // namespace

hipError_t launch_sweep_x(const SweepLaunch& L, hipStream_t s) {
    return L.physics == 0 ? dispatch_axis<NewtonianT, 0>(L, s) : dispatch_axis<RelativisticT, 0>(L, s);
}

hipError_t launch_sweep_y(const SweepLaunch& L, hipStream_t s) {
    return L.physics == 0 ? dispatch_axis<NewtonianT, 1>(L, s) : dispatch_axis<RelativisticT, 1>(L, s);
}

hipError_t launch_mask_writeback(const Planes& dst, const ConstPlanes& src, const unsigned char* mask, int nx, int ny,
                                 long pitch, const int* segs, int n_segs, hipStream_t s) {
    if (segs) {
        if (n_segs <= 0) return hipSuccess;
        hipLaunchKernelGGL(k_mask_writeback_segs, dim3((unsigned)((n_segs + 3) / 4)), dim3(256), 0, s, dst, src, mask,
                           nx, ny, pitch, segs, n_segs);
        return hipGetLastError();
    }
    const long n = (long)nx * ny;
    const long blocks = (n + 4095) / 4096;  // 4 waves of 1024 cells
    hipLaunchKernelGGL(k_mask_writeback, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, mask, nx, ny, pitch);
    return hipGetLastError();
}

int mask_tile_cols(const SweepLaunch& L) { return masked_tile_cols(L.y_tile); }

void mask_list_sizes(int nx, int ny, int dim, int order, int tc, long sizes[3]) {
    const int g = order == 2 ? 2 : 1;
    const int tw = kXThreads - 2 * g, trv = kXThreads / tc - 2 * g;
    sizes[0] = (long)((nx + tw - 1) / tw) * ny;
    sizes[1] = dim == 2 ? (long)((nx + tc - 1) / tc) * ((ny + trv - 1) / trv) : 0;
    sizes[2] = ((long)nx * ny + 63) / 64;
}

hipError_t launch_mask_lists(const unsigned char* mask, int nx, int ny, int dim, int order, int tc,
                             unsigned char* hit, int* xl, int* yl, int* wl, int* counts, hipStream_t s) {
    const int g = order == 2 ? 2 : 1;
    const int tw = kXThreads - 2 * g, trv = kXThreads / tc - 2 * g;
    long sz[3];
    mask_list_sizes(nx, ny, dim, order, tc, sz);
    unsigned char* hx = hit;
    unsigned char* hy = dim == 2 ? hit + sz[0] : nullptr;
    unsigned char* hw = hit + sz[0] + sz[1];
    const long n = (long)nx * ny;
    hipLaunchKernelGGL(k_mask_mark, dim3((unsigned)((n + 4095) / 4096)), dim3(256), 0, s, mask, n, nx, tw,
                       (nx + tw - 1) / tw, tc, trv, (nx + tc - 1) / tc, hx, hy, hw, counts);
    constexpr long kPerBlock = 256 * kCompactPer;
    const int bx = (int)((sz[0] + kPerBlock - 1) / kPerBlock), by = (int)((sz[1] + kPerBlock - 1) / kPerBlock),
              bw = (int)((sz[2] + kPerBlock - 1) / kPerBlock);
    hipLaunchKernelGGL(k_mask_compact, dim3((unsigned)(bx + by + bw)), dim3(256), 0, s, hx, hy, hw, sz[0], sz[1],
                       sz[2], bx, by, xl, yl, wl, counts);
    return hipGetLastError();
}

hipError_t launch_wavespeed(const ConstPlanes& in, int nx, int ny, long pitch, int physics, int dim,
                            const PhysParams& P, unsigned long long* slots, hipStream_t s) {
    dim3 grid((unsigned)ws_blocks(nx, ny)), block(256);
    if (physics == 0) {
        if (dim == 1) hipLaunchKernelGGL((k_wavespeed<Newtonian, 1>), grid, block, 0, s, in, nx, ny, pitch, P, slots);
        else hipLaunchKernelGGL((k_wavespeed<Newtonian, 2>), grid, block, 0, s, in, nx, ny, pitch, P, slots);
    } else {
        if (dim == 1) hipLaunchKernelGGL((k_wavespeed<Relativistic, 1>), grid, block, 0, s, in, nx, ny, pitch, P, slots);
        else hipLaunchKernelGGL((k_wavespeed<Relativistic, 2>), grid, block, 0, s, in, nx, ny, pitch, P, slots);
    }
    return hipGetLastError();
}

hipError_t launch_wavespeed_split(const ConstPlanes& in, int nx, int ny, long pitch, int physics, int dim,
                                  const PhysParams& P, const unsigned char* mask, unsigned long long* slots,
                                  unsigned long long* rest, hipStream_t s) {
    const dim3 grid((unsigned)ws_blocks(nx, ny)), block(256);
#define COSMOS_WS_SPLIT(PH, D) \
    hipLaunchKernelGGL((k_wavespeed_split<PH, D>), grid, block, 0, s, in, nx, ny, pitch, P, mask, slots, rest)
    if (physics == 0) {
        if (dim == 1) COSMOS_WS_SPLIT(Newtonian, 1);
        else COSMOS_WS_SPLIT(Newtonian, 2);
    } else {
        if (dim == 1) COSMOS_WS_SPLIT(Relativistic, 1);
        else COSMOS_WS_SPLIT(Relativistic, 2);
    }
#undef COSMOS_WS_SPLIT
    return hipGetLastError();
}

hipError_t launch_wavespeed_segs(const ConstPlanes& in, int nx, int ny, long pitch, int physics, int dim,
                                 const PhysParams& P, const unsigned char* mask, const int* segs, int n_segs,
                                 const unsigned long long* rest, unsigned long long* slots, hipStream_t s) {
    const int blocks = n_segs > 0 ? (n_segs + 3) / 4 : 1;  // one block at least: it merges rest
    const dim3 grid((unsigned)blocks), block(256);
#define COSMOS_WS_SEGS(PH, D) \
    hipLaunchKernelGGL((k_wavespeed_segs<PH, D>), grid, block, 0, s, in, nx, ny, pitch, P, mask, segs, n_segs, rest, \
                       slots)
    if (physics == 0) {
        if (dim == 1) COSMOS_WS_SEGS(Newtonian, 1);
        else COSMOS_WS_SEGS(Newtonian, 2);
    } else {
        if (dim == 1) COSMOS_WS_SEGS(Relativistic, 1);
        else COSMOS_WS_SEGS(Relativistic, 2);
    }
#undef COSMOS_WS_SEGS
    return hipGetLastError();
}

// Rows per block of the fused step: a block's time grows as about 3 RB + 4
// row passes (the first x sweep also runs on the 4 halo rows), and the blocks
// run in ceil(blocks / resident) rounds, so pick RB in [16, 64] minimising
// rounds * (3 RB + 4); ties go to the larger RB (fewer halo rows).
int fused_rows(int nx, int ny, int resident) {
    const long tx = (nx + kFuseCols - 1) / kFuseCols;
    int best = 64;
    long best_cost = -1;
    for (int rb = 64; rb >= 16; --rb) {
        const long blocks = tx * ((ny + rb - 1) / rb);
        const long rounds = (blocks + resident - 1) / resident;
        const long cost = rounds * (3L * rb + 4);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = rb;
        }
    }
    return best;
}

template <bool W0>
hipError_t launch_fused_w(const SweepLaunch& L, hipStream_t s) {
    // resident blocks of the kernel on the current device, once per device
    static int resident[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    if (resident[dev] == 0) {
        int cus = 0, per_cu = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_step_fused<NewtonianFastT<W0>, 2, W0>,
                                                             kXThreads, 0);
        if (e != hipSuccess) return e;
        resident[dev] = std::max(1, cus * per_cu);
    }
    const int rb = L.fuse_rows > 0 ? L.fuse_rows : fused_rows(L.nx, L.ny, resident[dev]);
    const int tx = (L.nx + kFuseCols - 1) / kFuseCols;
    const int ty = (L.ny + rb - 1) / rb;
    dim3 grid((unsigned)(tx * ty)), block(kXThreads);
    hipLaunchKernelGGL((k_step_fused<NewtonianFastT<W0>, 2, W0>), grid, block, 0, s, L, rb);
    return hipGetLastError();
}

hipError_t launch_step_fused(const SweepLaunch& L, hipStream_t s) {
    return L.w_zero ? launch_fused_w<true>(L, s) : launch_fused_w<false>(L, s);
}

// The dispatch order of an n-block sweep: the marked blocks first, then the
// others, each group in the XCD-contiguous order of xcd_tile; the marks are
// cleared for the next step.  One block of kOrderThreads threads, a chunk of
// slots at a time (ballot prefix counts within each wave, wave totals in LDS).
constexpr int kOrderThreads = 1024;

// With row > 0 a block counts as marked when it or one of its four neighbours
// in the launch's block grid (row width `row` blocks) was: the unphysical
// states whose solves run to a long cap sit at shocks that move less than a
// cell per step, so they cross into a neighbouring block from one step to the
// next.  row == 0: the block's own mark only.
__global__ __launch_bounds__(kOrderThreads) void k_hot_order(unsigned char* hot, int* order, int n, int row) {
    __shared__ int wave_hot[kOrderThreads / 64], wave_cold[kOrderThreads / 64];
    __shared__ int total_hot;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto marked = [&](int b) {
        if (hot[b]) return true;
        if (row <= 0) return false;
        return (b > 0 && hot[b - 1]) || (b + 1 < n && hot[b + 1]) || (b >= row && hot[b - row]) ||
               (b + row < n && hot[b + row]);
    };
    if (tid == 0) total_hot = 0;
    __syncthreads();
    int mine = 0;
    for (int k = tid; k < n; k += kOrderThreads) mine += marked(k) ? 1 : 0;
    if (mine) atomicAdd(&total_hot, mine);
    __syncthreads();
    const int nhot = total_hot;
    int base_hot = 0, base_cold = 0;
    for (int c = 0; c < n; c += kOrderThreads) {
        const int slot = c + tid;
        const bool valid = slot < n;
        const int b = valid ? xcd_tile(slot, n) : 0;
        const bool h = valid && marked(b);
        const bool cold = valid && !h;
        const unsigned long long mh = __ballot(h), mc = __ballot(cold);
        const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int ph = __popcll(mh & below), pc = __popcll(mc & below);
        if (lane == 0) {
            wave_hot[wave] = __popcll(mh);
            wave_cold[wave] = __popcll(mc);
        }
        __syncthreads();
        int oh = base_hot, oc = base_cold;
        for (int w = 0; w < wave; ++w) {
            oh += wave_hot[w];
            oc += wave_cold[w];
        }
        if (h) order[oh + ph] = b;
        if (cold) order[nhot + oc + pc] = b;
        for (int w = 0; w < kOrderThreads / 64; ++w) {
            base_hot += wave_hot[w];
            base_cold += wave_cold[w];
        }
        __syncthreads();
    }
    for (int k = tid; k < n; k += kOrderThreads) hot[k] = 0;
}

hipError_t launch_hot_order(unsigned char* hot, int* order, int n, int row, hipStream_t s) {
    hipLaunchKernelGGL(k_hot_order, dim3(1), dim3(kOrderThreads), 0, s, hot, order, n, row < 0 ? 0 : row);
    return hipGetLastError();
}

int sweep_row_blocks(const SweepLaunch& L, int axis) {
    if (axis == 0) {
        const int TW = kXThreads - 2 * (L.order == 2 ? 2 : 1);
        return (L.nx + TW - 1) / TW;
    }
    const int tc = y_tile_for(L);
    return tc == 8 || tc == 16 ? (L.nx + tc - 1) / tc : 0;
}

int sweep_x_grid(const SweepLaunch& L) {
    const int TW = kXThreads - 2 * (L.order == 2 ? 2 : 1);
    return (L.nx + TW - 1) / TW * L.ny;
}

int sweep_y_tile_grid(const SweepLaunch& L) {
    const int tc = y_tile_for(L);
    if (tc != 8 && tc != 16) return 0;
    const int trv = kXThreads / tc - 2 * (L.order == 2 ? 2 : 1);
    return (L.nx + tc - 1) / tc * ((L.ny + trv - 1) / trv);
}

hipError_t launch_step_begin(StepState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(kHeadThreads), 0, s, st);
    return hipGetLastError();
}

int sweep_x_blocks(int nx, int ny, int order) {
    const int TW = kXThreads - 2 * (order == 2 ? 2 : 1);
    return (nx + TW - 1) / TW * ny;
}

hipError_t launch_nan_cells(const ConstPlanes& in, int nx, int ny, long pitch, unsigned long long* count,
                            hipStream_t s) {
    long total = (long)nx * ny;
    long blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_nan_cells, dim3((unsigned)blocks), dim3(256), 0, s, in, nx, ny, pitch, count);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_zero_slots(unsigned long long* slots, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_slots, dim3(1), dim3(256), 0, s, slots);
    return hipGetLastError();
}

hipError_t launch_any_nonzero(const double* plane, int nx, int ny, long pitch, int* flag, hipStream_t s) {
    long total = (long)nx * ny;
    long blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_any_nonzero, dim3((unsigned)blocks), dim3(256), 0, s, plane, nx, ny, pitch, flag);
    return hipGetLastError();
}

}  // namespace cfv
