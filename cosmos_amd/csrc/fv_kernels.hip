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
#include "fv_kernels.hpp"

namespace cfv {
namespace {

constexpr int kXThreads = 256;
constexpr int kYThreads = 64;

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

// Half-step evolution of one cell from its conserved neighbours
// (EulerSolvers::evolveStateByHalf{X,Y}TimeStep, Solvers/EulerSolvers.cpp:191-223),
// then the p2c and flux of both evolved face states as the FORCE flux of
// the neighbouring interfaces will need them (Solvers/EulerFirstOrderSolver.cpp:52-59).
template <class Phys, int AXIS, int LIMMODE, bool W0>
__device__ __forceinline__ void half_step(const double* ul, const double* um, const double* ur, const SweepCoef& cf,
                                          double bias, int limiter, const PhysParams& P, double* Q0, double* G0,
                                          double* Q1, double* G1) {
    double xl[5], xr[5], fl[5], fr[5], a[5], b[5], p0[5], p1[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (W0 && k == 3) {
            xl[k] = xr[k] = 0.0;
            continue;
        }
        double s = LIMMODE == 2 ? slope_component_fast_vl(ul[k], um[k], ur[k])
                                : slope_component<LIMMODE>(ul[k], um[k], ur[k], bias, limiter);
        xl[k] = um[k] - (0.5 * s);
        xr[k] = um[k] + (0.5 * s);
    }
    Phys::template flux<AXIS>(xl, fl, Phys::inv_density(xl), P);
    Phys::template flux<AXIS>(xr, fr, Phys::inv_density(xr), P);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double ev = (W0 && k == 3) ? 0.0 : cf.hdt_dx * (fl[k] - fr[k]);
        a[k] = (W0 && k == 3) ? 0.0 : xl[k] + ev;
        b[k] = (W0 && k == 3) ? 0.0 : xr[k] + ev;
    }
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
        Phys::c2p(a, p0, ia, P);
        Phys::c2p(b, p1, ib, P);
        Phys::p2c(p0, Q0, P);
        Phys::p2c(p1, Q1, P);
        Phys::template flux<AXIS>(Q0, G0, ia, P);
        Phys::template flux<AXIS>(Q1, G1, ib, P);
    }
}

// FORCE flux of one interface from p2c and flux of its two states
// (computeLaxFriedrichsFlux, computeRichtmyerFlux, computeFORCEFlux:
// Solvers/EulerFirstOrderSolver.cpp:3-24).
template <class Phys, int AXIS, bool W0>
__device__ __forceinline__ void force_flux(const double* UL, const double* FL, const double* UR, const double* FR,
                                           const SweepCoef& cf, const PhysParams& P, double* F) {
    double us[5], fs[5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        us[k] = (W0 && k == 3) ? 0.0 : 0.5 * ((UL[k] + UR[k]) + (cf.dt_dx * (FL[k] - FR[k])));
    Phys::template flux<AXIS>(us, fs, Phys::inv_density(us), P);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (W0 && k == 3) {
            F[k] = 0.0;
            continue;
        }
        double lf = 0.5 * ((FL[k] + FR[k]) + (cf.dx_dt * (UL[k] - UR[k])));
        F[k] = 0.5 * (lf + fs[k]);
    }
}

// computeFORCEUpdate (Solvers/EulerFirstOrderSolver.cpp:96-101) then c2p.
template <class Phys, bool W0>
__device__ __forceinline__ void update_cell(const double* U, const double* Fl, const double* Fr, const SweepCoef& cf,
                                            const PhysParams& P, double* W) {
    double un[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) un[k] = (W0 && k == 3) ? 0.0 : U[k] + (cf.dt_dx * (Fl[k] - Fr[k]));
    Phys::c2p(un, W, Phys::inv_density(un), P);
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
    //   SLIC:  sA = rows 0-4   stage A U, then the interface fluxes
    //          sB = rows 5-14  stage B Q0, G0
    //   FORCE: sA = rows 0-9   U and F(U);  sB = rows 10-14 the interface fluxes
    __shared__ double lds[15][NT];
    double(*sA)[NT] = lds;
    double(*sB)[NT] = lds + (ORDER == 2 ? 5 : 10);

    SweepCoef cf;
    if (!resolve_coef(L, cf)) return;
    const PhysParams& P = L.P;
    const int nx = L.nx;
    const int tiles = (nx + TW - 1) / TW;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int row = blk / tiles;
    const int tile = blk - row * tiles;
    const int s = threadIdx.x;
    const int c = tile * TW - G + s;
    const int cc = ORDER == 2 ? mirror2(c, nx) : clamp1(c, nx);
    const long rowbase = (long)row * L.pitch;

    double W[5], U[5], F[5];
    load_cell<W0>(L.in, rowbase + cc, W);
    Phys::p2c(W, U, P);

    const int sl = s > 0 ? s - 1 : 0;
    const int sr = s < NT - 1 ? s + 1 : NT - 1;
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
        half_step<Phys, 0, LIMMODE, W0>(ul, U, ur, cf, L.bias, L.limiter, P, Q0, G0, Q1, G1);
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
        force_flux<Phys, 0, W0>(Q1, G1, UR, FR, cf, P, F);
        // interface (s, s+1) -> sA (stage A values are no longer read)
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (!(W0 && k == 3)) sA[k][s] = F[k];
        __syncthreads();
    } else {
        double FU[5];
        Phys::template flux<0>(U, FU, Phys::inv_density(U), P);
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
        force_flux<Phys, 0, W0>(U, FU, UR, FR, cf, P, F);
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (!(W0 && k == 3)) sB[k][s] = F[k];
        __syncthreads();
    }
    const bool valid = (s >= G) && (s < NT - G) && (c < nx);
    double Fl[5], Wn[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Fl[k] = (W0 && k == 3) ? 0.0 : (ORDER == 2 ? sA[k][sl] : sB[k][sl]);
    update_cell<Phys, W0>(U, Fl, F, cf, P, Wn);
    if (valid) store_cell<W0>(L.out, rowbase + c, Wn);
    if constexpr (WS) {
        double m = 0.0;
        if (valid) {
            double w = Phys::template wavespeed<DIM>(Wn, P);
            if (w > m) m = w;
        }
        if (L.step->ws_slot_atomics) {
            wave_max_to_slot(m, L.step->ws_slots, blockIdx.x * (NT / 64) + (threadIdx.x >> 6));
        } else {
            // rows no longer read in stage D: sB for SLIC, sA for FORCE
            block_max_store<NT>(m, L.step->ws_part, ORDER == 2 ? &sB[0][0] : &sA[0][0]);
        }
    }
}

// ---------------------------------------------------------------------------
// y sweep: thread per column, marching down ROWS rows
// ---------------------------------------------------------------------------
// Minimum waves per SIMD, which caps the VGPRs: 4 (128 VGPRs) for the
// HBM-bound fast Newtonian sweep, which fits without spilling; the exact
// Newtonian sweep would spill 44-68 B per lane at that cap and the
// special-relativistic one 172-200 B (its con2prim state), so they may use
// up to 168 (3 waves) and 256 (2 waves) VGPRs.
template <class Phys>
constexpr int y_min_waves() { return Phys::kPhysics == 1 ? 2 : (Phys::kFast ? 4 : 3); }

template <class Phys, int ORDER, int LIMMODE, bool WS, bool W0, int ROWS>
__global__ __launch_bounds__(kYThreads) __attribute__((amdgpu_waves_per_eu(y_min_waves<Phys>(), 8))) void k_sweep_y(
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
    const long pitch = L.pitch;
    double wsmax = 0.0;

    if constexpr (ORDER == 2) {
        double Ua[5], Ub[5], Q1p[5], G1p[5], Fp[5], Wn[5];
        load_cell<W0>(L.in, (long)mirror2(i0 - 2, ny) * pitch + jj, Wn);
        for (int r = i0 - 2; r <= i1 + 1; ++r) {
            double W[5], U[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) W[k] = Wn[k];
            if (r < i1 + 1) load_cell<W0>(L.in, (long)mirror2(r + 1, ny) * pitch + jj, Wn);
            Phys::p2c(W, U, P);
            if (r >= i0) {
                double Q0[5], G0[5], Q1[5], G1[5];
                half_step<Phys, 1, LIMMODE, W0>(Ua, Ub, U, cf, L.bias, L.limiter, P, Q0, G0, Q1, G1);
                if (r >= i0 + 1) {
                    double F[5];
                    force_flux<Phys, 1, W0>(Q1p, G1p, Q0, G0, cf, P, F);
                    if (r >= i0 + 2) {
                        double Wo[5];
                        update_cell<Phys, W0>(Ua, Fp, F, cf, P, Wo);
                        if (colok) store_cell<W0>(L.out, (long)(r - 2) * pitch + j, Wo);
                        if constexpr (WS) {
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
        load_cell<W0>(L.in, (long)clamp1(i0 - 1, ny) * pitch + jj, Wn);
        for (int r = i0 - 1; r <= i1; ++r) {
            double W[5], U[5], FU[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) W[k] = Wn[k];
            if (r < i1) load_cell<W0>(L.in, (long)clamp1(r + 1, ny) * pitch + jj, Wn);
            Phys::p2c(W, U, P);
            Phys::template flux<1>(U, FU, Phys::inv_density(U), P);
            if (r >= i0) {
                double F[5];
                force_flux<Phys, 1, W0>(Up, FUp, U, FU, cf, P, F);
                if (r >= i0 + 1) {
                    double Wo[5];
                    update_cell<Phys, W0>(Up, Fp, F, cf, P, Wo);
                    if (colok) store_cell<W0>(L.out, (long)(r - 1) * pitch + j, Wo);
                    if constexpr (WS) {
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
    if constexpr (WS) {
        if (L.step->ws_slot_atomics) wave_max_to_slot(wsmax, L.step->ws_slots, blockIdx.x);
        else block_max_store<kYThreads>(wsmax, L.step->ws_part, nullptr);
    }
}

// ---------------------------------------------------------------------------
// Stand-alone wave speed (computeMaximumWaveSpeed{,2D}) and the step driver
// ---------------------------------------------------------------------------
template <class Phys, int DIM>
__global__ __launch_bounds__(256) void k_wavespeed(ConstPlanes in, int nx, int ny, long pitch, PhysParams P,
                                                   unsigned long long* slots) {
    double m = 0.0;
    long total = (long)nx * ny;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        long i = idx / nx, j = idx - i * nx;
        double W[5];
        load_cell<false>(in, i * pitch + j, W);
        double w = Phys::template wavespeed<DIM>(W, P);
        if (w > m) m = w;
    }
    block_max_to<256>(m, slots);
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
    double dth = 0.5 * dt;
    st->half.dt_dx = dth / st->dx;
    st->half.dx_dt = st->dx / dth;
    st->half.hdt_dx = 0.5 * (dth / st->dx);
    st->full.dt_dx = dt / st->dx;
    st->full.dx_dt = st->dx / dt;
    st->full.hdt_dx = 0.5 * (dt / st->dx);
    st->t = st->t + dt;
    st->iter = st->iter + 1;
    st->active_steps = st->active_steps + 1;
}

__global__ __launch_bounds__(256) void k_zero_slots(unsigned long long* slots) {
    for (int k = threadIdx.x; k < kSlotWords; k += blockDim.x) slots[k] = 0ull;
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
    dim3 grid((unsigned)(tiles * L.ny)), block(kXThreads);
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

template <class Phys, int ORDER, int LIMMODE, bool W0>
hipError_t dispatch_y(const SweepLaunch& L, hipStream_t s) {
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

hipError_t launch_wavespeed(const ConstPlanes& in, int nx, int ny, long pitch, int physics, int dim,
                            const PhysParams& P, unsigned long long* slots, hipStream_t s) {
    long total = (long)nx * ny;
    long blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    dim3 grid((unsigned)blocks), block(256);
    if (physics == 0) {
        if (dim == 1) hipLaunchKernelGGL((k_wavespeed<Newtonian, 1>), grid, block, 0, s, in, nx, ny, pitch, P, slots);
        else hipLaunchKernelGGL((k_wavespeed<Newtonian, 2>), grid, block, 0, s, in, nx, ny, pitch, P, slots);
    } else {
        if (dim == 1) hipLaunchKernelGGL((k_wavespeed<Relativistic, 1>), grid, block, 0, s, in, nx, ny, pitch, P, slots);
        else hipLaunchKernelGGL((k_wavespeed<Relativistic, 2>), grid, block, 0, s, in, nx, ny, pitch, P, slots);
    }
    return hipGetLastError();
}

hipError_t launch_step_begin(StepState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(kHeadThreads), 0, s, st);
    return hipGetLastError();
}

int sweep_x_blocks(int nx, int ny, int order) {
    const int TW = kXThreads - 2 * (order == 2 ? 2 : 1);
    return (nx + TW - 1) / TW * ny;
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
