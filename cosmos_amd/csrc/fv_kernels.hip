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
#include "fv_kernels.hpp"

namespace cfv {
namespace {

constexpr int kXThreads = 256;
constexpr int kYThreads = 64;
constexpr int kYRows = 64;

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

__device__ __forceinline__ void load_cell(const ConstPlanes& in, long idx, double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) W[k] = in.p[k][idx];
}

__device__ __forceinline__ void store_cell(const Planes& out, long idx, const double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) out.p[k][idx] = W[k];
}

// Half-step evolution of one cell from its conserved neighbours
// (EulerSolvers::evolveStateByHalf{X,Y}TimeStep, Solvers/EulerSolvers.cpp:191-223),
// then the p2c and flux of both evolved face states as the FORCE flux of
// the neighbouring interfaces will need them (Solvers/EulerFirstOrderSolver.cpp:52-59).
template <class Phys, int AXIS, int LIMMODE>
__device__ __forceinline__ void half_step(const double* ul, const double* um, const double* ur, const SweepCoef& cf,
                                          double bias, int limiter, const PhysParams& P, double* Q0, double* G0,
                                          double* Q1, double* G1) {
    double xl[5], xr[5], fl[5], fr[5], a[5], b[5], p0[5], p1[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double s = slope_component<LIMMODE>(ul[k], um[k], ur[k], bias, limiter);
        xl[k] = um[k] - (0.5 * s);
        xr[k] = um[k] + (0.5 * s);
    }
    Phys::template flux<AXIS>(xl, fl, Phys::inv_density(xl), P);
    Phys::template flux<AXIS>(xr, fr, Phys::inv_density(xr), P);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double ev = cf.hdt_dx * (fl[k] - fr[k]);
        a[k] = xl[k] + ev;
        b[k] = xr[k] + ev;
    }
    // P0 = c2p(a) (side 0), P1 = c2p(b) (side 1); p2c(P) has density P[0] = a[0] (resp.
    // b[0]), so the flux of p2c(P) divides by the same density.
    double ia = Phys::inv_density(a), ib = Phys::inv_density(b);
    Phys::c2p(a, p0, ia, P);
    Phys::c2p(b, p1, ib, P);
    Phys::p2c(p0, Q0, P);
    Phys::p2c(p1, Q1, P);
    Phys::template flux<AXIS>(Q0, G0, ia, P);
    Phys::template flux<AXIS>(Q1, G1, ib, P);
}

// FORCE flux of one interface from p2c and flux of its two states
// (computeLaxFriedrichsFlux, computeRichtmyerFlux, computeFORCEFlux:
// Solvers/EulerFirstOrderSolver.cpp:3-24).
template <class Phys, int AXIS>
__device__ __forceinline__ void force_flux(const double* UL, const double* FL, const double* UR, const double* FR,
                                           const SweepCoef& cf, const PhysParams& P, double* F) {
    double us[5], fs[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) us[k] = 0.5 * ((UL[k] + UR[k]) + (cf.dt_dx * (FL[k] - FR[k])));
    Phys::template flux<AXIS>(us, fs, Phys::inv_density(us), P);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double lf = 0.5 * ((FL[k] + FR[k]) + (cf.dx_dt * (UL[k] - UR[k])));
        F[k] = 0.5 * (lf + fs[k]);
    }
}

// computeFORCEUpdate (Solvers/EulerFirstOrderSolver.cpp:96-101) then c2p.
template <class Phys>
__device__ __forceinline__ void update_cell(const double* U, const double* Fl, const double* Fr, const SweepCoef& cf,
                                            const PhysParams& P, double* W) {
    double un[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) un[k] = U[k] + (cf.dt_dx * (Fl[k] - Fr[k]));
    Phys::c2p(un, W, Phys::inv_density(un), P);
}

// Block-wide max of non-negative wave speeds, merged into *bits with an
// integer atomic max (the IEEE bit patterns of non-negative doubles order
// like the values).  The reference's running max skips NaN (ws > max is
// false), so NaN never enters.
template <int NT>
__device__ __forceinline__ void block_max_to(double m, unsigned long long* bits) {
    __shared__ double red[NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = (o > m) ? o : m;
    }
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) b = (red[w] > b) ? red[w] : b;
        if (b > 0.0) atomicMax(bits, (unsigned long long)__double_as_longlong(b));
    }
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
template <class Phys, int ORDER, int LIMMODE, int DIM, bool WS>
__global__ __launch_bounds__(kXThreads) void k_sweep_x(SweepLaunch L) {
    constexpr int NT = kXThreads;
    constexpr int G = ORDER == 2 ? 2 : 1;
    constexpr int TW = NT - 2 * G;
    __shared__ double sA[10][NT];   // stage A output (U; FORCE: U and F)
    __shared__ double sB[10][NT];   // stage B output (Q0, G0); then the interface fluxes

    SweepCoef cf;
    if (!resolve_coef(L, cf)) return;
    const PhysParams& P = L.P;
    const int nx = L.nx;
    const int tiles = (nx + TW - 1) / TW;
    const int row = blockIdx.x / tiles;
    const int tile = blockIdx.x - row * tiles;
    const int s = threadIdx.x;
    const int c = tile * TW - G + s;
    const int cc = ORDER == 2 ? mirror2(c, nx) : clamp1(c, nx);
    const long rowbase = (long)row * L.pitch;

    double W[5], U[5], F[5];
    load_cell(L.in, rowbase + cc, W);
    Phys::p2c(W, U, P);

    const int sl = s > 0 ? s - 1 : 0;
    const int sr = s < NT - 1 ? s + 1 : NT - 1;
    if constexpr (ORDER == 2) {
#pragma unroll
        for (int k = 0; k < 5; ++k) sA[k][s] = U[k];
        __syncthreads();
        double ul[5], ur[5], Q0[5], G0[5], Q1[5], G1[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            ul[k] = sA[k][sl];
            ur[k] = sA[k][sr];
        }
        half_step<Phys, 0, LIMMODE>(ul, U, ur, cf, L.bias, L.limiter, P, Q0, G0, Q1, G1);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            sB[k][s] = Q0[k];
            sB[5 + k][s] = G0[k];
        }
        __syncthreads();
        double UR[5], FR[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            UR[k] = sB[k][sr];
            FR[k] = sB[5 + k][sr];
        }
        force_flux<Phys, 0>(Q1, G1, UR, FR, cf, P, F);
        // interface (s, s+1) -> sA (stage A values are no longer read)
#pragma unroll
        for (int k = 0; k < 5; ++k) sA[k][s] = F[k];
        __syncthreads();
    } else {
        double FU[5];
        Phys::template flux<0>(U, FU, Phys::inv_density(U), P);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            sA[k][s] = U[k];
            sA[5 + k][s] = FU[k];
        }
        __syncthreads();
        double UR[5], FR[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            UR[k] = sA[k][sr];
            FR[k] = sA[5 + k][sr];
        }
        force_flux<Phys, 0>(U, FU, UR, FR, cf, P, F);
#pragma unroll
        for (int k = 0; k < 5; ++k) sB[k][s] = F[k];
        __syncthreads();
    }
    const bool valid = (s >= G) && (s < NT - G) && (c < nx);
    double Fl[5], Wn[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Fl[k] = ORDER == 2 ? sA[k][sl] : sB[k][sl];
    update_cell<Phys>(U, Fl, F, cf, P, Wn);
    if (valid) store_cell(L.out, rowbase + c, Wn);
    if constexpr (WS) {
        double m = 0.0;
        if (valid) {
            double w = Phys::template wavespeed<DIM>(Wn, P);
            if (w > m) m = w;
        }
        block_max_to<NT>(m, &L.step->ws_bits);
    }
}

// ---------------------------------------------------------------------------
// y sweep: thread per column, marching down kYRows rows
// ---------------------------------------------------------------------------
template <class Phys, int ORDER, int LIMMODE, bool WS>
__global__ __launch_bounds__(kYThreads) void k_sweep_y(SweepLaunch L) {
    SweepCoef cf;
    if (!resolve_coef(L, cf)) return;
    const PhysParams& P = L.P;
    const int nx = L.nx, ny = L.ny;
    const int tx = (nx + kYThreads - 1) / kYThreads;
    const int by = blockIdx.x / tx;
    const int bx = blockIdx.x - by * tx;
    const int j = bx * kYThreads + threadIdx.x;
    const bool colok = j < nx;
    const int jj = colok ? j : nx - 1;
    const int i0 = by * kYRows;
    const int i1 = min(i0 + kYRows, ny);
    const long pitch = L.pitch;
    double wsmax = 0.0;

    if constexpr (ORDER == 2) {
        double Ua[5], Ub[5], Q1p[5], G1p[5], Fp[5], Wn[5];
        load_cell(L.in, (long)mirror2(i0 - 2, ny) * pitch + jj, Wn);
        for (int r = i0 - 2; r <= i1 + 1; ++r) {
            double W[5], U[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) W[k] = Wn[k];
            if (r < i1 + 1) load_cell(L.in, (long)mirror2(r + 1, ny) * pitch + jj, Wn);
            Phys::p2c(W, U, P);
            if (r >= i0) {
                double Q0[5], G0[5], Q1[5], G1[5];
                half_step<Phys, 1, LIMMODE>(Ua, Ub, U, cf, L.bias, L.limiter, P, Q0, G0, Q1, G1);
                if (r >= i0 + 1) {
                    double F[5];
                    force_flux<Phys, 1>(Q1p, G1p, Q0, G0, cf, P, F);
                    if (r >= i0 + 2) {
                        double Wo[5];
                        update_cell<Phys>(Ua, Fp, F, cf, P, Wo);
                        if (colok) store_cell(L.out, (long)(r - 2) * pitch + j, Wo);
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
        load_cell(L.in, (long)clamp1(i0 - 1, ny) * pitch + jj, Wn);
        for (int r = i0 - 1; r <= i1; ++r) {
            double W[5], U[5], FU[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) W[k] = Wn[k];
            if (r < i1) load_cell(L.in, (long)clamp1(r + 1, ny) * pitch + jj, Wn);
            Phys::p2c(W, U, P);
            Phys::template flux<1>(U, FU, Phys::inv_density(U), P);
            if (r >= i0) {
                double F[5];
                force_flux<Phys, 1>(Up, FUp, U, FU, cf, P, F);
                if (r >= i0 + 1) {
                    double Wo[5];
                    update_cell<Phys>(Up, Fp, F, cf, P, Wo);
                    if (colok) store_cell(L.out, (long)(r - 1) * pitch + j, Wo);
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
    if constexpr (WS) block_max_to<kYThreads>(wsmax, &L.step->ws_bits);
}

// ---------------------------------------------------------------------------
// Stand-alone wave speed (computeMaximumWaveSpeed{,2D}) and the step driver
// ---------------------------------------------------------------------------
template <class Phys, int DIM>
__global__ __launch_bounds__(256) void k_wavespeed(ConstPlanes in, int nx, int ny, long pitch, PhysParams P,
                                                   unsigned long long* bits) {
    double m = 0.0;
    long total = (long)nx * ny;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        long i = idx / nx, j = idx - i * nx;
        double W[5];
        load_cell(in, i * pitch + j, W);
        double w = Phys::template wavespeed<DIM>(W, P);
        if (w > m) m = w;
    }
    block_max_to<256>(m, bits);
}

// One iteration of the solve loop's head (Solvers/EulerSecondOrderSolver.cpp:162-172 with
// EulerSolvers::computeStableTimeStep{,2D}, Solvers/EulerSolvers.cpp:130-161):
//   stop unless t < finalTime (and the step cap is not reached);
//   dt = CFL * (dx / S_max); dt *= 0.2 while iteration <= 5; clamp to finalTime - t;
//   t += dt, iteration += 1 (the sweeps of this step read dt from here).
__global__ void k_step_begin(StepState* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (st->done) return;
    if (!(st->t < st->tf) || st->iter >= st->max_steps) {
        st->done = 1;
        return;
    }
    double smax = __longlong_as_double((long long)st->ws_bits);
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
    st->ws_bits = 0ull;
    st->t = st->t + dt;
    st->iter = st->iter + 1;
    st->active_steps = st->active_steps + 1;
}

__global__ void k_zero_u64(unsigned long long* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

template <class Phys, int ORDER, int LIMMODE>
hipError_t dispatch_x(const SweepLaunch& L, hipStream_t s) {
    constexpr int G = ORDER == 2 ? 2 : 1;
    const int TW = kXThreads - 2 * G;
    const int tiles = (L.nx + TW - 1) / TW;
    dim3 grid((unsigned)(tiles * L.ny)), block(kXThreads);
    if (L.fuse_ws) {
        if (L.dim == 1) hipLaunchKernelGGL((k_sweep_x<Phys, ORDER, LIMMODE, 1, true>), grid, block, 0, s, L);
        else hipLaunchKernelGGL((k_sweep_x<Phys, ORDER, LIMMODE, 2, true>), grid, block, 0, s, L);
    } else {
        hipLaunchKernelGGL((k_sweep_x<Phys, ORDER, LIMMODE, 2, false>), grid, block, 0, s, L);
    }
    return hipGetLastError();
}

template <class Phys, int ORDER, int LIMMODE>
hipError_t dispatch_y(const SweepLaunch& L, hipStream_t s) {
    const int tx = (L.nx + kYThreads - 1) / kYThreads;
    const int ty = (L.ny + kYRows - 1) / kYRows;
    dim3 grid((unsigned)(tx * ty)), block(kYThreads);
    if (L.fuse_ws) hipLaunchKernelGGL((k_sweep_y<Phys, ORDER, LIMMODE, true>), grid, block, 0, s, L);
    else hipLaunchKernelGGL((k_sweep_y<Phys, ORDER, LIMMODE, false>), grid, block, 0, s, L);
    return hipGetLastError();
}

template <class Phys>
hipError_t dispatch_phys_x(const SweepLaunch& L, hipStream_t s) {
    const bool vl0 = (L.limiter == 1) && (L.bias == 0.0);
    if (L.order == 1) return dispatch_x<Phys, 1, 0>(L, s);
    return vl0 ? dispatch_x<Phys, 2, 1>(L, s) : dispatch_x<Phys, 2, 0>(L, s);
}

template <class Phys>
hipError_t dispatch_phys_y(const SweepLaunch& L, hipStream_t s) {
    const bool vl0 = (L.limiter == 1) && (L.bias == 0.0);
    if (L.order == 1) return dispatch_y<Phys, 1, 0>(L, s);
    return vl0 ? dispatch_y<Phys, 2, 1>(L, s) : dispatch_y<Phys, 2, 0>(L, s);
}

}  // namespace

hipError_t launch_sweep_x(const SweepLaunch& L, hipStream_t s) {
    return L.physics == 0 ? dispatch_phys_x<Newtonian>(L, s) : dispatch_phys_x<Relativistic>(L, s);
}

hipError_t launch_sweep_y(const SweepLaunch& L, hipStream_t s) {
    return L.physics == 0 ? dispatch_phys_y<Newtonian>(L, s) : dispatch_phys_y<Relativistic>(L, s);
}

hipError_t launch_wavespeed(const ConstPlanes& in, int nx, int ny, long pitch, int physics, int dim,
                            const PhysParams& P, unsigned long long* out_bits, hipStream_t s) {
    long total = (long)nx * ny;
    long blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    dim3 grid((unsigned)blocks), block(256);
    if (physics == 0) {
        if (dim == 1) hipLaunchKernelGGL((k_wavespeed<Newtonian, 1>), grid, block, 0, s, in, nx, ny, pitch, P, out_bits);
        else hipLaunchKernelGGL((k_wavespeed<Newtonian, 2>), grid, block, 0, s, in, nx, ny, pitch, P, out_bits);
    } else {
        if (dim == 1) hipLaunchKernelGGL((k_wavespeed<Relativistic, 1>), grid, block, 0, s, in, nx, ny, pitch, P, out_bits);
        else hipLaunchKernelGGL((k_wavespeed<Relativistic, 2>), grid, block, 0, s, in, nx, ny, pitch, P, out_bits);
    }
    return hipGetLastError();
}

hipError_t launch_step_begin(StepState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, s, st);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}

}  // namespace cfv
