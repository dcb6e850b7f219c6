// Batched persistent form of the FORCE / SLIC update (no counterpart in the
// reference): many independent problems (Newtonian and special relativistic,
// 1D and 2D, of uneven sizes) advanced to their step cap (or final time) by
// one persistent kernel per physics, captured with three small set-up
// kernels in one hipGraph.
//
// Each problem runs the solve / solve2D loop of the reference
// (Solvers/EulerSecondOrderSolver.cpp:133-178, Solvers/EulerFirstOrderSolver.cpp:165-216
// and the relativistic siblings): a step head (dt from the CFL condition with
// the 0.2x damping of iterations 0..5 and the final-time clamp), then one x
// sweep (1D) or X(dt/2) Y(dt) X(dt/2) (2D).  Every sweep is a "phase" that
// reads the buffer the previous phase wrote and writes the other one; the
// per-cell arithmetic is the per-sweep kernels' (fv_sweep_core.hpp), so each
// problem's result is bitwise the one cosmos_solve gives.
//
// Work decomposition (per physics):
//   whole task   one block runs every phase of a small problem in order, with
//                block barriers between phases and a block reduction for dt;
//   tiled task   one phase of a larger problem is split into tasks: runs of x
//                segments (a segment is 256 threads over one row piece, or
//                several short rows packed side by side), or 256 y marches
//                (a march is one column over ry rows).
// Tasks are claimed from a global ticket counter.  Tickets are issued whole
// tasks first (largest first), then step by step: every tiled problem's
// first-sweep tasks, then second, then third.  A task's inputs come only from
// tasks with smaller tickets, and a block waits only after claiming, so every
// task it waits for is held by a running block: the queue cannot deadlock,
// whatever the number of resident blocks.
//
// Per tiled problem, a 128-byte control record (BCtl):
//   done   tasks of the current phase completed.  The block completing the
//          last one runs the step head when the phase ends a step, then
//          publishes ready = phase + 1 with agent-scope release.
//   ready  phases whose inputs are complete; a task of phase k spins until
//          ready >= k, then acquires.  The spin is bounded: on expiry the run
//          is abandoned, every block leaves, and the host reports an error.
// The wave-speed maximum of each step's last sweep is merged with an integer
// atomic max on the bit patterns of the non-negative doubles.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cosmos_fv.h"
#include "fv_sweep_core.hpp"

namespace cfv {
namespace {

using namespace core;

constexpr int kBT = 256;  // threads per block of the persistent kernels

struct BProb {
    double* buf[2];   // the two state buffers: 5 planes of `plane` doubles each
    long plane;       // doubles per plane (ny * pitch)
    long pitch;       // doubles per row
    int physics, dim, nx, ny;
    int pps;          // phases (sweeps) per step: 1 or 3
    int whole;        // 1: run by one block from start to finish
    int xk;           // rows packed per x segment (0: rows longer than a segment)
    int xspr;         // x segments per row (1 when packed)
    int nxseg;        // x segments per phase
    int xtask;        // x segments per tiled task
    int ntx;          // tiled tasks per x phase
    int ry;           // rows per y march
    int nmarch;       // y marches per phase
    int ytc;          // y sweep by LDS tiles of ytc columns (0: column marches)
    int nytc;         // y tiles across a row
    int nytile;       // y tiles per phase
    int nty;          // tiled tasks per y phase
    int max_steps;
    double dx, cfl, tf;
    PhysParams P;
};

struct alignas(128) BCtl {
    unsigned long long ws;  // merged max wave speed (bit pattern)
    int done;
    int ready;
    int iter;               // steps begun in this run
    int cur;                // buffer holding the state when the run starts
    int stopped;
    int pad;
    double t, dt;
    SweepCoef half, full;
    long long t_end;        // device wall clock at completion
};

struct BatchArgs {
    const BProb* prob;
    BCtl* ctl;
    const int* whole_list;   // problems run whole, in ticket order
    const uint2* level;      // tiled tasks of one step, in ticket order: (problem, sub << 24 | task)
    unsigned* ticket;
    int* abort;
    long long spin_limit;
    int n_whole;
    int level_len;
    int levels;              // steps in the ticket schedule
    int limiter;
    double bias;
};

// A copy of *p held in scalar registers: p is the same for every lane of the
// wave (a problem record picked by the block's task), but the compiler cannot
// prove the record unaliased by the kernel's stores, so it would keep the
// fields in vector registers.
template <class T>
__device__ __forceinline__ T uniform_copy(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "word-sized records only");
    T v;
    const int* src = reinterpret_cast<const int*>(p);
    int* dst = reinterpret_cast<int*>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return v;
}

// (buffer chosen by a select, not an index: a register copy of the record
// must not be addressed)
__device__ __forceinline__ ConstPlanes cplanes(const BProb& pb, int b) {
    const double* base = b ? pb.buf[1] : pb.buf[0];
    ConstPlanes q;
#pragma unroll
    for (int k = 0; k < 5; ++k) q.p[k] = base + k * pb.plane;
    return q;
}
__device__ __forceinline__ Planes planes(const BProb& pb, int b) {
    double* base = b ? pb.buf[1] : pb.buf[0];
    Planes q;
#pragma unroll
    for (int k = 0; k < 5; ++k) q.p[k] = base + k * pb.plane;
    return q;
}

// The solve loop's head (Solvers/EulerSecondOrderSolver.cpp:162-172 with
// EulerSolvers::computeStableTimeStep{,2D}, Solvers/EulerSolvers.cpp:130-161),
// as k_step_begin runs it: false when the loop stops.
__device__ bool step_head(const BProb& pb, BCtl* c, double smax) {
    if (!(c->t < pb.tf) || c->iter >= pb.max_steps) return false;
    double dt = pb.cfl * (pb.dx / smax);
    if (c->iter <= 5) dt *= 0.2;
    if ((c->t + dt) > pb.tf) dt = pb.tf - c->t;
    c->dt = dt;
    const double dth = 0.5 * dt;
    c->half.dt_dx = dth / pb.dx;
    c->half.dx_dt = pb.dx / dth;
    c->half.hdt_dx = 0.5 * (dth / pb.dx);
    c->full.dt_dx = dt / pb.dx;
    c->full.dx_dt = pb.dx / dt;
    c->full.hdt_dx = 0.5 * (dt / pb.dx);
    c->t = c->t + dt;
    c->iter = c->iter + 1;
    return true;
}

__device__ void finish(const BProb& pb, BCtl* c) {
    c->stopped = 1;
    c->cur ^= (c->iter * pb.pps) & 1;
    c->t_end = (long long)wall_clock64();
}

__device__ __forceinline__ double wave_max(double m) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = (o > m) ? o : m;
    }
    return m;
}

// One x segment of problem pb: packed short rows (xk rows of nx + 2G threads
// each) or one piece of a long row.
template <class Phys, int ORDER, int LIMMODE>
__device__ __forceinline__ double x_seg(const BProb& pb, const ConstPlanes& in, const Planes& out, int sg, int ws_dim,
                                        const SweepCoef& cf, double bias, int limiter, double (*lds)[kBT]) {
    constexpr int G = ORDER == 2 ? 2 : 1;
    const int s = threadIdx.x;
    int row, c;
    bool valid;
    if (pb.xk > 0) {
        const int w = pb.nx + 2 * G;
        const int q = s / w;
        c = s - q * w - G;
        row = sg * pb.xk + q;
        valid = (q < pb.xk) && (row < pb.ny) && (c >= 0) && (c < pb.nx);
        if (row >= pb.ny) row = pb.ny - 1;  // idle threads still load inside the grid
    } else {
        row = sg / pb.xspr;
        const int tile = sg - row * pb.xspr;
        c = tile * (kBT - 2 * G) - G + s;
        valid = (s >= G) && (s < kBT - G) && (c < pb.nx);
    }
    const int cc = ORDER == 2 ? mirror2(c, pb.nx) : clamp1(c, pb.nx);
    const long rb = (long)row * pb.pitch;
    return x_cells<Phys, ORDER, LIMMODE, false, kBT>(in, out, rb + cc, rb + c, valid, ws_dim, cf, bias, limiter, pb.P,
                                                     lds);
}

// y march mi of problem pb: column mi mod nx over rows [ry * (mi / nx), + ry).
template <class Phys, int ORDER, int LIMMODE>
__device__ __forceinline__ double y_one(const BProb& pb, const ConstPlanes& in, const Planes& out, int mi, bool ws,
                                        const SweepCoef& cf, double bias, int limiter) {
    const int seg = mi / pb.nx;
    const int j = mi - seg * pb.nx;
    const int i0 = seg * pb.ry;
    const int i1 = min(i0 + pb.ry, pb.ny);
    return y_march<Phys, ORDER, LIMMODE, false>(in, out, pb.ny, pb.pitch, j, true, j, i0, i1, ws, cf, bias, limiter,
                                                pb.P);
}

// y tile yt of problem pb (YT columns by kBT / YT rows, 2G of them halo).
template <class Phys, int ORDER, int LIMMODE, int YT>
__device__ __forceinline__ double y_tile(const BProb& pb, const ConstPlanes& in, const Planes& out, int yt, bool ws,
                                         const SweepCoef& cf, double bias, int limiter, double (*lds)[kBT]) {
    constexpr int G = ORDER == 2 ? 2 : 1;
    constexpr int TRV = kBT / YT - 2 * G;
    const int br = yt / pb.nytc;
    const int bc = yt - br * pb.nytc;
    return y_tile_cells<Phys, ORDER, LIMMODE, false, kBT, YT>(in, out, pb.nx, pb.ny, pb.pitch, bc * YT, br * TRV,
                                                              ws ? 2 : 0, cf, bias, limiter, pb.P, lds);
}

struct WholeShared {
    SweepCoef half, full;
    double red[kBT / 64];
    int more;
    int cur;
};

template <class Phys, int ORDER, int LIMMODE, int YT>
__device__ void run_whole(const BatchArgs& A, int p, double (*lds)[kBT], WholeShared& sh) {
    const BProb pb = uniform_copy(A.prob + p);
    BCtl* c = A.ctl + p;
    if (threadIdx.x == 0) {
        sh.more = !c->stopped;
        sh.half = c->half;
        sh.full = c->full;
        sh.cur = c->cur;
    }
    __syncthreads();
    const int cur = sh.cur;
    int phase = 0;
    while (sh.more) {
        double m = 0.0;
        for (int sub = 0; sub < pb.pps; ++sub, ++phase) {
            const SweepCoef cf = (pb.dim == 1 || sub == 1) ? sh.full : sh.half;
            const int src = cur ^ (phase & 1);
            const ConstPlanes in = cplanes(pb, src);
            const Planes out = planes(pb, src ^ 1);
            const bool last = sub == pb.pps - 1;
            if (pb.dim == 1 || sub != 1) {
                for (int sg = 0; sg < pb.nxseg; ++sg) {
                    if (sg) __syncthreads();
                    const double w = x_seg<Phys, ORDER, LIMMODE>(pb, in, out, sg, last ? pb.dim : 0, cf, A.bias,
                                                                 A.limiter, lds);
                    m = (w > m) ? w : m;
                }
            } else if constexpr (YT == 0) {
                for (int mi = threadIdx.x; mi < pb.nmarch; mi += kBT) {
                    const double w = y_one<Phys, ORDER, LIMMODE>(pb, in, out, mi, last, cf, A.bias, A.limiter);
                    m = (w > m) ? w : m;
                }
            } else {
                for (int yt = 0; yt < pb.nytile; ++yt) {
                    if (yt) __syncthreads();
                    const double w = y_tile<Phys, ORDER, LIMMODE, YT>(pb, in, out, yt, last, cf, A.bias, A.limiter, lds);
                    m = (w > m) ? w : m;
                }
            }
            // the phase's outputs are complete before the next phase reads them
            __syncthreads();
        }
        m = wave_max(m);
        if ((threadIdx.x & 63) == 0) sh.red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double smax = sh.red[0];
#pragma unroll
            for (int w = 1; w < kBT / 64; ++w) smax = (sh.red[w] > smax) ? sh.red[w] : smax;
            const bool more = step_head(pb, c, smax);
            if (!more) finish(pb, c);
            sh.more = more;
            sh.half = c->half;
            sh.full = c->full;
        }
        __syncthreads();
    }
}

// One tiled task: phase `step * pps + sub` of problem p, task index `task`.
template <class Phys, int ORDER, int LIMMODE, int YT>
__device__ void run_task(const BatchArgs& A, int p, int sub, int task, int step, double (*lds)[kBT], int& s_go) {
    const BProb pb = uniform_copy(A.prob + p);
    BCtl* c = A.ctl + p;
    const int phase = step * pb.pps + sub;
    if (threadIdx.x == 0) {
        int go = sub < pb.pps;
        long long spins = 0;
        while (go && __hip_atomic_load(&c->ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < phase) {
            if (__hip_atomic_load(A.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                go = 0;
            } else if (++spins > A.spin_limit) {
                __hip_atomic_store(A.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                go = 0;
            } else {
                __builtin_amdgcn_s_sleep(8);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (go && c->stopped) go = 0;
        s_go = go;
    }
    __syncthreads();
    if (!s_go) return;
    const bool xphase = pb.dim == 1 || sub != 1;
    const SweepCoef cf = xphase && pb.dim == 2 ? c->half : c->full;
    const int src = c->cur ^ (phase & 1);
    const ConstPlanes in = cplanes(pb, src);
    const Planes out = planes(pb, src ^ 1);
    const bool last = sub == pb.pps - 1;
    double m = 0.0;
    if (xphase) {
        const int s0 = task * pb.xtask;
        const int s1 = min(s0 + pb.xtask, pb.nxseg);
        for (int sg = s0; sg < s1; ++sg) {
            if (sg > s0) __syncthreads();
            const double w =
                x_seg<Phys, ORDER, LIMMODE>(pb, in, out, sg, last ? pb.dim : 0, cf, A.bias, A.limiter, lds);
            m = (w > m) ? w : m;
        }
    } else if constexpr (YT == 0) {
        const int mi = task * kBT + threadIdx.x;
        if (mi < pb.nmarch) m = y_one<Phys, ORDER, LIMMODE>(pb, in, out, mi, last, cf, A.bias, A.limiter);
    } else {
        const int t0 = task * pb.xtask;
        const int t1 = min(t0 + pb.xtask, pb.nytile);
        for (int yt = t0; yt < t1; ++yt) {
            if (yt > t0) __syncthreads();
            const double w = y_tile<Phys, ORDER, LIMMODE, YT>(pb, in, out, yt, last, cf, A.bias, A.limiter, lds);
            m = (w > m) ? w : m;
        }
    }
    if (last) {
        m = wave_max(m);
        if ((threadIdx.x & 63) == 0 && m > 0.0)
            atomicMax(&c->ws, (unsigned long long)__double_as_longlong(m));
    }
    // every wave's stores are written back before the task is counted
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        const int ntask = xphase ? pb.ntx : pb.nty;
        const int old = __hip_atomic_fetch_add(&c->done, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == ntask - 1) {
            __hip_atomic_store(&c->done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bool more = true;
            if (last) {
                const unsigned long long bits =
                    __hip_atomic_exchange(&c->ws, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                more = step_head(pb, c, __longlong_as_double((long long)bits));
                if (!more) finish(pb, c);
            }
            __hip_atomic_store(&c->ready, more ? phase + 1 : INT_MAX, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Waves per SIMD: with the column march (YT = 0) the kernel holds its
// register window (the exact Newtonian march alone needs ~142 VGPRs, the
// special-relativistic ~180), so two waves (256 VGPRs); with LDS y tiles
// both sweeps are the x sweep's thread-per-cell stages, which fit four.
template <int YT>
constexpr int batch_min_waves() { return YT == 0 ? 2 : 3; }

template <class Phys, int ORDER, int LIMMODE, int YT>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(batch_min_waves<YT>(), 8))) void k_batch(
    BatchArgs A) {
    __shared__ double lds[15][kBT];
    __shared__ WholeShared sh;
    __shared__ int s_ticket;
    __shared__ int s_go;
    const long total = (long)A.n_whole + (long)A.levels * A.level_len;
    for (;;) {
        if (threadIdx.x == 0) {
            int t = (int)atomicAdd(A.ticket, 1u);
            if (__hip_atomic_load(A.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) t = INT_MAX;
            s_ticket = t;
        }
        __syncthreads();
        // uniform across the block: keep the task's indices (and so every
        // problem-record load) in scalar registers
        const int t = __builtin_amdgcn_readfirstlane(s_ticket);
        __syncthreads();
        if ((long)t >= total) return;
        if (t < A.n_whole) {
            run_whole<Phys, ORDER, LIMMODE, YT>(A, __builtin_amdgcn_readfirstlane(A.whole_list[t]), lds, sh);
        } else {
            const int u = t - A.n_whole;
            const int step = u / A.level_len;
            const uint2 e = A.level[u - step * A.level_len];
            const int p = __builtin_amdgcn_readfirstlane((int)e.x);
            const int y = __builtin_amdgcn_readfirstlane((int)e.y);
            run_task<Phys, ORDER, LIMMODE, YT>(A, p, y >> 24, y & 0xffffff, step, lds, s_go);
        }
        __syncthreads();
    }
}

// Set-up kernels of a run: reset the control records and tickets, merge each
// problem's initial max wave speed, run the first step head.
__global__ __launch_bounds__(256) void k_batch_reset(BCtl* ctl, int n, unsigned* tickets, int* abort,
                                                     long long* t0) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        BCtl* c = ctl + p;
        c->ws = 0ull;
        c->done = 0;
        c->ready = 0;
        c->iter = 0;
        c->stopped = 0;
        c->t = 0.0;
        c->dt = 0.0;
        c->t_end = 0;
    }
    if (p == 0) {
        tickets[0] = 0u;
        tickets[1] = 0u;
        *abort = 0;
        *t0 = (long long)wall_clock64();
    }
}

__global__ __launch_bounds__(256) void k_batch_init_ws(const BProb* prob, BCtl* ctl) {
    const BProb& pb = prob[blockIdx.x];
    BCtl* c = ctl + blockIdx.x;
    const ConstPlanes in = cplanes(pb, c->cur);
    const long cells = (long)pb.nx * pb.ny;
    double m = 0.0;
    for (long k = (long)blockIdx.y * blockDim.x + threadIdx.x; k < cells; k += (long)gridDim.y * blockDim.x) {
        const long i = k / pb.nx, j = k - i * pb.nx;
        double W[5];
        load_cell<false>(in, i * pb.pitch + j, W);
        double w;
        if (pb.physics == 0) w = pb.dim == 1 ? Newtonian::wavespeed<1>(W, pb.P) : Newtonian::wavespeed<2>(W, pb.P);
        else w = pb.dim == 1 ? Relativistic::wavespeed<1>(W, pb.P) : Relativistic::wavespeed<2>(W, pb.P);
        m = (w > m) ? w : m;
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(&c->ws, (unsigned long long)__double_as_longlong(m));
}

__global__ __launch_bounds__(256) void k_batch_init_head(const BProb* prob, BCtl* ctl, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const BProb& pb = prob[p];
    BCtl* c = ctl + p;
    const double smax = __longlong_as_double((long long)c->ws);
    c->ws = 0ull;
    const bool more = step_head(pb, c, smax);
    if (!more) finish(pb, c);
    c->ready = more ? 0 : INT_MAX;
}

__global__ __launch_bounds__(256) void k_batch_set_cur(BCtl* ctl, int p, int cur) { ctl[p].cur = cur; }

__global__ __launch_bounds__(256) void k_batch_aos_to_planes(const double* aos, Planes out, int nx, int ny,
                                                             long pitch) {
    const long total = (long)nx * ny;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (long)gridDim.x * blockDim.x) {
        const long i = c / nx, j = c - i * nx;
#pragma unroll
        for (int k = 0; k < 5; ++k) out.p[k][i * pitch + j] = aos[5 * c + k];
    }
}

__global__ __launch_bounds__(256) void k_batch_planes_to_aos(ConstPlanes in, double* aos, int nx, int ny,
                                                             long pitch) {
    const long total = (long)nx * ny;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (long)gridDim.x * blockDim.x) {
        const long i = c / nx, j = c - i * nx;
#pragma unroll
        for (int k = 0; k < 5; ++k) aos[5 * c + k] = in.p[k][i * pitch + j];
    }
}

typedef void (*BatchKernel)(BatchArgs);

template <class Phys, int YT>
BatchKernel pick_kernel_yt(int order, int limmode) {
    if (order == 1) return k_batch<Phys, 1, 0, YT>;
    if (limmode == 1) return k_batch<Phys, 2, 1, YT>;
    return k_batch<Phys, 2, 0, YT>;
}

template <class Phys>
BatchKernel pick_kernel(int order, int limmode, int ytc) {
    if (ytc == 16) return pick_kernel_yt<Phys, 16>(order, limmode);
    if (ytc == 8) return pick_kernel_yt<Phys, 8>(order, limmode);
    return pick_kernel_yt<Phys, 0>(order, limmode);
}

}  // namespace
}  // namespace cfv

using cfv::BCtl;
using cfv::BProb;

namespace {

struct PhysSchedule {
    cosmos_batch_schedule cfg;
    std::vector<int> whole;          // problems run whole, in ticket order
    std::vector<uint2> level;        // tiled tasks of one step
    int levels = 0;
    int* d_whole = nullptr;
    uint2* d_level = nullptr;
    int grid = 0;
    bool present = false;
};

}  // namespace

struct cosmos_batch {
    int n = 0;
    int order = 2, limiter = 1;
    double bias = 0.0;
    double c2p_tol = 1e-8;
    int c2p_maxit = 1000;
    std::vector<cosmos_batch_problem> desc;
    std::vector<BProb> prob;
    std::vector<int> cur;            // host mirror of each problem's current buffer
    BProb* d_prob = nullptr;
    BCtl* d_ctl = nullptr;
    double* arena = nullptr;
    size_t arena_doubles = 0;
    double* backup = nullptr;        // arena copy for autotuning trials
    double* stage = nullptr;
    size_t stage_doubles = 0;
    unsigned* d_tickets = nullptr;   // one counter per physics
    int* d_abort = nullptr;
    long long* d_t0 = nullptr;
    PhysSchedule sched[2];
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;               // schedule or parameters changed since the graph was built
    int cus = 0;
    double clock_khz = 0.0;          // device wall clock rate
    long long spin_limit = 1ll << 22;  // about 4 s of 1 us polls
    std::string err;
};

namespace {

thread_local std::string g_batch_create_error;

int bfail(cosmos_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_batch_create_error = msg;
    return code;
}

#define BCHK(b, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return bfail((b), COSMOS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class F>
int bguard(cosmos_batch* b, F&& f) {
    if (!b) return bfail(nullptr, COSMOS_ERR_ARGUMENT, "null batch handle");
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return bfail(b, COSMOS_ERR_DEVICE, "host allocation failed");
    } catch (...) {
        return bfail(b, COSMOS_ERR_STATE, "unexpected exception");
    }
}

long round_up(long v, long m) { return (v + m - 1) / m * m; }

cfv::PhysParams params_for(const cosmos_batch* b, double gamma) {
    cfv::PhysParams P = cfv::make_params(gamma);
    P.c2p_tol = b->c2p_tol;
    P.c2p_maxit = b->c2p_maxit;
    return P;
}

// Relative cost of a cell-sweep by physics (special-relativistic sweeps run
// a Newton con2prim per flux evaluation); only orders work.
double cell_cost(int physics) { return physics == 1 ? 8.0 : 1.0; }

double problem_cost(const cosmos_batch_problem& d) {
    const double phases = d.dim == 2 ? 3.0 : 1.0;
    return cell_cost(d.physics) * phases * (double)d.nx * (d.dim == 2 ? d.ny : 1) * (double)d.max_steps;
}

// Geometry of the segments, marches and tasks of problem p under the schedule of its physics.
void plan_problem(const cosmos_batch* b, BProb& pb, const cosmos_batch_schedule& cfg) {
    const int G = b->order == 2 ? 2 : 1;
    const int w = pb.nx + 2 * G;
    if (w <= cfv::kBT) {
        pb.xk = cfv::kBT / w;
        pb.xspr = 1;
        pb.nxseg = (pb.ny + pb.xk - 1) / pb.xk;
    } else {
        pb.xk = 0;
        pb.xspr = (pb.nx + (cfv::kBT - 2 * G) - 1) / (cfv::kBT - 2 * G);
        pb.nxseg = pb.ny * pb.xspr;
    }
    pb.xtask = std::max(1, cfg.x_segments_per_task);
    pb.ntx = (pb.nxseg + pb.xtask - 1) / pb.xtask;
    pb.ry = std::max(1, cfg.y_rows);
    pb.nmarch = pb.dim == 2 ? (pb.ny + pb.ry - 1) / pb.ry * pb.nx : 0;
    pb.ytc = cfg.y_tile_cols;
    if (pb.ytc > 0 && pb.dim == 2) {
        const int trv = cfv::kBT / pb.ytc - 2 * G;
        pb.nytc = (pb.nx + pb.ytc - 1) / pb.ytc;
        pb.nytile = pb.nytc * ((pb.ny + trv - 1) / trv);
        pb.nty = (pb.nytile + pb.xtask - 1) / pb.xtask;
    } else {
        pb.nytc = pb.nytile = 0;
        pb.nty = (pb.nmarch + cfv::kBT - 1) / cfv::kBT;
    }
    const long cells = (long)pb.nx * pb.ny;
    pb.whole = cells <= (long)cfg.whole_max_cells ? 1 : 0;
}

cfv::BatchKernel kernel_for(const cosmos_batch* b, int ph) {
    const int lm = (b->order == 2 && b->limiter == 1 && b->bias == 0.0) ? 1 : 0;
    const int ytc = b->sched[ph].cfg.y_tile_cols;
    return ph == 0 ? cfv::pick_kernel<cfv::Newtonian>(b->order, lm, ytc)
                   : cfv::pick_kernel<cfv::Relativistic>(b->order, lm, ytc);
}

int build_schedules(cosmos_batch* b) {
    for (int ph = 0; ph < 2; ++ph) {
        PhysSchedule& S = b->sched[ph];
        S.whole.clear();
        S.level.clear();
        S.levels = 0;
        S.present = false;
        std::vector<int> tiled;
        for (int p = 0; p < b->n; ++p) {
            if (b->desc[p].physics != ph) continue;
            S.present = true;
            plan_problem(b, b->prob[p], S.cfg);
            if (b->prob[p].whole) S.whole.push_back(p);
            else tiled.push_back(p);
        }
        auto by_cost = [&](int a, int c) {
            const double ca = problem_cost(b->desc[a]), cc = problem_cost(b->desc[c]);
            return ca != cc ? ca > cc : a < c;
        };
        std::sort(S.whole.begin(), S.whole.end(), by_cost);
        std::sort(tiled.begin(), tiled.end(), by_cost);
        for (int sub = 0; sub < 3; ++sub)
            for (int p : tiled) {
                const BProb& pb = b->prob[p];
                if (sub >= pb.pps) continue;
                const int nt = (pb.dim == 1 || sub != 1) ? pb.ntx : pb.nty;
                if (nt >= (1 << 24)) return bfail(b, COSMOS_ERR_ARGUMENT, "too many tasks per phase: raise x_segments_per_task");
                for (int t = 0; t < nt; ++t) S.level.push_back(make_uint2((unsigned)p, ((unsigned)sub << 24) | (unsigned)t));
                S.levels = std::max(S.levels, pb.max_steps);
            }
        const double total = (double)S.whole.size() + (double)S.levels * (double)S.level.size();
        if (total >= (double)INT_MAX)
            return bfail(b, COSMOS_ERR_ARGUMENT, "batch schedule exceeds 2^31 tasks: lower max_steps or raise x_segments_per_task");
        (void)hipFree(S.d_whole);
        (void)hipFree(S.d_level);
        S.d_whole = nullptr;
        S.d_level = nullptr;
        if (!S.whole.empty()) {
            BCHK(b, hipMalloc((void**)&S.d_whole, sizeof(int) * S.whole.size()));
            BCHK(b, hipMemcpy(S.d_whole, S.whole.data(), sizeof(int) * S.whole.size(), hipMemcpyHostToDevice));
        }
        if (!S.level.empty()) {
            BCHK(b, hipMalloc((void**)&S.d_level, sizeof(uint2) * S.level.size()));
            BCHK(b, hipMemcpy(S.d_level, S.level.data(), sizeof(uint2) * S.level.size(), hipMemcpyHostToDevice));
        }
        int occ = 0;
        cfv::BatchKernel k = kernel_for(b, ph);
        BCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k, cfv::kBT, 0));
        const int per_cu = S.cfg.blocks_per_cu > 0 ? S.cfg.blocks_per_cu : std::max(1, occ);
        S.grid = per_cu * b->cus;
    }
    for (int p = 0; p < b->n; ++p) b->prob[p].P = params_for(b, b->desc[p].adiabatic_index);
    BCHK(b, hipMemcpy(b->d_prob, b->prob.data(), sizeof(BProb) * b->n, hipMemcpyHostToDevice));
    return COSMOS_OK;
}

cfv::BatchArgs args_for(cosmos_batch* b, int ph) {
    PhysSchedule& S = b->sched[ph];
    cfv::BatchArgs A{};
    A.prob = b->d_prob;
    A.ctl = b->d_ctl;
    A.whole_list = S.d_whole;
    A.level = S.d_level;
    A.ticket = b->d_tickets + ph;
    A.abort = b->d_abort;
    A.spin_limit = b->spin_limit;
    A.n_whole = (int)S.whole.size();
    A.level_len = (int)S.level.size();
    A.levels = S.level.empty() ? 0 : S.levels;
    A.limiter = b->limiter;
    A.bias = b->bias;
    return A;
}

// Queue one run on b->stream (also used under stream capture).
int queue_run(cosmos_batch* b) {
    const dim3 blk(256);
    const dim3 gp((unsigned)((b->n + 255) / 256));
    hipLaunchKernelGGL(cfv::k_batch_reset, gp, blk, 0, b->stream, b->d_ctl, b->n, b->d_tickets, b->d_abort, b->d_t0);
    BCHK(b, hipGetLastError());
    hipLaunchKernelGGL(cfv::k_batch_init_ws, dim3((unsigned)b->n, 64), blk, 0, b->stream, (const BProb*)b->d_prob,
                       b->d_ctl);
    BCHK(b, hipGetLastError());
    hipLaunchKernelGGL(cfv::k_batch_init_head, gp, blk, 0, b->stream, (const BProb*)b->d_prob, b->d_ctl, b->n);
    BCHK(b, hipGetLastError());
    const bool both = b->sched[0].present && b->sched[1].present;
    if (both) {
        BCHK(b, hipEventRecord(b->fork, b->stream));
        BCHK(b, hipStreamWaitEvent(b->side, b->fork, 0));
    }
    for (int ph = 0; ph < 2; ++ph) {
        PhysSchedule& S = b->sched[ph];
        if (!S.present) continue;
        hipStream_t st = (both && ph == 1) ? b->side : b->stream;
        cfv::BatchKernel k = kernel_for(b, ph);
        cfv::BatchArgs A = args_for(b, ph);
        void* kargs[] = {&A};
        BCHK(b, hipLaunchKernel((const void*)k, dim3((unsigned)S.grid), blk, kargs, 0, st));
    }
    if (both) {
        BCHK(b, hipEventRecord(b->join, b->side));
        BCHK(b, hipStreamWaitEvent(b->stream, b->join, 0));
    }
    return COSMOS_OK;
}

int rebuild(cosmos_batch* b) {
    if (!b->dirty) return COSMOS_OK;
    int rc = build_schedules(b);
    if (rc) return rc;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = queue_run(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, COSMOS_ERR_DEVICE, std::string("batch graph capture: ") + hipGetErrorString(e));
    b->graph = g;
    BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->dirty = false;
    return COSMOS_OK;
}

int launch(cosmos_batch* b) {
    int rc = rebuild(b);
    if (rc) return rc;
    BCHK(b, hipGraphLaunch(b->exec, b->stream));
    return COSMOS_OK;
}

// After a run: wait, check the abort flag, refresh the host mirror of the buffers.
int settle_run(cosmos_batch* b) {
    BCHK(b, hipStreamSynchronize(b->stream));
    int ab = 0;
    BCHK(b, hipMemcpy(&ab, b->d_abort, sizeof(int), hipMemcpyDeviceToHost));
    if (ab) return bfail(b, COSMOS_ERR_STATE, "batch run abandoned: a task waited past the spin limit");
    std::vector<BCtl> ctl(b->n);
    BCHK(b, hipMemcpy(ctl.data(), b->d_ctl, sizeof(BCtl) * b->n, hipMemcpyDeviceToHost));
    for (int p = 0; p < b->n; ++p) b->cur[p] = ctl[p].cur;
    return COSMOS_OK;
}

int check_problem(const cosmos_batch_problem& d, int order, std::string& why) {
    if (d.physics != COSMOS_NEWTONIAN && d.physics != COSMOS_RELATIVISTIC) why = "physics must be Newtonian or relativistic";
    else if (d.dim != 1 && d.dim != 2) why = "dim must be 1 or 2";
    else if (d.nx < (order == 2 ? 2 : 1) || (d.dim == 2 && d.ny < (order == 2 ? 2 : 1)))
        why = order == 2 ? "SLIC needs at least 2 cells along each axis" : "grids need at least one cell";
    else if ((long)d.nx * (d.dim == 2 ? d.ny : 1) > (1l << 28)) why = "problem too large";
    else if (!(d.cell_spacing > 0.0) || !(d.cfl > 0.0)) why = "cell_spacing and cfl must be positive";
    else if (d.max_steps < 1 || d.max_steps > 1000000) why = "max_steps must be in [1, 1000000]";
    else if (!(d.adiabatic_index > 1.0)) why = "adiabatic_index must exceed 1";
    else return 0;
    return 1;
}

}  // namespace

extern "C" {

int cosmos_batch_default_schedule(int physics, cosmos_batch_schedule* out) {
    if (!out) return COSMOS_ERR_ARGUMENT;
    // the autotuner's choice on the 1024-problem Riemann workload
    out->whole_max_cells = physics == COSMOS_RELATIVISTIC ? 4096 : 65536;
    out->x_segments_per_task = physics == COSMOS_RELATIVISTIC ? 16 : 64;
    out->y_rows = physics == COSMOS_RELATIVISTIC ? 32 : 64;
    out->blocks_per_cu = 0;
    out->y_tile_cols = 8;
    return COSMOS_OK;
}

int cosmos_batch_create(cosmos_batch** out, const cosmos_batch_problem* problems, int n, int order, double bias,
                        int slope_limiter) {
    if (!out) return bfail(nullptr, COSMOS_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    if (!problems || n < 1) return bfail(nullptr, COSMOS_ERR_ARGUMENT, "need at least one problem");
    if (order != 1 && order != 2) return bfail(nullptr, COSMOS_ERR_ARGUMENT, "order must be 1 (FORCE) or 2 (SLIC)");
    for (int p = 0; p < n; ++p) {
        std::string why;
        if (check_problem(problems[p], order, why))
            return bfail(nullptr, COSMOS_ERR_ARGUMENT, "problem " + std::to_string(p) + ": " + why);
    }
    cosmos_batch* b = new (std::nothrow) cosmos_batch();
    if (!b) return bfail(nullptr, COSMOS_ERR_DEVICE, "host allocation failed");
    auto bail = [&](const std::string& msg) {
        cosmos_batch_destroy(b);
        return bfail(nullptr, COSMOS_ERR_DEVICE, msg);
    };
    b->n = n;
    b->order = order;
    b->bias = bias;
    b->limiter = slope_limiter;
    b->desc.assign(problems, problems + n);
    b->prob.resize(n);
    b->cur.assign(n, 0);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return bail("device setup: no ROCm-capable device is detected");
    b->cus = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
    b->clock_khz = khz;
    cosmos_batch_default_schedule(COSMOS_NEWTONIAN, &b->sched[0].cfg);
    cosmos_batch_default_schedule(COSMOS_RELATIVISTIC, &b->sched[1].cfg);
    size_t off = 0, max_cells = 0;
    std::vector<size_t> offs(n);
    for (int p = 0; p < n; ++p) {
        const cosmos_batch_problem& d = problems[p];
        BProb& pb = b->prob[p];
        pb.physics = d.physics;
        pb.dim = d.dim;
        pb.nx = d.nx;
        pb.ny = d.dim == 2 ? d.ny : 1;
        pb.pitch = round_up(d.nx, 16);
        pb.plane = (long)pb.ny * pb.pitch;
        pb.pps = d.dim == 2 ? 3 : 1;
        pb.max_steps = d.max_steps;
        pb.dx = d.cell_spacing;
        pb.cfl = d.cfl;
        pb.tf = d.final_time;
        offs[p] = off;
        off += 2 * 5 * (size_t)pb.plane;
        max_cells = std::max(max_cells, (size_t)pb.nx * pb.ny);
    }
    b->arena_doubles = off;
    b->stage_doubles = 5 * max_cells;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipMalloc((void**)&b->arena, sizeof(double) * b->arena_doubles);
    if (e == hipSuccess) e = hipMalloc((void**)&b->stage, sizeof(double) * b->stage_doubles);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_prob, sizeof(BProb) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_ctl, sizeof(BCtl) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_tickets, sizeof(unsigned) * 2);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_abort, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_t0, sizeof(long long));
    if (e == hipSuccess) e = hipMemset(b->arena, 0, sizeof(double) * b->arena_doubles);
    if (e == hipSuccess) e = hipMemset(b->d_ctl, 0, sizeof(BCtl) * n);
    if (e == hipSuccess) e = hipMemset(b->d_abort, 0, sizeof(int));
    if (e != hipSuccess) return bail(std::string("batch setup: ") + hipGetErrorString(e));
    for (int p = 0; p < n; ++p) {
        b->prob[p].buf[0] = b->arena + offs[p];
        b->prob[p].buf[1] = b->arena + offs[p] + 5 * (size_t)b->prob[p].plane;
    }
    *out = b;
    return COSMOS_OK;
}

int cosmos_batch_destroy(cosmos_batch* b) {
    if (!b) return COSMOS_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->side) (void)hipStreamSynchronize(b->side);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (PhysSchedule& S : b->sched) {
        (void)hipFree(S.d_whole);
        (void)hipFree(S.d_level);
    }
    (void)hipFree(b->arena);
    (void)hipFree(b->backup);
    (void)hipFree(b->stage);
    (void)hipFree(b->d_prob);
    (void)hipFree(b->d_ctl);
    (void)hipFree(b->d_tickets);
    (void)hipFree(b->d_abort);
    (void)hipFree(b->d_t0);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->side) (void)hipStreamDestroy(b->side);
    delete b;
    return COSMOS_OK;
}

const char* cosmos_batch_last_error(const cosmos_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int cosmos_batch_set_con2prim(cosmos_batch* b, double tolerance, int max_iterations) {
    return bguard(b, [&] {
        if (!(tolerance > 0.0) || max_iterations < 1)
            return bfail(b, COSMOS_ERR_ARGUMENT, "tolerance must be positive and max_iterations >= 1");
        b->c2p_tol = tolerance;
        b->c2p_maxit = max_iterations;
        b->dirty = true;
        return COSMOS_OK;
    });
}

int cosmos_batch_set_schedule(cosmos_batch* b, int physics, const cosmos_batch_schedule* s) {
    return bguard(b, [&] {
        if (physics != COSMOS_NEWTONIAN && physics != COSMOS_RELATIVISTIC)
            return bfail(b, COSMOS_ERR_ARGUMENT, "physics must be Newtonian or relativistic");
        if (!s || s->whole_max_cells < 0 || s->x_segments_per_task < 1 || s->y_rows < 1 || s->y_rows > 4096 ||
            s->blocks_per_cu < 0 || s->blocks_per_cu > 32 ||
            (s->y_tile_cols != 0 && s->y_tile_cols != 8 && s->y_tile_cols != 16))
            return bfail(b, COSMOS_ERR_ARGUMENT, "invalid schedule");
        b->sched[physics].cfg = *s;
        b->dirty = true;
        return COSMOS_OK;
    });
}

int cosmos_batch_get_schedule(cosmos_batch* b, int physics, cosmos_batch_schedule* s) {
    return bguard(b, [&] {
        if ((physics != COSMOS_NEWTONIAN && physics != COSMOS_RELATIVISTIC) || !s)
            return bfail(b, COSMOS_ERR_ARGUMENT, "invalid arguments");
        *s = b->sched[physics].cfg;
        return COSMOS_OK;
    });
}

int cosmos_batch_upload(cosmos_batch* b, int problem, const double* prim) {
    return bguard(b, [&] {
        if (problem < 0 || problem >= b->n || !prim) return bfail(b, COSMOS_ERR_ARGUMENT, "invalid problem or pointer");
        const BProb& pb = b->prob[problem];
        const long cells = (long)pb.nx * pb.ny;
        BCHK(b, hipStreamSynchronize(b->stream));
        BCHK(b, hipMemcpyAsync(b->stage, prim, sizeof(double) * 5 * cells, hipMemcpyHostToDevice, b->stream));
        cfv::Planes q;
        for (int k = 0; k < 5; ++k) q.p[k] = pb.buf[0] + k * pb.plane;
        hipLaunchKernelGGL(cfv::k_batch_aos_to_planes, dim3((unsigned)std::min<long>(1024, (cells + 255) / 256)),
                           dim3(256), 0, b->stream, (const double*)b->stage, q, pb.nx, pb.ny, pb.pitch);
        BCHK(b, hipGetLastError());
        hipLaunchKernelGGL(cfv::k_batch_set_cur, dim3(1), dim3(1), 0, b->stream, b->d_ctl, problem, 0);
        BCHK(b, hipGetLastError());
        BCHK(b, hipStreamSynchronize(b->stream));
        b->cur[problem] = 0;
        return COSMOS_OK;
    });
}

int cosmos_batch_download(cosmos_batch* b, int problem, double* prim) {
    return bguard(b, [&] {
        if (problem < 0 || problem >= b->n || !prim) return bfail(b, COSMOS_ERR_ARGUMENT, "invalid problem or pointer");
        const BProb& pb = b->prob[problem];
        const long cells = (long)pb.nx * pb.ny;
        BCHK(b, hipStreamSynchronize(b->stream));
        cfv::ConstPlanes q;
        for (int k = 0; k < 5; ++k) q.p[k] = pb.buf[b->cur[problem]] + k * pb.plane;
        hipLaunchKernelGGL(cfv::k_batch_planes_to_aos, dim3((unsigned)std::min<long>(1024, (cells + 255) / 256)),
                           dim3(256), 0, b->stream, q, b->stage, pb.nx, pb.ny, pb.pitch);
        BCHK(b, hipGetLastError());
        BCHK(b, hipMemcpyAsync(prim, b->stage, sizeof(double) * 5 * cells, hipMemcpyDeviceToHost, b->stream));
        BCHK(b, hipStreamSynchronize(b->stream));
        return COSMOS_OK;
    });
}

int cosmos_batch_run(cosmos_batch* b) {
    return bguard(b, [&] {
        int rc = launch(b);
        if (rc) return rc;
        return settle_run(b);
    });
}

int cosmos_batch_run_timed(cosmos_batch* b, double* ms) {
    return bguard(b, [&] {
        int rc = rebuild(b);
        if (rc) return rc;
        BCHK(b, hipEventRecord(b->ev0, b->stream));
        BCHK(b, hipGraphLaunch(b->exec, b->stream));
        BCHK(b, hipEventRecord(b->ev1, b->stream));
        rc = settle_run(b);
        if (rc) return rc;
        float f = 0.0f;
        BCHK(b, hipEventElapsedTime(&f, b->ev0, b->ev1));
        if (ms) *ms = f;
        return COSMOS_OK;
    });
}

int cosmos_batch_stats(cosmos_batch* b, double* time_reached, int* steps_taken, double* finish_us) {
    return bguard(b, [&] {
        BCHK(b, hipStreamSynchronize(b->stream));
        std::vector<BCtl> ctl(b->n);
        long long t0 = 0;
        BCHK(b, hipMemcpy(ctl.data(), b->d_ctl, sizeof(BCtl) * b->n, hipMemcpyDeviceToHost));
        BCHK(b, hipMemcpy(&t0, b->d_t0, sizeof t0, hipMemcpyDeviceToHost));
        for (int p = 0; p < b->n; ++p) {
            if (time_reached) time_reached[p] = ctl[p].t;
            if (steps_taken) steps_taken[p] = ctl[p].iter;
            if (finish_us) finish_us[p] = (double)(ctl[p].t_end - t0) * 1000.0 / b->clock_khz;
        }
        return COSMOS_OK;
    });
}

int cosmos_batch_autotune(cosmos_batch* b, int trial_steps, double* best_ms) {
    return bguard(b, [&] {
        if (trial_steps < 1) return bfail(b, COSMOS_ERR_ARGUMENT, "trial_steps must be >= 1");
        if (!b->backup) BCHK(b, hipMalloc((void**)&b->backup, sizeof(double) * b->arena_doubles));
        BCHK(b, hipStreamSynchronize(b->stream));
        BCHK(b, hipMemcpy(b->backup, b->arena, sizeof(double) * b->arena_doubles, hipMemcpyDeviceToDevice));
        const std::vector<int> cur0 = b->cur;
        std::vector<BCtl> ctl0(b->n);
        BCHK(b, hipMemcpy(ctl0.data(), b->d_ctl, sizeof(BCtl) * b->n, hipMemcpyDeviceToHost));
        std::vector<int> steps0(b->n);
        for (int p = 0; p < b->n; ++p) {
            steps0[p] = b->desc[p].max_steps;
            b->desc[p].max_steps = std::min(steps0[p], trial_steps);
            b->prob[p].max_steps = b->desc[p].max_steps;
        }
        auto restore = [&]() -> int {
            BCHK(b, hipMemcpy(b->arena, b->backup, sizeof(double) * b->arena_doubles, hipMemcpyDeviceToDevice));
            BCHK(b, hipMemcpy(b->d_ctl, ctl0.data(), sizeof(BCtl) * b->n, hipMemcpyHostToDevice));
            b->cur = cur0;
            return COSMOS_OK;
        };
        auto trial = [&](double& ms) -> int {
            b->dirty = true;
            double best = 1e300;
            for (int r = 0; r < 2; ++r) {
                double t = 0.0;
                int rc = cosmos_batch_run_timed(b, &t);
                if (rc) return rc;
                rc = restore();
                if (rc) return rc;
                best = std::min(best, t);
            }
            ms = best;
            return COSMOS_OK;
        };
        // coordinate descent over each present physics' schedule parameters
        double cur_ms = 0.0;
        int rc = trial(cur_ms);
        if (rc) return rc;
        const std::vector<int> cands[4] = {{0, 4096, 16384, 65536, 262144},
                                           {2, 4, 8, 16, 32, 64},
                                           {0, 8, 16},
                                           {8, 16, 32, 64, 128}};
        for (int ph = 0; ph < 2 && !rc; ++ph) {
            if (!b->sched[ph].present) continue;
            for (int param = 0; param < 4 && !rc; ++param) {
                const std::vector<int>& cand = cands[param];
                cosmos_batch_schedule best = b->sched[ph].cfg;
                if (param == 3 && best.y_tile_cols != 0) continue;  // y_rows only shapes the column march
                for (size_t k = 0; k < cand.size(); ++k) {
                    cosmos_batch_schedule c = best;
                    int& field = param == 0 ? c.whole_max_cells
                                            : (param == 1 ? c.x_segments_per_task
                                                          : (param == 2 ? c.y_tile_cols : c.y_rows));
                    if (field == cand[k]) continue;
                    field = cand[k];
                    const cosmos_batch_schedule keep = b->sched[ph].cfg;
                    b->sched[ph].cfg = c;
                    double ms = 0.0;
                    rc = trial(ms);
                    if (rc) break;
                    if (ms < cur_ms) {
                        cur_ms = ms;
                        best = c;
                    } else {
                        b->sched[ph].cfg = keep;
                    }
                }
                b->sched[ph].cfg = best;
            }
        }
        for (int p = 0; p < b->n; ++p) {
            b->desc[p].max_steps = steps0[p];
            b->prob[p].max_steps = steps0[p];
        }
        b->dirty = true;
        if (rc) return rc;
        if (best_ms) *best_ms = cur_ms;
        return COSMOS_OK;
    });
}

}  // extern "C"
