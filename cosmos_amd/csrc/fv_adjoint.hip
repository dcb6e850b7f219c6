// Reverse-mode (adjoint) finite-volume update on gfx950: the adjoint of one
// FORCE / SLIC sweep as one fused kernel, and a checkpointed driver for the
// adjoint of a whole trajectory of dimensionally split steps (Newtonian).
//
// Adjoint of one sweep (per line along the sweep axis; lambda = adjoint of
// the sweep output O, nu = (dc2p/dU)^T lambda at O):
//   U_c = p2c(W_c); faces (a_c, b_c) = half step of c from U_{c-1..c+1};
//   F_{c+1/2} = FORCE(b_c, a_{c+1});  O_c = c2p(U_c + k (F_{c-1/2} - F_{c+1/2}))
// so  phi_{c+1/2} = k (nu_{c+1} - nu_c)  is the adjoint of each interface flux,
// FORCE^T gives the adjoints of b_c and a_{c+1}, the half-step adjoint the
// contributions to U_{c-1..c+1}, and  gU_c = nu_c + sum of contributions,
// lambda_in_c = (dp2c/dW)^T gU_c.  Ghost cells of the transmissive boundary
// are copies of interior cells (mirror order for SLIC, edge copy for FORCE,
// Solvers/EulerSolvers.cpp:3-85), so their adjoints fold back onto those
// cells.
//
// Kernel layout: a 256-cell tile of one line per workgroup with a 3-cell
// halo (250 outputs) for x; for y, tiles of 4 neighbouring columns by 64
// rows (58 outputs per column), each thread's line neighbours 4 threads
// apart, so a wave reads 32 contiguous bytes per row; five LDS stages: U and nu; the half steps; the FORCE
// adjoints (written in place over the face states: each slot is read and
// written by one thread); the half-step adjoints; the per-cell totals and the
// ghost fold.  x lines are contiguous; y column groups are strided, with
// neighbouring groups in consecutive workgroups so that they share cache lines.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "../../include/cosmos_fv.h"
#include "fv_adjoint_physics.hpp"
#include "fv_gr_adjoint_launch.hpp"
#include "fv_kernels.hpp"

namespace cfv {
namespace {

constexpr int kAT = 256;  // tile (threads)
constexpr int kAH = 3;    // halo

struct AdjLaunch {
    ConstPlanes W, O, lam;  // sweep input, sweep output, adjoint of the output
    Planes lam_in;          // adjoint of the input
    int n, lines;           // cells per line, lines
    long pitch;
    int axis, order;
    adj::Coef C;
};

__device__ __forceinline__ int mirror_c(int c, int n) {
    if (c < 0) c = -c - 1;
    if (c >= n) c = 2 * n - 1 - c;
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}
__device__ __forceinline__ int clamp_c(int c, int n) { return c < 0 ? 0 : (c >= n ? n - 1 : c); }

__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int xcd = b % 8, k = b / 8, base = nb / 8, extra = nb % 8;
    return xcd * base + (xcd < extra ? xcd : extra) + k;
}

template <int AXIS, int ORDER, int TC>
__global__ __launch_bounds__(kAT) void k_adjoint_sweep(AdjLaunch L) {
    __shared__ double sU[5][kAT], sN[5][kAT], sA[5][kAT], sB[5][kAT], sM[5][kAT];
    constexpr int R = kAT / TC;    // positions along the line per tile
    constexpr int TW = R - 2 * kAH;
    const int tiles = (L.n + TW - 1) / TW;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int s = threadIdx.x;
    const int q = s / TC;          // position along the line within the tile
    int line, tile;
    if (AXIS == 0) {
        line = blk / tiles;
        tile = blk - line * tiles;
    } else {
        // TC neighbouring columns per tile; consecutive workgroups take
        // neighbouring column groups of one band, which share cache lines
        const int groups = (L.lines + TC - 1) / TC;
        tile = blk / groups;
        line = (blk - tile * groups) * TC + (s - q * TC);
    }
    const bool line_ok = line < L.lines;
    const long lc = line_ok ? line : L.lines - 1;
    const int c = tile * TW - kAH + q;
    const int n = L.n;
    const int cm = ORDER == 2 ? mirror_c(c, n) : clamp_c(c, n);
    const long addr = AXIS == 0 ? lc * L.pitch + cm : (long)cm * L.pitch + lc;
    const adj::Coef& C = L.C;

    // stage 1: U of every position; nu of real cells (0 for ghosts)
    double W[5], U[5], nu[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) W[k] = L.W.p[k][addr];
    adj::p2c(W, U, C);
    const bool real = c >= 0 && c < n && line_ok;
    if (real) {
        double O[5], g[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            O[k] = L.O.p[k][addr];
            g[k] = L.lam.p[k][addr];
        }
        adj::vjp_c2p(O, g, nu, C);
    } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) nu[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        sU[k][s] = U[k];
        sN[k][s] = nu[k];
    }
    __syncthreads();

    double ul[5], ur[5];
    if (q > 0 && q < R - 1) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            ul[k] = sU[k][s - TC];
            ur[k] = sU[k][s + TC];
        }
    }
    if constexpr (ORDER == 2) {
        // stage 2: the half step of each position -> (a, b)
        if (q > 0 && q < R - 1) {
            double a[5], b[5];
            adj::half_step<AXIS>(ul, U, ur, C, a, b);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                sA[k][s] = a[k];
                sB[k][s] = b[k];
            }
        }
        __syncthreads();
    }
    // stage 3: adjoint of interface (s, s+1), phi = k (nu_{s+1} - nu_s);
    // the state adjoints overwrite the slots this thread alone reads
    if (q > 0 && q < R - 2) {
        double UL[5], UR[5], phi[5], gl[5], gr[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            UL[k] = ORDER == 2 ? sB[k][s] : U[k];
            UR[k] = ORDER == 2 ? sA[k][s + TC] : ur[k];
            phi[k] = C.k * (sN[k][s + TC] - nu[k]);
        }
        adj::vjp_force<AXIS>(UL, UR, phi, C, gl, gr);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            sB[k][s] = gl[k];
            sA[k][s + TC] = gr[k];
        }
    }
    __syncthreads();
    if constexpr (ORDER == 2) {
        // stage 4: half-step adjoint of each position -> contributions to s-1, s, s+1
        if (q > 1 && q < R - 2) {
            double ga[5], gb[5], gL[5], gM[5], gR[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                ga[k] = sA[k][s];
                gb[k] = sB[k][s];
            }
            adj::vjp_half_step<AXIS>(ul, U, ur, ga, gb, C, gL, gM, gR);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                sA[k][s] = gL[k];
                sM[k][s] = gM[k];
                sB[k][s] = gR[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                sA[k][s] = 0.0;
                sM[k][s] = 0.0;
                sB[k][s] = 0.0;
            }
        }
        __syncthreads();
    }
    // stage 5: the total adjoint of each position's U (into sN, own slot)
    double tot[5];
    if (q > 0 && q < R - 1) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (ORDER == 2) tot[k] = ((nu[k] + sB[k][s - TC]) + sM[k][s]) + sA[k][s + TC];
            else tot[k] = (nu[k] + sB[k][s]) + sA[k][s];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) tot[k] = 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 5; ++k) sN[k][s] = tot[k];
    __syncthreads();
    // stage 6: outputs fold the ghost positions that copy them
    if (q >= kAH && q < R - kAH && real) {
        if (ORDER == 2) {
            if (c == 0) for (int k = 0; k < 5; ++k) tot[k] += sN[k][s - TC];
            if (c == 1) for (int k = 0; k < 5; ++k) tot[k] += sN[k][s - 3 * TC];
            if (c == n - 1) for (int k = 0; k < 5; ++k) tot[k] += sN[k][s + TC];
            if (c == n - 2) for (int k = 0; k < 5; ++k) tot[k] += sN[k][s + 3 * TC];
        } else {
            if (c == 0) for (int k = 0; k < 5; ++k) tot[k] += sN[k][s - TC];
            if (c == n - 1) for (int k = 0; k < 5; ++k) tot[k] += sN[k][s + TC];
        }
        double gW[5];
        adj::vjp_p2c(W, tot, gW, C);
#pragma unroll
        for (int k = 0; k < 5; ++k) L.lam_in.p[k][addr] = gW[k];
    }
}

// Columns per tile of the y reverse sweep (x lines are contiguous: 1).
constexpr int kAdjYCols = 4;

dim3 adj_grid(const AdjLaunch& L, int tc) {
    const int TW = kAT / tc - 2 * kAH;
    const long tiles = (L.n + TW - 1) / TW;
    return dim3((unsigned)(tiles * ((L.lines + tc - 1) / tc)));
}

hipError_t launch_adjoint_sweep(const AdjLaunch& L, hipStream_t s) {
    const dim3 b(kAT);
    if (L.axis == 0) {
        const dim3 g = adj_grid(L, 1);
        if (L.order == 2) hipLaunchKernelGGL((k_adjoint_sweep<0, 2, 1>), g, b, 0, s, L);
        else hipLaunchKernelGGL((k_adjoint_sweep<0, 1, 1>), g, b, 0, s, L);
    } else {
        const dim3 g = adj_grid(L, kAdjYCols);
        if (L.order == 2) hipLaunchKernelGGL((k_adjoint_sweep<1, 2, kAdjYCols>), g, b, 0, s, L);
        else hipLaunchKernelGGL((k_adjoint_sweep<1, 1, kAdjYCols>), g, b, 0, s, L);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_planes_to_aos(ConstPlanes in, double* aos, int nx, int ny, long pitch) {
    long total = (long)nx * ny;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (long)gridDim.x * blockDim.x) {
        long i = c / nx, j = c - i * nx;
        for (int k = 0; k < 5; ++k) aos[5 * c + k] = in.p[k][i * pitch + j];
    }
}
__global__ __launch_bounds__(256) void k_aos_to_planes(const double* aos, Planes out, int nx, int ny, long pitch) {
    long total = (long)nx * ny;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (long)gridDim.x * blockDim.x) {
        long i = c / nx, j = c - i * nx;
        for (int k = 0; k < 5; ++k) out.p[k][i * pitch + j] = aos[5 * c + k];
    }
}

}  // namespace
}  // namespace cfv

using cfv::ConstPlanes;
using cfv::Planes;

// One stored state: five planes of ny x pitch doubles.
struct AdjState {
    double* base = nullptr;
};

struct cosmos_adjoint {
    int physics = COSMOS_NEWTONIAN;
    int subcycles = 1;            // black hole: forcing sub-steps on each side of the sweeps
    double* gbuf = nullptr;       // black-hole sweeps: per-cell conserved adjoints (5 planes)
    int dim = 2, nx = 0, ny = 0;
    double gamma = 1.4;
    cosmos_solver* s = nullptr;
    long pitch = 0;
    size_t plane = 0;             // doubles per plane
    double* lam[2] = {nullptr, nullptr};  // adjoint ping-pong (5 planes each)
    double* stage = nullptr;      // AoS transfer buffer
    // trajectory
    int order = 2, limiter = 1;
    double bias = 0.0, dx = 0.0;
    std::vector<double> dts;      // recorded time steps
    int every = 20;               // checkpoint interval
    std::vector<double*> checkpoints;  // state at steps 0, every, 2 every, ...
    std::vector<double*> seg;     // per-sweep states of one segment during the reverse pass
    // black-hole forcing tapes (cosmos_adjoint_set_tape)
    int tape_mode = COSMOS_TAPE_SEGMENT;
    long long tape_reserve = 8LL << 30;
    double* tape = nullptr;            // the re-run tape
    double* carry = nullptr;           // running adjoints of the taped forcing reverse
    std::vector<double*> tapes;        // one per forcing step of a segment
    std::vector<double*> stapes;       // one per sweep of a segment: fluxes (5 planes) and face states (10)
    long fplane = 0;                   // doubles per flux plane
    // the segment whose states (and, with recorded_tapes, tapes) the segment
    // buffers hold: the forward pass leaves the last one there, so that the
    // reverse pass need not recompute it; -1: none
    int recorded_seg = -1;
    bool recorded_tapes = false;
    std::string err;
};

namespace {

thread_local std::string g_adj_create_error;

int adj_fail(cosmos_adjoint* a, int code, const std::string& msg) {
    if (a) a->err = msg;
    else g_adj_create_error = msg;
    return code;
}

#define ADJCHK(a, call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return adj_fail((a), COSMOS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int from_solver(cosmos_adjoint* a, int rc) { return rc ? adj_fail(a, rc, cosmos_last_error(a->s)) : rc; }

Planes planes_at(const cosmos_adjoint* a, double* base) {
    Planes p;
    for (int k = 0; k < 5; ++k) p.p[k] = base + k * a->plane;
    return p;
}
ConstPlanes cplanes_at(const cosmos_adjoint* a, const double* base) {
    ConstPlanes p;
    for (int k = 0; k < 5; ++k) p.p[k] = base + k * a->plane;
    return p;
}

hipStream_t stream_of(cosmos_adjoint* a) { return (hipStream_t)cosmos_stream(a->s); }

int current_base(cosmos_adjoint* a, double*& base) {
    double* planes[5];
    long pitch = 0;
    int rc = cosmos_state_planes(a->s, planes, &pitch);
    if (rc) return from_solver(a, rc);
    base = planes[0];
    return COSMOS_OK;
}

int alloc_state(cosmos_adjoint* a, double*& p) {
    ADJCHK(a, hipMalloc((void**)&p, sizeof(double) * 5 * a->plane));
    return COSMOS_OK;
}

// copy the handle's current state into dst (5 planes)
int save_state(cosmos_adjoint* a, double* dst) {
    double* cur;
    int rc = current_base(a, cur);
    if (rc) return rc;
    ADJCHK(a, hipMemcpyAsync(dst, cur, sizeof(double) * 5 * a->plane, hipMemcpyDeviceToDevice, stream_of(a)));
    return COSMOS_OK;
}

cfv::adj::Coef coef(const cosmos_adjoint* a, double dx, double dt, double bias, int limiter) {
    cfv::adj::Coef C{};
    C.k = dt / dx;
    C.kinv = dx / dt;
    C.c = 0.5 * (dt / dx);
    C.gm1 = a->gamma - 1.0;
    C.igm1 = 1.0 / (a->gamma - 1.0);
    C.bias = bias;
    C.limiter = limiter;
    return C;
}

// lam[dst] <- adjoint of the sweep from state W (planes at w_base) to O (o_base), given lam[src]
int adjoint_sweep_dev(cosmos_adjoint* a, int order, int axis, double dx, double dt, double bias, int limiter,
                      const double* w_base, const double* o_base, int src, int dst) {
    cfv::AdjLaunch L{};
    L.W = cplanes_at(a, w_base);
    L.O = cplanes_at(a, o_base);
    L.lam = cplanes_at(a, a->lam[src]);
    L.lam_in = planes_at(a, a->lam[dst]);
    L.n = axis == 0 ? a->nx : a->ny;
    L.lines = axis == 0 ? a->ny : a->nx;
    L.pitch = a->pitch;
    L.axis = axis;
    L.order = order;
    L.C = coef(a, dx, dt, bias, limiter);
    ADJCHK(a, cfv::launch_adjoint_sweep(L, stream_of(a)));
    return COSMOS_OK;
}

int lam_upload(cosmos_adjoint* a, const double* host, int which) {
    long cells = (long)a->nx * a->ny;
    hipStream_t st = stream_of(a);
    ADJCHK(a, hipMemcpyAsync(a->stage, host, sizeof(double) * 5 * cells, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(cfv::k_aos_to_planes, dim3(1024), dim3(256), 0, st, (const double*)a->stage,
                       planes_at(a, a->lam[which]), a->nx, a->ny, a->pitch);
    ADJCHK(a, hipGetLastError());
    return COSMOS_OK;
}

int lam_download(cosmos_adjoint* a, double* host, int which) {
    long cells = (long)a->nx * a->ny;
    hipStream_t st = stream_of(a);
    hipLaunchKernelGGL(cfv::k_planes_to_aos, dim3(1024), dim3(256), 0, st, cplanes_at(a, a->lam[which]), a->stage,
                       a->nx, a->ny, a->pitch);
    ADJCHK(a, hipGetLastError());
    ADJCHK(a, hipMemcpyAsync(host, a->stage, sizeof(double) * 5 * cells, hipMemcpyDeviceToHost, st));
    ADJCHK(a, hipStreamSynchronize(st));
    return COSMOS_OK;
}

// One exact forward sweep from the planes at `in` to the planes at `out`
// (the kernels and coefficients of cosmos_sweep, without its ping-pong).
int forward_sweep_dev(cosmos_adjoint* a, int axis, double dt, const double* in, double* out) {
    cfv::SweepLaunch L{};
    L.in = cplanes_at(a, in);
    L.out = planes_at(a, out);
    L.nx = a->nx;
    L.ny = a->ny;
    L.pitch = a->pitch;
    L.bias = a->bias;
    L.limiter = a->limiter;
    L.order = a->order;
    L.physics = COSMOS_NEWTONIAN;
    L.dim = a->dim;
    L.P = cfv::make_params(a->gamma);
    L.coef.dt_dx = dt / a->dx;
    L.coef.dx_dt = a->dx / dt;
    L.coef.hdt_dx = 0.5 * (dt / a->dx);
    L.y_tile = -1;  // the solvers' default y layout for this grid
    hipError_t e = axis == 0 ? cfv::launch_sweep_x(L, stream_of(a)) : cfv::launch_sweep_y(L, stream_of(a));
    ADJCHK(a, e);
    return COSMOS_OK;
}

// One operation of a step: a sweep along `axis` or (black hole) a forcing
// step, each with its own time step.
struct Op {
    int forcing;
    int axis;
    double dt;
};
constexpr int kMaxOps = 2 * 8 + 3;

// The operations of one solve-loop step with time step dt, in order:
// solve / solve2D (Solvers/EulerSecondOrderSolver.cpp:133-178) sweep x (1D) or
// X(dt * 0.5) Y(dt) X(dt * 0.5) (2D); the black-hole loops
// (Solvers/BlackHoleEulerSecondOrderSolver.cpp:145-218) wrap the sweeps in
// `subcycles` forcing steps of 0.5 * (dt / subcycles) on each side.
int step_ops(const cosmos_adjoint* a, double dt, Op* ops) {
    int n = 0;
    const bool bh = a->physics == COSMOS_BLACKHOLE;
    if (bh)
        for (int k = 0; k < a->subcycles; ++k) ops[n++] = {1, 0, 0.5 * (dt / a->subcycles)};
    if (a->dim == 1) {
        ops[n++] = {0, 0, dt};
    } else {
        ops[n++] = {0, 0, dt * 0.5};
        ops[n++] = {0, 1, dt};
        ops[n++] = {0, 0, dt * 0.5};
    }
    if (bh)
        for (int k = 0; k < a->subcycles; ++k) ops[n++] = {1, 0, 0.5 * (dt / a->subcycles)};
    return n;
}

// The operation through the public API (on the solver's own state).
int apply_op(cosmos_adjoint* a, const Op& op) {
    int rc = op.forcing ? cosmos_forcing(a->s, a->dx, op.dt, 0.0, 1)
                        : cosmos_sweep(a->s, a->order, op.axis, a->dx, op.dt, a->bias, a->limiter);
    return rc ? from_solver(a, rc) : COSMOS_OK;
}

int gr_launch_of(cosmos_adjoint* a, const Op& op, double bias, int limiter, cfv::GRLaunch& L) {
    int rc = op.forcing ? cfv::gr_forcing_launch(a->s, a->dx, op.dt, 0.0, 1, &L)
                        : cfv::gr_sweep_launch(a->s, a->order, a->dx, op.dt, bias, limiter, &L);
    if (rc) return adj_fail(a, rc, "black-hole launch set-up failed");
    return COSMOS_OK;
}

size_t tape_bytes(const cosmos_adjoint* a) { return sizeof(double) * cfv::kTapePlanes * a->plane; }
// a black-hole sweep's tape: its interface fluxes and the SLIC face states
constexpr int kSweepTapePlanes = 15;
size_t sweep_tape_bytes(const cosmos_adjoint* a) { return sizeof(double) * kSweepTapePlanes * a->fplane; }

void point_sweep_tape(const cosmos_adjoint* a, double* tape, cfv::GRLaunch& L) {
    for (int k = 0; k < 5; ++k) L.flux.p[k] = tape + k * a->fplane;
    L.half = L.order == 2 ? tape + 5 * a->fplane : nullptr;
    L.half_plane = a->fplane;
}

// The operation from the planes at `in` to the planes at `out` (the
// segment recompute of the reverse pass); a forcing step with a tape given
// writes it too.
int forward_op_dev(cosmos_adjoint* a, const Op& op, const double* in, double* out, double* tape) {
    if (a->physics != COSMOS_BLACKHOLE) return forward_sweep_dev(a, op.axis, op.dt, in, out);
    cfv::GRAdjLaunch A{};
    int rc = gr_launch_of(a, op, a->bias, a->limiter, A.L);
    if (rc) return rc;
    A.L.in = cplanes_at(a, in);
    A.L.out = planes_at(a, out);
    if (op.forcing && tape) {
        A.tape = tape;
        A.tape_plane = (long)a->plane;
        ADJCHK(a, cfv::launch_gr_forcing_tape(A, stream_of(a)));
        return COSMOS_OK;
    }
    if (!op.forcing && tape) point_sweep_tape(a, tape, A.L);
    ADJCHK(a, op.forcing ? cfv::launch_gr_forcing(A.L, stream_of(a)) : cfv::launch_gr_sweep(A.L, op.axis, stream_of(a)));
    return COSMOS_OK;
}

// lam[dst] <- the adjoint of operation op from state W (planes at w) to O
// (planes at o), given lam[src].  A black-hole sweep reads its interface
// fluxes and face states from `tape` when given (written by the segment
// recompute), else re-runs the forward sweep (rewriting O with the same
// values) for its fluxes; a forcing step reads `tape` when given, else re-runs
// its chain into the context's tape (rewriting O with the same values) or,
// with COSMOS_TAPE_NONE, inside the adjoint kernel.
int adjoint_op_dev(cosmos_adjoint* a, const Op& op, int order, double bias, int limiter, const double* w, double* o,
                   int src, int dst, double* tape = nullptr) {
    if (a->physics != COSMOS_BLACKHOLE)
        return adjoint_sweep_dev(a, order, op.axis, a->dx, op.dt, bias, limiter, w, o, src, dst);
    if (!a->gbuf) ADJCHK(a, hipMalloc((void**)&a->gbuf, sizeof(double) * 5 * a->plane));
    cfv::GRAdjLaunch A{};
    int rc = gr_launch_of(a, op, bias, limiter, A.L);
    if (rc) return rc;
    A.L.order = order;
    A.L.in = cplanes_at(a, w);
    A.L.out = planes_at(a, o);
    A.lam = cplanes_at(a, a->lam[src]);
    A.lam_in = planes_at(a, a->lam[dst]);
    A.gbuf = planes_at(a, a->gbuf);
    hipStream_t st = stream_of(a);
    if (op.forcing) {
        if (!tape && a->tape_mode != COSMOS_TAPE_NONE) {
            if (!a->tape) ADJCHK(a, hipMalloc((void**)&a->tape, tape_bytes(a)));
            tape = a->tape;
            A.tape = tape;
            A.tape_plane = (long)a->plane;
            ADJCHK(a, cfv::launch_gr_forcing_tape(A, st));
        }
        A.tape = tape;
        A.tape_plane = (long)a->plane;
        if (tape && !a->carry)
            ADJCHK(a, hipMalloc((void**)&a->carry, sizeof(double) * cfv::kCarryPlanes * a->plane));
        A.carry = a->carry;
        ADJCHK(a, hipMemsetAsync(a->lam[dst], 0, sizeof(double) * 5 * a->plane, st));
        ADJCHK(a, cfv::launch_gr_adjoint_forcing(A, st));
    } else {
        if (tape) point_sweep_tape(a, tape, A.L);
        else ADJCHK(a, cfv::launch_gr_sweep(A.L, op.axis, st));
        ADJCHK(a, cfv::launch_gr_adjoint_sweep(A, op.axis, st));
    }
    return COSMOS_OK;
}

template <class F>
int adj_guarded(cosmos_adjoint* a, F&& f) {
    if (!a) return adj_fail(nullptr, COSMOS_ERR_ARGUMENT, "null adjoint handle");
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return adj_fail(a, COSMOS_ERR_DEVICE, "host allocation failed");
    } catch (...) {
        return adj_fail(a, COSMOS_ERR_STATE, "unexpected exception");
    }
}

void free_list(std::vector<double*>& v) {
    for (double* p : v) hipFree(p);
    v.clear();
}

// The buffers of one segment: the output state of every operation (the
// segment's first input is its checkpoint) and, for a black hole in
// COSMOS_TAPE_SEGMENT mode, a tape per forcing step and per sweep where they
// leave tape_reserve bytes of device memory free (seg_tapes says whether).
int segment_buffers(cosmos_adjoint* a, bool& seg_tapes) {
    Op ops[kMaxOps];
    const int ns = step_ops(a, 1.0, ops);
    const size_t need = (size_t)a->every * ns;
    while (a->seg.size() < need) {
        double* p = nullptr;
        int rc = alloc_state(a, p);
        if (rc) return rc;
        a->seg.push_back(p);
    }
    int forcings = 0;
    for (int k = 0; k < ns; ++k) forcings += ops[k].forcing;
    const size_t nft = (size_t)a->every * forcings, nst = (size_t)a->every * (ns - forcings);
    seg_tapes = a->physics == COSMOS_BLACKHOLE && a->tape_mode == COSMOS_TAPE_SEGMENT;
    if (seg_tapes && (a->tapes.size() < nft || a->stapes.size() < nst)) {
        cfv::GRLaunch L;  // the flux layout (allocates the solver's flux planes)
        const Op sw{0, 0, 1.0};
        int rc = gr_launch_of(a, sw, a->bias, a->limiter, L);
        if (rc) return rc;
        a->fplane = (long)(a->ny + 1) * L.fpitch;
        size_t free_b = 0, total_b = 0;
        ADJCHK(a, hipMemGetInfo(&free_b, &total_b));
        const size_t more = (nft - std::min(nft, a->tapes.size())) * tape_bytes(a) +
                            (nst - std::min(nst, a->stapes.size())) * sweep_tape_bytes(a);
        if (free_b < more || free_b - more < (size_t)a->tape_reserve) {
            seg_tapes = false;
        } else {
            for (auto* v : {&a->tapes, &a->stapes}) {
                const size_t want = v == &a->tapes ? nft : nst;
                const size_t bytes = v == &a->tapes ? tape_bytes(a) : sweep_tape_bytes(a);
                while (v->size() < want) {
                    double* p = nullptr;
                    ADJCHK(a, hipMalloc((void**)&p, bytes));
                    v->push_back(p);
                }
            }
        }
    }
    return COSMOS_OK;
}

// The tape of operation k of the step_in_seg-th step of a segment (null
// without segment tapes).
double* seg_tape(const cosmos_adjoint* a, const Op* ops, int ns, int step_in_seg, int k, bool seg_tapes) {
    if (!seg_tapes) return nullptr;
    int forcings = 0, f = 0;
    for (int q = 0; q < ns; ++q) forcings += ops[q].forcing;
    for (int q = 0; q < k; ++q) f += ops[q].forcing;
    return ops[k].forcing ? a->tapes[(size_t)step_in_seg * forcings + f]
                          : a->stapes[(size_t)step_in_seg * (ns - forcings) + (k - f)];
}

}  // namespace

extern "C" {

int cosmos_adjoint_create(cosmos_adjoint** out, int physics, int dim, int nx, int ny, double adiabatic_index) {
    if (!out) return adj_fail(nullptr, COSMOS_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    if (physics != COSMOS_NEWTONIAN && physics != COSMOS_BLACKHOLE)
        return adj_fail(nullptr, COSMOS_ERR_ARGUMENT, "the adjoint supports COSMOS_NEWTONIAN and COSMOS_BLACKHOLE");
    cosmos_adjoint* a = new (std::nothrow) cosmos_adjoint();
    if (!a) return adj_fail(nullptr, COSMOS_ERR_DEVICE, "host allocation failed");
    int rc = cosmos_create(&a->s, physics, dim, nx, ny, adiabatic_index);
    if (rc) {
        std::string msg = cosmos_last_error(nullptr);
        delete a;
        return adj_fail(nullptr, rc, msg);
    }
    a->physics = physics;
    a->dim = dim;
    a->nx = nx;
    a->ny = dim == 1 ? 1 : ny;
    a->gamma = adiabatic_index;
    double* planes[5];
    rc = cosmos_state_planes(a->s, planes, &a->pitch);
    if (rc) {
        std::string msg = cosmos_last_error(a->s);
        cosmos_adjoint_destroy(a);
        return adj_fail(nullptr, rc, msg);
    }
    a->plane = (size_t)a->ny * a->pitch;
    hipError_t e = hipMalloc((void**)&a->lam[0], sizeof(double) * 5 * a->plane);
    if (e == hipSuccess) e = hipMalloc((void**)&a->lam[1], sizeof(double) * 5 * a->plane);
    if (e == hipSuccess) e = hipMalloc((void**)&a->stage, sizeof(double) * 5 * (size_t)a->nx * a->ny);
    if (e == hipSuccess) e = hipMemsetAsync(a->lam[0], 0, sizeof(double) * 5 * a->plane, stream_of(a));
    if (e == hipSuccess) e = hipMemsetAsync(a->lam[1], 0, sizeof(double) * 5 * a->plane, stream_of(a));
    if (e == hipSuccess) e = hipStreamSynchronize(stream_of(a));
    if (e != hipSuccess) {
        std::string msg = hipGetErrorString(e);
        cosmos_adjoint_destroy(a);
        return adj_fail(nullptr, COSMOS_ERR_DEVICE, "adjoint setup: " + msg);
    }
    *out = a;
    return COSMOS_OK;
}

int cosmos_adjoint_destroy(cosmos_adjoint* a) {
    if (!a) return COSMOS_OK;
    if (a->s) (void)cosmos_synchronize(a->s);
    hipFree(a->lam[0]);
    hipFree(a->lam[1]);
    hipFree(a->stage);
    hipFree(a->gbuf);
    hipFree(a->tape);
    hipFree(a->carry);
    free_list(a->tapes);
    free_list(a->stapes);
    free_list(a->checkpoints);
    free_list(a->seg);
    cosmos_destroy(a->s);
    delete a;
    return COSMOS_OK;
}

const char* cosmos_adjoint_last_error(const cosmos_adjoint* a) { return a ? a->err.c_str() : g_adj_create_error.c_str(); }

cosmos_solver* cosmos_adjoint_solver(cosmos_adjoint* a) { return a ? a->s : nullptr; }

int cosmos_adjoint_sweep(cosmos_adjoint* a, int order, int axis, double cell_spacing, double time_step, double bias,
                         int slope_limiter, const double* lambda_out, double* lambda_in) {
    return adj_guarded(a, [&] {
        if (!lambda_out || !lambda_in) return adj_fail(a, COSMOS_ERR_ARGUMENT, "null adjoint array");
        double *w, *o;
        int rc = current_base(a, w);
        if (rc) return rc;
        rc = cosmos_sweep(a->s, order, axis, cell_spacing, time_step, bias, slope_limiter);
        if (rc) return from_solver(a, rc);
        if ((rc = current_base(a, o))) return rc;
        if ((rc = lam_upload(a, lambda_out, 0))) return rc;
        a->dx = cell_spacing;
        a->order = order;
        const Op op{0, axis, time_step};
        if ((rc = adjoint_op_dev(a, op, order, bias, slope_limiter, w, o, 0, 1))) return rc;
        return lam_download(a, lambda_in, 1);
    });
}

int cosmos_adjoint_forcing(cosmos_adjoint* a, double cell_spacing, double time_step, const double* lambda_out,
                           double* lambda_in) {
    return adj_guarded(a, [&] {
        if (a->physics != COSMOS_BLACKHOLE)
            return adj_fail(a, COSMOS_ERR_ARGUMENT, "forcing needs a COSMOS_BLACKHOLE adjoint context");
        if (!lambda_out || !lambda_in) return adj_fail(a, COSMOS_ERR_ARGUMENT, "null adjoint array");
        double *w, *o;
        int rc = current_base(a, w);
        if (rc) return rc;
        rc = cosmos_forcing(a->s, cell_spacing, time_step, 0.0, 1);
        if (rc) return from_solver(a, rc);
        if ((rc = current_base(a, o))) return rc;
        if ((rc = lam_upload(a, lambda_out, 0))) return rc;
        a->dx = cell_spacing;
        const Op op{1, 0, time_step};
        if ((rc = adjoint_op_dev(a, op, a->order, 0.0, 1, w, o, 0, 1))) return rc;
        return lam_download(a, lambda_in, 1);
    });
}

int cosmos_adjoint_set_subcycles(cosmos_adjoint* a, int subcycles) {
    return adj_guarded(a, [&] {
        if (subcycles < 0 || subcycles > 8) return adj_fail(a, COSMOS_ERR_ARGUMENT, "subcycles must be in [0, 8]");
        a->subcycles = subcycles;
        a->dts.clear();
        return COSMOS_OK;
    });
}

int cosmos_adjoint_release(cosmos_adjoint* a) {
    return adj_guarded(a, [&] {
        ADJCHK(a, hipStreamSynchronize(stream_of(a)));
        free_list(a->seg);
        free_list(a->tapes);
        free_list(a->stapes);
        hipFree(a->tape);
        hipFree(a->carry);
        a->tape = a->carry = nullptr;
        a->recorded_seg = -1;
        return COSMOS_OK;
    });
}

int cosmos_adjoint_set_tape(cosmos_adjoint* a, int mode, long long reserve_bytes) {
    return adj_guarded(a, [&] {
        if (mode < COSMOS_TAPE_NONE || mode > COSMOS_TAPE_SEGMENT || reserve_bytes < 0)
            return adj_fail(a, COSMOS_ERR_ARGUMENT, "tape mode must be 0, 1 or 2 and the reserve >= 0");
        a->tape_mode = mode;
        a->tape_reserve = reserve_bytes;
        if (mode != COSMOS_TAPE_SEGMENT) {
            free_list(a->tapes);
            free_list(a->stapes);
            if (a->recorded_tapes) a->recorded_seg = -1;
        }
        return COSMOS_OK;
    });
}

int cosmos_adjoint_forward(cosmos_adjoint* a, int order, double cell_spacing, double cfl, int n_steps, double bias,
                           int slope_limiter, int checkpoint_every) {
    return adj_guarded(a, [&] {
        if (n_steps < 1 || checkpoint_every < 1)
            return adj_fail(a, COSMOS_ERR_ARGUMENT, "n_steps and checkpoint_every must be >= 1");
        a->order = order;
        a->dx = cell_spacing;
        a->bias = bias;
        a->limiter = slope_limiter;
        a->every = checkpoint_every;
        a->dts.clear();
        free_list(a->checkpoints);
        a->recorded_seg = -1;
        // the reverse pass recomputes segments with the exact kernels
        int rc0 = cosmos_set_mode(a->s, COSMOS_MODE_EXACT);
        if (rc0) return from_solver(a, rc0);
        // the last segment runs through the segment buffers (and tapes), where
        // the reverse pass finds it without recomputing it
        bool seg_tapes = false;
        if ((rc0 = segment_buffers(a, seg_tapes))) return rc0;
        const int last = (n_steps - 1) / a->every;
        Op ops[kMaxOps];
        for (int it = 0; it < n_steps; ++it) {
            if (it % a->every == 0) {
                double* cp = nullptr;
                int rc = alloc_state(a, cp);
                if (rc) return rc;
                a->checkpoints.push_back(cp);
                if ((rc = save_state(a, cp))) return rc;
            }
            double dt = 0.0;
            int rc = cosmos_stable_time_step(a->s, cell_spacing, cfl, 0.0, INFINITY, it, &dt);
            if (rc) return from_solver(a, rc);
            a->dts.push_back(dt);
            const int ns = step_ops(a, dt, ops);
            if (it / a->every != last) {
                for (int k = 0; k < ns; ++k)
                    if ((rc = apply_op(a, ops[k]))) return rc;
                continue;
            }
            const int sis = it - last * a->every;
            auto state = [&](int q) -> double* { return q == 0 ? a->checkpoints[last] : a->seg[q - 1]; };
            for (int k = 0; k < ns; ++k) {
                const int q = sis * ns + k;
                if ((rc = forward_op_dev(a, ops[k], state(q), state(q + 1), seg_tape(a, ops, ns, sis, k, seg_tapes))))
                    return rc;
            }
            // the step's output becomes the solver's state (the next dt reads it)
            double* cur = nullptr;
            if ((rc = current_base(a, cur))) return rc;
            ADJCHK(a, hipMemcpyAsync(cur, state((sis + 1) * ns), sizeof(double) * 5 * a->plane,
                                     hipMemcpyDeviceToDevice, stream_of(a)));
        }
        ADJCHK(a, hipStreamSynchronize(stream_of(a)));
        a->recorded_seg = last;
        a->recorded_tapes = seg_tapes;
        return COSMOS_OK;
    });
}

int cosmos_adjoint_replay(cosmos_adjoint* a, const double* prim_in, double* prim_out) {
    return adj_guarded(a, [&] {
        if (a->dts.empty()) return adj_fail(a, COSMOS_ERR_STATE, "no recorded trajectory (cosmos_adjoint_forward)");
        int rc = cosmos_upload(a->s, prim_in);
        if (rc) return from_solver(a, rc);
        Op ops[kMaxOps];
        for (double dt : a->dts) {
            const int ns = step_ops(a, dt, ops);
            for (int k = 0; k < ns; ++k)
                if ((rc = apply_op(a, ops[k]))) return rc;
        }
        rc = cosmos_download(a->s, prim_out);
        return from_solver(a, rc);
    });
}

int cosmos_adjoint_time_steps(cosmos_adjoint* a, double* out, int capacity, int* count) {
    return adj_guarded(a, [&] {
        if (!count) return adj_fail(a, COSMOS_ERR_ARGUMENT, "null count");
        *count = (int)a->dts.size();
        if (out)
            for (int k = 0; k < capacity && k < (int)a->dts.size(); ++k) out[k] = a->dts[k];
        return COSMOS_OK;
    });
}

int cosmos_adjoint_backward(cosmos_adjoint* a, const double* grad_out, double* grad_in) {
    return adj_guarded(a, [&] {
        if (a->dts.empty()) return adj_fail(a, COSMOS_ERR_STATE, "no recorded trajectory (cosmos_adjoint_forward)");
        if (!grad_out || !grad_in) return adj_fail(a, COSMOS_ERR_ARGUMENT, "null gradient array");
        Op ops[kMaxOps];
        const int ns = step_ops(a, 1.0, ops);  // operations per step (the count does not depend on dt)
        const int nsteps = (int)a->dts.size();
        bool seg_tapes = false;
        int rc = segment_buffers(a, seg_tapes);
        if (rc) return rc;
        auto tape_of = [&](int step_in_seg, int k) { return seg_tape(a, ops, ns, step_in_seg, k, seg_tapes); };
        if ((rc = lam_upload(a, grad_out, 0))) return rc;
        int cur = 0;
        for (int sgi = (int)a->checkpoints.size() - 1; sgi >= 0; --sgi) {
            const int s0 = sgi * a->every, s1 = std::min(nsteps, s0 + a->every);
            // recompute the segment from its checkpoint (state q + 1 = operation q
            // of state q), unless the forward pass left it in the buffers
            auto state = [&](int q) -> double* { return q == 0 ? a->checkpoints[sgi] : a->seg[q - 1]; };
            const bool kept = sgi == a->recorded_seg && a->recorded_tapes == seg_tapes;
            int q = 0;
            for (int it = s0; it < s1 && !kept; ++it) {
                step_ops(a, a->dts[it], ops);
                for (int k = 0; k < ns; ++k, ++q)
                    if ((rc = forward_op_dev(a, ops[k], state(q), state(q + 1), tape_of(it - s0, k)))) return rc;
            }
            if (!kept) {  // the buffers now hold segment sgi
                a->recorded_seg = sgi;
                a->recorded_tapes = seg_tapes;
            }
            // reverse through the segment
            for (int it = s1 - 1; it >= s0; --it) {
                step_ops(a, a->dts[it], ops);
                for (int k = ns - 1; k >= 0; --k) {
                    const int qi = (it - s0) * ns + k;
                    rc = adjoint_op_dev(a, ops[k], a->order, a->bias, a->limiter, state(qi), state(qi + 1), cur,
                                        cur ^ 1, tape_of(it - s0, k));
                    if (rc) return rc;
                    cur ^= 1;
                }
            }
        }
        return lam_download(a, grad_in, cur);
    });
}

}  // extern "C"
