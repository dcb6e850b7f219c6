// Two-level adaptive mesh refinement: the refinement flags, masked sweeps,
// restriction / injection and derefinement of EulerAMR::solve{,2D}Level1AMR
// (AMR/EulerAMR.cpp:169-590) and the special-relativistic copies
// (AMR/RelativisticEulerAMR.cpp), on gfx950.
//
// A cosmos_amr context holds two solver handles of the public C-ABI (the
// coarse nx x ny grid and the fine 2nx x 2ny grid) and the two flag grids as
// device bytes (row-major, no pitch).  Masked sweeps are the ordinary sweep
// kernels followed by a merge that restores the cells whose flag is clear:
// the reference's masked sweep reads the same boundary copy as its full
// sweep, so the two agree on every updated cell.  Flag kernels run one
// thread per coarse cell; each thread writes only its own coarse flag, its
// own children's fine flags and its own coarse state, so no launch races.
#include <climits>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cosmos_fv.h"
#include "fv_kernels.hpp"

namespace cfv {
namespace {

constexpr int kAmrThreads = 256;
typedef unsigned char u8;

struct Grid {
    double* p[5];
    long pitch;
    int nx, ny;
};

__device__ __forceinline__ int clampc(int c, int n) { return c < 0 ? 0 : (c >= n ? n - 1 : c); }

__device__ __forceinline__ void load(const Grid& g, int i, int j, double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) W[k] = g.p[k][(long)i * g.pitch + j];
}

__device__ __forceinline__ void store(const Grid& g, int i, int j, const double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) g.p[k][(long)i * g.pitch + j] = W[k];
}

// |(b - a) / dx| >= tol for any component, the test of the reference's flag
// loops (AMR/EulerAMR.cpp:215-216, :412-415).
__device__ __forceinline__ bool jump(const double* a, const double* b, double dx, double tol) {
    bool hit = false;
#pragma unroll
    for (int k = 0; k < 5; ++k) hit |= __builtin_fabs((b[k] - a[k]) / dx) >= tol;
    return hit;
}

template <class Phys>
__device__ __forceinline__ void cons_at(const Grid& g, int i, int j, const PhysParams& P, double* U) {
    double W[5];
    load(g, i, j, W);
    Phys::p2c(W, U, P);
}

// Mean of the children's conserved vectors, back to primitives: 1D 0.5 (L + R),
// 2D 0.25 (((TL + TR) + BL) + BR) (AMR/EulerAMR.cpp:290-296, :500-509).
template <class Phys>
__device__ __forceinline__ void restrict_children(const Grid& f, int dim, int i, int j, const PhysParams& P,
                                                  double* W) {
    double a[5], b[5], m[5];
    if (dim == 1) {
        cons_at<Phys>(f, 0, 2 * j, P, a);
        cons_at<Phys>(f, 0, 2 * j + 1, P, b);
#pragma unroll
        for (int k = 0; k < 5; ++k) m[k] = 0.5 * (a[k] + b[k]);
    } else {
        double c[5], d[5];
        cons_at<Phys>(f, 2 * i, 2 * j, P, a);
        cons_at<Phys>(f, 2 * i, 2 * j + 1, P, b);
        cons_at<Phys>(f, 2 * i + 1, 2 * j, P, c);
        cons_at<Phys>(f, 2 * i + 1, 2 * j + 1, P, d);
#pragma unroll
        for (int k = 0; k < 5; ++k) m[k] = 0.25 * (((a[k] + b[k]) + c[k]) + d[k]);
    }
    Phys::c2p(m, W, Phys::inv_density(m), P);
}

// A1: sticky refinement flags from the coarse state, 1-cell copy boundary.
// diag: the 2D "bottom" neighbour is the diagonal cell of the boundary-padded
// grid (AMR/RelativisticEulerAMR.cpp:422), whose bottom-right corner is a
// default state (1, 0, 0, 0, 1).
template <class Phys>
__global__ __launch_bounds__(kAmrThreads) void k_amr_refine(Grid c, int dim, double dx, double tol, int diag,
                                                            PhysParams P, u8* cflags, u8* fflags) {
    long idx = (long)blockIdx.x * kAmrThreads + threadIdx.x;
    if (idx >= (long)c.nx * c.ny) return;
    int i = (int)(idx / c.nx), j = (int)(idx - (long)i * c.nx);
    double um[5], un[5];
    cons_at<Phys>(c, i, j, P, um);
    cons_at<Phys>(c, i, clampc(j - 1, c.nx), P, un);
    bool hit = jump(un, um, dx, tol);
    cons_at<Phys>(c, i, clampc(j + 1, c.nx), P, un);
    hit |= jump(um, un, dx, tol);
    if (dim == 2) {
        cons_at<Phys>(c, clampc(i - 1, c.ny), j, P, un);
        hit |= jump(un, um, dx, tol);
        if (diag && i + 1 >= c.ny && j + 1 >= c.nx) {
            const double Wd[5] = {1.0, 0.0, 0.0, 0.0, 1.0};
            Phys::p2c(Wd, un, P);
        } else {
            cons_at<Phys>(c, clampc(i + 1, c.ny), diag ? clampc(j + 1, c.nx) : j, P, un);
        }
        hit |= jump(um, un, dx, tol);
    }
    if (hit) {
        const int fnx = 2 * c.nx;
        cflags[idx] = 0;
        if (dim == 1) {
            fflags[2 * j] = 1;
            fflags[2 * j + 1] = 1;
        } else {
            fflags[(long)(2 * i) * fnx + 2 * j] = 1;
            fflags[(long)(2 * i) * fnx + 2 * j + 1] = 1;
            fflags[(long)(2 * i + 1) * fnx + 2 * j] = 1;
            fflags[(long)(2 * i + 1) * fnx + 2 * j + 1] = 1;
        }
    }
}

// A5 after the fine sub-steps: refined coarse cells take the restriction of
// their children; children whose flag is clear take a copy of the coarse cell
// (AMR/EulerAMR.cpp:282-308, :494-536).
template <class Phys>
__global__ __launch_bounds__(kAmrThreads) void k_amr_restrict(Grid c, Grid f, int dim, PhysParams P,
                                                              const u8* cflags, const u8* fflags) {
    long idx = (long)blockIdx.x * kAmrThreads + threadIdx.x;
    if (idx >= (long)c.nx * c.ny) return;
    int i = (int)(idx / c.nx), j = (int)(idx - (long)i * c.nx);
    double W[5];
    if (!cflags[idx]) {
        restrict_children<Phys>(f, dim, i, j, P, W);
        store(c, i, j, W);
    } else {
        load(c, i, j, W);
    }
    const int rows = dim == 1 ? 1 : 2;
    for (int a = 0; a < rows; ++a)
        for (int b = 0; b < 2; ++b) {
            int fi = dim == 1 ? 0 : 2 * i + a, fj = 2 * j + b;
            if (!fflags[(long)fi * f.nx + fj]) store(f, fi, fj, W);
        }
}

// A2: derefinement of quiet refined cells, tested on fine cell 2j+1
// ((2i+1, 2j+1) in 2D) against its fine neighbours (1-cell copy boundary),
// over the coarse spacing at tol / 2 (AMR/EulerAMR.cpp:312-345, :540-586).
template <class Phys>
__global__ __launch_bounds__(kAmrThreads) void k_amr_derefine(Grid c, Grid f, int dim, double dx, double tol,
                                                              PhysParams P, u8* cflags, u8* fflags) {
    long idx = (long)blockIdx.x * kAmrThreads + threadIdx.x;
    if (idx >= (long)c.nx * c.ny) return;
    int i = (int)(idx / c.nx), j = (int)(idx - (long)i * c.nx);
    const double t2 = tol * 0.5;
    double um[5], un[5];
    bool marked;
    long first;
    if (dim == 1) {
        cons_at<Phys>(f, 0, 2 * j + 1, P, um);
        cons_at<Phys>(f, 0, 2 * j, P, un);
        marked = jump(un, um, dx, t2);
        cons_at<Phys>(f, 0, clampc(2 * j + 2, f.nx), P, un);
        marked |= jump(um, un, dx, t2);
        first = 2 * j;
    } else {
        const int fi = 2 * i + 1, fj = 2 * j + 1;
        cons_at<Phys>(f, fi, fj, P, um);
        cons_at<Phys>(f, fi - 1, fj, P, un);
        marked = jump(un, um, dx, t2);
        cons_at<Phys>(f, clampc(fi + 1, f.ny), fj, P, un);
        marked |= jump(um, un, dx, t2);
        cons_at<Phys>(f, fi, fj - 1, P, un);
        marked |= jump(un, um, dx, t2);
        cons_at<Phys>(f, fi, clampc(fj + 1, f.nx), P, un);
        marked |= jump(um, un, dx, t2);
        first = (long)(2 * i) * f.nx + 2 * j;
    }
    if (!marked && fflags[first]) {
        if (dim == 1) {
            fflags[2 * j] = 0;
            fflags[2 * j + 1] = 0;
        } else {
            fflags[first] = 0;
            fflags[first + 1] = 0;
            fflags[first + f.nx] = 0;
            fflags[first + f.nx + 1] = 0;
        }
        cflags[idx] = 1;
        double W[5];
        restrict_children<Phys>(f, dim, i, j, P, W);
        store(c, i, j, W);
    }
}

// A4: after a full sweep from `before` into `after`, cells whose flag is clear
// keep their previous state.
__global__ __launch_bounds__(kAmrThreads) void k_mask_merge(Grid before, Grid after, const u8* mask) {
    long idx = (long)blockIdx.x * kAmrThreads + threadIdx.x;
    if (idx >= (long)after.nx * after.ny) return;
    if (mask[idx]) return;
    int i = (int)(idx / after.nx), j = (int)(idx - (long)i * after.nx);
    double W[5];
    load(before, i, j, W);
    store(after, i, j, W);
}

__global__ __launch_bounds__(kAmrThreads) void k_fill(u8* a, long n, u8 v) {
    long idx = (long)blockIdx.x * kAmrThreads + threadIdx.x;
    if (idx < n) a[idx] = v;
}

// flag[0] = 1 if any byte of a is nonzero (the caller zeroes flag first).
__global__ __launch_bounds__(kAmrThreads) void k_any(const u8* a, long n, int* flag) {
    long idx = (long)blockIdx.x * kAmrThreads + threadIdx.x;
    bool v = idx < n && a[idx];
    if (__any(v) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// Initial fine state: every child a copy of its coarse cell (the loops before
// the time loop, AMR/EulerAMR.cpp:181-200).
__global__ __launch_bounds__(kAmrThreads) void k_inject_all(Grid c, Grid f, int dim) {
    long idx = (long)blockIdx.x * kAmrThreads + threadIdx.x;
    if (idx >= (long)c.nx * c.ny) return;
    int i = (int)(idx / c.nx), j = (int)(idx - (long)i * c.nx);
    double W[5];
    load(c, i, j, W);
    const int rows = dim == 1 ? 1 : 2;
    for (int a = 0; a < rows; ++a)
        for (int b = 0; b < 2; ++b) store(f, dim == 1 ? 0 : 2 * i + a, 2 * j + b, W);
}

dim3 blocks_of(long n) { return dim3((unsigned)((n + kAmrThreads - 1) / kAmrThreads)); }

}  // namespace
}  // namespace cfv

using cfv::Grid;

struct cosmos_amr {
    int physics = 0, dim = 2, nx = 0, ny = 0;
    cosmos_solver* coarse = nullptr;
    cosmos_solver* fine = nullptr;
    unsigned char* cflags = nullptr;  // device, [ny][nx]
    unsigned char* fflags = nullptr;  // device, [2ny][2nx] (1D: [2nx])
    int* any = nullptr;               // device scratch flag
    cfv::PhysParams P{};
    std::string err;
};

namespace {

thread_local std::string g_amr_create_error;

int amr_fail(cosmos_amr* a, int code, const std::string& msg) {
    if (a) a->err = msg;
    else g_amr_create_error = msg;
    return code;
}

#define AMRCHK(a, call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return amr_fail((a), COSMOS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// forward a solver failure with its message
int from_solver(cosmos_amr* a, cosmos_solver* s, int rc) {
    if (rc == COSMOS_OK) return rc;
    return amr_fail(a, rc, cosmos_last_error(s));
}

int grid_of(cosmos_amr* a, cosmos_solver* s, Grid& g) {
    double* planes[5];
    long pitch = 0;
    int rc = cosmos_state_planes(s, planes, &pitch);
    if (rc) return from_solver(a, s, rc);
    for (int k = 0; k < 5; ++k) g.p[k] = planes[k];
    g.pitch = pitch;
    g.nx = s == a->coarse ? a->nx : 2 * a->nx;
    g.ny = s == a->coarse ? a->ny : (a->dim == 1 ? 1 : 2 * a->ny);
    return COSMOS_OK;
}

long coarse_cells(const cosmos_amr* a) { return (long)a->nx * a->ny; }
long fine_cells(const cosmos_amr* a) { return (long)(2 * a->nx) * (a->dim == 1 ? 1 : 2 * a->ny); }

// Kernels that touch both levels (or the flags) run on the coarse handle's
// stream after both handles' queued work has finished; the solver calls of
// either level likewise start only after that stream is idle, so every
// access to a level's state or flags is ordered.
hipStream_t order_streams(cosmos_amr* a) {
    cosmos_synchronize(a->fine);
    cosmos_synchronize(a->coarse);
    return (hipStream_t)cosmos_stream(a->coarse);
}

template <class F>
int amr_guarded(cosmos_amr* a, F&& f) {
    if (!a) return amr_fail(nullptr, COSMOS_ERR_ARGUMENT, "null AMR handle");
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return amr_fail(a, COSMOS_ERR_DEVICE, "host allocation failed");
    } catch (...) {
        return amr_fail(a, COSMOS_ERR_STATE, "unexpected exception");
    }
}

template <template <class> class K, class... Args>
hipError_t launch_physics(int physics, dim3 g, hipStream_t s, Args... args) {
    if (physics == COSMOS_NEWTONIAN) hipLaunchKernelGGL(K<cfv::Newtonian>::fn, g, dim3(cfv::kAmrThreads), 0, s, args...);
    else hipLaunchKernelGGL(K<cfv::Relativistic>::fn, g, dim3(cfv::kAmrThreads), 0, s, args...);
    return hipGetLastError();
}

template <class Phys>
struct RefineK {
    static constexpr auto fn = cfv::k_amr_refine<Phys>;
};
template <class Phys>
struct RestrictK {
    static constexpr auto fn = cfv::k_amr_restrict<Phys>;
};
template <class Phys>
struct DerefineK {
    static constexpr auto fn = cfv::k_amr_derefine<Phys>;
};

int amr_refine(cosmos_amr* a, double dx, double tol) {
    Grid c;
    int rc = grid_of(a, a->coarse, c);
    if (rc) return rc;
    hipStream_t s = order_streams(a);
    AMRCHK(a, (launch_physics<RefineK>(a->physics, cfv::blocks_of(coarse_cells(a)), s, c, a->dim, dx, tol,
                                       a->physics == COSMOS_RELATIVISTIC ? 1 : 0, a->P, a->cflags, a->fflags)));
    return COSMOS_OK;
}

int amr_restrict(cosmos_amr* a) {
    Grid c, f;
    int rc = grid_of(a, a->coarse, c);
    if (!rc) rc = grid_of(a, a->fine, f);
    if (rc) return rc;
    hipStream_t s = order_streams(a);
    AMRCHK(a, (launch_physics<RestrictK>(a->physics, cfv::blocks_of(coarse_cells(a)), s, c, f, a->dim, a->P,
                                         (const unsigned char*)a->cflags, (const unsigned char*)a->fflags)));
    return COSMOS_OK;
}

int amr_derefine(cosmos_amr* a, double dx, double tol) {
    Grid c, f;
    int rc = grid_of(a, a->coarse, c);
    if (!rc) rc = grid_of(a, a->fine, f);
    if (rc) return rc;
    hipStream_t s = order_streams(a);
    AMRCHK(a, (launch_physics<DerefineK>(a->physics, cfv::blocks_of(coarse_cells(a)), s, c, f, a->dim, dx, tol,
                                         a->P, a->cflags, a->fflags)));
    return COSMOS_OK;
}

// One sweep of a level restricted to that level's flags.
int amr_sweep(cosmos_amr* a, int level, int order, int axis, double dx, double dt, double bias, int lim) {
    cosmos_solver* s = level == 0 ? a->coarse : a->fine;
    order_streams(a);
    Grid before, after;
    int rc = grid_of(a, s, before);
    if (rc) return rc;
    rc = cosmos_sweep(s, order, axis, dx, dt, bias, lim);
    if (rc) return from_solver(a, s, rc);
    rc = grid_of(a, s, after);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)cosmos_stream(s);
    long n = (long)after.nx * after.ny;
    hipLaunchKernelGGL(cfv::k_mask_merge, cfv::blocks_of(n), dim3(cfv::kAmrThreads), 0, st, before, after,
                       (const unsigned char*)(level == 0 ? a->cflags : a->fflags));
    AMRCHK(a, hipGetLastError());
    return COSMOS_OK;
}

// X(dt/2) Y(dt) X(dt/2) in 2D, one sweep in 1D, bias 0 and van Leer as the
// reference's AMR drivers hard-code (AMR/EulerAMR.cpp:231-238, :432-444).
int amr_flux_step(cosmos_amr* a, int level, int order, double dx, double dt) {
    if (a->dim == 1) return amr_sweep(a, level, order, 0, dx, dt, 0.0, 1);
    int rc = amr_sweep(a, level, order, 0, dx, dt * 0.5, 0.0, 1);
    if (!rc) rc = amr_sweep(a, level, order, 1, dx, dt, 0.0, 1);
    if (!rc) rc = amr_sweep(a, level, order, 0, dx, dt * 0.5, 0.0, 1);
    return rc;
}

int amr_any_refined(cosmos_amr* a, bool& out) {
    hipStream_t s = order_streams(a);
    int zero = 0, v = 0;
    AMRCHK(a, hipMemcpyAsync(a->any, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(cfv::k_any, cfv::blocks_of(fine_cells(a)), dim3(cfv::kAmrThreads), 0, s,
                       (const unsigned char*)a->fflags, fine_cells(a), a->any);
    AMRCHK(a, hipGetLastError());
    AMRCHK(a, hipMemcpyAsync(&v, a->any, sizeof(int), hipMemcpyDeviceToHost, s));
    AMRCHK(a, hipStreamSynchronize(s));
    out = v != 0;
    return COSMOS_OK;
}

}  // namespace

extern "C" {

int cosmos_amr_create(cosmos_amr** out, int physics, int dim, int nx, int ny, double adiabatic_index) {
    if (!out) return amr_fail(nullptr, COSMOS_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    if (physics != COSMOS_NEWTONIAN && physics != COSMOS_RELATIVISTIC)
        return amr_fail(nullptr, COSMOS_ERR_ARGUMENT, "AMR physics must be COSMOS_NEWTONIAN or COSMOS_RELATIVISTIC");
    if (dim != 1 && dim != 2) return amr_fail(nullptr, COSMOS_ERR_ARGUMENT, "dim must be 1 or 2");
    if (dim == 1) ny = 1;
    if (nx < 2 || ny < 1 || (dim == 2 && ny < 2)) return amr_fail(nullptr, COSMOS_ERR_ARGUMENT, "grid too small");
    if ((long)nx * ny > (1L << 29)) return amr_fail(nullptr, COSMOS_ERR_ARGUMENT, "grid too large");
    cosmos_amr* a = new (std::nothrow) cosmos_amr();
    if (!a) return amr_fail(nullptr, COSMOS_ERR_DEVICE, "host allocation failed");
    a->physics = physics;
    a->dim = dim;
    a->nx = nx;
    a->ny = ny;
    a->P = cfv::make_params(adiabatic_index);
    int rc = cosmos_create(&a->coarse, physics, dim, nx, ny, adiabatic_index);
    if (!rc) rc = cosmos_create(&a->fine, physics, dim, 2 * nx, dim == 1 ? 1 : 2 * ny, adiabatic_index);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMalloc((void**)&a->cflags, (size_t)coarse_cells(a));
    if (!rc && e == hipSuccess) e = hipMalloc((void**)&a->fflags, (size_t)fine_cells(a));
    if (!rc && e == hipSuccess) e = hipMalloc((void**)&a->any, sizeof(int));
    if (rc || e != hipSuccess) {
        std::string msg = rc ? std::string(cosmos_last_error(nullptr)) : std::string(hipGetErrorString(e));
        cosmos_amr_destroy(a);
        return amr_fail(nullptr, rc ? rc : COSMOS_ERR_DEVICE, "AMR setup: " + msg);
    }
    *out = a;
    return COSMOS_OK;
}

int cosmos_amr_destroy(cosmos_amr* a) {
    if (!a) return COSMOS_OK;
    if (a->coarse) cosmos_synchronize(a->coarse);
    if (a->fine) cosmos_synchronize(a->fine);
    hipFree(a->cflags);
    hipFree(a->fflags);
    hipFree(a->any);
    cosmos_destroy(a->coarse);
    cosmos_destroy(a->fine);
    delete a;
    return COSMOS_OK;
}

const char* cosmos_amr_last_error(const cosmos_amr* a) { return a ? a->err.c_str() : g_amr_create_error.c_str(); }

int cosmos_amr_set_con2prim(cosmos_amr* a, double tolerance, int max_iterations) {
    return amr_guarded(a, [&] {
        int rc = cosmos_set_con2prim(a->coarse, tolerance, max_iterations);
        if (rc) return from_solver(a, a->coarse, rc);
        rc = cosmos_set_con2prim(a->fine, tolerance, max_iterations);
        if (rc) return from_solver(a, a->fine, rc);
        a->P.c2p_tol = tolerance;
        a->P.c2p_maxit = max_iterations;
        return COSMOS_OK;
    });
}

int cosmos_amr_upload(cosmos_amr* a, const double* prim) {
    return amr_guarded(a, [&] {
        int rc = cosmos_upload(a->coarse, prim);
        if (rc) return from_solver(a, a->coarse, rc);
        Grid c, f;
        rc = grid_of(a, a->coarse, c);
        if (!rc) rc = grid_of(a, a->fine, f);
        if (rc) return rc;
        hipStream_t s = order_streams(a);
        hipLaunchKernelGGL(cfv::k_inject_all, cfv::blocks_of(coarse_cells(a)), dim3(cfv::kAmrThreads), 0, s, c, f,
                           a->dim);
        hipLaunchKernelGGL(cfv::k_fill, cfv::blocks_of(coarse_cells(a)), dim3(cfv::kAmrThreads), 0, s, a->cflags,
                           coarse_cells(a), (unsigned char)1);
        hipLaunchKernelGGL(cfv::k_fill, cfv::blocks_of(fine_cells(a)), dim3(cfv::kAmrThreads), 0, s, a->fflags,
                           fine_cells(a), (unsigned char)0);
        AMRCHK(a, hipGetLastError());
        AMRCHK(a, hipStreamSynchronize(s));
        return COSMOS_OK;
    });
}

int cosmos_amr_download(cosmos_amr* a, double* coarse, double* fine, unsigned char* coarse_flags,
                        unsigned char* fine_flags) {
    return amr_guarded(a, [&] {
        order_streams(a);
        int rc = coarse ? cosmos_download(a->coarse, coarse) : COSMOS_OK;
        if (rc) return from_solver(a, a->coarse, rc);
        rc = fine ? cosmos_download(a->fine, fine) : COSMOS_OK;
        if (rc) return from_solver(a, a->fine, rc);
        hipStream_t s = order_streams(a);
        if (coarse_flags)
            AMRCHK(a, hipMemcpyAsync(coarse_flags, a->cflags, (size_t)coarse_cells(a), hipMemcpyDeviceToHost, s));
        if (fine_flags)
            AMRCHK(a, hipMemcpyAsync(fine_flags, a->fflags, (size_t)fine_cells(a), hipMemcpyDeviceToHost, s));
        AMRCHK(a, hipStreamSynchronize(s));
        return COSMOS_OK;
    });
}

int cosmos_amr_refine_flags(cosmos_amr* a, double cell_spacing, double tolerance) {
    return amr_guarded(a, [&] { return amr_refine(a, cell_spacing, tolerance); });
}

int cosmos_amr_sweep(cosmos_amr* a, int level, int order, int axis, double cell_spacing, double time_step,
                     double bias, int slope_limiter) {
    return amr_guarded(a, [&] {
        if (level != 0 && level != 1) return amr_fail(a, COSMOS_ERR_ARGUMENT, "level must be 0 (coarse) or 1 (fine)");
        return amr_sweep(a, level, order, axis, cell_spacing, time_step, bias, slope_limiter);
    });
}

int cosmos_amr_restrict(cosmos_amr* a) {
    return amr_guarded(a, [&] { return amr_restrict(a); });
}

int cosmos_amr_derefine(cosmos_amr* a, double cell_spacing, double tolerance) {
    return amr_guarded(a, [&] { return amr_derefine(a, cell_spacing, tolerance); });
}

int cosmos_amr_solve_level1(cosmos_amr* a, int order, double cell_spacing, double cfl, double final_time,
                            double tolerance, int max_coarse_steps, double* time_out, int* steps_out) {
    return amr_guarded(a, [&] {
        if (order != 1 && order != 2) return amr_fail(a, COSMOS_ERR_ARGUMENT, "order must be 1 or 2");
        const int cap = max_coarse_steps < 0 ? INT_MAX : max_coarse_steps;
        double t = 0.0;
        int it = 0, fine_total = 0, rc = COSMOS_OK;
        while (t < final_time && it < cap) {
            double cdt = 0.0;
            rc = cosmos_stable_time_step(a->coarse, cell_spacing, cfl, t, final_time, fine_total, &cdt);
            if (rc) return from_solver(a, a->coarse, rc);
            if ((rc = amr_refine(a, cell_spacing, tolerance))) return rc;
            if ((rc = amr_flux_step(a, 0, order, cell_spacing, cdt))) return rc;
            bool refined = false;
            if ((rc = amr_any_refined(a, refined))) return rc;
            if (refined) {
                // fine sub-steps to the coarse time, less 1e-4 (pow(10.0, -4.0) in the reference)
                double ft = t;
                while (ft < ((t + cdt) - 1e-4)) {
                    double fdt = 0.0;
                    rc = cosmos_stable_time_step(a->fine, cell_spacing * 0.5, cfl, ft, t + cdt, fine_total, &fdt);
                    if (rc) return from_solver(a, a->fine, rc);
                    if ((rc = amr_flux_step(a, 1, order, cell_spacing * 0.5, fdt))) return rc;
                    ft += fdt;
                    fine_total += 1;
                }
            } else {
                fine_total += 1;
            }
            t += cdt;
            it += 1;
            if ((rc = amr_restrict(a))) return rc;
            if ((rc = amr_derefine(a, cell_spacing, tolerance))) return rc;
        }
        order_streams(a);
        if (time_out) *time_out = t;
        if (steps_out) *steps_out = it;
        return COSMOS_OK;
    });
}

}  // extern "C"
