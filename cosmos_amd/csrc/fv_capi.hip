// C ABI of the finite-volume update (declared in include/cosmos_fv.h).
//
// A cosmos_solver owns two device buffers of five SoA primitive planes and
// ping-pongs between them, one sweep at a time; the device-resident
// StepState carries t, dt and the iteration counter of a solve loop so that
// steps are queued without host round trips.  No C++ exception crosses the
// ABI: every entry point catches and converts to a status code.
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cosmos_fv.h"
#include "fv_gr_adjoint_launch.hpp"
#include "fv_kernels.hpp"

using cfv::ConstPlanes;
using cfv::Planes;
using cfv::StepState;
using cfv::SweepLaunch;

// The device status block (the cosmos_status_t layout): [0] con2prim solves
// at the iteration cap, [1] q floors, [2] density floors, [3] pressure floors,
// [4] NaN cells of the state (formed when read), [5] a solve loop took a NaN
// time step, [6] every cell excised from a black-hole wave speed, [7] unused.
constexpr int kStatusWords = 8;

struct cosmos_solver {
    int physics = 0;
    int dim = 2;
    int nx = 0, ny = 0;
    long pitch = 0;
    cfv::PhysParams P{};
    double* buf[2] = {nullptr, nullptr};
    int cur = 0;
    StepState* st = nullptr;   // device
    unsigned long long* st_slots = nullptr;    // device, max wave speed slots of the solve loop
    double* st_part = nullptr;                 // device, per-block maxima of the solve loop's last sweep
    unsigned long long* ws_scratch = nullptr;  // device, slots for cosmos_max_wave_speed
    unsigned long long* ws_host = nullptr;     // page-locked host copy of those slots (read back by DMA)
    int* flag = nullptr;                       // device, scratch flag
    unsigned long long* status = nullptr;      // device, the status block (cosmos_status_t layout, 8 words)
    int excised_known = 0;                     // black hole: excised_all computed for excised_key
    bool excised_all = false;
    double excised_key[6] = {0, 0, 0, 0, 0, 0};
    double* stage = nullptr;                   // device, AoS staging for transfers (lazily allocated)
    bool w_zero = false;                       // z-velocity plane identically zero (both buffers)
    int mode = COSMOS_MODE_EXACT;
    int y_rows = 0;                            // y-sweep rows per thread (0: by physics)
    int y_tile = -1;                           // y-sweep layout: 0 march, 8 / 16 tile columns (-1: by physics)
    int ws_mode = 0;                           // fused wave-speed hand-off (0: default)
    int fuse_step = 1;                         // fast mode: one fused kernel per 2D step where it applies
    int fuse_rows = 0;                         // its output rows per block (0: by occupancy)
    cfv::Hole hole{};                          // COSMOS_BLACKHOLE: the spacetime
    double excision = 0.0;                     // > 0: excision radius override
    double* gr_flux = nullptr;                 // device, interface fluxes of the GR sweeps (lazily allocated)
    long gr_fpitch = 0;
    // COSMOS_BLACKHOLE geometry cache (lazily allocated): the tables of one
    // (h, hole); geo_exact when h's half lattice is exact (Geo and neighbour
    // tables used), the source table otherwise alone
    int geo_cache = 1;
    double* geo_tab = nullptr;
    double* src_tab = nullptr;
    double* q_tab = nullptr;
    cfv::GeoTab geo{};
    bool geo_built = false, geo_exact = false;
    double geo_key[6] = {0, 0, 0, 0, 0, 0};
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false: the stream belongs to another handle (the levels of an AMR context)
    int device = 0;
    // solve-loop parameters set by cosmos_begin.  Steps are queued without
    // knowing on the host whether the device loop has stopped; a stopped step
    // writes nothing, so the current buffer is fixed up from the device's step
    // count (settle) before the host next touches the state.
    int run_order = 2;
    double run_bias = 0.0;
    int run_limiter = 1;
    bool run_ready = false;
    int cur_begin = 0;
    long queued = 0;
    int run_flips = 0;                         // buffer flips per step of a black-hole loop (0: sweeps_per_step)
    // special-relativistic solve loops: each sweep's dispatch order (blocks whose
    // solves hit the iteration cap in the previous step first) and marks
    unsigned char* hot_x = nullptr;
    unsigned char* hot_y = nullptr;
    int* order_x = nullptr;
    int* order_y = nullptr;
    int n_x = 0, n_y = 0;
    int row_x = 0, row_y = 0;                  // blocks per row of those block grids (0: marks not spread)
    bool unsettled = false;
    // optional per-kernel timing of queued steps (HIP events around each launch)
    bool profiling = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof;
    std::string err;
};

namespace {

thread_local std::string g_create_error;

int fail(cosmos_solver* s, int code, const std::string& msg) {
    if (s) s->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHK(s, call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((s), COSMOS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)


using cfv::make_hole;

Planes planes_of(cosmos_solver* s, int b) {
    Planes p;
    size_t plane = (size_t)s->ny * s->pitch;
    for (int k = 0; k < 5; ++k) p.p[k] = s->buf[b] + k * plane;
    return p;
}

ConstPlanes cplanes_of(cosmos_solver* s, int b) {
    ConstPlanes p;
    Planes q = planes_of(s, b);
    for (int k = 0; k < 5; ++k) p.p[k] = q.p[k];
    return p;
}

__global__ void k_aos_to_soa(const double* aos, Planes out, int nx, int ny, long pitch) {
    long total = (long)nx * ny;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (long)gridDim.x * blockDim.x) {
        long i = c / nx, j = c - i * nx;
        for (int k = 0; k < 5; ++k) out.p[k][i * pitch + j] = aos[5 * c + k];
    }
}

__global__ void k_soa_to_aos(ConstPlanes in, double* aos, int nx, int ny, long pitch) {
    long total = (long)nx * ny;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (long)gridDim.x * blockDim.x) {
        long i = c / nx, j = c - i * nx;
        for (int k = 0; k < 5; ++k) aos[5 * c + k] = in.p[k][i * pitch + j];
    }
}

dim3 grid_for(long cells) {
    long b = (cells + 255) / 256;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

SweepLaunch base_launch(cosmos_solver* s, int in_buf, int order, double bias, int limiter) {
    SweepLaunch L{};
    L.in = cplanes_of(s, in_buf);
    L.out = planes_of(s, in_buf ^ 1);
    L.nx = s->nx;
    L.ny = s->ny;
    L.pitch = s->pitch;
    L.bias = bias;
    L.limiter = limiter;
    L.order = order;
    L.physics = s->physics;
    L.dim = s->dim;
    L.w_zero = s->w_zero ? 1 : 0;
    L.fast = s->mode == COSMOS_MODE_FAST ? 1 : 0;
    L.P = s->P;
    L.step = nullptr;
    L.use_full = 0;
    L.fuse_ws = 0;
    L.y_rows = s->y_rows;
    L.y_tile = s->y_tile;
    L.fuse_rows = s->fuse_rows;
    L.status = s->status;
    return L;
}

// The special-relativistic loop's dispatch orders and marks for the sweep
// grids of this order (zero marks: the first step keeps the plain order).
int ensure_hot_order(cosmos_solver* s, int order) {
    if (s->physics != COSMOS_RELATIVISTIC) return COSMOS_OK;
    SweepLaunch L = base_launch(s, s->cur, order, 0.0, 1);
    const int nx = cfv::sweep_x_grid(L), ny = s->dim == 2 ? cfv::sweep_y_tile_grid(L) : 0;
    // marks spread to neighbouring blocks only under long iteration caps (the
    // reference's 1000), whose capped solves trail the launch; with short caps
    // the plain marks keep more of the XCD-contiguous order (measured faster)
    const bool spread = s->P.c2p_maxit > 100;
    s->row_x = spread ? cfv::sweep_row_blocks(L, 0) : 0;
    s->row_y = spread && s->dim == 2 ? cfv::sweep_row_blocks(L, 1) : 0;
    if (nx != s->n_x) {
        hipFree(s->hot_x);
        hipFree(s->order_x);
        s->hot_x = nullptr;
        s->order_x = nullptr;
        s->n_x = 0;
        HIPCHK(s, hipMalloc((void**)&s->hot_x, (size_t)nx));
        HIPCHK(s, hipMalloc((void**)&s->order_x, sizeof(int) * (size_t)nx));
        HIPCHK(s, hipMemsetAsync(s->hot_x, 0, (size_t)nx, s->stream));
        s->n_x = nx;
    }
    if (ny != s->n_y) {
        hipFree(s->hot_y);
        hipFree(s->order_y);
        s->hot_y = nullptr;
        s->order_y = nullptr;
        s->n_y = 0;
        if (ny > 0) {
            HIPCHK(s, hipMalloc((void**)&s->hot_y, (size_t)ny));
            HIPCHK(s, hipMalloc((void**)&s->order_y, sizeof(int) * (size_t)ny));
            HIPCHK(s, hipMemsetAsync(s->hot_y, 0, (size_t)ny, s->stream));
        }
        s->n_y = ny;
    }
    return COSMOS_OK;
}

int check_order(cosmos_solver* s, int order) {
    if (order != 1 && order != 2) return fail(s, COSMOS_ERR_ARGUMENT, "order must be 1 (FORCE) or 2 (SLIC)");
    if (order == 2 && (s->nx < 2 || (s->dim == 2 && s->ny < 2)))
        return fail(s, COSMOS_ERR_ARGUMENT, "SLIC needs at least 2 cells along each swept axis");
    return COSMOS_OK;
}

// Kernel kinds reported by cosmos_profile_read.
enum { kProfStep = 0, kProfSweepX = 1, kProfSweepY = 2, kProfSweepXWave = 3, kProfStepFused = 4, kProfKinds = 5 };

// Launches f() on the handle's stream, bracketed by HIP events when profiling.
template <class F>
hipError_t timed(cosmos_solver* s, int kind, F&& f) {
    if (!s->profiling) return f();
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, s->stream);
    if (e == hipSuccess) e = f();
    if (e == hipSuccess) e = hipEventRecord(b, s->stream);
    if (e != hipSuccess) {
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
        return e;
    }
    s->prof.push_back({kind, {a, b}});
    return hipSuccess;
}

int run_sweep(cosmos_solver* s, const SweepLaunch& L, int axis) {
    int kind = axis == 1 ? kProfSweepY : (L.fuse_ws ? kProfSweepXWave : kProfSweepX);
    hipError_t e = timed(s, kind, [&] {
        return axis == 0 ? cfv::launch_sweep_x(L, s->stream) : cfv::launch_sweep_y(L, s->stream);
    });
    if (e != hipSuccess) return fail(s, COSMOS_ERR_DEVICE, std::string("sweep launch: ") + hipGetErrorString(e));
    return COSMOS_OK;
}

int sweeps_per_step(const cosmos_solver* s) { return s->dim == 2 ? 3 : 1; }

// Whether the queued 2D step runs as one fused kernel: fast mode, Newtonian
// SLIC with van Leer and bias 0, at least 8 cells along each axis, wave speed
// handed over through the slots.  The buffer parity is the three sweeps'
// (one hand-off instead of three), so settle() is unchanged.
bool fused_step(const cosmos_solver* s) {
    return s->fuse_step && s->mode == COSMOS_MODE_FAST && s->physics == COSMOS_NEWTONIAN && s->dim == 2 &&
           s->run_order == 2 && s->run_limiter == 1 && s->run_bias == 0.0 && s->nx >= 8 && s->ny >= 8 &&
           s->ws_mode != 1;
}

// Queue one solve-loop iteration reading its dt from the device StepState.
int queue_step(cosmos_solver* s) {
    hipError_t e = timed(s, kProfStep, [&] {
        hipError_t r = cfv::launch_step_begin(s->st, s->stream);
        if (r == hipSuccess && s->n_x) r = cfv::launch_hot_order(s->hot_x, s->order_x, s->n_x, s->row_x, s->stream);
        if (r == hipSuccess && s->n_y) r = cfv::launch_hot_order(s->hot_y, s->order_y, s->n_y, s->row_y, s->stream);
        return r;
    });
    if (e != hipSuccess) return fail(s, COSMOS_ERR_DEVICE, std::string("step launch: ") + hipGetErrorString(e));
    const int spp = sweeps_per_step(s);
    const int b0 = (int)((s->cur_begin + (long)spp * s->queued) & 1);
    if (fused_step(s)) {
        SweepLaunch L = base_launch(s, b0, s->run_order, s->run_bias, s->run_limiter);
        L.step = s->st;
        e = timed(s, kProfStepFused, [&] { return cfv::launch_step_fused(L, s->stream); });
        if (e != hipSuccess)
            return fail(s, COSMOS_ERR_DEVICE, std::string("fused step launch: ") + hipGetErrorString(e));
        s->queued += 1;
        s->unsettled = true;
        return COSMOS_OK;
    }
    for (int k = 0; k < spp; ++k) {
        SweepLaunch L = base_launch(s, b0 ^ (k & 1), s->run_order, s->run_bias, s->run_limiter);
        L.step = s->st;
        L.use_full = (spp == 1) || (k == 1);
        L.fuse_ws = (k == spp - 1);
        const bool y = spp == 3 && k == 1;
        L.block_order = y ? s->order_y : s->order_x;
        L.order_n = y ? s->n_y : s->n_x;
        L.hot = y ? s->hot_y : s->hot_x;
        int rc = run_sweep(s, L, (spp == 3 && k == 1) ? 1 : 0);
        if (rc) return rc;
    }
    s->queued += 1;
    s->unsettled = true;
    return COSMOS_OK;
}

// After queued solve steps: wait, and point cur at the buffer that holds the
// state after the steps the device actually took.
int settle(cosmos_solver* s) {
    if (!s->unsettled) return COSMOS_OK;
    StepState h{};
    HIPCHK(s, hipMemcpyAsync(&h, s->st, sizeof h, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const int flips = s->run_flips ? s->run_flips : sweeps_per_step(s);
    s->cur = (int)((s->cur_begin + (long)flips * h.active_steps) & 1);
    s->unsettled = false;
    return COSMOS_OK;
}

template <class F>
int guarded(cosmos_solver* s, F&& f) {
    if (!s) return fail(nullptr, COSMOS_ERR_ARGUMENT, "null solver handle");
    try {
        hipError_t e = hipSetDevice(s->device);
        if (e != hipSuccess) return fail(s, COSMOS_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
        return f();
    } catch (const std::bad_alloc&) {
        return fail(s, COSMOS_ERR_DEVICE, "host allocation failed");
    } catch (...) {
        return fail(s, COSMOS_ERR_STATE, "unexpected exception");
    }
}

// ---------------------------------------------------------------------------
// General relativistic (COSMOS_BLACKHOLE) launches
// ---------------------------------------------------------------------------
cfv::GRLaunch gr_launch(cosmos_solver* s, double h) {
    cfv::GRLaunch L{};
    L.in = cplanes_of(s, s->cur);
    L.out = planes_of(s, s->cur ^ 1);
    L.nx = s->nx;
    L.ny = s->ny;
    L.pitch = s->pitch;
    L.dim = s->dim;
    L.h = h;
    L.P = s->P;
    L.hole = s->hole;
    if (s->gr_flux) {
        size_t plane = (size_t)(s->ny + 1) * s->gr_fpitch;
        for (int k = 0; k < 5; ++k) L.flux.p[k] = s->gr_flux + k * plane;
        L.fpitch = s->gr_fpitch;
    }
    return L;
}

// Whether every half-lattice position a h / 2, |a| <= 2 max(nx, ny) + 8, and
// every sum or difference of two of them is exact: h = m 2^e with m odd and
// a m < 2^53.
bool lattice_exact(double h, int nx, int ny) {
    if (!(h > 1e-290) || !std::isfinite(h)) return false;
    int e = 0;
    double f = std::frexp(h, &e);
    uint64_t m = (uint64_t)std::ldexp(f, 53);
    while (m && !(m & 1)) m >>= 1;
    const uint64_t amax = 2ull * (uint64_t)(nx > ny ? nx : ny) + 8ull;
    return m != 0 && m < (1ull << 53) / amax;
}

// The geometry cache for spacing h and the solver's hole (built on first use
// and whenever either changes).  With the cache off or for other physics the
// launch keeps an empty table and the kernels evaluate at the call sites.
int ensure_geo(cosmos_solver* s, double h, cfv::GeoTab& tab) {
    tab = cfv::GeoTab{};
    if (!s->geo_cache || s->physics != COSMOS_BLACKHOLE) return COSMOS_OK;
    const double key[6] = {h, s->hole.M, s->hole.a, s->hole.x, s->hole.y, s->hole.z};
    const bool same = s->geo_built && std::memcmp(key, s->geo_key, sizeof key) == 0;
    if (!same) {
        const long rows = s->dim == 2 ? s->ny + 4 : 1;
        const long pitch = ((long)s->nx + 4 + 15) / 16 * 16;
        const long spitch = s->pitch;
        const long qrows = s->dim == 2 ? s->ny + 2 : 1;
        const long qpitch = ((long)s->nx + 2 + 15) / 16 * 16;
        if (!s->src_tab) {
            HIPCHK(s, hipMalloc((void**)&s->src_tab, sizeof(double) * cfv::kSrcVals * (size_t)s->ny * spitch));
        }
        const bool exact = lattice_exact(h, s->nx, s->ny);
        if (exact && !s->geo_tab) {
            HIPCHK(s, hipMalloc((void**)&s->geo_tab, sizeof(double) * 3 * cfv::kGeoVals * (size_t)rows * pitch));
            HIPCHK(s, hipMalloc((void**)&s->q_tab, sizeof(double) * 5 * (size_t)qrows * qpitch));
        }
        cfv::GeoTab t{};
        t.pitch = pitch;
        t.plane = rows * pitch;
        t.spitch = spitch;
        t.splane = (long)s->ny * spitch;
        t.qpitch = qpitch;
        t.qplane = qrows * qpitch;
        cfv::GRLaunch L = gr_launch(s, h);
        L.tab = t;
        hipError_t e = cfv::launch_gr_geo_build(L, exact ? s->geo_tab : nullptr, s->src_tab, s->stream);
        if (e != hipSuccess) return fail(s, COSMOS_ERR_DEVICE, std::string("geometry cache: ") + hipGetErrorString(e));
        t.geo = exact ? s->geo_tab : nullptr;
        t.src = s->src_tab;
        t.q = exact ? s->q_tab : nullptr;
        s->geo = t;
        s->geo_exact = exact;
        s->geo_built = true;
        std::memcpy(s->geo_key, key, sizeof key);
    }
    tab = s->geo;
    return COSMOS_OK;
}

int ensure_gr_flux(cosmos_solver* s) {
    if (s->gr_flux) return COSMOS_OK;
    // x sweeps: ny rows of nx + 1 interfaces; y sweeps: ny + 1 rows of nx
    s->gr_fpitch = ((long)s->nx + 1 + 15) / 16 * 16;
    HIPCHK(s, hipMalloc((void**)&s->gr_flux, sizeof(double) * 5 * (size_t)(s->ny + 1) * s->gr_fpitch));
    return COSMOS_OK;
}

int gr_sweep_setup(cosmos_solver* s, int order, double h, double dt, double bias, int limiter, cfv::GRLaunch& L) {
    int rc = ensure_gr_flux(s);
    if (rc) return rc;
    L = gr_launch(s, h);
    if ((rc = ensure_geo(s, h, L.tab))) return rc;
    L.order = order;
    L.bias = bias;
    L.limiter = limiter;
    L.coef.dt_dx = dt / h;
    L.coef.dx_dt = h / dt;
    L.coef.hdt_dx = 0.5 * (dt / h);
    return COSMOS_OK;
}

int gr_sweep(cosmos_solver* s, int order, int axis, double h, double dt, double bias, int limiter) {
    cfv::GRLaunch L;
    int rc = gr_sweep_setup(s, order, h, dt, bias, limiter, L);
    if (rc) return rc;
    hipError_t e = timed(s, axis == 0 ? kProfSweepX : kProfSweepY, [&] { return cfv::launch_gr_sweep(L, axis, s->stream); });
    if (e != hipSuccess) return fail(s, COSMOS_ERR_DEVICE, std::string("GR sweep launch: ") + hipGetErrorString(e));
    s->cur ^= 1;
    return COSMOS_OK;
}

// The stage fractions of evolveConservedVariableVector{,2D}
// (Solvers/BlackHoleEulerForcingSolver.cpp:21-53, 98-167), each times dt / h as
// computeFractionalEvolutionVector forms it (Solvers/EulerSolvers.cpp:163-167).
int gr_forcing_setup(cosmos_solver* s, double h, double dt, double bias, int limiter, cfv::GRLaunch& L) {
    L = gr_launch(s, h);
    int rc = ensure_geo(s, h, L.tab);
    if (rc) return rc;
    L.bias = bias;
    L.limiter = limiter;
    L.dt = dt;
    static const double f1[3] = {0.5, 0.5, 1.0}, fx[3] = {0.25, 0.25, 0.5}, fy[3] = {0.5, 0.5, 1.0};
    for (int k = 0; k < 3; ++k) {
        L.frac_dt_h[k] = (s->dim == 1 ? f1[k] : fx[k]) * (dt / h);
        L.frac_dt_h_y[k] = fy[k] * (dt / h);
    }
    return COSMOS_OK;
}

int gr_forcing(cosmos_solver* s, double h, double dt, double bias, int limiter) {
    cfv::GRLaunch L;
    int rc = gr_forcing_setup(s, h, dt, bias, limiter, L);
    if (rc) return rc;
    hipError_t e = timed(s, kProfStep, [&] { return cfv::launch_gr_forcing(L, s->stream); });
    if (e != hipSuccess) return fail(s, COSMOS_ERR_DEVICE, std::string("GR forcing launch: ") + hipGetErrorString(e));
    s->cur ^= 1;
    return COSMOS_OK;
}

// computeMaximumWaveSpeed{,2D} skips excised cells at the unit-square
// positions (j / nx, i / ny); whether that is every cell depends on the
// spacetime and the grid only, so it is found once per spacetime and marked
// in status word 6 (a fully excised grid).
int gr_excised_status(cosmos_solver* s) {
    const double key[6] = {s->hole.M, s->hole.a, s->hole.x, s->hole.y, s->hole.z, s->hole.r_exc};
    if (!s->excised_known || std::memcmp(key, s->excised_key, sizeof key) != 0) {
        const double dr = 1.0 / (double)s->ny, dc = 1.0 / (double)s->nx;
        bool all = true;
        for (int i = 0; i < s->ny && all; ++i)
            for (int j = 0; j < s->nx && all; ++j) all = cfv::excised(s->hole, j * dc, s->dim == 2 ? i * dr : 0.0, 0.0);
        s->excised_all = all;
        s->excised_known = 1;
        std::memcpy(s->excised_key, key, sizeof key);
    }
    if (s->excised_all) HIPCHK(s, cfv::launch_fill_u64(s->status + 6, 1, 1ull, s->stream));
    return COSMOS_OK;
}

// One step of the device-resident black-hole loop (the solve loops of
// Solvers/BlackHoleEulerSecondOrderSolver.cpp:145-218 and ..FirstOrderSolver.cpp:183-253):
// the state's wave speed into the step state's slots, the time step and
// every coefficient formed on the device (k_step_begin), forcing(dt / 2N) x N
// with the forcing's bias 0 and van Leer as in the reference, the flux step,
// forcing(dt / 2N) x N.  Each launch takes its buffers from the step's place
// in the loop and does nothing once the loop has stopped.
int queue_gr_step(cosmos_solver* s, int order, double h, double bias, int limiter, int subcycles,
                  const cfv::GeoTab& tab) {
    int b = (int)((s->cur_begin + (long)s->run_flips * s->queued) & 1);
    auto at = [&](int buf) {
        cfv::GRLaunch L = gr_launch(s, h);
        L.in = cplanes_of(s, buf);
        L.out = planes_of(s, buf ^ 1);
        L.tab = tab;
        L.step = s->st;
        return L;
    };
    hipError_t e = timed(s, kProfSweepXWave, [&] {
        hipError_t r = cfv::launch_gr_wavespeed(at(b), s->st_slots, s->stream);
        return r == hipSuccess ? cfv::launch_step_begin(s->st, s->stream) : r;
    });
    auto forcing = [&]() {
        for (int k = 0; k < subcycles && e == hipSuccess; ++k) {
            cfv::GRLaunch L = at(b);
            L.bias = 0.0;
            L.limiter = 1;
            e = timed(s, kProfStep, [&] { return cfv::launch_gr_forcing(L, s->stream); });
            b ^= 1;
        }
    };
    auto sweep = [&](int axis, int full) {
        if (e != hipSuccess) return;
        cfv::GRLaunch L = at(b);
        L.order = order;
        L.bias = bias;
        L.limiter = limiter;
        L.use_full = full;
        e = timed(s, axis == 0 ? kProfSweepX : kProfSweepY, [&] { return cfv::launch_gr_sweep(L, axis, s->stream); });
        b ^= 1;
    };
    forcing();
    if (s->dim == 1) {
        sweep(0, 1);
    } else {
        sweep(0, 0);
        sweep(1, 1);
        sweep(0, 0);
    }
    forcing();
    if (e != hipSuccess) return fail(s, COSMOS_ERR_DEVICE, std::string("black-hole step launch: ") + hipGetErrorString(e));
    s->queued += 1;
    s->unsettled = true;
    return COSMOS_OK;
}

}  // namespace

namespace cfv {

int gr_sweep_launch(cosmos_solver* s, int order, double h, double dt, double bias, int limiter, GRLaunch* L) {
    if (!s || s->physics != COSMOS_BLACKHOLE) return COSMOS_ERR_ARGUMENT;
    return gr_sweep_setup(s, order, h, dt, bias, limiter, *L);
}

int gr_forcing_launch(cosmos_solver* s, double h, double dt, double bias, int limiter, GRLaunch* L) {
    if (!s || s->physics != COSMOS_BLACKHOLE) return COSMOS_ERR_ARGUMENT;
    return gr_forcing_setup(s, h, dt, bias, limiter, *L);
}

}  // namespace cfv

extern "C" {

int cosmos_create(cosmos_solver** out, int physics, int dim, int nx, int ny, double adiabatic_index) {
    if (!out) return fail(nullptr, COSMOS_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    if (physics != COSMOS_NEWTONIAN && physics != COSMOS_RELATIVISTIC && physics != COSMOS_BLACKHOLE)
        return fail(nullptr, COSMOS_ERR_ARGUMENT,
                    "physics must be COSMOS_NEWTONIAN, COSMOS_RELATIVISTIC or COSMOS_BLACKHOLE");
    if (dim != 1 && dim != 2) return fail(nullptr, COSMOS_ERR_ARGUMENT, "dim must be 1 or 2");
    if (dim == 1) ny = 1;
    if (nx < 1 || ny < 1) return fail(nullptr, COSMOS_ERR_ARGUMENT, "grid must have at least one cell");
    if ((long)nx * ny > (1L << 31)) return fail(nullptr, COSMOS_ERR_ARGUMENT, "grid too large");
    cosmos_solver* s = nullptr;
    try {
        s = new cosmos_solver();
    } catch (...) {
        return fail(nullptr, COSMOS_ERR_DEVICE, "host allocation failed");
    }
    s->physics = physics;
    s->dim = dim;
    s->nx = nx;
    s->ny = ny;
    s->pitch = ((long)nx + 15) / 16 * 16;
    s->P = cfv::make_params(adiabatic_index);
    s->hole = make_hole(1.0, 0.0, 0.5, 0.5, 0.0, 0.0);
    hipError_t e = hipGetDevice(&s->device);
    size_t bytes = sizeof(double) * 5 * (size_t)ny * s->pitch;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&s->buf[0], bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->buf[1], bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->st, sizeof(StepState));
    if (e == hipSuccess) e = hipMalloc((void**)&s->st_slots, sizeof(unsigned long long) * cfv::kSlotWords);
    if (e == hipSuccess) e = hipMalloc((void**)&s->ws_scratch, sizeof(unsigned long long) * cfv::kSlotWords);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->ws_host, sizeof(unsigned long long) * cfv::kSlotWords);
    if (e == hipSuccess) e = hipMalloc((void**)&s->flag, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&s->status, sizeof(unsigned long long) * kStatusWords);
    if (e == hipSuccess)
        e = hipMalloc((void**)&s->st_part, sizeof(double) * (size_t)cfv::sweep_x_blocks(s->nx, s->ny, 2));
    // Zero on the handle's own stream and wait: the stream is non-blocking, so
    // a fill queued on the null stream would not be ordered before the first
    // upload into these buffers.
    if (e == hipSuccess) e = hipMemsetAsync(s->buf[0], 0, bytes, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->buf[1], 0, bytes, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->st, 0, sizeof(StepState), s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->status, 0, sizeof(unsigned long long) * kStatusWords, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        std::string msg = std::string("device setup: ") + hipGetErrorString(e);
        cosmos_destroy(s);
        return fail(nullptr, COSMOS_ERR_DEVICE, msg);
    }
    *out = s;
    return COSMOS_OK;
}

int cosmos_destroy(cosmos_solver* s) {
    if (!s) return COSMOS_OK;
    hipSetDevice(s->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    for (auto& p : s->prof) {
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    hipFree(s->buf[0]);
    hipFree(s->buf[1]);
    hipFree(s->st);
    hipFree(s->st_slots);
    hipFree(s->st_part);
    hipFree(s->ws_scratch);
    hipHostFree(s->ws_host);
    hipFree(s->flag);
    hipFree(s->status);
    hipFree(s->stage);
    hipFree(s->hot_x);
    hipFree(s->hot_y);
    hipFree(s->order_x);
    hipFree(s->order_y);
    hipFree(s->gr_flux);
    hipFree(s->geo_tab);
    hipFree(s->src_tab);
    hipFree(s->q_tab);
    if (s->stream && s->own_stream) hipStreamDestroy(s->stream);
    delete s;
    return COSMOS_OK;
}

const char* cosmos_last_error(const cosmos_solver* s) {
    return s ? s->err.c_str() : g_create_error.c_str();
}

int cosmos_set_mode(cosmos_solver* s, int mode) {
    return guarded(s, [&] {
        if (mode != COSMOS_MODE_EXACT && mode != COSMOS_MODE_FAST)
            return fail(s, COSMOS_ERR_ARGUMENT, "mode must be COSMOS_MODE_EXACT or COSMOS_MODE_FAST");
        s->mode = mode;
        return COSMOS_OK;
    });
}

int cosmos_set_launch(cosmos_solver* s, int y_rows, int ws_mode) {
    return guarded(s, [&] {
        if (y_rows != 0 && y_rows != 16 && y_rows != 32 && y_rows != 64)
            return fail(s, COSMOS_ERR_ARGUMENT, "y_rows must be 0 (default), 16, 32 or 64");
        if (ws_mode < 0 || ws_mode > 2) return fail(s, COSMOS_ERR_ARGUMENT, "ws_mode must be 0, 1 or 2");
        s->y_rows = y_rows;
        s->ws_mode = ws_mode;
        s->run_ready = false;
        return COSMOS_OK;
    });
}

int cosmos_set_step_fusion(cosmos_solver* s, int enable, int rows_per_block) {
    return guarded(s, [&] {
        if (enable != 0 && enable != 1) return fail(s, COSMOS_ERR_ARGUMENT, "step fusion must be 0 or 1");
        if (rows_per_block != 0 && (rows_per_block < 4 || rows_per_block > 1024))
            return fail(s, COSMOS_ERR_ARGUMENT, "fused rows per block must be 0 (automatic) or in [4, 1024]");
        s->fuse_step = enable;
        s->fuse_rows = rows_per_block;
        s->run_ready = false;
        return COSMOS_OK;
    });
}

int cosmos_set_y_tile(cosmos_solver* s, int columns) {
    return guarded(s, [&] {
        if (columns != -1 && columns != 0 && columns != 8 && columns != 16)
            return fail(s, COSMOS_ERR_ARGUMENT, "y tile columns must be -1 (default), 0 (march), 8 or 16");
        s->y_tile = columns;
        s->run_ready = false;
        return COSMOS_OK;
    });
}

int cosmos_set_con2prim(cosmos_solver* s, double tolerance, int max_iterations) {
    return guarded(s, [&] {
        if (!(tolerance > 0.0) || max_iterations < 1)
            return fail(s, COSMOS_ERR_ARGUMENT, "con2prim tolerance must be > 0 and max_iterations >= 1");
        s->P.c2p_tol = tolerance;
        s->P.c2p_maxit = max_iterations;
        if (s->physics == COSMOS_BLACKHOLE) s->P.gr_tol = tolerance;
        return COSMOS_OK;
    });
}

int cosmos_set_spacetime(cosmos_solver* s, double mass, double spin, double x, double y, double z) {
    return guarded(s, [&] {
        if (s->physics != COSMOS_BLACKHOLE)
            return fail(s, COSMOS_ERR_ARGUMENT, "cosmos_set_spacetime needs a COSMOS_BLACKHOLE solver");
        s->hole = make_hole(mass, spin, x, y, z, s->excision);
        return COSMOS_OK;
    });
}

int cosmos_set_geometry_cache(cosmos_solver* s, int enable) {
    return guarded(s, [&] {
        s->geo_cache = enable ? 1 : 0;
        return COSMOS_OK;
    });
}

int cosmos_set_excision(cosmos_solver* s, double radius) {
    return guarded(s, [&] {
        if (s->physics != COSMOS_BLACKHOLE)
            return fail(s, COSMOS_ERR_ARGUMENT, "cosmos_set_excision needs a COSMOS_BLACKHOLE solver");
        s->excision = radius > 0.0 ? radius : 0.0;
        s->hole = make_hole(s->hole.M, s->hole.a, s->hole.x, s->hole.y, s->hole.z, s->excision);
        return COSMOS_OK;
    });
}

static int ensure_stage(cosmos_solver* s) {
    if (s->stage) return COSMOS_OK;
    HIPCHK(s, hipMalloc((void**)&s->stage, sizeof(double) * 5 * (size_t)s->nx * s->ny));
    return COSMOS_OK;
}

// Host transfers.  Page-locked host memory (cosmos_host_alloc, hipHostMalloc)
// is copied by DMA directly.  Pageable memory goes through two page-locked
// staging chunks: the CPU copy of one chunk (split over worker threads)
// overlaps the DMA of the other, instead of the runtime's own pageable path.
// The chunks belong to the device and are shared by its solvers (one
// transfer at a time holds them; 64 MiB of page-locked memory per device,
// allocated on the first pageable transfer and kept until the process ends).
constexpr size_t kChunkBytes = size_t(32) << 20;
constexpr int kCopyThreads = 8;
constexpr int kMaxStageDevices = 64;

struct StagePool {
    std::mutex mu;
    char* chunk[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};  // the last DMA of each chunk
};

StagePool g_stage[kMaxStageDevices];

static bool is_page_locked(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory is unknown to the runtime: not an error
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

static void par_memcpy(void* dst, const void* src, size_t bytes) {
    if (bytes < (size_t(4) << 20)) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes / kCopyThreads + 4095) & ~size_t(4095);
    std::thread th[kCopyThreads];
    int n = 0;
    for (size_t off = 0; off < bytes && n < kCopyThreads; off += part, ++n) {
        const size_t len = bytes - off < part ? bytes - off : part;
        th[n] = std::thread([=] { std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, len); });
    }
    for (int k = 0; k < n; ++k) th[k].join();
}

// The device's pool, its chunks allocated; called with pool.mu held.
static int stage_pool(cosmos_solver* s, StagePool** out) {
    if (s->device < 0 || s->device >= kMaxStageDevices)
        return fail(s, COSMOS_ERR_DEVICE, "device index beyond the supported 64");
    StagePool& p = g_stage[s->device];
    for (int k = 0; k < 2; ++k) {
        if (!p.chunk[k]) HIPCHK(s, hipHostMalloc((void**)&p.chunk[k], kChunkBytes, hipHostMallocDefault));
        if (!p.ev[k]) HIPCHK(s, hipEventCreateWithFlags(&p.ev[k], hipEventDisableTiming));
    }
    *out = &p;
    return COSMOS_OK;
}

// host -> device on the solver's stream, complete on return
static int h2d(cosmos_solver* s, void* dev, const void* host, size_t bytes) {
    if (is_page_locked(host)) {
        HIPCHK(s, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return COSMOS_OK;
    }
    if (s->device < 0 || s->device >= kMaxStageDevices)
        return fail(s, COSMOS_ERR_DEVICE, "device index beyond the supported 64");
    std::lock_guard<std::mutex> lock(g_stage[s->device].mu);
    StagePool* p = nullptr;
    int rc = stage_pool(s, &p);
    if (rc) return rc;
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += kChunkBytes, ++i) {
        const int b = (int)(i & 1);
        const size_t len = bytes - off < kChunkBytes ? bytes - off : kChunkBytes;
        HIPCHK(s, hipEventSynchronize(p->ev[b]));  // the chunk's previous DMA has drained
        par_memcpy(p->chunk[b], static_cast<const char*>(host) + off, len);
        HIPCHK(s, hipMemcpyAsync(static_cast<char*>(dev) + off, p->chunk[b], len, hipMemcpyHostToDevice, s->stream));
        HIPCHK(s, hipEventRecord(p->ev[b], s->stream));
    }
    HIPCHK(s, hipStreamSynchronize(s->stream));  // the chunks are free for the next transfer
    return COSMOS_OK;
}

// device -> host, complete on return
static int d2h(cosmos_solver* s, void* host, const void* dev, size_t bytes) {
    if (is_page_locked(host)) {
        HIPCHK(s, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return COSMOS_OK;
    }
    if (s->device < 0 || s->device >= kMaxStageDevices)
        return fail(s, COSMOS_ERR_DEVICE, "device index beyond the supported 64");
    std::lock_guard<std::mutex> lock(g_stage[s->device].mu);
    StagePool* p = nullptr;
    int rc = stage_pool(s, &p);
    if (rc) return rc;
    const size_t n = (bytes + kChunkBytes - 1) / kChunkBytes;
    auto issue = [&](size_t i) -> int {
        const int b = (int)(i & 1);
        const size_t off = i * kChunkBytes;
        const size_t len = bytes - off < kChunkBytes ? bytes - off : kChunkBytes;
        HIPCHK(s, hipMemcpyAsync(p->chunk[b], static_cast<const char*>(dev) + off, len, hipMemcpyDeviceToHost,
                                 s->stream));
        HIPCHK(s, hipEventRecord(p->ev[b], s->stream));
        return COSMOS_OK;
    };
    for (size_t i = 0; i < n && i < 2; ++i)
        if ((rc = issue(i))) return rc;
    for (size_t i = 0; i < n; ++i) {
        const int b = (int)(i & 1);
        const size_t off = i * kChunkBytes;
        const size_t len = bytes - off < kChunkBytes ? bytes - off : kChunkBytes;
        HIPCHK(s, hipEventSynchronize(p->ev[b]));
        par_memcpy(static_cast<char*>(host) + off, p->chunk[b], len);
        if (i + 2 < n && (rc = issue(i + 2))) return rc;
    }
    return COSMOS_OK;
}

static int upload_impl(cosmos_solver* s, const double* prim, hipMemcpyKind kind) {
    if (!prim) return fail(s, COSMOS_ERR_ARGUMENT, "null state pointer");
    int rc = settle(s);
    if (rc) return rc;
    rc = ensure_stage(s);
    if (rc) return rc;
    long cells = (long)s->nx * s->ny;
    if (kind == hipMemcpyHostToDevice) {
        if ((rc = h2d(s, s->stage, prim, sizeof(double) * 5 * cells))) return rc;
    } else {
        HIPCHK(s, hipMemcpyAsync(s->stage, prim, sizeof(double) * 5 * cells, kind, s->stream));
    }
    hipLaunchKernelGGL(k_aos_to_soa, grid_for(cells), dim3(256), 0, s->stream, s->stage, planes_of(s, s->cur), s->nx,
                       s->ny, s->pitch);
    HIPCHK(s, hipGetLastError());
    // Is the z-velocity plane identically zero?  Then the sweeps skip it (its
    // terms are exact zeros), and the other buffer's plane is zeroed to match.
    int nonzero = 1;
    HIPCHK(s, hipMemsetAsync(s->flag, 0, sizeof(int), s->stream));
    HIPCHK(s, cfv::launch_any_nonzero(planes_of(s, s->cur).p[3], s->nx, s->ny, s->pitch, s->flag, s->stream));
    HIPCHK(s, hipMemcpyAsync(&nonzero, s->flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->w_zero = (nonzero == 0);
    if (s->w_zero) {
        HIPCHK(s, hipMemsetAsync(planes_of(s, s->cur ^ 1).p[3], 0, sizeof(double) * (size_t)s->ny * s->pitch,
                                 s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
    }
    s->run_ready = false;
    return COSMOS_OK;
}

static int download_impl(cosmos_solver* s, double* prim, hipMemcpyKind kind) {
    if (!prim) return fail(s, COSMOS_ERR_ARGUMENT, "null state pointer");
    int rc = settle(s);
    if (rc) return rc;
    rc = ensure_stage(s);
    if (rc) return rc;
    long cells = (long)s->nx * s->ny;
    hipLaunchKernelGGL(k_soa_to_aos, grid_for(cells), dim3(256), 0, s->stream, cplanes_of(s, s->cur), s->stage,
                       s->nx, s->ny, s->pitch);
    HIPCHK(s, hipGetLastError());
    if (kind == hipMemcpyDeviceToHost) return d2h(s, prim, s->stage, sizeof(double) * 5 * cells);
    HIPCHK(s, hipMemcpyAsync(prim, s->stage, sizeof(double) * 5 * cells, kind, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return COSMOS_OK;
}

int cosmos_host_alloc(size_t bytes, void** out) {
    if (!out || bytes == 0) return COSMOS_ERR_ARGUMENT;
    *out = nullptr;
    return hipHostMalloc(out, bytes, hipHostMallocDefault) == hipSuccess ? COSMOS_OK : COSMOS_ERR_DEVICE;
}

int cosmos_host_free(void* p) {
    if (!p) return COSMOS_OK;
    return hipHostFree(p) == hipSuccess ? COSMOS_OK : COSMOS_ERR_DEVICE;
}

int cosmos_upload(cosmos_solver* s, const double* prim) {
    return guarded(s, [&] { return upload_impl(s, prim, hipMemcpyHostToDevice); });
}
int cosmos_download(cosmos_solver* s, double* prim) {
    return guarded(s, [&] { return download_impl(s, prim, hipMemcpyDeviceToHost); });
}
int cosmos_upload_device(cosmos_solver* s, const double* prim_device) {
    return guarded(s, [&] { return upload_impl(s, prim_device, hipMemcpyDeviceToDevice); });
}
int cosmos_download_device(cosmos_solver* s, double* prim_device) {
    return guarded(s, [&] { return download_impl(s, prim_device, hipMemcpyDeviceToDevice); });
}

int cosmos_sweep(cosmos_solver* s, int order, int axis, double cell_spacing, double time_step, double bias,
                 int slope_limiter) {
    return guarded(s, [&] {
        int rc = check_order(s, order);
        if (rc) return rc;
        if (axis != 0 && axis != 1) return fail(s, COSMOS_ERR_ARGUMENT, "axis must be 0 (x) or 1 (y)");
        if (axis == 1 && s->dim == 1) return fail(s, COSMOS_ERR_ARGUMENT, "a 1D solver has no y sweep");
        rc = settle(s);
        if (rc) return rc;
        s->run_ready = false;
        if (s->physics == COSMOS_BLACKHOLE)
            return gr_sweep(s, order, axis, cell_spacing, time_step, bias, slope_limiter);
        SweepLaunch L = base_launch(s, s->cur, order, bias, slope_limiter);
        L.coef.dt_dx = time_step / cell_spacing;
        L.coef.dx_dt = cell_spacing / time_step;
        L.coef.hdt_dx = 0.5 * (time_step / cell_spacing);
        s->run_ready = false;
        rc = run_sweep(s, L, axis);
        if (rc) return rc;
        s->cur ^= 1;
        return COSMOS_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// AMR support (called by fv_amr.hip, not part of the C-ABI)
// ---------------------------------------------------------------------------
namespace cfv {

// The masked sweep of a Newtonian or special-relativistic level
// (compute{X,Y}{SLIC,FORCE}TimeStepAMR: each flagged cell updated from the
// previous state through the full boundary copy, the others unchanged): only
// tiles that hold a flagged cell are swept, and only flagged cells are written,
// into the other buffer; amr_mask_writeback then copies them back, so the
// state stays in the current buffer.  Not counted in the status word.
// blocks (else null): the list_n blocks of this sweep that hold a flagged
// cell (amr_mask_lists for this mask and order); only those are launched.
int amr_masked_sweep(cosmos_solver* s, int order, int axis, double cell_spacing, double time_step, double bias,
                     int slope_limiter, const unsigned char* mask, const int* blocks, int list_n) {
    return guarded(s, [&] {
        int rc = check_order(s, order);
        if (rc) return rc;
        if (axis != 0 && axis != 1) return fail(s, COSMOS_ERR_ARGUMENT, "axis must be 0 (x) or 1 (y)");
        if (axis == 1 && s->dim == 1) return fail(s, COSMOS_ERR_ARGUMENT, "a 1D solver has no y sweep");
        if (s->physics == COSMOS_BLACKHOLE) return fail(s, COSMOS_ERR_ARGUMENT, "black-hole levels sweep unmasked");
        rc = settle(s);
        if (rc) return rc;
        s->run_ready = false;
        SweepLaunch L = base_launch(s, s->cur, order, bias, slope_limiter);
        L.coef.dt_dx = time_step / cell_spacing;
        L.coef.dx_dt = cell_spacing / time_step;
        L.coef.hdt_dx = 0.5 * (time_step / cell_spacing);
        L.status = nullptr;
        L.mask = mask;
        if (blocks) {
            L.block_order = blocks;
            L.order_n = list_n;
        }
        return run_sweep(s, L, axis);
    });
}

// The masked sweep's flagged cells from the other buffer back into the
// current one (segs: the n_segs 64-cell segments holding a flagged cell, or
// null for all).
hipError_t amr_mask_writeback(cosmos_solver* s, const unsigned char* mask, const int* segs, int n_segs) {
    return launch_mask_writeback(planes_of(s, s->cur), cplanes_of(s, s->cur ^ 1), mask, s->nx, s->ny, s->pitch,
                                 segs, n_segs, s->stream);
}

// Columns of s's masked y-sweep tiles (the lists of amr_mask_lists depend on them).
int amr_mask_tc(cosmos_solver* s) { return mask_tile_cols(base_launch(s, s->cur, 2, 0.0, 1)); }

// The blocks of s's masked sweeps of this order, and the write-back
// segments, that hold a flagged cell of mask, queued on s's stream
// (launch_mask_lists; hit: its zeroed mark maps).
hipError_t amr_mask_lists(cosmos_solver* s, const unsigned char* mask, int order, unsigned char* hit, int* xl,
                          int* yl, int* wl, int* counts) {
    return launch_mask_lists(mask, s->nx, s->ny, s->dim, order, amr_mask_tc(s), hit, xl, yl, wl, counts, s->stream);
}

// Queue s's work on owner's stream from now on (the levels of an AMR context
// then need no host synchronisation between them); s must not outlive owner.
int amr_share_stream(cosmos_solver* s, cosmos_solver* owner) {
    return guarded(s, [&] {
        HIPCHK(s, hipStreamSynchronize(s->stream));
        if (s->own_stream) HIPCHK(s, hipStreamDestroy(s->stream));
        s->stream = owner->stream;
        s->own_stream = false;
        return COSMOS_OK;
    });
}

}  // namespace cfv

extern "C" {

// EulerSolvers::computeStableTimeStep (Solvers/EulerSolvers.cpp:130-145)
static double stable_dt(double smax, double cell_spacing, double cfl, double current_time, double final_time,
                        int current_iteration) {
    double dt = cfl * (cell_spacing / smax);
    if (current_iteration <= 5) dt *= 0.2;
    if ((current_time + dt) > final_time) dt = final_time - current_time;
    return dt;
}

static int max_ws_impl(cosmos_solver* s, unsigned long long* slots_dev) {
    int rc = settle(s);
    if (rc) return rc;
    hipError_t e = cfv::launch_zero_slots(slots_dev, s->stream);
    if (e == hipSuccess && s->physics == COSMOS_BLACKHOLE && (rc = gr_excised_status(s))) return rc;
    if (e == hipSuccess) {
        if (s->physics == COSMOS_BLACKHOLE)
            e = cfv::launch_gr_wavespeed(gr_launch(s, 0.0), slots_dev, s->stream);
        else
            e = cfv::launch_wavespeed(cplanes_of(s, s->cur), s->nx, s->ny, s->pitch, s->physics, s->dim, s->P,
                                      slots_dev, s->stream);
    }
    if (e != hipSuccess) return fail(s, COSMOS_ERR_DEVICE, std::string("wave speed launch: ") + hipGetErrorString(e));
    return COSMOS_OK;
}

int cosmos_max_wave_speed(cosmos_solver* s, double* out) {
    return guarded(s, [&] {
        if (!out) return fail(s, COSMOS_ERR_ARGUMENT, "out is null");
        int rc = max_ws_impl(s, s->ws_scratch);
        if (rc) return rc;
        HIPCHK(s, hipMemcpyAsync(s->ws_host, s->ws_scratch, sizeof(unsigned long long) * cfv::kSlotWords,
                                 hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        *out = cfv::slots_max_host(s->ws_host);
        return COSMOS_OK;
    });
}

int cosmos_stable_time_step(cosmos_solver* s, double cell_spacing, double cfl, double current_time,
                            double final_time, int current_iteration, double* out) {
    return guarded(s, [&] {
        if (!out) return fail(s, COSMOS_ERR_ARGUMENT, "out is null");
        double smax = 0.0;
        int rc = cosmos_max_wave_speed(s, &smax);
        if (rc) return rc;
        *out = stable_dt(smax, cell_spacing, cfl, current_time, final_time, current_iteration);
        return COSMOS_OK;
    });
}

}  // extern "C"

namespace cfv {

// cosmos_stable_time_step of a masked Newtonian or special-relativistic AMR
// level, the same time step from fewer reads: full reads every cell and keeps
// the maximum over the unflagged ones in rest (kSlotWords device words); otherwise only
// the flagged cells of the n_segs listed 64-cell segments are read and rest
// joins them, which is the full maximum while no unflagged cell has changed
// since the full pass (between the masked sub-steps of one parent step).
int amr_stable_time_step(cosmos_solver* s, double cell_spacing, double cfl, double current_time, double final_time,
                         int current_iteration, const unsigned char* mask, const int* segs, int n_segs,
                         unsigned long long* rest, int full, double* out) {
    return guarded(s, [&] {
        if (!out) return fail(s, COSMOS_ERR_ARGUMENT, "out is null");
        if (s->physics == COSMOS_BLACKHOLE) return fail(s, COSMOS_ERR_ARGUMENT, "black-hole levels are unmasked");
        int rc = settle(s);
        if (rc) return rc;
        hipError_t e = launch_zero_slots(s->ws_scratch, s->stream);
        if (e == hipSuccess && full) e = launch_zero_slots(rest, s->stream);
        const ConstPlanes in = cplanes_of(s, s->cur);
        if (e == hipSuccess) {
            e = full ? launch_wavespeed_split(in, s->nx, s->ny, s->pitch, s->physics, s->dim, s->P, mask,
                                              s->ws_scratch, rest, s->stream)
                     : launch_wavespeed_segs(in, s->nx, s->ny, s->pitch, s->physics, s->dim, s->P, mask, segs,
                                             n_segs, rest, s->ws_scratch, s->stream);
        }
        if (e != hipSuccess)
            return fail(s, COSMOS_ERR_DEVICE, std::string("wave speed launch: ") + hipGetErrorString(e));
        HIPCHK(s, hipMemcpyAsync(s->ws_host, s->ws_scratch, sizeof(unsigned long long) * kSlotWords,
                                 hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        *out = stable_dt(slots_max_host(s->ws_host), cell_spacing, cfl, current_time, final_time, current_iteration);
        return COSMOS_OK;
    });
}

}  // namespace cfv

extern "C" {

int cosmos_begin(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time, int max_steps,
                 double bias, int slope_limiter) {
    return cosmos_begin_at(s, order, cell_spacing, cfl, final_time, max_steps, bias, slope_limiter, 0.0, 0);
}

int cosmos_begin_at(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time, int max_steps,
                    double bias, int slope_limiter, double start_time, int start_iteration) {
    return guarded(s, [&] {
        if (start_iteration < 0) return fail(s, COSMOS_ERR_ARGUMENT, "start_iteration must be >= 0");
        int rc = check_order(s, order);
        if (rc) return rc;
        if (s->physics == COSMOS_BLACKHOLE)
            return fail(s, COSMOS_ERR_ARGUMENT, "black-hole solvers step with cosmos_solve_forced");
        // the steps queued by an earlier run decide which buffer holds the state;
        // read that before the new run's StepState replaces their count
        rc = settle(s);
        if (rc) return rc;
        StepState h{};
        h.t = start_time;
        h.tf = final_time;
        h.cfl = cfl;
        h.dx = cell_spacing;
        h.iter = start_iteration;
        h.max_steps = max_steps < 0 || max_steps > INT_MAX - start_iteration ? INT_MAX : start_iteration + max_steps;
        h.done = 0;
        h.active_steps = 0;
        h.ws_slots = s->st_slots;
        h.ws_part = s->st_part;
        h.ws_nparts = 0;
        h.ws_slot_atomics = s->ws_mode != 1;
        h.status = s->status;
        h.ws_nparts_sweep = h.ws_slot_atomics ? 0 : cfv::sweep_x_blocks(s->nx, s->ny, order);
        HIPCHK(s, hipMemcpyAsync(s->st, &h, sizeof h, hipMemcpyHostToDevice, s->stream));
        rc = max_ws_impl(s, s->st_slots);
        if (rc) return rc;
        if ((rc = ensure_hot_order(s, order))) return rc;
        s->run_order = order;
        s->run_bias = bias;
        s->run_limiter = slope_limiter;
        s->run_ready = true;
        s->cur_begin = s->cur;
        s->queued = 0;
        s->run_flips = 0;
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return COSMOS_OK;
    });
}

int cosmos_advance(cosmos_solver* s, int n_steps) {
    return guarded(s, [&] {
        if (!s->run_ready) return fail(s, COSMOS_ERR_STATE, "cosmos_advance before cosmos_begin");
        for (int k = 0; k < n_steps; ++k) {
            int rc = queue_step(s);
            if (rc) return rc;
        }
        return COSMOS_OK;
    });
}

int cosmos_progress(cosmos_solver* s, double* time, int* steps, int* stopped) {
    return guarded(s, [&] {
        StepState h{};
        HIPCHK(s, hipMemcpyAsync(&h, s->st, sizeof h, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        if (time) *time = h.t;
        if (steps) *steps = h.active_steps;
        if (stopped) *stopped = h.done || !(h.t < h.tf) || h.iter >= h.max_steps;
        return COSMOS_OK;
    });
}

int cosmos_solve(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time, int max_steps,
                 double bias, int slope_limiter, double* time_out, int* steps_out) {
    return guarded(s, [&] {
        int rc = cosmos_begin(s, order, cell_spacing, cfl, final_time, max_steps, bias, slope_limiter);
        if (rc) return rc;
        int stopped = 0, steps = 0;
        double t = 0.0;
        const int chunk = max_steps >= 0 && max_steps < 64 ? (max_steps > 0 ? max_steps : 1) : 64;
        while (true) {
            rc = cosmos_advance(s, chunk);
            if (rc) return rc;
            rc = cosmos_progress(s, &t, &steps, &stopped);
            if (rc) return rc;
            if (stopped) break;
        }
        if (time_out) *time_out = t;
        if (steps_out) *steps_out = steps;
        return COSMOS_OK;
    });
}

int cosmos_forcing(cosmos_solver* s, double cell_spacing, double time_step, double bias, int slope_limiter) {
    return guarded(s, [&] {
        if (s->physics != COSMOS_BLACKHOLE)
            return fail(s, COSMOS_ERR_ARGUMENT, "cosmos_forcing needs a COSMOS_BLACKHOLE solver");
        int rc = settle(s);
        if (rc) return rc;
        s->run_ready = false;
        return gr_forcing(s, cell_spacing, time_step, bias, slope_limiter);
    });
}

int cosmos_solve_forced(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time,
                        int max_steps, double bias, int slope_limiter, int subcycles, double* time_out,
                        int* steps_out) {
    return cosmos_solve_forced_at(s, order, cell_spacing, cfl, final_time, max_steps, bias, slope_limiter, subcycles,
                                  0.0, 0, time_out, steps_out);
}

int cosmos_solve_forced_at(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time,
                           int max_steps, double bias, int slope_limiter, int subcycles, double start_time,
                           int start_iteration, double* time_out, int* steps_out) {
    return guarded(s, [&] {
        if (s->physics != COSMOS_BLACKHOLE)
            return fail(s, COSMOS_ERR_ARGUMENT, "cosmos_solve_forced needs a COSMOS_BLACKHOLE solver");
        int rc = check_order(s, order);
        if (rc) return rc;
        if (subcycles < 0) return fail(s, COSMOS_ERR_ARGUMENT, "subcycles must be >= 0");
        if (start_iteration < 0) return fail(s, COSMOS_ERR_ARGUMENT, "start_iteration must be >= 0");
        const int cap = max_steps < 0 || max_steps > INT_MAX - start_iteration ? INT_MAX : start_iteration + max_steps;
        if (!(start_time < final_time) || start_iteration >= cap) {
            if (time_out) *time_out = start_time;
            if (steps_out) *steps_out = 0;
            return COSMOS_OK;
        }
        if ((rc = settle(s))) return rc;
        if ((rc = ensure_gr_flux(s))) return rc;
        cfv::GeoTab tab;
        if ((rc = ensure_geo(s, cell_spacing, tab))) return rc;
        // the unit-square wave speed positions of every step are the same: a
        // fully excised grid is found once (status word 6)
        if ((rc = gr_excised_status(s))) return rc;
        StepState h{};
        h.t = start_time;
        h.tf = final_time;
        h.cfl = cfl;
        h.dx = cell_spacing;
        h.iter = start_iteration;
        h.max_steps = cap;
        h.ws_slots = s->st_slots;
        h.ws_slot_atomics = 1;
        h.status = s->status;
        h.gr = 1;
        h.gr_subcycles = subcycles;
        h.gr_dim = s->dim;
        HIPCHK(s, hipMemcpyAsync(s->st, &h, sizeof h, hipMemcpyHostToDevice, s->stream));
        HIPCHK(s, cfv::launch_zero_slots(s->st_slots, s->stream));
        s->run_ready = false;
        s->cur_begin = s->cur;
        s->queued = 0;
        s->run_flips = 2 * subcycles + sweeps_per_step(s);
        HIPCHK(s, hipStreamSynchronize(s->stream));  // h is read before it leaves scope
        // steps are queued in chunks of at most 16 (never past the step cap); the
        // loop's own stop test (t >= final_time, a NaN time step) turns the rest
        // of a chunk into launches that return at once
        long left = (long)cap - start_iteration;
        StepState now{};
        while (true) {
            const long chunk = left < 16 ? left : 16;
            for (long k = 0; k < chunk; ++k)
                if ((rc = queue_gr_step(s, order, cell_spacing, bias, slope_limiter, subcycles, tab))) return rc;
            left -= chunk;
            HIPCHK(s, hipMemcpyAsync(&now, s->st, sizeof now, hipMemcpyDeviceToHost, s->stream));
            HIPCHK(s, hipStreamSynchronize(s->stream));
            if (now.done || !(now.t < now.tf) || now.iter >= now.max_steps || left <= 0) break;
        }
        if ((rc = settle(s))) return rc;
        if (time_out) *time_out = now.t;
        if (steps_out) *steps_out = now.active_steps;
        return COSMOS_OK;
    });
}

int cosmos_status_async(cosmos_solver* s, cosmos_status_t* out) {
    return guarded(s, [&] {
        if (!out) return fail(s, COSMOS_ERR_ARGUMENT, "out is null");
        static_assert(sizeof(cosmos_status_t) == sizeof(unsigned long long) * kStatusWords, "status layout");
        int rc = settle(s);
        if (rc) return rc;
        HIPCHK(s, cfv::launch_fill_u64(s->status + 4, 1, 0ull, s->stream));
        HIPCHK(s, cfv::launch_nan_cells(cplanes_of(s, s->cur), s->nx, s->ny, s->pitch, s->status + 4, s->stream));
        HIPCHK(s, hipMemcpyAsync(out, s->status, sizeof(cosmos_status_t), hipMemcpyDeviceToHost, s->stream));
        return COSMOS_OK;
    });
}

int cosmos_status(cosmos_solver* s, cosmos_status_t* out) {
    return guarded(s, [&] {
        int rc = cosmos_status_async(s, out);
        if (rc) return rc;
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return COSMOS_OK;
    });
}

int cosmos_status_reset(cosmos_solver* s) {
    return guarded(s, [&] {
        HIPCHK(s, cfv::launch_fill_u64(s->status, kStatusWords, 0ull, s->stream));
        return COSMOS_OK;
    });
}

int cosmos_synchronize(cosmos_solver* s) {
    return guarded(s, [&] {
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return COSMOS_OK;
    });
}

void* cosmos_stream(cosmos_solver* s) { return s ? (void*)s->stream : nullptr; }

int cosmos_profile_enable(cosmos_solver* s, int on) {
    return guarded(s, [&] {
        s->profiling = on != 0;
        return COSMOS_OK;
    });
}

int cosmos_profile_read(cosmos_solver* s, double* total_ms, int* launches) {
    return guarded(s, [&] {
        if (!total_ms || !launches) return fail(s, COSMOS_ERR_ARGUMENT, "null output");
        HIPCHK(s, hipStreamSynchronize(s->stream));
        for (int k = 0; k < kProfKinds; ++k) {
            total_ms[k] = 0.0;
            launches[k] = 0;
        }
        for (auto& p : s->prof) {
            float ms = 0.0f;
            HIPCHK(s, hipEventElapsedTime(&ms, p.second.first, p.second.second));
            total_ms[p.first] += ms;
            launches[p.first] += 1;
            hipEventDestroy(p.second.first);
            hipEventDestroy(p.second.second);
        }
        s->prof.clear();
        return COSMOS_OK;
    });
}

namespace {
// A plain streaming copy (16-byte vectors), used to calibrate bandwidth and
// hardware byte counters against a kernel of known traffic.
// Each thread moves four 16-byte vectors, blockDim apart (coalesced), all
// loads issued before the stores.
// One 16-byte load and store per thread over as many blocks as that takes:
// of the streaming copies timed on MI355X (1 to 8 loads in flight per thread,
// non-temporal forms, grid-stride loops, hipMemcpyAsync) this one moves the
// most bytes per second, about 6.1-6.2 TB/s for 320 MiB to 1 GiB each way,
// against 6.1-6.4 TB/s for a read-only or write-only stream.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, long n2) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n2) dst[i] = src[i];
}
}  // namespace

int cosmos_device_copy(void* dst, const void* src, long bytes, void* stream) {
    if (!dst || !src || bytes <= 0 || (bytes % 16) != 0 || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
        return fail(nullptr, COSMOS_ERR_ARGUMENT, "copy needs 16-byte aligned pointers and a multiple of 16 bytes");
    long n2 = bytes / 16;
    long blocks = (n2 + 255) / 256;
    if (blocks > 0x7fffffffL) return fail(nullptr, COSMOS_ERR_ARGUMENT, "copy too large");
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (const double2*)src,
                       (double2*)dst, n2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, COSMOS_ERR_DEVICE, std::string("copy launch: ") + hipGetErrorString(e));
    return COSMOS_OK;
}

int cosmos_state_planes(cosmos_solver* s, double** planes5, long* pitch) {
    return guarded(s, [&] {
        if (!planes5) return fail(s, COSMOS_ERR_ARGUMENT, "planes5 is null");
        int rc = settle(s);
        if (rc) return rc;
        Planes p = planes_of(s, s->cur);
        for (int k = 0; k < 5; ++k) planes5[k] = p.p[k];
        if (pitch) *pitch = s->pitch;
        return COSMOS_OK;
    });
}

// ---------------------------------------------------------------------------
// Lossless checkpoint / restart and the reference's .dat text output
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

// Checkpoint file: a 128-byte little-endian header, the state as AoS
// primitives [ny][nx][5] FP64 (the cosmos_download layout), and the 64-bit
// FNV-1a hash of those bytes.
constexpr char kCkMagic[8] = {'C', 'O', 'S', 'M', 'O', 'S', 'C', 'K'};
constexpr uint32_t kCkVersion = 1;

struct CkHeader {
    char magic[8];
    uint32_t version;
    int32_t physics, dim, nx, ny, iteration;
    double gamma, time;
    double hole[5];   // M, a, x, y, z (black hole)
    double excision;  // explicit excision radius (0: the horizon)
    uint8_t pad[128 - 8 - 4 - 5 * 4 - 4 - 2 * 8 - 5 * 8 - 8];
};
static_assert(sizeof(CkHeader) == 128, "checkpoint header is 128 bytes");

uint64_t fnv1a(const unsigned char* p, size_t n) {
    uint64_t h = 1469598103934665603ULL;
    for (size_t k = 0; k < n; ++k) {
        h ^= p[k];
        h *= 1099511628211ULL;
    }
    return h;
}

}  // namespace

extern "C" {

int cosmos_checkpoint_save(cosmos_solver* s, const char* path, double time, int iteration) {
    return guarded(s, [&] {
        if (!path) return fail(s, COSMOS_ERR_ARGUMENT, "path is null");
        const size_t n = (size_t)s->nx * s->ny * 5;
        std::vector<double> state(n);
        int rc = download_impl(s, state.data(), hipMemcpyDeviceToHost);
        if (rc) return rc;
        CkHeader h{};
        std::memcpy(h.magic, kCkMagic, 8);
        h.version = kCkVersion;
        h.physics = s->physics;
        h.dim = s->dim;
        h.nx = s->nx;
        h.ny = s->ny;
        h.iteration = iteration;
        h.gamma = s->P.gamma;
        h.time = time;
        const double hole[5] = {s->hole.M, s->hole.a, s->hole.x, s->hole.y, s->hole.z};
        std::memcpy(h.hole, hole, sizeof hole);
        h.excision = s->excision;
        const uint64_t sum = fnv1a(reinterpret_cast<const unsigned char*>(state.data()), n * sizeof(double));
        const std::string tmp = std::string(path) + ".tmp";
        FILE* f = std::fopen(tmp.c_str(), "wb");
        if (!f) return fail(s, COSMOS_ERR_ARGUMENT, std::string("cannot open ") + tmp);
        bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(state.data(), sizeof(double), n, f) == n &&
                  std::fwrite(&sum, sizeof sum, 1, f) == 1;
        ok = (std::fclose(f) == 0) && ok;
        if (!ok || std::rename(tmp.c_str(), path) != 0) {
            std::remove(tmp.c_str());
            return fail(s, COSMOS_ERR_STATE, std::string("writing the checkpoint ") + path + " failed");
        }
        return COSMOS_OK;
    });
}

int cosmos_checkpoint_load(cosmos_solver* s, const char* path, double* time, int* iteration) {
    return guarded(s, [&] {
        if (!path) return fail(s, COSMOS_ERR_ARGUMENT, "path is null");
        FILE* f = std::fopen(path, "rb");
        if (!f) return fail(s, COSMOS_ERR_ARGUMENT, std::string("cannot open ") + path);
        CkHeader h{};
        const size_t n = (size_t)s->nx * s->ny * 5;
        std::vector<double> state(n);
        uint64_t sum = 0;
        const bool read_ok = std::fread(&h, sizeof h, 1, f) == 1;
        std::string why;
        if (!read_ok || std::memcmp(h.magic, kCkMagic, 8) != 0) why = "not a checkpoint file";
        else if (h.version != kCkVersion) why = "unsupported checkpoint version " + std::to_string(h.version);
        else if (h.physics != s->physics || h.dim != s->dim || h.nx != s->nx || h.ny != s->ny)
            why = "checkpoint grid or physics differs from the solver's";
        else if (h.gamma != s->P.gamma) why = "checkpoint adiabatic index differs from the solver's";
        else if (std::fread(state.data(), sizeof(double), n, f) != n || std::fread(&sum, sizeof sum, 1, f) != 1)
            why = "checkpoint file is truncated";
        else if (sum != fnv1a(reinterpret_cast<const unsigned char*>(state.data()), n * sizeof(double)))
            why = "checkpoint checksum mismatch";
        std::fclose(f);
        if (!why.empty()) return fail(s, COSMOS_ERR_STATE, why + " (" + path + ")");
        int rc = settle(s);
        if (rc) return rc;
        if (s->physics == COSMOS_BLACKHOLE) {
            s->excision = h.excision;
            s->hole = make_hole(h.hole[0], h.hole[1], h.hole[2], h.hole[3], h.hole[4], h.excision);
        }
        rc = upload_impl(s, state.data(), hipMemcpyHostToDevice);
        if (rc) return rc;
        if (time) *time = h.time;
        if (iteration) *iteration = h.iteration;
        return COSMOS_OK;
    });
}

// EulerTests::outputSolution{,2D} (Tests/EulerTests.cpp:189-252) and
// BlackHoleEulerTests::outputSolution{,2D} (Tests/BlackHoleEulerTests.cpp:201-286):
// one file per primitive, "x value" (1D) or "y x value" rows with a blank line
// after each grid row (2D), coordinates on the unit square (1 / cell count), the
// stream's default formatting; black-hole cells inside the excision region
// (tested at (j / nx, i / ny)) written as 0.
int cosmos_write_dat(cosmos_solver* s, const char* directory) {
    return guarded(s, [&] {
        if (!directory) return fail(s, COSMOS_ERR_ARGUMENT, "directory is null");
        const size_t n = (size_t)s->nx * s->ny * 5;
        std::vector<double> W(n);
        int rc = download_impl(s, W.data(), hipMemcpyDeviceToHost);
        if (rc) return rc;
        static const char* names[5] = {"density", "xVelocity", "yVelocity", "zVelocity", "pressure"};
        const bool bh = s->physics == COSMOS_BLACKHOLE;
        for (int k = 0; k < 5; ++k) {
            const std::string fn = std::string(directory) + "/" + names[k] + ".dat";
            std::ofstream out(fn);
            if (!out) return fail(s, COSMOS_ERR_ARGUMENT, "cannot open " + fn);
            if (s->dim == 1) {
                const double h = 1.0 / s->nx;
                for (int i = 0; i < s->nx; ++i) {
                    const bool cut = bh && cfv::excised(s->hole, i * h, 0.0, 0.0);
                    out << (h * i) << " " << (cut ? 0.0 : W[5 * (size_t)i + k]) << std::endl;
                }
            } else {
                const double hr = 1.0 / s->ny, hc = 1.0 / s->nx;
                for (int i = 0; i < s->ny; ++i) {
                    for (int j = 0; j < s->nx; ++j) {
                        const bool cut = bh && cfv::excised(s->hole, j * hc, i * hr, 0.0);
                        out << (hr * i) << " " << (hc * j) << " "
                            << (cut ? 0.0 : W[5 * ((size_t)i * s->nx + j) + k]) << std::endl;
                    }
                    out << std::endl;
                }
            }
            if (!out) return fail(s, COSMOS_ERR_STATE, "writing " + fn + " failed");
        }
        return COSMOS_OK;
    });
}

}  // extern "C"
