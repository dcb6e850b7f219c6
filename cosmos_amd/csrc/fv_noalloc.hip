// cosmos_step_noalloc: one step of the solve loop on caller-owned device
// planes that allocates nothing (no hipMalloc, hipFree or allocating library
// call), for callers that cannot allocate: code under HIP graph capture and
// tight serving loops.
//
// The step is the one cosmos_solve / cosmos_advance queue (the reference's
// solve / solve2D iteration, Solvers/EulerSecondOrderSolver.cpp:133-178 and
// Solvers/EulerFirstOrderSolver.cpp:220-265): the maximum wave speed of the
// input, the time step with the six-step damping and the final-time clamp,
// then X(dt/2) Y(dt) X(dt/2) in 2D or one sweep in 1D, in exact arithmetic.
// Its device working memory (the step state, the wave-speed slots and, in
// 2D, the y sweep's output grid) lives in `scratch`, a file-scope device
// array of kScratchBytes that every GPU the library runs on holds until the
// process ends.  Its address on each device comes from hipGetSymbolAddress,
// taken once per device on the first call there; host code uses only that
// address, never the array's name.  Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU (on different streams or threads) race.
#include <mutex>
#include <string>

#include "../../include/cosmos_fv.h"
#include "fv_kernels.hpp"

namespace cfv {
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr size_t kStepOffset = 0;       // StepState
constexpr size_t kSlotsOffset = 256;    // kSlotWords wave-speed slots
constexpr size_t kGridOffset = 8192;    // 2D: five planes of the y sweep's output
constexpr size_t kGridDoubles = (kScratchBytes - kGridOffset) / sizeof(double);
static_assert(sizeof(StepState) <= kSlotsOffset - kStepOffset, "StepState must fit its scratch slot");
static_assert(kSlotsOffset + sizeof(unsigned long long) * kSlotWords <= kGridOffset, "slots must fit their slot");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};
thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

// The scratch array's device address on the current device (taken once).
int scratch_of(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(COSMOS_ERR_DEVICE, "device index beyond the supported 64");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(COSMOS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return COSMOS_OK;
}

// The step state from its initial value (a kernel argument: no host copy),
// and the wave-speed slots cleared (a fill, not a memset node).
__global__ __launch_bounds__(256) void k_noalloc_head(StepState* st, StepState init, unsigned long long* slots) {
    if (threadIdx.x == 0) *st = init;
    for (int k = threadIdx.x; k < kSlotWords; k += blockDim.x) slots[k] = 0ull;
}

__global__ void k_noalloc_dt(const StepState* st, double* out) {
    if (threadIdx.x == 0) *out = st->dt;
}

}  // namespace
}  // namespace cfv

using cfv::StepState;
using cfv::SweepLaunch;

extern "C" const char* cosmos_noalloc_last_error(void) { return cfv::g_error.c_str(); }

extern "C" size_t cosmos_noalloc_grid_limit(void) { return cfv::kGridDoubles; }

extern "C" int cosmos_step_noalloc(int physics, int dim, int nx, int ny, long pitch, const double* in, double* out,
                                   int order, double cell_spacing, double cfl, double current_time,
                                   double final_time, int current_iteration, double bias, int slope_limiter,
                                   double adiabatic_index, double* time_step_out, void* stream) {
    using namespace cfv;
    try {
        if (physics != COSMOS_NEWTONIAN && physics != COSMOS_RELATIVISTIC)
            return fail(COSMOS_ERR_ARGUMENT, "physics must be COSMOS_NEWTONIAN or COSMOS_RELATIVISTIC");
        if (dim != 1 && dim != 2) return fail(COSMOS_ERR_ARGUMENT, "dim must be 1 or 2");
        if (dim == 1) ny = 1;
        if (nx < 1 || ny < 1 || pitch < nx) return fail(COSMOS_ERR_ARGUMENT, "need nx, ny >= 1 and pitch >= nx");
        if (order != 1 && order != 2) return fail(COSMOS_ERR_ARGUMENT, "order must be 1 (FORCE) or 2 (SLIC)");
        if (order == 2 && (nx < 2 || (dim == 2 && ny < 2)))
            return fail(COSMOS_ERR_ARGUMENT, "SLIC needs at least 2 cells along each swept axis");
        if (!in || !out || in == out) return fail(COSMOS_ERR_ARGUMENT, "need distinct input and output planes");
        if (current_iteration < 0) return fail(COSMOS_ERR_ARGUMENT, "current_iteration must be >= 0");
        if (!(current_time < final_time))
            return fail(COSMOS_ERR_ARGUMENT, "current_time must be below final_time (the loop takes no step)");
        const size_t plane = (size_t)pitch * (size_t)ny;
        if (dim == 2 && 5 * plane > kGridDoubles)
            return fail(COSMOS_ERR_ARGUMENT, "grid of " + std::to_string(5 * plane) + " doubles (5 x pitch x ny) "
                                             "exceeds the no-allocation limit of " + std::to_string(kGridDoubles));
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return fail(COSMOS_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
        unsigned char* base = nullptr;
        int rc = scratch_of(dev, &base);
        if (rc) return rc;
        hipStream_t s = static_cast<hipStream_t>(stream);
        StepState* st = reinterpret_cast<StepState*>(base + kStepOffset);
        unsigned long long* slots = reinterpret_cast<unsigned long long*>(base + kSlotsOffset);
        double* grid = reinterpret_cast<double*>(base + kGridOffset);

        StepState h{};
        h.t = current_time;
        h.tf = final_time;
        h.cfl = cfl;
        h.dx = cell_spacing;
        h.iter = current_iteration;
        h.max_steps = current_iteration + 1;
        h.ws_slots = slots;
        h.ws_part = nullptr;
        h.ws_nparts = 0;
        h.ws_nparts_sweep = 0;
        h.ws_slot_atomics = 1;
        const PhysParams P = make_params(adiabatic_index);
        hipLaunchKernelGGL(k_noalloc_head, dim3(1), dim3(256), 0, s, st, h, slots);
        e = hipGetLastError();
        ConstPlanes cin;
        Planes pout, pgrid;
        ConstPlanes cout_, cgrid;
        for (int k = 0; k < 5; ++k) {
            cin.p[k] = in + k * plane;
            pout.p[k] = out + k * plane;
            cout_.p[k] = out + k * plane;
            pgrid.p[k] = grid + k * plane;
            cgrid.p[k] = grid + k * plane;
        }
        if (e == hipSuccess) e = launch_wavespeed(cin, nx, ny, pitch, physics, dim, P, slots, s);
        if (e == hipSuccess) e = launch_step_begin(st, s);
        SweepLaunch L{};
        L.nx = nx;
        L.ny = ny;
        L.pitch = pitch;
        L.bias = bias;
        L.limiter = slope_limiter;
        L.order = order;
        L.physics = physics;
        L.dim = dim;
        L.P = P;
        L.step = st;
        L.y_tile = -1;
        if (dim == 1) {
            L.in = cin;
            L.out = pout;
            L.use_full = 1;
            if (e == hipSuccess) e = launch_sweep_x(L, s);
        } else {
            L.in = cin;
            L.out = pout;
            L.use_full = 0;
            if (e == hipSuccess) e = launch_sweep_x(L, s);
            L.in = cout_;
            L.out = pgrid;
            L.use_full = 1;
            if (e == hipSuccess) e = launch_sweep_y(L, s);
            L.in = cgrid;
            L.out = pout;
            L.use_full = 0;
            if (e == hipSuccess) e = launch_sweep_x(L, s);
        }
        if (e == hipSuccess && time_step_out) {
            hipLaunchKernelGGL(k_noalloc_dt, dim3(1), dim3(64), 0, s, st, time_step_out);
            e = hipGetLastError();
        }
        if (e != hipSuccess) return fail(COSMOS_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e));
        return COSMOS_OK;
    } catch (...) {
        return fail(COSMOS_ERR_STATE, "unexpected exception");
    }
}
