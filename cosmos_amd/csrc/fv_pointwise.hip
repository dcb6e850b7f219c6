// Batched point-wise entry points: the slope limiter and the limited slope
// vector of SlopeLimiters (Solvers/SlopeLimiters.cpp:3-101) evaluated on the
// device for n independent inputs.  They serve the drop-in SlopeLimiters
// class and table checks; the sweeps inline the same device functions.
#include <string>

#include "../../include/cosmos_fv.h"
#include "fv_physics.hpp"

namespace {

__global__ void k_limiter(const double* r, long n, double bias, int limiter, double* out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cfv::limiter_general(r[i], bias, limiter);
}

__global__ void k_slope(const double* ul, const double* um, const double* ur, long n, double bias, int limiter,
                        double* out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 5 * n) out[i] = cfv::slope_component<0>(ul[i], um[i], ur[i], bias, limiter);
}

struct DeviceBuffer {
    double* p = nullptr;
    ~DeviceBuffer() {
        if (p) hipFree(p);
    }
};

}  // namespace

extern "C" {

int cosmos_slope_limiter(const double* steepness, long n, double bias, int slope_limiter, double* out) {
    if (!steepness || !out || n < 0) return COSMOS_ERR_ARGUMENT;
    if (n == 0) return COSMOS_OK;
    DeviceBuffer d;
    size_t bytes = sizeof(double) * (size_t)n;
    if (hipMalloc((void**)&d.p, 2 * bytes) != hipSuccess) return COSMOS_ERR_DEVICE;
    if (hipMemcpy(d.p, steepness, bytes, hipMemcpyHostToDevice) != hipSuccess) return COSMOS_ERR_DEVICE;
    hipLaunchKernelGGL(k_limiter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, d.p, n, bias, slope_limiter,
                       d.p + n);
    if (hipGetLastError() != hipSuccess) return COSMOS_ERR_DEVICE;
    if (hipMemcpy(out, d.p + n, bytes, hipMemcpyDeviceToHost) != hipSuccess) return COSMOS_ERR_DEVICE;
    return COSMOS_OK;
}

int cosmos_slope_vector(const double* left, const double* middle, const double* right, long n, double bias,
                        int slope_limiter, double* out) {
    if (!left || !middle || !right || !out || n < 0) return COSMOS_ERR_ARGUMENT;
    if (n == 0) return COSMOS_OK;
    DeviceBuffer d;
    size_t bytes = sizeof(double) * 5 * (size_t)n;
    if (hipMalloc((void**)&d.p, 4 * bytes) != hipSuccess) return COSMOS_ERR_DEVICE;
    const size_t m = 5 * (size_t)n;
    if (hipMemcpy(d.p, left, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.p + m, middle, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.p + 2 * m, right, bytes, hipMemcpyHostToDevice) != hipSuccess)
        return COSMOS_ERR_DEVICE;
    hipLaunchKernelGGL(k_slope, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, 0, d.p, d.p + m, d.p + 2 * m, (long)n,
                       bias, slope_limiter, d.p + 3 * m);
    if (hipGetLastError() != hipSuccess) return COSMOS_ERR_DEVICE;
    if (hipMemcpy(out, d.p + 3 * m, bytes, hipMemcpyDeviceToHost) != hipSuccess) return COSMOS_ERR_DEVICE;
    return COSMOS_OK;
}

}  // extern "C"
