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
#include "fv_kernels.hpp"

namespace cfv {
namespace {

constexpr int kGRThreads = 256;

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

// position (x, y) moved by off along the sweep axis
template <int AXIS>
__device__ __forceinline__ Geo geo_off(const Hole& b, double x, double y, double off) {
    return AXIS == 0 ? geo_at(b, x + off, y, 0.0) : geo_at(b, x, y + off, 0.0);
}

// Limited slope of the conserved vectors (SlopeLimiters::computeSlopeVector{X,Y},
// Solvers/SlopeLimiters.cpp:123-147, on top of computeSlopeVector :71-101).
__device__ __forceinline__ void slope5(const double* ul, const double* um, const double* ur, double bias, int limiter,
                                       double* s) {
    if (bias == 0.0 && limiter == 1) {
#pragma unroll
        for (int k = 0; k < 5; ++k) s[k] = slope_component<1>(ul[k], um[k], ur[k], bias, limiter);
    } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) s[k] = slope_component<0>(ul[k], um[k], ur[k], bias, limiter);
    }
}

// evolveStateByHalf{X,Y}TimeStep (Solvers/BlackHoleEulerSolvers.cpp:219-307) of the
// middle cell centred at (x, y): slope from p2c at x - h, x, x + h; face fluxes
// at x -/+ h/2; the selected face evolved by 0.5 (dt/h) (F_L - F_R) and
// converted back at that face.
template <int AXIS>
__device__ void gr_half_step(const double* WL, const double* WM, const double* WR, double x, double y, int side,
                             const GRLaunch& L, double* out) {
    const Hole& b = L.hole;
    const double h = L.h;
    double ul[5], um[5], ur[5], s[5];
    {
        Geo g = geo_off<AXIS>(b, x, y, -h);
        gr_p2c(WL, g, ul, L.P);
    }
    {
        Geo g = geo_at(b, x, y, 0.0);
        gr_p2c(WM, g, um, L.P);
    }
    {
        Geo g = geo_off<AXIS>(b, x, y, h);
        gr_p2c(WR, g, ur, L.P);
    }
    slope5(ul, um, ur, L.bias, L.limiter, s);
    double xl[5], xr[5], fl[5], fr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    const double hh = h * 0.5;
    Geo gl = geo_off<AXIS>(b, x, y, -hh);
    gr_flux<AXIS>(xl, gl, L.P, fl);
    Geo gr = geo_off<AXIS>(b, x, y, hh);
    gr_flux<AXIS>(xr, gr, L.P, fr);
    double e[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) e[k] = (side == 0 ? xl[k] : xr[k]) + (L.coef.hdt_dx * (fl[k] - fr[k]));
    gr_c2p(e, side == 0 ? gl.sdet : gr.sdet, L.P, out);
}

// compute{X,Y}FORCEFlux (Solvers/BlackHoleEulerFirstOrderSolver.cpp:3-99) at the
// interface (x, y): the states' p2c and fluxes at the interface -/+ h/2, the
// Richtmyer state's flux at the interface.
template <int AXIS>
__device__ void gr_force(const double* WL, const double* WR, double x, double y, const GRLaunch& L, double* F) {
    const Hole& b = L.hole;
    const double hh = L.h * 0.5;
    double ul[5], ur[5], fl[5], fr[5], us[5], fs[5];
    {
        Geo g = geo_off<AXIS>(b, x, y, -hh);
        gr_p2c(WL, g, ul, L.P);
        gr_flux<AXIS>(ul, g, L.P, fl);
    }
    {
        Geo g = geo_off<AXIS>(b, x, y, hh);
        gr_p2c(WR, g, ur, L.P);
        gr_flux<AXIS>(ur, g, L.P, fr);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) us[k] = 0.5 * ((ul[k] + ur[k]) + (L.coef.dt_dx * (fl[k] - fr[k])));
    {
        Geo g = geo_at(b, x, y, 0.0);
        gr_flux<AXIS>(us, g, L.P, fs);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double lf = 0.5 * ((fl[k] + fr[k]) + (L.coef.dx_dt * (ul[k] - ur[k])));
        F[k] = 0.5 * (lf + fs[k]);
    }
}

// One flux per interface.  x sweep: thread -> (row i, interface c), c in [0, nx];
// y sweep: thread -> (interface c, column j), c in [0, ny].  Interface c lies
// between cells c - 1 and c at (c - 0.5) * h.
template <int ORDER, int AXIS>
__global__ __launch_bounds__(kGRThreads) void k_gr_iface(GRLaunch L) {
    const int n = AXIS == 0 ? L.nx : L.ny;
    const long lines = AXIS == 0 ? L.ny : L.nx;
    const long per = AXIS == 0 ? (long)(n + 1) : lines;
    const long total = (long)(n + 1) * lines;
    long idx = (long)blockIdx.x * kGRThreads + threadIdx.x;
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
    if (AXIS == 0) {
        x = (c - 0.5) * L.h;
        y = L.dim == 2 ? i * L.h : 0.0;
    } else {
        x = j * L.h;
        y = (c - 0.5) * L.h;
    }
    double F[5];
    if (ORDER == 1) {
        double WL[5], WR[5];
        load5(L.in, at(c - 1), WL);
        load5(L.in, at(c), WR);
        gr_force<AXIS>(WL, WR, x, y, L, F);
    } else {
        // computeXSLICFlux (Solvers/BlackHoleEulerSecondOrderSolver.cpp:3-35): the left
        // cell's right-face state evolved about the interface - h/2, the right cell's
        // left-face state about the interface + h/2, then FORCE at the interface.
        const double hh = L.h * 0.5;
        double W0[5], W1[5], W2[5], pr[5], pl[5];
        load5(L.in, at(c - 2), W0);
        load5(L.in, at(c - 1), W1);
        load5(L.in, at(c), W2);
        if (AXIS == 0) gr_half_step<AXIS>(W0, W1, W2, x - hh, y, 1, L, pr);
        else gr_half_step<AXIS>(W0, W1, W2, x, y - hh, 1, L, pr);
        load5(L.in, at(c + 1), W0);
        if (AXIS == 0) gr_half_step<AXIS>(W1, W2, W0, x + hh, y, 0, L, pl);
        else gr_half_step<AXIS>(W1, W2, W0, x, y + hh, 0, L, pl);
        gr_force<AXIS>(pr, pl, x, y, L, F);
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
template <int ORDER, int AXIS>
__global__ __launch_bounds__(kGRThreads) void k_gr_update(GRLaunch L) {
    long idx = (long)blockIdx.x * kGRThreads + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return;
    int i = (int)(idx / L.nx), j = (int)(idx - (long)i * L.nx);
    long cell = (long)i * L.pitch + j;
    double W[5];
    load5(L.in, cell, W);
    if (stencil_clear(L.hole, L.dim, i, j, ORDER == 2 ? 2 : 1, L.h)) {
        Geo g = geo_at(L.hole, j * L.h, L.dim == 2 ? i * L.h : 0.0, 0.0);
        double U[5], un[5];
        gr_p2c(W, g, U, L.P);
        // x: interfaces j, j + 1 of row i; y: interfaces i, i + 1 of column j
        long fl = (long)i * L.fpitch + j;
        long fr = AXIS == 0 ? fl + 1 : fl + L.fpitch;
#pragma unroll
        for (int k = 0; k < 5; ++k) un[k] = U[k] + (L.coef.dt_dx * (L.flux.p[k][fl] - L.flux.p[k][fr]));
        gr_c2p(un, g.sdet, L.P, W);
    }
    store5(L.out, cell, W);
}

// evolveStateByFractional{X,Y}TimeStep (Solvers/BlackHoleEulerSolvers.cpp:171-217):
// slope from p2c at x - h, x, x + h; face fluxes at x -/+ h (a whole cell,
// as in the reference); U_M + frac (dt/h) (F_L - F_R) converted back at x.
template <int AXIS>
__device__ void gr_fractional(double fdt_h, const double* WL, const double* WM, const double* WR, double x, double y,
                              const GRLaunch& L, double* out) {
    const Hole& b = L.hole;
    const double h = L.h;
    double ul[5], um[5], ur[5], s[5];
    // the neighbours' p2c and the face fluxes share their positions x -/+ h
    const Geo gl = geo_off<AXIS>(b, x, y, -h);
    const Geo gc = geo_at(b, x, y, 0.0);
    const Geo gr = geo_off<AXIS>(b, x, y, h);
    gr_p2c(WL, gl, ul, L.P);
    gr_p2c(WM, gc, um, L.P);
    gr_p2c(WR, gr, ur, L.P);
    slope5(ul, um, ur, L.bias, L.limiter, s);
    double xl[5], xr[5], fl[5], fr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        xl[k] = um[k] - (0.5 * s[k]);
        xr[k] = um[k] + (0.5 * s[k]);
    }
    gr_flux<AXIS>(xl, gl, L.P, fl);
    gr_flux<AXIS>(xr, gr, L.P, fr);
    double e[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) e[k] = um[k] + (fdt_h * (fl[k] - fr[k]));
    gr_c2p(e, gc.sdet, L.P, out);
}

// computeRungeKuttaTimeStep{,2D} (Solvers/BlackHoleEulerForcingSolver.cpp:187-250) with
// evolveConservedVariableVector{,2D} (:3-185): per cell, RK4 on the source term,
// each stage's middle state first evolved by fractional flux sweeps against
// the step-start neighbours (1D: x by 1/2, 1/2, 1; 2D: x 1/4 y 1/2 x 1/4,
// twice, then x 1/2 y 1 x 1/2), U + (k1 + 2 k2 + 2 k3 + k4) / 6.
template <int DIM>
__global__ __launch_bounds__(kGRThreads) void k_gr_forcing(GRLaunch L) {
    long idx = (long)blockIdx.x * kGRThreads + threadIdx.x;
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
    // neighbours' conserved vectors, each formed at its own centre
    // (the grid loop of computeRungeKuttaTimeStep{,2D}) ...
    double uL[5], uM[5], uR[5], uT[5], uB[5];
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
    // ... and back to primitives at x -/+ h (y -/+ h) inside evolveConservedVariableVector.
    double WLn[5], WRn[5], WTn[5], WBn[5];
    gr_c2p(uL, geo_at(b, x - h, y, 0.0).sdet, L.P, WLn);
    gr_c2p(uR, geo_at(b, x + h, y, 0.0).sdet, L.P, WRn);
    if (DIM == 2) {
        gr_c2p(uT, geo_at(b, x, y - h, 0.0).sdet, L.P, WTn);
        gr_c2p(uB, geo_at(b, x, y + h, 0.0).sdet, L.P, WBn);
    }
    SourceGeo sg;
    source_geo(b, x, y, 0.0, h, sg);
    double kk[4][5];
    double src[5];
    gr_source(uM, sg, L.P, src);
#pragma unroll
    for (int k = 0; k < 5; ++k) kk[0][k] = L.dt * src[k];
    for (int st = 0; st < 3; ++st) {
        double tmp[5], Wm[5], e1[5], um[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) tmp[k] = st < 2 ? uM[k] + (0.5 * kk[st][k]) : uM[k] + kk[st][k];
        gr_c2p(tmp, gc.sdet, L.P, Wm);
        if (DIM == 1) {
            gr_fractional<0>(L.frac_dt_h[st], WLn, Wm, WRn, x, y, L, e1);
        } else {
            double e2[5], e3[5];
            gr_fractional<0>(L.frac_dt_h[st], WLn, Wm, WRn, x, y, L, e2);
            gr_fractional<1>(L.frac_dt_h_y[st], WTn, e2, WBn, x, y, L, e3);
            gr_fractional<0>(L.frac_dt_h[st], WLn, e3, WRn, x, y, L, e1);
        }
        gr_p2c(e1, gc, um, L.P);
        gr_source(um, sg, L.P, src);
#pragma unroll
        for (int k = 0; k < 5; ++k) kk[st + 1][k] = L.dt * src[k];
    }
    double un[5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        un[k] = uM[k] + ((1.0 / 6.0) * (kk[0][k] + ((2.0 * kk[1][k]) + ((2.0 * kk[2][k]) + kk[3][k]))));
    gr_c2p(un, gc.sdet, L.P, W);
    store5(L.out, cell, W);
}

// computeMaximumWaveSpeed{,2D} (Solvers/BlackHoleEulerSolvers.cpp:88-140): the
// reference places cells on the unit square whatever the grid spacing
// (1/N per cell), tests excision at (j/nx, i/ny) but evaluates the sound
// speed at (i/ny, j/nx), and takes max |c| over non-excised cells
// (NaN never enters the running max).
template <int DIM>
__global__ __launch_bounds__(kGRThreads) void k_gr_wavespeed(GRLaunch L, double dr, double dc,
                                                             unsigned long long* slots) {
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

}  // namespace

hipError_t launch_gr_sweep(const GRLaunch& L, int axis, hipStream_t s) {
    const dim3 block(kGRThreads);
    const long ifaces = axis == 0 ? (long)(L.nx + 1) * L.ny : (long)(L.ny + 1) * L.nx;
    if (L.order == 2) {
        if (axis == 0) hipLaunchKernelGGL((k_gr_iface<2, 0>), blocks_for(ifaces), block, 0, s, L);
        else hipLaunchKernelGGL((k_gr_iface<2, 1>), blocks_for(ifaces), block, 0, s, L);
    } else {
        if (axis == 0) hipLaunchKernelGGL((k_gr_iface<1, 0>), blocks_for(ifaces), block, 0, s, L);
        else hipLaunchKernelGGL((k_gr_iface<1, 1>), blocks_for(ifaces), block, 0, s, L);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long cells = (long)L.nx * L.ny;
    if (L.order == 2) {
        if (axis == 0) hipLaunchKernelGGL((k_gr_update<2, 0>), blocks_for(cells), block, 0, s, L);
        else hipLaunchKernelGGL((k_gr_update<2, 1>), blocks_for(cells), block, 0, s, L);
    } else {
        if (axis == 0) hipLaunchKernelGGL((k_gr_update<1, 0>), blocks_for(cells), block, 0, s, L);
        else hipLaunchKernelGGL((k_gr_update<1, 1>), blocks_for(cells), block, 0, s, L);
    }
    return hipGetLastError();
}

hipError_t launch_gr_forcing(const GRLaunch& L, hipStream_t s) {
    const long cells = (long)L.nx * L.ny;
    if (L.dim == 1) hipLaunchKernelGGL((k_gr_forcing<1>), blocks_for(cells), dim3(kGRThreads), 0, s, L);
    else hipLaunchKernelGGL((k_gr_forcing<2>), blocks_for(cells), dim3(kGRThreads), 0, s, L);
    return hipGetLastError();
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
