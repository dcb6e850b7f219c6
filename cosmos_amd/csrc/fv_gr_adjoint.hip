// gfx950 kernels of the adjoint (reverse mode) of the general-relativistic
// update: the FORCE / SLIC sweep with excision
// (Solvers/BlackHoleEulerSecondOrderSolver.cpp:3-143, Solvers/BlackHoleEulerFirstOrderSolver.cpp:3-191)
// and the RK4 source-term forcing step
// (Solvers/BlackHoleEulerForcingSolver.cpp:3-250).
//
// Sweep: the forward fluxes are in the flux planes (the caller re-runs the
// forward sweep first, or points them at the ones a forward sweep kept, with
// its SLIC face states, which the interface adjoint then reads instead of
// re-running the half steps).  k_gra_cells gives each cell's adjoint through its
// update (c2p, then p2c back to the input primitives; excised stencils pass
// their adjoint through unchanged) and stores the adjoint of the updated
// conserved vector; k_gra_ifaces forms each interface flux's adjoint from its
// two cells' and scatters it through the FORCE / SLIC stencil onto the input
// cells with FP64 atomic adds (accumulation order is unspecified, so results
// may differ in the last bits from run to run).
//
// Forcing: one thread per cell re-runs the cell's RK4 chain, keeping every
// stage's states, then walks it back; its own and its four neighbours'
// adjoints are accumulated with atomic adds into a zeroed output.  With a
// tape (written by k_gra_forcing_tape) the walk back runs as a chain of
// smaller kernels instead (k_grr_*).
#include "fv_gr_adjoint.hpp"
#include "fv_gr_adjoint_launch.hpp"

namespace cfv {
namespace {

constexpr int kGAT = 64;  // one-wave blocks: the reverse kernels run at one wave per SIMD

__device__ __forceinline__ int mirror2_a(int c, int n) {
    if (c < 0) c = -c - 1;
    if (c >= n) c = 2 * n - 1 - c;
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}

__device__ __forceinline__ int clamp1_a(int c, int n) {
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}

__device__ __forceinline__ void ld5(const ConstPlanes& in, long idx, double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) W[k] = in.p[k][idx];
}

__device__ __forceinline__ void st5(const Planes& out, long idx, const double* W) {
#pragma unroll
    for (int k = 0; k < 5; ++k) out.p[k][idx] = W[k];
}

__device__ __forceinline__ void add5(const Planes& out, long idx, const double* g) {
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (g[k] != 0.0) atomicAdd(out.p[k] + idx, g[k]);
}

__device__ __forceinline__ void zero5(double* g) {
#pragma unroll
    for (int k = 0; k < 5; ++k) g[k] = 0.0;
}

// Position-only values from the launch's geometry cache where it has one
// (bitwise what evaluating at the call site gives), else evaluated there.
// (ax, ay): the half-lattice point of (x, y).
__device__ __forceinline__ Geo geo_l(const GRLaunch& L, double x, double y, int ax, int ay) {
    return L.tab.geo ? geo_from_table(L.tab, L.dim, ax, ay) : geo_at(L.hole, x, y, 0.0);
}

template <int AXIS>
__device__ __forceinline__ Geo geo_shift_l(const GRLaunch& L, double x, double y, double off, int ax, int ay, int d) {
    return AXIS == 0 ? geo_l(L, x + off, y, ax + d, ay) : geo_l(L, x, y + off, ax, ay + d);
}

template <int AXIS>
__device__ __forceinline__ adjgr::ForceGeo force_geo_l(const GRLaunch& L, double x, double y, int ax, int ay) {
    const double hh = L.h * 0.5;
    return adjgr::ForceGeo{geo_shift_l<AXIS>(L, x, y, -hh, ax, ay, -1), geo_shift_l<AXIS>(L, x, y, hh, ax, ay, 1),
                           geo_l(L, x, y, ax, ay)};
}

template <int AXIS>
__device__ __forceinline__ adjgr::HalfGeo half_geo_l(const GRLaunch& L, double x, double y, int ax, int ay) {
    const double h = L.h;
    return adjgr::HalfGeo{geo_shift_l<AXIS>(L, x, y, -h, ax, ay, -2), geo_l(L, x, y, ax, ay),
                          geo_shift_l<AXIS>(L, x, y, h, ax, ay, 2), geo_shift_l<AXIS>(L, x, y, -h * 0.5, ax, ay, -1),
                          geo_shift_l<AXIS>(L, x, y, h * 0.5, ax, ay, 1)};
}

// The source term's position part at cell (i, j) = (x, y).
__device__ __forceinline__ void source_l(const GRLaunch& L, int i, int j, double x, double y, SourceGeo& sg) {
    if (L.tab.src) source_from_table(L.tab.src + (long)i * L.tab.spitch + j, L.tab.splane, sg);
    else source_geo(L.hole, x, y, 0.0, L.h, sg);
}

// The forward's excision test (stencil_clear in the forward kernels).
__device__ __forceinline__ bool clear_of(const Hole& b, int dim, int i, int j, int w, double h) {
    double y = dim == 2 ? i * h : 0.0;
    if (excised(b, j * h, y, 0.0)) return false;
    for (int d = 1; d <= w; ++d) {
        if (excised(b, (j - d) * h, y, 0.0) || excised(b, (j + d) * h, y, 0.0)) return false;
        if (dim == 2 && (excised(b, j * h, (i - d) * h, 0.0) || excised(b, j * h, (i + d) * h, 0.0))) return false;
    }
    return true;
}

// Per cell of a sweep: the adjoint through the conservative update.
template <int ORDER, int AXIS>
__global__ __launch_bounds__(kGAT) void k_gra_cells(GRAdjLaunch A) {
    const GRLaunch& L = A.L;
    long idx = (long)blockIdx.x * kGAT + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return;
    int i = (int)(idx / L.nx), j = (int)(idx - (long)i * L.nx);
    long cell = (long)i * L.pitch + j;
    double lam[5];
    ld5(A.lam, cell, lam);
    double gun[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (clear_of(L.hole, L.dim, i, j, ORDER == 2 ? 2 : 1, L.h)) {
        double W[5], U[5], un[5];
        ld5(L.in, cell, W);
        const Geo g = geo_l(L, j * L.h, L.dim == 2 ? i * L.h : 0.0, 2 * j, L.dim == 2 ? 2 * i : 0);
        gr_p2c(W, g, U, L.P);
        long fl = (long)i * L.fpitch + j;
        long fr = AXIS == 0 ? fl + 1 : fl + L.fpitch;
#pragma unroll
        for (int k = 0; k < 5; ++k) un[k] = U[k] + (L.coef.dt_dx * (L.flux.p[k][fl] - L.flux.p[k][fr]));
        adjgr::vjp_c2p(un, g.sdet, gr_tol(L.P, 1e-8), L.P, lam, gun);
        double gW[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        adjgr::vjp_p2c(W, g, L.P, gun, gW);
        st5(A.lam_in, cell, gW);
    } else {
        st5(A.lam_in, cell, lam);
    }
    st5(A.gbuf, cell, gun);
}

// Per interface: the flux adjoint dt/dx (g_right - g_left) of the updated
// conserved vectors of its two cells, back through the FORCE / SLIC stencil.
template <int ORDER, int AXIS>
__global__ __launch_bounds__(kGAT) void k_gra_ifaces(GRAdjLaunch A) {
    const GRLaunch& L = A.L;
    const int n = AXIS == 0 ? L.nx : L.ny;
    const long lines = AXIS == 0 ? L.ny : L.nx;
    const long per = AXIS == 0 ? (long)(n + 1) : lines;
    const long total = (long)(n + 1) * lines;
    long idx = (long)blockIdx.x * kGAT + threadIdx.x;
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
    auto cell_at = [&](int cc) -> long { return AXIS == 0 ? (long)i * L.pitch + cc : (long)cc * L.pitch + j; };
    // un_c = U_c + k (F_c - F_{c+1}): dF_c gets k (g_c - g_{c-1})
    double gF[5];
    bool any = false;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double gr = c < n ? A.gbuf.p[k][cell_at(c)] : 0.0;
        const double gl = c > 0 ? A.gbuf.p[k][cell_at(c - 1)] : 0.0;
        gF[k] = L.coef.dt_dx * (gr - gl);
        any |= gF[k] != 0.0;
    }
    if (!any) return;
    auto at = [&](int cc) -> long { return cell_at(ORDER == 2 ? mirror2_a(cc, n) : clamp1_a(cc, n)); };
    double x, y;
    int ax, ay;
    if (AXIS == 0) {
        x = (c - 0.5) * L.h;
        y = L.dim == 2 ? i * L.h : 0.0;
        ax = 2 * c - 1;
        ay = L.dim == 2 ? 2 * i : 0;
    } else {
        x = j * L.h;
        y = (c - 0.5) * L.h;
        ax = 2 * j;
        ay = 2 * c - 1;
    }
    if (ORDER == 1) {
        double WL[5], WR[5], gL[5], gR[5];
        ld5(L.in, at(c - 1), WL);
        ld5(L.in, at(c), WR);
        zero5(gL);
        zero5(gR);
        adjgr::vjp_force_g<AXIS>(WL, WR, force_geo_l<AXIS>(L, x, y, ax, ay), L.coef, L.P, gF, gL, gR);
        add5(A.lam_in, at(c - 1), gL);
        add5(A.lam_in, at(c), gR);
        return;
    }
    const double hh = L.h * 0.5;
    const double xl = AXIS == 0 ? x - hh : x, yl = AXIS == 0 ? y : y - hh;
    const double xr = AXIS == 0 ? x + hh : x, yr = AXIS == 0 ? y : y + hh;
    double W0[5], W1[5], W2[5], W3[5], pr[5], pl[5];
    ld5(L.in, at(c - 2), W0);
    ld5(L.in, at(c - 1), W1);
    ld5(L.in, at(c), W2);
    ld5(L.in, at(c + 1), W3);
    // each metric point evaluated once for the forward re-run and the adjoint
    const adjgr::HalfGeo HL = half_geo_l<AXIS>(L, xl, yl, AXIS == 0 ? ax - 1 : ax, AXIS == 0 ? ay : ay - 1),
                         HR = half_geo_l<AXIS>(L, xr, yr, AXIS == 0 ? ax + 1 : ax, AXIS == 0 ? ay : ay + 1);
    if (L.half) {
        // the face states the forward sweep kept
        const long f = AXIS == 0 ? (long)i * L.fpitch + c : (long)c * L.fpitch + j;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            pr[k] = L.half[k * L.half_plane + f];
            pl[k] = L.half[(5 + k) * L.half_plane + f];
        }
    } else {
        adjgr::half_fwd_g<AXIS>(W0, W1, W2, 1, HL, L.coef, L.bias, L.limiter, L.P, pr);
        adjgr::half_fwd_g<AXIS>(W1, W2, W3, 0, HR, L.coef, L.bias, L.limiter, L.P, pl);
    }
    double gpr[5], gpl[5];
    zero5(gpr);
    zero5(gpl);
    adjgr::vjp_force_g<AXIS>(pr, pl, force_geo_l<AXIS>(L, x, y, ax, ay), L.coef, L.P, gF, gpr, gpl);
    double g0[5], g1[5], g2[5], g3[5];
    zero5(g0);
    zero5(g1);
    zero5(g2);
    zero5(g3);
    adjgr::vjp_half_g<AXIS>(W0, W1, W2, 1, HL, L.coef, L.bias, L.limiter, L.P, gpr, g0, g1, g2);
    adjgr::vjp_half_g<AXIS>(W1, W2, W3, 0, HR, L.coef, L.bias, L.limiter, L.P, gpl, g1, g2, g3);
    add5(A.lam_in, at(c - 2), g0);
    add5(A.lam_in, at(c - 1), g1);
    add5(A.lam_in, at(c), g2);
    add5(A.lam_in, at(c + 1), g3);
}

// Per cell: the adjoint of the RK4 forcing (gr forcing kernel), into a zeroed lam_in.
template <int DIM>
__global__ __launch_bounds__(kGAT) void k_gra_forcing(GRAdjLaunch A) {
    const GRLaunch& L = A.L;
    long idx = (long)blockIdx.x * kGAT + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return;
    int i = (int)(idx / L.nx), j = (int)(idx - (long)i * L.nx);
    long cell = (long)i * L.pitch + j;
    double lam[5];
    ld5(A.lam, cell, lam);
    const Hole& b = L.hole;
    const double h = L.h;
    if (!clear_of(b, DIM, i, j, 1, h)) {
        add5(A.lam_in, cell, lam);
        return;
    }
    const PhysParams& P = L.P;
    const double tol = gr_tol(P, 1e-8);
    const double x = j * h, y = DIM == 2 ? i * h : 0.0;
    // forward chain, every stage kept
    const long nb[4] = {(long)i * L.pitch + clamp1_a(j - 1, L.nx), (long)i * L.pitch + clamp1_a(j + 1, L.nx),
                        DIM == 2 ? (long)clamp1_a(i - 1, L.ny) * L.pitch + j : cell,
                        DIM == 2 ? (long)clamp1_a(i + 1, L.ny) * L.pitch + j : cell};
    const double nbx[4] = {(j - 1) * h, (j + 1) * h, x, x};
    const double nby[4] = {y, y, DIM == 2 ? (i - 1) * h : y, DIM == 2 ? (i + 1) * h : y};
    // the metric at the centre and one cell either side, shared by every
    // fractional sweep of the four stages (the positions the forward forms)
    const int ax = 2 * j, ay = DIM == 2 ? 2 * i : 0;
    const Geo gc = geo_l(L, x, y, ax, ay);
    const Geo gxm = geo_l(L, x - h, y, ax - 2, ay), gxp = geo_l(L, x + h, y, ax + 2, ay);
    Geo gym, gyp;
    if (DIM == 2) {
        gym = geo_l(L, x, y - h, ax, ay - 2);
        gyp = geo_l(L, x, y + h, ax, ay + 2);
    }
    const int nn = DIM == 2 ? 4 : 2;
    double Wnb[4][5], unb[4][5], Wn[4][5];
    for (int q = 0; q < nn; ++q) {
        ld5(L.in, nb[q], Wnb[q]);
        gr_p2c(Wnb[q], geo_l(L, nbx[q], nby[q], q == 0 ? ax - 2 : q == 1 ? ax + 2 : ax,
                             q == 2 ? ay - 2 : q == 3 ? ay + 2 : ay), unb[q], P);
        gr_c2p(unb[q], (q == 0 ? gxm : q == 1 ? gxp : q == 2 ? gym : gyp).sdet, P, Wn[q]);
    }
    double W[5], uM[5];
    ld5(L.in, cell, W);
    gr_p2c(W, gc, uM, P);
    SourceGeo sg;
    source_l(L, i, j, x, y, sg);
    double kk[4][5], tmp[3][5], Wm[3][5], e2[3][5], e3[3][5], e1[3][5], um[3][5], src[5];
    gr_source(uM, sg, P, src);
#pragma unroll
    for (int k = 0; k < 5; ++k) kk[0][k] = L.dt * src[k];
    for (int st = 0; st < 3; ++st) {
#pragma unroll
        for (int k = 0; k < 5; ++k) tmp[st][k] = st < 2 ? uM[k] + (0.5 * kk[st][k]) : uM[k] + kk[st][k];
        gr_c2p(tmp[st], gc.sdet, P, Wm[st]);
        if (DIM == 1) {
            adjgr::frac_fwd_g<0>(L.frac_dt_h[st], Wn[0], Wm[st], Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, e1[st]);
        } else {
            adjgr::frac_fwd_g<0>(L.frac_dt_h[st], Wn[0], Wm[st], Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, e2[st]);
            adjgr::frac_fwd_g<1>(L.frac_dt_h_y[st], Wn[2], e2[st], Wn[3], gym, gc, gyp, L.bias, L.limiter, P, e3[st]);
            adjgr::frac_fwd_g<0>(L.frac_dt_h[st], Wn[0], e3[st], Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, e1[st]);
        }
        gr_p2c(e1[st], gc, um[st], P);
        gr_source(um[st], sg, P, src);
#pragma unroll
        for (int k = 0; k < 5; ++k) kk[st + 1][k] = L.dt * src[k];
    }
    double un[5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        un[k] = uM[k] + ((1.0 / 6.0) * (kk[0][k] + ((2.0 * kk[1][k]) + ((2.0 * kk[2][k]) + kk[3][k]))));
    // reverse
    double gun[5], guM[5], gkk[4][5], gWn[4][5];
    zero5(gun);
    adjgr::vjp_c2p(un, gc.sdet, tol, P, lam, gun);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        guM[k] = gun[k];
        gkk[0][k] = (1.0 / 6.0) * gun[k];
        gkk[1][k] = (1.0 / 6.0) * 2.0 * gun[k];
        gkk[2][k] = (1.0 / 6.0) * 2.0 * gun[k];
        gkk[3][k] = (1.0 / 6.0) * gun[k];
    }
    for (int q = 0; q < 4; ++q) zero5(gWn[q]);
    for (int st = 2; st >= 0; --st) {
        double gsrc[5], gum[5], ge1[5], gWm[5], gtmp[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) gsrc[k] = L.dt * gkk[st + 1][k];
        zero5(gum);
        adjgr::vjp_source(um[st], sg, P, gsrc, gum);
        zero5(ge1);
        adjgr::vjp_p2c(e1[st], gc, P, gum, ge1);
        zero5(gWm);
        if (DIM == 1) {
            adjgr::vjp_frac_g<0>(L.frac_dt_h[st], Wn[0], Wm[st], Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, ge1,
                                 gWn[0], gWm, gWn[1]);
        } else {
            double ge3[5], ge2[5];
            zero5(ge3);
            zero5(ge2);
            adjgr::vjp_frac_g<0>(L.frac_dt_h[st], Wn[0], e3[st], Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, ge1,
                                 gWn[0], ge3, gWn[1]);
            adjgr::vjp_frac_g<1>(L.frac_dt_h_y[st], Wn[2], e2[st], Wn[3], gym, gc, gyp, L.bias, L.limiter, P, ge3,
                                 gWn[2], ge2, gWn[3]);
            adjgr::vjp_frac_g<0>(L.frac_dt_h[st], Wn[0], Wm[st], Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, ge2,
                                 gWn[0], gWm, gWn[1]);
        }
        zero5(gtmp);
        adjgr::vjp_c2p(tmp[st], gc.sdet, tol, P, gWm, gtmp);
        const double f = st < 2 ? 0.5 : 1.0;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            guM[k] += gtmp[k];
            gkk[st][k] += f * gtmp[k];
        }
    }
    {
        double gsrc[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) gsrc[k] = L.dt * gkk[0][k];
        adjgr::vjp_source(uM, sg, P, gsrc, guM);
    }
    double gW[5];
    zero5(gW);
    adjgr::vjp_p2c(W, gc, P, guM, gW);
    add5(A.lam_in, cell, gW);
    for (int q = 0; q < nn; ++q) {
        double gu[5], gw[5];
        zero5(gu);
        zero5(gw);
        adjgr::vjp_c2p(unb[q], (q == 0 ? gxm : q == 1 ? gxp : q == 2 ? gym : gyp).sdet, tol, P, gWn[q], gu);
        adjgr::vjp_p2c(Wnb[q], geo_l(L, nbx[q], nby[q], q == 0 ? ax - 2 : q == 1 ? ax + 2 : ax,
                                     q == 2 ? ay - 2 : q == 3 ? ay + 2 : ay), P, gu, gw);
        add5(A.lam_in, nb[q], gw);
    }
}

// Taped forcing adjoint.  k_gra_forcing_tape runs the forward RK4 chain of
// every cell (bitwise the forward forcing kernel; it writes the step's output
// too) and keeps the values its reverse needs in the tape (one SoA plane per
// value; layout in fv_gr_adjoint_launch.hpp); the k_grr_* kernels walk the chain
// back from the tape, so the reverse pass evaluates no forward stage again.
template <int DIM>
__global__ __launch_bounds__(kGAT) void k_gra_forcing_tape(GRAdjLaunch A) {
    const GRLaunch& L = A.L;
    long idx = (long)blockIdx.x * kGAT + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return;
    int i = (int)(idx / L.nx), j = (int)(idx - (long)i * L.nx);
    long cell = (long)i * L.pitch + j;
    double* const tp = A.tape + cell;
    const long pl = A.tape_plane;
    auto put = [&](int slot, const double* v) {
#pragma unroll
        for (int k = 0; k < 5; ++k) tp[(long)(slot + k) * pl] = v[k];
    };
    double W[5];
    ld5(L.in, cell, W);
    const Hole& b = L.hole;
    const double h = L.h;
    if (!clear_of(b, DIM, i, j, 1, h)) {
        st5(L.out, cell, W);
        return;
    }
    const PhysParams& P = L.P;
    const double x = j * h, y = DIM == 2 ? i * h : 0.0;
    const int ax = 2 * j, ay = DIM == 2 ? 2 * i : 0;
    const Geo gc = geo_l(L, x, y, ax, ay);
    const Geo gxm = geo_l(L, x - h, y, ax - 2, ay), gxp = geo_l(L, x + h, y, ax + 2, ay);
    Geo gym, gyp;
    if (DIM == 2) {
        gym = geo_l(L, x, y - h, ax, ay - 2);
        gyp = geo_l(L, x, y + h, ax, ay + 2);
    }
    // neighbours: p2c at their own centres, c2p at x -/+ h (y -/+ h)
    double Wn[4][5];
    {
        const long nb[4] = {(long)i * L.pitch + clamp1_a(j - 1, L.nx), (long)i * L.pitch + clamp1_a(j + 1, L.nx),
                            DIM == 2 ? (long)clamp1_a(i - 1, L.ny) * L.pitch + j : cell,
                            DIM == 2 ? (long)clamp1_a(i + 1, L.ny) * L.pitch + j : cell};
        const double nbx[4] = {(j - 1) * h, (j + 1) * h, x, x};
        const double nby[4] = {y, y, DIM == 2 ? (i - 1) * h : y, DIM == 2 ? (i + 1) * h : y};
        const int nn = DIM == 2 ? 4 : 2;
        for (int q = 0; q < nn; ++q) {
            double Wq[5], uq[5];
            ld5(L.in, nb[q], Wq);
            gr_p2c(Wq, geo_l(L, nbx[q], nby[q], q == 0 ? ax - 2 : q == 1 ? ax + 2 : ax,
                             q == 2 ? ay - 2 : q == 3 ? ay + 2 : ay), uq, P);
            gr_c2p(uq, (q == 0 ? gxm : q == 1 ? gxp : q == 2 ? gym : gyp).sdet, P, Wn[q]);
            put(kTapeWn + 5 * q, Wn[q]);
        }
    }
    double uM[5];
    gr_p2c(W, gc, uM, P);
    SourceGeo sg;
    source_l(L, i, j, x, y, sg);
    double kk[4][5], src[5];
    gr_source(uM, sg, P, src);
#pragma unroll
    for (int k = 0; k < 5; ++k) kk[0][k] = L.dt * src[k];
    for (int st = 0; st < 3; ++st) {
        const int base = kTapeStage + kTapePerStage * st;
        double tmp[5], Wm[5], e1[5], um[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) tmp[k] = st < 2 ? uM[k] + (0.5 * kk[st][k]) : uM[k] + kk[st][k];
        gr_c2p(tmp, gc.sdet, P, Wm);
        put(base + kTapeTmp, tmp);
        put(base + kTapeWm, Wm);
        if (DIM == 1) {
            adjgr::frac_fwd_g<0>(L.frac_dt_h[st], Wn[0], Wm, Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, e1);
        } else {
            double e2[5], e3[5];
            adjgr::frac_fwd_g<0>(L.frac_dt_h[st], Wn[0], Wm, Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, e2);
            adjgr::frac_fwd_g<1>(L.frac_dt_h_y[st], Wn[2], e2, Wn[3], gym, gc, gyp, L.bias, L.limiter, P, e3);
            adjgr::frac_fwd_g<0>(L.frac_dt_h[st], Wn[0], e3, Wn[1], gxm, gc, gxp, L.bias, L.limiter, P, e1);
            put(base + kTapeE2, e2);
            put(base + kTapeE3, e3);
        }
        put(base + kTapeE1, e1);
        gr_p2c(e1, gc, um, P);
        gr_source(um, sg, P, src);
#pragma unroll
        for (int k = 0; k < 5; ++k) kk[st + 1][k] = L.dt * src[k];
    }
    double un[5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        un[k] = uM[k] + ((1.0 / 6.0) * (kk[0][k] + ((2.0 * kk[1][k]) + ((2.0 * kk[2][k]) + kk[3][k]))));
    put(kTapeUn, un);
    gr_c2p(un, gc.sdet, P, W);
    st5(L.out, cell, W);
}

// The reverse of a taped forcing step as a chain of small kernels, one per
// piece of the RK4 chain, so that no kernel holds more than one fractional
// sweep's adjoint in registers; the running adjoints pass through the carry
// planes (layout in fv_gr_adjoint_launch.hpp).  Per stage st = 2, 1, 0:
//   src: gk -> the source's adjoint at the stage's state -> ge (= adjoint of e1)
//   fx1: (2D) the last x fraction back to e3; fy: the y fraction back to e2
//   fx2: the first x fraction back to Wm, then through c2p to tmp: adds to
//        guM, and sets gk for the stage before.
// The neighbour-state adjoints gWn accumulate over all stages; head starts
// the chain from the output adjoint, tail ends it at the input state and the
// neighbours (atomic adds into a zeroed lam_in).
struct RevCell {
    int i, j;
    long cell;
    double x, y;
};

template <int DIM>
__device__ __forceinline__ bool rev_cell(const GRAdjLaunch& A, RevCell& c) {
    const GRLaunch& L = A.L;
    long idx = (long)blockIdx.x * kGAT + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return false;
    c.i = (int)(idx / L.nx);
    c.j = (int)(idx - (long)c.i * L.nx);
    c.cell = (long)c.i * L.pitch + c.j;
    c.x = c.j * L.h;
    c.y = DIM == 2 ? c.i * L.h : 0.0;
    return clear_of(L.hole, DIM, c.i, c.j, 1, L.h);
}

__device__ __forceinline__ void tget(const GRAdjLaunch& A, int slot, long cell, double* v) {
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = A.tape[(long)(slot + k) * A.tape_plane + cell];
}
__device__ __forceinline__ void cget(const GRAdjLaunch& A, int slot, long cell, double* v) {
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = A.carry[(long)(slot + k) * A.tape_plane + cell];
}
__device__ __forceinline__ void cput(const GRAdjLaunch& A, int slot, long cell, const double* v) {
#pragma unroll
    for (int k = 0; k < 5; ++k) A.carry[(long)(slot + k) * A.tape_plane + cell] = v[k];
}

template <int DIM>
__global__ __launch_bounds__(kGAT) void k_grr_head(GRAdjLaunch A) {
    const GRLaunch& L = A.L;
    long idx = (long)blockIdx.x * kGAT + threadIdx.x;
    if (idx >= (long)L.nx * L.ny) return;
    RevCell c;
    const bool clear = rev_cell<DIM>(A, c);
    double lam[5];
    ld5(A.lam, c.cell, lam);
    if (!clear) {
        add5(A.lam_in, c.cell, lam);
        return;
    }
    const Geo gc = geo_l(L, c.x, c.y, 2 * c.j, DIM == 2 ? 2 * c.i : 0);
    double un[5], gun[5], gk[5], z[5];
    tget(A, kTapeUn, c.cell, un);
    zero5(gun);
    adjgr::vjp_c2p(un, gc.sdet, gr_tol(L.P, 1e-8), L.P, lam, gun);
#pragma unroll
    for (int k = 0; k < 5; ++k) gk[k] = (1.0 / 6.0) * gun[k];
    zero5(z);
    cput(A, kCarryGun, c.cell, gun);
    cput(A, kCarryGuM, c.cell, gun);
    cput(A, kCarryGk, c.cell, gk);
    for (int q = 0; q < 2 * DIM; ++q) cput(A, kCarryGWn + 5 * q, c.cell, z);
}

template <int DIM>
__global__ __launch_bounds__(kGAT) void k_grr_src(GRAdjLaunch A, int st) {
    const GRLaunch& L = A.L;
    RevCell c;
    if (!rev_cell<DIM>(A, c)) return;
    const Geo gc = geo_l(L, c.x, c.y, 2 * c.j, DIM == 2 ? 2 * c.i : 0);
    SourceGeo sg;
    source_l(L, c.i, c.j, c.x, c.y, sg);
    double gk[5], e1[5], um[5], gsrc[5], gum[5], ge1[5];
    cget(A, kCarryGk, c.cell, gk);
    tget(A, kTapeStage + kTapePerStage * st + kTapeE1, c.cell, e1);
    gr_p2c(e1, gc, um, L.P);
#pragma unroll
    for (int k = 0; k < 5; ++k) gsrc[k] = L.dt * gk[k];
    zero5(gum);
    adjgr::vjp_source(um, sg, L.P, gsrc, gum);
    zero5(ge1);
    adjgr::vjp_p2c(e1, gc, L.P, gum, ge1);
    cput(A, kCarryGe, c.cell, ge1);
}

// One fractional sweep's adjoint: carry ge (the adjoint of its output) back
// to its middle input (slot `mid` of the stage) and the neighbour pair of its
// axis; last == true: the stage's first x fraction, continued through c2p.
template <int DIM, int AXIS, bool LAST>
__global__ __launch_bounds__(kGAT) void k_grr_frac(GRAdjLaunch A, int st) {
    const GRLaunch& L = A.L;
    RevCell c;
    if (!rev_cell<DIM>(A, c)) return;
    const int base = kTapeStage + kTapePerStage * st;
    const int mid = LAST ? kTapeWm : (AXIS == 0 ? kTapeE3 : kTapeE2);
    const int ax = 2 * c.j, ay = DIM == 2 ? 2 * c.i : 0;
    const Geo gc = geo_l(L, c.x, c.y, ax, ay);
    const Geo ga = geo_shift_l<AXIS>(L, c.x, c.y, -L.h, ax, ay, -2), gb = geo_shift_l<AXIS>(L, c.x, c.y, L.h, ax, ay, 2);
    double ge[5], Wa[5], Wm[5], Wb[5], gWa[5], gWb[5], gm[5];
    cget(A, kCarryGe, c.cell, ge);
    tget(A, base + mid, c.cell, Wm);
    tget(A, kTapeWn + 10 * AXIS, c.cell, Wa);
    tget(A, kTapeWn + 10 * AXIS + 5, c.cell, Wb);
    cget(A, kCarryGWn + 10 * AXIS, c.cell, gWa);
    cget(A, kCarryGWn + 10 * AXIS + 5, c.cell, gWb);
    zero5(gm);
    const double fdt_h = AXIS == 0 ? L.frac_dt_h[st] : L.frac_dt_h_y[st];
    adjgr::vjp_frac_g<AXIS>(fdt_h, Wa, Wm, Wb, ga, gc, gb, L.bias, L.limiter, L.P, ge, gWa, gm, gWb);
    cput(A, kCarryGWn + 10 * AXIS, c.cell, gWa);
    cput(A, kCarryGWn + 10 * AXIS + 5, c.cell, gWb);
    if (!LAST) {
        cput(A, kCarryGe, c.cell, gm);
        return;
    }
    // Wm = c2p(tmp), tmp = uM + f k_st: the stage's adjoint goes to uM and to k_st
    double tmp[5], gtmp[5], guM[5], gun[5], gk[5];
    tget(A, base + kTapeTmp, c.cell, tmp);
    zero5(gtmp);
    adjgr::vjp_c2p(tmp, gc.sdet, gr_tol(L.P, 1e-8), L.P, gm, gtmp);
    cget(A, kCarryGuM, c.cell, guM);
    cget(A, kCarryGun, c.cell, gun);
    const double f = st < 2 ? 0.5 : 1.0;
    const double cf = st == 0 ? (1.0 / 6.0) : (1.0 / 6.0) * 2.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        guM[k] += gtmp[k];
        gk[k] = cf * gun[k] + f * gtmp[k];
    }
    cput(A, kCarryGuM, c.cell, guM);
    cput(A, kCarryGk, c.cell, gk);
}

template <int DIM>
__global__ __launch_bounds__(kGAT) void k_grr_tail(GRAdjLaunch A) {
    const GRLaunch& L = A.L;
    RevCell c;
    if (!rev_cell<DIM>(A, c)) return;
    const Hole& b = L.hole;
    const double h = L.h;
    {
        const Geo gc = geo_l(L, c.x, c.y, 2 * c.j, DIM == 2 ? 2 * c.i : 0);
        SourceGeo sg;
        source_l(L, c.i, c.j, c.x, c.y, sg);
        double W[5], uM[5], gk[5], gsrc[5], guM[5], gW[5];
        ld5(L.in, c.cell, W);
        gr_p2c(W, gc, uM, L.P);
        cget(A, kCarryGk, c.cell, gk);
        cget(A, kCarryGuM, c.cell, guM);
#pragma unroll
        for (int k = 0; k < 5; ++k) gsrc[k] = L.dt * gk[k];
        adjgr::vjp_source(uM, sg, L.P, gsrc, guM);
        zero5(gW);
        adjgr::vjp_p2c(W, gc, L.P, guM, gW);
        add5(A.lam_in, c.cell, gW);
    }
    const double tol = gr_tol(L.P, 1e-8);
    for (int q = 0; q < 2 * DIM; ++q) {
        const int i = c.i, j = c.j;
        const long nbq = q == 0 ? (long)i * L.pitch + clamp1_a(j - 1, L.nx)
                         : q == 1 ? (long)i * L.pitch + clamp1_a(j + 1, L.nx)
                         : q == 2 ? (long)clamp1_a(i - 1, L.ny) * L.pitch + j
                                  : (long)clamp1_a(i + 1, L.ny) * L.pitch + j;
        const double px = q == 0 ? (j - 1) * h : q == 1 ? (j + 1) * h : c.x;
        const double py = q == 2 ? (i - 1) * h : q == 3 ? (i + 1) * h : c.y;
        const double sx = q == 0 ? c.x - h : q == 1 ? c.x + h : c.x;
        const double sy = q == 2 ? c.y - h : q == 3 ? c.y + h : c.y;
        double Wq[5], uq[5], gq[5], gu[5], gw[5];
        cget(A, kCarryGWn + 5 * q, c.cell, gq);
        ld5(L.in, nbq, Wq);
        const int qax = 2 * j + (q == 0 ? -2 : q == 1 ? 2 : 0), qay = DIM == 2 ? 2 * i + (q == 2 ? -2 : q == 3 ? 2 : 0) : 0;
        const Geo gp = geo_l(L, px, py, qax, qay);
        gr_p2c(Wq, gp, uq, L.P);
        zero5(gu);
        zero5(gw);
        adjgr::vjp_c2p(uq, geo_l(L, sx, sy, qax, qay).sdet, tol, L.P, gq, gu);
        adjgr::vjp_p2c(Wq, gp, L.P, gu, gw);
        add5(A.lam_in, nbq, gw);
    }
}

dim3 gat_blocks(long work) { return dim3((unsigned)((work + kGAT - 1) / kGAT)); }

}  // namespace

hipError_t launch_gr_adjoint_sweep(const GRAdjLaunch& A, int axis, hipStream_t s) {
    const GRLaunch& L = A.L;
    const long cells = (long)L.nx * L.ny;
    const long ifaces = axis == 0 ? (long)(L.nx + 1) * L.ny : (long)(L.ny + 1) * L.nx;
    const dim3 blk(kGAT);
    if (L.order == 2) {
        if (axis == 0) {
            hipLaunchKernelGGL((k_gra_cells<2, 0>), gat_blocks(cells), blk, 0, s, A);
            hipLaunchKernelGGL((k_gra_ifaces<2, 0>), gat_blocks(ifaces), blk, 0, s, A);
        } else {
            hipLaunchKernelGGL((k_gra_cells<2, 1>), gat_blocks(cells), blk, 0, s, A);
            hipLaunchKernelGGL((k_gra_ifaces<2, 1>), gat_blocks(ifaces), blk, 0, s, A);
        }
    } else {
        if (axis == 0) {
            hipLaunchKernelGGL((k_gra_cells<1, 0>), gat_blocks(cells), blk, 0, s, A);
            hipLaunchKernelGGL((k_gra_ifaces<1, 0>), gat_blocks(ifaces), blk, 0, s, A);
        } else {
            hipLaunchKernelGGL((k_gra_cells<1, 1>), gat_blocks(cells), blk, 0, s, A);
            hipLaunchKernelGGL((k_gra_ifaces<1, 1>), gat_blocks(ifaces), blk, 0, s, A);
        }
    }
    return hipGetLastError();
}

hipError_t launch_gr_adjoint_forcing(const GRAdjLaunch& A, hipStream_t s) {
    const long cells = (long)A.L.nx * A.L.ny;
    if (A.tape) {
        const dim3 g = gat_blocks(cells), blk(kGAT);
        if (A.L.dim == 1) {
            hipLaunchKernelGGL((k_grr_head<1>), g, blk, 0, s, A);
            for (int st = 2; st >= 0; --st) {
                hipLaunchKernelGGL((k_grr_src<1>), g, blk, 0, s, A, st);
                hipLaunchKernelGGL((k_grr_frac<1, 0, true>), g, blk, 0, s, A, st);
            }
            hipLaunchKernelGGL((k_grr_tail<1>), g, blk, 0, s, A);
        } else {
            hipLaunchKernelGGL((k_grr_head<2>), g, blk, 0, s, A);
            for (int st = 2; st >= 0; --st) {
                hipLaunchKernelGGL((k_grr_src<2>), g, blk, 0, s, A, st);
                hipLaunchKernelGGL((k_grr_frac<2, 0, false>), g, blk, 0, s, A, st);
                hipLaunchKernelGGL((k_grr_frac<2, 1, false>), g, blk, 0, s, A, st);
                hipLaunchKernelGGL((k_grr_frac<2, 0, true>), g, blk, 0, s, A, st);
            }
            hipLaunchKernelGGL((k_grr_tail<2>), g, blk, 0, s, A);
        }
    } else {
        if (A.L.dim == 1) hipLaunchKernelGGL((k_gra_forcing<1>), gat_blocks(cells), dim3(kGAT), 0, s, A);
        else hipLaunchKernelGGL((k_gra_forcing<2>), gat_blocks(cells), dim3(kGAT), 0, s, A);
    }
    return hipGetLastError();
}

hipError_t launch_gr_forcing_tape(const GRAdjLaunch& A, hipStream_t s) {
    const long cells = (long)A.L.nx * A.L.ny;
    if (A.L.dim == 1) hipLaunchKernelGGL((k_gra_forcing_tape<1>), gat_blocks(cells), dim3(kGAT), 0, s, A);
    else hipLaunchKernelGGL((k_gra_forcing_tape<2>), gat_blocks(cells), dim3(kGAT), 0, s, A);
    return hipGetLastError();
}

}  // namespace cfv
