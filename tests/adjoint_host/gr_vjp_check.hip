// Host-side check of the general-relativistic vector-Jacobian products
// (cosmos_amd/csrc/fv_gr_adjoint.hpp) against central finite differences of
// the forward functions (cosmos_amd/csrc/fv_gr.hpp), compiled for the host:
// for random states u, directions e and output adjoints g,
//     <g, (f(u + h e) - f(u - h e)) / 2h>  ==  <vjp_f(u, g), e>.
// Con2prim runs with a tight tolerance so that its iteration count does not
// change across the finite-difference stencil.  Prints one line per product
// and exits non-zero if any relative difference exceeds 1e-6.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <functional>
#include <vector>

#include "fv_gr_adjoint.hpp"
#include "fv_kernels.hpp"

using namespace cfv;

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static double uni(double a, double b) {
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return a + (b - a) * ((g_state >> 11) * (1.0 / 9007199254740992.0));
}

static Hole make_hole(double M, double a, double x, double y) {
    Hole b{};
    b.M = M;
    b.a = a;
    b.x = x;
    b.y = y;
    b.z = 0.0;
    b.ma = (M * M) * (a * a);
    b.aM = a * M;
    b.r_exc = M * (1.0 + std::sqrt(1.0 - a * a));
    return b;
}

static void random_prim(double* W) {
    W[0] = uni(0.5, 2.0);
    W[1] = uni(-0.3, 0.3);
    W[2] = uni(-0.3, 0.3);
    W[3] = uni(-0.1, 0.1);
    W[4] = uni(0.05, 1.0);
}

static double dot(const double* a, const double* b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static int failures = 0;

// f maps nin inputs to nout outputs; vjp accumulates into gin.
static void check(const char* name, int nin, int nout, const std::vector<double>& x0,
                  const std::function<void(const double*, double*)>& f,
                  const std::function<void(const double*, const double*, double*)>& vjp, double hrel = 1e-6) {
    std::vector<double> e(nin), g(nout), xp(nin), xm(nin), fp(nout), fm(nout), gin(nin, 0.0);
    for (int i = 0; i < nin; ++i) e[i] = uni(-1.0, 1.0) * (std::fabs(x0[i]) + 0.1);
    for (int i = 0; i < nout; ++i) g[i] = uni(-1.0, 1.0);
    const double h = hrel;
    for (int i = 0; i < nin; ++i) {
        xp[i] = x0[i] + h * e[i];
        xm[i] = x0[i] - h * e[i];
    }
    f(xp.data(), fp.data());
    f(xm.data(), fm.data());
    double fd = 0.0;
    for (int i = 0; i < nout; ++i) fd += g[i] * (fp[i] - fm[i]) / (2.0 * h);
    vjp(x0.data(), g.data(), gin.data());
    const double ad = dot(gin.data(), e.data(), nin);
    const double rel = std::fabs(fd - ad) / std::fmax(std::fmax(std::fabs(fd), std::fabs(ad)), 1e-300);
    const bool ok = rel < 1e-6;
    if (!ok) ++failures;
    std::printf("%-28s fd % .12e  adjoint % .12e  rel %.2e %s\n", name, fd, ad, rel, ok ? "ok" : "FAIL");
}

int main() {
    PhysParams P = make_params(4.0 / 3.0);
    P.gr_tol = 1e-14;
    P.c2p_maxit = 200;
    const Hole holes[2] = {make_hole(1.0, 0.0, 20.0, 20.0), make_hole(0.05, 0.6, 0.5, 0.5)};
    const double hgrid = 40.0 / 1024.0;
    SweepCoef cf{};
    const double dt = 0.3 * hgrid;
    cf.dt_dx = dt / hgrid;
    cf.dx_dt = hgrid / dt;
    cf.hdt_dx = 0.5 * (dt / hgrid);
    for (int trial = 0; trial < 6; ++trial) {
        const Hole& b = holes[trial % 2];
        const double scale = trial % 2 == 0 ? 1.0 : 0.05;
        const double x = b.x + scale * uni(3.0, 9.0) * (trial % 3 == 0 ? -1.0 : 1.0);
        const double y = b.y + scale * uni(-6.0, 6.0);
        const double hh = trial % 2 == 0 ? hgrid : 0.01;
        const Geo G = geo_at(b, x, y, 0.0);
        std::vector<double> W(5);
        random_prim(W.data());
        std::printf("-- trial %d at (%.4f, %.4f)\n", trial, x, y);
        check("p2c", 5, 5, W, [&](const double* w, double* u) { gr_p2c(w, G, u, P); },
              [&](const double* w, const double* g, double* gw) { adjgr::vjp_p2c(w, G, P, g, gw); });
        std::vector<double> U(5);
        gr_p2c(W.data(), G, U.data(), P);
        check("c2p", 5, 5, U, [&](const double* u, double* w) { gr_c2p(u, G.sdet, P, w); },
              [&](const double* u, const double* g, double* gu) { adjgr::vjp_c2p(u, G.sdet, P.gr_tol, P, g, gu); });
        check("flux x", 5, 5, U, [&](const double* u, double* f) { gr_flux<0>(u, G, P, f); },
              [&](const double* u, const double* g, double* gu) { adjgr::vjp_flux<0>(u, G, P, g, gu); });
        check("flux y", 5, 5, U, [&](const double* u, double* f) { gr_flux<1>(u, G, P, f); },
              [&](const double* u, const double* g, double* gu) { adjgr::vjp_flux<1>(u, G, P, g, gu); });
        SourceGeo S;
        source_geo(b, x, y, 0.0, hh, S);
        check("source", 5, 5, U, [&](const double* u, double* s) { gr_source(u, S, P, s); },
              [&](const double* u, const double* g, double* gu) { adjgr::vjp_source(u, S, P, g, gu); });
        // three neighbouring states for the stencils
        std::vector<double> W3(15);
        for (int c = 0; c < 3; ++c) {
            random_prim(&W3[5 * c]);
            for (int k = 0; k < 5; ++k) W3[5 * c + k] = 0.5 * (W3[5 * c + k] + W[k]);
        }
        for (int limiter = 0; limiter < 3; ++limiter)
            for (int side = 0; side < 2; ++side) {
                char name[64];
                std::snprintf(name, sizeof name, "half x side %d lim %d", side, limiter);
                const double bias = limiter == 1 ? 0.0 : 0.25;
                check(name, 15, 5, W3,
                      [&](const double* w, double* o) {
                          adjgr::half_fwd<0>(w, w + 5, w + 10, x, y, side, b, hh, cf, bias, limiter, P, o);
                      },
                      [&](const double* w, const double* g, double* gw) {
                          adjgr::vjp_half<0>(w, w + 5, w + 10, x, y, side, b, hh, cf, bias, limiter, P, g, gw, gw + 5,
                                             gw + 10);
                      });
                std::snprintf(name, sizeof name, "half y side %d lim %d", side, limiter);
                check(name, 15, 5, W3,
                      [&](const double* w, double* o) {
                          adjgr::half_fwd<1>(w, w + 5, w + 10, x, y, side, b, hh, cf, bias, limiter, P, o);
                      },
                      [&](const double* w, const double* g, double* gw) {
                          adjgr::vjp_half<1>(w, w + 5, w + 10, x, y, side, b, hh, cf, bias, limiter, P, g, gw, gw + 5,
                                             gw + 10);
                      });
            }
        std::vector<double> W2(W3.begin(), W3.begin() + 10);
        check("force x", 10, 5, W2, [&](const double* w, double* F) { adjgr::force_fwd<0>(w, w + 5, x, y, b, hh, cf, P, F); },
              [&](const double* w, const double* g, double* gw) {
                  adjgr::vjp_force<0>(w, w + 5, x, y, b, hh, cf, P, g, gw, gw + 5);
              });
        check("force y", 10, 5, W2, [&](const double* w, double* F) { adjgr::force_fwd<1>(w, w + 5, x, y, b, hh, cf, P, F); },
              [&](const double* w, const double* g, double* gw) {
                  adjgr::vjp_force<1>(w, w + 5, x, y, b, hh, cf, P, g, gw, gw + 5);
              });
        const double fdt = 0.25 * (dt / hh);
        check("frac x", 15, 5, W3,
              [&](const double* w, double* o) { adjgr::frac_fwd<0>(fdt, w, w + 5, w + 10, x, y, b, hh, 0.0, 1, P, o); },
              [&](const double* w, const double* g, double* gw) {
                  adjgr::vjp_frac<0>(fdt, w, w + 5, w + 10, x, y, b, hh, 0.0, 1, P, g, gw, gw + 5, gw + 10);
              });
        check("frac y", 15, 5, W3,
              [&](const double* w, double* o) { adjgr::frac_fwd<1>(fdt, w, w + 5, w + 10, x, y, b, hh, 0.0, 1, P, o); },
              [&](const double* w, const double* g, double* gw) {
                  adjgr::vjp_frac<1>(fdt, w, w + 5, w + 10, x, y, b, hh, 0.0, 1, P, g, gw, gw + 5, gw + 10);
              });
    }
    std::printf("%s: %d failure(s)\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
