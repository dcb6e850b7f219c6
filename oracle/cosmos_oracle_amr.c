/* CPU restatement of the Cosmos two-level AMR driver and its refinement flags.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): tests use it as the checker; the
 * product path never calls it.
 *
 * Follows EulerAMR::solveLevel1AMR / solve2DLevel1AMR (AMR/EulerAMR.cpp:169-590)
 * and the special-relativistic copies (AMR/RelativisticEulerAMR.cpp), which
 * differ only in the 2D refinement test's "bottom" neighbour, taken from the
 * diagonal cell (i + 1, j + 1) (AMR/RelativisticEulerAMR.cpp:422).  Per coarse
 * step: the sticky refinement flags (A1), the masked coarse sweep(s), masked
 * fine sub-steps to the coarse time, restriction of refined coarse cells and
 * injection into unflagged fine cells (A5), then derefinement of quiet
 * coarse cells (A2).
 */
#include "cosmos_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* A sweep where only cells with mask[c] != 0 take the updated value
 * (compute{,X,Y}{FORCE,SLIC}TimeStep{,2D}AMR, AMR/EulerAMR.cpp:3-167): the
 * masked sweep's fluxes read the same boundary copy as the full sweep, so it
 * is the full sweep restricted to the mask. */
static int sweep_masked(int physics, int order, int axis, int dim, double* prim, int nx, int ny, double dx, double dt,
                        double gamma, const unsigned char* mask) {
    long cells = (long)nx * (dim == 1 ? 1 : ny);
    double* upd = (double*)malloc(sizeof(double) * 5 * (size_t)cells);
    if (!upd) return -1;
    memcpy(upd, prim, sizeof(double) * 5 * (size_t)cells);
    int rc = orc_sweep(physics, order, axis, dim, upd, nx, ny, dx, dt, 0.0, 1, gamma);
    for (long c = 0; c < cells; ++c)
        if (mask[c]) memcpy(prim + 5 * c, upd + 5 * c, 40);
    free(upd);
    return rc;
}

static void flux_step_masked(int physics, int order, int dim, double* prim, int nx, int ny, double dx, double dt,
                             double gamma, const unsigned char* mask) {
    if (dim == 1) {
        sweep_masked(physics, order, 0, 1, prim, nx, 1, dx, dt, gamma, mask);
        return;
    }
    sweep_masked(physics, order, 0, 2, prim, nx, ny, dx, dt * 0.5, gamma, mask);
    sweep_masked(physics, order, 1, 2, prim, nx, ny, dx, dt, gamma, mask);
    sweep_masked(physics, order, 0, 2, prim, nx, ny, dx, dt * 0.5, gamma, mask);
}

static int clampi(int c, int n) { return c < 0 ? 0 : (c >= n ? n - 1 : c); }

/* |(U_b - U_a) / dx| >= tol for any component (abs after the division, as in
 * the reference's tests, :215-216). */
static int jump(const double* a, const double* b, double dx, double tol) {
    for (int k = 0; k < 5; ++k)
        if (fabs((b[k] - a[k]) / dx) >= tol) return 1;
    return 0;
}

/* coarse <- c2p(mean of the children's conserved vectors): 1D 0.5 (L + R),
 * 2D 0.25 (((TL + TR) + BL) + BR) (:290-296, :500-509). */
static void restrict_cell(int physics, int dim, const double* fine, int fnx, int i, int j, double gamma,
                          double* out) {
    double u[4][5], m[5];
    if (dim == 1) {
        orc_p2c(physics, fine + 5 * (2L * j), u[0], gamma);
        orc_p2c(physics, fine + 5 * (2L * j + 1), u[1], gamma);
        for (int k = 0; k < 5; ++k) m[k] = 0.5 * (u[0][k] + u[1][k]);
    } else {
        orc_p2c(physics, fine + 5 * ((2L * i) * fnx + 2 * j), u[0], gamma);
        orc_p2c(physics, fine + 5 * ((2L * i) * fnx + 2 * j + 1), u[1], gamma);
        orc_p2c(physics, fine + 5 * ((2L * i + 1) * fnx + 2 * j), u[2], gamma);
        orc_p2c(physics, fine + 5 * ((2L * i + 1) * fnx + 2 * j + 1), u[3], gamma);
        for (int k = 0; k < 5; ++k) m[k] = 0.25 * (((u[0][k] + u[1][k]) + u[2][k]) + u[3][k]);
    }
    orc_c2p(physics, m, out, gamma);
}

/* A1 (AMR/EulerAMR.cpp:202-229, 397-431): a coarse cell whose conserved vector
 * jumps to a neighbour (1-cell copy boundary) by |dU_k| / dx >= tol clears its
 * flag and sets its children's; nothing is ever cleared (sticky).  The
 * special-relativistic 2D copy reads the diagonal cell (i + 1, j + 1) of the
 * boundary-padded grid as "bottom" (AMR/RelativisticEulerAMR.cpp:422), whose
 * bottom-right corner is a default-constructed state (1, 0, 0, 0, 1). */
void orc_amr_refine_flags(int physics, int dim, const double* coarse, int nx, int ny, double dx, double tol,
                          double gamma, unsigned char* cflags, unsigned char* fflags) {
    if (dim == 1) ny = 1;
    const int fnx = 2 * nx;
    const long nc = (long)nx * ny;
    double* U = (double*)malloc(sizeof(double) * 5 * (size_t)nc);
    if (!U) return;
    const double Wdef[5] = {1.0, 0.0, 0.0, 0.0, 1.0};
    double Udef[5];
    orc_p2c(physics, Wdef, Udef, gamma);
    for (long c = 0; c < nc; ++c) orc_p2c(physics, coarse + 5 * c, U + 5 * c, gamma);
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            const double* m = U + 5 * ((long)i * nx + j);
            const double* l = U + 5 * ((long)i * nx + clampi(j - 1, nx));
            const double* r = U + 5 * ((long)i * nx + clampi(j + 1, nx));
            int hit = jump(l, m, dx, tol) || jump(m, r, dx, tol);
            if (dim == 2) {
                const double* top = U + 5 * ((long)clampi(i - 1, ny) * nx + j);
                const double* bot = U + 5 * ((long)clampi(i + 1, ny) * nx + j);
                if (physics == 1) {
                    /* the special-relativistic copy reads the diagonal cell of the
                     * boundary-padded grid, whose bottom-right corner is a
                     * default-constructed state (1, 0, 0, 0, 1) */
                    bot = (i + 1 >= ny && j + 1 >= nx) ? Udef
                                                       : U + 5 * ((long)clampi(i + 1, ny) * nx + clampi(j + 1, nx));
                }
                hit = hit || jump(top, m, dx, tol) || jump(m, bot, dx, tol);
            }
            if (hit) {
                cflags[(long)i * nx + j] = 0;
                for (int a = 0; a < (dim == 1 ? 1 : 2); ++a)
                    for (int b = 0; b < 2; ++b) fflags[(long)(2 * i + a) * fnx + 2 * j + b] = 1;
            }
        }
    free(U);
}

int orc_amr_level1(int physics, int dim, int order, const double* prim, int nx, int ny, double dx, double cfl,
                   double tf, double tol, double gamma, int max_coarse_steps, double* coarse, double* fine,
                   unsigned char* cflags, unsigned char* fflags, double* t_out) {
    if (dim == 1) ny = 1;
    const int fnx = 2 * nx, fny = dim == 1 ? 1 : 2 * ny;
    const long nc = (long)nx * ny, nf = (long)fnx * fny;
    memcpy(coarse, prim, sizeof(double) * 5 * (size_t)nc);
    for (long c = 0; c < nc; ++c) cflags[c] = 1;
    for (long f = 0; f < nf; ++f) fflags[f] = 0;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j)
            for (int a = 0; a < (dim == 1 ? 1 : 2); ++a)
                for (int b = 0; b < 2; ++b)
                    memcpy(fine + 5 * ((long)(2 * i + a) * fnx + 2 * j + b), coarse + 5 * ((long)i * nx + j), 40);
    double t = 0.0;
    int it = 0, fine_total = 0;
    const int cap = max_coarse_steps < 0 ? 0x7fffffff : max_coarse_steps;
    while (t < tf && it < cap) {
        double cdt = orc_stable_dt(physics, dim, coarse, nx, ny, dx, cfl, t, tf, fine_total, gamma);
        /* A1: sticky refinement flags from the coarse state */
        orc_amr_refine_flags(physics, dim, coarse, nx, ny, dx, tol, gamma, cflags, fflags);
        flux_step_masked(physics, order, dim, coarse, nx, ny, dx, cdt, gamma, cflags);
        int refined = 0;
        for (long f = 0; f < nf; ++f) refined |= fflags[f];
        if (refined) {
            double ft = t;
            while (ft < ((t + cdt) - pow(10.0, -4.0))) {
                double fdt = orc_stable_dt(physics, dim, fine, fnx, fny, dx * 0.5, cfl, ft, t + cdt, fine_total, gamma);
                flux_step_masked(physics, order, dim, fine, fnx, fny, dx * 0.5, fdt, gamma, fflags);
                ft += fdt;
                fine_total += 1;
            }
        } else {
            fine_total += 1;
        }
        t += cdt;
        it += 1;
        /* A5: restriction of refined coarse cells, injection into unflagged fine cells */
        for (int i = 0; i < ny; ++i)
            for (int j = 0; j < nx; ++j) {
                double* cc = coarse + 5 * ((long)i * nx + j);
                if (!cflags[(long)i * nx + j]) restrict_cell(physics, dim, fine, fnx, i, j, gamma, cc);
                for (int a = 0; a < (dim == 1 ? 1 : 2); ++a)
                    for (int b = 0; b < 2; ++b) {
                        long f = (long)(2 * i + a) * fnx + 2 * j + b;
                        if (!fflags[f]) memcpy(fine + 5 * f, cc, 40);
                    }
            }
        /* A2: derefinement tested on fine cell (2i+1) [(2i+1, 2j+1)] against its fine
         * neighbours (1-cell copy boundary), over the coarse spacing, at tol / 2 */
        double* FU = (double*)malloc(sizeof(double) * 5 * (size_t)nf);
        if (!FU) return -1;
        for (long f = 0; f < nf; ++f) orc_p2c(physics, fine + 5 * f, FU + 5 * f, gamma);
        for (int i = 0; i < ny; ++i)
            for (int j = 0; j < nx; ++j) {
                int marked;
                long first;
                if (dim == 1) {
                    const double* m = FU + 5 * (2L * j + 1);
                    marked = jump(FU + 5 * (2L * j), m, dx, tol * 0.5) ||
                             jump(m, FU + 5 * clampi(2 * j + 2, fnx), dx, tol * 0.5);
                    first = 2L * j;
                } else {
                    int fi = 2 * i + 1, fj = 2 * j + 1;
                    const double* m = FU + 5 * ((long)fi * fnx + fj);
                    const double* top = FU + 5 * ((long)(fi - 1) * fnx + fj);
                    const double* bot = FU + 5 * ((long)clampi(fi + 1, fny) * fnx + fj);
                    const double* l = FU + 5 * ((long)fi * fnx + fj - 1);
                    const double* r = FU + 5 * ((long)fi * fnx + clampi(fj + 1, fnx));
                    marked = jump(top, m, dx, tol * 0.5) || jump(m, bot, dx, tol * 0.5) ||
                             jump(l, m, dx, tol * 0.5) || jump(m, r, dx, tol * 0.5);
                    first = (2L * i) * fnx + 2 * j;
                }
                if (!marked && fflags[first]) {
                    for (int a = 0; a < (dim == 1 ? 1 : 2); ++a)
                        for (int b = 0; b < 2; ++b) fflags[(long)(2 * i + a) * fnx + 2 * j + b] = 0;
                    cflags[(long)i * nx + j] = 1;
                    restrict_cell(physics, dim, fine, fnx, i, j, gamma, coarse + 5 * ((long)i * nx + j));
                }
            }
        free(FU);
    }
    if (t_out) *t_out = t;
    return it;
}

/* Three-level AMR: EulerAMR::solveLevel2AMR / solve2DLevel2AMR
 * (AMR/EulerAMR.cpp:593-917, 918-1381) and the special-relativistic copies
 * (AMR/RelativisticEulerAMR.cpp:605-1394), which are the same loops (the 2D
 * copy reads the true bottom neighbour, unlike its level-1 driver).  Levels:
 * coarse nx x ny, intermediate 2nx x 2ny, fine 4nx x 4ny (1D: nx, 2nx, 4nx).
 * Per coarse step: coarse refinement flags, the masked coarse flux step, then,
 * if any intermediate flag is set, intermediate sub-steps to the coarse time
 * less 1e-4, each with: intermediate refinement flags (only for flagged cells,
 * at tol * 2 * order), the masked intermediate step, fine sub-steps to the
 * intermediate time less 1e-4 (if any fine flag is set), restriction /
 * injection between intermediate and fine, and fine derefinement (at tol);
 * finally restriction / injection between coarse and intermediate and
 * intermediate derefinement (at tol * 0.5 * (1 / order)).  Every test divides
 * by the coarse spacing.  The iteration counter passed to every stable time
 * step is the reference's fineTotalIteration, advanced as it is there. */

/* A level of nx x ny cells (1D: ny = 1) in AoS primitives, its flags, and its
 * conserved vectors (scratch). */
typedef struct {
    double* W;
    unsigned char* f;
    int nx, ny;
} Level;

static void level_cons(int physics, const Level* L, double gamma, double* U) {
    for (long c = 0; c < (long)L->nx * L->ny; ++c) orc_p2c(physics, L->W + 5 * c, U + 5 * c, gamma);
}

/* parent cell (i, j) <- c2p(mean of its children's conserved vectors) */
static void restrict_to(int physics, int dim, const Level* P, const Level* C, int i, int j, double gamma) {
    restrict_cell(physics, dim, C->W, C->nx, i, j, gamma, P->W + 5 * ((long)i * P->nx + j));
}

static int children_all(int dim, const Level* C, int i, int j) {
    if (dim == 1) return C->f[2 * j] && C->f[2 * j + 1];
    const long a = (long)(2 * i) * C->nx + 2 * j;
    return C->f[a] && C->f[a + 1] && C->f[a + C->nx] && C->f[a + C->nx + 1];
}

static void children_set(int dim, const Level* C, int i, int j, unsigned char v) {
    for (int a = 0; a < (dim == 1 ? 1 : 2); ++a)
        for (int b = 0; b < 2; ++b) C->f[(long)(2 * i + a) * C->nx + 2 * j + b] = v;
}

/* Refinement flags of P from its own state (copy boundary, the four (two)
 * neighbours); only_flagged: cells whose flag is clear are not tested. */
static void l2_refine(int physics, int dim, const Level* P, const Level* C, double dx, double tol, int only_flagged,
                      double gamma, double* U) {
    level_cons(physics, P, gamma, U);
    for (int i = 0; i < P->ny; ++i)
        for (int j = 0; j < P->nx; ++j) {
            const long c = (long)i * P->nx + j;
            if (only_flagged && !P->f[c]) continue;
            const double* m = U + 5 * c;
            int hit = jump(U + 5 * ((long)i * P->nx + clampi(j - 1, P->nx)), m, dx, tol) ||
                      jump(m, U + 5 * ((long)i * P->nx + clampi(j + 1, P->nx)), dx, tol);
            if (dim == 2)
                hit = hit || jump(U + 5 * ((long)clampi(i - 1, P->ny) * P->nx + j), m, dx, tol) ||
                      jump(m, U + 5 * ((long)clampi(i + 1, P->ny) * P->nx + j), dx, tol);
            if (hit) {
                P->f[c] = 0;
                children_set(dim, C, i, j, 1);
            }
        }
}

/* Restriction of refined parents and injection into unflagged children;
 * strict: only parents with all children flagged, only under flagged parents. */
static void l2_exchange(int physics, int dim, const Level* P, const Level* C, int strict, double gamma) {
    for (int i = 0; i < P->ny; ++i)
        for (int j = 0; j < P->nx; ++j) {
            const long c = (long)i * P->nx + j;
            if (!P->f[c] && (!strict || children_all(dim, C, i, j))) restrict_to(physics, dim, P, C, i, j, gamma);
            for (int a = 0; a < (dim == 1 ? 1 : 2); ++a)
                for (int b = 0; b < 2; ++b) {
                    const long f = (long)(2 * i + a) * C->nx + 2 * j + b;
                    if (!C->f[f] && (!strict || P->f[c])) memcpy(C->W + 5 * f, P->W + 5 * c, 40);
                }
        }
}

/* Derefinement: child (2i+1, 2j+1) (1D: 2j+1) against its neighbours (copy
 * boundary) at tol; quiet parents with all children flagged take their
 * restriction back. */
static void l2_derefine(int physics, int dim, const Level* P, const Level* C, double dx, double tol, double gamma,
                        double* U) {
    level_cons(physics, C, gamma, U);
    for (int i = 0; i < P->ny; ++i)
        for (int j = 0; j < P->nx; ++j) {
            int marked;
            if (dim == 1) {
                const double* m = U + 5 * (2L * j + 1);
                marked = jump(U + 5 * (2L * j), m, dx, tol) || jump(m, U + 5L * clampi(2 * j + 2, C->nx), dx, tol);
            } else {
                const int fi = 2 * i + 1, fj = 2 * j + 1;
                const double* m = U + 5 * ((long)fi * C->nx + fj);
                marked = jump(U + 5 * ((long)(fi - 1) * C->nx + fj), m, dx, tol) ||
                         jump(m, U + 5 * ((long)clampi(fi + 1, C->ny) * C->nx + fj), dx, tol) ||
                         jump(U + 5 * ((long)fi * C->nx + fj - 1), m, dx, tol) ||
                         jump(m, U + 5 * ((long)fi * C->nx + clampi(fj + 1, C->nx)), dx, tol);
            }
            if (!marked && children_all(dim, C, i, j)) {
                children_set(dim, C, i, j, 0);
                P->f[(long)i * P->nx + j] = 1;
                restrict_to(physics, dim, P, C, i, j, gamma);
            }
        }
}

static int any_flag(const Level* L) {
    for (long c = 0; c < (long)L->nx * L->ny; ++c)
        if (L->f[c]) return 1;
    return 0;
}

int orc_amr_level2(int physics, int dim, int order, const double* prim, int nx, int ny, double dx, double cfl,
                   double tf, double tol, double gamma, int max_coarse_steps, double* coarse, double* inter,
                   double* fine, unsigned char* cflags, unsigned char* iflags, unsigned char* fflags, double* t_out) {
    if (dim == 1) ny = 1;
    const int s = dim == 1 ? 1 : 2; /* rows per child row pair */
    Level C = {coarse, cflags, nx, ny}, M = {inter, iflags, 2 * nx, s * ny}, F = {fine, fflags, 4 * nx, s * s * ny};
    const long nc = (long)nx * ny, nm = (long)M.nx * M.ny, nf = (long)F.nx * F.ny;
    memcpy(coarse, prim, sizeof(double) * 5 * (size_t)nc);
    for (long c = 0; c < nc; ++c) cflags[c] = 1;
    memset(iflags, 0, (size_t)nm);
    memset(fflags, 0, (size_t)nf);
    for (int i = 0; i < M.ny; ++i)
        for (int j = 0; j < M.nx; ++j)
            memcpy(inter + 5 * ((long)i * M.nx + j), coarse + 5 * ((long)(i / s) * nx + j / 2), 40);
    for (int i = 0; i < F.ny; ++i)
        for (int j = 0; j < F.nx; ++j)
            memcpy(fine + 5 * ((long)i * F.nx + j), coarse + 5 * ((long)(i / (s * s)) * nx + j / 4), 40);
    double* U = (double*)malloc(sizeof(double) * 5 * (size_t)nf);
    if (!U) return -1;
    const double itol = (tol * 2.0) * order;         /* AMRTolerance * 2.0 * order */
    const double dtol = (tol * 0.5) * (1.0 / order); /* AMRTolerance * 0.5 * (1.0 / order) */
    double t = 0.0;
    int it = 0, fine_total = 0;
    const int cap = max_coarse_steps < 0 ? 0x7fffffff : max_coarse_steps;
    while (t < tf && it < cap) {
        const double cdt = orc_stable_dt(physics, dim, coarse, nx, ny, dx, cfl, t, tf, fine_total, gamma);
        l2_refine(physics, dim, &C, &M, dx, tol, 0, gamma, U);
        flux_step_masked(physics, order, dim, coarse, nx, ny, dx, cdt, gamma, cflags);
        if (any_flag(&M)) {
            double mt = t;
            while (mt < ((t + cdt) - pow(10.0, -4.0))) {
                const double mdt = orc_stable_dt(physics, dim, inter, M.nx, M.ny, dx * 0.5, cfl, mt, t + cdt,
                                                 fine_total, gamma);
                l2_refine(physics, dim, &M, &F, dx, itol, 1, gamma, U);
                flux_step_masked(physics, order, dim, inter, M.nx, M.ny, dx * 0.5, mdt, gamma, iflags);
                if (any_flag(&F)) {
                    double ft = mt;
                    while (ft < ((mt + mdt) - pow(10.0, -4.0))) {
                        const double fdt = orc_stable_dt(physics, dim, fine, F.nx, F.ny, dx * 0.25, cfl, ft, mt + mdt,
                                                         fine_total, gamma);
                        flux_step_masked(physics, order, dim, fine, F.nx, F.ny, dx * 0.25, fdt, gamma, fflags);
                        ft += fdt;
                        fine_total += 1;
                    }
                } else {
                    fine_total += 1;
                }
                mt += mdt;
                fine_total += 1;
                l2_exchange(physics, dim, &M, &F, 1, gamma);
                l2_derefine(physics, dim, &M, &F, dx, tol, gamma, U);
            }
        } else {
            fine_total += 1;
        }
        t += cdt;
        it += 1;
        l2_exchange(physics, dim, &C, &M, 0, gamma);
        l2_derefine(physics, dim, &C, &M, dx, dtol, gamma, U);
    }
    free(U);
    if (t_out) *t_out = t;
    return it;
}
