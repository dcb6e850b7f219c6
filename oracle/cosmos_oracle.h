/* CPU restatement of the Cosmos finite-volume arithmetic (test infrastructure).
 *
 * Only tests, the golden-fixture scripts and the CPU-baseline leg of bench.py
 * load this library, as the checker.  The product path never calls it.
 *
 * States are flat row-major AoS arrays of primitives,
 * prim[(i * nx + j) * 5 + k], i = row (y), j = column (x),
 * k in (rho, u, v, w, p).  physics: 0 = Newtonian, 1 = special relativistic.
 * Every function has the same signature as its ref_* counterpart in
 * ref_driver.cpp so that the two can be compared call for call.
 */
#ifndef COSMOS_ORACLE_H
#define COSMOS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

int orc_sweep(int physics, int order, int axis, int dim, double* prim, int nx, int ny, double dx,
              double dt, double bias, int limiter, double gamma);
double orc_max_wavespeed(int physics, int dim, const double* prim, int nx, int ny, double gamma);
double orc_stable_dt(int physics, int dim, const double* prim, int nx, int ny, double dx, double cfl,
                     double t, double tf, int iteration, double gamma);
double orc_stable_dt_scalar(double dt, double t, double tf, int iteration);
int orc_run(int physics, int order, int dim, double* prim, int nx, int ny, double dx, double cfl,
            double tf, int max_steps, double bias, int limiter, double gamma, double* t_out);
void orc_slic_flux(int physics, int axis, const double* ll, const double* l, const double* r,
                   const double* rr, double dx, double dt, double bias, int limiter, double gamma,
                   double* out);
void orc_force_flux(int physics, int axis, const double* l, const double* r, double dx, double dt,
                    double gamma, double* out);
void orc_half_step(int physics, int axis, const double* l, const double* m, const double* r, double dx,
                   double dt, double bias, int limiter, int side, double gamma, double* out);
void orc_p2c(int physics, const double* w, double* u, double gamma);
void orc_c2p(int physics, const double* u, double* w, double gamma);
void orc_flux(int physics, int axis, const double* u, double* f, double gamma);
double orc_limiter(double r, double bias, int limiter);
void orc_slope(const double* ul, const double* um, const double* ur, double bias, int limiter, double* out);

/* General relativistic (Kerr-Schild black hole of mass M, spin a, at (bx, by, bz)).
 * Cell (i, j) of a 2D grid sits at (j*h, i*h); a 1D cell i at (i*h, 0). */
int orc_gr_sweep(int order, int axis, int dim, double* prim, int nx, int ny, double h, double dt, double bias,
                 int limiter, double gamma, double M, double a, double bx, double by, double bz);
int orc_gr_forcing(int dim, double* prim, int nx, int ny, double h, double dt, double bias, int limiter, double gamma,
                   double M, double a, double bx, double by, double bz);
double orc_gr_max_wavespeed(int dim, const double* prim, int nx, int ny, double gamma, double M, double a, double bx,
                            double by, double bz);
int orc_gr_run(int order, int dim, double* prim, int nx, int ny, double h, double cfl, double tf, int max_steps,
               double bias, int limiter, int subcycles, double gamma, double M, double a, double bx, double by,
               double bz, double* t_out);
/* BlackHoleEulerAMR::solve{,2D}Level1AMR (two levels, 1D/2D) and
 * solve{,2D}Level2AMR (three levels; the 2D one completed past the
 * reference's truncation); outputs as orc_amr_level1 / orc_amr_level2. */
int orc_gr_amr_level1(int order, int dim, const double* prim, int nx, int ny, double dx, double cfl, double tf,
                      double tol, int max_coarse_steps, double gamma, double M, double a, double bx, double by,
                      double bz, double* coarse, double* fine, unsigned char* cflags, unsigned char* fflags,
                      double* t_out);
int orc_gr_amr_level2(int order, int dim, const double* prim, int nx, int ny, double dx, double cfl, double tf,
                      double tol, int max_coarse_steps, double gamma, double M, double a, double bx, double by,
                      double bz, double* coarse, double* inter, double* fine, unsigned char* cflags,
                      unsigned char* iflags, unsigned char* fflags, double* t_out);
/* what: 0 p2c, 1 c2p, 2 x flux, 3 y flux, 4 source term, 5 sound speed, 6 metric pieces (44 values) */
void orc_gr_point(int what, const double* in, double x, double y, double z, double h, double gamma, double M,
                  double a, double bx, double by, double bz, double* out);

/* Single interface fluxes of the first-order solvers (Solvers/EulerFirstOrderSolver.cpp:27-85). */
void orc_lf_flux(int physics, int axis, const double* l, const double* r, double dx, double dt, double gamma,
                 double* out);
void orc_richtmyer_flux(int physics, int axis, const double* l, const double* r, double dx, double dt,
                        double gamma, double* out);

/* Point-level statics of the solver classes: s = up to five 5-vectors (primitive
 * states unless noted), par = {dx, dt, bias, step fraction, gamma}; physics 0, 1
 * or 2 (black hole: spacetime = {M, a, bx, by, bz}, pos = {x, y, z}).
 *   P2C, C2P (s0 conserved), FLUX (s0 conserved): the state conversions
 *   LF, RICHTMYER, FORCE (s0, s1): compute{X,Y}...Flux(left, right) (pos = the interface)
 *   SLIC (s0..s3), HALF (s0..s2, side; pos = the middle centre), SLOPE (s0..s2)
 *   FRACTIONAL (s0..s2; black hole), FORCING (s0..s4 conserved L, M, R, T, B; black
 *   hole, dimension axis + 1)
 *   VEC_*: the vector arithmetic on s0.. (conserved / flux vectors), LIMITER: phi(s0[0]) */
enum {
    ORC_PT_P2C, ORC_PT_C2P, ORC_PT_FLUX, ORC_PT_LF, ORC_PT_RICHTMYER, ORC_PT_FORCE, ORC_PT_SLIC, ORC_PT_HALF,
    ORC_PT_SLOPE, ORC_PT_FRACTIONAL, ORC_PT_FORCING,
    ORC_PT_VEC_LF, ORC_PT_VEC_RICHTMYER, ORC_PT_VEC_FORCE, ORC_PT_VEC_UPDATE, ORC_PT_VEC_EVOLUTION,
    ORC_PT_VEC_SLOPE, ORC_PT_LIMITER
};
void orc_point(int fn, int physics, int axis, const double* s, const double* par, int lim, int side,
               const double* spacetime, const double* pos, double* out);

/* Two-level AMR (EulerAMR::solve{,2D}Level1AMR and the special-relativistic copies):
 * coarse nx x ny, fine 2nx x 2ny, flags as 0/1 bytes; at most max_coarse_steps coarse
 * steps (< 0: none).  Returns the coarse steps taken. */
int orc_amr_level1(int physics, int dim, int order, const double* prim, int nx, int ny, double dx, double cfl,
                   double tf, double tol, double gamma, int max_coarse_steps, double* coarse, double* fine,
                   unsigned char* cflags, unsigned char* fflags, double* t_out);

/* Three-level AMR (EulerAMR::solve{,2D}Level2AMR and the special-relativistic copies):
 * coarse nx x ny, intermediate 2nx x 2ny, fine 4nx x 4ny (1D: nx, 2nx, 4nx), flags as
 * 0/1 bytes.  Returns the coarse steps taken. */
int orc_amr_level2(int physics, int dim, int order, const double* prim, int nx, int ny, double dx, double cfl,
                   double tf, double tol, double gamma, int max_coarse_steps, double* coarse, double* inter,
                   double* fine, unsigned char* cflags, unsigned char* iflags, unsigned char* fflags, double* t_out);

/* One sticky refinement-flag pass over a coarse state (flags in/out). */
void orc_amr_refine_flags(int physics, int dim, const double* coarse, int nx, int ny, double dx, double tol,
                          double gamma, unsigned char* cflags, unsigned char* fflags);

/* Excision radius override (<= 0: the reference's horizon radius M (1 + sqrt(1 - a^2))). */
void orc_gr_set_excision(double radius);

/* Special-relativistic con2prim Newton controls.  The reference hard-codes a
 * stopping tolerance of 1e-8 and at most 1000 iterations
 * (RelativisticEuler/RelativisticEulerStateVector.cpp:59-80); other values
 * select the tighter "config mode".  Returns the previous settings. */
void orc_set_c2p(double tol, int max_iterations);
/* Con2prim event counters of the special-relativistic solves since the last
 * reset: out[0] iteration-cap hits, out[1] q floors, out[2]
 * density floors, out[3] pressure floors (each distinct solve counted once). */
void orc_events_reset(void);
void orc_events_read(long long* out);
/* Black-hole con2prim controls: tolerance > 0 replaces every tolerance of the
 * reference's (1e-10 in the x flux, 1e-8 elsewhere); 0 restores them.  The
 * iteration cap (reference: 1000) applies to every black-hole solve. */
void orc_gr_set_c2p(double tolerance, int max_iterations);
/* Number of OpenMP threads used by sweeps (0 = runtime default). */
void orc_set_threads(int n);
int orc_thread_count(void);
/* Per-step NaN cell counts of orc_gr_run (buf[k] = cells holding a NaN after
 * step k; NULL stops recording). */
void orc_set_step_nan_counts(int* buf, int capacity);
/* Black-hole solve loop steps with a given time-step schedule dts[0 .. n). */
void orc_gr_run_dts(int order, int dim, double* prim, int nx, int ny, double h, const double* dts, int n, double bias,
                    int limiter, int subcycles, double gamma, double M, double a, double bx, double by, double bz);
/* Branch trace of the slope limiter: while on, each slope component evaluated
 * appends one byte to buf (bit 0: |dl| < 1e-5 clamped, bit 1: |dr| clamped,
 * bits 2-4: the limiter's piece); meaningful with one thread.  orc_trace_end
 * returns the number of components evaluated (bytes beyond capacity are
 * counted, not stored). */
void orc_trace_begin(unsigned char* buf, long capacity);
/* Replay a recorded trace: each slope component takes the clamps and limiter
 * piece recorded for it (the forward map on the recorded run's pieces). */
void orc_trace_replay(unsigned char* buf, long length);
long orc_trace_end(void);

#ifdef __cplusplus
}
#endif
#endif
