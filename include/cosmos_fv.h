/* cosmos_fv.h -- C ABI of the MI355X (gfx950) finite-volume update.
 *
 * The reference (Cosmos) exposes this path as static C++ class methods, with
 * no FFI.  Each entry point below replaces one of those methods; the comment
 * names the reference interface (file:line).  The C++ drop-in classes in
 * cosmos_amd/dropin/ keep the reference's class names and signatures and are
 * implemented on top of these functions.
 *
 * Conventions
 *   - A solver handle owns the device state of one grid: five SoA FP64
 *     planes of primitives (rho, u, v, w, p), the state of record as in the
 *     reference, plus a second buffer for the sweep output.
 *   - Host arrays are row-major AoS primitives prim[(i * nx + j) * 5 + k],
 *     i = row (y), j = column (x): vector<vector<StateVector>> flattened.
 *   - physics: COSMOS_NEWTONIAN (Euler/), COSMOS_RELATIVISTIC (RelativisticEuler/),
 *     COSMOS_BLACKHOLE (BlackHoleEuler/: general relativistic on a fixed
 *     Kerr-Schild background; cell (i, j) sits at (j*h, i*h), a 1D cell j at
 *     (j*h, 0), with h the cell_spacing argument, as in the reference).
 *   - limiter: 0 SuperBee, 1 van Leer, anything else MinBee
 *     (SlopeLimiters::computeSlopeLimiter, Solvers/SlopeLimiters.cpp:55-69).
 *   - Every function returns COSMOS_OK (0) or a negative code and never
 *     throws; cosmos_last_error() describes the last failure.  Numerical
 *     faults (NaN, con2prim cap) propagate as values, as in the reference.
 *   - A handle is not thread-safe; separate handles are independent.  Work is
 *     queued on the handle's HIP stream; functions that return host values
 *     synchronise it.
 *   - All arithmetic is FP64 with the reference's operation order; results
 *     are bitwise equal to the reference CPU path.
 */
#ifndef COSMOS_FV_H
#define COSMOS_FV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COSMOS_OK 0
#define COSMOS_ERR_ARGUMENT (-1)
#define COSMOS_ERR_DEVICE (-2)
#define COSMOS_ERR_STATE (-3)

#define COSMOS_NEWTONIAN 0
#define COSMOS_RELATIVISTIC 1
#define COSMOS_BLACKHOLE 2

#define COSMOS_MODE_EXACT 0
#define COSMOS_MODE_FAST 1

typedef struct cosmos_solver cosmos_solver;

/* Create a solver for an nx-by-ny grid (dim 2) or an nx-cell line (dim 1,
 * ny ignored).  adiabatic_index plays EulerMaterialParameters
 * (Euler/EulerMaterialParameters.cpp:3-21). */
int cosmos_create(cosmos_solver** out, int physics, int dim, int nx, int ny, double adiabatic_index);
int cosmos_destroy(cosmos_solver* s);

/* Message for the last failure on s (s == NULL: the last cosmos_create
 * failure on this thread). */
const char* cosmos_last_error(const cosmos_solver* s);

/* Arithmetic mode (no counterpart in the reference).  COSMOS_MODE_EXACT (the
 * default) keeps the reference's operation order: results are bitwise equal
 * to the reference CPU path.  COSMOS_MODE_FAST re-associates the Newtonian
 * second-order van Leer (bias 0) sweeps for throughput (one reciprocal per
 * primitive recovery, no prim->cons->prim round trips); results then differ
 * from the reference by rounding only, bounded by the tests at 1e-11 relative
 * after the 100-step 4096^2 workload.  Other physics/scheme combinations
 * always run exact. */
int cosmos_set_mode(cosmos_solver* s, int mode);

/* Launch shape (no counterpart in the reference; results do not depend on it):
 *   y_rows   rows of the grid each y-sweep thread marches down (16, 32 or 64;
 *            0: the default, 64 for Newtonian and 16 for special-relativistic)
 *   ws_mode  how the solve loop's fused wave-speed maximum leaves the last
 *            sweep of a step: 1 one partial per workgroup, reduced by the
 *            next step's head; 2 one atomic maximum per wavefront into 32
 *            slots; 0: the default (2). */
int cosmos_set_launch(cosmos_solver* s, int y_rows, int ws_mode);
/* y-sweep layout (no counterpart in the reference; results do not depend on
 * it): 0 one thread per column marching down y_rows rows with a register
 * window; 8 or 16: one thread per cell over LDS tiles of that many columns
 * (re-evaluating 4 halo rows per tile for SLIC, at the x sweep's occupancy);
 * -1: the default (for Newtonian the march where it runs at least 1024
 * blocks, else 16-column tiles; 8-column tiles for special relativistic). */
int cosmos_set_y_tile(cosmos_solver* s, int columns);
/* Step fusion (no counterpart in the reference): with 1 (the default), a 2D
 * fast-mode Newtonian SLIC step with van Leer and bias 0 queued by cosmos_begin
 * / cosmos_advance / cosmos_solve on a grid of at least 8 x 8 cells runs its
 * X(dt/2) Y(dt) X(dt/2) sweeps as one kernel that reads and writes each cell
 * once and hands conserved vectors between its sweeps (the result differs from
 * the three sweeps' by rounding; both stay within the fast mode's tolerance of
 * the reference); 0 queues the three sweeps.  rows_per_block:
 * output rows of each block of the fused kernel, a launch shape only (0: chosen
 * from the grid and the kernel's occupancy; else 4 .. 1024). */
int cosmos_set_step_fusion(cosmos_solver* s, int enable, int rows_per_block);

/* Con2prim controls.  Special relativistic: the stopping tolerance and
 * iteration cap (default: the reference's 1e-8 and 1000,
 * RelativisticEuler/RelativisticEulerStateVector.cpp:59-80).  Black hole: every
 * con2prim (fluxes, source, state) stops at `tolerance` instead of the
 * reference's fixed 1e-10 (x fluxes) / 1e-8 (elsewhere), with the same cap;
 * without this call the reference's values apply.  A tight tolerance keeps
 * the Newton iteration count from changing under small perturbations, which
 * finite-difference checks of gradients need (no counterpart in the
 * reference for the black-hole case). */
int cosmos_set_con2prim(cosmos_solver* s, double tolerance, int max_iterations);

/* The black hole of a COSMOS_BLACKHOLE solver: BlackHoleSpacetime(mass, spin,
 * x, y, z) (BlackHoleEuler/BlackHoleSpacetime.cpp:43-52; the default, as the
 * reference's default constructor :3-11, is mass 1, spin 0 at (0.5, 0.5, 0)).
 * spin is dimensionless (the metric uses mass * spin). */
int cosmos_set_spacetime(cosmos_solver* s, double mass, double spin, double x, double y, double z);

/* Excision radius of a COSMOS_BLACKHOLE solver (no counterpart in the
 * reference, which always excises r <= M (1 + sqrt(1 - spin^2)),
 * BlackHoleEuler/BlackHoleSpacetime.cpp:372-386).  radius > 0 replaces that
 * radius (same <= test); radius <= 0 restores it. */
int cosmos_set_excision(cosmos_solver* s, double radius);
/* Black-hole geometry cache (no counterpart in the reference; on by default):
 * the position-only metric values (lapse, shift, spatial metric and its
 * determinant at every cell centre and interface, and the source term's
 * derivatives and extrinsic curvature at every centre) tabulated once per
 * spacing and spacetime in device memory, about 1 KB per cell, instead of
 * re-derived at every use.  Results are bitwise those of the uncached
 * evaluation: the interface and neighbour tables are used only when every
 * position the kernels form is exact (a spacing with a short binary
 * significand, e.g. 40 / 1024), the source table always.  0 disables it. */
int cosmos_set_geometry_cache(cosmos_solver* s, int enable);

/* State transfer.  *_device variants take device pointers to the same AoS
 * layout (for callers that already hold the grid in GPU memory).  Host
 * arrays in page-locked memory (cosmos_host_alloc) are copied by DMA
 * directly; pageable arrays are staged through two page-locked 32 MiB chunks
 * per solver, the CPU copy of one overlapping the DMA of the other. */
int cosmos_upload(cosmos_solver* s, const double* prim);
int cosmos_download(cosmos_solver* s, double* prim);
int cosmos_upload_device(cosmos_solver* s, const double* prim_device);
int cosmos_download_device(cosmos_solver* s, double* prim_device);
/* Page-locked host memory (hipHostMalloc) for state arrays; no counterpart in
 * the reference, whose grids are std::vector.  Free with cosmos_host_free. */
int cosmos_host_alloc(size_t bytes, void** out);
int cosmos_host_free(void* p);

/* One dimensionally split sweep in place on the device state.
 *   order 2, dim 2: EulerSecondOrderSolver::compute{X,Y}SLICTimeStep2D
 *                   (Solvers/EulerSecondOrderSolver.cpp:48-100)
 *   order 2, dim 1: EulerSecondOrderSolver::computeSLICTimeStep (:27-46)
 *   order 1, dim 2: EulerFirstOrderSolver::compute{X,Y}FORCETimeStep2D
 *                   (Solvers/EulerFirstOrderSolver.cpp:120-174)
 *   order 1, dim 1: EulerFirstOrderSolver::computeFORCETimeStep (:103-118)
 * and the RelativisticEuler* and BlackHoleEuler* counterparts
 * (Solvers/RelativisticEulerSecondOrderSolver.cpp, Solvers/RelativisticEulerFirstOrderSolver.cpp,
 * Solvers/BlackHoleEulerSecondOrderSolver.cpp:37-143, Solvers/BlackHoleEulerFirstOrderSolver.cpp:101-181;
 * black-hole cells whose stencil touches the excision region keep their state).
 * axis 0 = x, 1 = y.  time_step is the sweep's step (the caller passes
 * 0.5 * dt for the outer Strang sweeps, as solve2D does). */
int cosmos_sweep(cosmos_solver* s, int order, int axis, double cell_spacing, double time_step, double bias,
                 int slope_limiter);

/* EulerSolvers::computeMaximumWaveSpeed{,2D} (Solvers/EulerSolvers.cpp:87-128),
 * RelativisticEulerSolvers::computeMaximumWaveSpeed{,2D}
 * (Solvers/RelativisticEulerSolvers.cpp:89-130),
 * BlackHoleEulerSolvers::computeMaximumWaveSpeed{,2D} (Solvers/BlackHoleEulerSolvers.cpp:88-140;
 * like the reference it places the cells on the unit square for this, whatever
 * the cell spacing, with x and y exchanged in the eigenvalue evaluation). */
int cosmos_max_wave_speed(cosmos_solver* s, double* out);

/* EulerSolvers::computeStableTimeStep{,2D} (Solvers/EulerSolvers.cpp:130-161) and its
 * RelativisticEuler / BlackHoleEuler counterparts (Solvers/BlackHoleEulerSolvers.cpp:142-157). */
int cosmos_stable_time_step(cosmos_solver* s, double cell_spacing, double cfl, double current_time,
                            double final_time, int current_iteration, double* out);

/* The loop of solve / solve2D (Solvers/EulerSecondOrderSolver.cpp:133-178,
 * Solvers/EulerFirstOrderSolver.cpp:220-265 and the RelativisticEuler*
 * counterparts) run on the device: while t < final_time, dt from the CFL
 * condition (with the reference's 0.2x damping of iterations 0..5 and the
 * clamp to final_time), then X(dt/2) Y(dt) X(dt/2) in 2D or one sweep in 1D.
 *   cosmos_begin   resets t and the iteration counter (solve starts at 0)
 *   cosmos_advance queues up to n more steps, asynchronously
 *   cosmos_progress synchronises and reports t, steps taken and whether the
 *                  loop has stopped (t >= final_time or max_steps reached)
 * cosmos_solve = begin + advance until stopped.  max_steps < 0 means no cap;
 * final_time = +inf with a cap gives a fixed number of steps. */
int cosmos_begin(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time, int max_steps,
                 double bias, int slope_limiter);
int cosmos_advance(cosmos_solver* s, int n_steps);
int cosmos_progress(cosmos_solver* s, double* time, int* steps, int* stopped);
int cosmos_solve(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time, int max_steps,
                 double bias, int slope_limiter, double* time_out, int* steps_out);
/* cosmos_begin for a resumed run (no counterpart in the reference, whose loops
 * always start at t = 0, iteration 0): the loop starts at start_time with the
 * iteration counter at start_iteration (the first six iterations' 0.2x dt
 * damping keys on it, Solvers/EulerSolvers.cpp:134-137), and max_steps caps the
 * steps taken from there.  Resuming from a checkpoint of step k with the same
 * final time continues the run bitwise. */
int cosmos_begin_at(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time, int max_steps,
                    double bias, int slope_limiter, double start_time, int start_iteration);

/* COSMOS_BLACKHOLE only.  One RK4 source-term step in place:
 * BlackHoleEulerForcingSolver::computeRungeKuttaTimeStep{,2D}
 * (Solvers/BlackHoleEulerForcingSolver.cpp:187-250). */
int cosmos_forcing(cosmos_solver* s, double cell_spacing, double time_step, double bias, int slope_limiter);

/* COSMOS_BLACKHOLE only.  The loop of BlackHoleEuler{First,Second}OrderSolver::solve{,2D}
 * (Solvers/BlackHoleEulerFirstOrderSolver.cpp:183-253,
 * Solvers/BlackHoleEulerSecondOrderSolver.cpp:145-218): per step dt from the CFL
 * condition, `subcycles` forcing steps of dt / (2 subcycles) (bias 0 and van
 * Leer, as the reference hard-codes), the flux step (Strang X Y X in 2D), and
 * the same forcing again.  max_steps < 0 means no cap.  The host reads the
 * wave speed once per step. */
int cosmos_solve_forced(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time,
                        int max_steps, double bias, int slope_limiter, int subcycles, double* time_out,
                        int* steps_out);
/* The same loop resumed at start_time / start_iteration (as cosmos_begin_at);
 * steps_out counts the steps taken by this call. */
int cosmos_solve_forced_at(cosmos_solver* s, int order, double cell_spacing, double cfl, double final_time,
                           int max_steps, double bias, int slope_limiter, int subcycles, double start_time,
                           int start_iteration, double* time_out, int* steps_out);

/* ------------------------------------------------------------------------
 * Checkpoint / restart and text output.
 * cosmos_checkpoint_save writes the solver's state losslessly (FP64 bits) with
 * the run's time and iteration counter, the adiabatic index and, for a black
 * hole, the spacetime and excision radius; cosmos_checkpoint_load checks the
 * file (magic, version, physics, grid, adiabatic index, checksum) against the
 * solver, restores the state (and spacetime) and returns the time and
 * iteration to pass to cosmos_begin_at / cosmos_solve_forced_at.  No
 * counterpart in the reference, which keeps no restart state.
 * cosmos_write_dat writes density.dat, xVelocity.dat, yVelocity.dat,
 * zVelocity.dat and pressure.dat into `directory` in the layout of
 * EulerTests::outputSolution{,2D} (Tests/EulerTests.cpp:189-252), shared by the
 * relativistic tests, and, for a black hole, of
 * BlackHoleEulerTests::outputSolution{,2D} (Tests/BlackHoleEulerTests.cpp:201-286):
 * excised cells written as 0.
 * ------------------------------------------------------------------------ */
int cosmos_checkpoint_save(cosmos_solver* s, const char* path, double time, int iteration);
int cosmos_checkpoint_load(cosmos_solver* s, const char* path, double* time, int* iteration);
int cosmos_write_dat(cosmos_solver* s, const char* directory);

/* ------------------------------------------------------------------------
 * Two-level adaptive mesh refinement (AMR/EulerAMR.cpp, AMR/RelativisticEulerAMR.cpp).
 * A cosmos_amr context holds a coarse nx-by-ny grid, a fine 2nx-by-2ny grid (1D:
 * nx and 2nx cells) and one refinement flag per cell of each (bytes 0/1, row
 * major): a coarse cell whose flag is set is advanced on the coarse grid, a
 * fine cell whose flag is set on the fine grid.
 * ------------------------------------------------------------------------ */
typedef struct cosmos_amr cosmos_amr;

int cosmos_amr_create(cosmos_amr** out, int physics, int dim, int nx, int ny, double adiabatic_index);
int cosmos_amr_destroy(cosmos_amr* a);
const char* cosmos_amr_last_error(const cosmos_amr* a);
/* as cosmos_set_con2prim, for both levels */
int cosmos_amr_set_con2prim(cosmos_amr* a, double tolerance, int max_iterations);
/* No counterpart in the reference (its drivers print status lines instead).
 * While enabled, every flag, restriction / injection, derefinement and masked-
 * sweep merge kernel is bracketed by HIP events on the stream it runs on;
 * cosmos_amr_profile_read synchronises and returns, per kind (0 refinement
 * flags, 1 restriction / injection, 2 derefinement, 3 masked-sweep merge), the
 * summed kernel milliseconds and launch counts since the last read. */
int cosmos_amr_profile_enable(cosmos_amr* a, int on);
int cosmos_amr_profile_read(cosmos_amr* a, double* total_ms, int* launches);
/* The solver handle holding level 0 (coarse), 1 or 2 (the three-level run's
 * finest; null before cosmos_amr_solve_level2) of the context, for its stream,
 * state planes and the solver API; owned by the context.  No counterpart in the
 * reference (its drivers return the levels as vectors). */
cosmos_solver* cosmos_amr_level_solver(cosmos_amr* a, int level);
/* COSMOS_BLACKHOLE contexts: as cosmos_set_spacetime / cosmos_set_excision,
 * for every level (the blackHole argument of BlackHoleEulerAMR's drivers,
 * AMR/BlackHoleEulerAMR.cpp:239-240).  The flag, restriction and
 * derefinement passes convert each state at the cell coordinate the
 * reference's call site names (index times the level's spacing, 1D y = 0). */
int cosmos_amr_set_spacetime(cosmos_amr* a, double mass, double spin, double x, double y, double z);
int cosmos_amr_set_excision(cosmos_amr* a, double radius);
/* The coarse state (AoS primitives); the fine grid becomes a copy of each coarse
 * cell in its children, coarse flags all set, fine flags all clear
 * (the set-up of solve{,2D}Level1AMR, AMR/EulerAMR.cpp:174-200, 356-392). */
int cosmos_amr_upload(cosmos_amr* a, const double* prim);
/* Any of the outputs may be NULL. */
int cosmos_amr_download(cosmos_amr* a, double* coarse, double* fine, unsigned char* coarse_flags,
                        unsigned char* fine_flags);
/* Replace the coarse and / or fine flags (0/1 bytes, row-major; a null pointer
 * leaves that level's flags): the AMRStructure argument of the reference's
 * masked sweeps (computeFORCETimeStepAMR and siblings, AMR/EulerAMR.cpp:3-167). */
int cosmos_amr_set_flags(cosmos_amr* a, const unsigned char* coarse_flags, const unsigned char* fine_flags);
/* Refinement flags from the coarse state: a coarse cell with any conserved
 * component k jumping by |dU_k| / cell_spacing >= tolerance to a neighbour
 * (left, right; 2D also top and bottom, 1-cell copy boundary) clears its own
 * flag and sets its children's; flags are never cleared here
 * (AMR/EulerAMR.cpp:202-229, 397-431; the special-relativistic 2D test reads
 * the diagonal cell as "bottom", AMR/RelativisticEulerAMR.cpp:422). */
int cosmos_amr_refine_flags(cosmos_amr* a, double cell_spacing, double tolerance);
/* One sweep of a level (0 coarse, 1 fine) updating only the cells whose flag
 * is set (compute{,X,Y}{FORCE,SLIC}TimeStep{,2D}AMR, AMR/EulerAMR.cpp:3-167). */
int cosmos_amr_sweep(cosmos_amr* a, int level, int order, int axis, double cell_spacing, double time_step,
                     double bias, int slope_limiter);
/* Coarse cells whose flag is clear take the mean of their children's conserved
 * vectors; fine cells whose flag is clear take a copy of their parent
 * (AMR/EulerAMR.cpp:282-308, 494-536).  Black-hole contexts place the cells
 * with the coarse spacing last passed to cosmos_amr_refine_flags,
 * cosmos_amr_derefine or a solve call (COSMOS_ERR_STATE before any). */
int cosmos_amr_restrict(cosmos_amr* a);
/* Derefinement: a refined coarse cell whose centre child (1D: child 2j+1; 2D:
 * (2i+1, 2j+1)) jumps by less than tolerance / 2 over cell_spacing to each fine
 * neighbour is made coarse again and restricted (AMR/EulerAMR.cpp:310-345, 538-586). */
int cosmos_amr_derefine(cosmos_amr* a, double cell_spacing, double tolerance);
/* EulerAMR::solveLevel1AMR / solve2DLevel1AMR and the RelativisticEulerAMR and
 * BlackHoleEulerAMR copies (AMR/EulerAMR.cpp:169-590,
 * AMR/BlackHoleEulerAMR.cpp:238-694): per coarse step the CFL step (damped by
 * the fine-step counter), the refinement flags, the masked coarse flux step,
 * masked fine sub-steps to the coarse time (less 1e-4; the 2D black-hole
 * driver sub-steps even with no fine flag set, :533), restriction and
 * derefinement; bias 0 and van Leer, as the reference fixes them; no
 * black-hole forcing (the reference's AMR drivers call none).
 * max_coarse_steps < 0 means no cap. */
int cosmos_amr_solve_level1(cosmos_amr* a, int order, double cell_spacing, double cfl, double final_time,
                            double tolerance, int max_coarse_steps, double* time_out, int* steps_out);
/* EulerAMR::solve{,2D}Level2AMR and the special-relativistic copies
 * (AMR/EulerAMR.cpp:593-1381, AMR/RelativisticEulerAMR.cpp:605-1394): three
 * levels of nx x ny, 2nx x 2ny and 4nx x 4ny cells (1D: nx, 2nx, 4nx) started
 * from the uploaded coarse state (finer levels copies of it, only the coarse
 * flags set; the 2D tests read the true bottom neighbour); per coarse step the
 * coarse flags and masked step, intermediate sub-steps (flags of flagged cells
 * at tolerance * 2 * order, masked step, fine sub-steps, restriction /
 * injection, fine derefinement) to the coarse time less 1e-4, then coarse
 * restriction / injection and intermediate derefinement at
 * tolerance * 0.5 * (1 / order); every test over the coarse spacing, every dt
 * damped by the reference's fine-step counter.  Black hole:
 * BlackHoleEulerAMR::solveLevel2AMR (AMR/BlackHoleEulerAMR.cpp:695-1055) and
 * solve2DLevel2AMR (:1057-1334), whose source is truncated at :1319-1325
 * inside the intermediate/fine exchange; past that point the 2D run follows
 * the 1D and special-relativistic drivers (strict exchange, fine
 * derefinement, coarse exchange, intermediate derefinement, natural cell
 * coordinates), a completion of this library's, not reference behaviour. */
int cosmos_amr_solve_level2(cosmos_amr* a, int order, double cell_spacing, double cfl, double final_time,
                            double tolerance, int max_coarse_steps, double* time_out, int* steps_out);
/* The three levels after cosmos_amr_solve_level2 (any pointer may be null),
 * row-major, flags as 0/1 bytes. */
int cosmos_amr_download_level2(cosmos_amr* a, double* coarse, double* intermediate, double* fine,
                               unsigned char* coarse_flags, unsigned char* intermediate_flags,
                               unsigned char* fine_flags);

/* ------------------------------------------------------------------------
 * Adjoint (reverse mode) of the Newtonian and black-hole updates (no
 * counterpart in the reference).  A cosmos_adjoint context owns a solver
 * handle for an nx-by-ny grid (dim 2) or an nx-cell line (dim 1).  It
 * differentiates the update the reference computes (SLIC or FORCE sweeps,
 * dimensionally split X(dt/2) Y(dt) X(dt/2) in 2D; for the black hole the
 * RK4 forcing steps on either side, Solvers/BlackHoleEulerSecondOrderSolver.cpp:181-218)
 * with respect to the primitive state, with the time steps held at the
 * values the forward pass chose (a frozen dt schedule), the limiters
 * differentiated on the branch the forward pass took (one-sided derivatives
 * at their kinks) and the con2prim Newton solves through their roots.
 * Host arrays are AoS primitives, as for cosmos_upload.
 * ------------------------------------------------------------------------ */
typedef struct cosmos_adjoint cosmos_adjoint;

int cosmos_adjoint_create(cosmos_adjoint** out, int physics, int dim, int nx, int ny, double adiabatic_index);
int cosmos_adjoint_destroy(cosmos_adjoint* a);
const char* cosmos_adjoint_last_error(const cosmos_adjoint* a);
/* The context's solver: upload the initial state through it. */
cosmos_solver* cosmos_adjoint_solver(cosmos_adjoint* a);
/* One sweep and its adjoint: the solver's state is advanced by the sweep
 * (as cosmos_sweep), and lambda_in = (d output / d input)^T lambda_out. */
int cosmos_adjoint_sweep(cosmos_adjoint* a, int order, int axis, double cell_spacing, double time_step, double bias,
                         int slope_limiter, const double* lambda_out, double* lambda_in);
/* Advance the solver n_steps steps of the solve / solve2D loop (dt from the CFL
 * condition with the 0.2x damping of the first six steps, no final time),
 * recording every dt and a device checkpoint of the state every
 * checkpoint_every steps.  The solver is set to COSMOS_MODE_EXACT: the
 * reverse pass recomputes segments with the exact kernels.  The last segment
 * runs through the context's segment buffers (and, for a black hole, its
 * tapes, cosmos_adjoint_set_tape), where the reverse pass uses it without
 * recomputing it. */
int cosmos_adjoint_forward(cosmos_adjoint* a, int order, double cell_spacing, double cfl, int n_steps, double bias,
                           int slope_limiter, int checkpoint_every);
/* grad_in = (d W_N / d W_0)^T grad_out for the recorded trajectory: each
 * segment between checkpoints is recomputed forward (except the one the
 * segment buffers still hold: after cosmos_adjoint_forward the last one),
 * then traversed in reverse one operation at a time, in buffers of the
 * context: the solver's state is not touched and stays at W_N. */
int cosmos_adjoint_backward(cosmos_adjoint* a, const double* grad_out, double* grad_in);
/* The recorded dt schedule applied to another initial state (for finite-
 * difference checks); the solver's state is left at the result. */
int cosmos_adjoint_replay(cosmos_adjoint* a, const double* prim_in, double* prim_out);
/* The recorded dt schedule: *count = steps recorded, out[0 .. min(capacity,
 * count)) = their time steps (out may be NULL to query the count).  No
 * counterpart in the reference (its solve loops keep dt local). */
int cosmos_adjoint_time_steps(cosmos_adjoint* a, double* out, int capacity, int* count);
/* Black hole (COSMOS_BLACKHOLE contexts; set the spacetime and excision on
 * the context's solver).  One RK4 forcing step of the solver's state
 * (bias 0 and van Leer, as the solve loop fixes them) and its adjoint, as
 * cosmos_adjoint_sweep does for a sweep. */
int cosmos_adjoint_forcing(cosmos_adjoint* a, double cell_spacing, double time_step, const double* lambda_out,
                           double* lambda_in);
/* Forcing sub-steps on each side of a step's sweeps (default 1; 0 to 8), as
 * the subcyclingIterations argument of the black-hole solve loops. */
int cosmos_adjoint_set_subcycles(cosmos_adjoint* a, int subcycles);
/* Black hole: how the reverse pass gets the values of each forcing step's RK4
 * chain.  COSMOS_TAPE_NONE: the adjoint kernel re-runs the chain in
 * registers; COSMOS_TAPE_RERUN: a forward kernel writes the chain to a device
 * tape (about 100 doubles per cell) just before the step's adjoint;
 * COSMOS_TAPE_SEGMENT (default): the segment recompute writes one tape per
 * forcing step of the segment, and one per sweep holding its interface
 * fluxes and SLIC face states (15 doubles per interface), and the reverse
 * reads them instead of re-running the steps; it falls back to
 * COSMOS_TAPE_RERUN when those tapes would leave less than reserve_bytes of
 * device memory free (default 8 GiB).  The gradients agree to rounding (the
 * adjoint kernels accumulate with atomic adds) across the modes. */
enum { COSMOS_TAPE_NONE = 0, COSMOS_TAPE_RERUN = 1, COSMOS_TAPE_SEGMENT = 2 };
int cosmos_adjoint_set_tape(cosmos_adjoint* a, int mode, long long reserve_bytes);
/* Free the segment buffers and tapes (the recorded trajectory stays): the
 * next cosmos_adjoint_backward allocates them again and recomputes every
 * segment. */
int cosmos_adjoint_release(cosmos_adjoint* a);

/* Batched point-wise limiter functions on host arrays (evaluated on the device):
 *   cosmos_slope_limiter: out[i] = SlopeLimiters::computeSlopeLimiter(steepness[i], bias, limiter)
 *                         (Solvers/SlopeLimiters.cpp:3-69)
 *   cosmos_slope_vector:  out[5i..5i+4] = SlopeLimiters::computeSlopeVector(L_i, M_i, R_i, bias, limiter)
 *                         on conserved vectors (Solvers/SlopeLimiters.cpp:71-101) */
int cosmos_slope_limiter(const double* steepness, long n, double bias, int slope_limiter, double* out);
int cosmos_slope_vector(const double* left, const double* middle, const double* right, long n, double bias,
                        int slope_limiter, double* out);

/* ---- Point-level functions, evaluated on the host ----------------------------
 * The public statics of the reference's solver classes that work on one state,
 * one interface or one cell, computed by the same physics code the kernels run
 * (compiled for the host), so their results are the kernels' bits without a
 * device round trip per call.  Primitive states are (rho, vx, vy, vz, p),
 * conserved and flux vectors the reference's five components.  For
 * COSMOS_BLACKHOLE, spacetime = (mass, spin, x, y, z) of
 * BlackHoleSpacetime(mass, spin, x, y, z) and position = the (x, y, z) the
 * reference method takes; both are ignored (may be NULL) otherwise.  Every
 * entry point returns COSMOS_OK or COSMOS_ERR_ARGUMENT (message:
 * cosmos_point_last_error, per thread). */
const char* cosmos_point_last_error(void);

/* Vector arithmetic on n components:
 *   lax_friedrichs  0.5((F_L + F_R) + (dx/dt)(U_L - U_R))  EulerFirstOrderSolver::computeLaxFriedrichsFlux
 *                                                         (Solvers/EulerFirstOrderSolver.cpp:3-11)
 *   richtmyer       0.5((U_L + U_R) + (dt/dx)(F_L - F_R))  ::computeRichtmyerFlux, the intermediate state (:13-19)
 *   force           0.5(F_LF + F_Ri)                       ::computeFORCEFlux (:21-24)
 *   force_update    U + (dt/dx)(F_L - F_R)                 ::computeFORCEUpdate (:96-101)
 *   evolution       (f (dt/dx))(F_L - F_R)                 EulerSolvers::computeFractionalEvolutionVector;
 *                   f = 0.5 is computeEvolutionVector     (Solvers/EulerSolvers.cpp:163-172)
 *   slope           SlopeLimiters::computeSlopeVector on conserved vectors (Solvers/SlopeLimiters.cpp:71-101) */
int cosmos_vector_lax_friedrichs(const double* left_cons, const double* right_cons, const double* left_flux,
                                 const double* right_flux, long n, double cell_spacing, double time_step,
                                 double* out);
int cosmos_vector_richtmyer(const double* left_cons, const double* right_cons, const double* left_flux,
                            const double* right_flux, long n, double cell_spacing, double time_step, double* out);
int cosmos_vector_force(const double* lax_friedrichs, const double* richtmyer, long n, double* out);
int cosmos_vector_force_update(const double* cons, const double* left_flux, const double* right_flux, long n,
                               double cell_spacing, double time_step, double* out);
int cosmos_vector_evolution(double step_fraction, const double* left_flux, const double* right_flux, long n,
                            double cell_spacing, double time_step, double* out);
int cosmos_vector_slope(const double* left, const double* middle, const double* right, long n, double bias,
                        int slope_limiter, double* out);
/* SlopeLimiters::computeSlopeLimiter (Solvers/SlopeLimiters.cpp:3-69). */
int cosmos_point_slope_limiter(double steepness, double bias, int slope_limiter, double* out);

/* State conversions: computeConservedVariableVector, setConservedVariableVector
 * (the primitives it stores) and the static compute{X,Y}FluxVector(cons)
 * (Euler/EulerStateVector.cpp:30-117, RelativisticEuler/RelativisticEulerStateVector.cpp:30-309,
 * BlackHoleEuler/BlackHoleEulerStateVector.cpp:30-357). */
int cosmos_point_conserved(int physics, const double* prim, double adiabatic_index, const double* spacetime,
                           const double* position, double* cons);
int cosmos_point_primitive(int physics, const double* cons, double adiabatic_index, const double* spacetime,
                           const double* position, double* prim);
int cosmos_point_flux(int physics, int axis, const double* cons, double adiabatic_index, const double* spacetime,
                      const double* position, double* flux);

#define COSMOS_FLUX_LAX_FRIEDRICHS 0
#define COSMOS_FLUX_RICHTMYER 1
#define COSMOS_FLUX_FORCE 2
/* compute{X,Y}{LaxFriedrichs,Richtmyer,FORCE}Flux(left, right, ...) of
 * EulerFirstOrderSolver (Solvers/EulerFirstOrderSolver.cpp:27-94),
 * RelativisticEulerFirstOrderSolver (Solvers/RelativisticEulerFirstOrderSolver.cpp) and
 * BlackHoleEulerFirstOrderSolver (Solvers/BlackHoleEulerFirstOrderSolver.cpp:3-99; position =
 * the interface, the states taken at position -/+ cell_spacing * 0.5 along the axis). */
int cosmos_point_interface_flux(int physics, int kind, int axis, const double* left, const double* right,
                                double cell_spacing, double time_step, double adiabatic_index,
                                const double* spacetime, const double* position, double* out);
/* compute{X,Y}SLICFlux(leftLeft, left, right, rightRight, ...) of the second-order solvers
 * (Solvers/EulerSecondOrderSolver.cpp:3-27, Solvers/RelativisticEulerSecondOrderSolver.cpp,
 * Solvers/BlackHoleEulerSecondOrderSolver.cpp:3-35; position = the interface). */
int cosmos_point_slic_flux(int physics, int axis, const double* left_left, const double* left, const double* right,
                           const double* right_right, double cell_spacing, double time_step, double bias,
                           int slope_limiter, double adiabatic_index, const double* spacetime,
                           const double* position, double* out);
/* evolveStateByHalf{X,Y}TimeStep(left, middle, right, ..., side) (Solvers/EulerSolvers.cpp:191-223,
 * Solvers/RelativisticEulerSolvers.cpp, Solvers/BlackHoleEulerSolvers.cpp:250-307; position = the
 * middle cell's centre): the primitives of the evolved face state (side 0: left / top,
 * 1: right / bottom). */
int cosmos_point_half_step(int physics, int axis, const double* left, const double* middle, const double* right,
                           double cell_spacing, double time_step, double bias, int slope_limiter, int side,
                           double adiabatic_index, const double* spacetime, const double* position,
                           double* prim_out);
/* BlackHoleEulerSolvers::evolveStateByFractional{X,Y}TimeStep(stepFraction, left, middle, right, ...)
 * (Solvers/BlackHoleEulerSolvers.cpp:171-217): COSMOS_BLACKHOLE only. */
int cosmos_point_fractional_step(int axis, double step_fraction, const double* left, const double* middle,
                                 const double* right, double cell_spacing, double time_step, double bias,
                                 int slope_limiter, double adiabatic_index, const double* spacetime,
                                 const double* position, double* prim_out);
/* SlopeLimiters::computeSlopeVector(left, middle, right state, ...) (Solvers/SlopeLimiters.cpp:103-121;
 * axis and cell_spacing unused) and computeSlopeVector{X,Y} (:123-147, COSMOS_BLACKHOLE: the states
 * taken at position -/+ cell_spacing along the axis). */
int cosmos_point_slope_vector(int physics, int axis, const double* left, const double* middle, const double* right,
                              double cell_spacing, double bias, int slope_limiter, double adiabatic_index,
                              const double* spacetime, const double* position, double* out);
/* BlackHoleEulerForcingSolver::evolveConservedVariableVector (dim 1; top and bottom unused) and
 * evolveConservedVariableVector2D (dim 2) on conserved vectors (Solvers/BlackHoleEulerForcingSolver.cpp:3-185):
 * the RK4 source-term update of the middle cell at position. */
int cosmos_point_forcing(int dim, const double* left, const double* middle, const double* right, const double* top,
                         const double* bottom, double cell_spacing, double time_step, double bias, int slope_limiter,
                         double adiabatic_index, const double* spacetime, const double* position,
                         double* cons_out);

/* Wait for all queued work on the handle's stream. */
/* One step of the solve loop on caller-owned device memory that allocates
 * nothing: no counterpart in the reference; for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops).  The step is the one
 * cosmos_solve takes (computeStableTimeStep{,2D}'s dt with the six-step
 * damping and the final-time clamp, then X(dt/2) Y(dt) X(dt/2) in 2D or one
 * sweep in 1D; exact arithmetic, Newtonian or special relativistic).
 * in / out: five SoA FP64 planes of primitives (rho, u, v, w, p), plane k at
 * base + k * pitch * ny, row i at + i * pitch (pitch >= nx doubles); distinct
 * buffers; in is not written.  time_step_out: a device double receiving dt
 * (may be NULL).  stream: a hipStream_t (NULL: the null stream) on the
 * current device.  Work is queued, not waited for.  Device working memory is
 * a fixed 32 MiB array per GPU held until the process ends: a 2D grid needs
 * 5 * pitch * ny <= cosmos_noalloc_grid_limit() doubles (about 838k cells at
 * pitch = nx), larger grids are rejected with COSMOS_ERR_ARGUMENT.  Calls
 * share that array: calls that overlap on one GPU race.  Errors:
 * cosmos_noalloc_last_error() (this thread). */
int cosmos_step_noalloc(int physics, int dim, int nx, int ny, long pitch, const double* in, double* out, int order,
                        double cell_spacing, double cfl, double current_time, double final_time,
                        int current_iteration, double bias, int slope_limiter, double adiabatic_index,
                        double* time_step_out, void* stream);
size_t cosmos_noalloc_grid_limit(void);
const char* cosmos_noalloc_last_error(void);

/* Numerical-fault status (no counterpart in the reference, which accepts
 * these silently: RelativisticEuler/RelativisticEulerStateVector.cpp:45-50,
 * 59-80, 93-100; Solvers/EulerSecondOrderSolver.cpp:109;
 * Solvers/BlackHoleEulerSolvers.cpp:88-140).  Counters accumulate on the
 * device from the solver's creation or last cosmos_status_reset:
 *   c2p_cap .. c2p_pressure_floor: special-relativistic con2prim solves of
 *     the per-solver sweeps (cosmos_sweep and the solve loops; not the batched
 *     or AMR kernels) that ran to the iteration cap, took the q floor, or had
 *     their density / pressure raised to 1e-8; each distinct solve of a sweep
 *     counted once (every half step, interface flux and cell update of a line
 *     once, though the reference evaluates some twice and the kernels
 *     recompute tile halos);
 *   nan_cells: cells of the current state holding a NaN, formed when read;
 *   nan_time_step: 1 once a solve loop took a NaN time step (after which the
 *     loop stops, as the reference's does);
 *   excised_grid: 1 once a black-hole wave speed found every cell excised
 *     (its maximum is then 0 and dt infinite, clamped to the final time). */
typedef struct {
    long long c2p_cap;
    long long c2p_q_floor;
    long long c2p_density_floor;
    long long c2p_pressure_floor;
    long long nan_cells;
    long long nan_time_step;
    long long excised_grid;
    long long reserved;
} cosmos_status_t;
/* Queue the NaN count and a copy of the status into out (host memory that
 * must stay valid until the stream reaches the copy: pinned memory makes the
 * copy asynchronous); cosmos_status does the same and waits. */
int cosmos_status_async(cosmos_solver* s, cosmos_status_t* out);
int cosmos_status(cosmos_solver* s, cosmos_status_t* out);
int cosmos_status_reset(cosmos_solver* s);

int cosmos_synchronize(cosmos_solver* s);
/* The handle's HIP stream (hipStream_t), for callers that time or order work. */
void* cosmos_stream(cosmos_solver* s);
/* Instrumentation (no counterpart in the reference): when enabled, every
 * kernel queued by cosmos_advance is bracketed by HIP events on the handle's
 * stream.  cosmos_profile_read synchronises and returns, for kinds
 * 0 step head (dt), 1 x sweep, 2 y sweep, 3 x sweep with the fused wave-speed
 * maximum, 4 fused whole step (cosmos_set_step_fusion), the summed
 * milliseconds and launch counts since the last read (arrays of
 * COSMOS_PROFILE_KINDS entries). */
#define COSMOS_PROFILE_KINDS 5
int cosmos_profile_enable(cosmos_solver* s, int on);
int cosmos_profile_read(cosmos_solver* s, double* total_ms, int* launches);

/* A plain device copy of `bytes` (a multiple of 16, 16-byte aligned) on a HIP
 * stream (NULL: the default stream), for bandwidth and counter calibration
 * (no counterpart in the reference). */
int cosmos_device_copy(void* dst, const void* src, long bytes, void* stream);

/* Device pointers of the current state planes (rho, u, v, w, p) and their
 * row pitch in elements; valid until the next sweep, step or destroy. */
int cosmos_state_planes(cosmos_solver* s, double** planes5, long* pitch);


/* ------------------------------------------------------------------------
 * Batched persistent form (no counterpart in the reference: the reference
 * solves one problem per call).  Many independent Newtonian and
 * special-relativistic problems, 1D and 2D, of uneven sizes, each advanced
 * by the solve / solve2D loop (dt from the CFL condition, the 0.2x damping
 * of iterations 0..5, the final-time clamp) until max_steps steps or
 * final_time, all in one launch of a captured hipGraph: set-up kernels, then
 * one persistent kernel per physics that takes tasks from a ticket queue.
 * Each problem's result is bitwise the one cosmos_solve gives for it (same
 * order, limiter and bias for the whole batch).  Host arrays are AoS
 * primitives, as for cosmos_upload.
 * ------------------------------------------------------------------------ */
typedef struct cosmos_batch cosmos_batch;

typedef struct {
    int physics;             /* COSMOS_NEWTONIAN or COSMOS_RELATIVISTIC */
    int dim;                 /* 1 or 2 */
    int nx, ny;              /* ny is ignored in 1D */
    double cell_spacing;     /* dx (both axes in 2D, as solve2D) */
    double cfl;
    double final_time;       /* INFINITY: run exactly max_steps steps */
    int max_steps;           /* 1 .. 1000000 */
    double adiabatic_index;
} cosmos_batch_problem;

/* How a physics' problems are cut into tasks (the autotuner's knobs):
 * problems of at most whole_max_cells cells run start to finish in one
 * block; larger ones run phase by phase as tasks of x_segments_per_task
 * x segments (256-thread row pieces, short rows packed side by side) and,
 * for the y sweep, of as many LDS tiles of y_tile_cols columns (8 or 16)
 * or, with y_tile_cols 0, of 256 column marches of y_rows rows.
 * blocks_per_cu: persistent blocks per compute unit (0: the kernel's
 * occupancy). */
typedef struct {
    int whole_max_cells;
    int x_segments_per_task;
    int y_rows;
    int blocks_per_cu;
    int y_tile_cols;
} cosmos_batch_schedule;

int cosmos_batch_default_schedule(int physics, cosmos_batch_schedule* out);
int cosmos_batch_create(cosmos_batch** out, const cosmos_batch_problem* problems, int n, int order, double bias,
                        int slope_limiter);
int cosmos_batch_destroy(cosmos_batch* b);
const char* cosmos_batch_last_error(const cosmos_batch* b);
/* Special-relativistic con2prim stopping tolerance and iteration cap for the
 * batch (default: the reference's 1e-8 and 1000). */
int cosmos_batch_set_con2prim(cosmos_batch* b, double tolerance, int max_iterations);
int cosmos_batch_set_schedule(cosmos_batch* b, int physics, const cosmos_batch_schedule* s);
int cosmos_batch_get_schedule(cosmos_batch* b, int physics, cosmos_batch_schedule* s);
int cosmos_batch_upload(cosmos_batch* b, int problem, const double* prim);
int cosmos_batch_download(cosmos_batch* b, int problem, double* prim);
/* Advance every problem from its current state (t = 0, iteration 0) to its
 * stop, then wait.  cosmos_batch_run_timed also returns the graph's
 * duration from HIP events around its launch. */
int cosmos_batch_run(cosmos_batch* b);
int cosmos_batch_run_timed(cosmos_batch* b, double* ms);
/* Per problem after a run (each array n entries, any may be NULL): the time
 * reached, the steps taken and the completion time in microseconds after
 * the run's first kernel started (device wall clock). */
int cosmos_batch_stats(cosmos_batch* b, double* time_reached, int* steps_taken, double* finish_us);
/* Coordinate descent over each physics' schedule (whole_max_cells,
 * x_segments_per_task, y_tile_cols, y_rows) on timed runs capped at
 * trial_steps steps;
 * keeps the fastest.  The problems' states are restored afterwards. */
int cosmos_batch_autotune(cosmos_batch* b, int trial_steps, double* best_ms);

#ifdef __cplusplus
}
#endif
#endif
