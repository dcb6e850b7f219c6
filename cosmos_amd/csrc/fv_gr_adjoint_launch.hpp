// Launch interface of the general-relativistic adjoint kernels and the
// solver-side builders of the forward GR launches they mirror (internal to
// the library; not part of the C-ABI).
#pragma once

#include "../../include/cosmos_fv.h"
#include "fv_kernels.hpp"

namespace cfv {

struct GRAdjLaunch {
    GRLaunch L;       // the forward operation: L.in its input state, L.flux its interface fluxes (sweeps)
    ConstPlanes lam;  // adjoint of the operation's output
    Planes lam_in;    // adjoint of its input (forcing: accumulated, so zero it first)
    Planes gbuf;      // sweeps: per-cell adjoint of the updated conserved vector (state layout)
    double* tape;     // forcing: kTapePlanes planes of tape_plane doubles (state layout), or null
    long tape_plane;
    double* carry;    // forcing with a tape: kCarryPlanes planes of tape_plane doubles
};

// Forcing tape, one plane per value: the final conserved vector un, the four
// neighbour states Wn (x-, x+, y-, y+; 1D uses the first two), then per RK
// stage the stage's conserved state tmp, its primitives Wm and the fractional
// sweep outputs e2 (x), e3 (y), e1 (last x; the only one in 1D).
constexpr int kTapeUn = 0, kTapeWn = 5, kTapeStage = 25, kTapePerStage = 25;
constexpr int kTapeTmp = 0, kTapeWm = 5, kTapeE2 = 10, kTapeE3 = 15, kTapeE1 = 20;
constexpr int kTapePlanes = kTapeStage + 3 * kTapePerStage;

// Running adjoints of the taped forcing reverse: of the output's conserved
// vector (gun), of the current stage increment (gk), of the step-start
// conserved state (guM, accumulated), of the fractional sweep output being
// walked back (ge), and of the four neighbour states (gWn, accumulated).
constexpr int kCarryGun = 0, kCarryGk = 5, kCarryGuM = 10, kCarryGe = 15, kCarryGWn = 20;
constexpr int kCarryPlanes = 40;

// Sweep: the flux planes must hold the forward sweep's interface fluxes.
hipError_t launch_gr_adjoint_sweep(const GRAdjLaunch& A, int axis, hipStream_t s);
// Forcing: with A.tape set, the tape must hold the operation's chain
// (launch_gr_forcing_tape on the same input); otherwise the kernel re-runs it.
hipError_t launch_gr_adjoint_forcing(const GRAdjLaunch& A, hipStream_t s);
// The forward forcing (output bitwise that of launch_gr_forcing) writing the tape.
hipError_t launch_gr_forcing_tape(const GRAdjLaunch& A, hipStream_t s);

// The GR launch that cosmos_sweep / cosmos_forcing of a black-hole solver
// would issue (coefficients, spacetime, physics parameters, flux planes),
// with its input and output planes left to the caller.
int gr_sweep_launch(cosmos_solver* s, int order, double h, double dt, double bias, int limiter, GRLaunch* L);
int gr_forcing_launch(cosmos_solver* s, double h, double dt, double bias, int limiter, GRLaunch* L);

}  // namespace cfv
