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
};

// Sweep: the flux planes must hold the forward sweep's interface fluxes.
hipError_t launch_gr_adjoint_sweep(const GRAdjLaunch& A, int axis, hipStream_t s);
hipError_t launch_gr_adjoint_forcing(const GRAdjLaunch& A, hipStream_t s);

// The GR launch that cosmos_sweep / cosmos_forcing of a black-hole solver
// would issue (coefficients, spacetime, physics parameters, flux planes),
// with its input and output planes left to the caller.
int gr_sweep_launch(cosmos_solver* s, int order, double h, double dt, double bias, int limiter, GRLaunch* L);
int gr_forcing_launch(cosmos_solver* s, double h, double dt, double bias, int limiter, GRLaunch* L);

}  // namespace cfv
