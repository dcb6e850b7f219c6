// Standalone stand-in for the host application's general-relativistic
// primitive state record (see EulerStateVector.hpp in this directory).
#ifndef COSMOS_COMPAT_BLACK_HOLE_EULER_STATE_VECTOR_HPP
#define COSMOS_COMPAT_BLACK_HOLE_EULER_STATE_VECTOR_HPP

#include "BlackHoleSpacetime.hpp"
#include "EulerStateVector.hpp"

class BlackHoleEulerStateVector : public cosmos_compat::PrimitiveState {
public:
    using cosmos_compat::PrimitiveState::PrimitiveState;
};

#endif
