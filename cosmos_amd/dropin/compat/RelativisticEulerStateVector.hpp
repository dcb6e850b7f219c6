// Standalone stand-in for the host application's special-relativistic
// primitive state record (see EulerStateVector.hpp in this directory).
#ifndef COSMOS_COMPAT_RELATIVISTIC_EULER_STATE_VECTOR_HPP
#define COSMOS_COMPAT_RELATIVISTIC_EULER_STATE_VECTOR_HPP

#include "EulerStateVector.hpp"

class RelativisticEulerStateVector : public cosmos_compat::PrimitiveState {
public:
    using cosmos_compat::PrimitiveState::PrimitiveState;
};

#endif
