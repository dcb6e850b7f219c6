// Standalone stand-in for the host application's primitive state records,
// exposing the constructor, getters and setters the drop-in solver classes
// use (the reference's are Euler/EulerStateVector.hpp and
// RelativisticEuler/RelativisticEulerStateVector.hpp).  Inside the host
// application its own headers take precedence on the include path.
#ifndef COSMOS_COMPAT_EULER_STATE_VECTOR_HPP
#define COSMOS_COMPAT_EULER_STATE_VECTOR_HPP

#include "EulerMaterialParameters.hpp"

namespace cosmos_compat {

// (rho, u, v, w, p); the default state is (1, 0, 0, 0, 1).
class PrimitiveState {
public:
    PrimitiveState() : density(1.0), xVelocity(0.0), yVelocity(0.0), zVelocity(0.0), pressure(1.0) {}
    PrimitiveState(double rho, double u, double v, double w, double p)
        : density(rho), xVelocity(u), yVelocity(v), zVelocity(w), pressure(p) {}

    void setDensity(double x) { density = x; }
    void setXVelocity(double x) { xVelocity = x; }
    void setYVelocity(double x) { yVelocity = x; }
    void setZVelocity(double x) { zVelocity = x; }
    void setPressure(double x) { pressure = x; }
    double getDensity() { return density; }
    double getXVelocity() { return xVelocity; }
    double getYVelocity() { return yVelocity; }
    double getZVelocity() { return zVelocity; }
    double getPressure() { return pressure; }

private:
    double density, xVelocity, yVelocity, zVelocity, pressure;
};

}  // namespace cosmos_compat

class EulerStateVector : public cosmos_compat::PrimitiveState {
public:
    using cosmos_compat::PrimitiveState::PrimitiveState;
};

#endif
