// Standalone stand-in for the host application's black-hole description:
// mass, dimensionless spin and position, with the constructors, getters and
// setters the drop-in solver classes use (the reference's is
// BlackHoleEuler/BlackHoleSpacetime.hpp; its default hole is mass 1, spin 0
// at (0.5, 0.5, 0)).  Inside the host application its own header takes
// precedence on the include path.
#ifndef COSMOS_COMPAT_BLACK_HOLE_SPACETIME_HPP
#define COSMOS_COMPAT_BLACK_HOLE_SPACETIME_HPP

class BlackHoleSpacetime {
public:
    BlackHoleSpacetime() : BlackHoleSpacetime(1.0, 0.0, 0.5, 0.5, 0.0) {}
    explicit BlackHoleSpacetime(double mass) : BlackHoleSpacetime(mass, 0.0, 0.5, 0.5, 0.0) {}
    BlackHoleSpacetime(double mass, double x, double y, double z) : BlackHoleSpacetime(mass, 0.0, x, y, z) {}
    BlackHoleSpacetime(double mass, double spin) : BlackHoleSpacetime(mass, spin, 0.5, 0.5, 0.0) {}
    BlackHoleSpacetime(double mass, double spin, double x, double y, double z)
        : blackHoleMass(mass), blackHoleSpin(spin), blackHoleXPosition(x), blackHoleYPosition(y),
          blackHoleZPosition(z) {}

    void setBlackHoleMass(double v) { blackHoleMass = v; }
    void setBlackHoleSpin(double v) { blackHoleSpin = v; }
    void setBlackHoleXPosition(double v) { blackHoleXPosition = v; }
    void setBlackHoleYPosition(double v) { blackHoleYPosition = v; }
    void setBlackHoleZPosition(double v) { blackHoleZPosition = v; }
    double getBlackHoleMass() { return blackHoleMass; }
    double getBlackHoleSpin() { return blackHoleSpin; }
    double getBlackHoleXPosition() { return blackHoleXPosition; }
    double getBlackHoleYPosition() { return blackHoleYPosition; }
    double getBlackHoleZPosition() { return blackHoleZPosition; }

private:
    double blackHoleMass, blackHoleSpin, blackHoleXPosition, blackHoleYPosition, blackHoleZPosition;
};

#endif
