// Standalone stand-in for the host application's material-parameter type,
// with the public interface the drop-in solver classes use
// (the reference's is Euler/EulerMaterialParameters.hpp:3-17).  When the
// drop-in is compiled inside the host application, the application's own
// header is found first on the include path and this one is not used.
#ifndef COSMOS_COMPAT_EULER_MATERIAL_PARAMETERS_HPP
#define COSMOS_COMPAT_EULER_MATERIAL_PARAMETERS_HPP

class EulerMaterialParameters {
public:
    EulerMaterialParameters() : adiabaticIndex(1.0) {}
    explicit EulerMaterialParameters(double index) : adiabaticIndex(index) {}
    void setAdiabaticIndex(double index) { adiabaticIndex = index; }
    double getAdiabaticIndex() { return adiabaticIndex; }

private:
    double adiabaticIndex;
};

#endif
