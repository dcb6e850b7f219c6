// Drop-in replacements for the reference's solver classes on the hot path,
// with the same class names, static method names, argument lists and
// by-value / in-place semantics, implemented over the C-ABI in
// include/cosmos_fv.h (so every sweep runs as gfx950 HIP kernels).
//
//   reference class (file)                                  here
//   EulerSolvers (Solvers/EulerSolvers.hpp:10-39)           wave speed, stable time step
//   EulerFirstOrderSolver (Solvers/EulerFirstOrderSolver.hpp:8-47)
//   EulerSecondOrderSolver (Solvers/EulerSecondOrderSolver.hpp:6-29)
//   RelativisticEulerSolvers / ...FirstOrderSolver / ...SecondOrderSolver
//   BlackHoleEulerSolvers / ...FirstOrderSolver / ...SecondOrderSolver / ...ForcingSolver
//     (Solvers/BlackHoleEuler{Solvers,FirstOrderSolver,SecondOrderSolver,ForcingSolver}.hpp)
//   SlopeLimiters (Solvers/SlopeLimiters.hpp:10-35)         limiter functions and slope vectors
//   EulerAMR, RelativisticEulerAMR (AMR/EulerAMR.hpp:8-44,
//     AMR/RelativisticEulerAMR.hpp)                         masked sweeps, level-1 and level-2 drivers
//   BlackHoleEulerAMR (AMR/BlackHoleEulerAMR.hpp:8-53)      masked sweeps, level-1 and level-2 drivers
//                                                           (solve2DLevel2AMR: the reference's definition
//                                                           is truncated; completed as the 1D one goes on)
//
// The state types come from the host application's headers
// (EulerStateVector.hpp, RelativisticEulerStateVector.hpp,
// BlackHoleEulerStateVector.hpp, BlackHoleSpacetime.hpp,
// EulerMaterialParameters.hpp) when the drop-in is built inside it; the
// compat/ directory supplies minimal stand-ins for standalone builds.
//
// Differences, all deliberate: device or argument failures throw
// std::runtime_error (the reference has no error path); solve/solve2D run
// the whole loop on the device and do not print the per-iteration status
// line (EulerSolvers::outputStatus is still provided).
#ifndef COSMOS_DROPIN_HPP
#define COSMOS_DROPIN_HPP

#include <tuple>
#include <vector>

#include "BlackHoleEulerStateVector.hpp"
#include "BlackHoleSpacetime.hpp"
#include "EulerMaterialParameters.hpp"
#include "EulerStateVector.hpp"
#include "RelativisticEulerStateVector.hpp"

using std::tuple;
using std::vector;

class SlopeLimiters {
public:
    static double computeGradientRatio(double steepness, double bias);
    static double computeSuperBeeSlopeLimiter(double steepness, double bias);
    static double computeVanLeerSlopeLimiter(double steepness, double bias);
    static double computeMinBeeSlopeLimiter(double steepness, double bias);
    static double computeSlopeLimiter(double steepness, double bias, int slopeLimiter);
    static vector<double> computeSlopeVector(vector<double> leftConservedVariableVector,
                                             vector<double> middleConservedVariableVector,
                                             vector<double> rightConservedVariableVector, double bias, int slopeLimiter);
};

class EulerSolvers {
public:
    static double computeMaximumWaveSpeed(vector<EulerStateVector>& currentCells, EulerMaterialParameters materialParameters);
    static double computeMaximumWaveSpeed2D(vector<vector<EulerStateVector> >& currentCells,
                                            EulerMaterialParameters materialParameters);
    static double computeStableTimeStep(double timeStep, double currentTime, double finalTime, int currentIteration);
    static double computeStableTimeStep(vector<EulerStateVector> currentCells, double cellSpacing, double CFLCoefficient,
                                        double currentTime, double finalTime, int currentIteration,
                                        EulerMaterialParameters materialParameters);
    static double computeStableTimeStep2D(vector<vector<EulerStateVector> > currentCells, double cellSpacing,
                                          double CFLCoefficient, double currentTime, double finalTime,
                                          int currentIteration, EulerMaterialParameters materialParameters);
    static void outputStatus(int currentIteration, double currentTime, double timeStep);
};

class EulerFirstOrderSolver {
public:
    static void computeFORCETimeStep(vector<EulerStateVector>& currentCells, double cellSpacing, double timeStep,
                                     EulerMaterialParameters materialParameters);
    static void computeXFORCETimeStep2D(vector<vector<EulerStateVector> >& currentCells, double cellSpacing,
                                        double timeStep, EulerMaterialParameters materialParameters);
    static void computeYFORCETimeStep2D(vector<vector<EulerStateVector> >& currentCells, double cellSpacing,
                                        double timeStep, EulerMaterialParameters materialParameters);
    static vector<EulerStateVector> solve(vector<EulerStateVector> initialCells, double cellSpacing,
                                          double CFLCoefficient, double finalTime,
                                          EulerMaterialParameters materialParameters);
    static vector<vector<EulerStateVector> > solve2D(vector<vector<EulerStateVector> > initialCells,
                                                     double cellSpacing, double CFLCoefficient, double finalTime,
                                                     EulerMaterialParameters materialParameters);
};

class EulerSecondOrderSolver {
public:
    static void computeSLICTimeStep(vector<EulerStateVector>& currentCells, double cellSpacing, double timeStep,
                                    double bias, int slopeLimiter, EulerMaterialParameters materialParameters);
    static void computeXSLICTimeStep2D(vector<vector<EulerStateVector> >& currentCells, double cellSpacing,
                                       double timeStep, double bias, int slopeLimiter,
                                       EulerMaterialParameters materialParameters);
    static void computeYSLICTimeStep2D(vector<vector<EulerStateVector> >& currentCells, double cellSpacing,
                                       double timeStep, double bias, int slopeLimiter,
                                       EulerMaterialParameters materialParameters);
    static vector<EulerStateVector> solve(vector<EulerStateVector> initialCells, double cellSpacing,
                                          double CFLCoefficient, double finalTime, double bias, int slopeLimiter,
                                          EulerMaterialParameters materialParameters);
    static vector<vector<EulerStateVector> > solve2D(vector<vector<EulerStateVector> > initialCells,
                                                     double cellSpacing, double CFLCoefficient, double finalTime,
                                                     double bias, int slopeLimiter,
                                                     EulerMaterialParameters materialParameters);
};

class RelativisticEulerSolvers {
public:
    static double computeMaximumWaveSpeed(vector<RelativisticEulerStateVector>& currentCells,
                                          EulerMaterialParameters materialParameters);
    static double computeMaximumWaveSpeed2D(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                            EulerMaterialParameters materialParameters);
    static double computeStableTimeStep(vector<RelativisticEulerStateVector> currentCells, double cellSpacing,
                                        double CFLCoefficient, double currentTime, double finalTime,
                                        int currentIteration, EulerMaterialParameters materialParameters);
    static double computeStableTimeStep2D(vector<vector<RelativisticEulerStateVector> > currentCells,
                                          double cellSpacing, double CFLCoefficient, double currentTime,
                                          double finalTime, int currentIteration,
                                          EulerMaterialParameters materialParameters);
};

class RelativisticEulerFirstOrderSolver {
public:
    static void computeFORCETimeStep(vector<RelativisticEulerStateVector>& currentCells, double cellSpacing,
                                     double timeStep, EulerMaterialParameters materialParameters);
    static void computeXFORCETimeStep2D(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                        double cellSpacing, double timeStep, EulerMaterialParameters materialParameters);
    static void computeYFORCETimeStep2D(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                        double cellSpacing, double timeStep, EulerMaterialParameters materialParameters);
    static vector<RelativisticEulerStateVector> solve(vector<RelativisticEulerStateVector> initialCells,
                                                      double cellSpacing, double CFLCoefficient, double finalTime,
                                                      EulerMaterialParameters materialParameters);
    static vector<vector<RelativisticEulerStateVector> > solve2D(
        vector<vector<RelativisticEulerStateVector> > initialCells, double cellSpacing, double CFLCoefficient,
        double finalTime, EulerMaterialParameters materialParameters);
};

class RelativisticEulerSecondOrderSolver {
public:
    static void computeSLICTimeStep(vector<RelativisticEulerStateVector>& currentCells, double cellSpacing,
                                    double timeStep, double bias, int slopeLimiter,
                                    EulerMaterialParameters materialParameters);
    static void computeXSLICTimeStep2D(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                       double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                       EulerMaterialParameters materialParameters);
    static void computeYSLICTimeStep2D(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                       double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                       EulerMaterialParameters materialParameters);
    static vector<RelativisticEulerStateVector> solve(vector<RelativisticEulerStateVector> initialCells,
                                                      double cellSpacing, double CFLCoefficient, double finalTime,
                                                      double bias, int slopeLimiter,
                                                      EulerMaterialParameters materialParameters);
    static vector<vector<RelativisticEulerStateVector> > solve2D(
        vector<vector<RelativisticEulerStateVector> > initialCells, double cellSpacing, double CFLCoefficient,
        double finalTime, double bias, int slopeLimiter, EulerMaterialParameters materialParameters);
};

class BlackHoleEulerSolvers {
public:
    static double computeMaximumWaveSpeed(vector<BlackHoleEulerStateVector>& currentCells,
                                          EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static double computeMaximumWaveSpeed2D(vector<vector<BlackHoleEulerStateVector> >& currentCells,
                                            EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static double computeStableTimeStep(vector<BlackHoleEulerStateVector> currentCells, double cellSpacing,
                                        double CFLCoefficient, double currentTime, double finalTime,
                                        int currentIteration, EulerMaterialParameters materialParameters,
                                        BlackHoleSpacetime blackHole);
    static double computeStableTimeStep2D(vector<vector<BlackHoleEulerStateVector> > currentCells, double cellSpacing,
                                          double CFLCoefficient, double currentTime, double finalTime,
                                          int currentIteration, EulerMaterialParameters materialParameters,
                                          BlackHoleSpacetime blackHole);
};

class BlackHoleEulerForcingSolver {
public:
    static void computeRungeKuttaTimeStep(vector<BlackHoleEulerStateVector>& currentCells, double cellSpacing,
                                          double timeStep, double bias, int slopeLimiter,
                                          EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeRungeKuttaTimeStep2D(vector<vector<BlackHoleEulerStateVector> >& currentCells,
                                            double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                            EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
};

class BlackHoleEulerFirstOrderSolver {
public:
    static void computeFORCETimeStep(vector<BlackHoleEulerStateVector>& currentCells, double cellSpacing,
                                     double timeStep, EulerMaterialParameters materialParameters,
                                     BlackHoleSpacetime blackHole);
    static void computeXFORCETimeStep2D(vector<vector<BlackHoleEulerStateVector> >& currentCells, double cellSpacing,
                                        double timeStep, EulerMaterialParameters materialParameters,
                                        BlackHoleSpacetime blackHole);
    static void computeYFORCETimeStep2D(vector<vector<BlackHoleEulerStateVector> >& currentCells, double cellSpacing,
                                        double timeStep, EulerMaterialParameters materialParameters,
                                        BlackHoleSpacetime blackHole);
    static vector<BlackHoleEulerStateVector> solve(vector<BlackHoleEulerStateVector> initialCells, double cellSpacing,
                                                   double CFLCoefficient, double finalTime, int subcyclingIterations,
                                                   EulerMaterialParameters materialParameters,
                                                   BlackHoleSpacetime blackHole);
    static vector<vector<BlackHoleEulerStateVector> > solve2D(vector<vector<BlackHoleEulerStateVector> > initialCells,
                                                              double cellSpacing, double CFLCoefficient,
                                                              double finalTime, int subcyclingIterations,
                                                              EulerMaterialParameters materialParameters,
                                                              BlackHoleSpacetime blackHole);
};

class BlackHoleEulerSecondOrderSolver {
public:
    static void computeSLICTimeStep(vector<BlackHoleEulerStateVector>& currentCells, double cellSpacing,
                                    double timeStep, double bias, int slopeLimiter,
                                    EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeXSLICTimeStep2D(vector<vector<BlackHoleEulerStateVector> >& currentCells, double cellSpacing,
                                       double timeStep, double bias, int slopeLimiter,
                                       EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeYSLICTimeStep2D(vector<vector<BlackHoleEulerStateVector> >& currentCells, double cellSpacing,
                                       double timeStep, double bias, int slopeLimiter,
                                       EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static vector<BlackHoleEulerStateVector> solve(vector<BlackHoleEulerStateVector> initialCells, double cellSpacing,
                                                   double CFLCoefficient, double finalTime, double bias,
                                                   int slopeLimiter, int subcyclingIterations,
                                                   EulerMaterialParameters materialParameters,
                                                   BlackHoleSpacetime blackHole);
    static vector<vector<BlackHoleEulerStateVector> > solve2D(vector<vector<BlackHoleEulerStateVector> > initialCells,
                                                              double cellSpacing, double CFLCoefficient,
                                                              double finalTime, double bias, int slopeLimiter,
                                                              int subcyclingIterations,
                                                              EulerMaterialParameters materialParameters,
                                                              BlackHoleSpacetime blackHole);
};

// Two- and three-level AMR (the drivers run on the device; the status lines
// of AMRHelper are not printed).  Masked sweeps update only the cells whose
// AMRStructure entry is true.
class EulerAMR {
public:
    static void computeFORCETimeStepAMR(vector<EulerStateVector>& currentCells, double cellSpacing, double timeStep,
                                        vector<bool> AMRStructure, EulerMaterialParameters materialParameters);
    static void computeXFORCETimeStep2DAMR(vector<vector<EulerStateVector> >& currentCells, double cellSpacing,
                                           double timeStep, vector<vector<bool> > AMRStructure,
                                           EulerMaterialParameters materialParameters);
    static void computeYFORCETimeStep2DAMR(vector<vector<EulerStateVector> >& currentCells, double cellSpacing,
                                           double timeStep, vector<vector<bool> > AMRStructure,
                                           EulerMaterialParameters materialParameters);
    static void computeSLICTimeStepAMR(vector<EulerStateVector>& currentCells, double cellSpacing, double timeStep,
                                       double bias, int slopeLimiter, vector<bool> AMRStructure,
                                       EulerMaterialParameters materialParameters);
    static void computeXSLICTimeStep2DAMR(vector<vector<EulerStateVector> >& currentCells, double cellSpacing,
                                          double timeStep, double bias, int slopeLimiter,
                                          vector<vector<bool> > AMRStructure, EulerMaterialParameters materialParameters);
    static void computeYSLICTimeStep2DAMR(vector<vector<EulerStateVector> >& currentCells, double cellSpacing,
                                          double timeStep, double bias, int slopeLimiter,
                                          vector<vector<bool> > AMRStructure, EulerMaterialParameters materialParameters);
    static tuple<vector<EulerStateVector>, vector<EulerStateVector>, vector<bool>, vector<bool> > solveLevel1AMR(
        vector<EulerStateVector> initialCells, double cellSpacing, double CFLCoefficient, double finalTime,
        double AMRTolerance, int order, EulerMaterialParameters materialParameters);
    static tuple<vector<vector<EulerStateVector> >, vector<vector<EulerStateVector> >, vector<vector<bool> >,
                 vector<vector<bool> > >
    solve2DLevel1AMR(vector<vector<EulerStateVector> > initialCells, double cellSpacing, double CFLCoefficient,
                     double finalTime, double AMRTolerance, int order, EulerMaterialParameters materialParameters);
    static tuple<vector<EulerStateVector>, vector<EulerStateVector>, vector<EulerStateVector>, vector<bool>,
                 vector<bool>, vector<bool> >
    solveLevel2AMR(vector<EulerStateVector> initialCells, double cellSpacing, double CFLCoefficient, double finalTime,
                   double AMRTolerance, int order, EulerMaterialParameters materialParameters);
    static tuple<vector<vector<EulerStateVector> >, vector<vector<EulerStateVector> >, vector<vector<EulerStateVector> >,
                 vector<vector<bool> >, vector<vector<bool> >, vector<vector<bool> > >
    solve2DLevel2AMR(vector<vector<EulerStateVector> > initialCells, double cellSpacing, double CFLCoefficient,
                     double finalTime, double AMRTolerance, int order, EulerMaterialParameters materialParameters);
};

class RelativisticEulerAMR {
public:
    static void computeFORCETimeStepAMR(vector<RelativisticEulerStateVector>& currentCells, double cellSpacing,
                                        double timeStep, vector<bool> AMRStructure,
                                        EulerMaterialParameters materialParameters);
    static void computeXFORCETimeStep2DAMR(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                           double cellSpacing, double timeStep, vector<vector<bool> > AMRStructure,
                                           EulerMaterialParameters materialParameters);
    static void computeYFORCETimeStep2DAMR(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                           double cellSpacing, double timeStep, vector<vector<bool> > AMRStructure,
                                           EulerMaterialParameters materialParameters);
    static void computeSLICTimeStepAMR(vector<RelativisticEulerStateVector>& currentCells, double cellSpacing,
                                       double timeStep, double bias, int slopeLimiter, vector<bool> AMRStructure,
                                       EulerMaterialParameters materialParameters);
    static void computeXSLICTimeStep2DAMR(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                          double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                          vector<vector<bool> > AMRStructure, EulerMaterialParameters materialParameters);
    static void computeYSLICTimeStep2DAMR(vector<vector<RelativisticEulerStateVector> >& currentCells,
                                          double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                          vector<vector<bool> > AMRStructure, EulerMaterialParameters materialParameters);
    static tuple<vector<RelativisticEulerStateVector>, vector<RelativisticEulerStateVector>, vector<bool>, vector<bool> >
    solveLevel1AMR(vector<RelativisticEulerStateVector> initialCells, double cellSpacing, double CFLCoefficient,
                   double finalTime, double AMRTolerance, int order, EulerMaterialParameters materialParameters);
    static tuple<vector<vector<RelativisticEulerStateVector> >, vector<vector<RelativisticEulerStateVector> >,
                 vector<vector<bool> >, vector<vector<bool> > >
    solve2DLevel1AMR(vector<vector<RelativisticEulerStateVector> > initialCells, double cellSpacing,
                     double CFLCoefficient, double finalTime, double AMRTolerance, int order,
                     EulerMaterialParameters materialParameters);
    static tuple<vector<RelativisticEulerStateVector>, vector<RelativisticEulerStateVector>,
                 vector<RelativisticEulerStateVector>, vector<bool>, vector<bool>, vector<bool> >
    solveLevel2AMR(vector<RelativisticEulerStateVector> initialCells, double cellSpacing, double CFLCoefficient,
                   double finalTime, double AMRTolerance, int order, EulerMaterialParameters materialParameters);
    static tuple<vector<vector<RelativisticEulerStateVector> >, vector<vector<RelativisticEulerStateVector> >,
                 vector<vector<RelativisticEulerStateVector> >, vector<vector<bool> >, vector<vector<bool> >,
                 vector<vector<bool> > >
    solve2DLevel2AMR(vector<vector<RelativisticEulerStateVector> > initialCells, double cellSpacing,
                     double CFLCoefficient, double finalTime, double AMRTolerance, int order,
                     EulerMaterialParameters materialParameters);
};

class BlackHoleEulerAMR {
public:
    static void computeFORCETimeStepAMR(vector<BlackHoleEulerStateVector>& currentCells, double cellSpacing,
                                        double timeStep, vector<bool> AMRStructure,
                                        EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeXFORCETimeStep2DAMR(vector<vector<BlackHoleEulerStateVector> >& currentCells,
                                           double cellSpacing, double timeStep, vector<vector<bool> > AMRStructure,
                                           EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeYFORCETimeStep2DAMR(vector<vector<BlackHoleEulerStateVector> >& currentCells,
                                           double cellSpacing, double timeStep, vector<vector<bool> > AMRStructure,
                                           EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeSLICTimeStepAMR(vector<BlackHoleEulerStateVector>& currentCells, double cellSpacing,
                                       double timeStep, double bias, int slopeLimiter, vector<bool> AMRStructure,
                                       EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeXSLICTimeStep2DAMR(vector<vector<BlackHoleEulerStateVector> >& currentCells,
                                          double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                          vector<vector<bool> > AMRStructure,
                                          EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeYSLICTimeStep2DAMR(vector<vector<BlackHoleEulerStateVector> >& currentCells,
                                          double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                          vector<vector<bool> > AMRStructure,
                                          EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static tuple<vector<BlackHoleEulerStateVector>, vector<BlackHoleEulerStateVector>, vector<bool>, vector<bool> >
    solveLevel1AMR(vector<BlackHoleEulerStateVector> initialCells, double cellSpacing, double CFLCoefficient,
                   double finalTime, double AMRTolerance, int order, EulerMaterialParameters materialParameters,
                   BlackHoleSpacetime blackHole);
    static tuple<vector<vector<BlackHoleEulerStateVector> >, vector<vector<BlackHoleEulerStateVector> >,
                 vector<vector<bool> >, vector<vector<bool> > >
    solve2DLevel1AMR(vector<vector<BlackHoleEulerStateVector> > initialCells, double cellSpacing,
                     double CFLCoefficient, double finalTime, double AMRTolerance, int order,
                     EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static tuple<vector<BlackHoleEulerStateVector>, vector<BlackHoleEulerStateVector>, vector<BlackHoleEulerStateVector>,
                 vector<bool>, vector<bool>, vector<bool> >
    solveLevel2AMR(vector<BlackHoleEulerStateVector> initialCells, double cellSpacing, double CFLCoefficient,
                   double finalTime, double AMRTolerance, int order, EulerMaterialParameters materialParameters,
                   BlackHoleSpacetime blackHole);
    static tuple<vector<vector<BlackHoleEulerStateVector> >, vector<vector<BlackHoleEulerStateVector> >,
                 vector<vector<BlackHoleEulerStateVector> >, vector<vector<bool> >, vector<vector<bool> >,
                 vector<vector<bool> > >
    solve2DLevel2AMR(vector<vector<BlackHoleEulerStateVector> > initialCells, double cellSpacing,
                     double CFLCoefficient, double finalTime, double AMRTolerance, int order,
                     EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
};

#endif
