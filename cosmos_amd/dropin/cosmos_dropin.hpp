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
// Grid-level methods run on the device; the point-level statics (one state,
// interface or cell: fluxes, half and fractional steps, slope vectors, the
// black-hole forcing of one cell, the vector arithmetic) are evaluated on the
// host by the library's cosmos_point_* / cosmos_vector_* entry points, which
// compile the kernels' own physics code for the host (same bits, no device
// round trip per call).
//
// Differences, all deliberate: device or argument failures throw
// std::runtime_error (the reference has no error path); solve/solve2D run
// the whole loop on the device and do not print the per-iteration status
// line (EulerSolvers::outputStatus is still provided); vector arguments of
// the point-level statics must have the reference's five components.
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
    static vector<double> computeSlopeVector(EulerStateVector leftStateVector, EulerStateVector middleStateVector,
                                             EulerStateVector rightStateVector, double bias, int slopeLimiter,
                                             EulerMaterialParameters materialParameters);
    static vector<double> computeSlopeVector(RelativisticEulerStateVector leftStateVector,
                                             RelativisticEulerStateVector middleStateVector,
                                             RelativisticEulerStateVector rightStateVector, double bias,
                                             int slopeLimiter, EulerMaterialParameters materialParameters);
    static vector<double> computeSlopeVectorX(BlackHoleEulerStateVector leftStateVector,
                                              BlackHoleEulerStateVector middleStateVector,
                                              BlackHoleEulerStateVector rightStateVector, double cellSpacing,
                                              double bias, int slopeLimiter, double xCoordinate, double yCoordinate,
                                              double zCoordinate, EulerMaterialParameters materialParameters,
                                              BlackHoleSpacetime blackHole);
    static vector<double> computeSlopeVectorY(BlackHoleEulerStateVector topStateVector,
                                              BlackHoleEulerStateVector middleStateVector,
                                              BlackHoleEulerStateVector bottomStateVector, double cellSpacing,
                                              double bias, int slopeLimiter, double xCoordinate, double yCoordinate,
                                              double zCoordinate, EulerMaterialParameters materialParameters,
                                              BlackHoleSpacetime blackHole);
};

class EulerSolvers {
public:
    static vector<EulerStateVector> insertBoundaryCells(vector<EulerStateVector>& currentCells, int boundarySize);
    static vector<vector<EulerStateVector> > insertBoundaryCells2D(vector<vector<EulerStateVector> >& currentCells,
                                                                   int boundarySize);
    static vector<double> computeFractionalEvolutionVector(double stepFraction, vector<double> leftFluxVector,
                                                           vector<double> rightFluxVector, double cellSpacing,
                                                           double timeStep);
    static vector<double> computeEvolutionVector(vector<double> leftFluxVector, vector<double> rightFluxVector,
                                                 double cellSpacing, double timeStep);
    static EulerStateVector evolveStateByHalfTimeStep(vector<double> leftExtrapolatedValue,
                                                      vector<double> rightExtrapolatedValue,
                                                      vector<double> evolutionVector, int side,
                                                      EulerMaterialParameters materialParameters);
    static EulerStateVector evolveStateByHalfXTimeStep(EulerStateVector leftStateVector,
                                                       EulerStateVector middleStateVector,
                                                       EulerStateVector rightStateVector, double cellSpacing,
                                                       double timeStep, double bias, int slopeLimiter, int side,
                                                       EulerMaterialParameters materialParameters);
    static EulerStateVector evolveStateByHalfYTimeStep(EulerStateVector topStateVector,
                                                       EulerStateVector middleStateVector,
                                                       EulerStateVector bottomStateVector, double cellSpacing,
                                                       double timeStep, double bias, int slopeLimiter, int side,
                                                       EulerMaterialParameters materialParameters);
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
    static vector<double> computeLaxFriedrichsFlux(vector<double> leftConservedVariableVector,
                                                   vector<double> rightConservedVariableVector,
                                                   vector<double> leftFluxVector, vector<double> rightFluxVector,
                                                   double cellSpacing, double timeStep);
    static vector<double> computeRichtmyerFlux(vector<double> leftConservedVariableVector,
                                               vector<double> rightConservedVariableVector,
                                               vector<double> leftFluxVector, vector<double> rightFluxVector,
                                               double cellSpacing, double timeStep);
    static vector<double> computeFORCEFlux(vector<double> laxFriedrichsFlux, vector<double> richtmyerFlux);
    static vector<double> computeXLaxFriedrichsFlux(EulerStateVector leftStateVector, EulerStateVector rightStateVector,
                                                    double cellSpacing, double timeStep,
                                                    EulerMaterialParameters materialParameters);
    static vector<double> computeXRichtmyerFlux(EulerStateVector leftStateVector, EulerStateVector rightStateVector,
                                                double cellSpacing, double timeStep,
                                                EulerMaterialParameters materialParameters);
    static vector<double> computeXFORCEFlux(EulerStateVector leftStateVector, EulerStateVector rightStateVector,
                                            double cellSpacing, double timeStep,
                                            EulerMaterialParameters materialParameters);
    static vector<double> computeYLaxFriedrichsFlux(EulerStateVector topStateVector, EulerStateVector bottomStateVector,
                                                    double cellSpacing, double timeStep,
                                                    EulerMaterialParameters materialParameters);
    static vector<double> computeYRichtmyerFlux(EulerStateVector topStateVector, EulerStateVector bottomStateVector,
                                                double cellSpacing, double timeStep,
                                                EulerMaterialParameters materialParameters);
    static vector<double> computeYFORCEFlux(EulerStateVector topStateVector, EulerStateVector bottomStateVector,
                                            double cellSpacing, double timeStep,
                                            EulerMaterialParameters materialParameters);
    static vector<double> computeFORCEUpdate(vector<double> conservedVariableVector, vector<double> leftFluxVector,
                                             vector<double> rightFluxVector, double cellSpacing, double timeStep);
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
    static vector<double> computeXSLICFlux(EulerStateVector leftLeftStateVector, EulerStateVector leftStateVector,
                                           EulerStateVector rightStateVector, EulerStateVector rightRightStateVector,
                                           double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                           EulerMaterialParameters materialParameters);
    static vector<double> computeYSLICFlux(EulerStateVector topTopStateVector, EulerStateVector topStateVector,
                                           EulerStateVector bottomStateVector, EulerStateVector bottomBottomStateVector,
                                           double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                           EulerMaterialParameters materialParameters);
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
    static vector<RelativisticEulerStateVector> insertBoundaryCells(vector<RelativisticEulerStateVector>& currentCells,
                                                                    int boundarySize);
    static vector<vector<RelativisticEulerStateVector> > insertBoundaryCells2D(
        vector<vector<RelativisticEulerStateVector> >& currentCells, int boundarySize);
    static RelativisticEulerStateVector evolveStateByHalfTimeStep(vector<double> leftExtrapolatedValue,
                                                                  vector<double> rightExtrapolatedValue,
                                                                  vector<double> evolutionVector, int side,
                                                                  EulerMaterialParameters materialParameters);
    static RelativisticEulerStateVector evolveStateByHalfXTimeStep(
        RelativisticEulerStateVector leftStateVector, RelativisticEulerStateVector middleStateVector,
        RelativisticEulerStateVector rightStateVector, double cellSpacing, double timeStep, double bias,
        int slopeLimiter, int side, EulerMaterialParameters materialParameters);
    static RelativisticEulerStateVector evolveStateByHalfYTimeStep(
        RelativisticEulerStateVector topStateVector, RelativisticEulerStateVector middleStateVector,
        RelativisticEulerStateVector bottomStateVector, double cellSpacing, double timeStep, double bias,
        int slopeLimiter, int side, EulerMaterialParameters materialParameters);
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
    static vector<double> computeXLaxFriedrichsFlux(RelativisticEulerStateVector leftStateVector,
                                                    RelativisticEulerStateVector rightStateVector, double cellSpacing,
                                                    double timeStep, EulerMaterialParameters materialParameters);
    static vector<double> computeXRichtmyerFlux(RelativisticEulerStateVector leftStateVector,
                                                RelativisticEulerStateVector rightStateVector, double cellSpacing,
                                                double timeStep, EulerMaterialParameters materialParameters);
    static vector<double> computeXFORCEFlux(RelativisticEulerStateVector leftStateVector,
                                            RelativisticEulerStateVector rightStateVector, double cellSpacing,
                                            double timeStep, EulerMaterialParameters materialParameters);
    static vector<double> computeYLaxFriedrichsFlux(RelativisticEulerStateVector topStateVector,
                                                    RelativisticEulerStateVector bottomStateVector, double cellSpacing,
                                                    double timeStep, EulerMaterialParameters materialParameters);
    static vector<double> computeYRichtmyerFlux(RelativisticEulerStateVector topStateVector,
                                                RelativisticEulerStateVector bottomStateVector, double cellSpacing,
                                                double timeStep, EulerMaterialParameters materialParameters);
    static vector<double> computeYFORCEFlux(RelativisticEulerStateVector topStateVector,
                                            RelativisticEulerStateVector bottomStateVector, double cellSpacing,
                                            double timeStep, EulerMaterialParameters materialParameters);
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
    static vector<double> computeXSLICFlux(RelativisticEulerStateVector leftLeftStateVector,
                                           RelativisticEulerStateVector leftStateVector,
                                           RelativisticEulerStateVector rightStateVector,
                                           RelativisticEulerStateVector rightRightStateVector, double cellSpacing,
                                           double timeStep, double bias, int slopeLimiter,
                                           EulerMaterialParameters materialParameters);
    static vector<double> computeYSLICFlux(RelativisticEulerStateVector topTopStateVector,
                                           RelativisticEulerStateVector topStateVector,
                                           RelativisticEulerStateVector bottomStateVector,
                                           RelativisticEulerStateVector bottomBottomStateVector, double cellSpacing,
                                           double timeStep, double bias, int slopeLimiter,
                                           EulerMaterialParameters materialParameters);
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
    static vector<BlackHoleEulerStateVector> insertBoundaryCells(vector<BlackHoleEulerStateVector>& currentCells,
                                                                 int boundarySize);
    static vector<vector<BlackHoleEulerStateVector> > insertBoundaryCells2D(
        vector<vector<BlackHoleEulerStateVector> >& currentCells, int boundarySize);
    static BlackHoleEulerStateVector evolveStateByFractionalTimeStep(vector<double> middleConservedVariableVector,
                                                                     vector<double> conservedVariableVectorEvolution,
                                                                     double xCoordinate, double yCoordinate,
                                                                     double zCoordinate,
                                                                     EulerMaterialParameters materialParameters,
                                                                     BlackHoleSpacetime blackHole);
    static BlackHoleEulerStateVector evolveStateByFractionalXTimeStep(
        double stepFraction, BlackHoleEulerStateVector leftStateVector, BlackHoleEulerStateVector middleStateVector,
        BlackHoleEulerStateVector rightStateVector, double cellSpacing, double timeStep, double bias, int slopeLimiter,
        double xCoordinate, double yCoordinate, double zCoordinate, EulerMaterialParameters materialParameters,
        BlackHoleSpacetime blackHole);
    static BlackHoleEulerStateVector evolveStateByFractionalYTimeStep(
        double stepFraction, BlackHoleEulerStateVector topStateVector, BlackHoleEulerStateVector middleStateVector,
        BlackHoleEulerStateVector bottomStateVector, double cellSpacing, double timeStep, double bias, int slopeLimiter,
        double xCoordinate, double yCoordinate, double zCoordinate, EulerMaterialParameters materialParameters,
        BlackHoleSpacetime blackHole);
    static BlackHoleEulerStateVector evolveStateByHalfXTimeStep(vector<double> leftExtrapolatedValue,
                                                                vector<double> rightExtrapolatedValue,
                                                                vector<double> evolutionVector, int side,
                                                                double cellSpacing, double xCoordinate,
                                                                double yCoordinate, double zCoordinate,
                                                                EulerMaterialParameters materialParameters,
                                                                BlackHoleSpacetime blackHole);
    static BlackHoleEulerStateVector evolveStateByHalfXTimeStep(
        BlackHoleEulerStateVector leftStateVector, BlackHoleEulerStateVector middleStateVector,
        BlackHoleEulerStateVector rightStateVector, double cellSpacing, double timeStep, double bias, int slopeLimiter,
        int side, double xCoordinate, double yCoordinate, double zCoordinate, EulerMaterialParameters materialParameters,
        BlackHoleSpacetime blackHole);
    static BlackHoleEulerStateVector evolveStateByHalfYTimeStep(vector<double> topExtrapolatedValue,
                                                                vector<double> bottomExtrapolatedValue,
                                                                vector<double> evolutionVector, int side,
                                                                double cellSpacing, double xCoordinate,
                                                                double yCoordinate, double zCoordinate,
                                                                EulerMaterialParameters materialParameters,
                                                                BlackHoleSpacetime blackHole);
    static BlackHoleEulerStateVector evolveStateByHalfYTimeStep(
        BlackHoleEulerStateVector topStateVector, BlackHoleEulerStateVector middleStateVector,
        BlackHoleEulerStateVector bottomStateVector, double cellSpacing, double timeStep, double bias, int slopeLimiter,
        int side, double xCoordinate, double yCoordinate, double zCoordinate, EulerMaterialParameters materialParameters,
        BlackHoleSpacetime blackHole);
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
    static vector<double> evolveConservedVariableVector(vector<double> leftConservedVariableVector,
                                                        vector<double> middleConservedVariableVector,
                                                        vector<double> rightConservedVaraibleVector,
                                                        double xCoordinate, double yCoordinate, double zCoordinate,
                                                        double cellSpacing, double timeStep, double bias,
                                                        int slopeLimiter, EulerMaterialParameters materialParameters,
                                                        BlackHoleSpacetime blackHole);
    static vector<double> evolveConservedVariableVector2D(
        vector<double> leftConservedVariableVector, vector<double> middleConservedVariableVector,
        vector<double> rightConservedVariableVector, vector<double> topConservedVariableVector,
        vector<double> bottomConservedVariableVector, double xCoordinate, double yCoordinate, double zCoordinate,
        double cellSpacing, double timeStep, double bias, int slopeLimiter, EulerMaterialParameters materialParameters,
        BlackHoleSpacetime blackHole);
    static void computeRungeKuttaTimeStep(vector<BlackHoleEulerStateVector>& currentCells, double cellSpacing,
                                          double timeStep, double bias, int slopeLimiter,
                                          EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static void computeRungeKuttaTimeStep2D(vector<vector<BlackHoleEulerStateVector> >& currentCells,
                                            double cellSpacing, double timeStep, double bias, int slopeLimiter,
                                            EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
};

class BlackHoleEulerFirstOrderSolver {
public:
    static vector<double> computeXLaxFriedrichsFlux(BlackHoleEulerStateVector leftStateVector,
                                                    BlackHoleEulerStateVector rightStateVector, double xCoordinate,
                                                    double yCoordinate, double zCoordinate, double cellSpacing,
                                                    double timeStep, EulerMaterialParameters materialParameters,
                                                    BlackHoleSpacetime blackHole);
    static vector<double> computeXRichtmyerFlux(BlackHoleEulerStateVector leftStateVector,
                                                BlackHoleEulerStateVector rightStateVector, double xCoordinate,
                                                double yCoordinate, double zCoordinate, double cellSpacing,
                                                double timeStep, EulerMaterialParameters materialParameters,
                                                BlackHoleSpacetime blackHole);
    static vector<double> computeXFORCEFlux(BlackHoleEulerStateVector leftStateVector,
                                            BlackHoleEulerStateVector rightStateVector, double xCoordinate,
                                            double yCoordinate, double zCoordinate, double cellSpacing,
                                            double timeStep, EulerMaterialParameters materialParameters,
                                            BlackHoleSpacetime blackHole);
    static vector<double> computeYLaxFriedrichsFlux(BlackHoleEulerStateVector topStateVector,
                                                    BlackHoleEulerStateVector bottomStateVector, double xCoordinate,
                                                    double yCoordinate, double zCoordinate, double cellSpacing,
                                                    double timeStep, EulerMaterialParameters materialParameters,
                                                    BlackHoleSpacetime blackHole);
    static vector<double> computeYRichtmyerFlux(BlackHoleEulerStateVector topStateVector,
                                                BlackHoleEulerStateVector bottomStateVector, double xCoordinate,
                                                double yCoordinate, double zCoordinate, double cellSpacing,
                                                double timeStep, EulerMaterialParameters materialParameters,
                                                BlackHoleSpacetime blackHole);
    static vector<double> computeYFORCEFlux(BlackHoleEulerStateVector topStateVector,
                                            BlackHoleEulerStateVector bottomStateVector, double xCoordinate,
                                            double yCoordinate, double zCoordinate, double cellSpacing,
                                            double timeStep, EulerMaterialParameters materialParameters,
                                            BlackHoleSpacetime blackHole);
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
    static vector<double> computeXSLICFlux(BlackHoleEulerStateVector leftLeftStateVector,
                                           BlackHoleEulerStateVector leftStateVector,
                                           BlackHoleEulerStateVector rightStateVector,
                                           BlackHoleEulerStateVector rightRightStateVector, double cellSpacing,
                                           double timeStep, double bias, int slopeLimiter, double xCoordinate,
                                           double yCoordinate, double zCoordinate,
                                           EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
    static vector<double> computeYSLICFlux(BlackHoleEulerStateVector topTopStateVector,
                                           BlackHoleEulerStateVector topStateVector,
                                           BlackHoleEulerStateVector bottomStateVector,
                                           BlackHoleEulerStateVector bottomBottomStateVector, double cellSpacing,
                                           double timeStep, double bias, int slopeLimiter, double xCoordinate,
                                           double yCoordinate, double zCoordinate,
                                           EulerMaterialParameters materialParameters, BlackHoleSpacetime blackHole);
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
