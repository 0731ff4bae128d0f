// Common pieces of the registered per-thread systems.
//
// A system is a struct with the compile-time sizes of the reference's
// ProblemSolver template (SD, NCP, NSP, NISP, NE, NA, NIA; see
// SourceCodes/SingleSystem_PerThread_Interface.cuh:117) and the six user
// callbacks of the reference, as static member templates with the reference's
// argument lists (Tutorials_SingleSystem_PerThread/T1_Reference/
// Reference_SystemDefinition.cuh:5-76):
//   OdeFunction, EventFunction, ActionAfterEventDetection,
//   ActionAfterSuccessfulTimeStep, Initialization, Finalization.
// Systems that support the discrete adjoint also provide
//   OdeVJP(tid, NT, lam, X, T, cPAR, sPAR, sPARi, ACC, ACCi, gX, gP)
// setting gX = (dF/dX)^T lam and gP = (dF/dcPAR)^T lam.
// Every callback is MPGOS_HD so the same definition serves the gfx950 kernels
// and the host oracle; elementary functions come from mpgos::math so both
// sides evaluate identical bits.
#pragma once

#include "../mpgos_math.h"

namespace mpgos {
namespace systems {

// Callbacks that do nothing, for systems without events or accessories.
struct NoCallbacks {
    template <class P> MPGOS_HD static void EventFunction(int, int, P*, P, P, P*, P*, P*, P*, int*, P*, int*) {}
    template <class P> MPGOS_HD static void ActionAfterEventDetection(int, int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}
    template <class P> MPGOS_HD static void ActionAfterSuccessfulTimeStep(int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}
    template <class P> MPGOS_HD static void Initialization(int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}
    template <class P> MPGOS_HD static void Finalization(int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}
};

} // namespace systems
} // namespace mpgos
