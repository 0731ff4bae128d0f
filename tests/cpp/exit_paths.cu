// Exit paths of the drop-in per-thread facade: a driver written against the
// reference's interface that makes one invalid call, chosen by argv[1].  The
// reference terminates the process on each of them with a message and
// exit(EXIT_FAILURE) (index checks: SingleSystem_PerThread_Interface.cuh:565-581;
// wrong SolverOption argument count: :659-663; CUDA errors: exit(1), :13-22);
// the facade must do the same.  Mode "ok" makes valid calls only and returns 0.
#include <cstdio>
#include <cstring>

#include "systems/lorenz.h"

template <class P>
__forceinline__ __device__ void PerThread_OdeFunction(int tid, int NT, P* F, P* X, P T, P* cPAR, P* sPAR, int* sPARi, P* ACC, int* ACCi)
{
    mpgos::systems::Lorenz::OdeFunction<P>(tid, NT, F, X, T, cPAR, sPAR, sPARi, ACC, ACCi);
}
template <class P>
__forceinline__ __device__ void PerThread_EventFunction(int, int, P*, P, P, P*, P*, P*, P*, int*, P*, int*) {}
template <class P>
__forceinline__ __device__ void PerThread_ActionAfterEventDetection(int, int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}
template <class P>
__forceinline__ __device__ void PerThread_ActionAfterSuccessfulTimeStep(int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}
template <class P>
__forceinline__ __device__ void PerThread_Initialization(int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}
template <class P>
__forceinline__ __device__ void PerThread_Finalization(int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}

#include "SingleSystem_PerThread_Interface.cuh"

const int NT = 256, SD = 3, NCP = 1, NDO = 4;

int main(int argc, char** argv)
{
    const char* mode = argc > 1 ? argv[1] : "ok";
    ProblemSolver<NT, SD, NCP, 0, 0, 0, 0, 0, NDO, RKCK45, double> S(0);
    if (!std::strcmp(mode, "set_problem")) S.SetHost(NT, ActualState, 0, 1.0);
    else if (!std::strcmp(mode, "set_component")) S.SetHost(0, ActualState, SD, 1.0);
    else if (!std::strcmp(mode, "get_parameter")) (void)S.GetHost<double>(0, ControlParameters, NCP);
    else if (!std::strcmp(mode, "dense_index")) S.SetHost(0, DenseIndex, NDO);
    else if (!std::strcmp(mode, "dense_state")) (void)S.GetHost<double>(0, DenseState, 0, NDO);
    else if (!std::strcmp(mode, "option_needs_index")) S.SolverOption(RelativeTolerance, 1e-9);
    else if (!std::strcmp(mode, "option_takes_no_index")) S.SolverOption(InitialTimeStep, 0, 1e-3);
    else if (!std::strcmp(mode, "tolerance_index")) S.SolverOption(AbsoluteTolerance, SD, 1e-9);
    else if (!std::strcmp(mode, "print_thread")) S.Print(DenseOutput, NT);
    else {
        S.SetHost(NT - 1, ActualState, SD - 1, 1.0);
        S.SolverOption(RelativeTolerance, SD - 1, 1e-9);
        std::printf("ok %g\n", S.GetHost<double>(NT - 1, ActualState, SD - 1));
        return 0;
    }
    std::printf("returned\n");   // not reached: every invalid call above terminates
    return 0;
}
