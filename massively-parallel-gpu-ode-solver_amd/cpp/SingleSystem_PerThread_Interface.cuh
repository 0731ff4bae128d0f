// Drop-in C++ facade of the reference's per-thread module for gfx950.
//
// Same file name, class template, enumerations and member functions as
// SourceCodes/SingleSystem_PerThread_Interface.cuh (ProblemSolver at :117-208,
// enums at :33-62, device helpers at :218-304), so a driver written for the
// reference compiles with hipcc against this header unchanged: the user's
// system-definition header (PerThread_OdeFunction, PerThread_EventFunction,
// PerThread_ActionAfterEventDetection, PerThread_ActionAfterSuccessfulTimeStep,
// PerThread_Initialization, PerThread_Finalization as __device__ function
// templates over Precision) is included first, exactly as in
// Tutorials_SingleSystem_PerThread/T1_Reference/Reference.cu:7-8.
//
// Underneath, the object is the library's native engine (csrc/pt_engine.h) over
// the gfx950 kernels (csrc/pt_device.h), instantiated with the user callbacks.
// Errors keep the reference's behaviour: a message and exit(EXIT_FAILURE)
// (the C ABI in include/mpgos.h returns codes instead).  Native() exposes the
// engine for the options that have no counterpart in the reference
// (persistent form, graph replay, event log, launch variants).
//
// Build:  hipcc --offload-arch=gfx950 -O3 -std=c++17 -ffp-contract=off \
//           -I<repo>/include -I<pkg>/csrc -I<pkg>/cpp driver.cu
// The header is self-contained (no link against libmpgos.so).  It defines the
// engine's error sink; a program with several translation units including it
// defines MPGOS_EXTERNAL_ERROR_SINK in all but one of them.
#ifndef MPGOS_SINGLESYSTEM_PERTHREAD_INTERFACE_CUH
#define MPGOS_SINGLESYSTEM_PERTHREAD_INTERFACE_CUH

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <iostream>
#include <string>

#include "mpgos.h"
#include "pt_engine.h"

#if !defined(MPGOS_EXTERNAL_ERROR_SINK) && !defined(MPGOS_FACADE_ERROR_SINK_DEFINED)
#define MPGOS_FACADE_ERROR_SINK_DEFINED
namespace mpgos {
// Engine errors are reported the way the reference reports them (stderr).
void set_error(const std::string& msg) { std::cerr << "ERROR: " << msg << std::endl; }
} // namespace mpgos
#endif

enum Algorithms { RK4, RKCK45 };

enum ListOfVariables { All, TimeDomain, ActualState, ActualTime, ControlParameters, SharedParameters,
                       IntegerSharedParameters, Accessories, IntegerAccessories, DenseOutput, DenseIndex,
                       DenseTime, DenseState };

enum ListOfSolverOptions { ThreadsPerBlock, InitialTimeStep, ActiveNumberOfThreads, MaximumTimeStep,
                           MinimumTimeStep, TimeStepGrowLimit, TimeStepShrinkLimit, RelativeTolerance,
                           AbsoluteTolerance, EventTolerance, EventDirection, DenseOutputMinimumTimeStep,
                           DenseOutputSaveFrequency, PreferSharedMemory };

inline void ListCUDADevices()
{
    int n = 0;
    (void)hipGetDeviceCount(&n);
    for (int i = 0; i < n; i++) {
        hipDeviceProp_t p;
        (void)hipGetDeviceProperties(&p, i);
        std::cout << "\nDevice number: " << i << "\nDevice name:   " << p.name << "\n"
                  << "Total global memory:        " << p.totalGlobalMem / 1024 / 1024 << " Mb\n"
                  << "Total shared memory:        " << p.sharedMemPerBlock / 1024 << " Kb\n"
                  << "Number of multiprocessors:  " << p.multiProcessorCount << "\n"
                  << "Architecture:               " << p.gcnArchName << "\n"
                  << "Warp size:                  " << p.warpSize << "\n";
    }
    std::cout << std::endl;
}

// The reference picks the CUDA device closest to a compute capability
// (PT-IF:256-270); on this platform every device is a gfx950, so the first one
// is returned.
inline int SelectDeviceByClosestRevision(int, int) { return 0; }

inline void PrintPropertiesOfSpecificDevice(int device)
{
    hipDeviceProp_t p;
    (void)hipGetDeviceProperties(&p, device);
    std::cout << "Selected device number: " << device << "\nSelected device name:   " << p.name << " ("
              << p.gcnArchName << ")\n" << std::endl;
}

namespace mpgos {
namespace facade {

// The user's free callback templates as a system struct for the engine.
template <int SD_, int NCP_, int NSP_, int NISP_, int NE_, int NA_, int NIA_>
struct UserSystem {
    static constexpr const char* name = "user";
    static constexpr int SD = SD_, NCP = NCP_, NSP = NSP_, NISP = NISP_, NE = NE_, NA = NA_, NIA = NIA_;
    static constexpr bool HAS_VJP = false;

    template <class P> __device__ static void OdeFunction(int tid, int NT, P* F, P* X, P T, P* cPAR, P* sPAR, int* sPARi, P* ACC, int* ACCi)
    { PerThread_OdeFunction<P>(tid, NT, F, X, T, cPAR, sPAR, sPARi, ACC, ACCi); }
    template <class P> __device__ static void EventFunction(int tid, int NT, P* EF, P T, P dT, P* TD, P* X, P* cPAR, P* sPAR, int* sPARi, P* ACC, int* ACCi)
    { PerThread_EventFunction<P>(tid, NT, EF, T, dT, TD, X, cPAR, sPAR, sPARi, ACC, ACCi); }
    template <class P> __device__ static void ActionAfterEventDetection(int tid, int NT, int IDX, int& UDT, P& T, P& dT, P* TD, P* X, P* cPAR, P* sPAR, int* sPARi, P* ACC, int* ACCi)
    { PerThread_ActionAfterEventDetection<P>(tid, NT, IDX, UDT, T, dT, TD, X, cPAR, sPAR, sPARi, ACC, ACCi); }
    template <class P> __device__ static void ActionAfterSuccessfulTimeStep(int tid, int NT, int& UDT, P& T, P& dT, P* TD, P* X, P* cPAR, P* sPAR, int* sPARi, P* ACC, int* ACCi)
    { PerThread_ActionAfterSuccessfulTimeStep<P>(tid, NT, UDT, T, dT, TD, X, cPAR, sPAR, sPARi, ACC, ACCi); }
    template <class P> __device__ static void Initialization(int tid, int NT, int& DOIDX, P& T, P& dT, P* TD, P* X, P* cPAR, P* sPAR, int* sPARi, P* ACC, int* ACCi)
    { PerThread_Initialization<P>(tid, NT, DOIDX, T, dT, TD, X, cPAR, sPAR, sPARi, ACC, ACCi); }
    template <class P> __device__ static void Finalization(int tid, int NT, int& DOIDX, P& T, P& dT, P* TD, P* X, P* cPAR, P* sPAR, int* sPARi, P* ACC, int* ACCi)
    { PerThread_Finalization<P>(tid, NT, DOIDX, T, dT, TD, X, cPAR, sPAR, sPARi, ACC, ACCi); }
};

inline void die(const char* where, int rc)
{
    if (rc == MPGOS_OK) return;
    std::cerr << "ERROR: In solver member function " << where << " (code " << rc << ")" << std::endl;
    std::exit(rc == MPGOS_E_HIP ? 1 : EXIT_FAILURE);
}

} // namespace facade
} // namespace mpgos

template <int NT, int SD, int NCP, int NSP, int NISP, int NE, int NA, int NIA, int NDO, Algorithms Algorithm, class Precision>
class ProblemSolver {
    using Sys = mpgos::facade::UserSystem<SD, NCP, NSP, NISP, NE, NA, NIA>;
    using Engine = mpgos::pt::Solver<Sys, Precision, false>;
    Engine engine;

public:
    explicit ProblemSolver(int AssociatedDevice)
    {
        mpgos::facade::die("ProblemSolver", engine.init(Algorithm == RK4 ? MPGOS_RK4 : MPGOS_RKCK45, NT, NDO, AssociatedDevice));
    }
    ~ProblemSolver() {}

    template <typename T> void SolverOption(ListOfSolverOptions Option, T Value)
    {
        if (Option == RelativeTolerance || Option == AbsoluteTolerance || Option == EventTolerance || Option == EventDirection) {
            std::cerr << "ERROR: In solver member function SolverOption!\n       This option needs 2 input arguments or not applicable!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        mpgos::facade::die("SolverOption", engine.option((int)Option, 0, (double)Value));
    }
    template <typename T> void SolverOption(ListOfSolverOptions Option, int Index, T Value)
    {
        if (!(Option == RelativeTolerance || Option == AbsoluteTolerance || Option == EventTolerance || Option == EventDirection)) {
            std::cerr << "ERROR: In solver member function SolverOption!\n       This option needs 1 input arguments or not applicable!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        mpgos::facade::die("SolverOption", engine.option((int)Option, Index, (double)Value));
    }

    template <typename T> void SetHost(int ProblemNumber, ListOfVariables Variable, int SerialNumber, T Value)
    { mpgos::facade::die("SetHost", engine.set_host((int)Variable, ProblemNumber, SerialNumber, 0, (double)Value)); }
    template <typename T> void SetHost(int ProblemNumber, ListOfVariables Variable, T Value)
    { mpgos::facade::die("SetHost", engine.set_host((int)Variable, ProblemNumber, 0, 0, (double)Value)); }
    template <typename T> void SetHost(ListOfVariables Variable, int SerialNumber, T Value)
    { mpgos::facade::die("SetHost", engine.set_host((int)Variable, 0, SerialNumber, 0, (double)Value)); }
    template <typename T> void SetHost(int ProblemNumber, ListOfVariables Variable, int SerialNumber, int TimeStepNumber, T Value)
    { mpgos::facade::die("SetHost", engine.set_host((int)Variable, ProblemNumber, SerialNumber, TimeStepNumber, (double)Value)); }

    void SynchroniseFromHostToDevice(ListOfVariables Variable) { mpgos::facade::die("SynchroniseFromHostToDevice", engine.sync((int)Variable, true)); }
    void SynchroniseFromDeviceToHost(ListOfVariables Variable) { mpgos::facade::die("SynchroniseFromDeviceToHost", engine.sync((int)Variable, false)); }

    template <typename T> T GetHost(int ProblemNumber, ListOfVariables Variable, int SerialNumber)
    { double v = 0; mpgos::facade::die("GetHost", engine.get_host((int)Variable, ProblemNumber, SerialNumber, 0, &v)); return (T)v; }
    template <typename T> T GetHost(int ProblemNumber, ListOfVariables Variable)
    { double v = 0; mpgos::facade::die("GetHost", engine.get_host((int)Variable, ProblemNumber, 0, 0, &v)); return (T)v; }
    template <typename T> T GetHost(ListOfVariables Variable, int SerialNumber)
    { double v = 0; mpgos::facade::die("GetHost", engine.get_host((int)Variable, 0, SerialNumber, 0, &v)); return (T)v; }
    template <typename T> T GetHost(int ProblemNumber, ListOfVariables Variable, int SerialNumber, int TimeStepNumber)
    { double v = 0; mpgos::facade::die("GetHost", engine.get_host((int)Variable, ProblemNumber, SerialNumber, TimeStepNumber, &v)); return (T)v; }

    void Print(ListOfVariables Variable) { mpgos::facade::die("Print", engine.print((int)Variable, 0, nullptr)); }
    void Print(ListOfVariables Variable, int ThreadID) { mpgos::facade::die("Print", engine.print((int)Variable, ThreadID, nullptr)); }

    void Solve() { mpgos::facade::die("Solve", engine.solve()); }
    void SynchroniseDevice() { mpgos::facade::die("SynchroniseDevice", engine.synchronise_device()); }
    void InsertSynchronisationPoint() { mpgos::facade::die("InsertSynchronisationPoint", engine.insert_sync_point()); }
    void SynchroniseSolver() { mpgos::facade::die("SynchroniseSolver", engine.synchronise_solver()); }

    // Options and outputs that have no counterpart in the reference.
    Engine& Native() { return engine; }
};

#endif
