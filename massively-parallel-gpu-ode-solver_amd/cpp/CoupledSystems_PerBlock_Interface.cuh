// Drop-in C++ facade of the reference's per-block (coupled systems) module
// for gfx950.
//
// Same file name, class template, enumerations and member functions as
// SourceCodes/CoupledSystems_PerBlock_Interface.cuh (ProblemSolver at
// :135-248, enums at :33-68), so a driver written for the reference compiles
// with hipcc against this header unchanged: the user's system-definition
// header (CoupledSystems_PerBlock_OdeFunction, _EventFunction,
// _ActionAfterEventDetection, _ActionAfterSuccessfulTimeStep,
// _Initialization, _Finalization as __device__ function templates over
// Precision) is included first, as in
// Tutorials_CoupledSystems_PerBlock/T1_RingOfDuffingOscillators/RingOfDuffingOscillators.cu:10-11.
//
// Underneath, the object is the library's per-block engine (csrc/pb_engine.h)
// over the gfx950 kernels (csrc/pb_device.h) instantiated with the user
// callbacks.  TPB is accepted and only used to lay out the unit-scope Print
// files as the reference does (its thread padding); the launch shape is the
// engine's (SPB * UPS lanes rounded up to whole wavefronts, or the
// dense-contraction kernel for a full FP64 coupling matrix).  Errors keep the
// reference's behaviour: a message and exit(EXIT_FAILURE).  Native() exposes
// the engine for options without a reference counterpart (coupling kernel
// choice, timing).
//
// Build:  hipcc --offload-arch=gfx950 -O3 -std=c++17 -ffp-contract=off \
//           -I<repo>/include -I<pkg>/csrc -I<pkg>/cpp driver.cu
#ifndef MPGOS_COUPLEDSYSTEMS_PERBLOCK_INTERFACE_CUH
#define MPGOS_COUPLEDSYSTEMS_PERBLOCK_INTERFACE_CUH

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "mpgos.h"
#include "pb_engine.h"

#if !defined(MPGOS_EXTERNAL_ERROR_SINK) && !defined(MPGOS_FACADE_ERROR_SINK_DEFINED)
#define MPGOS_FACADE_ERROR_SINK_DEFINED
namespace mpgos {
// Engine errors are reported the way the reference reports them (stderr).
void set_error(const std::string& msg) { std::cerr << "ERROR: " << msg << std::endl; }
} // namespace mpgos
#endif

enum Algorithms { RK4, RKCK45 };

enum ListOfVariables { All, TimeDomain, ActualTime, ActualState, UnitParameters, SystemParameters, GlobalParameters,
                       IntegerGlobalParameters, UnitAccessories, IntegerUnitAccessories, SystemAccessories,
                       IntegerSystemAccessories, CouplingMatrix, CouplingStrength, CouplingIndex, DenseOutput,
                       DenseIndex, DenseTime, DenseState };

enum ListOfSolverOptions { InitialTimeStep, ActiveSystems, MaximumTimeStep, MinimumTimeStep, TimeStepGrowLimit,
                           TimeStepShrinkLimit, RelativeTolerance, AbsoluteTolerance, EventTolerance, EventDirection,
                           DenseOutputSaveFrequency, DenseOutputMinimumTimeStep, SharedGlobalVariables,
                           SharedCouplingMatrices };

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

// Every device here is a gfx950: the first one is returned.
inline int SelectDeviceByClosestRevision(int, int) { return 0; }

inline void PrintPropertiesOfSpecificDevice(int device)
{
    hipDeviceProp_t p;
    (void)hipGetDeviceProperties(&p, device);
    std::cout << "Selected device number: " << device << "\nSelected device name:   " << p.name << " ("
              << p.gcnArchName << ")\n" << std::endl;
}

namespace mpgos {
namespace facade_pb {

// The user's free callback templates as a per-block system struct.
template <int UD_, int NC_, int NUP_, int NSP_, int NGP_, int NiGP_, int NUA_, int NiUA_, int NSA_, int NiSA_, int NE_>
struct UserSystem {
    static constexpr const char* name = "user";
    static constexpr int UD = UD_, NC = NC_, NUP = NUP_, NSP = NSP_, NGP = NGP_, NiGP = NiGP_, NUA = NUA_, NiUA = NiUA_,
                         NSA = NSA_, NiSA = NiSA_, NE = NE_;
#define MPGOS_PBF_ARGS P* uPAR, P* sPAR, P* gPAR, int* igPAR, P* uACC, int* iuACC, P* sACC, int* isACC
#define MPGOS_PBF_PASS uPAR, sPAR, gPAR, igPAR, uACC, iuACC, sACC, isACC
    template <class P> __device__ static void OdeFunction(int sid, int uid, P* F, P* X, P T, MPGOS_PBF_ARGS, P* CPT, P* CPF)
    { CoupledSystems_PerBlock_OdeFunction<P>(sid, uid, F, X, T, MPGOS_PBF_PASS, CPT, CPF); }
    template <class P> __device__ static void EventFunction(int sid, int uid, P* EF, P T, P dT, P* TD, P* X, MPGOS_PBF_ARGS)
    { CoupledSystems_PerBlock_EventFunction<P>(sid, uid, EF, T, dT, TD, X, MPGOS_PBF_PASS); }
    template <class P> __device__ static void ActionAfterEventDetection(int sid, int uid, int IDX, int& UDT, P& T, P& dT, P* TD, P* X, MPGOS_PBF_ARGS)
    { CoupledSystems_PerBlock_ActionAfterEventDetection<P>(sid, uid, IDX, UDT, T, dT, TD, X, MPGOS_PBF_PASS); }
    template <class P> __device__ static void ActionAfterSuccessfulTimeStep(int sid, int uid, int& UDT, P& T, P& dT, P* TD, P* X, MPGOS_PBF_ARGS)
    { CoupledSystems_PerBlock_ActionAfterSuccessfulTimeStep<P>(sid, uid, UDT, T, dT, TD, X, MPGOS_PBF_PASS); }
    template <class P> __device__ static void Initialization(int sid, int uid, int& DOIDX, P& T, P& dT, P* TD, P* X, MPGOS_PBF_ARGS)
    { CoupledSystems_PerBlock_Initialization<P>(sid, uid, DOIDX, T, dT, TD, X, MPGOS_PBF_PASS); }
    template <class P> __device__ static void Finalization(int sid, int uid, int& DOIDX, P& T, P& dT, P* TD, P* X, MPGOS_PBF_ARGS)
    { CoupledSystems_PerBlock_Finalization<P>(sid, uid, DOIDX, T, dT, TD, X, MPGOS_PBF_PASS); }
#undef MPGOS_PBF_ARGS
#undef MPGOS_PBF_PASS
};

inline void die(const char* where, int rc)
{
    if (rc == MPGOS_OK) return;
    std::cerr << "ERROR: In solver member function " << where << " (code " << rc << ")" << std::endl;
    std::exit(EXIT_FAILURE);
}

} // namespace facade_pb
} // namespace mpgos

template <int NS, int UPS, int UD, int TPB, int SPB, int NC, int CBW, int CCI, int NUP, int NSP, int NGP, int NiGP, int NUA,
          int NiUA, int NSA, int NiSA, int NE, int NDO, Algorithms Algorithm, class Precision>
class ProblemSolver {
    using Sys = mpgos::facade_pb::UserSystem<UD, NC, NUP, NSP, NGP, NiGP, NUA, NiUA, NSA, NiSA, NE>;
    using Engine = mpgos::pb::Solver<Sys, Precision>;
    Engine engine;

    static void die(const char* w, int rc) { mpgos::facade_pb::die(w, rc); }
    template <class T> T* host(ListOfVariables v) { size_t n = 0; return static_cast<T*>(engine.buffer((int)v, true, &n)); }

    // the reference's logical-thread layout (PB-IF:386-391): blocks of SPB*UPS
    // logical threads padded to a multiple of TPB
    static constexpr int kLogical = SPB * UPS;
    static constexpr int kPadded = (kLogical / TPB + (kLogical % TPB == 0 ? 0 : 1)) * TPB;
    static constexpr int kGrid = NS / SPB + (NS % SPB == 0 ? 0 : 1);
    static constexpr long kTotalLogical = (long)kPadded * kGrid;

    static void open(std::ofstream& f, const std::string& name)
    {
        f.open(name);
        f.precision(10);
        f.flags(std::ios::scientific);
    }
    // unit scope: one row per logical thread of the reference, PADDING where it pads (PB-IF:1533-1571)
    template <class T> void print_unit(const std::string& name, int columns, const T* data)
    {
        std::ofstream f;
        open(f, name);
        const long limit = (long)(NS / SPB) * kPadded + (long)(NS % SPB) * UPS;
        const long NU = (long)NS * UPS;
        for (long i = 0; i < kTotalLogical; i++) {
            const long local = i % kPadded, block = i / kPadded;
            const bool real = local < kLogical && i < limit;
            const long unit = block * kLogical + local;      // (system * UPS + unit) in this library's layout
            for (int j = 0; j < columns; j++) {
                f.width(18);
                if (real) f << data[unit + j * NU] << ',';
                else f << "PADDING" << ',';
            }
            f << '\n';
        }
    }
    // system and global scope (PB-IF:1576-1597)
    template <class T> void print_plain(const std::string& name, int rows, int columns, const T* data)
    {
        std::ofstream f;
        open(f, name);
        for (int i = 0; i < rows; i++) {
            for (int j = 0; j < columns; j++) { f.width(18); f << data[i + (long)j * rows] << ','; }
            f << '\n';
        }
    }

public:
    explicit ProblemSolver(int AssociatedDevice)
    {
        static_assert(TPB > 0 && SPB > 0 && UPS > 0, "TPB, SPB and UPS must be positive");
        die("ProblemSolver", engine.init(Algorithm == RK4 ? MPGOS_RK4 : MPGOS_RKCK45, NS, UPS, SPB, CBW, CCI, NDO, AssociatedDevice));
    }
    ~ProblemSolver() {}

    template <typename T> void SolverOption(ListOfSolverOptions Option, T Value)
    {
        if (Option == RelativeTolerance || Option == AbsoluteTolerance || Option == EventTolerance || Option == EventDirection) {
            std::cerr << "ERROR: In solver member function SolverOption!\n       This option needs 2 input arguments or not applicable!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        die("SolverOption", engine.option((int)Option, 0, (double)Value));
    }
    template <typename T> void SolverOption(ListOfSolverOptions Option, int Index, T Value)
    {
        if (!(Option == RelativeTolerance || Option == AbsoluteTolerance || Option == EventTolerance || Option == EventDirection)) {
            std::cerr << "ERROR: In solver member function SolverOption!\n       This option needs 1 input arguments or not applicable!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        die("SolverOption", engine.option((int)Option, Index, (double)Value));
    }

    // unit scope
    template <typename T> void SetHost(int SystemNumber, int UnitNumber, ListOfVariables Variable, int SerialNumber, T Value)
    { die("SetHost", engine.set_host((int)Variable, SystemNumber, UnitNumber, SerialNumber, 0, (double)Value)); }
    // system scope and dense time
    template <typename T> void SetHost(int SystemNumber, ListOfVariables Variable, int SerialNumber, T Value)
    { die("SetHost", engine.set_host((int)Variable, SystemNumber, 0, SerialNumber, 0, (double)Value)); }
    // global scope
    template <typename T> void SetHost(ListOfVariables Variable, int SerialNumber, T Value)
    { die("SetHost", engine.set_host((int)Variable, 0, 0, SerialNumber, 0, (double)Value)); }
    // coupling matrix
    template <typename T> void SetHost(int CouplingNumber, ListOfVariables Variable, int Row, int Col, T Value)
    { die("SetHost", engine.set_host((int)Variable, CouplingNumber, Row, Col, 0, (double)Value)); }
    // dense index and actual time
    template <typename T> void SetHost(int SystemNumber, ListOfVariables Variable, T Value)
    { die("SetHost", engine.set_host((int)Variable, SystemNumber, 0, 0, 0, (double)Value)); }
    // dense state
    template <typename T> void SetHost(int SystemNumber, int UnitNumber, ListOfVariables Variable, int ComponentNumber, int TimeStep, T Value)
    { die("SetHost", engine.set_host((int)Variable, SystemNumber, UnitNumber, ComponentNumber, TimeStep, (double)Value)); }

    void SynchroniseFromHostToDevice(ListOfVariables Variable) { die("SynchroniseFromHostToDevice", engine.sync((int)Variable, true)); }
    void SynchroniseFromDeviceToHost(ListOfVariables Variable) { die("SynchroniseFromDeviceToHost", engine.sync((int)Variable, false)); }

    template <typename T> T GetHost(int SystemNumber, int UnitNumber, ListOfVariables Variable, int SerialNumber)
    { double v = 0; die("GetHost", engine.get_host((int)Variable, SystemNumber, UnitNumber, SerialNumber, 0, &v)); return (T)v; }
    template <typename T> T GetHost(int SystemNumber, ListOfVariables Variable, int SerialNumber)
    { double v = 0; die("GetHost", engine.get_host((int)Variable, SystemNumber, 0, SerialNumber, 0, &v)); return (T)v; }
    template <typename T> T GetHost(ListOfVariables Variable, int SerialNumber)
    { double v = 0; die("GetHost", engine.get_host((int)Variable, 0, 0, SerialNumber, 0, &v)); return (T)v; }
    template <typename T> T GetHost(int CouplingNumber, ListOfVariables Variable, int Row, int Col)
    { double v = 0; die("GetHost", engine.get_host((int)Variable, CouplingNumber, Row, Col, 0, &v)); return (T)v; }
    template <typename T> T GetHost(int SystemNumber, ListOfVariables Variable)
    { double v = 0; die("GetHost", engine.get_host((int)Variable, SystemNumber, 0, 0, 0, &v)); return (T)v; }
    template <typename T> T GetHost(int SystemNumber, int UnitNumber, ListOfVariables Variable, int ComponentNumber, int TimeStep)
    { double v = 0; die("GetHost", engine.get_host((int)Variable, SystemNumber, UnitNumber, ComponentNumber, TimeStep, &v)); return (T)v; }

    // PB-IF:1808-1912: file names and number format of the reference
    void Print(ListOfVariables Variable)
    {
        switch (Variable) {
        case TimeDomain: print_plain("TimeDomainInSolverObject.txt", NS, 2, host<Precision>(Variable)); break;
        case ActualTime: print_plain("ActualTimeInSolverObject.txt", NS, 1, host<Precision>(Variable)); break;
        case ActualState: print_unit("ActualStateInSolverObject.txt", UD, host<Precision>(Variable)); break;
        case UnitParameters: print_unit("UnitParametersInSolverObject.txt", NUP, host<Precision>(Variable)); break;
        case SystemParameters: print_plain("SystemParametersInSolverObject.txt", NS, NSP, host<Precision>(Variable)); break;
        case GlobalParameters: print_plain("GlobalParametersInSolverObject.txt", NGP, 1, host<Precision>(Variable)); break;
        case IntegerGlobalParameters: print_plain("IntegerGlobalParametersInSolverObject.txt", NiGP, 1, host<int>(Variable)); break;
        case UnitAccessories: print_unit("UnitAccessoriesInSolverObject.txt", NUA, host<Precision>(Variable)); break;
        case IntegerUnitAccessories: print_unit("IntegerUnitAccessoriesInSolverObject.txt", NiUA, host<int>(Variable)); break;
        case SystemAccessories: print_plain("SystemAccessoriesInSolverObject.txt", NS, NSA, host<Precision>(Variable)); break;
        case IntegerSystemAccessories: print_plain("IntegerSystemAccessoriesInSolverObject.txt", NS, NiSA, host<int>(Variable)); break;
        case CouplingStrength: print_plain("CouplingStrengthInSolverObject.txt", NS, NC, host<Precision>(Variable)); break;
        case CouplingIndex: print_plain("CouplingIndexInSolverObject.txt", NC, 1, host<int>(Variable)); break;
        default:
            std::cerr << "ERROR: In solver member function Print!\n       This option needs different argument configuration or not applicable!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
    }
    // PB-IF:1917-1945 (coupling matrix as set; dense output of one system, format of PB-IF:1602-1700)
    void Print(ListOfVariables Variable, int SerialNumber)
    {
        if (Variable == CouplingMatrix) {
            if (SerialNumber < 0 || SerialNumber >= NC) die("Print", MPGOS_E_ARG);
            std::vector<double> full((size_t)UPS * UPS);
            for (int r = 0; r < UPS; r++)
                for (int c = 0; c < UPS; c++) full[r + (size_t)c * UPS] = GetHost<double>(SerialNumber, CouplingMatrix, r, c);
            print_plain("CouplingMatrixInSolverObject_" + std::to_string(SerialNumber) + ".txt", UPS, UPS, full.data());
            return;
        }
        if (Variable != DenseOutput) {
            std::cerr << "ERROR: In solver member function Print!\n       This option needs different argument configuration or not applicable!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        if (SerialNumber < 0 || SerialNumber >= NS) die("Print", MPGOS_E_ARG);
        const long NU = (long)NS * UPS;
        const Precision* up = host<Precision>(UnitParameters);
        const Precision* dt = host<Precision>(DenseTime);
        const Precision* dx = host<Precision>(DenseState);
        const int rows = host<int>(DenseIndex)[SerialNumber];
        std::ofstream f;
        open(f, "DenseOutput_" + std::to_string(SerialNumber) + ".txt");
        auto header = [&]() {
            f.width(18); f << "Serial No." << ',';
            for (int i = 0; i < UPS; i++) { f.width(18); f << "U" + std::to_string(i); if (i < UPS - 1) f << ','; }
            f << '\n';
        };
        f << "UnitParameters:\n";
        header();
        for (int i = 0; i < NUP; i++) {
            f.width(18); f << i << ',';
            for (int j = 0; j < UPS; j++) {
                f.width(18); f << up[(long)SerialNumber * UPS + j + i * NU];
                if (j < UPS - 1) f << ',';
            }
            f << '\n';
        }
        f << '\n';
        f << "UnitAccessories:\n";
        header();
        f << '\n' << '\n';
        f << "Time series:\n";
        f.width(18); f << "t" << ',';
        for (int j = 0; j < UPS; j++)
            for (int k = 0; k < UD; k++) {
                f.width(18); f << "U" + std::to_string(j) + "/C" + std::to_string(k);
                if (j < UPS - 1 || k < UD - 1) f << ',';
            }
        f << '\n';
        for (int i = 0; i < rows && i < NDO; i++) {
            f.width(18); f << dt[SerialNumber + (long)i * NS] << ',';
            for (int j = 0; j < UPS; j++)
                for (int k = 0; k < UD; k++) {
                    f.width(18); f << dx[(long)SerialNumber * UPS + j + k * NU + (long)i * NU * UD];
                    if (j < UPS - 1 || k < UD - 1) f << ',';
                }
            f << '\n';
        }
    }

    void Solve() { die("Solve", engine.solve()); }
    void SynchroniseDevice() { die("SynchroniseDevice", engine.synchronise_device()); }
    void InsertSynchronisationPoint() { die("InsertSynchronisationPoint", engine.insert_sync_point()); }
    void SynchroniseSolver() { die("SynchroniseSolver", engine.synchronise_solver()); }

    Engine& Native() { return engine; }
};

#endif
