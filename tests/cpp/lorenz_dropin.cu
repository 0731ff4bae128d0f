// A driver written against the reference's per-thread interface (the shape of
// Tutorials_SingleSystem_PerThread/T2_Lorenz/Lorenz.cu), compiled unchanged
// against the drop-in header.  Its callbacks forward to the registered Lorenz
// system so the GPU test can compare the result with the oracle bit for bit.
// It also keeps a dense output (32 points, at least 0.5 time units apart) and
// writes the reference's Print files (ActualState, TimeDomain, ControlParameters,
// ActualTime, and DenseOutput for instances 0 and 1234) into the working
// directory, so the test can compare their text with the oracle's values.
//   usage: lorenz_dropin <inputs.bin> <outputs.bin>
// inputs.bin : rho[NT], x[3][NT] (float64);  outputs.bin: x[3][NT], t[NT], accepted[NT] (float64)
#include <cstdio>
#include <vector>

#include "systems/lorenz.h"

template <class Precision>
__forceinline__ __device__ void PerThread_OdeFunction(int tid, int NT, Precision* F, Precision* X, Precision T,
                                                      Precision* cPAR, Precision* sPAR, int* sPARi, Precision* ACC, int* ACCi)
{
    mpgos::systems::Lorenz::OdeFunction<Precision>(tid, NT, F, X, T, cPAR, sPAR, sPARi, ACC, ACCi);
}
template <class Precision>
__forceinline__ __device__ void PerThread_EventFunction(int tid, int NT, Precision* EF, Precision T, Precision dT, Precision* TD,
                                                        Precision* X, Precision* cPAR, Precision* sPAR, int* sPARi, Precision* ACC, int* ACCi) {}
template <class Precision>
__forceinline__ __device__ void PerThread_ActionAfterEventDetection(int tid, int NT, int IDX, int& UDT, Precision& T, Precision& dT,
                                                                    Precision* TD, Precision* X, Precision* cPAR, Precision* sPAR,
                                                                    int* sPARi, Precision* ACC, int* ACCi) {}
template <class Precision>
__forceinline__ __device__ void PerThread_ActionAfterSuccessfulTimeStep(int tid, int NT, int& UDT, Precision& T, Precision& dT,
                                                                        Precision* TD, Precision* X, Precision* cPAR, Precision* sPAR,
                                                                        int* sPARi, Precision* ACC, int* ACCi) {}
template <class Precision>
__forceinline__ __device__ void PerThread_Initialization(int tid, int NT, int& DOIDX, Precision& T, Precision& dT, Precision* TD,
                                                         Precision* X, Precision* cPAR, Precision* sPAR, int* sPARi, Precision* ACC, int* ACCi) {}
template <class Precision>
__forceinline__ __device__ void PerThread_Finalization(int tid, int NT, int& DOIDX, Precision& T, Precision& dT, Precision* TD,
                                                       Precision* X, Precision* cPAR, Precision* sPAR, int* sPARi, Precision* ACC, int* ACCi) {}

#include "SingleSystem_PerThread_Interface.cuh"

#define SOLVER RKCK45
#define PRECISION double
const int NT = 4096;
const int SD = 3;
const int NCP = 1;
const int NSP = 0;
const int NISP = 0;
const int NE = 0;
const int NA = 0;
const int NIA = 0;
const int NDO = 32;

int main(int argc, char** argv)
{
    if (argc < 3) { std::fprintf(stderr, "usage: %s inputs.bin outputs.bin\n", argv[0]); return 2; }
    std::vector<double> in(4 * NT);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(in.data(), sizeof(double), in.size(), f) != in.size()) { std::fprintf(stderr, "bad input\n"); return 2; }
    std::fclose(f);

    ProblemSolver<NT, SD, NCP, NSP, NISP, NE, NA, NIA, NDO, SOLVER, PRECISION> ScanLorenz(SelectDeviceByClosestRevision(9, 5));
    ScanLorenz.SolverOption(ThreadsPerBlock, 64);
    ScanLorenz.SolverOption(DenseOutputMinimumTimeStep, 0.5);
    ScanLorenz.SolverOption(DenseOutputSaveFrequency, 1);
    for (int i = 0; i < SD; i++) {
        ScanLorenz.SolverOption(RelativeTolerance, i, 1e-10);
        ScanLorenz.SolverOption(AbsoluteTolerance, i, 1e-10);
    }
    for (int tid = 0; tid < NT; tid++) {
        ScanLorenz.SetHost(tid, TimeDomain, 0, 0.0);
        ScanLorenz.SetHost(tid, TimeDomain, 1, 10.0);
        ScanLorenz.SetHost(tid, ActualTime, 0.0);
        ScanLorenz.SetHost(tid, ControlParameters, 0, in[tid]);
        for (int c = 0; c < SD; c++) ScanLorenz.SetHost(tid, ActualState, c, in[NT + c * NT + tid]);
        ScanLorenz.SetHost(tid, DenseIndex, 0);
    }
    ScanLorenz.SynchroniseFromHostToDevice(All);
    ScanLorenz.Solve();
    ScanLorenz.SynchroniseFromDeviceToHost(All);
    ScanLorenz.SynchroniseFromDeviceToHost(DenseOutput);
    ScanLorenz.InsertSynchronisationPoint();
    ScanLorenz.SynchroniseSolver();

    ScanLorenz.Print(ActualState);
    ScanLorenz.Print(TimeDomain);
    ScanLorenz.Print(ControlParameters);
    ScanLorenz.Print(ActualTime);
    ScanLorenz.Print(DenseOutput, 0);
    ScanLorenz.Print(DenseOutput, 1234);

    std::vector<double> out(5 * NT);
    for (int tid = 0; tid < NT; tid++) {
        for (int c = 0; c < SD; c++) out[c * NT + tid] = ScanLorenz.GetHost<PRECISION>(tid, ActualState, c);
        out[3 * NT + tid] = ScanLorenz.GetHost<PRECISION>(tid, ActualTime);
        double acc = 0;
        ScanLorenz.Native().get_host(MPGOS_PT_ACCEPTED_STEPS, tid, 0, 0, &acc);
        out[4 * NT + tid] = acc;
    }
    f = std::fopen(argv[2], "wb");
    if (!f || std::fwrite(out.data(), sizeof(double), out.size(), f) != out.size()) return 3;
    std::fclose(f);
    std::printf("Test finished!\n");
    return 0;
}
