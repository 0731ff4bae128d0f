// A driver written against the reference's per-block interface (the shape of
// Tutorials_CoupledSystems_PerBlock/T1_RingOfDuffingOscillators/RingOfDuffingOscillators.cu),
// compiled unchanged against the drop-in header.  Its callbacks forward to the
// registered double-well ring system (the tutorial's system definition with
// this library's cos), so the GPU test can compare with the oracle bit for bit.
//   usage: ring_dropin <outputs.bin>   (run in a scratch directory: Print writes files there)
// outputs.bin: x[2][NS*UPS], t[NS], uACC[3][NS*UPS], iuACC[NS*UPS], sACC[2][NS], isACC[NS], DOidx[NS] (float64)
#include <cstdio>
#include <vector>

#include "systems/pb_ring.h"

using mpgos::systems::RingDW;

template <class Precision>
__forceinline__ __device__ void CoupledSystems_PerBlock_OdeFunction(int sid, int uid, Precision* F, Precision* X, Precision T,
    Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC, int* iuACC, Precision* sACC, int* isACC,
    Precision* CPT, Precision* CPF)
{ RingDW::OdeFunction<Precision>(sid, uid, F, X, T, uPAR, sPAR, gPAR, igPAR, uACC, iuACC, sACC, isACC, CPT, CPF); }

template <class Precision>
__forceinline__ __device__ void CoupledSystems_PerBlock_EventFunction(int sid, int uid, Precision* EF, Precision T, Precision dT,
    Precision* TD, Precision* X, Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC, int* iuACC,
    Precision* sACC, int* isACC)
{ RingDW::EventFunction<Precision>(sid, uid, EF, T, dT, TD, X, uPAR, sPAR, gPAR, igPAR, uACC, iuACC, sACC, isACC); }

template <class Precision>
__forceinline__ __device__ void CoupledSystems_PerBlock_ActionAfterEventDetection(int sid, int uid, int IDX, int& UDT,
    Precision& T, Precision& dT, Precision* TD, Precision* X, Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR,
    Precision* uACC, int* iuACC, Precision* sACC, int* isACC)
{ RingDW::ActionAfterEventDetection<Precision>(sid, uid, IDX, UDT, T, dT, TD, X, uPAR, sPAR, gPAR, igPAR, uACC, iuACC, sACC, isACC); }

template <class Precision>
__forceinline__ __device__ void CoupledSystems_PerBlock_ActionAfterSuccessfulTimeStep(int sid, int uid, int& UDT, Precision& T,
    Precision& dT, Precision* TD, Precision* X, Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC,
    int* iuACC, Precision* sACC, int* isACC)
{ RingDW::ActionAfterSuccessfulTimeStep<Precision>(sid, uid, UDT, T, dT, TD, X, uPAR, sPAR, gPAR, igPAR, uACC, iuACC, sACC, isACC); }

template <class Precision>
__forceinline__ __device__ void CoupledSystems_PerBlock_Initialization(int sid, int uid, int& DOIDX, Precision& T, Precision& dT,
    Precision* TD, Precision* X, Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC, int* iuACC,
    Precision* sACC, int* isACC)
{ RingDW::Initialization<Precision>(sid, uid, DOIDX, T, dT, TD, X, uPAR, sPAR, gPAR, igPAR, uACC, iuACC, sACC, isACC); }

template <class Precision>
__forceinline__ __device__ void CoupledSystems_PerBlock_Finalization(int sid, int uid, int& DOIDX, Precision& T, Precision& dT,
    Precision* TD, Precision* X, Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC, int* iuACC,
    Precision* sACC, int* isACC)
{ RingDW::Finalization<Precision>(sid, uid, DOIDX, T, dT, TD, X, uPAR, sPAR, gPAR, igPAR, uACC, iuACC, sACC, isACC); }

#include "CoupledSystems_PerBlock_Interface.cuh"

#define PI 3.14159265358979323846
const int NumberOfInitialConditionsX = 5;
const int NumberOfInitialConditionsY = 5;
const int NumberOfUnitsPerSystem = 14;
#define TIMESTEP 1.0e-2
#define TOLERANCE 1.0e-8
#define PERIOD 10.0
#define SOLVER RKCK45
#define PRECISION double
const int NS = NumberOfInitialConditionsX * NumberOfInitialConditionsY;
const int UPS = NumberOfUnitsPerSystem;
const int UD = 2, TPB = 32, SPB = 3, NC = 2, CBW = 1, CCI = 1, NUP = 3, NSP = 0, NGP = 1, NiGP = 0, NUA = 3, NiUA = 1,
          NSA = 2, NiSA = 1, NE = 1, NDO = 500;
typedef ProblemSolver<NS, UPS, UD, TPB, SPB, NC, CBW, CCI, NUP, NSP, NGP, NiGP, NUA, NiUA, NSA, NiSA, NE, NDO, SOLVER, PRECISION> Solver;

void Linspace(std::vector<PRECISION>& x, PRECISION B, PRECISION E, int N)
{
    x[0] = B;
    if (N > 1) {
        x[N - 1] = E;
        PRECISION inc = (E - B) / (N - 1);
        for (int i = 1; i < N - 1; i++) x[i] = B + i * inc;
    }
}

void FillSolverObject(Solver& S, const std::vector<PRECISION>& ICX, const std::vector<PRECISION>& ICY)
{
    PRECISION CPS = 0.004 / (2.0 * (double)CBW);
    int sid = 0;
    for (auto const& X0 : ICX)
        for (auto const& Y0 : ICY) {
            S.SetHost(sid, TimeDomain, 0, 0.0);
            S.SetHost(sid, TimeDomain, 1, PERIOD * 2.0 * PI / 0.5);
            for (int i = 0; i < NC; i++) S.SetHost(sid, CouplingStrength, i, CPS);
            S.SetHost(sid, SystemAccessories, 0, TIMESTEP);
            S.SetHost(sid, SystemAccessories, 1, TIMESTEP);
            for (int u = 0; u < UPS; u++) {
                PRECISION a = u == 11 ? X0 : -5.0, b = u == 11 ? Y0 : -5.0;
                S.SetHost(sid, u, ActualState, 0, a);
                S.SetHost(sid, u, ActualState, 1, b);
                S.SetHost(sid, u, UnitAccessories, 0, a);
                S.SetHost(sid, u, UnitAccessories, 1, b);
                S.SetHost(sid, u, UnitParameters, 0, 0.5);
                S.SetHost(sid, u, UnitParameters, 1, 0.24);
                S.SetHost(sid, u, UnitParameters, 2, 13.633);
            }
            sid++;
        }
    for (int i = 0; i < NC; i++) S.SetHost(CouplingIndex, i, i);
    S.SetHost(GlobalParameters, 0, 0.004);
}

void FillCouplingMatrix(Solver& S)
{
    for (int Row = 0; Row < UPS; Row++)
        for (int Diag = -CBW; Diag <= CBW; Diag++) {
            int Col = Row + Diag;
            if (Col >= UPS) Col -= UPS;
            if (Col < 0) Col += UPS;
            S.SetHost(0, CouplingMatrix, Row, Col, Diag != 0 ? 1.0 : 0.0);
            S.SetHost(1, CouplingMatrix, Row, Col, Diag != 0 ? 1.0 : 0.0);
        }
}

int main(int argc, char** argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: %s outputs.bin\n", argv[0]); return 2; }
    std::vector<PRECISION> ICX(NumberOfInitialConditionsX), ICY(NumberOfInitialConditionsY);
    Linspace(ICX, -8.0, 8.0, NumberOfInitialConditionsX);
    Linspace(ICY, -20.0, 40.0, NumberOfInitialConditionsY);
    ListCUDADevices();
    int SelectedDevice = SelectDeviceByClosestRevision(3, 5);
    PrintPropertiesOfSpecificDevice(SelectedDevice);

    Solver ScanSystem(SelectedDevice);
    ScanSystem.SolverOption(InitialTimeStep, TIMESTEP);
    ScanSystem.SolverOption(RelativeTolerance, 0, TOLERANCE);
    ScanSystem.SolverOption(AbsoluteTolerance, 0, TOLERANCE);
    ScanSystem.SolverOption(RelativeTolerance, 1, TOLERANCE);
    ScanSystem.SolverOption(AbsoluteTolerance, 1, TOLERANCE);
    ScanSystem.SolverOption(EventDirection, 0, -1);
    ScanSystem.SolverOption(DenseOutputMinimumTimeStep, 1e-6);
    ScanSystem.SolverOption(DenseOutputSaveFrequency, 1);
    ScanSystem.SolverOption(SharedGlobalVariables, 0);

    FillSolverObject(ScanSystem, ICX, ICY);
    FillCouplingMatrix(ScanSystem);
    ScanSystem.SynchroniseFromHostToDevice(All);
    ScanSystem.InsertSynchronisationPoint();
    ScanSystem.SynchroniseSolver();
    ScanSystem.Solve();
    ScanSystem.InsertSynchronisationPoint();
    ScanSystem.SynchroniseSolver();
    ScanSystem.SynchroniseFromDeviceToHost(All);
    ScanSystem.InsertSynchronisationPoint();
    ScanSystem.SynchroniseSolver();

    std::vector<double> out;
    for (int c = 0; c < UD; c++)
        for (int s = 0; s < NS; s++)
            for (int u = 0; u < UPS; u++) out.push_back(ScanSystem.GetHost<PRECISION>(s, u, ActualState, c));
    for (int s = 0; s < NS; s++) out.push_back(ScanSystem.GetHost<PRECISION>(s, ActualTime));
    for (int c = 0; c < NUA; c++)
        for (int s = 0; s < NS; s++)
            for (int u = 0; u < UPS; u++) out.push_back(ScanSystem.GetHost<PRECISION>(s, u, UnitAccessories, c));
    for (int s = 0; s < NS; s++)
        for (int u = 0; u < UPS; u++) out.push_back(ScanSystem.GetHost<int>(s, u, IntegerUnitAccessories, 0));
    for (int c = 0; c < NSA; c++)
        for (int s = 0; s < NS; s++) out.push_back(ScanSystem.GetHost<PRECISION>(s, SystemAccessories, c));
    for (int s = 0; s < NS; s++) out.push_back(ScanSystem.GetHost<int>(s, IntegerSystemAccessories, 0));
    for (int s = 0; s < NS; s++) out.push_back(ScanSystem.GetHost<int>(s, DenseIndex));
    FILE* f = std::fopen(argv[1], "wb");
    if (!f || std::fwrite(out.data(), sizeof(double), out.size(), f) != out.size()) { std::fprintf(stderr, "write failed\n"); return 2; }
    std::fclose(f);

    ScanSystem.Print(ActualState);
    ScanSystem.Print(TimeDomain);
    ScanSystem.Print(CouplingMatrix, 0);
    for (int s = 0; s < NS; s++) ScanSystem.Print(DenseOutput, s);
    return 0;
}
