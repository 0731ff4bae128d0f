// Per-block coupled systems (one system = UPS coupled units, unit dimension UD).
//
// A per-block system is a struct with the compile-time sizes of the
// reference's per-block ProblemSolver (UD, NC, NUP, NSP, NGP, NiGP, NUA, NiUA,
// NSA, NiSA, NE; SourceCodes/CoupledSystems_PerBlock_Interface.cuh:135) and the
// six per-block callbacks with the reference's argument lists
// (Tutorials_CoupledSystems_PerBlock/T1_RingOfDuffingOscillators/
// RingOfDuffingOscillators_SystemDefinition.cuh:7-104).  T, dT, DOIDX, TD and
// the system-scope arrays are shared by the units of a system: writes to them
// are guarded by uid == 0 in user code, as in the reference tutorials.
#pragma once
#include "system_base.h"

namespace mpgos {
namespace systems {

struct PBNoCallbacks {
#define MPGOS_PB_ARGS P*, P*, P*, int*, P*, int*, P*, int*
    template <class P> MPGOS_HD static void EventFunction(int, int, P*, P, P, P*, P*, MPGOS_PB_ARGS) {}
    template <class P> MPGOS_HD static void ActionAfterEventDetection(int, int, int, int&, P&, P&, P*, P*, MPGOS_PB_ARGS) {}
    template <class P> MPGOS_HD static void ActionAfterSuccessfulTimeStep(int, int, int&, P&, P&, P*, P*, MPGOS_PB_ARGS) {}
    template <class P> MPGOS_HD static void Finalization(int, int, int&, P&, P&, P*, P*, MPGOS_PB_ARGS) {}
#undef MPGOS_PB_ARGS
};

// Ring of forced hardening Duffing units with a dense coupling matrix on the
// velocity equation:  x_u'' + k_u x_u' + x_u + x_u^3 = B cos t + sum_v C_uv x_v.
// uPAR = (k), sPAR = (B); coupling 0 takes CPT = x (coupling terms) with
// CPF = 1 and is added to F[CouplingIndex[0]] (= 1) scaled by the system's
// coupling strength.  Workload of the per-block configuration (U = 256 units,
// 4096 systems, C_uv ~ N(0, 0.05/U), zero diagonal).
struct RingDuffing : PBNoCallbacks {
    static constexpr const char* name = "ring_duffing";
    static constexpr int UD = 2, NC = 1, NUP = 1, NSP = 1, NGP = 0, NiGP = 0, NUA = 0, NiUA = 0, NSA = 0, NiSA = 0, NE = 0;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P T, P* uPAR, P* sPAR, P*, int*, P*, int*, P*, int*, P* CPT,
                                     P* CPF)
    {
        using mpgos::math::cos;
        using mpgos::math::fma;
        const P x = X[0], v = X[1];
        F[0] = v;
        F[1] = fma(-uPAR[0], v, sPAR[0] * cos(T)) - fma(x * x, x, x);
        CPT[0] = x;
        CPF[0] = P(1.0);
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int uid, int&, P& T, P&, P* TD, P*, P*, P*, P*, int*, P*, int*, P*, int*)
    {
        if (uid == 0) T = TD[0];
    }
};

// RingDuffing with a quadrature state q' = x^2 per unit (UD = 3), for losses
// of the form mean over a window of x^2; with the vector-Jacobian product
// that the discrete adjoint needs (gradients with respect to the state, the
// unit parameter k and the system parameter B).
struct RingDuffingQ : PBNoCallbacks {
    static constexpr const char* name = "ring_duffing_q";
    static constexpr int UD = 3, NC = 1, NUP = 1, NSP = 1, NGP = 0, NiGP = 0, NUA = 0, NiUA = 0, NSA = 0, NiSA = 0, NE = 0;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P T, P* uPAR, P* sPAR, P*, int*, P*, int*, P*, int*, P* CPT,
                                     P* CPF)
    {
        using mpgos::math::cos;
        using mpgos::math::fma;
        const P x = X[0], v = X[1];
        F[0] = v;
        F[1] = fma(-uPAR[0], v, sPAR[0] * cos(T)) - fma(x * x, x, x);
        F[2] = x * x;
        CPT[0] = x;
        CPF[0] = P(1.0);
    }

    template <class P>
    MPGOS_HD static void OdeVJP(int, int, const P* kbF, const P* kbCPT, const P*, P* X, P T, P* uPAR, P*, P*, int*, P*, int*,
                                P*, int*, P* gX, P* gU, P* gS)
    {
        using mpgos::math::cos;
        const P x = X[0], v = X[1];
        gX[0] = -(P(1) + P(3) * x * x) * kbF[1] + P(2) * x * kbF[2] + kbCPT[0];
        gX[1] = kbF[0] - uPAR[0] * kbF[1];
        gX[2] = P(0);
        gU[0] = -v * kbF[1];
        gS[0] = cos(T) * kbF[1];
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int uid, int&, P& T, P&, P* TD, P*, P*, P*, P*, int*, P*, int*, P*, int*)
    {
        if (uid == 0) T = TD[0];
    }
};

// The reference's per-block tutorial system (double-well Duffing ring, two
// couplings, unit events and user termination), same expressions and callback
// behaviour as RingOfDuffingOscillators_SystemDefinition.cuh:7-104:
// uPAR = (omega, k, B), gPAR = (damping of the ring), NE = 1 (x' = 0),
// uACC[2] = x at the second event, iuACC[0] counts events, the second event
// of any unit terminates its system; sACC = (min dT, max dT), isACC = steps.
struct RingDW {
    static constexpr const char* name = "ring_dw";
    static constexpr int UD = 2, NC = 2, NUP = 3, NSP = 0, NGP = 1, NiGP = 0, NUA = 3, NiUA = 1, NSA = 2, NiSA = 1, NE = 1;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P T, P* uPAR, P*, P* gPAR, int*, P*, int*, P*, int*, P* CPT,
                                     P* CPF)
    {
        using mpgos::math::cos;
        F[0] = X[1] - gPAR[0] * X[0];
        F[1] = -uPAR[1] * X[1] + X[0] - X[0] * X[0] * X[0] + uPAR[2] * cos(uPAR[0] * T) - gPAR[0] * X[1];
        CPT[0] = X[0];
        CPF[0] = P(1.0);
        CPT[1] = X[1];
        CPF[1] = P(1.0);
    }

    template <class P>
    MPGOS_HD static void EventFunction(int, int, P* EF, P, P, P*, P* X, P*, P*, P*, int*, P*, int*, P*, int*)
    {
        EF[0] = X[1];
    }

    template <class P>
    MPGOS_HD static void ActionAfterEventDetection(int, int, int, int& UDT, P&, P&, P*, P* X, P*, P*, P*, int*, P* uACC,
                                                   int* iuACC, P*, int*)
    {
        iuACC[0]++;
        if (iuACC[0] == 2) {
            UDT = 1;
            uACC[2] = X[0];
        }
    }

    template <class P>
    MPGOS_HD static void ActionAfterSuccessfulTimeStep(int, int uid, int&, P&, P& dT, P*, P*, P*, P*, P*, int*, P*, int*,
                                                       P* sACC, int* isACC)
    {
        if (uid == 0) {
            if (sACC[0] > dT) sACC[0] = dT;
            if (sACC[1] < dT) sACC[1] = dT;
            isACC[0]++;
        }
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int uid, int& DOIDX, P& T, P& dT, P* TD, P* X, P*, P*, P*, int*, P* uACC,
                                        int* iuACC, P* sACC, int* isACC)
    {
        if (uid == 0) {
            sACC[0] = dT;
            sACC[1] = dT;
            isACC[0] = 0;
            T = TD[0];
            DOIDX = 0;
        }
        uACC[2] = X[0];
        iuACC[0] = 0;
    }

    template <class P>
    MPGOS_HD static void Finalization(int, int, int&, P&, P&, P*, P*, P*, P*, P*, int*, P*, int*, P*, int*) {}
};

} // namespace systems
} // namespace mpgos

namespace mpgos {
namespace systems {

// Linear test system x_u' = a_u x_u + s sum_v M_uv x_v (UD = 1, uPAR = (a)):
// the whole system is x' = (diag(a) + s M) x, so a fixed-step integration is
// a closed-form matrix polynomial for every coupling storage form.
struct PBLinear : PBNoCallbacks {
    static constexpr const char* name = "pb_linear";
    static constexpr int UD = 1, NC = 1, NUP = 1, NSP = 0, NGP = 0, NiGP = 0, NUA = 0, NiUA = 0, NSA = 0, NiSA = 0, NE = 0;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P, P* uPAR, P*, P*, int*, P*, int*, P*, int*, P* CPT, P* CPF)
    {
        F[0] = uPAR[0] * X[0];
        CPT[0] = X[0];
        CPF[0] = P(1.0);
    }

    // vector-Jacobian product: adjoints kbF of F, kbCPT of CPT and kbCPF of CPF
    // to adjoints of X, of the unit parameters and of the system parameters
    template <class P>
    MPGOS_HD static void OdeVJP(int, int, const P* kbF, const P* kbCPT, const P*, P* X, P, P* uPAR, P*, P*, int*, P*, int*,
                                P*, int*, P* gX, P* gU, P*)
    {
        gX[0] = uPAR[0] * kbF[0] + kbCPT[0];
        gU[0] = X[0] * kbF[0];
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int uid, int&, P& T, P&, P* TD, P*, P*, P*, P*, int*, P*, int*, P*, int*)
    {
        if (uid == 0) T = TD[0];
    }
};

} // namespace systems
} // namespace mpgos
