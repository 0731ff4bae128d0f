// Keller-Miksis bubble under dual-frequency (quasi-periodic) forcing, the
// system of the reference's quasi-periodic-forcing tutorial, with its
// callbacks: user termination at the first event, minimum tracking.
// Same expressions and callback behaviour as
// Tutorials_SingleSystem_PerThread/T4_QuasiperiodicForcing/QuasiperiodicForcing_SystemDefinition.cuh:7-94
// (dimensionless radius y1 = R / R_E, time scaled by the first forcing period):
//   y1' = y2,  y2' = N / D,
//   N = (c0 + c1 y2) y1^-c10 - c2 (1 + c9 y2) - c3 / y1 - c4 y2 / y1
//       - 1.5 (1 - c9 y2 / 3) y2^2 - (c5 sin 2 pi t + c6 sin(2 pi c11 t + c12)) (1 + c9 y2)
//       - y1 (c7 cos 2 pi t + c8 cos(2 pi c11 t + c12)),
//   D = y1 - c9 y1 y2 + c4 c9;
// cPAR = (c0 .. c12, then reference-only quantities 13 .. 20); event y2 = 0
// (a local extremum); the first one terminates the instance (ACCi[0] counts
// events); ACC = (y1, t at the start, minimum y1 and its time).  The
// reference evaluates sin 2 pi t, cos 2 pi t with sincospi(2t); here with the
// library's sin and cos of 2 pi t (bit-reproducible on the host).
#pragma once
#include "system_base.h"

namespace mpgos {
namespace systems {

struct KellerMiksis : NoCallbacks {
    static constexpr const char* name = "keller_miksis";
    static constexpr int SD = 2, NCP = 21, NSP = 0, NISP = 0, NE = 1, NA = 4, NIA = 1;
    static constexpr bool HAS_VJP = false;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P T, P* c, P*, int*, P*, int*)
    {
        using mpgos::math::cos;
        using mpgos::math::pow;
        using mpgos::math::sin;
        const P pi = P(3.14159265358979323846);
        const P rx1 = P(1) / X[0];
        const P p = pow(rx1, c[10]);
        const P a1 = P(2) * pi * T;
        const P s1 = sin(a1), c1 = cos(a1);
        const P a2 = P(2) * c[11] * pi * T + c[12];
        const P s2 = sin(a2), c2 = cos(a2);
        const P y2 = X[1];
        const P N = (c[0] + c[1] * y2) * p - c[2] * (P(1) + c[9] * y2) - c[3] * rx1 - c[4] * y2 * rx1
                    - P(1.5) * (P(1) - c[9] * y2 * (P(1) / P(3))) * y2 * y2
                    - (c[5] * s1 + c[6] * s2) * (P(1) + c[9] * y2) - X[0] * (c[7] * c1 + c[8] * c2);
        const P D = X[0] - c[9] * X[0] * y2 + c[4] * c[9];
        F[0] = y2;
        F[1] = N * (P(1) / D);
    }

    template <class P>
    MPGOS_HD static void EventFunction(int, int, P* EF, P, P, P*, P* X, P*, P*, int*, P*, int*)
    {
        EF[0] = X[1];
    }

    template <class P>
    MPGOS_HD static void ActionAfterEventDetection(int, int, int, int& UDT, P&, P&, P*, P*, P*, P*, int*, P*, int* ACCi)
    {
        ACCi[0]++;
        if (ACCi[0] == 1) UDT = 1;
    }

    template <class P>
    MPGOS_HD static void ActionAfterSuccessfulTimeStep(int, int, int&, P& T, P&, P*, P* X, P*, P*, int*, P* ACC, int*)
    {
        if (X[0] < ACC[2]) { ACC[2] = X[0]; ACC[3] = T; }
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int, int&, P& T, P&, P*, P* X, P*, P*, int*, P* ACC, int* ACCi)
    {
        ACC[0] = X[0];
        ACC[1] = T;
        ACC[2] = X[0];
        ACC[3] = T;
        ACCi[0] = 0;
    }
};

} // namespace systems
} // namespace mpgos

namespace mpgos {
namespace systems {

// The same bubble under one forcing frequency, as the reference's large time
// scale differences tutorial runs it
// (Tutorials_SingleSystem_PerThread/T5_LargeTimeScaleDifferences/KellerMiksis_SystemDefinition.cuh:7-94):
// 13 control parameters (c6 = c8 = c11 = 0), no events, every Solve restarts
// the time at TimeDomain[0] (one forcing period per Solve) and ACC[0] keeps
// the largest radius of the Solve.
struct KellerMiksisMax : NoCallbacks {
    static constexpr const char* name = "keller_miksis_max";
    static constexpr int SD = 2, NCP = 13, NSP = 0, NISP = 0, NE = 0, NA = 1, NIA = 0;
    static constexpr bool HAS_VJP = false;

    template <class P>
    MPGOS_HD static void OdeFunction(int tid, int NT, P* F, P* X, P T, P* c, P* s, int* si, P* a, int* ai)
    {
        KellerMiksis::OdeFunction<P>(tid, NT, F, X, T, c, s, si, a, ai);
    }

    template <class P>
    MPGOS_HD static void ActionAfterSuccessfulTimeStep(int, int, int&, P&, P&, P*, P* X, P*, P*, int*, P* ACC, int*)
    {
        if (X[0] > ACC[0]) ACC[0] = X[0];
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int, int&, P& T, P&, P* TD, P* X, P*, P*, int*, P* ACC, int*)
    {
        T = TD[0];
        ACC[0] = X[0];
    }
};

} // namespace systems
} // namespace mpgos
