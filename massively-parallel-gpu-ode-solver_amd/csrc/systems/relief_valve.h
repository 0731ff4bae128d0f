// The pressure relief valve of the reference's impact-dynamics tutorial, with
// its full callback set: a state jump at impact and two user terminations.
// Same expressions and callback behaviour as
// Tutorials_SingleSystem_PerThread/T6_ImpactDynamics/ImpactDynamics_SystemDefinition.cuh:7-84:
//   x' = v,  v' = -kappa v - (x + delta) + p,  p' = beta (q - x sqrt(p)),
// sPAR = (kappa, delta, beta, r), cPAR = (q); event functions EF = (v, x);
// at x = 0 (impact) v -> -r v; the first v = 0 crossing (counted in ACCi[0])
// terminates the instance; ACC = (min x, max x, last "moving" x); ACCi[1]
// counts accepted steps with |x - ACC[2]| < 1e-6 and 50 of them terminate
// (the valve rests on its seat).  A regression fixture for the callback
// surface: state changes in ActionAfterEventDetection, user termination from
// both actions, integer accessories.
#pragma once
#include "system_base.h"

namespace mpgos {
namespace systems {

struct ReliefValve : NoCallbacks {
    static constexpr const char* name = "relief_valve";
    static constexpr int SD = 3, NCP = 1, NSP = 4, NISP = 0, NE = 2, NA = 3, NIA = 2;
    static constexpr bool HAS_VJP = false;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P, P* cPAR, P* sPAR, int*, P*, int*)
    {
        using mpgos::math::sqrt;
        F[0] = X[1];
        F[1] = -sPAR[0] * X[1] - (X[0] + sPAR[1]) + X[2];
        F[2] = sPAR[2] * (cPAR[0] - X[0] * sqrt(X[2]));
    }

    template <class P>
    MPGOS_HD static void EventFunction(int, int, P* EF, P, P, P*, P* X, P*, P*, int*, P*, int*)
    {
        EF[0] = X[1];    // Poincare section v = 0
        EF[1] = X[0];    // impact x = 0
    }

    template <class P>
    MPGOS_HD static void ActionAfterEventDetection(int, int, int IDX, int& UDT, P&, P&, P*, P* X, P*, P* sPAR, int*, P*,
                                                   int* ACCi)
    {
        if (IDX == 0) ACCi[0]++;
        if (ACCi[0] == 1 && IDX == 0) UDT = 1;
        if (IDX == 1) X[1] = -sPAR[3] * X[1];
    }

    template <class P>
    MPGOS_HD static void ActionAfterSuccessfulTimeStep(int, int, int& UDT, P&, P&, P*, P* X, P*, P*, int*, P* ACC, int* ACCi)
    {
        using mpgos::math::fabs;
        if (X[0] < ACC[0]) ACC[0] = X[0];
        if (X[0] > ACC[1]) ACC[1] = X[0];
        if (fabs(X[0] - ACC[2]) < P(1e-6)) ACCi[1]++;
        else ACC[2] = X[0];
        if (ACCi[1] == 50) UDT = 1;
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int, int&, P& T, P&, P* TD, P* X, P*, P*, int*, P* ACC, int* ACCi)
    {
        T = TD[0];
        ACC[0] = X[0];
        ACC[1] = X[0];
        ACCi[0] = 0;
        ACC[2] = X[0];
        ACCi[1] = 0;
    }
};

} // namespace systems
} // namespace mpgos
