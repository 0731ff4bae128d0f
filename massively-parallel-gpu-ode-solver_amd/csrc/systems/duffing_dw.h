// The double-well Duffing system of the reference's first tutorial, with its
// full callback set: two event functions (x' = 0 and x = 0), three accessories
// and one integer accessory, dense output.  Same expressions and callback
// behaviour as Tutorials_SingleSystem_PerThread/T1_Reference/
// Reference_SystemDefinition.cuh:5-76: cPAR = (k), sPAR = (B),
//   F0 = x', F1 = x - x^3 - k x' + B cos t,
//   EF = (x', x); at detection ACC[0] = max x, events of EF[1] counted in
//   ACCi[0] and ACC[1] = x' at the second one; ACC[2] = max x over accepted
//   steps; Initialization resets T = TD[0] and the dense-output index.
// A regression fixture for the callback surface (events, accessories, shared
// parameters, dense output).
#pragma once
#include "system_base.h"

namespace mpgos {
namespace systems {

struct DuffingDW {
    static constexpr const char* name = "duffing_dw";
    static constexpr int SD = 2, NCP = 1, NSP = 1, NISP = 0, NE = 2, NA = 3, NIA = 1;
    static constexpr bool HAS_VJP = false;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P T, P* cPAR, P* sPAR, int*, P*, int*)
    {
        using mpgos::math::cos;
        F[0] = X[1];
        F[1] = X[0] - X[0] * X[0] * X[0] - cPAR[0] * X[1] + sPAR[0] * cos(T);
    }

    template <class P>
    MPGOS_HD static void EventFunction(int, int, P* EF, P, P, P*, P* X, P*, P*, int*, P*, int*)
    {
        EF[0] = X[1];
        EF[1] = X[0];
    }

    template <class P>
    MPGOS_HD static void ActionAfterEventDetection(int, int, int IDX, int&, P&, P&, P*, P* X, P*, P*, int*, P* ACC, int* ACCi)
    {
        if (X[0] > ACC[0]) ACC[0] = X[0];
        if (IDX == 1) ACCi[0]++;
        if ((IDX == 1) && (ACCi[0] == 2)) ACC[1] = X[1];
    }

    template <class P>
    MPGOS_HD static void ActionAfterSuccessfulTimeStep(int, int, int&, P&, P&, P*, P* X, P*, P*, int*, P* ACC, int*)
    {
        if (X[0] > ACC[2]) ACC[2] = X[0];
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int, int& DOIDX, P& T, P&, P* TD, P* X, P*, P*, int*, P* ACC, int* ACCi)
    {
        T = TD[0];
        DOIDX = 0;
        ACC[0] = X[0];
        ACC[1] = X[1];
        ACC[2] = X[0];
        ACCi[0] = 0;
    }

    template <class P>
    MPGOS_HD static void Finalization(int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int*) {}
};

} // namespace systems
} // namespace mpgos
