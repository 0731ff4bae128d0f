// Affine test system F = A X + b (2-dim), cPAR = (a00, a01, a10, a11, b0, b1),
// one event function EF = x0 - sPAR[0] whose detections are counted in
// ACCi[0].  Its solutions are known in closed form (exponential, polynomial,
// harmonic), which pins the steppers, the controller, the event step control
// and the adjoint against mathematics rather than against another program.
#pragma once
#include "system_base.h"

namespace mpgos {
namespace systems {

struct Affine2 : NoCallbacks {
    static constexpr const char* name = "affine2";
    static constexpr int SD = 2, NCP = 6, NSP = 1, NISP = 0, NE = 1, NA = 0, NIA = 1;
    static constexpr bool HAS_VJP = true;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P, P* cPAR, P*, int*, P*, int*)
    {
        using mpgos::math::fma;
        F[0] = fma(cPAR[0], X[0], fma(cPAR[1], X[1], cPAR[4]));
        F[1] = fma(cPAR[2], X[0], fma(cPAR[3], X[1], cPAR[5]));
    }

    template <class P>
    MPGOS_HD static void OdeVJP(int, int, const P* lam, P* X, P, P* cPAR, P*, int*, P*, int*, P* gX, P* gP)
    {
        gX[0] = cPAR[0] * lam[0] + cPAR[2] * lam[1];
        gX[1] = cPAR[1] * lam[0] + cPAR[3] * lam[1];
        gP[0] = X[0] * lam[0];
        gP[1] = X[1] * lam[0];
        gP[2] = X[0] * lam[1];
        gP[3] = X[1] * lam[1];
        gP[4] = lam[0];
        gP[5] = lam[1];
    }

    template <class P>
    MPGOS_HD static void EventFunction(int, int, P* EF, P, P, P*, P* X, P*, P* sPAR, int*, P*, int*)
    {
        EF[0] = X[0] - sPAR[0];
    }

    template <class P>
    MPGOS_HD static void ActionAfterEventDetection(int, int, int, int&, P&, P&, P*, P*, P*, P*, int*, P*, int* ACCi)
    {
        ACCi[0]++;
    }
};

} // namespace systems
} // namespace mpgos
