// Lorenz system  x' = 10 (y - x),  y' = rho x - y - x z,  z' = x y - (8/3) z,
// cPAR = (rho).  Workload of the adaptive RKCK45 ensemble (262144 instances,
// rho ~ U[0, 21]).  Follows Tutorials_SingleSystem_PerThread/T2_Lorenz/
// Lorenz_SystemDefinition.cuh:5-14, with b = 8/3 instead of the tutorial's
// literal 2.666.
#pragma once
#include "system_base.h"

namespace mpgos {
namespace systems {

struct Lorenz : NoCallbacks {
    static constexpr const char* name = "lorenz";
    static constexpr int SD = 3, NCP = 1, NSP = 0, NISP = 0, NE = 0, NA = 0, NIA = 0;
    static constexpr bool HAS_VJP = true;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P, P* cPAR, P*, int*, P*, int*)
    {
        using mpgos::math::fma;
        const P x = X[0], y = X[1], z = X[2];
        F[0] = P(10.0) * (y - x);
        F[1] = fma(cPAR[0], x, -y) - x * z;
        F[2] = fma(x, y, -P(8.0 / 3.0) * z);
    }

    template <class P>
    MPGOS_HD static void OdeVJP(int, int, const P* lam, P* X, P, P* cPAR, P*, int*, P*, int*, P* gX, P* gP)
    {
        const P x = X[0], y = X[1], z = X[2];
        gX[0] = -P(10.0) * lam[0] + (cPAR[0] - z) * lam[1] + y * lam[2];
        gX[1] = P(10.0) * lam[0] - lam[1] + x * lam[2];
        gX[2] = -x * lam[1] - P(8.0 / 3.0) * lam[2];
        gP[0] = x * lam[1];
    }
};

} // namespace systems
} // namespace mpgos
