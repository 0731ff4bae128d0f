// Forced hardening Duffing oscillator  x'' + k x' + x + x^3 = B cos(t)
// as a 2-dim first-order system, cPAR = (k, B); accessory ACC[0] = max |x|
// over the accepted steps of a Solve.  Workload of the fixed-step RK4 ensemble
// (16384 instances, B = linspace(0.3, 10)) and of the adjoint ensemble.
// Callback surface as in Tutorials_SingleSystem_PerThread/T1_Reference/
// Reference_SystemDefinition.cuh:5-76 (that tutorial integrates the
// double-well variant x'' + k x' - x + x^3 = B cos t; see duffing_dw.h).
#pragma once
#include "system_base.h"

namespace mpgos {
namespace systems {

struct Duffing : NoCallbacks {
    static constexpr const char* name = "duffing";
    static constexpr int SD = 2, NCP = 2, NSP = 0, NISP = 0, NE = 0, NA = 1, NIA = 0;
    static constexpr bool HAS_VJP = true;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P T, P* cPAR, P*, int*, P*, int*)
    {
        using mpgos::math::cos;
        using mpgos::math::fma;
        const P x = X[0], v = X[1];
        F[0] = v;
        F[1] = fma(-cPAR[0], v, cPAR[1] * cos(T)) - fma(x * x, x, x);
    }

    template <class P>
    MPGOS_HD static void OdeVJP(int, int, const P* lam, P* X, P T, P* cPAR, P*, int*, P*, int*, P* gX, P* gP)
    {
        using mpgos::math::cos;
        const P x = X[0], v = X[1];
        gX[0] = -(P(1) + P(3) * x * x) * lam[1];
        gX[1] = lam[0] - cPAR[0] * lam[1];
        gP[0] = -v * lam[1];
        gP[1] = cos(T) * lam[1];
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int, int&, P& T, P&, P* TD, P* X, P*, P*, int*, P* ACC, int*)
    {
        using mpgos::math::fabs;
        T = TD[0];
        ACC[0] = fabs(X[0]);
    }

    template <class P>
    MPGOS_HD static void ActionAfterSuccessfulTimeStep(int, int, int&, P&, P&, P*, P* X, P*, P*, int*, P* ACC, int*)
    {
        using mpgos::math::fabs;
        const P ax = fabs(X[0]);
        if (ax > ACC[0]) ACC[0] = ax;
    }
};

// Duffing with a quadrature state q' = x^2 (X = (x, v, q)), cPAR = (k, B).
// The loss "mean of x^2 over the last forcing period" is (q(H) - q(H-2pi))/2pi.
struct DuffingQ : NoCallbacks {
    static constexpr const char* name = "duffing_q";
    static constexpr int SD = 3, NCP = 2, NSP = 0, NISP = 0, NE = 0, NA = 0, NIA = 0;
    static constexpr bool HAS_VJP = true;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P T, P* cPAR, P*, int*, P*, int*)
    {
        using mpgos::math::cos;
        using mpgos::math::fma;
        const P x = X[0], v = X[1];
        F[0] = v;
        F[1] = fma(-cPAR[0], v, cPAR[1] * cos(T)) - fma(x * x, x, x);
        F[2] = x * x;
    }

    template <class P>
    MPGOS_HD static void OdeVJP(int, int, const P* lam, P* X, P T, P* cPAR, P*, int*, P*, int*, P* gX, P* gP)
    {
        using mpgos::math::cos;
        const P x = X[0], v = X[1];
        gX[0] = -(P(1) + P(3) * x * x) * lam[1] + P(2) * x * lam[2];
        gX[1] = lam[0] - cPAR[0] * lam[1];
        gX[2] = P(0);
        gP[0] = -v * lam[1];
        gP[1] = cos(T) * lam[1];
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int, int&, P& T, P&, P* TD, P*, P*, P*, int*, P*, int*)
    {
        T = TD[0];
    }
};

// Uneven time scales: x'' + k x' + x + x^3 = B cos(eps t), cPAR = (k, B, eps),
// one event function EF = x' (Poincare section at turning points).  Hits are
// recorded by the solver's event log.
struct DuffingEps : NoCallbacks {
    static constexpr const char* name = "duffing_eps";
    static constexpr int SD = 2, NCP = 3, NSP = 0, NISP = 0, NE = 1, NA = 0, NIA = 0;
    static constexpr bool HAS_VJP = true;

    template <class P>
    MPGOS_HD static void OdeFunction(int, int, P* F, P* X, P T, P* cPAR, P*, int*, P*, int*)
    {
        using mpgos::math::cos;
        using mpgos::math::fma;
        const P x = X[0], v = X[1];
        F[0] = v;
        F[1] = fma(-cPAR[0], v, cPAR[1] * cos(cPAR[2] * T)) - fma(x * x, x, x);
    }

    template <class P>
    MPGOS_HD static void OdeVJP(int, int, const P* lam, P* X, P T, P* cPAR, P*, int*, P*, int*, P* gX, P* gP)
    {
        using mpgos::math::cos;
        using mpgos::math::sin;
        const P x = X[0], v = X[1];
        gX[0] = -(P(1) + P(3) * x * x) * lam[1];
        gX[1] = lam[0] - cPAR[0] * lam[1];
        gP[0] = -v * lam[1];
        gP[1] = cos(cPAR[2] * T) * lam[1];
        gP[2] = -cPAR[1] * T * sin(cPAR[2] * T) * lam[1];
    }

    template <class P>
    MPGOS_HD static void EventFunction(int, int, P* EF, P, P, P*, P* X, P*, P*, int*, P*, int*)
    {
        EF[0] = X[1];
    }

    template <class P>
    MPGOS_HD static void Initialization(int, int, int&, P& T, P&, P* TD, P*, P*, P*, int*, P*, int*)
    {
        T = TD[0];
    }
};

} // namespace systems
} // namespace mpgos
