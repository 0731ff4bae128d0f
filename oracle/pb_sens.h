// ORACLE -- test infrastructure only (see pt_oracle.h).
//
// Forward sensitivities of the per-block integration with frozen step sizes:
// the FP64 oracle (pb_oracle.h) runs a system and records (t_n, h_n) of every
// accepted step; the same steps are then replayed in forward-mode dual numbers
// (dual.h), stages and coupling exactly as pb_oracle.h writes them.  The
// result is the Jacobian of the final unit states of the system with respect
// to N = NSP + UD + NUP directions: the system parameters, the initial state
// of unit 0 and the unit parameters of unit 0 -- the quantities the discrete
// adjoint returns as the system-parameter gradient, dL/dx(0) of unit 0 and
// the unit-parameter gradient of unit 0.  Events and user actions are not
// replayed (the systems used here have none that change the state).
#pragma once

#include <vector>

#include "dual.h"
#include "pb_oracle.h"

namespace oracle {

template <class Sys>
struct PBSens {
    static constexpr int UD = Sys::UD, NC = Sys::NC, NUP = Sys::NUP, NSP = Sys::NSP, NUA = Sys::NUA, NiUA = Sys::NiUA;
    static constexpr int NSA = Sys::NSA, NiSA = Sys::NiSA;
    static constexpr int N = NSP + UD + NUP;
    using D = dual::D<N>;
    template <int M> struct L1 { static constexpr int v = M > 0 ? M : 1; };

    // jac[((i * UD + c) * UPS + u)] = d x_u,c(end) / d direction i, for system sid
    static void run(const PBJob& J, int sid, double* jac)
    {
        using dual::fma;
        const int U = J.ups;
        const size_t nu = (size_t)J.ns * U;
        std::vector<double> tr;
        {
            PBSystem<Sys> fwd(J, sid);
            fwd.trace = &tr;
            // the forward run writes its outputs into J; keep the inputs
            std::vector<double> x0((size_t)UD * U), ua((size_t)L1<NUA>::v * U);
            for (int u = 0; u < U; u++)
                for (int c = 0; c < UD; c++) x0[(size_t)c * U + u] = J.X[(size_t)sid * U + u + c * nu];
            fwd.run();
            for (int u = 0; u < U; u++)
                for (int c = 0; c < UD; c++) J.X[(size_t)sid * U + u + c * nu] = x0[(size_t)c * U + u];
        }
        std::vector<D> x((size_t)U * UD), up((size_t)U * L1<NUP>::v);
        std::vector<double> ua((size_t)U * L1<NUA>::v);
        std::vector<int> uai((size_t)U * L1<NiUA>::v);
        D sp[L1<NSP>::v];
        double sa[L1<NSA>::v] = {0};
        int sai[L1<NiSA>::v] = {0};
        std::vector<double> gp(Sys::NGP > 0 ? Sys::NGP : 1, 0.0);
        std::vector<int> gpi(Sys::NiGP > 0 ? Sys::NiGP : 1, 0);
        for (int i = 0; i < Sys::NGP; i++) gp[i] = J.gPAR[i];
        for (int i = 0; i < Sys::NiGP; i++) gpi[i] = J.gPARi[i];
        for (int i = 0; i < NSP; i++) { sp[i] = D(J.sPAR[sid + i * J.ns]); sp[i].d[i] = 1.0; }
        for (int i = 0; i < NSA; i++) sa[i] = J.sACC[sid + i * J.ns];
        for (int i = 0; i < NiSA; i++) sai[i] = J.sACCi[sid + i * J.ns];
        for (int u = 0; u < U; u++) {
            const size_t g = (size_t)sid * U + u;
            for (int c = 0; c < UD; c++) x[(size_t)u * UD + c] = D(J.X[g + c * nu]);
            for (int c = 0; c < NUP; c++) up[(size_t)u * L1<NUP>::v + c] = D(J.uPAR[g + c * nu]);
            for (int c = 0; c < NUA; c++) ua[(size_t)u * L1<NUA>::v + c] = J.uACC[g + c * nu];
            for (int c = 0; c < NiUA; c++) uai[(size_t)u * L1<NiUA>::v + c] = J.uACCi[g + c * nu];
        }
        for (int c = 0; c < UD; c++) x[c].d[NSP + c] = 1.0;
        for (int c = 0; c < NUP; c++) up[c].d[NSP + UD + c] = 1.0;
        double strength[L1<NC>::v];
        for (int i = 0; i < NC; i++) strength[i] = J.strength[sid + i * J.ns];

        std::vector<D> cpt((size_t)U * L1<NC>::v), cpf((size_t)U * L1<NC>::v);
        // all units' right-hand sides and the coupling terms (pb_oracle.h rhs)
        auto rhs = [&](std::vector<D>& F, const std::vector<D>& Y, const D& T) {
            // the dual system definitions take non-const parameter arrays of the
            // scalar type; accessories are not differentiated
            std::vector<D> uad(ua.size()), sad(L1<NSA>::v), gpd(gp.size());
            for (size_t i = 0; i < ua.size(); i++) uad[i] = D(ua[i]);
            for (int i = 0; i < L1<NSA>::v; i++) sad[i] = D(sa[i]);
            for (size_t i = 0; i < gp.size(); i++) gpd[i] = D(gp[i]);
            for (int u = 0; u < U; u++)
                Sys::template OdeFunction<D>(sid, u, &F[(size_t)u * UD], const_cast<D*>(&Y[(size_t)u * UD]), T,
                                             &up[(size_t)u * L1<NUP>::v], sp, gpd.data(), gpi.data(),
                                             &uad[(size_t)u * L1<NUA>::v], &uai[(size_t)u * L1<NiUA>::v], sad.data(), sai,
                                             &cpt[(size_t)u * L1<NC>::v], &cpf[(size_t)u * L1<NC>::v]);
            for (int u = 0; u < U; u++)
                for (int i = 0; i < NC; i++) {
                    D v(0.0);
                    if (J.cci == 0 && J.cbw == 0) {
                        for (int col = 0; col < U; col++)
                            v = fma(D(J.cMat[(size_t)u + (size_t)col * U + (size_t)i * U * U]), cpt[(size_t)col * L1<NC>::v + i], v);
                    } else {
                        const int W = J.cbw > 0 ? 2 * J.cbw + 1 : U;
                        for (int d = 0; d < W; d++) {
                            int col = u + d - J.cbw;
                            if (col >= U) col -= U;
                            if (col < 0) col += U;
                            const double m = J.cci == 0 ? J.cMat[(size_t)u + (size_t)d * U + (size_t)i * U * W] : J.cMat[(size_t)d + (size_t)i * W];
                            v = fma(D(m), cpt[(size_t)col * L1<NC>::v + i], v);
                        }
                    }
                    D& f = F[(size_t)u * UD + J.cIndex[i]];
                    f = fma(D(strength[i]) * cpf[(size_t)u * L1<NC>::v + i], v, f);
                }
        };

        const size_t n = (size_t)U * UD;
        std::vector<D> k1(n), k2(n), k3(n), k4(n), k5(n), k6(n), Y(n), next(n);
        for (size_t s = 0; s + 1 < tr.size(); s += 2) {
            const double t = tr[s], hh = tr[s + 1];
            const D h(hh);
            if (J.algorithm == 0) {
                rhs(next, x, D(t));
                for (size_t e = 0; e < n; e++) Y[e] = fma(next[e] * D(0.5), h, x[e]);
                const D t2(t + 0.5 * hh);
                rhs(k1, Y, t2);
                for (size_t e = 0; e < n; e++) { next[e] = fma(D(2.0), k1[e], next[e]); Y[e] = fma(k1[e] * D(0.5), h, x[e]); }
                rhs(k1, Y, t2);
                for (size_t e = 0; e < n; e++) { next[e] = fma(D(2.0), k1[e], next[e]); Y[e] = fma(k1[e], h, x[e]); }
                rhs(k1, Y, D(t + hh));
                for (size_t e = 0; e < n; e++) next[e] = fma(h * (next[e] + k1[e]), D(1.0 / 6.0), x[e]);
            } else {
                rhs(k1, x, D(t));
                for (size_t e = 0; e < n; e++) Y[e] = fma(k1[e] * D(1.0 / 5.0), h, x[e]);
                rhs(k2, Y, D(std::fma(1.0 / 5.0, hh, t)));
                for (size_t e = 0; e < n; e++) Y[e] = fma(fma(k2[e], D(9.0 / 40.0), k1[e] * D(3.0 / 40.0)), h, x[e]);
                rhs(k3, Y, D(std::fma(3.0 / 10.0, hh, t)));
                for (size_t e = 0; e < n; e++)
                    Y[e] = fma(fma(k3[e], D(6.0 / 5.0), fma(k2[e], D(-9.0 / 10.0), k1[e] * D(3.0 / 10.0))), h, x[e]);
                rhs(k4, Y, D(std::fma(3.0 / 5.0, hh, t)));
                for (size_t e = 0; e < n; e++)
                    Y[e] = fma(fma(k4[e], D(35.0 / 27.0), fma(k3[e], D(-70.0 / 27.0), fma(k2[e], D(5.0 / 2.0), k1[e] * D(-11.0 / 54.0)))), h, x[e]);
                rhs(k5, Y, D(t + hh));
                for (size_t e = 0; e < n; e++)
                    Y[e] = fma(fma(k5[e], D(253.0 / 4096.0), fma(k4[e], D(44275.0 / 110592.0), fma(k3[e], D(575.0 / 13824.0),
                               fma(k2[e], D(175.0 / 512.0), k1[e] * D(1631.0 / 55296.0))))), h, x[e]);
                rhs(k6, Y, D(std::fma(7.0 / 8.0, hh, t)));
                for (size_t e = 0; e < n; e++)
                    next[e] = fma(fma(k6[e], D(512.0 / 1771.0), fma(k4[e], D(125.0 / 594.0), fma(k3[e], D(250.0 / 621.0), k1[e] * D(37.0 / 378.0)))), h, x[e]);
            }
            x = next;
        }
        for (int i = 0; i < N; i++)
            for (int c = 0; c < UD; c++)
                for (int u = 0; u < U; u++) jac[((size_t)i * UD + c) * U + u] = x[(size_t)u * UD + c].d[i];
    }
};

} // namespace oracle
