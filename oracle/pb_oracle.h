// ORACLE -- test infrastructure only (see pt_oracle.h).
//
// FP64 CPU restatement of the reference's per-block coupled-systems solver,
// one system at a time, its units processed in order (the reference runs them
// in parallel between barriers; callbacks that write system scope do so from
// unit 0 only, as the reference's tutorials do):
//   driver        SourceCodes/CoupledSystmes_PerBlock_Solver.cuh:2214-2888
//   RK4, RKCK45   SourceCodes/CoupledSystems_PerBlock_ExplicitRungeKutta_Steppers.cuh:2016-2544
//   coupling      same file :2634-2713 (all four storage forms)
//   controllers   SourceCodes/CoupledSystems_PerBlock_ExplicitRungeKutta_ErrorControllers.cuh:424-542
//   events        SourceCodes/CoupledSystems_PerBlock_EventHandling.cuh:215-268
//   dense output  SourceCodes/CoupledSystems_PerBlock_DenseOutput.cuh:312-395
// Stage multiply-adds as fma() in the contracted form of the reference's
// expressions (per-block RK4 ends with x + h (sum) / 6, unlike the per-thread
// form); coupling sums accumulate column by column with fma.
#pragma once

#include <cmath>
#include <vector>

#include "mpgos_math.h"

namespace oracle {

struct PBJob {
    int ns, ups, n_active, ndo, algorithm, cbw, cci, cmat_size;
    double dt0, dtmax, dtmin, grow, shrink, do_min_dt;
    int do_freq;
    const double* rtol; const double* atol; const double* etol; const int* edir;
    double* X; double* uPAR; double* uACC; int* uACCi;                       // stride NS*UPS
    double* TD; double* T; double* sPAR; double* sACC; int* sACCi; double* strength; int* DOidx; double* DOt;
    double* DOx;
    double* gPAR; int* gPARi; int* cIndex; double* cMat;
    long long* nacc; long long* nrej; int* status;
    int nthreads;
};

template <class Sys>
struct PBSystem {
    static constexpr int UD = Sys::UD, NC = Sys::NC, NUP = Sys::NUP, NSP = Sys::NSP, NUA = Sys::NUA, NiUA = Sys::NiUA;
    static constexpr int NSA = Sys::NSA, NiSA = Sys::NiSA, NE = Sys::NE;
    template <int N> struct L1 { static constexpr int v = N > 0 ? N : 1; };
    const PBJob& J;
    const int sid;
    const int U;
    std::vector<double> gPAR;
    std::vector<int> gPARi;

    // system scope
    double td[2], t, dt, dtNew, doNext;
    double sPAR[L1<NSP>::v], sACC[L1<NSA>::v], strength[L1<NC>::v];
    int sACCi[L1<NiSA>::v];
    int doIdx, skipped, stop, userStop;
    long long accepted = 0, rejected = 0;
    int flags = 0;
    std::vector<double>* trace = nullptr;   // when set: (t, h) of every accepted step
    // unit scope
    std::vector<double> x, up, ua, ef, cpf, cpt;
    std::vector<int> uai;

    PBSystem(const PBJob& j, int s) : J(j), sid(s), U(j.ups) {}

    double* X(int u) { return &x[(size_t)u * UD]; }
    double* UP(int u) { return &up[(size_t)u * L1<NUP>::v]; }
    double* UA(int u) { return &ua[(size_t)u * L1<NUA>::v]; }
    int* UAI(int u) { return &uai[(size_t)u * L1<NiUA>::v]; }
    double* EF(int u) { return &ef[(size_t)u * L1<NE>::v]; }
    double* CPF(int u) { return &cpf[(size_t)u * L1<NC>::v]; }
    double* CPT(int u) { return &cpt[(size_t)u * L1<NC>::v]; }

    size_t gu(int u) const { return (size_t)sid * U + u; }
    size_t NU() const { return (size_t)J.ns * U; }

    double coupling_value(int u, int i)
    {
        double v = 0.0;
        if (J.cci == 0 && J.cbw == 0) {
            for (int col = 0; col < U; col++)
                v = std::fma(J.cMat[(size_t)u + (size_t)col * U + (size_t)i * U * U], CPT(col)[i], v);
        } else {
            const int W = J.cbw > 0 ? 2 * J.cbw + 1 : U;
            for (int d = 0; d < W; d++) {
                int col = u + d - J.cbw;
                if (col >= U) col -= U;
                if (col < 0) col += U;
                const double m = J.cci == 0 ? J.cMat[(size_t)u + (size_t)d * U + (size_t)i * U * W] : J.cMat[(size_t)d + (size_t)i * W];
                v = std::fma(m, CPT(col)[i], v);
            }
        }
        return v;
    }

    // all units' right-hand sides at (Y, T), then the coupling terms
    void rhs(std::vector<double>& F, const std::vector<double>& Y, double T)
    {
        for (int u = 0; u < U; u++)
            Sys::template OdeFunction<double>(sid, u, &F[(size_t)u * UD], const_cast<double*>(&Y[(size_t)u * UD]), T, UP(u),
                                              sPAR, gPAR.data(), gPARi.data(), UA(u), UAI(u), sACC, sACCi, CPT(u), CPF(u));
        for (int u = 0; u < U; u++)
            for (int i = 0; i < NC; i++) {
                const double v = coupling_value(u, i);
                double& f = F[(size_t)u * UD + J.cIndex[i]];
                f = std::fma(strength[i] * CPF(u)[i], v, f);
            }
    }

#define MPGOS_PBO_USER(u, TT, DTT) TT, DTT, td, X(u), UP(u), sPAR, gPAR.data(), gPARi.data(), UA(u), UAI(u), sACC, sACCi

    void dense_store()
    {
        for (int u = 0; u < U; u++)
            for (int c = 0; c < UD; c++) J.DOx[gu(u) + (size_t)c * NU() + (size_t)doIdx * NU() * UD] = X(u)[c];
        J.DOt[sid + (size_t)doIdx * J.ns] = t;
        doIdx++;
        skipped = 0;
        doNext = std::fmin(t + J.do_min_dt, td[1]);
    }

    void run()
    {
        const size_t nu = NU();
        x.assign((size_t)U * UD, 0); up.assign((size_t)U * L1<NUP>::v, 0); ua.assign((size_t)U * L1<NUA>::v, 0);
        uai.assign((size_t)U * L1<NiUA>::v, 0); ef.assign((size_t)U * L1<NE>::v, 0); cpf.assign((size_t)U * L1<NC>::v, 0);
        cpt.assign((size_t)U * L1<NC>::v, 0);
        gPAR.assign(Sys::NGP > 0 ? Sys::NGP : 1, 0); gPARi.assign(Sys::NiGP > 0 ? Sys::NiGP : 1, 0);
        for (int i = 0; i < Sys::NGP; i++) gPAR[i] = J.gPAR[i];
        for (int i = 0; i < Sys::NiGP; i++) gPARi[i] = J.gPARi[i];
        for (int u = 0; u < U; u++) {
            for (int c = 0; c < UD; c++) X(u)[c] = J.X[gu(u) + c * nu];
            for (int c = 0; c < NUP; c++) UP(u)[c] = J.uPAR[gu(u) + c * nu];
            for (int c = 0; c < NUA; c++) UA(u)[c] = J.uACC[gu(u) + c * nu];
            for (int c = 0; c < NiUA; c++) UAI(u)[c] = J.uACCi[gu(u) + c * nu];
        }
        td[0] = J.TD[sid]; td[1] = J.TD[sid + J.ns];
        t = J.T[sid];
        dt = J.dt0; dtNew = J.dt0;
        for (int i = 0; i < NSP; i++) sPAR[i] = J.sPAR[sid + i * J.ns];
        for (int i = 0; i < NSA; i++) sACC[i] = J.sACC[sid + i * J.ns];
        for (int i = 0; i < NiSA; i++) sACCi[i] = J.sACCi[sid + i * J.ns];
        for (int i = 0; i < NC; i++) strength[i] = J.strength[sid + i * J.ns];
        doIdx = J.DOidx[sid];
        skipped = 0; stop = 0; userStop = 0; doNext = 0.0;

        for (int u = 0; u < U; u++) Sys::template Initialization<double>(sid, u, doIdx, MPGOS_PBO_USER(u, t, dt));
        if (NE > 0)
            for (int u = 0; u < U; u++) Sys::template EventFunction<double>(sid, u, EF(u), MPGOS_PBO_USER(u, t, dt));
        if (J.ndo > 0 && doIdx < J.ndo) dense_store();

        const size_t n = (size_t)U * UD;
        std::vector<double> k1(n), k2(n), k3(n), k4(n), k5(n), k6(n), Y(n), next(n), err(n);
        std::vector<double> efn((size_t)U * L1<NE>::v);
        while (stop == 0) {
            int accept = 1, finite = 1, atEnd = 0;
            dt = dtNew;
            if (dt > td[1] - t) { dt = td[1] - t; atEnd = 1; }
            const double h = dt;
            if (J.algorithm == 0) {
                rhs(next, x, t);
                for (size_t e = 0; e < n; e++) Y[e] = std::fma(next[e] * 0.5, h, x[e]);
                const double t2 = t + 0.5 * h;
                rhs(k1, Y, t2);
                for (size_t e = 0; e < n; e++) { next[e] = std::fma(2.0, k1[e], next[e]); Y[e] = std::fma(k1[e] * 0.5, h, x[e]); }
                rhs(k1, Y, t2);
                for (size_t e = 0; e < n; e++) { next[e] = std::fma(2.0, k1[e], next[e]); Y[e] = std::fma(k1[e], h, x[e]); }
                rhs(k1, Y, t + h);
                for (size_t e = 0; e < n; e++) {
                    next[e] = std::fma(h * (next[e] + k1[e]), 1.0 / 6.0, x[e]);
                    if (!std::isfinite(next[e])) finite = 0;
                }
                // PB-CTL:424-450
                if (!finite) { stop = 1; accept = 0; flags |= 1; }
                dtNew = J.dt0;
            } else {
                rhs(k1, x, t);
                for (size_t e = 0; e < n; e++) Y[e] = std::fma(k1[e] * (1.0 / 5.0), h, x[e]);
                rhs(k2, Y, std::fma(1.0 / 5.0, h, t));
                for (size_t e = 0; e < n; e++) Y[e] = std::fma(std::fma(k2[e], 9.0 / 40.0, k1[e] * (3.0 / 40.0)), h, x[e]);
                rhs(k3, Y, std::fma(3.0 / 10.0, h, t));
                for (size_t e = 0; e < n; e++)
                    Y[e] = std::fma(std::fma(k3[e], 6.0 / 5.0, std::fma(k2[e], -9.0 / 10.0, k1[e] * (3.0 / 10.0))), h, x[e]);
                rhs(k4, Y, std::fma(3.0 / 5.0, h, t));
                for (size_t e = 0; e < n; e++)
                    Y[e] = std::fma(std::fma(k4[e], 35.0 / 27.0, std::fma(k3[e], -70.0 / 27.0, std::fma(k2[e], 5.0 / 2.0, k1[e] * (-11.0 / 54.0)))), h, x[e]);
                rhs(k5, Y, t + h);
                for (size_t e = 0; e < n; e++)
                    Y[e] = std::fma(std::fma(k5[e], 253.0 / 4096.0, std::fma(k4[e], 44275.0 / 110592.0, std::fma(k3[e], 575.0 / 13824.0,
                                    std::fma(k2[e], 175.0 / 512.0, k1[e] * (1631.0 / 55296.0))))), h, x[e]);
                rhs(k6, Y, std::fma(7.0 / 8.0, h, t));
                double ratio = 1e30;
                for (size_t e = 0; e < n; e++) {
                    next[e] = std::fma(std::fma(k6[e], 512.0 / 1771.0, std::fma(k4[e], 125.0 / 594.0, std::fma(k3[e], 250.0 / 621.0, k1[e] * (37.0 / 378.0)))), h, x[e]);
                    double q = k1[e] * (37.0 / 378.0 - 2825.0 / 27648.0);
                    q = std::fma(k3[e], 250.0 / 621.0 - 18575.0 / 48384.0, q);
                    q = std::fma(k4[e], 125.0 / 594.0 - 13525.0 / 55296.0, q);
                    q = std::fma(k5[e], 0.0 - 277.0 / 14336.0, q);
                    q = std::fma(k6[e], 512.0 / 1771.0 - 1.0 / 4.0, q);
                    err[e] = std::fma(h, std::fabs(q), 1e-18);
                    if (!std::isfinite(next[e]) || !std::isfinite(err[e])) finite = 0;
                    const int c = (int)(e % UD);
                    const double tol = std::fmax(J.rtol[c] * std::fmax(std::fabs(next[e]), std::fabs(x[e])), J.atol[c]);
                    accept = accept & (err[e] < tol);
                    ratio = std::fmin(ratio, tol / err[e]);
                }
                // PB-CTL:500-541 (safety factor 0.8)
                double fac = 0.8 * mpgos::math::pow(ratio, accept == 1 ? 1.0 / 5.0 : 1.0 / 4.0);
                if (!std::isfinite(fac)) finite = 0;
                if (!finite) {
                    fac = J.shrink;
                    accept = 0;
                    if (dt < J.dtmin * 1.01) { stop = 1; flags |= 2; }
                } else if (dt < J.dtmin * 1.01) {
                    accept = 1;
                    flags |= 4;
                }
                fac = std::fmax(std::fmin(fac, J.grow), J.shrink);
                dtNew = std::fmax(std::fmin(dt * fac, J.dtmax), J.dtmin);
            }
            if (NE > 0 && stop == 0) {
                const double tn = t + dt;
                for (int u = 0; u < U; u++)
                    Sys::template EventFunction<double>(sid, u, &efn[(size_t)u * L1<NE>::v], tn, dtNew, td, &next[(size_t)u * UD], UP(u),
                                                        sPAR, gPAR.data(), gPARi.data(), UA(u), UAI(u), sACC, sACCi);
                if (accept == 1) {
                    double evDt = dt;
                    bool corrected = false;
                    for (int u = 0; u < U; u++)
                        for (int i = 0; i < NE; i++) {
                            const double a = EF(u)[i], b = efn[(size_t)u * L1<NE>::v + i], tol = J.etol[i];
                            if ((a > tol && b < -tol && J.edir[i] <= 0) || (a < -tol && b > tol && J.edir[i] >= 0)) {
                                evDt = std::fmin(evDt, -a / (b - a) * dt);
                                corrected = true;
                            }
                        }
                    if (corrected) {
                        if (evDt < J.dtmin) flags |= 8;
                        else { dtNew = evDt; accept = 0; }
                    }
                }
            }
            if (stop != 0 || accept != 1) {
                if (stop == 0) rejected++;
                continue;
            }
            if (trace) { trace->push_back(t); trace->push_back(dt); }
            x = next;
            t += dt;
            for (int u = 0; u < U; u++) Sys::template ActionAfterSuccessfulTimeStep<double>(sid, u, userStop, MPGOS_PBO_USER(u, t, dt));
            if (NE > 0) {
                for (int u = 0; u < U; u++)
                    for (int i = 0; i < NE; i++) {
                        const double a = EF(u)[i], b = efn[(size_t)u * L1<NE>::v + i], tol = J.etol[i];
                        if ((a > tol && std::fabs(b) < tol && J.edir[i] <= 0) || (a < -tol && std::fabs(b) < tol && J.edir[i] >= 0))
                            Sys::template ActionAfterEventDetection<double>(sid, u, i, userStop, MPGOS_PBO_USER(u, t, dt));
                    }
                for (int u = 0; u < U; u++) Sys::template EventFunction<double>(sid, u, EF(u), MPGOS_PBO_USER(u, t, dtNew));
            }
            if (J.ndo > 0) {
                bool store = doIdx < J.ndo && doNext < t && skipped >= J.do_freq - 1;
                if (doIdx < J.ndo && (atEnd == 1 || userStop == 1)) store = true;
                if (store) dense_store();
                else skipped++;
            }
            accepted++;
            if (atEnd == 1 || userStop == 1) stop = 1;
        }
        for (int u = 0; u < U; u++) Sys::template Finalization<double>(sid, u, doIdx, MPGOS_PBO_USER(u, t, dt));
        for (int u = 0; u < U; u++) {
            for (int c = 0; c < UD; c++) J.X[gu(u) + c * nu] = X(u)[c];
            for (int c = 0; c < NUA; c++) J.uACC[gu(u) + c * nu] = UA(u)[c];
            for (int c = 0; c < NiUA; c++) J.uACCi[gu(u) + c * nu] = UAI(u)[c];
        }
        J.T[sid] = t;
        J.TD[sid] = td[0]; J.TD[sid + J.ns] = td[1];
        for (int i = 0; i < NSA; i++) J.sACC[sid + i * J.ns] = sACC[i];
        for (int i = 0; i < NiSA; i++) J.sACCi[sid + i * J.ns] = sACCi[i];
        J.DOidx[sid] = doIdx;
        J.nacc[sid] = accepted;
        J.nrej[sid] = rejected;
        J.status[sid] = flags;
    }
#undef MPGOS_PBO_USER
};

} // namespace oracle
