// ORACLE -- test infrastructure only.  Nothing in the product links or calls
// this code; it is used by tests/, by __graft_entry__.smoke() as the checker
// and by bench.py's cpu_baseline leg.
//
// A CPU restatement, in plain C++ over double (or over the op-counting /
// dual-number scalars below), of the reference's per-thread solver path:
//   driver loop      SourceCodes/SingleSystem_PerThread_Solver.cuh:91-435
//   RK4 stepper      SourceCodes/SingleSystem_PerThread_ExplicitRungeKutta_Steppers.cuh:6-117
//   RKCK45 stepper   SourceCodes/SingleSystem_PerThread_ExplicitRungeKutta_Steppers.cuh:121-290
//   controllers      SourceCodes/SingleSystem_PerThread_ExplicitRungeKutta_ErrorControllers.cuh:5-86
//   event control    SourceCodes/SingleSystem_PerThread_EventHandling.cuh:5-45
//   event detection  SourceCodes/SingleSystem_PerThread_Solver.cuh:323-345
//   dense output     SourceCodes/SingleSystem_PerThread_DenseOutput.cuh:5-58
//   fmin/fmax/fabs   SourceCodes/MPGOS_Overloaded_MathFunction.cuh:7-39 (libm semantics)
// Multiply-adds of the Runge-Kutta stage combinations are written as fma()
// (the contracted form of the reference's expressions); everything else is
// compiled with -ffp-contract=off.  The elementary functions cos/sin/exp/log/pow
// and the user systems are the library's own (shared, MPGOS_HD) definitions,
// because the reference calls CUDA libm there and bit-identical parity needs the
// same function on both sides; tests/test_math.py pins those functions against
// the platform libm independently.
//
// Deliberate deviations from the reference, shared with the GPU path:
//  * the event-option staging bug (PT-SOL:60, `ltid < SD` instead of `< NE`)
//    is not reproduced: every event uses its own tolerance and direction;
//  * the first dense-output store of a Solve is skipped when the dense index is
//    already NDO (the reference writes past the buffer, PT-SOL:179-193);
//  * device printf is replaced by per-instance status bits and counters.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "mpgos_math.h"

namespace oracle {

// Status bits (same values as the library's MPGOS_STATUS_*).
enum { ST_NONFINITE_RK4 = 1, ST_NONFINITE_MIN_STEP = 2, ST_MIN_STEP = 4, ST_EVENT_OMITTED = 8 };

struct Job {
    int nt, ndo, algorithm, n_active;       // algorithm: 0 RK4, 1 RKCK45
    double dt0, dtmax, dtmin, grow, shrink, do_min_dt;
    int do_freq;
    const double* rtol; const double* atol; const double* etol; const int* edir;
    int evcap;
    double* TD; double* X; double* T; double* cPAR; double* sPAR; int* sPARi; double* ACC; int* ACCi;
    int* DOidx; double* DOt; double* DOx;
    long long* nacc; long long* nrej; int* status; int* nev; double* evlog; int* evlogidx;
    int nthreads;
};

template <int N> struct AtLeast1 { static constexpr int v = N > 0 ? N : 1; };

inline double dmin(double a, double b) { return std::fmin(a, b); }
inline double dmax(double a, double b) { return std::fmax(a, b); }
// Step sizes carry no derivative (dual.h overloads this for dual numbers).
inline double frozen(double a) { return a; }
// Seeds the differentiation directions after the inputs are read (dual.h).
template <class S> inline void seed_directions(S*, int, S*, int) {}

// The scalar type S is double for the oracle proper; the op-counting scalar
// (flops.h) and the dual numbers (dual.h) reuse the same code.
template <class Sys, class S>
struct Instance {
    static constexpr int SD = Sys::SD, NCP = Sys::NCP, NE = Sys::NE, NA = Sys::NA, NIA = Sys::NIA;
    const Job& J;
    const int tid;
    S* sPAR;          // shared parameters (read-only copies)
    int* sPARi;

    S td[2], x[SD], cp[AtLeast1<NCP>::v], acc[AtLeast1<NA>::v];
    int acci[AtLeast1<NIA>::v];
    S ef[AtLeast1<NE>::v];
    S t, dt, dtNew, doNext;
    int doIdx, doSkipped, stop, userStop;
    long long accepted = 0, rejected = 0;
    int flags = 0, events = 0;
    // optional per-attempt trace (tests): time, step, accepted, tol/err ratio
    struct Trace { int cap; int n; double* t; double* h; int* acc; double* ratio; };
    Trace* trace = nullptr;
    S lastRatio = S(0.0);

    Instance(const Job& j, int id, S* sp, int* spi) : J(j), tid(id), sPAR(sp), sPARi(spi) {}

    static S fmn(const S& a, const S& b) { using std::fmin; return fmin(a, b); }
    static S fmx(const S& a, const S& b) { using std::fmax; return fmax(a, b); }
    static S fab(const S& a) { using std::fabs; return fabs(a); }
    static bool fin(const S& a) { using std::isfinite; return isfinite(a); }
    static S fm(const S& a, const S& b, const S& c) { using std::fma; return fma(a, b, c); }
    static S pw(const S& a, double e) { using mpgos::math::pow; return pow(a, S(e)); }
    static S cr(const S& a, bool fifth) { using mpgos::math::ctrl_root; return ctrl_root(a, fifth); }
    static S frz(const S& a) { using oracle::frozen; return frozen(a); }

    void read()
    {
        const int NT = J.nt;
        for (int i = 0; i < 2; i++) td[i] = J.TD[tid + i * NT];
        for (int i = 0; i < SD; i++) x[i] = J.X[tid + i * NT];
        for (int i = 0; i < NCP; i++) cp[i] = J.cPAR[tid + i * NT];
        for (int i = 0; i < NA; i++) acc[i] = J.ACC[tid + i * NT];
        for (int i = 0; i < NIA; i++) acci[i] = J.ACCi[tid + i * NT];
        t = J.T[tid];
        doIdx = J.DOidx[tid];
    }

    void write()
    {
        const int NT = J.nt;
        for (int i = 0; i < 2; i++) J.TD[tid + i * NT] = value(td[i]);
        for (int i = 0; i < SD; i++) J.X[tid + i * NT] = value(x[i]);
        for (int i = 0; i < NCP; i++) J.cPAR[tid + i * NT] = value(cp[i]);
        for (int i = 0; i < NA; i++) J.ACC[tid + i * NT] = value(acc[i]);
        for (int i = 0; i < NIA; i++) J.ACCi[tid + i * NT] = acci[i];
        J.T[tid] = value(t);
        J.DOidx[tid] = doIdx;
        J.nacc[tid] = accepted;
        J.nrej[tid] = rejected;
        J.status[tid] = flags;
        J.nev[tid] = events;
    }
    static double value(const S& s) { return (double)s; }

    void dense_store()
    {
        const int NT = J.nt;
        J.DOt[tid + doIdx * NT] = value(t);
        for (int i = 0; i < SD; i++) J.DOx[tid + i * NT + doIdx * NT * SD] = value(x[i]);
        doIdx += 1;
        doSkipped = 0;
        doNext = fmn(t + S(J.do_min_dt), td[1]);
    }

    void rhs(S* f, S* y, S time) { Sys::template OdeFunction<S>(tid, J.nt, f, y, time, cp, sPAR, sPARi, acc, acci); }

    // Classic RK4 with the reference's accumulation of k1 + 2 k2 + 2 k3 in the
    // output array (PT-STP:31-116).
    bool step_rk4(S h, S* out)
    {
        S k[SD], y[SD];
        const S half = S(0.5) * h;
        const S sixth = S(1.0 / 6.0) * h;
        rhs(out, x, t);
        S tm = t + half;
        for (int i = 0; i < SD; i++) y[i] = fm(out[i], half, x[i]);
        rhs(k, y, tm);
        for (int i = 0; i < SD; i++) { out[i] = fm(S(2.0), k[i], out[i]); y[i] = fm(k[i], half, x[i]); }
        rhs(k, y, tm);
        tm = t + h;
        for (int i = 0; i < SD; i++) { out[i] = fm(S(2.0), k[i], out[i]); y[i] = fm(k[i], h, x[i]); }
        rhs(k, y, tm);
        bool ok = true;
        for (int i = 0; i < SD; i++) {
            out[i] = fm(sixth, out[i] + k[i], x[i]);
            if (!fin(out[i])) ok = false;
        }
        return ok;
    }

    // Cash-Karp 4(5) (PT-STP:148-289).  Stage coefficients (a_ij, c_i) and
    // weights (b, b - b*) are those of the reference's expressions.
    bool step_rkck45(S h, S* out, S* err)
    {
        S K[6][SD], y[SD];
        rhs(K[0], x, t);
        // stage 2
        for (int i = 0; i < SD; i++) y[i] = fm(h, S(1.0 / 5.0) * K[0][i], x[i]);
        rhs(K[1], y, fm(h, S(1.0 / 5.0), t));
        // stage 3
        for (int i = 0; i < SD; i++) {
            S a = S(3.0 / 40.0) * K[0][i];
            a = fm(S(9.0 / 40.0), K[1][i], a);
            y[i] = fm(h, a, x[i]);
        }
        rhs(K[2], y, fm(h, S(3.0 / 10.0), t));
        // stage 4
        for (int i = 0; i < SD; i++) {
            S a = S(3.0 / 10.0) * K[0][i];
            a = fm(S(-9.0 / 10.0), K[1][i], a);
            a = fm(S(6.0 / 5.0), K[2][i], a);
            y[i] = fm(h, a, x[i]);
        }
        rhs(K[3], y, fm(h, S(3.0 / 5.0), t));
        // stage 5
        for (int i = 0; i < SD; i++) {
            S a = S(-11.0 / 54.0) * K[0][i];
            a = fm(S(5.0 / 2.0), K[1][i], a);
            a = fm(S(-70.0 / 27.0), K[2][i], a);
            a = fm(S(35.0 / 27.0), K[3][i], a);
            y[i] = fm(h, a, x[i]);
        }
        rhs(K[4], y, t + h);
        // stage 6
        for (int i = 0; i < SD; i++) {
            S a = S(1631.0 / 55296.0) * K[0][i];
            a = fm(S(175.0 / 512.0), K[1][i], a);
            a = fm(S(575.0 / 13824.0), K[2][i], a);
            a = fm(S(44275.0 / 110592.0), K[3][i], a);
            a = fm(S(253.0 / 4096.0), K[4][i], a);
            y[i] = fm(h, a, x[i]);
        }
        rhs(K[5], y, fm(h, S(7.0 / 8.0), t));
        bool ok = true;
        for (int i = 0; i < SD; i++) {
            S b = S(37.0 / 378.0) * K[0][i];
            b = fm(S(250.0 / 621.0), K[2][i], b);
            b = fm(S(125.0 / 594.0), K[3][i], b);
            b = fm(S(512.0 / 1771.0), K[5][i], b);
            out[i] = fm(h, b, x[i]);
            S e = S(37.0 / 378.0 - 2825.0 / 27648.0) * K[0][i];
            e = fm(S(250.0 / 621.0 - 18575.0 / 48384.0), K[2][i], e);
            e = fm(S(125.0 / 594.0 - 13525.0 / 55296.0), K[3][i], e);
            e = fm(S(0.0 - 277.0 / 14336.0), K[4][i], e);
            e = fm(S(512.0 / 1771.0 - 1.0 / 4.0), K[5][i], e);
            err[i] = fm(h, fab(e), S(1e-18));
            if (!fin(out[i]) || !fin(err[i])) ok = false;
        }
        return ok;
    }

    // PT-CTL:38-85
    void control(const S* out, const S* err, int& accept, bool& finite)
    {
        S ratio = S(1e30);
        for (int i = 0; i < SD; i++) {
            const S tol = fmx(S(J.rtol[i]) * fmx(fab(out[i]), fab(x[i])), S(J.atol[i]));
            accept = accept && (err[i] < tol);
            ratio = fmn(ratio, tol / err[i]);
        }
        lastRatio = ratio;
        S factor = S(0.9) * cr(ratio, accept == 1);   // R^(1/5) or R^(1/4), as the kernel computes it
        if (!fin(factor)) finite = false;
        const bool tiny = dt < S(J.dtmin) * S(1.01);
        if (!finite) {
            factor = S(J.shrink);
            accept = 0;
            if (tiny) { stop = 1; flags |= ST_NONFINITE_MIN_STEP; }
        } else if (tiny) {
            accept = 1;
            flags |= ST_MIN_STEP;
        }
        factor = fmx(fmn(factor, S(J.grow)), S(J.shrink));
        dtNew = frz(fmx(fmn(dt * factor, S(J.dtmax)), S(J.dtmin)));
    }

    static bool crosses(const S& a, const S& b, double tol, int dir)
    {
        return (a > S(tol) && b < S(-tol) && dir <= 0) || (a < S(-tol) && b > S(tol) && dir >= 0);
    }
    static bool lands(const S& a, const S& b, double tol, int dir)
    {
        return (a > S(tol) && fab(b) < S(tol) && dir <= 0) || (a < S(-tol) && fab(b) < S(tol) && dir >= 0);
    }

    void run()
    {
        read();
        seed_directions(x, SD, cp, NCP);
        dt = S(J.dt0);
        dtNew = S(J.dt0);
        doSkipped = 0;
        stop = 0;
        userStop = 0;
        doNext = S(0.0);
        const int NT = J.nt;
        Sys::template Initialization<S>(tid, NT, doIdx, t, dt, td, x, cp, sPAR, sPARi, acc, acci);
        if (NE > 0) Sys::template EventFunction<S>(tid, NT, ef, t, dt, td, x, cp, sPAR, sPARi, acc, acci);
        if (J.ndo > 0 && doIdx < J.ndo) dense_store();

        S out[SD], err[SD], efNext[AtLeast1<NE>::v];
        while (stop == 0) {
            int accept = 1;
            bool finite = true;
            int atEnd = 0;
            dt = dtNew;
            if (dt > td[1] - t) { dt = td[1] - t; atEnd = 1; }

            if (J.algorithm == 0) {
                finite = step_rk4(dt, out);
                if (!finite) { stop = 1; flags |= ST_NONFINITE_RK4; }
                dtNew = S(J.dt0);
            } else {
                finite = step_rkck45(dt, out, err);
                control(out, err, accept, finite);
            }

            if (NE > 0) {
                Sys::template EventFunction<S>(tid, NT, efNext, t + dt, dt, td, out, cp, sPAR, sPARi, acc, acci);
                // PT-EVT:5-45
                S dtEvent = dt;
                bool shorten = false;
                if (accept == 1 && stop == 0) {
                    for (int i = 0; i < NE; i++) {
                        if (crosses(ef[i], efNext[i], J.etol[i], J.edir[i])) {
                            dtEvent = fmn(dtEvent, -ef[i] / (efNext[i] - ef[i]) * dt);
                            shorten = true;
                        }
                    }
                }
                if (shorten) {
                    if (dtEvent < S(J.dtmin)) flags |= ST_EVENT_OMITTED;
                    else { dtNew = frz(dtEvent); accept = 0; }
                }
            }

            if (trace && trace->n < trace->cap) {
                trace->t[trace->n] = value(t);
                trace->h[trace->n] = value(dt);
                trace->acc[trace->n] = accept;
                trace->ratio[trace->n] = value(lastRatio);
                trace->n++;
            }
            if (accept != 1) { rejected++; continue; }

            t = t + dt;
            for (int i = 0; i < SD; i++) x[i] = out[i];
            Sys::template ActionAfterSuccessfulTimeStep<S>(tid, NT, userStop, t, dt, td, x, cp, sPAR, sPARi, acc, acci);
            if (NE > 0) {
                for (int i = 0; i < NE; i++) {
                    if (!lands(ef[i], efNext[i], J.etol[i], J.edir[i])) continue;
                    if (events < J.evcap) {
                        const long long row = (long long)events * (SD + 1);
                        J.evlog[tid + row * NT] = value(t);
                        for (int c = 0; c < SD; c++) J.evlog[tid + (row + 1 + c) * NT] = value(x[c]);
                        J.evlogidx[tid + (long long)events * NT] = i;
                    }
                    events++;
                    Sys::template ActionAfterEventDetection<S>(tid, NT, i, userStop, t, dt, td, x, cp, sPAR, sPARi, acc, acci);
                }
                Sys::template EventFunction<S>(tid, NT, efNext, t, dt, td, x, cp, sPAR, sPARi, acc, acci);
                for (int i = 0; i < NE; i++) ef[i] = efNext[i];
            }
            if (J.ndo > 0) {
                bool store = doIdx < J.ndo && doNext < t && doSkipped >= J.do_freq - 1;
                if (doIdx < J.ndo && (atEnd == 1 || userStop == 1)) store = true;
                if (store) dense_store();
                else doSkipped++;
            }
            if (atEnd == 1 || userStop == 1) stop = 1;
            accepted++;
        }
        Sys::template Finalization<S>(tid, NT, doIdx, t, dt, td, x, cp, sPAR, sPARi, acc, acci);
        write();
    }
};

} // namespace oracle
