// Per-block coupled-systems solver for gfx950: one workgroup integrates SPB
// systems of UPS coupled units; one lane owns one (system, unit) pair.
//
// Semantics follow the reference's per-block kernels
// (SourceCodes/CoupledSystmes_PerBlock_Solver.cuh, single-system form at
// :2214-2888, multi-system masking at :359 and :2532): system scope (time,
// step, system parameters and accessories, dense-output bookkeeping, flags)
// lives in LDS, unit scope in VGPRs; every right-hand-side evaluation is
// followed by the coupling terms
//   F[CouplingIndex[i]] += CouplingStrength[i] * CPF[i] * sum_col M_i(u, col) CPT_col[i]
// (ComputeCouplingValue_SingleSystem, CoupledSystems_PerBlock_ExplicitRungeKutta_Steppers.cuh:2634-2713,
// all four storage forms of the coupling matrix); the step is controlled per
// system by reductions over its units: acceptance = AND over units, ratio =
// min over units of min_i tol_i/err_i, safety factor 0.8
// (CoupledSystems_PerBlock_ExplicitRungeKutta_ErrorControllers.cuh:424-542);
// event step control takes the minimum interpolated step over units
// (CoupledSystems_PerBlock_EventHandling.cuh:215-268); dense output is stored
// only after accepted steps (CoupledSystems_PerBlock_DenseOutput.cuh:312-395).
//
// MI355X design: the reductions are LDS atomics on ordered bit patterns
// (min of non-negative doubles, AND of flags), which are order-independent and
// therefore deterministic; coupling terms are double-buffered in LDS so one
// barrier per stage suffices; the coupling matrix is read from L2 (or staged
// in LDS when SharedCouplingMatrices is on); multiply-adds are explicit fma in
// the contracted form of the reference's expressions, so the host oracle
// reproduces every bit.
#pragma once

#include <hip/hip_runtime.h>

#include "mpgos.h"
#include "mpgos_math.h"

namespace mpgos {
namespace pb {

template <int N> struct Pos { static constexpr int value = N > 0 ? N : 1; };

constexpr int kMaxBlock = 1024;

template <class Sys, class P>
struct Args {
    static constexpr int UD = Sys::UD, NE = Sys::NE;
    // unit scope, index (system * UPS + unit) + component * NU, NU = NS * UPS
    P* X; P* uPAR; P* uACC; int* uACCi;
    // system scope, index system + component * NS
    P* TD; P* T; P* sPAR; P* sACC; int* sACCi; P* strength; int* DOidx; P* DOt;
    P* DOx;                   // unit index + component * NU + k * NU * UD
    // global scope
    P* gPAR; int* gPARi; int* cIndex; P* cMat;
    long long* nacc; long long* nrej; int* status;
    int NS, UPS, SPB, nActive, NDO, CBW, CCI, cMatSize, sharedGlobals, sharedMatrix;
    int doFreq;
    P dt0, dtmax, dtmin, grow, shrink, doMinDt;
    P rtol[UD], atol[UD];
    P etol[Pos<NE>::value];
    int edir[Pos<NE>::value];
};

// System scope of one system (LDS).
template <class Sys, class P>
struct SysScope {
    P TD[2];
    P t, dt, newDt, doT;
    P sPAR[Pos<Sys::NSP>::value];
    P sACC[Pos<Sys::NSA>::value];
    P strength[Pos<Sys::NC>::value];
    unsigned long long ratioBits;     // min over units of tol/err (non-negative double bits)
    unsigned long long eventBits;     // min over units of the interpolated event step
    int sACCi[Pos<Sys::NiSA>::value];
    int DOidx, updDO, skipped;
    int update, isfin, end, term, udt, corrected;
    int nacc, nrej, status;
};

template <class P> __device__ __forceinline__ unsigned long long order_bits(P v);
template <> __device__ __forceinline__ unsigned long long order_bits<double>(double v) { return (unsigned long long)math::bits(v); }
template <> __device__ __forceinline__ unsigned long long order_bits<float>(float v) { return (unsigned long long)math::bits(v); }
template <class P> __device__ __forceinline__ P from_order_bits(unsigned long long b);
template <> __device__ __forceinline__ double from_order_bits<double>(unsigned long long b) { return math::from_bits(b); }
template <> __device__ __forceinline__ float from_order_bits<float>(unsigned long long b) { return math::from_bits32((uint32_t)b); }

template <class P> __device__ __forceinline__ P dmin(P a, P b);
template <> __device__ __forceinline__ double dmin<double>(double a, double b) { return __builtin_fmin(a, b); }
template <> __device__ __forceinline__ float dmin<float>(float a, float b) { return __builtin_fminf(a, b); }
template <class P> __device__ __forceinline__ P dmax(P a, P b);
template <> __device__ __forceinline__ double dmax<double>(double a, double b) { return __builtin_fmax(a, b); }
template <> __device__ __forceinline__ float dmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }

// LDS bytes: system scopes, two coupling-term buffers, optional global
// parameters and coupling matrix copies.
template <class Sys, class P>
__host__ __device__ inline size_t lds_layout(int SPB, int UPS, int sharedGlobals, int sharedMatrix, int cMatSize,
                                             size_t* offCPT, size_t* offG, size_t* offGi, size_t* offM)
{
    size_t o = 0;
    o += (size_t)SPB * sizeof(SysScope<Sys, P>);
    o = (o + 15) & ~(size_t)15;
    *offCPT = o;
    o += 2 * (size_t)SPB * UPS * Pos<Sys::NC>::value * sizeof(P);
    *offG = o;
    if (sharedGlobals) o += (size_t)Sys::NGP * sizeof(P);
    o = (o + 15) & ~(size_t)15;
    *offM = o;
    if (sharedMatrix) o += (size_t)Sys::NC * cMatSize * sizeof(P);
    *offGi = o;
    if (sharedGlobals) o += (size_t)Sys::NiGP * sizeof(int);
    return o;
}

template <class Sys, class P, int ALG>
struct Unit {
    static constexpr int UD = Sys::UD, NC = Sys::NC, NE = Sys::NE;
    using A = Args<Sys, P>;

    // coupling: F[CouplingIndex[i]] += strength_i CPF_i sum_col M_i(u, col) CPT_col[i]
    static __device__ __forceinline__ void couple(const A& a, const P* M, const int* cIndex, const P* strength,
                                                  const P* cptSys, int u, const P* CPF, P* F)
    {
        const int UPS = a.UPS;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            P v = P(0);
            if (a.CCI == 0 && a.CBW == 0) {
                const P* Mi = M + (size_t)i * UPS * UPS + u;
                for (int col = 0; col < UPS; col++) v = math::fma(Mi[(size_t)col * UPS], cptSys[col * NC + i], v);
            } else if (a.CCI == 0) {
                const int W = 2 * a.CBW + 1;
                const P* Mi = M + (size_t)i * UPS * W + u;
                for (int d = 0; d < W; d++) {
                    int col = u + d - a.CBW;
                    if (col >= UPS) col -= UPS;
                    if (col < 0) col += UPS;
                    v = math::fma(Mi[(size_t)d * UPS], cptSys[col * NC + i], v);
                }
            } else {
                const int W = a.CBW > 0 ? 2 * a.CBW + 1 : UPS;
                const P* Mi = M + (size_t)i * W;
                for (int d = 0; d < W; d++) {
                    int col = u + d - a.CBW;
                    if (col >= UPS) col -= UPS;
                    if (col < 0) col += UPS;
                    v = math::fma(Mi[d], cptSys[col * NC + i], v);
                }
            }
            const int ci = cIndex[i];
            const P sc = strength[i] * CPF[i];
#pragma unroll
            for (int c = 0; c < UD; c++)
                if (c == ci) F[c] = math::fma(sc, v, F[c]);
        }
    }
};

// One workgroup = SPB systems.  Lane (lsys, u) with lsys = tid / UPS, u = tid % UPS.
template <class Sys, class P, int ALG>
__device__ void pb_body(const Args<Sys, P>& a)
{
    constexpr int UD = Sys::UD, NC = Sys::NC, NUP = Sys::NUP, NSP = Sys::NSP, NGP = Sys::NGP, NiGP = Sys::NiGP;
    constexpr int NUA = Sys::NUA, NiUA = Sys::NiUA, NSA = Sys::NSA, NiSA = Sys::NiSA, NE = Sys::NE;
    using S = SysScope<Sys, P>;
    using math::fma;

    extern __shared__ __align__(16) unsigned char mpgos_pb_lds[];
    size_t offCPT, offG, offGi, offM;
    lds_layout<Sys, P>(a.SPB, a.UPS, a.sharedGlobals, a.sharedMatrix, a.cMatSize, &offCPT, &offG, &offGi, &offM);
    S* sys = reinterpret_cast<S*>(mpgos_pb_lds);
    P* cptBuf = reinterpret_cast<P*>(mpgos_pb_lds + offCPT);
    const int cptStride = a.SPB * a.UPS * Pos<NC>::value;

    const int UPS = a.UPS, NS = a.NS;
    const int tid = threadIdx.x;
    const int lsys = tid / UPS;
    const int u = tid - lsys * UPS;
    const bool lane = tid < a.SPB * UPS;                       // lane owns a (system, unit) pair
    const int gsys = blockIdx.x * a.SPB + lsys;
    const bool exists = lane && gsys < NS;
    const bool active = exists && gsys < a.nActive;
    const bool leader = lane && u == 0;
    const size_t NU = (size_t)NS * UPS;
    const size_t gu = (size_t)gsys * UPS + u;                 // unit-scope index

    // ---- global scope (LDS when SharedGlobalVariables / SharedCouplingMatrices)
    P* gPAR = a.gPAR;
    int* gPARi = a.gPARi;
    const P* M = a.cMat;
    if (a.sharedGlobals) {
        P* lg = reinterpret_cast<P*>(mpgos_pb_lds + offG);
        int* lgi = reinterpret_cast<int*>(mpgos_pb_lds + offGi);
        for (int i = tid; i < NGP; i += blockDim.x) lg[i] = a.gPAR[i];
        for (int i = tid; i < NiGP; i += blockDim.x) lgi[i] = a.gPARi[i];
        gPAR = lg;
        gPARi = lgi;
    }
    if (a.sharedMatrix) {
        P* lm = reinterpret_cast<P*>(mpgos_pb_lds + offM);
        for (int i = tid; i < NC * a.cMatSize; i += blockDim.x) lm[i] = a.cMat[i];
        M = lm;
    }
    int cIndex[Pos<NC>::value];
#pragma unroll
    for (int i = 0; i < NC; i++) cIndex[i] = a.cIndex[i];

    // ---- system scope (PB-SOL:2285-2335)
    if (leader) {
        S& s = sys[lsys];
        if (active) {
            s.TD[0] = a.TD[gsys]; s.TD[1] = a.TD[gsys + NS];
            s.t = a.T[gsys];
            s.dt = a.dt0; s.newDt = a.dt0;
            for (int i = 0; i < NSP; i++) s.sPAR[i] = a.sPAR[gsys + i * NS];
            for (int i = 0; i < NSA; i++) s.sACC[i] = a.sACC[gsys + i * NS];
            for (int i = 0; i < NiSA; i++) s.sACCi[i] = a.sACCi[gsys + i * NS];
            for (int i = 0; i < NC; i++) s.strength[i] = a.strength[gsys + i * NS];
            s.DOidx = a.DOidx[gsys];
            s.updDO = 1; s.skipped = 0; s.term = 0; s.udt = 0; s.update = 0; s.doT = P(0);
        } else {
            s.term = 1; s.update = 0; s.updDO = 0;
        }
        s.nacc = 0; s.nrej = 0; s.status = 0;
    }

    // ---- unit scope in registers
    P X[UD], uPAR[Pos<NUP>::value], uACC[Pos<NUA>::value], EF[Pos<NE>::value], CPF[Pos<NC>::value];
    int uACCi[Pos<NiUA>::value];
    if (exists) {
#pragma unroll
        for (int c = 0; c < UD; c++) X[c] = a.X[gu + c * NU];
#pragma unroll
        for (int c = 0; c < NUP; c++) uPAR[c] = a.uPAR[gu + c * NU];
#pragma unroll
        for (int c = 0; c < NUA; c++) uACC[c] = a.uACC[gu + c * NU];
#pragma unroll
        for (int c = 0; c < NiUA; c++) uACCi[c] = a.uACCi[gu + c * NU];
    }
#pragma unroll
    for (int i = 0; i < NC; i++) CPF[i] = P(0);
    __syncthreads();

    S* my = lane ? &sys[lsys] : &sys[0];
#define MPGOS_PB_USER(TT, DTT) TT, DTT, my->TD, X, uPAR, my->sPAR, gPAR, gPARi, uACC, uACCi, my->sACC, my->sACCi
    // ---- initialisation (PB-SOL:2465-2528)
    if (active) Sys::template Initialization<P>(gsys, u, my->DOidx, MPGOS_PB_USER(my->t, my->dt));
    __syncthreads();
    if (NE > 0) {
        if (active) Sys::template EventFunction<P>(gsys, u, EF, MPGOS_PB_USER(my->t, my->dt));
        __syncthreads();
    }
    if (a.NDO > 0) {
        // first store (the reference stores unconditionally; skipped once the buffer is full)
        if (active && my->DOidx < a.NDO) {
            const int k = my->DOidx;
#pragma unroll
            for (int c = 0; c < UD; c++) a.DOx[gu + c * NU + (size_t)k * NU * UD] = X[c];
            if (u == 0) a.DOt[gsys + (size_t)k * NS] = my->t;
        }
        __syncthreads();
        if (leader && active && my->DOidx < a.NDO) {
            my->DOidx++;
            my->skipped = 0;
            my->doT = dmin<P>(my->t + a.doMinDt, my->TD[1]);
        }
        __syncthreads();
    }

    int buf = 0;
    auto rhs = [&](P* F, P* Y, P T) {
        // RHS of this unit, coupling terms published, one barrier, coupling added
        P* cpt = cptBuf + (size_t)buf * cptStride;
        if (lane && my->term == 0)
            Sys::template OdeFunction<P>(gsys, u, F, Y, T, uPAR, my->sPAR, gPAR, gPARi, uACC, uACCi, my->sACC, my->sACCi,
                                         cpt + (size_t)(lsys * UPS + u) * Pos<NC>::value, CPF);
        __syncthreads();
        if (lane && my->term == 0)
            Unit<Sys, P, ALG>::couple(a, M, cIndex, my->strength, cpt + (size_t)lsys * UPS * Pos<NC>::value, u, CPF, F);
        buf ^= 1;
    };

    for (;;) {
        // ---- step initialisation (PB-SOL:2535-2549)
        if (leader && my->term == 0) {
            my->update = 1; my->isfin = 1; my->end = 0;
            my->dt = my->newDt;
            if (my->dt > (my->TD[1] - my->t)) { my->dt = my->TD[1] - my->t; my->end = 1; }
            my->ratioBits = order_bits<P>(P(1e30));
            my->eventBits = ~0ull;            // no candidate yet (above every finite step)
            my->corrected = 0;
        }
        __syncthreads();
        const P h = my->dt, t = my->t;
        P next[UD], err[UD];
        bool isfin = true;
        if (ALG == MPGOS_RK4) {
            // PB-STP:2016-2224
            P k[UD], Y[UD];
            rhs(next, X, t);
#pragma unroll
            for (int c = 0; c < UD; c++) Y[c] = fma(next[c] * P(0.5), h, X[c]);
            const P t2 = t + P(0.5) * h;
            rhs(k, Y, t2);
#pragma unroll
            for (int c = 0; c < UD; c++) { next[c] = fma(P(2), k[c], next[c]); Y[c] = fma(k[c] * P(0.5), h, X[c]); }
            rhs(k, Y, t2);
#pragma unroll
            for (int c = 0; c < UD; c++) { next[c] = fma(P(2.0), k[c], next[c]); Y[c] = fma(k[c], h, X[c]); }
            rhs(k, Y, t + h);
#pragma unroll
            for (int c = 0; c < UD; c++) {
                next[c] = fma(h * (next[c] + k[c]), P(1.0 / 6.0), X[c]);
                if (!math::isfinite(next[c])) isfin = false;
            }
        } else {
            // PB-STP:2227-2544
            P k1[UD], k2[UD], k3[UD], k4[UD], k5[UD], k6[UD], Y[UD];
            rhs(k1, X, t);
#pragma unroll
            for (int c = 0; c < UD; c++) Y[c] = fma(k1[c] * P(1.0 / 5.0), h, X[c]);
            rhs(k2, Y, fma(P(1.0 / 5.0), h, t));
#pragma unroll
            for (int c = 0; c < UD; c++) Y[c] = fma(fma(k2[c], P(9.0 / 40.0), k1[c] * P(3.0 / 40.0)), h, X[c]);
            rhs(k3, Y, fma(P(3.0 / 10.0), h, t));
#pragma unroll
            for (int c = 0; c < UD; c++)
                Y[c] = fma(fma(k3[c], P(6.0 / 5.0), fma(k2[c], P(-9.0 / 10.0), k1[c] * P(3.0 / 10.0))), h, X[c]);
            rhs(k4, Y, fma(P(3.0 / 5.0), h, t));
#pragma unroll
            for (int c = 0; c < UD; c++)
                Y[c] = fma(fma(k4[c], P(35.0 / 27.0), fma(k3[c], P(-70.0 / 27.0), fma(k2[c], P(5.0 / 2.0), k1[c] * P(-11.0 / 54.0)))), h, X[c]);
            rhs(k5, Y, t + h);
#pragma unroll
            for (int c = 0; c < UD; c++)
                Y[c] = fma(fma(k5[c], P(253.0 / 4096.0), fma(k4[c], P(44275.0 / 110592.0), fma(k3[c], P(575.0 / 13824.0),
                           fma(k2[c], P(175.0 / 512.0), k1[c] * P(1631.0 / 55296.0))))), h, X[c]);
            rhs(k6, Y, fma(P(7.0 / 8.0), h, t));
#pragma unroll
            for (int c = 0; c < UD; c++) {
                next[c] = fma(fma(k6[c], P(512.0 / 1771.0), fma(k4[c], P(125.0 / 594.0), fma(k3[c], P(250.0 / 621.0), k1[c] * P(37.0 / 378.0)))), h, X[c]);
                P e = k1[c] * P(37.0 / 378.0 - 2825.0 / 27648.0);
                e = fma(k3[c], P(250.0 / 621.0 - 18575.0 / 48384.0), e);
                e = fma(k4[c], P(125.0 / 594.0 - 13525.0 / 55296.0), e);
                e = fma(k5[c], P(0.0 - 277.0 / 14336.0), e);
                e = fma(k6[c], P(512.0 / 1771.0 - 1.0 / 4.0), e);
                err[c] = fma(h, math::fabs(e), P(1e-18));
                if (!math::isfinite(next[c]) || !math::isfinite(err[c])) isfin = false;
            }
        }
        if (lane && my->term == 0 && !isfin) my->isfin = 0;
        // ---- controller (PB-CTL:424-542)
        if (ALG == MPGOS_RKCK45) {
            if (lane && my->term == 0) {
                P R = P(1e30);
                int upd = 1;
#pragma unroll
                for (int c = 0; c < UD; c++) {
                    const P tol = dmax<P>(a.rtol[c] * dmax<P>(math::fabs(next[c]), math::fabs(X[c])), a.atol[c]);
                    upd = upd & (err[c] < tol);
                    R = dmin<P>(R, tol / err[c]);
                }
                atomicAnd(&my->update, upd);
                atomicMin(&my->ratioBits, order_bits<P>(R));
            }
            __syncthreads();
            if (leader && my->term == 0) {
                const P R = from_order_bits<P>(my->ratioBits);
                P fac;
                if (my->update == 1) fac = P(0.8) * math::pow(R, P(1.0 / 5.0));
                else fac = P(0.8) * math::pow(R, P(1.0 / 4.0));
                if (!math::isfinite(fac)) my->isfin = 0;
                if (my->isfin == 0) {
                    fac = a.shrink;
                    my->update = 0;
                    if (my->dt < a.dtmin * P(1.01)) { my->term = 1; my->status |= MPGOS_STATUS_NONFINITE_MIN_STEP; }
                } else if (my->dt < a.dtmin * P(1.01)) {
                    my->update = 1;
                    my->status |= MPGOS_STATUS_MIN_STEP_REACHED;
                }
                fac = dmin<P>(fac, a.grow);
                fac = dmax<P>(fac, a.shrink);
                P nd = my->dt * fac;
                nd = dmin<P>(nd, a.dtmax);
                nd = dmax<P>(nd, a.dtmin);
                my->newDt = nd;
            }
        } else {
            __syncthreads();
            if (leader && my->term == 0) {
                if (my->isfin == 0) { my->term = 1; my->update = 0; my->status |= MPGOS_STATUS_NONFINITE_RK4; }
                my->newDt = a.dt0;
            }
        }
        __syncthreads();
        // ---- events (PB-SOL:2618-2655, PB-EVT:215-268)
        P EFn[Pos<NE>::value];
        if (NE > 0) {
            P tn = my->t + my->dt;
            // the event function of the candidate state, at the end of the step
            if (lane && my->term == 0)
                Sys::template EventFunction<P>(gsys, u, EFn, tn, my->newDt, my->TD, next, uPAR, my->sPAR, gPAR, gPARi, uACC,
                                               uACCi, my->sACC, my->sACCi);
            if (lane && my->update == 1 && my->term == 0) {
                bool corr = false;
                P ev = my->dt;
#pragma unroll
                for (int i = 0; i < NE; i++) {
                    const P tol = a.etol[i];
                    if ((EF[i] > tol && EFn[i] < -tol && a.edir[i] <= 0) || (EF[i] < -tol && EFn[i] > tol && a.edir[i] >= 0)) {
                        ev = dmin<P>(ev, -EF[i] / (EFn[i] - EF[i]) * my->dt);
                        corr = true;
                    }
                }
                if (corr) {
                    atomicMin(&my->eventBits, order_bits<P>(ev));
                    atomicOr(&my->corrected, 1);
                }
            }
            __syncthreads();
            if (leader && my->corrected == 1) {
                const P ev = dmin<P>(my->dt, from_order_bits<P>(my->eventBits));   // corrected: one candidate at least
                if (ev < a.dtmin) my->status |= MPGOS_STATUS_EVENT_OMITTED;
                else { my->newDt = ev; my->update = 0; }
            }
            __syncthreads();
        }
        // ---- update (PB-SOL:2670-2681)
        const int upd = my->update;
        if (lane && upd == 1) {
#pragma unroll
            for (int c = 0; c < UD; c++) X[c] = next[c];
        }
        __syncthreads();
        if (leader && upd == 1) my->t += my->dt;
        __syncthreads();
        if (lane && upd == 1) Sys::template ActionAfterSuccessfulTimeStep<P>(gsys, u, my->udt, MPGOS_PB_USER(my->t, my->dt));
        __syncthreads();
        if (NE > 0) {
            if (lane && upd == 1) {
#pragma unroll
                for (int i = 0; i < NE; i++) {
                    const P tol = a.etol[i];
                    if ((EF[i] > tol && math::fabs(EFn[i]) < tol && a.edir[i] <= 0) ||
                        (EF[i] < -tol && math::fabs(EFn[i]) < tol && a.edir[i] >= 0))
                        Sys::template ActionAfterEventDetection<P>(gsys, u, i, my->udt, MPGOS_PB_USER(my->t, my->dt));
                }
            }
            __syncthreads();
            if (lane && upd == 1) Sys::template EventFunction<P>(gsys, u, EFn, MPGOS_PB_USER(my->t, my->newDt));
            __syncthreads();
            if (lane && upd == 1) {
#pragma unroll
                for (int i = 0; i < NE; i++) EF[i] = EFn[i];
            }
        }
        // ---- dense output (PB-DO:312-395)
        if (a.NDO > 0) {
            if (leader) {
                if (upd == 1) {
                    int s = (my->DOidx < a.NDO && my->doT < my->t && my->skipped >= a.doFreq - 1) ? 1 : 0;
                    if (my->DOidx < a.NDO && (my->end == 1 || my->udt == 1)) s = 1;
                    my->updDO = s;
                } else {
                    my->updDO = 0;
                }
            }
            __syncthreads();
            if (lane && my->updDO == 1 && upd == 1) {
                const int k = my->DOidx;
#pragma unroll
                for (int c = 0; c < UD; c++) a.DOx[gu + c * NU + (size_t)k * NU * UD] = X[c];
                if (u == 0) a.DOt[gsys + (size_t)k * NS] = my->t;
            }
            __syncthreads();
            if (leader && upd == 1) {
                if (my->updDO == 1) {
                    my->DOidx++;
                    my->skipped = 0;
                    my->doT = dmin<P>(my->t + a.doMinDt, my->TD[1]);
                } else {
                    my->skipped++;
                }
            }
        }
        // ---- statistics and termination (PB-SOL:2797-2807)
        if (leader && my->term == 0) {
            if (upd == 1) {
                my->nacc++;
                if (my->end == 1 || my->udt == 1) { my->term = 1; my->update = 0; }
            } else {
                my->nrej++;
            }
        }
        __syncthreads();
        int done = 1;
        for (int s = 0; s < a.SPB; s++) done &= sys[s].term;
        if (done) break;
        __syncthreads();
    }
    __syncthreads();
    // ---- finalisation and write-back (PB-SOL:2814-2887)
    if (active) Sys::template Finalization<P>(gsys, u, my->DOidx, MPGOS_PB_USER(my->t, my->dt));
    __syncthreads();
    if (exists) {
#pragma unroll
        for (int c = 0; c < UD; c++) a.X[gu + c * NU] = X[c];
#pragma unroll
        for (int c = 0; c < NUA; c++) a.uACC[gu + c * NU] = uACC[c];
#pragma unroll
        for (int c = 0; c < NiUA; c++) a.uACCi[gu + c * NU] = uACCi[c];
    }
    if (leader && active) {
        S& s = sys[lsys];
        a.T[gsys] = s.t;
        a.TD[gsys] = s.TD[0]; a.TD[gsys + NS] = s.TD[1];
        for (int i = 0; i < NSA; i++) a.sACC[gsys + i * NS] = s.sACC[i];
        for (int i = 0; i < NiSA; i++) a.sACCi[gsys + i * NS] = s.sACCi[i];
        a.DOidx[gsys] = s.DOidx;
        a.nacc[gsys] = s.nacc;
        a.nrej[gsys] = s.nrej;
        a.status[gsys] = s.status;
    }
#undef MPGOS_PB_USER
}

template <class Sys, class P, int ALG>
__global__ void __launch_bounds__(kMaxBlock) pb_kernel(Args<Sys, P> a)
{
    pb_body<Sys, P, ALG>(a);
}

} // namespace pb
} // namespace mpgos
