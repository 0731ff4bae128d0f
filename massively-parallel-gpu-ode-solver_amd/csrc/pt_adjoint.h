// Discrete adjoint of the per-thread RK4 / RKCK45 integration (FP64).
//
// No counterpart in the reference: it differentiates the map the reference's
// solver computes (SourceCodes/SingleSystem_PerThread_Solver.cuh:197-394 with
// the steppers of SingleSystem_PerThread_ExplicitRungeKutta_Steppers.cuh),
// taking the accepted step sizes as given (the error controller and event step
// control are not differentiated -- "frozen" step sequence).
//
// Record pass (record_body): the ordinary lane loop, plus a checkpoint
// (x, t, proposed dt, event-function values, accessories and integer
// accessories) before accepted step 0 and after every K-th accepted step, SoA
// in HBM: ckpt[((k*(SD+2+NE+NA+NIA)) + c)*NT + i].  The reverse pass restores
// the accessories with the state, so callbacks and right-hand sides that read
// them replay the recorded values; control parameters and time domains are
// taken as the Solve left them (callbacks that change them during a Solve are
// outside what the adjoint differentiates).
// Reverse pass (reverse_body), one lane per instance, segments last to first:
//   1. restore the segment's checkpoint and re-run the same adaptive loop
//      (controller, rejections and all) -- identical arithmetic, so the same
//      accepted steps -- storing (x_n, t_n, h_n) of each accepted step in a
//      per-lane segment buffer, in LDS when K segment records of the
//      workgroup fit there (per wavefront: lds[(j*(SD+2) + c)*64 + lane]: the records
//      never leave the compute unit), else in HBM seg[((j*(SD+2)) + c)*NT + i],
//      j < K;
//   2. walk the segment backwards (while the next segment down is re-stepped
//      into the record slots already read: two independent chains in one
//      loop body, see reverse_body): recompute the step's stage values from x_n
//      and propagate the adjoint through the stages with the system's
//      vector-Jacobian product (Sys::OdeVJP):
//        x_{n+1} = x_n + h sum_j b_j k_j,  k_j = f(y_j, t_n + c_j h),
//        y_j = x_n + h sum_{l<j} a_jl k_l
//      kb_j = h b_j lam';  for j = s..1: (gy, gp) = VJP(kb_j at y_j);
//      lam += gy; mu += gp; kb_l += h a_jl gy (l < j);  lam_n = lam' + sum gy.
// Event step control is replayed as recorded (the event values are part of
// the checkpoint), so shortened steps recur; user actions at events are not
// differentiated.  Cost per accepted step: one forward step (recompute) +
// stage recompute + one VJP per stage; memory: NT * ceil(N/K) * (SD+2+NE+NA+NIA)
// doubles of checkpoints plus K * (SD+2) doubles of segment records per lane
// (LDS, or HBM when they do not fit).
#pragma once

#include "pt_device.h"

namespace mpgos {
namespace pt {

constexpr int MPGOS_STATUS_CHECKPOINT_OVERFLOW_BIT = 16;

template <class Sys>
struct AdjArgs {
    Args<Sys, double, double> a;   // forward buffers and options (NDO and event log disabled in the reverse pass)
    double* ckpt;
    double* seg;
    double* lam;                   // in: dL/dx(T); out: dL/dx(0)      [SD][NT]
    double* gp;                    // out: dL/dcPAR                    [NCP][NT]
    int K;
    int ckcap;
};

template <class Sys>
constexpr int checkpoint_words() { return Sys::SD + 2 + Sys::NE + Sys::NA + Sys::NIA; }

template <class Sys, int ALG>
__device__ __forceinline__ void store_checkpoint(const AdjArgs<Sys>& A, const Lane<Sys, double, false, ALG>& L, int k)
{
    constexpr int SD = Sys::SD, NE = Sys::NE;
    const int NT = A.a.NT;
    const size_t base = (size_t)k * checkpoint_words<Sys>();
#pragma unroll
    for (int c = 0; c < SD; c++) A.ckpt[(base + c) * NT + L.tid] = L.X[c];
    A.ckpt[(base + SD) * NT + L.tid] = L.t;
    A.ckpt[(base + SD + 1) * NT + L.tid] = L.newDt;
#pragma unroll
    for (int e = 0; e < NE; e++) A.ckpt[(base + SD + 2 + e) * NT + L.tid] = L.EF[e];
#pragma unroll
    for (int i = 0; i < Sys::NA; i++) A.ckpt[(base + SD + 2 + NE + i) * NT + L.tid] = L.ACC[i];
#pragma unroll
    for (int i = 0; i < Sys::NIA; i++) A.ckpt[(base + SD + 2 + NE + Sys::NA + i) * NT + L.tid] = (double)L.ACCi[i];
}

template <class Sys, int ALG>
__device__ __forceinline__ void record_body(const AdjArgs<Sys>& A)
{
    static_assert(offsetof(AdjArgs<Sys>, a) == 0, "Lane's rare paths read the Args at the start of the kernel arguments");
    const Args<Sys, double, double>& a = A.a;
    double* sPAR; int* sPARi;
    stage_shared(a, sPAR, sPARi);
    const int tid = threadIdx.x + blockIdx.x * blockDim.x;
    if (tid >= a.nActive) return;
    Lane<Sys, double, false, ALG> L;
    L.load(a, tid, sPAR, sPARi);
    store_checkpoint<Sys, ALG>(A, L, 0);
    int untilCk = A.K, k = 0;      // accepted steps to the next checkpoint (no division in the loop)
    while (L.term == 0) {
        const int before = L.nacc;
        L.step(a, sPAR, sPARi);
        if (L.nacc != before && --untilCk == 0) {
            untilCk = A.K;
            ++k;
            if (L.term == 0) {
                if (k < A.ckcap) store_checkpoint<Sys, ALG>(A, L, k);
                else L.status |= MPGOS_STATUS_CHECKPOINT_OVERFLOW_BIT;
            }
        }
    }
    L.finish(a, sPAR, sPARi);
}

// ---------------------------------------------------------------- reverse steps
template <class Sys>
struct Vjp {
    static constexpr int SD = Sys::SD, NCP = Sys::NCP;
    template <int NCPA>
    static __device__ __forceinline__ void apply(int tid, int NT, const double* kb, double* y, double T, double* cPAR,
                                                 double* sPAR, int* sPARi, double* ACC, int* ACCi, double* lam, double* mu,
                                                 double* gy)
    {
        double gp[NCPA];
        Sys::template OdeVJP<double>(tid, NT, kb, y, T, cPAR, sPAR, sPARi, ACC, ACCi, gy, gp);
#pragma unroll
        for (int c = 0; c < SD; c++) lam[c] += gy[c];
#pragma unroll
        for (int c = 0; c < NCP; c++) mu[c] += gp[c];
    }
};

// RK4 (stages as in Lane::rk4).  lam: in dL/dx_{n+1}, out dL/dx_n.
template <class Sys>
__device__ __forceinline__ void reverse_rk4(int tid, int NT, const double* x, double t, double h, double* cPAR, double* sPAR,
                                            int* sPARi, double* ACC, int* ACCi, double* lam, double* mu)
{
    using math::fma;
    constexpr int SD = Sys::SD, NCPA = Pos<Sys::NCP>::value;
    double k[SD], y2[SD], y3[SD], y4[SD], xx[SD];
#pragma unroll
    for (int c = 0; c < SD; c++) xx[c] = x[c];
    const double half = 0.5 * h;
    const double sixth = (1.0 / 6.0) * h;
    const double T2 = t + half, T4 = t + h;
    Sys::template OdeFunction<double>(tid, NT, k, xx, t, cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
    for (int c = 0; c < SD; c++) y2[c] = fma(k[c], half, xx[c]);
    Sys::template OdeFunction<double>(tid, NT, k, y2, T2, cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
    for (int c = 0; c < SD; c++) y3[c] = fma(k[c], half, xx[c]);
    Sys::template OdeFunction<double>(tid, NT, k, y3, T2, cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
    for (int c = 0; c < SD; c++) y4[c] = fma(k[c], h, xx[c]);

    double kb[SD], gy[SD], lo[SD];
#pragma unroll
    for (int c = 0; c < SD; c++) lo[c] = lam[c];
    // stage 4 (weight 1/6)
#pragma unroll
    for (int c = 0; c < SD; c++) kb[c] = sixth * lo[c];
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb, y4, T4, cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
    // stage 3 (weight 2/6, plus h * dy4/dk3)
#pragma unroll
    for (int c = 0; c < SD; c++) kb[c] = fma(2.0 * sixth, lo[c], h * gy[c]);
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb, y3, T2, cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
    // stage 2
#pragma unroll
    for (int c = 0; c < SD; c++) kb[c] = fma(2.0 * sixth, lo[c], half * gy[c]);
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb, y2, T2, cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
    // stage 1
#pragma unroll
    for (int c = 0; c < SD; c++) kb[c] = fma(sixth, lo[c], half * gy[c]);
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb, xx, t, cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
}

// Cash-Karp (stages as in Lane::rkck45, 5th-order weights b).
template <class Sys>
__device__ __forceinline__ void reverse_rkck45(int tid, int NT, const double* x, double t, double h, double* cPAR,
                                               double* sPAR, int* sPARi, double* ACC, int* ACCi, double* lam, double* mu)
{
    using math::fma;
    constexpr int SD = Sys::SD, NCPA = Pos<Sys::NCP>::value;
    constexpr double a21 = 1.0 / 5.0;
    constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
    constexpr double a41 = 3.0 / 10.0, a42 = -9.0 / 10.0, a43 = 6.0 / 5.0;
    constexpr double a51 = -11.0 / 54.0, a52 = 5.0 / 2.0, a53 = -70.0 / 27.0, a54 = 35.0 / 27.0;
    constexpr double a61 = 1631.0 / 55296.0, a62 = 175.0 / 512.0, a63 = 575.0 / 13824.0, a64 = 44275.0 / 110592.0,
                     a65 = 253.0 / 4096.0;
    constexpr double b1 = 37.0 / 378.0, b3 = 250.0 / 621.0, b4 = 125.0 / 594.0, b6 = 512.0 / 1771.0;

    double y[6][SD], k[5][SD];
    const double T[6] = {t, fma(h, 1.0 / 5.0, t), fma(h, 3.0 / 10.0, t), fma(h, 3.0 / 5.0, t), t + h, fma(h, 7.0 / 8.0, t)};
#pragma unroll
    for (int c = 0; c < SD; c++) y[0][c] = x[c];
    Sys::template OdeFunction<double>(tid, NT, k[0], y[0], T[0], cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
    for (int c = 0; c < SD; c++) y[1][c] = fma(h, a21 * k[0][c], x[c]);
    Sys::template OdeFunction<double>(tid, NT, k[1], y[1], T[1], cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
    for (int c = 0; c < SD; c++) y[2][c] = fma(h, fma(a32, k[1][c], a31 * k[0][c]), x[c]);
    Sys::template OdeFunction<double>(tid, NT, k[2], y[2], T[2], cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
    for (int c = 0; c < SD; c++) y[3][c] = fma(h, fma(a43, k[2][c], fma(a42, k[1][c], a41 * k[0][c])), x[c]);
    Sys::template OdeFunction<double>(tid, NT, k[3], y[3], T[3], cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
    for (int c = 0; c < SD; c++)
        y[4][c] = fma(h, fma(a54, k[3][c], fma(a53, k[2][c], fma(a52, k[1][c], a51 * k[0][c]))), x[c]);
    Sys::template OdeFunction<double>(tid, NT, k[4], y[4], T[4], cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
    for (int c = 0; c < SD; c++)
        y[5][c] = fma(h, fma(a65, k[4][c], fma(a64, k[3][c], fma(a63, k[2][c], fma(a62, k[1][c], a61 * k[0][c])))), x[c]);

    // adjoints of the stage derivatives, seeded by the output weights
    double kb[6][SD], gy[SD];
#pragma unroll
    for (int c = 0; c < SD; c++) {
        const double hl = h * lam[c];
        kb[0][c] = b1 * hl; kb[1][c] = 0.0; kb[2][c] = b3 * hl; kb[3][c] = b4 * hl; kb[4][c] = 0.0; kb[5][c] = b6 * hl;
    }
    // stage 6 -> contributes to k1..k5
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb[5], y[5], T[5], cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
#pragma unroll
    for (int c = 0; c < SD; c++) {
        const double g = h * gy[c];
        kb[0][c] = fma(a61, g, kb[0][c]); kb[1][c] = fma(a62, g, kb[1][c]); kb[2][c] = fma(a63, g, kb[2][c]);
        kb[3][c] = fma(a64, g, kb[3][c]); kb[4][c] = fma(a65, g, kb[4][c]);
    }
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb[4], y[4], T[4], cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
#pragma unroll
    for (int c = 0; c < SD; c++) {
        const double g = h * gy[c];
        kb[0][c] = fma(a51, g, kb[0][c]); kb[1][c] = fma(a52, g, kb[1][c]); kb[2][c] = fma(a53, g, kb[2][c]);
        kb[3][c] = fma(a54, g, kb[3][c]);
    }
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb[3], y[3], T[3], cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
#pragma unroll
    for (int c = 0; c < SD; c++) {
        const double g = h * gy[c];
        kb[0][c] = fma(a41, g, kb[0][c]); kb[1][c] = fma(a42, g, kb[1][c]); kb[2][c] = fma(a43, g, kb[2][c]);
    }
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb[2], y[2], T[2], cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
#pragma unroll
    for (int c = 0; c < SD; c++) {
        const double g = h * gy[c];
        kb[0][c] = fma(a31, g, kb[0][c]); kb[1][c] = fma(a32, g, kb[1][c]);
    }
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb[1], y[1], T[1], cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
#pragma unroll
    for (int c = 0; c < SD; c++) kb[0][c] = fma(a21, h * gy[c], kb[0][c]);
    Vjp<Sys>::template apply<NCPA>(tid, NT, kb[0], y[0], T[0], cPAR, sPAR, sPARi, ACC, ACCi, lam, mu, gy);
}

// LDS bytes of the shared-parameter staging (16-byte aligned); the segment
// records of the LDS form follow.
template <class Sys>
__host__ __device__ constexpr size_t adjoint_shared_bytes(int preferShared)
{
    return preferShared ? (((size_t)Sys::NSP * sizeof(double) + (size_t)Sys::NISP * sizeof(int)) + 15) & ~(size_t)15 : 0;
}

template <class Sys, int ALG, bool LDS_SEG>
__device__ __forceinline__ void reverse_body(const AdjArgs<Sys>& A)
{
    constexpr int SD = Sys::SD, NCP = Sys::NCP, NA = Sys::NA, NIA = Sys::NIA;
    static_assert(offsetof(AdjArgs<Sys>, a) == 0, "Lane's rare paths read the Args at the start of the kernel arguments");
    const Args<Sys, double, double>& a = A.a;
    double* sPAR; int* sPARi;
    stage_shared(a, sPAR, sPARi);
    const int tid = threadIdx.x + blockIdx.x * blockDim.x;
    if (tid >= a.nActive) return;
    const int NT = a.NT;
    // segment record word w of this lane: in LDS one region per wavefront,
    // word-major with a lane stride of 64 doubles (a record's words are then
    // constant offsets from one address); in HBM word-major over all lanes
    extern __shared__ __align__(16) unsigned char mpgos_dyn_lds[];
    double* const segBase =
        LDS_SEG ? reinterpret_cast<double*>(mpgos_dyn_lds + adjoint_shared_bytes<Sys>(a.preferShared)) +
                      (size_t)(threadIdx.x >> 6) * ((size_t)A.K * (SD + 2) * 64) + (threadIdx.x & 63)
                : A.seg + tid;
    auto seg = [&](size_t w) -> double& {
        if constexpr (LDS_SEG) return segBase[w * 64];
        else return segBase[w * NT];
    };
    const long long total = a.nacc[tid];
    double lam[SD], mu[Pos<NCP>::value];
#pragma unroll
    for (int c = 0; c < SD; c++) lam[c] = A.lam[c * NT + tid];
#pragma unroll
    for (int c = 0; c < NCP; c++) mu[c] = 0.0;

    Lane<Sys, double, false, ALG> L;
    L.tid = tid;
#pragma unroll
    for (int i = 0; i < 2; i++) L.TD[i] = a.TD[tid + i * NT];
#pragma unroll
    for (int i = 0; i < NCP; i++) L.cPAR[i] = a.cPAR[tid + i * NT];

    const bool ok = (a.status[tid] & MPGOS_STATUS_CHECKPOINT_OVERFLOW_BIT) == 0;
    const long long nseg = ok ? (total + A.K - 1) / A.K : 0;
    const int K = A.K;
    // Restore segment s's checkpoint into the lane.
    auto restore = [&](long long s) {
        const size_t cb = (size_t)s * checkpoint_words<Sys>();
#pragma unroll
        for (int c = 0; c < SD; c++) L.X[c] = A.ckpt[(cb + c) * NT + tid];
        L.t = A.ckpt[(cb + SD) * NT + tid];
        L.newDt = A.ckpt[(cb + SD + 1) * NT + tid];
#pragma unroll
        for (int e = 0; e < Sys::NE; e++) L.EF[e] = A.ckpt[(cb + SD + 2 + e) * NT + tid];
#pragma unroll
        for (int i = 0; i < NA; i++) L.ACC[i] = A.ckpt[(cb + SD + 2 + Sys::NE + i) * NT + tid];
#pragma unroll
        for (int i = 0; i < NIA; i++) L.ACCi[i] = (int)A.ckpt[(cb + SD + 2 + Sys::NE + NA + i) * NT + tid];
        L.dt = L.newDt;
        L.tc = 0.0;
        L.term = 0; L.udt = 0; L.nacc = 0; L.nrej = 0; L.status = 0; L.nev = 0;
        L.doIdx = 0; L.skipped = 0; L.doT = 0.0;
    };
    // accepted steps of segment s
    auto count = [&](long long s) { return (int)((total - s * K) < K ? (total - s * K) : K); };
    // One record buffer of K slots serves two segments at once: segment s is
    // differentiated (its records read last to first) while segment s - 1 is
    // re-stepped into the slots already read.  Segments alternate the slot
    // order -- step j in slot j, or in slot K - 1 - j -- so that the m-th
    // record written (step m of segment s - 1) goes to the slot the m-th read
    // (step cnt - 1 - m of segment s) has just freed; rejected attempts write
    // nothing, so writes never overtake reads.
    auto slot = [&](bool flip, int j) { return flip ? K - 1 - j : j; };
    // One attempted re-step of the lane; an accepted step stores (x_n, t_n, h_n)
    // of step j in its slot and advances j.
    auto attempt = [&](bool flip, int& j) {
        double xp[SD];
#pragma unroll
        for (int c = 0; c < SD; c++) xp[c] = L.X[c];
        const double tp = L.t;
        const int before = L.nacc;
        // the Cash-Karp step with its controller holds more uniform values
        // than fit in SGPRs next to the adjoint chain: re-read the argument
        // block at every attempt instead of spilling it
        if constexpr (ALG == MPGOS_RKCK45) L.template step<false>(kernel_args<Args<Sys, double, double>>(), sPAR, sPARi);
        else L.template step<false>(a, sPAR, sPARi);
        if (L.nacc != before) {
            const size_t sb = (size_t)slot(flip, j) * (SD + 2);
#pragma unroll
            for (int c = 0; c < SD; c++) seg(sb + c) = xp[c];
            seg(sb + SD) = tp;
            seg(sb + SD + 1) = L.dt;
            j++;
        }
    };
    // Adjoint through step j of the segment whose records use slot order flip,
    // with the parameters and accessories the lane held at the end of that
    // segment's re-step (kept apart while the next segment is re-stepped)
    double bPAR[Pos<NCP>::value], bACC[Pos<NA>::value];
    int bACCi[Pos<NIA>::value];
    auto keep = [&]() {
#pragma unroll
        for (int i = 0; i < NCP; i++) bPAR[i] = L.cPAR[i];
#pragma unroll
        for (int i = 0; i < NA; i++) bACC[i] = L.ACC[i];
#pragma unroll
        for (int i = 0; i < NIA; i++) bACCi[i] = L.ACCi[i];
    };
    auto back = [&](bool flip, int j) {
        const size_t sb = (size_t)slot(flip, j) * (SD + 2);
        double x[SD];
#pragma unroll
        for (int c = 0; c < SD; c++) x[c] = seg(sb + c);
        const double tn = seg(sb + SD), hn = seg(sb + SD + 1);
        if (ALG == MPGOS_RK4)
            reverse_rk4<Sys>(tid, NT, x, tn, hn, bPAR, sPAR, sPARi, bACC, bACCi, lam, mu);
        else
            reverse_rkck45<Sys>(tid, NT, x, tn, hn, bPAR, sPAR, sPARi, bACC, bACCi, lam, mu);
    };
    if (nseg > 0) {
        // the last segment alone
        bool flip = false;
        restore(nseg - 1);
        int have = 0;
        const int cntLast = count(nseg - 1);
        while (have < cntLast && L.term == 0) attempt(flip, have);
        for (long long s = nseg - 1; s >= 0; s--) {
            // `have` records of segment s are in the buffer (slot order flip)
            keep();
            if (s == 0) {
                for (int jj = have - 1; jj >= 0; jj--) back(flip, jj);
                break;
            }
            // segment s - 1 is full (K accepted steps, at least K attempts),
            // so the first `have` <= K iterations always hold one attempt and
            // one adjoint step: two independent chains in one block
            restore(s - 1);
            int next = 0;
            for (int m = 0; m < have; m++) {
                back(flip, have - 1 - m);
                attempt(!flip, next);
            }
            while (next < K && L.term == 0) attempt(!flip, next);
            have = next;
            flip = !flip;
        }
    }
#pragma unroll
    for (int c = 0; c < SD; c++) A.lam[c * NT + tid] = ok ? lam[c] : __builtin_nan("");
#pragma unroll
    for (int c = 0; c < NCP; c++) A.gp[c * NT + tid] = ok ? mu[c] : __builtin_nan("");
}

template <class Sys, int ALG>
__global__ void __launch_bounds__(kMaxBlock) adjoint_record_kernel(AdjArgs<Sys> A) { record_body<Sys, ALG>(A); }

template <class Sys, int ALG, bool LDS_SEG>
__global__ void __launch_bounds__(kMaxBlock) adjoint_reverse_kernel(AdjArgs<Sys> A) { reverse_body<Sys, ALG, LDS_SEG>(A); }

} // namespace pt
} // namespace mpgos
