// Discrete adjoint of the per-block RK4 / RKCK45 integration (FP64).
//
// No counterpart in the reference: it differentiates the map the reference's
// per-block solver computes (SourceCodes/CoupledSystmes_PerBlock_Solver.cuh:
// 2214-2888 with the steppers of
// CoupledSystems_PerBlock_ExplicitRungeKutta_Steppers.cuh:2016-2544 and the
// coupling terms at :2634-2713), with the accepted step sizes taken as given
// (frozen step sequence; the error controller and event step control are not
// differentiated; user actions that change the state are not replayed).
//
// Record pass (pb_record_kernel in pb_device.h): the forward integration
// storing the unit states every K accepted steps and (t_n, h_n) of every
// accepted step.
// Reverse pass (pb_reverse_kernel), lane per (system, unit), SPB systems per
// workgroup, segments last to first:
//   1. restore the segment's checkpoint and recompute its accepted steps with
//      the recorded (t_n, h_n) -- the stage arithmetic of the forward pass, so
//      the states are the recorded ones bit for bit -- keeping the state at
//      the start of each step in a segment buffer seg[(j * UD + c) * NU + unit];
//   2. walk the segment backwards: recompute the step's stage values Y_j,
//      their coupling values v_j and factors CPF_j, and propagate the adjoint
//      through the stages, next = x + h sum_j b_j k_j, Y_j = x + h sum_l a_jl k_l,
//      k_j = F(Y_j) + s CPF_j (M CPT(Y))_j:
//        kb_j = h b_j lam';  for j = s..1:
//          w = s CPF_j kb_j[ci]                       (per unit, into LDS)
//          kbCPT = M^T w,  kbCPF = s v_j kb_j[ci]      (transposed coupling)
//          (gy, gU, gS) = Sys::OdeVJP(kb_j, kbCPT, kbCPF at Y_j)
//          lam += gy;  kb_l += h a_jl gy (l < j);  muU += gU;  muS += gS
// The system-parameter gradient is summed over the units of a system in a
// fixed order at the end (deterministic).  Cost per accepted step: one
// recompute of the step (segment pass) + stage recompute + one transposed
// coupling product and one VJP per stage.
#pragma once

#include "pb_device.h"

namespace mpgos {
namespace pb {

template <class Sys>
struct AdjArgs {
    double* seg;          // [K][UD][NU]
    double* lam;          // in: dL/dx(end), out: dL/dx(start)      [UD][NU]
    double* gU;           // out: dL/d(unit parameters)             [NUP][NU]
    double* gS;           // out: dL/d(system parameters)           [NSP][NS]
    const double* cMatT;  // transposed coupling matrices in the stored form of cMat
};

// sum over the matrix row u of M_i (any storage form) against vals[col * NCp + i]
template <class Sys>
__device__ __forceinline__ double coupling_dot(const Args<Sys, double>& a, const double* M, const double* vals, int u, int i)
{
    constexpr int NCp = Pos<Sys::NC>::value;
    const int UPS = a.UPS;
    double v = 0.0;
    if (a.CCI == 0 && a.CBW == 0) {
        const double* Mi = M + (size_t)i * UPS * UPS + u;
        for (int col = 0; col < UPS; col++) v = math::fma(Mi[(size_t)col * UPS], vals[col * NCp + i], v);
    } else {
        const int W = a.CBW > 0 ? 2 * a.CBW + 1 : UPS;
        for (int d = 0; d < W; d++) {
            int col = u + d - a.CBW;
            if (col >= UPS) col -= UPS;
            if (col < 0) col += UPS;
            const double m = a.CCI == 0 ? M[(size_t)i * UPS * W + u + (size_t)d * UPS] : M[(size_t)i * W + d];
            v = math::fma(m, vals[col * NCp + i], v);
        }
    }
    return v;
}

// System scope of the reverse pass (LDS) and the LDS layout: per-system
// scope, two coupling-term buffers, two adjoint-weight buffers, the
// system-parameter gradient partials, two block-wide counters.
template <class Sys>
struct RevScope {
    double sPAR[Pos<Sys::NSP>::value]; double sACC[Pos<Sys::NSA>::value]; double strength[Pos<Sys::NC>::value];
    int sACCi[Pos<Sys::NiSA>::value];
    long long nacc;
    int nseg, len, k, status;
};
struct RevLayout { size_t cpt, wb, red, blk, total; };

template <class Sys>
__host__ __device__ inline RevLayout reverse_layout(int SPB, int UPS)
{
    constexpr int NCp = Pos<Sys::NC>::value, NSPp = Pos<Sys::NSP>::value;
    RevLayout L;
    size_t o = ((size_t)SPB * sizeof(RevScope<Sys>) + 15) & ~(size_t)15;
    L.cpt = o;
    o += 2 * (size_t)SPB * UPS * NCp * sizeof(double);
    L.wb = o;
    o += 2 * (size_t)SPB * UPS * NCp * sizeof(double);
    L.red = o;
    o += (size_t)SPB * UPS * NSPp * sizeof(double);
    L.blk = o;
    o += 2 * sizeof(int);
    L.total = o;
    return L;
}

template <class Sys, int ALG>
__device__ void pb_reverse_body(const Args<Sys, double>& a, const AdjArgs<Sys>& A)
{
    using P = double;
    constexpr int UD = Sys::UD, NC = Sys::NC, NUP = Sys::NUP, NSP = Sys::NSP, NGP = Sys::NGP, NiGP = Sys::NiGP;
    constexpr int NUA = Sys::NUA, NiUA = Sys::NiUA, NSA = Sys::NSA, NiSA = Sys::NiSA;
    constexpr int NCp = Pos<NC>::value, NUPp = Pos<NUP>::value, NSPp = Pos<NSP>::value;
    constexpr int ST = ALG == MPGOS_RK4 ? 4 : 6;       // stages
    using math::fma;

    using Scope = RevScope<Sys>;
    extern __shared__ __align__(16) unsigned char mpgos_pb_lds[];
    const int UPS = a.UPS, NS = a.NS, SPB = a.SPB, K = a.recK;
    const RevLayout L = reverse_layout<Sys>(SPB, UPS);
    Scope* sys = reinterpret_cast<Scope*>(mpgos_pb_lds);
    P* cpt = reinterpret_cast<P*>(mpgos_pb_lds + L.cpt);
    P* wb = reinterpret_cast<P*>(mpgos_pb_lds + L.wb);
    P* red = reinterpret_cast<P*>(mpgos_pb_lds + L.red);            // [SPB][UPS][NSPp]
    int* blk = reinterpret_cast<int*>(mpgos_pb_lds + L.blk);         // [0] rounds, [1] longest segment
    const int stride = SPB * UPS * NCp;

    const int tid = threadIdx.x;
    const int lsys = tid / UPS;
    const int u = tid - lsys * UPS;
    const bool lane = lsys < SPB;
    const bool leader = lane && u == 0;
    const int gsys = blockIdx.x * SPB + lsys;
    const bool active = lane && gsys < a.nActive && gsys < NS;
    const size_t NU = (size_t)NS * UPS;
    const size_t gu = (size_t)gsys * UPS + u;
    Scope* my = lane ? &sys[lsys] : &sys[0];

    int cIndex[NCp];
#pragma unroll
    for (int i = 0; i < NC; i++) cIndex[i] = a.cIndex[i];
    if (tid < 2) blk[tid] = 0;
    if (leader) {
        my->status = 0;
        my->nseg = 0;
        if (active) {
            for (int i = 0; i < NSP; i++) my->sPAR[i] = a.sPAR[gsys + i * NS];
            for (int i = 0; i < NSA; i++) my->sACC[i] = a.sACC[gsys + i * NS];
            for (int i = 0; i < NiSA; i++) my->sACCi[i] = a.sACCi[gsys + i * NS];
            for (int i = 0; i < NC; i++) my->strength[i] = a.strength[gsys + i * NS];
            my->nacc = a.nacc[gsys];
            my->status = a.status[gsys];
            const bool ok = (my->status & MPGOS_STATUS_CHECKPOINT_OVERFLOW) == 0;
            my->nseg = ok ? (int)((my->nacc + K - 1) / K) : 0;
        }
    }
    P uPAR[NUPp], uACC[Pos<NUA>::value], lam[UD], muU[NUPp], muS[NSPp];
    int uACCi[Pos<NiUA>::value];
#pragma unroll
    for (int c = 0; c < NUP; c++) { uPAR[c] = active ? a.uPAR[gu + c * NU] : P(0); muU[c] = P(0); }
#pragma unroll
    for (int c = 0; c < NUA; c++) uACC[c] = active ? a.uACC[gu + c * NU] : P(0);
#pragma unroll
    for (int c = 0; c < NiUA; c++) uACCi[c] = active ? a.uACCi[gu + c * NU] : 0;
#pragma unroll
    for (int c = 0; c < UD; c++) lam[c] = active ? A.lam[gu + c * NU] : P(0);
#pragma unroll
    for (int c = 0; c < NSP; c++) muS[c] = P(0);
    __syncthreads();
    if (leader) atomicMax(&blk[0], my->nseg);
    __syncthreads();
    const int rounds = blk[0];

    P* gPAR = a.gPAR;
    int* gPARi = a.gPARi;
    int buf = 0, wbuf = 0;
    // forward right-hand sides (with coupling) at Y, time T; also the coupling
    // values v and factors CPF of this unit
    auto rhs = [&](bool run, P* F, P* Y, P T, P* v, P* CPF) {
        P* c = cpt + (size_t)buf * stride;
        if (run)
            Sys::template OdeFunction<P>(gsys, u, F, Y, T, uPAR, my->sPAR, gPAR, gPARi, uACC, uACCi, my->sACC, my->sACCi,
                                         c + (size_t)(lsys * UPS + u) * NCp, CPF);
        __syncthreads();
        if (run) {
#pragma unroll
            for (int i = 0; i < NC; i++) {
                v[i] = coupling_dot<Sys>(a, a.cMat, c + (size_t)lsys * UPS * NCp, u, i);
                Unit<Sys, P>::add(cIndex[i], my->strength[i] * CPF[i], v[i], F);
            }
        }
        buf ^= 1;
    };
    // adjoint of one stage: kb (adjoint of the stage derivative) at Y -> gy
    auto vjp = [&](bool run, const P* kb, P* Y, P T, const P* v, const P* CPF, P* gy) {
        P* w = wb + (size_t)wbuf * stride;
        P kbCPF[NCp], kbCPT[NCp];
        if (run) {
#pragma unroll
            for (int i = 0; i < NC; i++) {
                P g = P(0);
#pragma unroll
                for (int c = 0; c < UD; c++)
                    if (c == cIndex[i]) g = kb[c];
                w[(size_t)(lsys * UPS + u) * NCp + i] = my->strength[i] * CPF[i] * g;
                kbCPF[i] = my->strength[i] * v[i] * g;
            }
        }
        __syncthreads();
        if (run) {
#pragma unroll
            for (int i = 0; i < NC; i++) kbCPT[i] = coupling_dot<Sys>(a, A.cMatT, w + (size_t)lsys * UPS * NCp, u, i);
            P gU[NUPp], gS[NSPp];
            Sys::template OdeVJP<P>(gsys, u, kb, kbCPT, kbCPF, Y, T, uPAR, my->sPAR, gPAR, gPARi, uACC, uACCi, my->sACC,
                                    my->sACCi, gy, gU, gS);
#pragma unroll
            for (int c = 0; c < NUP; c++) muU[c] += gU[c];
#pragma unroll
            for (int c = 0; c < NSP; c++) muS[c] += gS[c];
        } else {
#pragma unroll
            for (int c = 0; c < UD; c++) gy[c] = P(0);
        }
        wbuf ^= 1;
    };

    // Butcher coefficients of the steppers as the forward pass evaluates them
    // (pb_device.h): Y_j = x + h sum_l A[j][l] k_l, next = x + h sum_j B[j] k_j
    constexpr int S6 = 6;
    P Acoef[S6][S6] = {};
    P Bcoef[S6] = {};
    P Ccoef[S6] = {};
    if (ALG == MPGOS_RK4) {
        Acoef[1][0] = 0.5; Acoef[2][1] = 0.5; Acoef[3][2] = 1.0;
        Bcoef[0] = 1.0 / 6.0; Bcoef[1] = 2.0 / 6.0; Bcoef[2] = 2.0 / 6.0; Bcoef[3] = 1.0 / 6.0;
    } else {
        Acoef[1][0] = 1.0 / 5.0;
        Acoef[2][0] = 3.0 / 40.0; Acoef[2][1] = 9.0 / 40.0;
        Acoef[3][0] = 3.0 / 10.0; Acoef[3][1] = -9.0 / 10.0; Acoef[3][2] = 6.0 / 5.0;
        Acoef[4][0] = -11.0 / 54.0; Acoef[4][1] = 5.0 / 2.0; Acoef[4][2] = -70.0 / 27.0; Acoef[4][3] = 35.0 / 27.0;
        Acoef[5][0] = 1631.0 / 55296.0; Acoef[5][1] = 175.0 / 512.0; Acoef[5][2] = 575.0 / 13824.0;
        Acoef[5][3] = 44275.0 / 110592.0; Acoef[5][4] = 253.0 / 4096.0;
        Bcoef[0] = 37.0 / 378.0; Bcoef[2] = 250.0 / 621.0; Bcoef[3] = 125.0 / 594.0; Bcoef[5] = 512.0 / 1771.0;
        (void)Ccoef;
    }

    // one forward step with the forward pass's exact arithmetic (pb_device.h);
    // optionally keeps stage values, coupling values and factors
    auto step = [&](bool run, P* X, P t, P h, P (*Ys)[UD], P (*vs)[NCp], P (*cpfs)[NCp]) {
        P next[UD];
        if (ALG == MPGOS_RK4) {
            P k[UD], Y[UD];
#pragma unroll
            for (int c = 0; c < UD; c++) Ys[0][c] = X[c];
            rhs(run, next, X, t, vs[0], cpfs[0]);
#pragma unroll
            for (int c = 0; c < UD; c++) Y[c] = fma(next[c] * P(0.5), h, X[c]);
            const P t2 = t + P(0.5) * h;
#pragma unroll
            for (int c = 0; c < UD; c++) Ys[1][c] = Y[c];
            rhs(run, k, Y, t2, vs[1], cpfs[1]);
#pragma unroll
            for (int c = 0; c < UD; c++) { next[c] = fma(P(2), k[c], next[c]); Y[c] = fma(k[c] * P(0.5), h, X[c]); }
#pragma unroll
            for (int c = 0; c < UD; c++) Ys[2][c] = Y[c];
            rhs(run, k, Y, t2, vs[2], cpfs[2]);
#pragma unroll
            for (int c = 0; c < UD; c++) { next[c] = fma(P(2.0), k[c], next[c]); Y[c] = fma(k[c], h, X[c]); }
#pragma unroll
            for (int c = 0; c < UD; c++) Ys[3][c] = Y[c];
            rhs(run, k, Y, t + h, vs[3], cpfs[3]);
#pragma unroll
            for (int c = 0; c < UD; c++) next[c] = fma(h * (next[c] + k[c]), P(1.0 / 6.0), X[c]);
        } else {
            P k1[UD], k2[UD], k3[UD], k4[UD], k5[UD], k6[UD], Y[UD];
#pragma unroll
            for (int c = 0; c < UD; c++) Ys[0][c] = X[c];
            rhs(run, k1, X, t, vs[0], cpfs[0]);
#pragma unroll
            for (int c = 0; c < UD; c++) Ys[1][c] = Y[c] = fma(k1[c] * P(1.0 / 5.0), h, X[c]);
            rhs(run, k2, Y, fma(P(1.0 / 5.0), h, t), vs[1], cpfs[1]);
#pragma unroll
            for (int c = 0; c < UD; c++) Ys[2][c] = Y[c] = fma(fma(k2[c], P(9.0 / 40.0), k1[c] * P(3.0 / 40.0)), h, X[c]);
            rhs(run, k3, Y, fma(P(3.0 / 10.0), h, t), vs[2], cpfs[2]);
#pragma unroll
            for (int c = 0; c < UD; c++)
                Ys[3][c] = Y[c] = fma(fma(k3[c], P(6.0 / 5.0), fma(k2[c], P(-9.0 / 10.0), k1[c] * P(3.0 / 10.0))), h, X[c]);
            rhs(run, k4, Y, fma(P(3.0 / 5.0), h, t), vs[3], cpfs[3]);
#pragma unroll
            for (int c = 0; c < UD; c++)
                Ys[4][c] = Y[c] = fma(fma(k4[c], P(35.0 / 27.0), fma(k3[c], P(-70.0 / 27.0), fma(k2[c], P(5.0 / 2.0), k1[c] * P(-11.0 / 54.0)))), h, X[c]);
            rhs(run, k5, Y, t + h, vs[4], cpfs[4]);
#pragma unroll
            for (int c = 0; c < UD; c++)
                Ys[5][c] = Y[c] = fma(fma(k5[c], P(253.0 / 4096.0), fma(k4[c], P(44275.0 / 110592.0), fma(k3[c], P(575.0 / 13824.0),
                                      fma(k2[c], P(175.0 / 512.0), k1[c] * P(1631.0 / 55296.0))))), h, X[c]);
            rhs(run, k6, Y, fma(P(7.0 / 8.0), h, t), vs[5], cpfs[5]);
#pragma unroll
            for (int c = 0; c < UD; c++)
                next[c] = fma(fma(k6[c], P(512.0 / 1771.0), fma(k4[c], P(125.0 / 594.0), fma(k3[c], P(250.0 / 621.0), k1[c] * P(37.0 / 378.0)))), h, X[c]);
        }
        if (run) {
#pragma unroll
            for (int c = 0; c < UD; c++) X[c] = next[c];
        }
    };
    auto stage_time = [&](int j, P t, P h) -> P {
        if (ALG == MPGOS_RK4) return j == 0 ? t : (j == 3 ? t + h : t + P(0.5) * h);
        switch (j) {
        case 0: return t;
        case 1: return fma(P(1.0 / 5.0), h, t);
        case 2: return fma(P(3.0 / 10.0), h, t);
        case 3: return fma(P(3.0 / 5.0), h, t);
        case 4: return t + h;
        default: return fma(P(7.0 / 8.0), h, t);
        }
    };

    P Ys[S6][UD], vs[S6][NCp], cpfs[S6][NCp];
    for (int r = 0; r < rounds; r++) {
        // this round's segment of each system: the r-th from the end
        if (leader) {
            const int k = my->nseg - 1 - r;
            my->k = k;
            my->len = k >= 0 ? (int)((my->nacc - (long long)k * K) < K ? (my->nacc - (long long)k * K) : K) : 0;
        }
        if (tid == 0) blk[1] = 0;
        __syncthreads();
        if (leader) atomicMax(&blk[1], my->len);
        __syncthreads();
        const int maxlen = blk[1];
        const int k = my->k, len = lane ? my->len : 0;
        P X[UD];
        if (active && len > 0) {
#pragma unroll
            for (int c = 0; c < UD; c++) X[c] = a.recX[((size_t)k * UD + c) * NU + gu];
        } else {
#pragma unroll
            for (int c = 0; c < UD; c++) X[c] = P(0);
        }
        // 1. recompute the segment, keeping the state at the start of each step
        for (int j = 0; j < maxlen; j++) {
            const bool run = active && j < len;
            P t = P(0), h = P(0);
            if (run) {
                const size_t n = (size_t)k * K + j;
#pragma unroll
                for (int c = 0; c < UD; c++) A.seg[((size_t)j * UD + c) * NU + gu] = X[c];
                t = a.recT[n * NS + gsys];
                h = a.recH[n * NS + gsys];
            }
            step(run, X, t, h, Ys, vs, cpfs);
        }
        // 2. reverse through the segment
        for (int j = maxlen - 1; j >= 0; j--) {
            const bool run = active && j < len;
            P t = P(0), h = P(0);
            if (run) {
                const size_t n = (size_t)k * K + j;
#pragma unroll
                for (int c = 0; c < UD; c++) X[c] = A.seg[((size_t)j * UD + c) * NU + gu];
                t = a.recT[n * NS + gsys];
                h = a.recH[n * NS + gsys];
            }
            P Xs[UD];
#pragma unroll
            for (int c = 0; c < UD; c++) Xs[c] = X[c];
            step(run, Xs, t, h, Ys, vs, cpfs);         // stage values of this step
            P kb[S6][UD];
#pragma unroll
            for (int s = 0; s < S6; s++)
#pragma unroll
                for (int c = 0; c < UD; c++) kb[s][c] = h * Bcoef[s] * lam[c];
#pragma unroll
            for (int s = ST - 1; s >= 0; s--) {
                P gy[UD];
                vjp(run, kb[s], Ys[s], stage_time(s, t, h), vs[s], cpfs[s], gy);
#pragma unroll
                for (int c = 0; c < UD; c++) lam[c] += gy[c];
#pragma unroll
                for (int l = 0; l < S6; l++)
                    if (l < s && Acoef[s][l] != P(0)) {
#pragma unroll
                        for (int c = 0; c < UD; c++) kb[l][c] = fma(h * Acoef[s][l], gy[c], kb[l][c]);
                    }
            }
        }
        __syncthreads();
    }

    // outputs: dL/dx(start), dL/d(unit parameters), dL/d(system parameters)
    const bool bad = active && (my->status & MPGOS_STATUS_CHECKPOINT_OVERFLOW) != 0;
    if (active) {
        const P nan = __builtin_nan("");
#pragma unroll
        for (int c = 0; c < UD; c++) A.lam[gu + c * NU] = bad ? nan : lam[c];
#pragma unroll
        for (int c = 0; c < NUP; c++) A.gU[gu + c * NU] = bad ? nan : muU[c];
    }
    if (lane) {
#pragma unroll
        for (int c = 0; c < NSP; c++) red[((size_t)lsys * UPS + u) * NSPp + c] = muS[c];
    }
    __syncthreads();
    if (leader && active) {
        for (int c = 0; c < NSP; c++) {
            P s = P(0);
            for (int v = 0; v < UPS; v++) s += red[((size_t)lsys * UPS + v) * NSPp + c];
            A.gS[gsys + (size_t)c * NS] = bad ? __builtin_nan("") : s;
        }
    }
}

template <class Sys, int ALG, int MAXT>
__global__ void __launch_bounds__(MAXT) pb_reverse_kernel(Args<Sys, double> a, AdjArgs<Sys> A)
{
    pb_reverse_body<Sys, ALG>(a, A);
}

} // namespace pb
} // namespace mpgos
