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
// The dense-contraction form (full FP64 matrix; below) runs the coupling
// products on MFMA and replaces the stage recompute by a stage buffer; it
// returns the same bits as the lane form.
#pragma once

#include "pb_device.h"

namespace mpgos {
namespace pb {

template <class Sys>
struct AdjArgs {
    double* seg;          // [K][UD][NU] (lane form)
    double* stg;          // [K][stages][UD + NC][NU] stage inputs and coupling values (dense-contraction form)
    double* gSp;          // [NSP][NU] per-unit partials of dL/d(system parameters) (dense-contraction form)
    double* lam;          // in: dL/dx(end), out: dL/dx(start)      [UD][NU]
    double* gU;           // out: dL/d(unit parameters)             [NUP][NU]
    double* gS;           // out: dL/d(system parameters)           [NSP][NS]
    const double* cMatT;  // transposed coupling matrices in the stored form of cMat
    const double* cTileT; // their pair-tiled copy (dense-contraction form; null: read cMatT)
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

// ---------------------------------------------------------------- dense-contraction form
// Butcher coefficients of the steppers (the adjoint's coefficients; the forward
// arithmetic itself is written out in the steps below) and the stage times
// exactly as the forward pass computes them.
template <int ALG>
__device__ __forceinline__ double rk_b(int s)
{
    if (ALG == MPGOS_RK4) return (s == 0 || s == 3) ? 1.0 / 6.0 : 2.0 / 6.0;
    switch (s) {
    case 0: return 37.0 / 378.0;
    case 2: return 250.0 / 621.0;
    case 3: return 125.0 / 594.0;
    case 5: return 512.0 / 1771.0;
    default: return 0.0;
    }
}

template <int ALG>
__device__ __forceinline__ double rk_a(int s, int l)
{
    if (ALG == MPGOS_RK4) return (l == s - 1) ? (s == 3 ? 1.0 : 0.5) : 0.0;
    constexpr double A[6][5] = {{0, 0, 0, 0, 0},
                                {1.0 / 5.0, 0, 0, 0, 0},
                                {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0},
                                {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0, 0},
                                {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0},
                                {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
    return l < s ? A[s][l] : 0.0;
}

template <int ALG>
__device__ __forceinline__ double rk_stage_time(int s, double t, double h)
{
    using math::fma;
    if (ALG == MPGOS_RK4) return s == 0 ? t : (s == 3 ? t + h : t + 0.5 * h);
    switch (s) {
    case 0: return t;
    case 1: return fma(1.0 / 5.0, h, t);
    case 2: return fma(3.0 / 10.0, h, t);
    case 3: return fma(3.0 / 5.0, h, t);
    case 4: return t + h;
    default: return fma(7.0 / 8.0, h, t);
    }
}

// LDS of the dense-contraction reverse pass: system scopes, the B operand and
// the contraction results (rows of mfma_row(UPS) doubles per coupling and
// block system, mfma_index), the system-parameter gradient partials and two
// block-wide counters.
struct RevMfmaLayout { size_t cptB, coup, red, blk, total; };

template <class Sys>
__host__ __device__ inline RevMfmaLayout reverse_mfma_layout(int UPS)
{
    constexpr int NCp = Pos<Sys::NC>::value, NSPp = Pos<Sys::NSP>::value, SPB = kMfmaSystems;
    RevMfmaLayout L;
    size_t o = ((size_t)SPB * sizeof(RevScope<Sys>) + 15) & ~(size_t)15;
    L.cptB = o;
    o += (size_t)NCp * SPB * mfma_row(UPS) * sizeof(double);
    L.coup = o;
    o += (size_t)NCp * SPB * mfma_row(UPS) * sizeof(double);
    L.red = o;
    o += (size_t)SPB * UPS * NSPp * sizeof(double);
    L.blk = o;
    o += 2 * sizeof(int);
    L.total = o;
    return L;
}

// Reverse pass in the dense-contraction form (full FP64 coupling matrix, UPS a
// multiple of 32 up to 256): 8 systems per workgroup, 4 units per lane
// (2 * UPS threads), every coupling product on v_mfma_f64_4x4x4f64 (mfma_contract44) --
// M CPT(Y) in the segment replay, M^T w (cMatT) in the adjoint sweep.  Round r
// handles the r-th segment from the end of every system in two launches:
//   pb_replay_mfma_kernel: restore checkpoint k = nseg - 1 - r and replay the
//     segment's recorded steps, keeping each stage's input Y_s and coupling
//     values v_s in the stage buffer
//       stg[(((j * ST + s) * (UD + NC)) + c) * NU + unit],  j < K;
//   pb_sweep_mfma_kernel: the adjoint sweep through the segment from the
//     stage buffer (no forward coupling product again; CPF_s from the unit's
//     right-hand side at (Y_s, T_s)); dL/dx, the unit-parameter gradient and
//     the per-unit system-parameter partials carry over between rounds in HBM
//     (lam, gU, gSp);
// pb_reverse_finish_kernel sums the partials per system in unit order.  Two
// kernels instead of one keep the stage-vector registers of the replay and the
// adjoint registers of the sweep from being live together.  Sums run in the
// same order as in the lane form (the accumulation chain of the MFMA
// instruction is a sequential fused multiply-add in column order), so both
// forms return the same bits.
template <class Sys, int PAIRS>
struct MfmaRevCtx {
    static constexpr int SPB = kMfmaSystems;
    // tiles a wave contracts together in the replay and sweep kernels: all of
    // them for every system (these kernels hold fewer values across the
    // contraction than the forward step; ring with a quadrature state,
    // reverse 352 -> 321 ms against one tile at a time)
    static constexpr int TPW = PAIRS >= 4 ? PAIRS / 2 : 1;
    using Scope = RevScope<Sys>;
    Scope* sys; double* cptB; double* coup; int* blk;
    int lanesPerSys, lsys, u0, gsys, k, len, maxlen;
    bool leader, active;
    size_t NU;
    unsigned gu0;
    Scope* my;

    // lane mapping, system scope and this round's segment of each system
    // (k = nseg - 1 - r, len steps; maxlen over the workgroup)
    __device__ __forceinline__ void setup(const Args<Sys, double>& a, unsigned char* lds, int r)
    {
        const RevMfmaLayout L = reverse_mfma_layout<Sys>(a.UPS);
        sys = reinterpret_cast<Scope*>(lds);
        cptB = reinterpret_cast<double*>(lds + L.cptB);
        coup = reinterpret_cast<double*>(lds + L.coup);
        blk = reinterpret_cast<int*>(lds + L.blk);
        const int tid = threadIdx.x, UPS = a.UPS, NS = a.NS, K = a.recK;
        lanesPerSys = UPS / PAIRS;
        lsys = tid / lanesPerSys;                     // < SPB (blockDim.x = SPB * UPS / PAIRS)
        u0 = tid - lsys * lanesPerSys;
        leader = u0 == 0;
        gsys = blockIdx.x * SPB + lsys;
        active = gsys < NS && gsys < a.nActive;
        NU = (size_t)NS * UPS;
        gu0 = (unsigned)gsys * (unsigned)UPS + (unsigned)u0;
        my = &sys[lsys];
        for (int i = tid; i < Pos<Sys::NC>::value * kMfmaSystems * mfma_row(UPS); i += blockDim.x) cptB[i] = 0.0;
        if (tid == 0) blk[1] = 0;
        if (leader) {
            my->status = 0; my->k = -1; my->len = 0;
            if (active) {
                for (int i = 0; i < Sys::NSP; i++) my->sPAR[i] = a.sPAR[gsys + i * NS];
                for (int i = 0; i < Sys::NSA; i++) my->sACC[i] = a.sACC[gsys + i * NS];
                for (int i = 0; i < Sys::NiSA; i++) my->sACCi[i] = a.sACCi[gsys + i * NS];
                for (int i = 0; i < Sys::NC; i++) my->strength[i] = a.strength[gsys + i * NS];
                my->nacc = a.nacc[gsys];
                my->status = a.status[gsys];
                const bool ok = (my->status & MPGOS_STATUS_CHECKPOINT_OVERFLOW) == 0;
                const int nseg = ok ? (int)((my->nacc + K - 1) / K) : 0;
                const int kk = nseg - 1 - r;
                my->k = kk;
                my->len = kk >= 0 ? (int)((my->nacc - (long long)kk * K) < K ? (my->nacc - (long long)kk * K) : K) : 0;
            }
        }
        __syncthreads();
        if (leader) atomicMax(&blk[1], my->len);
        __syncthreads();
        maxlen = blk[1];
        k = my->k;
        len = my->len;
    }
    __device__ __forceinline__ int uid(int p) const { return u0 + p * lanesPerSys; }
    __device__ __forceinline__ unsigned at(int p) const { return gu0 + (unsigned)(p * lanesPerSys); }
    // row r of a unit-scope array [rows][NU]: wave-uniform row base, 32-bit lane offset
    __device__ __forceinline__ double* row(double* base, size_t r) const { return base + r * NU; }
};

template <class Sys, int ALG, int PAIRS>
__device__ void pb_replay_mfma_body(const Args<Sys, double>& a, const AdjArgs<Sys>& A, int r)
{
    using P = double;
    constexpr int UD = Sys::UD, NC = Sys::NC, NUP = Sys::NUP, NUA = Sys::NUA, NiUA = Sys::NiUA;
    constexpr int NCp = Pos<NC>::value, NUPp = Pos<NUP>::value;
    constexpr int ST = ALG == MPGOS_RK4 ? 4 : 6, SW = UD + NC, SPB = kMfmaSystems;
    using math::fma;
    using U = Unit<Sys, P>;
    extern __shared__ __align__(16) unsigned char mpgos_pb_lds[];
    MfmaRevCtx<Sys, PAIRS> C;
    C.setup(a, mpgos_pb_lds, r);
    if (C.maxlen == 0) return;                         // workgroup-uniform
    const int UPS = a.UPS, NS = a.NS, K = a.recK;
    const int gsys = C.gsys, lsys = C.lsys, k = C.k, len = C.len;
    const bool active = C.active;
    auto* my = C.my;
    double* cptB = C.cptB;
    double* coup = C.coup;

    int cIndex[NCp];
#pragma unroll
    for (int i = 0; i < NC; i++) cIndex[i] = a.cIndex[i];
    P uPAR[PAIRS][NUPp], uACC[PAIRS][Pos<NUA>::value], X[PAIRS][UD];
    int uACCi[PAIRS][Pos<NiUA>::value];
#pragma unroll
    for (int p = 0; p < PAIRS; p++) {
        const bool ld = active && len > 0;
#pragma unroll
        for (int c = 0; c < NUP; c++) uPAR[p][c] = ld ? C.row(a.uPAR, c)[C.at(p)] : P(0);
#pragma unroll
        for (int c = 0; c < NUA; c++) uACC[p][c] = ld ? C.row(a.uACC, c)[C.at(p)] : P(0);
#pragma unroll
        for (int c = 0; c < NiUA; c++) uACCi[p][c] = ld ? a.uACCi[C.at(p) + c * C.NU] : 0;
#pragma unroll
        for (int c = 0; c < UD; c++) X[p][c] = ld ? C.row(a.recX, (size_t)k * UD + c)[C.at(p)] : P(0);
    }
    P* gPAR = a.gPAR;
    int* gPARi = a.gPARi;

    // right-hand sides of the lane's units at (Y, T) plus the coupling terms;
    // v receives the coupling values (M CPT)_u.  The barriers order LDS only,
    // so the stage-buffer stores stay in flight across them.
    auto rhs = [&](bool run, P (*F)[UD], P (*Y)[UD], const P T, P (*v)[NCp]) {
        P CPF[PAIRS][NCp];
        if (run) {
#pragma unroll
            for (int p = 0; p < PAIRS; p++) {
                P cp[NCp];
                Sys::template OdeFunction<P>(gsys, C.uid(p), F[p], Y[p], T, uPAR[p], my->sPAR, gPAR, gPARi, uACC[p], uACCi[p],
                                             my->sACC, my->sACCi, cp, CPF[p]);
#pragma unroll
                for (int i = 0; i < NC; i++) cptB[mfma_index(i, lsys, C.uid(p), UPS)] = cp[i];
            }
        }
        lds_barrier();
        mfma_contract44<NC, true, 8, MfmaRevCtx<Sys, PAIRS>::TPW>(a.cMat, cptB, coup, UPS, SPB, a.cTile);
        lds_barrier();
        if (run) {
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int i = 0; i < NC; i++) {
                    v[p][i] = coup[mfma_index(i, lsys, C.uid(p), UPS)];
                    U::add(cIndex[i], my->strength[i] * CPF[p][i], v[p][i], F[p]);
                }
        }
    };
    for (int j = 0; j < C.maxlen; j++) {
        const bool run = active && j < len;
        P t = P(0), h = P(0);
        if (run) {
            const size_t n = (size_t)k * K + j;
            t = a.recT[n * NS + gsys];
            h = a.recH[n * NS + gsys];
        }
        auto keep = [&](int s, P (*Y)[UD], P (*v)[NCp]) {
            if (run) {
                const size_t base = ((size_t)j * ST + s) * SW;
#pragma unroll
                for (int p = 0; p < PAIRS; p++) {
#pragma unroll
                    for (int c = 0; c < UD; c++) C.row(A.stg, base + c)[C.at(p)] = Y[p][c];
#pragma unroll
                    for (int i = 0; i < NC; i++) C.row(A.stg, base + UD + i)[C.at(p)] = v[p][i];
                }
            }
        };
        // the forward pass's stage arithmetic (pb_body)
        P v[PAIRS][NCp], next[PAIRS][UD], Y[PAIRS][UD];
        if (ALG == MPGOS_RK4) {
            P k1[PAIRS][UD];
            rhs(run, next, X, t, v);
            keep(0, X, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) Y[p][c] = fma(next[p][c] * P(0.5), h, X[p][c]);
            const P t2 = t + P(0.5) * h;
            rhs(run, k1, Y, t2, v);
            keep(1, Y, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) { next[p][c] = fma(P(2), k1[p][c], next[p][c]); Y[p][c] = fma(k1[p][c] * P(0.5), h, X[p][c]); }
            rhs(run, k1, Y, t2, v);
            keep(2, Y, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) { next[p][c] = fma(P(2.0), k1[p][c], next[p][c]); Y[p][c] = fma(k1[p][c], h, X[p][c]); }
            rhs(run, k1, Y, t + h, v);
            keep(3, Y, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) next[p][c] = fma(h * (next[p][c] + k1[p][c]), P(1.0 / 6.0), X[p][c]);
        } else {
            P k1[PAIRS][UD], k2[PAIRS][UD], k3[PAIRS][UD], k4[PAIRS][UD], k5[PAIRS][UD], k6[PAIRS][UD];
            rhs(run, k1, X, t, v);
            keep(0, X, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) Y[p][c] = fma(k1[p][c] * P(1.0 / 5.0), h, X[p][c]);
            rhs(run, k2, Y, fma(P(1.0 / 5.0), h, t), v);
            keep(1, Y, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) Y[p][c] = fma(fma(k2[p][c], P(9.0 / 40.0), k1[p][c] * P(3.0 / 40.0)), h, X[p][c]);
            rhs(run, k3, Y, fma(P(3.0 / 10.0), h, t), v);
            keep(2, Y, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++)
                    Y[p][c] = fma(fma(k3[p][c], P(6.0 / 5.0), fma(k2[p][c], P(-9.0 / 10.0), k1[p][c] * P(3.0 / 10.0))), h, X[p][c]);
            rhs(run, k4, Y, fma(P(3.0 / 5.0), h, t), v);
            keep(3, Y, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++)
                    Y[p][c] = fma(fma(k4[p][c], P(35.0 / 27.0), fma(k3[p][c], P(-70.0 / 27.0), fma(k2[p][c], P(5.0 / 2.0),
                                  k1[p][c] * P(-11.0 / 54.0)))), h, X[p][c]);
            rhs(run, k5, Y, t + h, v);
            keep(4, Y, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++)
                    Y[p][c] = fma(fma(k5[p][c], P(253.0 / 4096.0), fma(k4[p][c], P(44275.0 / 110592.0), fma(k3[p][c], P(575.0 / 13824.0),
                                  fma(k2[p][c], P(175.0 / 512.0), k1[p][c] * P(1631.0 / 55296.0))))), h, X[p][c]);
            rhs(run, k6, Y, fma(P(7.0 / 8.0), h, t), v);
            keep(5, Y, v);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++)
                    next[p][c] = fma(fma(k6[p][c], P(512.0 / 1771.0), fma(k4[p][c], P(125.0 / 594.0), fma(k3[p][c], P(250.0 / 621.0),
                                     k1[p][c] * P(37.0 / 378.0)))), h, X[p][c]);
        }
        if (run) {
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) X[p][c] = next[p][c];
        }
    }
}

template <class Sys, int ALG, int PAIRS>
__device__ void pb_sweep_mfma_body(const Args<Sys, double>& a, const AdjArgs<Sys>& A, int r)
{
    using P = double;
    constexpr int UD = Sys::UD, NC = Sys::NC, NUP = Sys::NUP, NSP = Sys::NSP, NUA = Sys::NUA, NiUA = Sys::NiUA;
    constexpr int NCp = Pos<NC>::value, NUPp = Pos<NUP>::value, NSPp = Pos<NSP>::value;
    constexpr int ST = ALG == MPGOS_RK4 ? 4 : 6, SW = UD + NC, SPB = kMfmaSystems;
    using math::fma;
    extern __shared__ __align__(16) unsigned char mpgos_pb_lds[];
    MfmaRevCtx<Sys, PAIRS> C;
    C.setup(a, mpgos_pb_lds, r);
    if (C.maxlen == 0) return;                         // workgroup-uniform
    const int UPS = a.UPS, NS = a.NS, K = a.recK;
    const int gsys = C.gsys, lsys = C.lsys, k = C.k, len = C.len;
    const bool active = C.active;
    const bool mine = active && len > 0;               // this round has a segment of this system
    auto* my = C.my;
    double* cptB = C.cptB;
    double* coup = C.coup;

    int cIndex[NCp];
#pragma unroll
    for (int i = 0; i < NC; i++) cIndex[i] = a.cIndex[i];
    // adjoint state carried between rounds; round 0 starts the gradients at zero
    P uPAR[PAIRS][NUPp], uACC[PAIRS][Pos<NUA>::value], lam[PAIRS][UD], muU[PAIRS][NUPp], muS[PAIRS][NSPp];
    int uACCi[PAIRS][Pos<NiUA>::value];
#pragma unroll
    for (int p = 0; p < PAIRS; p++) {
#pragma unroll
        for (int c = 0; c < NUP; c++) {
            uPAR[p][c] = mine ? C.row(a.uPAR, c)[C.at(p)] : P(0);
            muU[p][c] = (mine && r > 0) ? C.row(A.gU, c)[C.at(p)] : P(0);
        }
#pragma unroll
        for (int c = 0; c < NUA; c++) uACC[p][c] = mine ? C.row(a.uACC, c)[C.at(p)] : P(0);
#pragma unroll
        for (int c = 0; c < NiUA; c++) uACCi[p][c] = mine ? a.uACCi[C.at(p) + c * C.NU] : 0;
#pragma unroll
        for (int c = 0; c < UD; c++) lam[p][c] = mine ? C.row(A.lam, c)[C.at(p)] : P(0);
#pragma unroll
        for (int c = 0; c < NSP; c++) muS[p][c] = (mine && r > 0) ? C.row(A.gSp, c)[C.at(p)] : P(0);
    }
    P* gPAR = a.gPAR;
    int* gPARi = a.gPARi;

    for (int j = C.maxlen - 1; j >= 0; j--) {
        const bool run = active && j < len;
        P t = P(0), h = P(0);
        if (run) {
            const size_t n = (size_t)k * K + j;
            t = a.recT[n * NS + gsys];
            h = a.recH[n * NS + gsys];
        }
        P kb[ST][PAIRS][UD];
#pragma unroll
        for (int s = 0; s < ST; s++)
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) kb[s][p][c] = h * rk_b<ALG>(s) * lam[p][c];
#pragma unroll
        for (int s = ST - 1; s >= 0; s--) {
            const P Ts = rk_stage_time<ALG>(s, t, h);
            const size_t base = ((size_t)j * ST + s) * SW;
            P kbCPF[PAIRS][NCp];
            if (run) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++) {
                    P Y[UD], F[UD], cp[NCp], cpf[NCp];
#pragma unroll
                    for (int c = 0; c < UD; c++) Y[c] = C.row(A.stg, base + c)[C.at(p)];
                    Sys::template OdeFunction<P>(gsys, C.uid(p), F, Y, Ts, uPAR[p], my->sPAR, gPAR, gPARi, uACC[p], uACCi[p],
                                                 my->sACC, my->sACCi, cp, cpf);
#pragma unroll
                    for (int i = 0; i < NC; i++) {
                        P g = P(0);
#pragma unroll
                        for (int c = 0; c < UD; c++)
                            if (c == cIndex[i]) g = kb[s][p][c];
                        cptB[mfma_index(i, lsys, C.uid(p), UPS)] = my->strength[i] * cpf[i] * g;
                        kbCPF[p][i] = my->strength[i] * C.row(A.stg, base + UD + i)[C.at(p)] * g;
                    }
                }
            }
            __syncthreads();
            mfma_contract44<NC, true, 8, MfmaRevCtx<Sys, PAIRS>::TPW>(A.cMatT, cptB, coup, UPS, SPB, A.cTileT);
            __syncthreads();
#pragma unroll
            for (int p = 0; p < PAIRS; p++) {
                P gy[UD];
                if (run) {
                    P Y[UD], kbCPT[NCp], gU[NUPp], gS[NSPp];
#pragma unroll
                    for (int c = 0; c < UD; c++) Y[c] = C.row(A.stg, base + c)[C.at(p)];
#pragma unroll
                    for (int i = 0; i < NC; i++) kbCPT[i] = coup[mfma_index(i, lsys, C.uid(p), UPS)];
                    Sys::template OdeVJP<P>(gsys, C.uid(p), kb[s][p], kbCPT, kbCPF[p], Y, Ts, uPAR[p], my->sPAR, gPAR, gPARi,
                                            uACC[p], uACCi[p], my->sACC, my->sACCi, gy, gU, gS);
#pragma unroll
                    for (int c = 0; c < NUP; c++) muU[p][c] += gU[c];
#pragma unroll
                    for (int c = 0; c < NSP; c++) muS[p][c] += gS[c];
                } else {
#pragma unroll
                    for (int c = 0; c < UD; c++) gy[c] = P(0);
                }
#pragma unroll
                for (int c = 0; c < UD; c++) lam[p][c] += gy[c];
#pragma unroll
                for (int l = 0; l < ST; l++)
                    if (l < s && rk_a<ALG>(s, l) != P(0)) {
#pragma unroll
                        for (int c = 0; c < UD; c++) kb[l][p][c] = fma(h * rk_a<ALG>(s, l), gy[c], kb[l][p][c]);
                    }
            }
        }
    }
    if (mine) {
#pragma unroll
        for (int p = 0; p < PAIRS; p++) {
#pragma unroll
            for (int c = 0; c < UD; c++) C.row(A.lam, c)[C.at(p)] = lam[p][c];
#pragma unroll
            for (int c = 0; c < NUP; c++) C.row(A.gU, c)[C.at(p)] = muU[p][c];
#pragma unroll
            for (int c = 0; c < NSP; c++) C.row(A.gSp, c)[C.at(p)] = muS[p][c];
        }
    }
}

template <class Sys, int ALG, int PAIRS>
__global__ void __launch_bounds__(kMfmaSystems * 256 / PAIRS) pb_replay_mfma_kernel(Args<Sys, double> a, AdjArgs<Sys> A, int r)
{
    pb_replay_mfma_body<Sys, ALG, PAIRS>(a, A, r);
}

template <class Sys, int ALG, int PAIRS>
__global__ void __launch_bounds__(kMfmaSystems * 256 / PAIRS) pb_sweep_mfma_kernel(Args<Sys, double> a, AdjArgs<Sys> A, int r)
{
    pb_sweep_mfma_body<Sys, ALG, PAIRS>(a, A, r);
}

// Outputs of the dense-contraction reverse pass, one thread per unit: systems
// without accepted steps get zero gradients (dL/dx unchanged), systems whose
// record overflowed NaN; dL/d(system parameters) sums the per-unit partials in
// unit order (thread 0 of the system's units).
template <class Sys>
__global__ void pb_reverse_finish_kernel(Args<Sys, double> a, AdjArgs<Sys> A)
{
    constexpr int UD = Sys::UD, NUP = Sys::NUP, NSP = Sys::NSP;
    const size_t NU = (size_t)a.NS * a.UPS;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NU) return;
    const int gsys = (int)(g / a.UPS), u = (int)(g - (size_t)gsys * a.UPS);
    if (gsys >= a.nActive) return;
    const bool bad = (a.status[gsys] & MPGOS_STATUS_CHECKPOINT_OVERFLOW) != 0;
    const bool none = a.nacc[gsys] == 0;
    const double nan = __builtin_nan("");
    if (bad) {
        for (int c = 0; c < UD; c++) A.lam[g + c * NU] = nan;
        for (int c = 0; c < NUP; c++) A.gU[g + c * NU] = nan;
    } else if (none) {
        for (int c = 0; c < NUP; c++) A.gU[g + c * NU] = 0.0;
    }
    if (u == 0) {
        for (int c = 0; c < NSP; c++) {
            double s = 0.0;
            if (!none)
                for (int v = 0; v < a.UPS; v++) s += A.gSp[(size_t)c * NU + g + v];
            A.gS[gsys + (size_t)c * a.NS] = bad ? nan : s;
        }
    }
}

} // namespace pb
} // namespace mpgos
