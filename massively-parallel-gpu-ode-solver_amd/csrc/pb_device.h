// Per-block coupled-systems solver for gfx950: one workgroup integrates SPB
// systems of UPS coupled units; a lane owns PAIRS (system, unit) pairs of one
// unit (PAIRS = 1 in the lane-per-pair form, 2 in the dense-contraction form).
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
// therefore deterministic; multiply-adds are explicit fma in the contracted
// form of the reference's expressions, so the host oracle reproduces every bit.
//
// Two coupling forms:
//  * lane-per-pair (any storage form): each lane forms its own dot product,
//    coupling matrix from L2 (or LDS), coupling terms double-buffered in LDS
//    (one barrier per stage);
//  * dense contraction (full matrix, FP64, UPS a multiple of 32 up to 256, 8
//    systems per workgroup, 4 pairs per lane): the products C . [CPT_s0 ..
//    CPT_s7] of all 8 systems run on v_mfma_f64_4x4x4f64 (mfma_contract44:
//    four 4 x 4 x 4 blocks per instruction take the four 4-row slices of a
//    16-unit tile against 4 systems; two chains per tile cover the 8
//    systems).  The instruction accumulates its four products as a chain of
//    fused multiply-adds in k order, so chaining the column blocks in order
//    equals the sequential fma(M[u][col], CPT[col], v) over col = 0 .. UPS-1
//    of the lane-per-pair form bit for bit.  Results go through LDS to the
//    lane that owns the pair (two barriers per stage).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mpgos.h"
#include "mpgos_math.h"

namespace mpgos {
namespace pb {

template <int N> struct Pos { static constexpr int value = N > 0 ? N : 1; };

// Workgroup barrier that orders LDS only: waits for the wave's own LDS
// operations (lgkmcnt) and not for its global loads and stores, which stay in
// flight across the barrier (__syncthreads also drains vmcnt).  For barriers
// whose only purpose is LDS communication; global data exchanged between lanes
// still needs __syncthreads.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int kMaxBlock = 1024;
constexpr int kMfmaSystems = 8;          // systems per workgroup of the dense-contraction form
constexpr int kMfmaPairs = 4;            // default pairs per lane of the dense-contraction form (2 * UPS threads)

// Tiles of the dense-contraction form that a wave contracts together: all of
// its PAIRS / 2 tiles when the lane's unit states and stage vectors leave room
// in the registers (UD * PAIRS <= 8), else one at a time.  Only the first way
// keeps leading matrix pairs in LDS.
template <class Sys, int PAIRS>
__host__ __device__ constexpr int mfma_tiles_together() { return PAIRS >= 4 && Sys::UD * PAIRS <= 8 ? PAIRS / 2 : 1; }
template <class Sys, int PAIRS>
__host__ __device__ constexpr bool mfma_lds_matrix() { return mfma_tiles_together<Sys, PAIRS>() > 1; }

// LDS of the dense-contraction form: the B operand (coupling terms) and the
// contraction results, one row of UPS + 2 doubles per (coupling, block
// system).  A wave's writes and reads of one system's consecutive units fall
// in consecutive banks; rows 4 banks apart (UPS a multiple of 16) put the B
// reads of a column block (4 columns x 4 systems, two columns per half-wave)
// in 8 distinct bank pairs per half-wave and spread the result stores (4
// systems x 16 rows of a tile) evenly over the 32 banks (UPS + 4, 8 banks
// apart, stored them 4 ways conflicting: twice the measured bank conflicts,
// 1.4 % slower).
__host__ __device__ constexpr int mfma_row(int UPS) { return UPS + 2; }
__host__ __device__ inline size_t mfma_index(int i, int s, int u, int UPS)
{
    return ((size_t)i * kMfmaSystems + s) * mfma_row(UPS) + u;
}

enum Coupling { kLaneCoupling = 0, kMfmaCoupling = 1 };

// CP: precision of the step size, tolerances and controller (FP64 in the
// FLOAT_KAHAN mode, else P).
template <class Sys, class P, class CP = P>
struct Args {
    static constexpr int UD = Sys::UD, NE = Sys::NE;
    // unit scope, index (system * UPS + unit) + component * NU, NU = NS * UPS
    P* X; P* uPAR; P* uACC; int* uACCi;
    // system scope, index system + component * NS
    P* TD; P* T; P* sPAR; P* sACC; int* sACCi; P* strength; int* DOidx; P* DOt;
    P* DOx;                   // unit index + component * NU + k * NU * UD
    // global scope
    P* gPAR; int* gPARi; int* cIndex; P* cMat;
    const P* cTile;           // dense-contraction form: the matrix in column-block pairs per lane (null: read cMat)
    int ldsPairs;             // dense-contraction form: leading column-block pairs of every tile of cTile held in LDS
    long long* nacc; long long* nrej; int* status;
    int NS, UPS, SPB, nActive, NDO, CBW, CCI, cMatSize, sharedGlobals, sharedMatrix;
    int doFreq;
    CP dt0, dtmax, dtmin, grow, shrink;
    P doMinDt;
    CP rtol[UD], atol[UD];
    P etol[Pos<NE>::value];
    int edir[Pos<NE>::value];
    // record pass of the discrete adjoint (unused otherwise): unit states every
    // recK accepted steps, recX[(k * UD + c) * NU + unit], k < recCkCap; time
    // and size of every accepted step, recT/recH[n * NS + system], n < recStepCap
    P* recX; P* recT; P* recH;
    int recK, recCkCap, recStepCap;
};

// System scope of one system (LDS).  KAHAN: FP32 time with its compensation
// tc, FP64 step sizes, and dtP, the FP32 copy of the step size that the user
// callbacks see (and may change).
template <class Sys, class P, bool KAHAN = false>
struct SysScope {
    using CP = typename std::conditional<KAHAN, double, P>::type;
    P TD[2];
    P t;
    CP dt, newDt;
    P doT;
    P tc, dtP;
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

// LDS bytes: system scopes, coupling terms (two [system][unit][coupling]
// buffers in the lane-per-pair form; one [coupling][unit][16] B-operand
// buffer plus a [coupling][unit][system] result buffer in the dense form),
// optional global parameters and coupling matrix copies.
struct LdsLayout {
    size_t cpt, coup, g, gi, m, mt, total;
};

template <class Sys, class P, bool KAHAN = false>
__host__ __device__ inline LdsLayout lds_layout(int cpl, int SPB, int UPS, int sharedGlobals, int sharedMatrix,
                                                int cMatSize, int ldsPairs = 0)
{
    constexpr int NCp = Pos<Sys::NC>::value;
    LdsLayout L;
    size_t o = (size_t)SPB * sizeof(SysScope<Sys, P, KAHAN>);
    o = (o + 15) & ~(size_t)15;
    L.cpt = o;
    if (cpl == kMfmaCoupling) {
        o += (size_t)NCp * kMfmaSystems * mfma_row(UPS) * sizeof(P);
        L.coup = o;
        o += (size_t)NCp * kMfmaSystems * mfma_row(UPS) * sizeof(P);
    } else {
        o += 2 * (size_t)SPB * UPS * NCp * sizeof(P);
        L.coup = o;
    }
    L.g = o;
    if (sharedGlobals) o += (size_t)Sys::NGP * sizeof(P);
    o = (o + 15) & ~(size_t)15;
    L.m = o;
    if (sharedMatrix) o += (size_t)Sys::NC * cMatSize * sizeof(P);
    L.gi = o;
    if (sharedGlobals) o += (size_t)Sys::NiGP * sizeof(int);
    o = (o + 15) & ~(size_t)15;
    L.mt = o;
    if (cpl == kMfmaCoupling) o += (size_t)NCp * (UPS / 16) * ldsPairs * 64 * 16;
    L.total = o;
    return L;
}

template <class Sys, class P>
struct Unit {
    static constexpr int UD = Sys::UD, NC = Sys::NC, NE = Sys::NE;

    // coupling: F[CouplingIndex[i]] += strength_i CPF_i sum_col M_i(u, col) CPT_col[i]
    template <class A>
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
            add(cIndex[i], strength[i] * CPF[i], v, F);
        }
    }

    // F[ci] = fma(sc, v, F[ci]) for a run-time component ci, written as a
    // select per component so that F stays in registers (a conditional store
    // per component is merged by the compiler into one dynamically indexed
    // store, which moves F to scratch memory)
    static __device__ __forceinline__ void add(int ci, P sc, P v, P* F)
    {
#pragma unroll
        for (int c = 0; c < UD; c++) {
            const P f = math::fma(sc, v, F[c]);
            F[c] = c == ci ? f : F[c];
        }
    }
};

// Dense contraction of the coupling terms of the block's systems: for every
// coupling i, unit u and system s < SPB (at most 8),
//   coup[mfma_index(i, s, u)] = sum_col M_i[u + col * UPS] * cptB[mfma_index(i, s, col)]
// accumulated in column order, on v_mfma_f64_4x4x4f64 (called by every lane of
// the workgroup: the instruction needs all 64 lanes of a wave; UPS is a
// multiple of 16).  The instruction computes four independent
// 4 x 4 x 4 products per instruction (block b = lanes with (lane / 4) % 4 == b:
// A[i][k] in lane i + 4b + 16k, B[k][j] in lane j + 4b + 16k, D[i][j] in lane
// j + 4b + 16i) and accumulates its four products as a fused multiply-add
// chain in k order (both probed on the MI355X).  The four blocks take the four
// 4-row slices of a 16-unit tile against the same 4 systems, so the A operand
// is lane (row = lane % 16, column = lane / 16) and every B column is a system:
// two chains per tile (systems 0..3 and 4..7, independent, so one tile per
// pass keeps two in flight).  Output (unit 4 ((lane / 4) % 4) + lane / 16 of
// the tile, system lane % 4 (+ 4)) in one register per chain.
//
// The M tile values of a pass stream from L2 (a 256-unit matrix is 512 KiB) in
// groups of G column blocks.  GLOBAL_M: M is in global memory, read with
// global loads (a generic pointer gives flat loads, which also count as LDS
// operations, so the wait for a group's B operands from LDS would wait for all
// of its M loads too); otherwise M is in LDS.  Mt: the pair-tiled copy of M
// (tile_pair_matrix), read instead of M; Ml, ldsPairs: the leading ldsPairs
// column-block pairs of every tile of Mt, copied into LDS.  TPW: the tiles a
// wave contracts together when the launch gives every wave exactly TPW tiles
// (2 * UPS threads at TPW = 2).
template <int NC, bool GLOBAL_M = true, int G = 8, int TPW = 1>
__device__ __forceinline__ void mfma_contract44(const double* M, const double* cptB, double* coup, int UPS, int SPB,
                                                const double* Mt = nullptr, const double* Ml = nullptr, int ldsPairs = 0)
{
    using GPtr = const __attribute__((address_space(1))) double*;
    struct alignas(16) Pair { double x, y; };
    using GP = const __attribute__((address_space(1))) Pair*;
    using LP = const __attribute__((address_space(3))) Pair*;
    const int l = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int tiles = UPS >> 4;
    const int r16 = l & 15, k4 = l >> 4, sys = l & 3;
    const int drow = 4 * ((l >> 2) & 3) + (l >> 4);
    const int blocks = UPS >> 2;                        // a multiple of G (UPS is a multiple of 32)
    const size_t colStride = (size_t)4 * UPS;
    const int RS = mfma_row(UPS);
    auto store = [&](int i, int t, double acc0, double acc1) {
        if (sys < SPB) coup[mfma_index(i, sys, t * 16 + drow, UPS)] = acc0;
        if (sys + 4 < SPB) coup[mfma_index(i, sys + 4, t * 16 + drow, UPS)] = acc1;
    };
    // pair-tiled copy (tile_pair_matrix): one 16-byte load per lane brings the
    // lane's A values of two column blocks, half the vector-memory
    // instructions of column-major reads.  Per tile, row and system the MFMAs
    // run in column order, as in the column-major loop at the end.
    if (GLOBAL_M && Mt && TPW > 1 && tiles == TPW * nw) {
        // every tile of the wave (w, w + nw, ...) at once, G / TPW column
        // blocks per tile and group: one B read per column block and chain
        // serves all of them, and the 2 TPW accumulator chains are
        // independent.  The leading ldsPairs pairs of every tile (a multiple
        // of G / (2 TPW)) come from the workgroup's LDS copy (Ml, same
        // layout), which takes their loads off the vector-memory path; the
        // global groups are read two buffers deep, the first during the LDS
        // part.
        constexpr int TT = TPW > 1 ? TPW : 2;
        constexpr int GB = G / TT;                      // column blocks per tile and group
        static_assert(GB % 2 == 0, "whole pairs per group");
        double res[Pos<NC>::value][TT][2];
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const double* Bi = cptB + mfma_index(i, sys, k4, UPS);
            double (&acc)[TT][2] = res[i];
#pragma unroll
            for (int k = 0; k < TT; k++) acc[k][0] = acc[k][1] = 0.0;
            auto group = [&](const double (*a)[GB], int c0) {
#pragma unroll
                for (int d = 0; d < GB; d++) {
                    const double b0 = Bi[(c0 + d) * 4], b1 = Bi[(c0 + d) * 4 + 4 * RS];
#pragma unroll
                    for (int k = 0; k < TT; k++) {
                        acc[k][0] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[k][d], b0, acc[k][0], 0, 0, 0);
                        acc[k][1] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[k][d], b1, acc[k][1], 0, 0, 0);
                    }
                }
            };
            const Pair* At = reinterpret_cast<const Pair*>(Mt) + ((size_t)i * tiles + w) * (size_t)blocks * 32 + l;
            const size_t gs = (size_t)nw * blocks * 32;          // the wave's next tile
            auto load = [&](double (*a)[GB], int c0) {
#pragma unroll
                for (int k = 0; k < TT; k++)
#pragma unroll
                    for (int d = 0; d < GB; d += 2) {
                        const GP q = (GP)(At + k * gs + (size_t)((c0 + d) >> 1) * 64);
                        a[k][d] = q->x;
                        a[k][d + 1] = q->y;
                    }
            };
            const int cg = Ml ? 2 * ldsPairs : 0;
            const int n = (blocks - cg) / GB;
            double a0[TT][GB], a1[TT][GB];
            if (n > 0) load(a0, cg);
            if (Ml) {
                const Pair* Lt = reinterpret_cast<const Pair*>(Ml) + ((size_t)i * tiles + w) * (size_t)ldsPairs * 64 + l;
                const size_t ls = (size_t)nw * ldsPairs * 64;
                for (int cp = 0; cp < ldsPairs; cp += GB / 2) {
                    double a[TT][GB];
#pragma unroll
                    for (int k = 0; k < TT; k++)
#pragma unroll
                        for (int d = 0; d < GB; d += 2) {
                            const LP q = (LP)(Lt + k * ls + (size_t)(cp + (d >> 1)) * 64);
                            a[k][d] = q->x;
                            a[k][d + 1] = q->y;
                        }
                    group(a, 2 * cp);
                }
            }
            if (n > 0) {
                // every load in the loop body is unconditional, so the wait
                // before a group's MFMAs counts only that group's loads
                int g = 0;
                for (; g + 2 < n; g += 2) {
                    load(a1, cg + (g + 1) * GB);
                    group(a0, cg + g * GB);
                    load(a0, cg + (g + 2) * GB);
                    group(a1, cg + (g + 1) * GB);
                }
                if (g + 1 < n) {
                    load(a1, cg + (g + 1) * GB);
                    group(a0, cg + g * GB);
                    group(a1, cg + (g + 1) * GB);
                } else {
                    group(a0, cg + g * GB);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < NC; i++)
#pragma unroll
            for (int k = 0; k < TT; k++) store(i, w + k * nw, res[i][k][0], res[i][k][1]);
        return;
    }
    if (GLOBAL_M && Mt) {
        // one tile at a time, G column blocks per group (coup and cptB are
        // distinct here)
        for (int t = w; t < tiles; t += nw) {
#pragma unroll
            for (int i = 0; i < NC; i++) {
                const double* Bi = cptB + mfma_index(i, sys, k4, UPS);
                const Pair* At = reinterpret_cast<const Pair*>(Mt + ((size_t)i * tiles + t) * (size_t)blocks * 64) + l;
                double acc0 = 0.0, acc1 = 0.0;          // systems 0..3, 4..7
                for (int c0 = 0; c0 < blocks; c0 += G) {
                    double a0[G];
#pragma unroll
                    for (int d = 0; d < G; d += 2) {
                        const GP q = (GP)(At + (size_t)((c0 + d) >> 1) * 64);
                        a0[d] = q->x;
                        a0[d + 1] = q->y;
                    }
#pragma unroll
                    for (int d = 0; d < G; d++) {
                        const double b0 = Bi[(c0 + d) * 4], b1 = Bi[(c0 + d) * 4 + 4 * RS];
                        acc0 = __builtin_amdgcn_mfma_f64_4x4x4f64(a0[d], b0, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f64_4x4x4f64(a0[d], b1, acc1, 0, 0, 0);
                    }
                }
                store(i, t, acc0, acc1);
            }
        }
        return;
    }
    // one 16-unit tile per pass; its two system halves are two independent chains
    for (int t = w; t < tiles; t += nw) {
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const double* Ai = M + (size_t)i * UPS * UPS + t * 16 + r16 + (size_t)k4 * UPS;
            const double* Bi = cptB + mfma_index(i, sys, k4, UPS);
            auto ldA = [&](int blk) -> double {
                if constexpr (GLOBAL_M) return *(GPtr)(Ai + (size_t)blk * colStride);
                else return Ai[(size_t)blk * colStride];
            };
            double acc0 = 0.0, acc1 = 0.0;              // systems 0..3, 4..7
            for (int c0 = 0; c0 < blocks; c0 += G) {
                double a0[G];
#pragma unroll
                for (int d = 0; d < G; d++) a0[d] = ldA(c0 + d);
#pragma unroll
                for (int d = 0; d < G; d++) {
                    const double b0 = Bi[(c0 + d) * 4], b1 = Bi[(c0 + d) * 4 + 4 * RS];
                    acc0 = __builtin_amdgcn_mfma_f64_4x4x4f64(a0[d], b0, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f64_4x4x4f64(a0[d], b1, acc1, 0, 0, 0);
                }
            }
            store(i, t, acc0, acc1);
        }
    }
}

// The pair-tiled copy of the dense coupling matrices read by mfma_contract44:
// for matrix i, 16-unit tile t, column-block pair cp, lane l and j in {0, 1},
//   Mt[(((i * tiles + t) * blocks / 2 + cp) * 64 + l) * 2 + j]
//     = M_i[t * 16 + l % 16 + (4 * (2 cp + j) + l / 16) * UPS],
// i.e. the A operand of lane l for column blocks 2 cp and 2 cp + 1 side by
// side (the same values, so the contraction's results are unchanged).
template <int Unused = 0>
__global__ void tile_pair_matrix(const double* M, double* Mt, int UPS, int NC)
{
    const size_t n = (size_t)NC * UPS * UPS;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int tiles = UPS >> 4, pairs = UPS >> 3;
    const int j = (int)(k & 1), l = (int)((k >> 1) & 63);
    size_t r = k >> 7;
    const int cp = (int)(r % pairs); r /= pairs;
    const int t = (int)(r % tiles);
    const int i = (int)(r / tiles);
    Mt[k] = M[(size_t)i * UPS * UPS + t * 16 + (l & 15) + (size_t)(4 * (2 * cp + j) + (l >> 4)) * UPS];
}

// One workgroup = SPB systems.  A system has UPS / PAIRS lanes; lane t belongs
// to block system lsys = t / (UPS / PAIRS) and owns the PAIRS units
// u0 + p * (UPS / PAIRS), u0 = t % (UPS / PAIRS), so the system-scope scalars
// (time, step) are shared by the lane's pairs and global accesses to unit-scope
// arrays stay coalesced.
//
// KAHAN (P = float): FLOAT_KAHAN mode.  Unit states, stages, right-hand sides
// and coupling terms in FP32; the state and the time are accumulated with Kahan
// compensation (x += dt * sum b_j k_j and t += dt as compensated sums); the
// step size, the error norm (from the FP32 stages in FP64 arithmetic), the
// controller and the event step control in FP64.
template <class Sys, class P, int ALG, int PAIRS, int CPL, bool REC = false, bool KAHAN = false>
__device__ void pb_body(const Args<Sys, P, typename SysScope<Sys, P, KAHAN>::CP>& a)
{
    constexpr int UD = Sys::UD, NC = Sys::NC, NUP = Sys::NUP, NSP = Sys::NSP, NGP = Sys::NGP, NiGP = Sys::NiGP;
    constexpr int NUA = Sys::NUA, NiUA = Sys::NiUA, NSA = Sys::NSA, NiSA = Sys::NiSA, NE = Sys::NE;
    constexpr int NCp = Pos<NC>::value;
    using S = SysScope<Sys, P, KAHAN>;
    using CP = typename S::CP;
    using U = Unit<Sys, P>;
    using math::fma;

    extern __shared__ __align__(16) unsigned char mpgos_pb_lds[];
    const int ldsPairs = mfma_lds_matrix<Sys, PAIRS>() ? a.ldsPairs : 0;
    const LdsLayout L = lds_layout<Sys, P, KAHAN>(CPL, a.SPB, a.UPS, a.sharedGlobals, a.sharedMatrix, a.cMatSize, ldsPairs);
    S* sys = reinterpret_cast<S*>(mpgos_pb_lds);
    P* cptBuf = reinterpret_cast<P*>(mpgos_pb_lds + L.cpt);
    P* coupBuf = reinterpret_cast<P*>(mpgos_pb_lds + L.coup);
    const int cptStride = a.SPB * a.UPS * NCp;

    const int UPS = a.UPS, NS = a.NS, SPB = a.SPB;
    const int lanesPerSys = UPS / PAIRS;
    const int tid = threadIdx.x;
    const int lsys = tid / lanesPerSys;
    const int u0 = tid - lsys * lanesPerSys;
    const bool lane = lsys < SPB;                              // lane owns PAIRS (system, unit) pairs
    const bool leader = lane && u0 == 0;
    const size_t NU = (size_t)NS * UPS;
    const int gsys = blockIdx.x * SPB + lsys;
    const bool exists = lane && gsys < NS;
    const bool active = exists && gsys < a.nActive;
    S* const my = lane ? &sys[lsys] : &sys[0];
    auto uid = [&](int p) { return u0 + p * lanesPerSys; };
    auto gu = [&](int p) { return (size_t)gsys * UPS + uid(p); };   // unit-scope index

    // ---- global scope (LDS when SharedGlobalVariables / SharedCouplingMatrices)
    P* gPAR = a.gPAR;
    int* gPARi = a.gPARi;
    const P* M = a.cMat;
    if (a.sharedGlobals) {
        P* lg = reinterpret_cast<P*>(mpgos_pb_lds + L.g);
        int* lgi = reinterpret_cast<int*>(mpgos_pb_lds + L.gi);
        for (int i = tid; i < NGP; i += blockDim.x) lg[i] = a.gPAR[i];
        for (int i = tid; i < NiGP; i += blockDim.x) lgi[i] = a.gPARi[i];
        gPAR = lg;
        gPARi = lgi;
    }
    if (a.sharedMatrix) {
        P* lm = reinterpret_cast<P*>(mpgos_pb_lds + L.m);
        for (int i = tid; i < NC * a.cMatSize; i += blockDim.x) lm[i] = a.cMat[i];
        M = lm;
    }
    if (CPL == kMfmaCoupling)   // B operand: rows of absent systems stay zero
        for (int i = tid; i < NCp * kMfmaSystems * mfma_row(UPS); i += blockDim.x) cptBuf[i] = P(0);
    const double* Ml = nullptr;
    if (CPL == kMfmaCoupling && ldsPairs > 0) {
        // the leading ldsPairs column-block pairs of every tile of the tiled
        // matrix, for the whole Solve (16 bytes per lane and pair)
        using Q = double2;
        Q* lm = reinterpret_cast<Q*>(mpgos_pb_lds + L.mt);
        const Q* src = reinterpret_cast<const Q*>(a.cTile);
        const int lp = ldsPairs, pairs = UPS >> 3;
        const int n = NCp * (UPS >> 4) * lp * 64;
        for (int k = tid; k < n; k += blockDim.x) {
            const int r = k >> 6;
            lm[k] = src[((size_t)(r / lp) * pairs + r % lp) * 64 + (k & 63)];
        }
        Ml = reinterpret_cast<const double*>(lm);
    }
    int cIndex[NCp];
#pragma unroll
    for (int i = 0; i < NC; i++) cIndex[i] = a.cIndex[i];

    // ---- system scope (PB-SOL:2285-2335)
    if (leader) {
        S& s = sys[lsys];
        if (active) {
            s.TD[0] = a.TD[gsys]; s.TD[1] = a.TD[gsys + NS];
            s.t = a.T[gsys];
            s.dt = a.dt0; s.newDt = a.dt0;
            s.tc = P(0); s.dtP = (P)a.dt0;
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
    P X[PAIRS][UD], uPAR[PAIRS][Pos<NUP>::value], uACC[PAIRS][Pos<NUA>::value], EF[PAIRS][Pos<NE>::value];
    P Xc[PAIRS][KAHAN ? UD : 1];                 // Kahan compensation of the unit states
    P CPF[PAIRS][NCp];
    int uACCi[PAIRS][Pos<NiUA>::value];
#pragma unroll
    for (int p = 0; p < PAIRS; p++) {
        if (exists) {
            const size_t g = gu(p);
#pragma unroll
            for (int c = 0; c < UD; c++) { X[p][c] = a.X[g + c * NU]; if (KAHAN) Xc[p][c] = P(0); }
#pragma unroll
            for (int c = 0; c < NUP; c++) uPAR[p][c] = a.uPAR[g + c * NU];
#pragma unroll
            for (int c = 0; c < NUA; c++) uACC[p][c] = a.uACC[g + c * NU];
#pragma unroll
            for (int c = 0; c < NiUA; c++) uACCi[p][c] = a.uACCi[g + c * NU];
        }
#pragma unroll
        for (int i = 0; i < NC; i++) CPF[p][i] = P(0);
    }
    __syncthreads();

    // the step size as the user callbacks see it (FP32 copy in the Kahan mode)
    auto udt = [&]() -> P& {
        if constexpr (KAHAN) return my->dtP;
        else return my->dt;
    };
#define MPGOS_PB_USER(p, TT, DTT) TT, DTT, my->TD, X[p], uPAR[p], my->sPAR, gPAR, gPARi, uACC[p], uACCi[p], my->sACC, my->sACCi
    // ---- initialisation (PB-SOL:2465-2528)
    if (active) {
#pragma unroll
        for (int p = 0; p < PAIRS; p++) Sys::template Initialization<P>(gsys, uid(p), my->DOidx, MPGOS_PB_USER(p, my->t, udt()));
    }
    __syncthreads();
    if (NE > 0) {
        if (active) {
#pragma unroll
            for (int p = 0; p < PAIRS; p++) Sys::template EventFunction<P>(gsys, uid(p), EF[p], MPGOS_PB_USER(p, my->t, udt()));
        }
        __syncthreads();
    }
    if (a.NDO > 0) {
        // first store (the reference stores unconditionally; skipped once the buffer is full)
        if (active && my->DOidx < a.NDO) {
            const int k = my->DOidx;
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) a.DOx[gu(p) + c * NU + (size_t)k * NU * UD] = X[p][c];
            if (leader) a.DOt[gsys + (size_t)k * NS] = my->t;
        }
        __syncthreads();
        if (leader && active && my->DOidx < a.NDO) {
            my->DOidx++;
            my->skipped = 0;
            my->doT = dmin<P>(my->t + a.doMinDt, my->TD[1]);
        }
        __syncthreads();
    }

    if (REC && active) {        // checkpoint 0: the initial unit states
#pragma unroll
        for (int p = 0; p < PAIRS; p++)
#pragma unroll
            for (int c = 0; c < UD; c++) a.recX[(size_t)c * NU + gu(p)] = X[p][c];
    }

    int buf = 0;
    // right-hand sides of this lane's pairs at (Y[p], T), then the coupling terms
    auto rhs = [&](P (*F)[UD], P (*Y)[UD], const P T) {
        const bool run = lane && my->term == 0;
        if constexpr (CPL == kMfmaCoupling) {
            if (run) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++) {
                    P cp[NCp];
                    Sys::template OdeFunction<P>(gsys, uid(p), F[p], Y[p], T, uPAR[p], my->sPAR, gPAR, gPARi, uACC[p], uACCi[p],
                                                 my->sACC, my->sACCi, cp, CPF[p]);
#pragma unroll
                    for (int i = 0; i < NC; i++) cptBuf[mfma_index(i, lsys, uid(p), UPS)] = cp[i];
                }
            }
            __syncthreads();
            if (a.sharedMatrix) mfma_contract44<NC, false>(M, cptBuf, coupBuf, UPS, SPB);
            else mfma_contract44<NC, true, 8, mfma_tiles_together<Sys, PAIRS>()>(M, cptBuf, coupBuf, UPS, SPB, (const double*)a.cTile, Ml,
                                                                                 ldsPairs);
            __syncthreads();
            if (run) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++)
#pragma unroll
                    for (int i = 0; i < NC; i++)
                        U::add(cIndex[i], my->strength[i] * CPF[p][i], coupBuf[mfma_index(i, lsys, uid(p), UPS)], F[p]);
            }
        } else {
            P* cpt = cptBuf + (size_t)buf * cptStride;
            if (run) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++)
                    Sys::template OdeFunction<P>(gsys, uid(p), F[p], Y[p], T, uPAR[p], my->sPAR, gPAR, gPARi, uACC[p], uACCi[p],
                                                 my->sACC, my->sACCi, cpt + (size_t)(lsys * UPS + uid(p)) * NCp, CPF[p]);
            }
            __syncthreads();
            if (run) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++)
                    U::couple(a, M, cIndex, my->strength, cpt + (size_t)lsys * UPS * NCp, uid(p), CPF[p], F[p]);
            }
            buf ^= 1;
        }
    };

    for (;;) {
        // ---- step initialisation (PB-SOL:2535-2549)
        if (leader && my->term == 0) {
            my->update = 1; my->isfin = 1; my->end = 0;
            my->dt = my->newDt;
            CP remaining;
            if (KAHAN) remaining = (CP)my->TD[1] - ((CP)my->t - (CP)my->tc);
            else remaining = my->TD[1] - my->t;
            if (my->dt > remaining) { my->dt = remaining; my->end = 1; }
            if (KAHAN) my->dtP = (P)my->dt;
            my->ratioBits = order_bits<CP>(CP(1e30));
            my->eventBits = ~0ull;            // no candidate yet (above every finite step)
            my->corrected = 0;
        }
        __syncthreads();
        const P h = (P)my->dt, t = my->t;
        P next[PAIRS][UD], inc[PAIRS][KAHAN ? UD : 1];
        CP err[PAIRS][UD];
        bool isfin = true;
        if (ALG == MPGOS_RK4) {
            // PB-STP:2016-2224
            P k[PAIRS][UD], Y[PAIRS][UD];
            rhs(next, X, t);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) Y[p][c] = fma(next[p][c] * P(0.5), h, X[p][c]);
            const P t2 = t + P(0.5) * h;
            rhs(k, Y, t2);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) { next[p][c] = fma(P(2), k[p][c], next[p][c]); Y[p][c] = fma(k[p][c] * P(0.5), h, X[p][c]); }
            rhs(k, Y, t2);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) { next[p][c] = fma(P(2.0), k[p][c], next[p][c]); Y[p][c] = fma(k[p][c], h, X[p][c]); }
            rhs(k, Y, t + h);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) {
                    if (KAHAN) { inc[p][c] = h * (next[p][c] + k[p][c]) * P(1.0 / 6.0); next[p][c] = X[p][c] + inc[p][c]; }
                    else next[p][c] = fma(h * (next[p][c] + k[p][c]), P(1.0 / 6.0), X[p][c]);
                    if (!math::isfinite(next[p][c])) isfin = false;
                }
        } else {
            // PB-STP:2227-2544
            P k1[PAIRS][UD], k2[PAIRS][UD], k3[PAIRS][UD], k4[PAIRS][UD], k5[PAIRS][UD], k6[PAIRS][UD], Y[PAIRS][UD];
            rhs(k1, X, t);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) Y[p][c] = fma(k1[p][c] * P(1.0 / 5.0), h, X[p][c]);
            rhs(k2, Y, fma(P(1.0 / 5.0), h, t));
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) Y[p][c] = fma(fma(k2[p][c], P(9.0 / 40.0), k1[p][c] * P(3.0 / 40.0)), h, X[p][c]);
            rhs(k3, Y, fma(P(3.0 / 10.0), h, t));
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++)
                    Y[p][c] = fma(fma(k3[p][c], P(6.0 / 5.0), fma(k2[p][c], P(-9.0 / 10.0), k1[p][c] * P(3.0 / 10.0))), h, X[p][c]);
            rhs(k4, Y, fma(P(3.0 / 5.0), h, t));
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++)
                    Y[p][c] = fma(fma(k4[p][c], P(35.0 / 27.0), fma(k3[p][c], P(-70.0 / 27.0), fma(k2[p][c], P(5.0 / 2.0),
                                  k1[p][c] * P(-11.0 / 54.0)))), h, X[p][c]);
            rhs(k5, Y, t + h);
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++)
                    Y[p][c] = fma(fma(k5[p][c], P(253.0 / 4096.0), fma(k4[p][c], P(44275.0 / 110592.0), fma(k3[p][c], P(575.0 / 13824.0),
                                  fma(k2[p][c], P(175.0 / 512.0), k1[p][c] * P(1631.0 / 55296.0))))), h, X[p][c]);
            rhs(k6, Y, fma(P(7.0 / 8.0), h, t));
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) {
                    const P sum = fma(k6[p][c], P(512.0 / 1771.0), fma(k4[p][c], P(125.0 / 594.0), fma(k3[p][c], P(250.0 / 621.0),
                                      k1[p][c] * P(37.0 / 378.0))));
                    if (KAHAN) {
                        inc[p][c] = sum * h;
                        next[p][c] = X[p][c] + inc[p][c];
                        // FP64 error norm over the FP32 stages
                        double e = (double)k1[p][c] * (37.0 / 378.0 - 2825.0 / 27648.0);
                        e = math::fma((double)k3[p][c], 250.0 / 621.0 - 18575.0 / 48384.0, e);
                        e = math::fma((double)k4[p][c], 125.0 / 594.0 - 13525.0 / 55296.0, e);
                        e = math::fma((double)k5[p][c], 0.0 - 277.0 / 14336.0, e);
                        e = math::fma((double)k6[p][c], 512.0 / 1771.0 - 1.0 / 4.0, e);
                        err[p][c] = math::fma((double)h, math::fabs(e), 1e-18);
                    } else {
                        next[p][c] = fma(sum, h, X[p][c]);
                        P e = k1[p][c] * P(37.0 / 378.0 - 2825.0 / 27648.0);
                        e = fma(k3[p][c], P(250.0 / 621.0 - 18575.0 / 48384.0), e);
                        e = fma(k4[p][c], P(125.0 / 594.0 - 13525.0 / 55296.0), e);
                        e = fma(k5[p][c], P(0.0 - 277.0 / 14336.0), e);
                        e = fma(k6[p][c], P(512.0 / 1771.0 - 1.0 / 4.0), e);
                        err[p][c] = (CP)fma(h, math::fabs(e), P(1e-18));
                    }
                    if (!math::isfinite(next[p][c]) || !math::isfinite(err[p][c])) isfin = false;
                }
        }
        if (lane && my->term == 0 && !isfin) my->isfin = 0;
        // ---- controller (PB-CTL:424-542)
        if (ALG == MPGOS_RKCK45) {
            const bool part = lane && my->term == 0;
            CP R = CP(1e30);                    // the reduction's start value, also for lanes not taking part
            int upd = 1;
            if (part) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++)
#pragma unroll
                    for (int c = 0; c < UD; c++) {
                        const CP tol = dmax<CP>(a.rtol[c] * dmax<CP>(math::fabs((CP)next[p][c]), math::fabs((CP)X[p][c])), a.atol[c]);
                        upd = upd & (err[p][c] < tol);
                        R = dmin<CP>(R, tol / err[p][c]);
                    }
            }
            // systems on aligned power-of-two groups of lanes (or whole waves):
            // reduce in the group with lane shuffles, then one LDS atomic per
            // group instead of one per lane (same-address atomics serialise).
            // min and AND do not depend on the order, so the bits are those of
            // the per-lane atomics.  Only where the lane's stage vectors leave
            // registers for it (the two-tile dense-contraction kernels: ring of
            // 256 two-component units 107.2 -> 106.5 ms; with three components
            // the extra registers spilled, 125.8 -> 131.3 ms).
            const int grp = !(CPL == kMfmaCoupling && mfma_tiles_together<Sys, PAIRS>() > 1) ? 0
                            : lanesPerSys >= 64 ? (lanesPerSys % 64 == 0 ? 64 : 0)
                                                : ((lanesPerSys & (lanesPerSys - 1)) == 0 ? lanesPerSys : 0);
            if (grp > 1) {
                for (int m = 1; m < grp; m <<= 1) {
                    R = dmin<CP>(R, __shfl_xor(R, m));
                    upd &= __shfl_xor(upd, m);
                }
            }
            if (part && (grp == 0 || (tid & (grp - 1)) == 0)) {
                atomicAnd(&my->update, upd);
                atomicMin(&my->ratioBits, order_bits<CP>(R));
            }
            __syncthreads();
            if (leader && my->term == 0) {
                const CP R = from_order_bits<CP>(my->ratioBits);
                CP fac = CP(0.8) * math::pow_ctrl(R, my->update == 1 ? CP(1.0 / 5.0) : CP(1.0 / 4.0));
                if (!math::isfinite(fac)) my->isfin = 0;
                if (my->isfin == 0) {
                    fac = a.shrink;
                    my->update = 0;
                    if (my->dt < a.dtmin * CP(1.01)) { my->term = 1; my->status |= MPGOS_STATUS_NONFINITE_MIN_STEP; }
                } else if (my->dt < a.dtmin * CP(1.01)) {
                    my->update = 1;
                    my->status |= MPGOS_STATUS_MIN_STEP_REACHED;
                }
                fac = dmin<CP>(fac, a.grow);
                fac = dmax<CP>(fac, a.shrink);
                CP nd = my->dt * fac;
                nd = dmin<CP>(nd, a.dtmax);
                nd = dmax<CP>(nd, a.dtmin);
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
        P EFn[PAIRS][Pos<NE>::value];
        if (NE > 0) {
            const P tn = my->t + (P)my->dt;
            if (lane && my->term == 0) {
                bool corr = false;
                CP ev = my->dt;
#pragma unroll
                for (int p = 0; p < PAIRS; p++) {
                    // the event function of the candidate state, at the end of the step
                    Sys::template EventFunction<P>(gsys, uid(p), EFn[p], tn, (P)my->newDt, my->TD, next[p], uPAR[p], my->sPAR, gPAR,
                                                   gPARi, uACC[p], uACCi[p], my->sACC, my->sACCi);
                    if (my->update == 1) {
#pragma unroll
                        for (int i = 0; i < NE; i++) {
                            const P tol = a.etol[i];
                            if ((EF[p][i] > tol && EFn[p][i] < -tol && a.edir[i] <= 0) ||
                                (EF[p][i] < -tol && EFn[p][i] > tol && a.edir[i] >= 0)) {
                                ev = dmin<CP>(ev, -(CP)EF[p][i] / ((CP)EFn[p][i] - (CP)EF[p][i]) * my->dt);
                                corr = true;
                            }
                        }
                    }
                }
                if (corr) {
                    atomicMin(&my->eventBits, order_bits<CP>(ev));
                    atomicOr(&my->corrected, 1);
                }
            }
            __syncthreads();
            if (leader && my->corrected == 1) {
                const CP ev = dmin<CP>(my->dt, from_order_bits<CP>(my->eventBits));   // corrected: one candidate at least
                if (ev < a.dtmin) my->status |= MPGOS_STATUS_EVENT_OMITTED;
                else { my->newDt = ev; my->update = 0; }
            }
            __syncthreads();
        }
        // ---- update (PB-SOL:2670-2681)
        const int upd = my->update;
        if (lane && upd == 1) {
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
#pragma unroll
                for (int c = 0; c < UD; c++) {
                    if (KAHAN) {   // compensated x += inc
                        const P y = inc[p][c] - Xc[p][c];
                        const P sx = X[p][c] + y;
                        Xc[p][c] = (sx - X[p][c]) - y;
                        X[p][c] = sx;
                    } else {
                        X[p][c] = next[p][c];
                    }
                }
            if (REC) {
                const long long n = my->nacc;            // accepted steps before this one
                if (leader) {
                    if (n < a.recStepCap) { a.recT[(size_t)n * NS + gsys] = my->t; a.recH[(size_t)n * NS + gsys] = my->dt; }
                    else my->status |= MPGOS_STATUS_CHECKPOINT_OVERFLOW;
                }
                if ((n + 1) % a.recK == 0) {       // once per accepted step of the block, not per stage
                    const long long k = (n + 1) / a.recK;
                    if (k < a.recCkCap) {
#pragma unroll
                        for (int p = 0; p < PAIRS; p++)
#pragma unroll
                            for (int c = 0; c < UD; c++) a.recX[((size_t)k * UD + c) * NU + gu(p)] = X[p][c];
                    } else if (leader) {
                        my->status |= MPGOS_STATUS_CHECKPOINT_OVERFLOW;
                    }
                }
            }
        }
        __syncthreads();
        if (leader && upd == 1) {
            if (KAHAN) {   // compensated t += dt
                const P y = (P)my->dt - my->tc;
                const P st = my->t + y;
                my->tc = (st - my->t) - y;
                my->t = st;
            } else {
                my->t += my->dt;
            }
        }
        __syncthreads();
        if (lane && upd == 1) {
#pragma unroll
            for (int p = 0; p < PAIRS; p++)
                Sys::template ActionAfterSuccessfulTimeStep<P>(gsys, uid(p), my->udt, MPGOS_PB_USER(p, my->t, udt()));
        }
        __syncthreads();
        if (NE > 0) {
            if (lane && upd == 1) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++)
#pragma unroll
                    for (int i = 0; i < NE; i++) {
                        const P tol = a.etol[i];
                        if ((EF[p][i] > tol && math::fabs(EFn[p][i]) < tol && a.edir[i] <= 0) ||
                            (EF[p][i] < -tol && math::fabs(EFn[p][i]) < tol && a.edir[i] >= 0))
                            Sys::template ActionAfterEventDetection<P>(gsys, uid(p), i, my->udt, MPGOS_PB_USER(p, my->t, udt()));
                    }
            }
            __syncthreads();
            if (lane && upd == 1) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++) Sys::template EventFunction<P>(gsys, uid(p), EFn[p], MPGOS_PB_USER(p, my->t, (P)my->newDt));
            }
            __syncthreads();
            if (lane && upd == 1) {
#pragma unroll
                for (int p = 0; p < PAIRS; p++)
#pragma unroll
                    for (int i = 0; i < NE; i++) EF[p][i] = EFn[p][i];
            }
        }
        // ---- dense output (PB-DO:312-395)
        if (a.NDO > 0) {
            if (leader) {
                if (upd == 1) {
                    int st = (my->DOidx < a.NDO && my->doT < my->t && my->skipped >= a.doFreq - 1) ? 1 : 0;
                    if (my->DOidx < a.NDO && (my->end == 1 || my->udt == 1)) st = 1;
                    my->updDO = st;
                } else {
                    my->updDO = 0;
                }
            }
            __syncthreads();
            if (lane && my->updDO == 1 && upd == 1) {
                const int k = my->DOidx;
#pragma unroll
                for (int p = 0; p < PAIRS; p++)
#pragma unroll
                    for (int c = 0; c < UD; c++) a.DOx[gu(p) + c * NU + (size_t)k * NU * UD] = X[p][c];
                if (leader) a.DOt[gsys + (size_t)k * NS] = my->t;
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
        for (int q = 0; q < SPB; q++) done &= sys[q].term;
        if (done) break;
        __syncthreads();
    }
    __syncthreads();
    // ---- finalisation and write-back (PB-SOL:2814-2887)
    if (active) {
#pragma unroll
        for (int p = 0; p < PAIRS; p++) Sys::template Finalization<P>(gsys, uid(p), my->DOidx, MPGOS_PB_USER(p, my->t, udt()));
    }
    __syncthreads();
    if (exists) {
#pragma unroll
        for (int p = 0; p < PAIRS; p++) {
            const size_t g = gu(p);
#pragma unroll
            for (int c = 0; c < UD; c++) a.X[g + c * NU] = X[p][c];
#pragma unroll
            for (int c = 0; c < NUA; c++) a.uACC[g + c * NU] = uACC[p][c];
#pragma unroll
            for (int c = 0; c < NiUA; c++) a.uACCi[g + c * NU] = uACCi[p][c];
        }
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

// Lane-per-pair form; MAXT bounds the block size so that blocks of up to 256
// threads get the larger register budget.
template <class Sys, class P, int ALG, int MAXT, bool KAHAN = false>
__global__ void __launch_bounds__(MAXT) pb_kernel(Args<Sys, P, typename SysScope<Sys, P, KAHAN>::CP> a)
{
    pb_body<Sys, P, ALG, 1, kLaneCoupling, false, KAHAN>(a);
}

// Record pass of the discrete adjoint (FP64): the forward integration plus
// checkpoints (lane-per-pair or dense-contraction form).
template <class Sys, int ALG, int MAXT>
__global__ void __launch_bounds__(MAXT) pb_record_kernel(Args<Sys, double> a)
{
    pb_body<Sys, double, ALG, 1, kLaneCoupling, true>(a);
}
template <class Sys, int ALG>
__global__ void __launch_bounds__(512) pb_record_mfma_kernel(Args<Sys, double> a)
{
    pb_body<Sys, double, ALG, 4, kMfmaCoupling, true>(a);
}

// Dense-contraction form: SPB = 8, PAIRS units per lane, 8 * UPS / PAIRS
// threads (UPS a multiple of 32 up to 256, so every wave is full).
template <class Sys, class P, int ALG, int PAIRS>
__global__ void __launch_bounds__(8 * 256 / PAIRS) pb_mfma_kernel(Args<Sys, P> a)
{
    if constexpr (std::is_same<P, double>::value)
        pb_body<Sys, P, ALG, PAIRS, kMfmaCoupling>(a);
}

} // namespace pb
} // namespace mpgos
