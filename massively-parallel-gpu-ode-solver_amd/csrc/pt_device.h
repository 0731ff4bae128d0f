// Per-thread solver for gfx950: one lane integrates one instance.
//
// Semantics follow the reference kernel SingleSystem_PerThread
// (SourceCodes/SingleSystem_PerThread_Solver.cuh:11-437) step for step:
// prologue (loads, Initialization, first event evaluation, first dense store),
// the loop  clip-to-end -> stepper -> controller -> event step control ->
// [accept: t,x update, ActionAfterSuccessfulTimeStep, event detection and
// ActionAfterEventDetection, event re-evaluation, dense output, termination],
// then Finalization and write-back.  The numerics are
//   RK4     SingleSystem_PerThread_ExplicitRungeKutta_Steppers.cuh:6-117
//   RKCK45  SingleSystem_PerThread_ExplicitRungeKutta_Steppers.cuh:121-290
//   controllers  SingleSystem_PerThread_ExplicitRungeKutta_ErrorControllers.cuh:5-86
//   event step control  SingleSystem_PerThread_EventHandling.cuh:5-45
//   dense output  SingleSystem_PerThread_DenseOutput.cuh:5-58
// with every multiply-add of the stage combinations written as an explicit
// fused multiply-add (the form a contracting compiler gives the reference's
// expressions) and the rest compiled with -ffp-contract=off, so the FP64 result
// is reproducible bit for bit by the host oracle.
//
// Design for MI355X (not present in the reference):
//  * instance state lives in VGPRs; SoA HBM layout (index tid + c*NT) keeps
//    every load/store of a wavefront contiguous;
//  * tolerances and solver options arrive by value in the kernel argument
//    block, so they sit in SGPRs; shared parameters are staged in LDS when
//    PreferSharedMemory is on (the reference default);
//  * device printf is replaced by per-instance status bits and counters;
//  * a persistent form pulls instances from a device work queue and refills
//    lanes whose instance finished, found with a wavefront ballot;
//  * FLOAT_KAHAN mode: FP32 state and stages, compensated accumulation of
//    time and state, FP64 error norm, controller and event step control.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <type_traits>

#include "mpgos_math.h"
#include "mpgos.h"

namespace mpgos {
namespace pt {

constexpr int kMaxBlock = 512;
constexpr int kMaxWaves = kMaxBlock / 64;

template <int N> struct Pos { static constexpr int value = N > 0 ? N : 1; };

// Kernel argument block.  CP is the precision of the step-size control.
template <class Sys, class P, class CP>
struct Args {
    static constexpr int SD = Sys::SD;
    static constexpr int NE = Sys::NE;
    // SoA buffers (reference layout)
    P* TD; P* X; P* T; P* cPAR; P* sPAR; int* sPARi; P* ACC; int* ACCi;
    int* DOidx; P* DOt; P* DOx;
    // per-instance outputs
    long long* nacc; long long* nrej; int* status; int* nev; P* evlog; int* evlogidx;
    unsigned int* queue;      // persistent form: [0] = next list entry past the solo ones; [1] = suspended count
                              // of a slice; [2] = next solo entry
    // time-sliced solve: the instances to work on are list[0, listLen) (the
    // identity 0..nActive-1 when list is null); with slice > 0 an instance
    // that has attempted `slice` steps in this launch is suspended (in the
    // relaunch form's last launch, listOut null, it stops there with
    // MPGOS_STATUS_STEP_BUDGET instead): its lane
    // state goes to its record saved[id * W .. id * W + W) (W 32-bit words,
    // 16-byte aligned), phase[id] = 1, and its id is appended to listOut
    // (count in queue[1]); phase[id] = 2 once finished
    const int* list; int* listOut; uint32_t* saved; int* phase;
    int listLen, slice;
    const unsigned* listLenDev;   // relaunch form: the list length is the previous launch's count (device)
    unsigned* outCount;           // relaunch form: count of this launch's suspended instances (else queue[1])
    float* sortKey;               // sorted form's pilot launch: estimated remaining attempted steps of each
                                  // suspended instance, at its slot of listOut
    int NT, nActive, NDO, evcap, preferShared;
    int doFreq;
    int roundSteps;           // persistent form: attempted steps between compactions
    int solo;                 // persistent form: list entries [0, solo) run one per workgroup (workgroups [0, solo))
    CP dt0, dtmax, dtmin, grow, shrink;
    P doMinDt;
    CP rtol[Pos<SD>::value], atol[Pos<SD>::value];
    P etol[Pos<NE>::value];
    int edir[Pos<NE>::value];
};

// The kernel's argument block read afresh from the kernel-argument segment,
// through a pointer the compiler cannot see through.  Used on the rare paths
// (finishing or suspending an instance, dense-output and event-log stores):
// the fields they read are then loaded where they are used instead of being
// held in SGPRs, and spilled to VGPR lanes, across the whole step loop.  Valid
// only in kernels whose argument block begins with the Args: solve_kernel and
// the adjoint kernels (whose AdjArgs start with it, asserted there).
template <class A>
__device__ __forceinline__ const A& kernel_args()
{
    typedef __attribute__((address_space(4))) const A* KP;
    KP p = (KP)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const A*)p;
}

template <class Sys, class P, bool KAHAN, int ALG>
struct Lane {
    using CP = typename std::conditional<KAHAN, double, P>::type;
    using A = Args<Sys, P, CP>;
    using Sys_t = Sys;
    static constexpr bool kKahan = KAHAN;
    static constexpr int SD = Sys::SD, NCP = Sys::NCP, NE = Sys::NE, NA = Sys::NA, NIA = Sys::NIA;

    P TD[2];
    P X[SD];
    P Xc[KAHAN ? SD : 1];                 // Kahan compensation of the state
    P cPAR[Pos<NCP>::value];
    P ACC[Pos<NA>::value];
    int ACCi[Pos<NIA>::value];
    P EF[Pos<NE>::value];
    P t;
    P tc;                                 // Kahan compensation of the time
    CP dt, newDt;
    P doT;
    int doIdx, skipped, term, udt;
    int nacc, nrej, status, nev;
    int tid;
    int left;                             // time-sliced solve: attempted steps left in this launch

    // whole lane state to and from HBM (time-sliced solve), bit for bit
    // one 16-byte aligned record per instance (rare: once per slice), moved
    // with 16-byte accesses from one base address per lane
    static constexpr int kSavedWords = (int)((sizeof(Lane) / 4 + 3) & ~(size_t)3);
    __device__ __forceinline__ void save(const A& a) const
    {
        uint32_t w[kSavedWords] = {};
        __builtin_memcpy(w, this, sizeof(Lane));
        uint4* dst = reinterpret_cast<uint4*>(a.saved + (size_t)tid * kSavedWords);
#pragma unroll
        for (int i = 0; i < kSavedWords / 4; i++) dst[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
    }
    __device__ __forceinline__ void restore(const A& a, int id)
    {
        uint32_t w[kSavedWords];
        const uint4* src = reinterpret_cast<const uint4*>(a.saved + (size_t)id * kSavedWords);
#pragma unroll
        for (int i = 0; i < kSavedWords / 4; i++) {
            const uint4 v = src[i];
            w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
        }
        __builtin_memcpy(this, w, sizeof(Lane));
    }
    // the instance's first launch initialises it, later launches resume it
    __device__ __forceinline__ void begin(const A& a, int id, P* sPAR, int* sPARi)
    {
        if (a.phase[id] == 1) restore(a, id);
        else load(a, id, sPAR, sPARi);
        left = a.slice;
    }
    // after a step: true when the lane is to be released (finished, or its
    // slice budget spent); release() then finishes or suspends it
    __device__ __forceinline__ bool stop() { return term || --left <= 0; }
    __device__ __forceinline__ void release(const A& a, P* sPAR, int* sPARi)
    {
        if (term) {
            finish(a, sPAR, sPARi);
            a.phase[tid] = 2;
            return;
        }
        if (!a.listOut) {   // the relaunch form's last launch: no later launch resumes the instance
            status |= MPGOS_STATUS_STEP_BUDGET;
            finish(a, sPAR, sPARi);
            a.phase[tid] = 2;
            return;
        }
        save(a);
        a.phase[tid] = 1;
        const unsigned slot = atomicAdd(a.outCount ? a.outCount : &a.queue[1], 1u);
        a.listOut[slot] = tid;
        if (a.sortKey) {
            // remaining attempts at the rate so far: attempts * (TD1 - t) / (t - TD0)
            const double done = (double)t - (double)TD[0], rest = (double)TD[1] - (double)t;
            const double att = (double)(nacc + nrej);
            a.sortKey[slot] = done > 0.0 ? (float)__builtin_fmin(att * rest / done, 3.0e38) : 3.0e38f;
        }
    }

    // ------------------------------------------------------------ load / prologue
    __device__ __forceinline__ void load(const A& a, int id, P* sPAR, int* sPARi)
    {
        tid = id;
        const int NT = a.NT;
#pragma unroll
        for (int i = 0; i < 2; i++) TD[i] = a.TD[tid + i * NT];
#pragma unroll
        for (int i = 0; i < SD; i++) { X[i] = a.X[tid + i * NT]; if (KAHAN) Xc[i] = P(0); }
#pragma unroll
        for (int i = 0; i < NCP; i++) cPAR[i] = a.cPAR[tid + i * NT];
#pragma unroll
        for (int i = 0; i < NA; i++) ACC[i] = a.ACC[tid + i * NT];
#pragma unroll
        for (int i = 0; i < NIA; i++) ACCi[i] = a.ACCi[tid + i * NT];
        t = a.T[tid];
        tc = P(0);
        dt = a.dt0;
        newDt = a.dt0;
        doIdx = a.DOidx[tid];
        skipped = 0;
        term = 0;
        udt = 0;
        nacc = 0; nrej = 0; status = 0; nev = 0;
        doT = P(0);

        // PT-SOL:148-193
        P dtP = (P)dt;
        Sys::template Initialization<P>(tid, NT, doIdx, t, dtP, TD, X, cPAR, sPAR, sPARi, ACC, ACCi);
        dt = (CP)dtP;
        if (NE > 0)
            Sys::template EventFunction<P>(tid, NT, EF, t, dtP, TD, X, cPAR, sPAR, sPARi, ACC, ACCi);
        if (a.NDO > 0 && doIdx < a.NDO) store_dense(a);
    }

    __device__ __forceinline__ void store_dense(const A& a)
    {
        // PT-DO:5-37 (store branch)
        const int NT = a.NT;
        a.DOt[tid + doIdx * NT] = t;
        int k = tid + doIdx * NT * SD;
#pragma unroll
        for (int i = 0; i < SD; i++) { a.DOx[k] = X[i]; k += NT; }
        doIdx++;
        skipped = 0;
        doT = mpgos_fmin(t + a.doMinDt, TD[1]);
    }

    static __device__ __forceinline__ double mpgos_fmin(double x, double y) { return __builtin_fmin(x, y); }
    static __device__ __forceinline__ float mpgos_fmin(float x, float y) { return __builtin_fminf(x, y); }
    static __device__ __forceinline__ double mpgos_fmax(double x, double y) { return __builtin_fmax(x, y); }
    static __device__ __forceinline__ float mpgos_fmax(float x, float y) { return __builtin_fmaxf(x, y); }

    // ------------------------------------------------------------ steppers
    // RK4: PT-STP:6-117.  Returns next state in `nx`; in Kahan mode `inc` gets
    // the increment dt/6 (k1 + 2k2 + 2k3 + k4).
    __device__ __forceinline__ void rk4(const A& a, P h, P* nx, P* inc, int& isfin, P* sPAR, int* sPARi)
    {
        using math::fma;
        P Xs[SD], k[SD];
        const P dTp2 = P(0.5) * h;
        const P dTp6 = P(1.0 / 6.0) * h;
        Sys::template OdeFunction<P>(tid, a.NT, nx, X, t, cPAR, sPAR, sPARi, ACC, ACCi);
        P T = t + dTp2;
#pragma unroll
        for (int i = 0; i < SD; i++) Xs[i] = fma(nx[i], dTp2, X[i]);
        Sys::template OdeFunction<P>(tid, a.NT, k, Xs, T, cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
        for (int i = 0; i < SD; i++) { nx[i] = fma(P(2.0), k[i], nx[i]); Xs[i] = fma(k[i], dTp2, X[i]); }
        Sys::template OdeFunction<P>(tid, a.NT, k, Xs, T, cPAR, sPAR, sPARi, ACC, ACCi);
        T = t + h;
#pragma unroll
        for (int i = 0; i < SD; i++) { nx[i] = fma(P(2.0), k[i], nx[i]); Xs[i] = fma(k[i], h, X[i]); }
        Sys::template OdeFunction<P>(tid, a.NT, k, Xs, T, cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
        for (int i = 0; i < SD; i++) {
            const P s = nx[i] + k[i];
            if (KAHAN) { inc[i] = dTp6 * s; nx[i] = X[i] + inc[i]; }
            else nx[i] = fma(dTp6, s, X[i]);
            if (!math::isfinite(nx[i])) isfin = 0;
        }
    }

    // RKCK45: PT-STP:121-290.  err[] in control precision.
    __device__ __forceinline__ void rkck45(const A& a, P h, P* nx, P* inc, CP* err, int& isfin, P* sPAR, int* sPARi)
    {
        using math::fma;
        P Xs[SD], k1[SD], k2[SD], k3[SD], k4[SD], k5[SD], k6[SD];
        Sys::template OdeFunction<P>(tid, a.NT, k1, X, t, cPAR, sPAR, sPARi, ACC, ACCi);

        P T = fma(h, P(1.0 / 5.0), t);
#pragma unroll
        for (int i = 0; i < SD; i++) Xs[i] = fma(h, P(1.0 / 5.0) * k1[i], X[i]);
        Sys::template OdeFunction<P>(tid, a.NT, k2, Xs, T, cPAR, sPAR, sPARi, ACC, ACCi);

        T = fma(h, P(3.0 / 10.0), t);
#pragma unroll
        for (int i = 0; i < SD; i++) {
            P s = P(3.0 / 40.0) * k1[i];
            s = fma(P(9.0 / 40.0), k2[i], s);
            Xs[i] = fma(h, s, X[i]);
        }
        Sys::template OdeFunction<P>(tid, a.NT, k3, Xs, T, cPAR, sPAR, sPARi, ACC, ACCi);

        T = fma(h, P(3.0 / 5.0), t);
#pragma unroll
        for (int i = 0; i < SD; i++) {
            P s = P(3.0 / 10.0) * k1[i];
            s = fma(P(-9.0 / 10.0), k2[i], s);
            s = fma(P(6.0 / 5.0), k3[i], s);
            Xs[i] = fma(h, s, X[i]);
        }
        Sys::template OdeFunction<P>(tid, a.NT, k4, Xs, T, cPAR, sPAR, sPARi, ACC, ACCi);

        T = t + h;
#pragma unroll
        for (int i = 0; i < SD; i++) {
            P s = P(-11.0 / 54.0) * k1[i];
            s = fma(P(5.0 / 2.0), k2[i], s);
            s = fma(P(-70.0 / 27.0), k3[i], s);
            s = fma(P(35.0 / 27.0), k4[i], s);
            Xs[i] = fma(h, s, X[i]);
        }
        Sys::template OdeFunction<P>(tid, a.NT, k5, Xs, T, cPAR, sPAR, sPARi, ACC, ACCi);

        T = fma(h, P(7.0 / 8.0), t);
#pragma unroll
        for (int i = 0; i < SD; i++) {
            P s = P(1631.0 / 55296.0) * k1[i];
            s = fma(P(175.0 / 512.0), k2[i], s);
            s = fma(P(575.0 / 13824.0), k3[i], s);
            s = fma(P(44275.0 / 110592.0), k4[i], s);
            s = fma(P(253.0 / 4096.0), k5[i], s);
            Xs[i] = fma(h, s, X[i]);
        }
        Sys::template OdeFunction<P>(tid, a.NT, k6, Xs, T, cPAR, sPAR, sPARi, ACC, ACCi);

#pragma unroll
        for (int i = 0; i < SD; i++) {
            P s = P(37.0 / 378.0) * k1[i];
            s = fma(P(250.0 / 621.0), k3[i], s);
            s = fma(P(125.0 / 594.0), k4[i], s);
            s = fma(P(512.0 / 1771.0), k6[i], s);
            if (KAHAN) { inc[i] = h * s; nx[i] = X[i] + inc[i]; }
            else nx[i] = fma(h, s, X[i]);

            if (KAHAN) {
                // FP64 error norm over the FP32 stages
                double e = (37.0 / 378.0 - 2825.0 / 27648.0) * (double)k1[i];
                e = math::fma(250.0 / 621.0 - 18575.0 / 48384.0, (double)k3[i], e);
                e = math::fma(125.0 / 594.0 - 13525.0 / 55296.0, (double)k4[i], e);
                e = math::fma(0.0 - 277.0 / 14336.0, (double)k5[i], e);
                e = math::fma(512.0 / 1771.0 - 1.0 / 4.0, (double)k6[i], e);
                err[i] = math::fma((double)h, math::fabs(e), 1e-18);
            } else {
                P e = P(37.0 / 378.0 - 2825.0 / 27648.0) * k1[i];
                e = fma(P(250.0 / 621.0 - 18575.0 / 48384.0), k3[i], e);
                e = fma(P(125.0 / 594.0 - 13525.0 / 55296.0), k4[i], e);
                e = fma(P(0.0 - 277.0 / 14336.0), k5[i], e);
                e = fma(P(512.0 / 1771.0 - 1.0 / 4.0), k6[i], e);
                if (std::is_same<P, double>::value) err[i] = (CP)math::fma((double)h, (double)math::fabs(e), 1e-18);
                else err[i] = (CP)((double)(h * math::fabs(e)) + 1e-18);   // float*float, then + double literal
            }
            if (!math::isfinite(nx[i]) || !math::isfinite(err[i])) isfin = 0;
        }
    }

    // ------------------------------------------------------------ controllers
    // PT-CTL:23-86
    __device__ __forceinline__ void control_rkck45(const A& a, const P* nx, const CP* err, int& update, int& isfin)
    {
        CP R = CP(1e30);
#pragma unroll
        for (int i = 0; i < SD; i++) {
            const CP tol = mpgos_fmax(a.rtol[i] * mpgos_fmax(math::fabs((CP)nx[i]), math::fabs((CP)X[i])), a.atol[i]);
            update = update && (err[i] < tol);
            R = mpgos_fmin(R, tol / err[i]);
        }
        // one power for both outcomes: lanes of a wavefront that accept and
        // lanes that reject then share one evaluation
        CP fac = CP(0.9) * math::ctrl_root(R, update == 1);
        if (!math::isfinite(fac)) isfin = 0;

        if (isfin == 0) {
            fac = a.shrink;
            update = 0;
            if (dt < a.dtmin * CP(1.01)) { term = 1; status |= MPGOS_STATUS_NONFINITE_MIN_STEP; }
        } else {
            if (dt < a.dtmin * CP(1.01)) { update = 1; status |= MPGOS_STATUS_MIN_STEP_REACHED; }
        }
        fac = mpgos_fmin(fac, a.grow);
        fac = mpgos_fmax(fac, a.shrink);
        newDt = dt * fac;
        newDt = mpgos_fmin(newDt, a.dtmax);
        newDt = mpgos_fmax(newDt, a.dtmin);
    }

    // PT-EVT:5-45
    __device__ __forceinline__ void event_control(const A& a, const P* EFn, int& update)
    {
        CP evDt = dt;
        int corrected = 0;
        if (update == 1 && term == 0) {
#pragma unroll
            for (int i = 0; i < NE; i++) {
                const P tol = a.etol[i];
                if ((EF[i] > tol && EFn[i] < -tol && a.edir[i] <= 0) ||
                    (EF[i] < -tol && EFn[i] > tol && a.edir[i] >= 0)) {
                    evDt = mpgos_fmin(evDt, (-(CP)EF[i] / ((CP)EFn[i] - (CP)EF[i])) * dt);
                    corrected = 1;
                }
            }
        }
        if (corrected == 1) {
            if (evDt < a.dtmin) status |= MPGOS_STATUS_EVENT_OMITTED;
            else { newDt = evDt; update = 0; }
        }
    }

    // ------------------------------------------------------------ one attempted step
    // One iteration of the loop at PT-SOL:197-394.  DENSE = false compiles the
    // dense-output bookkeeping out (the adjoint's re-steps, which run with
    // NDO = 0), so that the step has no branch on it.
    template <bool DENSE = true>
    __device__ __forceinline__ void step(const A& a, P* sPAR, int* sPARi)
    {
        const int NT = a.NT;
        int update = 1, isfin = 1, endReached = 0;
        dt = newDt;
        CP remaining;
        if (KAHAN) remaining = (CP)TD[1] - ((CP)t - (CP)tc);
        else remaining = (CP)(TD[1] - t);
        if (dt > remaining) { dt = remaining; endReached = 1; }

        const P h = (P)dt;
        P nx[SD], inc[KAHAN ? SD : 1];
        if (ALG == MPGOS_RK4) {
            rk4(a, h, nx, inc, isfin, sPAR, sPARi);
            // PT-CTL:5-20
            if (isfin == 0) { term = 1; status |= MPGOS_STATUS_NONFINITE_RK4; }
            newDt = a.dt0;
        } else {
            CP err[SD];
            rkck45(a, h, nx, inc, err, isfin, sPAR, sPARi);
            control_rkck45(a, nx, err, update, isfin);
        }

        P EFn[Pos<NE>::value];
        P dtP = h;
        if (NE > 0) {
            Sys::template EventFunction<P>(tid, NT, EFn, t + h, dtP, TD, nx, cPAR, sPAR, sPARi, ACC, ACCi);
            event_control(a, EFn, update);
        }

        if (update == 1) {
            if (KAHAN) {
                // compensated t += h and x += inc
                P y = h - tc; P s = t + y; tc = (s - t) - y; t = s;
#pragma unroll
                for (int i = 0; i < SD; i++) { P yi = inc[i] - Xc[i]; P si = X[i] + yi; Xc[i] = (si - X[i]) - yi; X[i] = si; }
            } else {
                t += h;
#pragma unroll
                for (int i = 0; i < SD; i++) X[i] = nx[i];
            }
            Sys::template ActionAfterSuccessfulTimeStep<P>(tid, NT, udt, t, dtP, TD, X, cPAR, sPAR, sPARi, ACC, ACCi);
            if (NE > 0) {
#pragma unroll
                for (int i = 0; i < NE; i++) {
                    const P tol = a.etol[i];
                    if ((EF[i] > tol && math::fabs(EFn[i]) < tol && a.edir[i] <= 0) ||
                        (EF[i] < -tol && math::fabs(EFn[i]) < tol && a.edir[i] >= 0)) {
                        const A& c = kernel_args<A>();   // rare path: fields loaded here
                        if (nev < c.evcap) {
                            const long long base = (long long)nev * (SD + 1);
                            c.evlog[tid + base * NT] = t;
#pragma unroll
                            for (int j = 0; j < SD; j++) c.evlog[tid + (base + 1 + j) * NT] = X[j];
                            c.evlogidx[tid + (long long)nev * NT] = i;
                        }
                        nev++;
                        Sys::template ActionAfterEventDetection<P>(tid, NT, i, udt, t, dtP, TD, X, cPAR, sPAR, sPARi, ACC, ACCi);
                    }
                }
                Sys::template EventFunction<P>(tid, NT, EFn, t, dtP, TD, X, cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
                for (int i = 0; i < NE; i++) EF[i] = EFn[i];
            }
            if (DENSE && a.NDO > 0) {
                // PT-DO:40-58 then PT-DO:5-37
                int upd;
                if (doIdx < a.NDO && doT < t && skipped >= a.doFreq - 1) upd = 1;
                else upd = 0;
                if (doIdx < a.NDO && (endReached == 1 || udt == 1)) upd = 1;
                if (upd) store_dense(kernel_args<A>());
                else skipped++;
            }
            if (endReached == 1 || udt == 1) term = 1;
            dt = (CP)dtP;   // a callback may have changed dT; it reaches Finalization (PT-SOL:398-410)
            nacc++;
        } else {
            nrej++;
        }
    }

    // ------------------------------------------------------------ epilogue
    __device__ __forceinline__ void finish(const A& a, P* sPAR, int* sPARi)
    {
        const int NT = a.NT;
        P dtP = (P)dt;
        Sys::template Finalization<P>(tid, NT, doIdx, t, dtP, TD, X, cPAR, sPAR, sPARi, ACC, ACCi);
#pragma unroll
        for (int i = 0; i < 2; i++) a.TD[tid + i * NT] = TD[i];
#pragma unroll
        for (int i = 0; i < SD; i++) a.X[tid + i * NT] = X[i];
#pragma unroll
        for (int i = 0; i < NCP; i++) a.cPAR[tid + i * NT] = cPAR[i];
#pragma unroll
        for (int i = 0; i < NA; i++) a.ACC[tid + i * NT] = ACC[i];
#pragma unroll
        for (int i = 0; i < NIA; i++) a.ACCi[tid + i * NT] = ACCi[i];
        a.T[tid] = t;
        a.DOidx[tid] = doIdx;
        a.nacc[tid] = nacc;
        a.nrej[tid] = nrej;
        a.status[tid] = status;
        a.nev[tid] = nev;
    }
};

// Words of the lane state that an instantiation never uses (Kahan terms
// without Kahan, accessories, event values and the slice budget of systems or
// forms that have none): the persistent form does not move them between
// lanes, so they hold no registers.
template <class L_t, bool SLICED>
__host__ __device__ constexpr bool lane_word_live(int w)
{
    using Sys = typename L_t::Sys_t;
    const int b = 4 * w;
    auto in = [&](size_t off, size_t bytes) { return (size_t)b >= off && (size_t)b < off + bytes; };
    if (!L_t::kKahan && (in(offsetof(L_t, Xc), sizeof(L_t::Xc)) || in(offsetof(L_t, tc), sizeof(L_t::tc)))) return false;
    if (Sys::NA == 0 && in(offsetof(L_t, ACC), sizeof(L_t::ACC))) return false;
    if (Sys::NIA == 0 && in(offsetof(L_t, ACCi), sizeof(L_t::ACCi))) return false;
    if (Sys::NE == 0 && in(offsetof(L_t, EF), sizeof(L_t::EF))) return false;
    if (!SLICED && in(offsetof(L_t, left), sizeof(L_t::left))) return false;
    return true;
}

// Stage shared parameters into LDS (PT-SOL:67-88) or point at global memory.
template <class Sys, class P, class CP>
__device__ __forceinline__ void stage_shared(const Args<Sys, P, CP>& a, P*& sPAR, int*& sPARi)
{
    constexpr int NSP = Sys::NSP, NISP = Sys::NISP;
    extern __shared__ __align__(16) unsigned char mpgos_dyn_lds[];
    if (a.preferShared && (NSP + NISP) > 0) {
        P* lp = reinterpret_cast<P*>(mpgos_dyn_lds);
        int* li = reinterpret_cast<int*>(lp + NSP);
        for (int i = threadIdx.x; i < NSP; i += blockDim.x) lp[i] = a.sPAR[i];
        for (int i = threadIdx.x; i < NISP; i += blockDim.x) li[i] = a.sPARi[i];
        __syncthreads();
        sPAR = lp;
        sPARi = li;
    } else {
        sPAR = a.sPAR;
        sPARi = a.sPARi;
    }
}

// Reference form: lane = instance, whole integration in one launch.
template <class Sys, class P, bool KAHAN, int ALG, bool SLICED = false>
__device__ __forceinline__ void plain_body(const Args<Sys, P, typename Lane<Sys, P, KAHAN, ALG>::CP>& a)
{
    P* sPAR; int* sPARi;
    stage_shared(a, sPAR, sPARi);
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    const int n = (SLICED && a.listLenDev) ? (int)*a.listLenDev : a.listLen;
    if (i < n) {
        using L_t = Lane<Sys, P, KAHAN, ALG>;
        L_t L;
        const int id = a.list ? a.list[i] : i;
        // With event functions the step keeps more uniform values live (event
        // tolerances, direction tests) than there are SGPRs, and the
        // compiler spills argument fields to VGPR lanes and reloads them as
        // whole 16-register tuples inside the loop; re-reading the argument
        // block at every step instead keeps the loop free of spills (fewer
        // instructions per step, which bounds a lone lane).  Without events
        // the loop does not spill and keeps its hoisted arguments.
        auto step_args = [&]() -> const typename L_t::A& {
            if constexpr (Sys::NE > 0) return kernel_args<typename L_t::A>();
            else return a;
        };
        if constexpr (SLICED) {
            // the loop only steps; finishing or suspending follows it, so
            // the step loop keeps the plain form's registers
            L.begin(a, id, sPAR, sPARi);
            do L.step(step_args(), sPAR, sPARi);
            while (!L.stop());
            L.release(kernel_args<typename L_t::A>(), sPAR, sPARi);
        } else {
            L.load(a, id, sPAR, sPARi);
            while (L.term == 0) L.step(step_args(), sPAR, sPARi);
            L.finish(kernel_args<typename L_t::A>(), sPAR, sPARi);
        }
    }
}

// Persistent form with lane compaction.  Lanes pull instances from a device
// work queue; every `roundSteps` attempted steps the workgroup
//   1. counts its live lanes (wavefront ballots, prefix over waves in LDS),
//   2. reserves as many fresh instances as it has idle lanes (one atomic per
//      workgroup) and starts them in the idle lanes where they are (no state
//      moves while the queue lasts),
//   3. once nothing new arrives, packs the live instances into the lowest
//      lanes through LDS -- only when that empties at least one wave, so the
//      waves above the packed range have nothing to do and sleep at the
//      barrier instead of issuing instructions for idle lanes.
// Every instance's arithmetic is independent of the lane that runs it, so the
// results are bit-identical to the plain form.  The workgroup leaves the loop
// together once the queue is drained and no lane is live (block-uniform test).
template <class Sys, class P, bool KAHAN, int ALG, bool SLICED = false>
__device__ __forceinline__ void persistent_body(const Args<Sys, P, typename Lane<Sys, P, KAHAN, ALG>::CP>& a)
{
    using L_t = Lane<Sys, P, KAHAN, ALG>;
    constexpr int NW = (int)(sizeof(L_t) / 4);       // lane state in 32-bit words
    constexpr int CW = 4;                             // words moved per LDS pass
    constexpr int NCH = (NW + CW - 1) / CW;
    P* sPAR; int* sPARi;
    stage_shared(a, sPAR, sPARi);

    extern __shared__ __align__(16) unsigned char mpgos_dyn_lds[];
    const size_t sharedBytes = ((size_t)Sys::NSP * sizeof(P) + (size_t)Sys::NISP * sizeof(int) + 15) & ~(size_t)15;
    unsigned* ctl = reinterpret_cast<unsigned*>(mpgos_dyn_lds + sharedBytes);   // [0..15] wave counts, [16] base, [17] fresh, [18] drained
    unsigned* xch = ctl + 32;                                                  // CW * blockDim.x words

    const int B = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int nwaves = (B + 63) >> 6;
    const unsigned lane = __lane_id();

    L_t L;
    bool busy = false;
    const unsigned D = (unsigned)a.solo;
    if (tid == 0) {
        ctl[18] = 0u;
        ctl[20] = blockIdx.x < D ? 1u : 0u;   // solo phase: lane 0 alone runs one of the first D entries
    }
    for (;;) {
        // ---- 1. count live lanes
        const unsigned long long live_mask = __ballot(busy);
        if (lane == 0) ctl[wave] = (unsigned)__popcll(live_mask);
        __syncthreads();
        unsigned before = 0, live = 0, occupied = 0;
        for (int w = 0; w < nwaves; w++) {
            const unsigned c = ctl[w];
            if (w < wave) before += c;
            live += c;
            occupied += c > 0 ? 1u : 0u;
        }
        const unsigned long long below = (1ull << lane) - 1ull;
        const unsigned rank = before + (unsigned)__popcll(live_mask & below);             // among live lanes
        const unsigned idleRank = (unsigned)(wave * 64) - before + (unsigned)__popcll(~live_mask & below);   // among idle lanes
        // ---- 2. reserve fresh instances for the idle lanes
        if (tid == 0) {
            unsigned base = 0, fresh = 0;
            if (ctl[20] && live == 0) {
                const unsigned q = atomicAdd(&a.queue[2], 1u);
                if (q < D) { base = q; fresh = 1u; }
                else ctl[20] = 0u;                  // solo entries taken: a full workgroup from now on
            }
            if (!ctl[20]) {
                const unsigned need = (unsigned)B - live;
                if (!ctl[18] && need > 0) {
                    base = D + atomicAdd(a.queue, need);
                    if (base < (unsigned)a.listLen) fresh = min(need, (unsigned)a.listLen - base);
                    if (base + need >= (unsigned)a.listLen) ctl[18] = 1u;
                }
            }
            ctl[16] = base;
            ctl[17] = fresh;
        }
        __syncthreads();
        const unsigned base = ctl[16], fresh = ctl[17];
        if (live == 0 && fresh == 0) break;           // block-uniform: queue drained, nothing live
        // ---- 3. fresh instances start in the idle lanes where they are; with
        // nothing new, live states are packed into lanes [0, live) when that
        // empties a wave (block-uniform decision)
        if (fresh != 0) {
            if (!busy && idleRank < fresh) {
                const int q = (int)(base + idleRank);
                const int id = a.list ? a.list[q] : q;
                if constexpr (SLICED) L.begin(a, id, sPAR, sPARi);
                else L.load(a, id, sPAR, sPARi);
                busy = true;
            }
        } else if (occupied > (live + 63u) / 64u) {
            // the lane state moves in place, CW words per pass (no second copy of it)
            unsigned char* lw = reinterpret_cast<unsigned char*>(&L);
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                constexpr int last = NW - (NCH - 1) * CW;
                const int nw = c == NCH - 1 ? last : CW;
                bool any = false;
#pragma unroll
                for (int j = 0; j < CW; j++) any = any || (j < nw && lane_word_live<L_t, SLICED>(c * CW + j));
                if (!any) continue;                   // compile-time: a chunk of unused words
                uint32_t w[CW];
                __builtin_memcpy(w, lw + c * CW * 4, nw * 4);
                if (busy) {
#pragma unroll
                    for (int j = 0; j < CW; j++)
                        if (j < nw && lane_word_live<L_t, SLICED>(c * CW + j)) xch[j * B + rank] = w[j];
                }
                __syncthreads();
                if ((unsigned)tid < live) {
#pragma unroll
                    for (int j = 0; j < CW; j++)
                        if (j < nw && lane_word_live<L_t, SLICED>(c * CW + j)) w[j] = xch[j * B + tid];
                    __builtin_memcpy(lw + c * CW * 4, w, nw * 4);
                }
                __syncthreads();
            }
            busy = (unsigned)tid < live;
        }
        // ---- 4. integrate
        for (int m = 0; m < a.roundSteps; m++) {
            if (!__any(busy)) break;
            if (busy) {
                L.step(a, sPAR, sPARi);
                if constexpr (SLICED) {
                    if (L.stop()) {
                        L.release(kernel_args<typename L_t::A>(), sPAR, sPARi);
                        busy = false;
                    }
                } else if (L.term) {
                    L.finish(kernel_args<typename L_t::A>(), sPAR, sPARi);
                    busy = false;
                }
            }
        }
    }
}

// LDS bytes of the persistent form beyond the shared parameters.
template <class Sys, class P, bool KAHAN, int ALG>
constexpr size_t persistent_lds_bytes(int block)
{
    return 32 * 4 + (size_t)4 * 4 * block;
}

// Clears the per-instance phases and the per-launch counts of the relaunch
// form (a kernel, so the launch sequence can be captured in a graph).
template <int Unused = 0>
__global__ void reset_relaunch_kernel(int* phase, int n, unsigned* counts, int nc)
{
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i < n) phase[i] = 0;
    if (i < nc) counts[i] = 0u;
}

// Sorted form: after the pilot launch the suspended instances are ordered by
// their estimated remaining attempted steps, largest first, so that every
// wavefront of the following launch holds instances of similar length (lane
// coherence; the order only schedules, results are bit-identical).  A
// counting sort on 256 buckets of 1/8 octave (float exponent and the top three
// mantissa bits of the estimate): histogram, scan, scatter.
__device__ __forceinline__ int sort_bucket(float key)
{
    const float k = key < 1.0f ? 1.0f : (key > 4.0e9f ? 4.0e9f : key);
    const unsigned b = __float_as_uint(k);
    const int v = ((int)(b >> 23) - 127) * 8 + (int)((b >> 20) & 7u);   // 0 .. 255 for keys in [1, 2^32)
    return 255 - v;                                                       // largest estimates first
}

template <int Unused = 0>
__global__ void sort_histogram_kernel(const float* key, const unsigned* n, unsigned* hist)
{
    __shared__ unsigned h[256];
    for (int j = threadIdx.x; j < 256; j += blockDim.x) h[j] = 0u;
    __syncthreads();
    const unsigned len = *n;
    for (unsigned i = threadIdx.x + blockIdx.x * blockDim.x; i < len; i += gridDim.x * blockDim.x)
        atomicAdd(&h[sort_bucket(key[i])], 1u);
    __syncthreads();
    for (int j = threadIdx.x; j < 256; j += blockDim.x)
        if (h[j]) atomicAdd(&hist[j], h[j]);
}

// exclusive scan of the 256 bucket counts into bucket offsets (one workgroup)
template <int Unused = 0>
__global__ void sort_scan_kernel(unsigned* hist)
{
    if (threadIdx.x == 0) {
        unsigned run = 0u;
        for (int j = 0; j < 256; j++) { const unsigned c = hist[j]; hist[j] = run; run += c; }
    }
}

// one workgroup per 256 list entries: the workgroup counts its entries per
// bucket in LDS, reserves each bucket's range with one global atomic, and
// places its entries with LDS atomics (no global atomics per entry)
template <int Unused = 0>
__global__ void sort_scatter_kernel(const float* key, const int* listIn, const unsigned* n, unsigned* offset, int* listOut)
{
    __shared__ unsigned cnt[256], base[256];
    const unsigned len = *n;
    const unsigned i = threadIdx.x + blockIdx.x * blockDim.x;
    if (blockIdx.x * blockDim.x >= len) return;          // workgroup-uniform
    cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int b = i < len ? sort_bucket(key[i]) : -1;
    if (b >= 0) atomicAdd(&cnt[b], 1u);
    __syncthreads();
    const unsigned c = cnt[threadIdx.x];
    if (c) base[threadIdx.x] = atomicAdd(&offset[threadIdx.x], c);
    cnt[threadIdx.x] = 0u;
    __syncthreads();
    if (b >= 0) listOut[base[b] + atomicAdd(&cnt[b], 1u)] = listIn[i];
}

template <int Unused = 0>
__global__ void zero_kernel(unsigned* p, int n)
{
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i < n) p[i] = 0u;
}

// Resets the work-queue head on the stream (a kernel, so the sequence can be
// captured in a graph without a memset node).
template <int Unused = 0>
__global__ void reset_queue_kernel(unsigned int* q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) { q[0] = 0u; q[2] = 0u; }
}

} // namespace pt
} // namespace mpgos
