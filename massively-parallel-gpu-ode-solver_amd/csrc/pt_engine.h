// Host side of the per-thread module: owns the pinned host mirrors, the device
// buffers, the HIP stream and events of one solver object and launches the
// kernels of pt_device.h.  It implements the member functions of the
// reference's ProblemSolver (SourceCodes/SingleSystem_PerThread_Interface.cuh:
// constructor 309-518, SolverOption 610-703, SetHost 705-830, Synchronise*
// 832-965, GetHost 967-1080, Print 1111-1276, Solve 1279-1285, synchronisation
// 1288-1312) with return codes instead of exit().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cstring>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

#include "mpgos.h"
#include "pt_adjoint.h"
#include "pt_device.h"

namespace mpgos {

void set_error(const std::string& msg);   // thread-local message for mpgos_last_error()

#define MPGOS_HIP_TRY(call)                                                              \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::mpgos::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " \
                               + __FILE__ + ":" + std::to_string(__LINE__));             \
            return MPGOS_E_HIP;                                                          \
        }                                                                                \
    } while (0)

// Index bound check with the reference's message (PT-IF:565-581), returning
// MPGOS_E_ARG instead of exiting.
inline int bound_check(const char* fn, const char* what, int v, int lo, int hi)
{
    if (v < lo || v > hi) {
        char buf[256];
        if (lo > hi) std::snprintf(buf, sizeof buf, "%s: %s index %d out of range (acceptable index: none)", fn, what, v);
        else std::snprintf(buf, sizeof buf, "%s: %s index %d out of range (acceptable index: %d-%d)", fn, what, v, lo, hi);
        set_error(buf);
        return MPGOS_E_ARG;
    }
    return MPGOS_OK;
}

namespace pt {

// Kernel entry points.  WPE is the waves-per-EU hint of the register-budget
// variant (1 = no constraint); MODE 0 = plain, 1 = persistent.
template <class Sys, class P, bool KAHAN, int ALG, int MODE, int WPE>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(WPE)))
solve_kernel(Args<Sys, P, typename Lane<Sys, P, KAHAN, ALG>::CP> a)
{
    // MODE + 2: the same forms with the time-sliced lane lifecycle
    if (MODE == MPGOS_KERNEL_PLAIN) plain_body<Sys, P, KAHAN, ALG>(a);
    else if (MODE == MPGOS_KERNEL_PERSISTENT) persistent_body<Sys, P, KAHAN, ALG>(a);
    else if (MODE == MPGOS_KERNEL_PLAIN + 2) plain_body<Sys, P, KAHAN, ALG, true>(a);
    else persistent_body<Sys, P, KAHAN, ALG, true>(a);
}

// Virtual interface the C ABI talks to; one implementation per (system, precision).
class SolverBase {
public:
    virtual ~SolverBase() {}
    virtual int init(int algorithm, int nt, int ndo, int device) = 0;
    virtual int dims(mpgos_pt_dims* d) const = 0;
    virtual int option(int opt, int idx, double v) = 0;
    virtual int set_host(int var, int problem, int serial, int step, double v) = 0;
    virtual int get_host(int var, int problem, int serial, int step, double* v) const = 0;
    virtual void* host_buffer(int var, size_t* nbytes) = 0;
    virtual void* device_buffer(int var, size_t* nbytes) = 0;
    virtual int sync(int var, bool to_device) = 0;
    virtual int solve() = 0;
    virtual int synchronise_device() = 0;
    virtual int insert_sync_point() = 0;
    virtual int synchronise_solver() = 0;
    virtual int print(int var, int tid, const char* dir) = 0;
    virtual int last_solve_ms(double* ms) const = 0;
    virtual void* stream() = 0;
    virtual int solve_slice(long long max_attempted_steps, int* unfinished) = 0;
    virtual int solve_noalloc(void* scratch, size_t bytes) = 0;
    virtual int device_index() const = 0;
    virtual int adjoint_record(int K, int capacity) = 0;
    virtual int adjoint_reverse() = 0;
};

template <class T>
struct Buffer {
    T* h = nullptr;   // pinned host mirror
    T* d = nullptr;   // device
    size_t n = 0;
    int alloc(size_t count)
    {
        n = count;
        if (count == 0) return MPGOS_OK;
        if (hipHostMalloc((void**)&h, count * sizeof(T), hipHostMallocDefault) != hipSuccess) {
            set_error("failed to allocate pinned host memory");
            return MPGOS_E_MEMORY;
        }
        std::memset(h, 0, count * sizeof(T));
        if (hipMalloc((void**)&d, count * sizeof(T)) != hipSuccess) {
            set_error("failed to allocate device memory");
            return MPGOS_E_MEMORY;
        }
        return MPGOS_OK;
    }
    void release()
    {
        if (h) (void)hipHostFree(h);
        if (d) (void)hipFree(d);
        h = nullptr; d = nullptr; n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

template <class Sys, class P, bool KAHAN>
class Solver : public SolverBase {
public:
    using CP = typename std::conditional<KAHAN, double, P>::type;
    static constexpr int SD = Sys::SD, NCP = Sys::NCP, NSP = Sys::NSP, NISP = Sys::NISP;
    static constexpr int NE = Sys::NE, NA = Sys::NA, NIA = Sys::NIA;

    int algorithm = MPGOS_RKCK45;
    int NT = 0, NDO = 0, device = 0;
    hipStream_t strm = nullptr;
    hipEvent_t syncEvent = nullptr, evStart = nullptr, evStop = nullptr;
    int numCU = 0;

    // options (SingleSystem_PerThread_Interface.cuh:340-345, 475-496 defaults)
    int blockSize = 64, nActive = 0, preferShared = 1;
    bool blockSet = false;                 // ThreadsPerBlock given explicitly
    static constexpr int kPersistentBlock = 256;
    int block() const { return (kernelMode != MPGOS_KERNEL_PERSISTENT || blockSet) ? blockSize : kPersistentBlock; }
    double dt0 = 1e-2, dtmax = 1.0e6, dtmin = 1.0e-12, grow = 5.0, shrink = 0.1, doMinDt = 0.0;
    int doFreq = 1;
    std::vector<double> rtol, atol, etol;
    std::vector<int> edir;
    int kernelMode = MPGOS_KERNEL_PLAIN, wavesPerSimd = 0, graphCapture = 0, evcap = 0, variant = 0, roundSteps = 32;
    int queueOrder = MPGOS_QUEUE_INDEX, soloInstances = 0;
    int relaunchSteps = 300, relaunchLaunches = 16;
    unsigned* relaunchCounts = nullptr;   // per-launch suspended counts of the relaunch form
    float* sortKey = nullptr;             // sorted form: estimated remaining steps per suspended instance
    unsigned* sortHist = nullptr;         // sorted form: bucket counts, then offsets
    int relaunchCountsN = 0;
    int* order = nullptr;            // work-queue order of the persistent form (device), when not index order
    std::vector<int> orderHost, rankHost;
    bool orderDirty = true;

    Buffer<P> TD, X, T, cPAR, sPAR, ACC, DOt, DOx, evlog;
    Buffer<int> sPARi, ACCi, DOidx, status, nev, evlogidx;
    Buffer<long long> nacc, nrej;
    Buffer<P> lam, gp;                      // adjoint state and parameter gradient
    double* ckpt = nullptr;                 // adjoint checkpoints (device only)
    double* seg = nullptr;                  // adjoint segment buffer (device only)
    size_t ckptCount = 0, segCount = 0;
    int adjK = 0, adjCap = 0;
    unsigned int* queue = nullptr;

    // graph replay
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::vector<unsigned char> graphArgs;
    long long graphKey = -1;

    ~Solver() override
    {
        if (strm) (void)hipStreamSynchronize(strm);
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        TD.release(); X.release(); T.release(); cPAR.release(); sPAR.release(); ACC.release();
        DOt.release(); DOx.release(); evlog.release();
        sPARi.release(); ACCi.release(); DOidx.release(); status.release(); nev.release(); evlogidx.release();
        nacc.release(); nrej.release(); lam.release(); gp.release();
        if (ckpt) (void)hipFree(ckpt);
        if (seg) (void)hipFree(seg);
        if (queue) (void)hipFree(queue);
        for (void* q : {(void*)saved, (void*)phase, (void*)listA, (void*)listB, (void*)order, (void*)relaunchCounts,
                        (void*)sortKey, (void*)sortHist})
            if (q) (void)hipFree(q);
        if (syncEvent) (void)hipEventDestroy(syncEvent);
        if (evStart) (void)hipEventDestroy(evStart);
        if (evStop) (void)hipEventDestroy(evStop);
        if (strm) (void)hipStreamDestroy(strm);
    }

    int init(int alg, int nt, int ndo, int dev) override
    {
        if (alg != MPGOS_RK4 && alg != MPGOS_RKCK45) { set_error("unknown algorithm"); return MPGOS_E_ARG; }
        if (nt <= 0 || ndo < 0) { set_error("NT must be positive and NDO non-negative"); return MPGOS_E_ARG; }
        algorithm = alg; NT = nt; NDO = ndo; device = dev; nActive = nt;
        MPGOS_HIP_TRY(hipSetDevice(device));
        MPGOS_HIP_TRY(hipStreamCreateWithFlags(&strm, hipStreamNonBlocking));
        MPGOS_HIP_TRY(hipEventCreateWithFlags(&syncEvent, hipEventDisableTiming));
        MPGOS_HIP_TRY(hipEventCreate(&evStart));
        MPGOS_HIP_TRY(hipEventCreate(&evStop));
        hipDeviceProp_t prop;
        MPGOS_HIP_TRY(hipGetDeviceProperties(&prop, device));
        numCU = prop.multiProcessorCount;
        blockSize = prop.warpSize;   // reference default: ThreadsPerBlock = warp size (PT-IF:343)

        const size_t nt_ = (size_t)nt;
        int rc;
        if ((rc = TD.alloc(2 * nt_)) || (rc = X.alloc(SD * nt_)) || (rc = T.alloc(nt_)) ||
            (rc = cPAR.alloc(NCP * nt_)) || (rc = sPAR.alloc(NSP)) || (rc = sPARi.alloc(NISP)) ||
            (rc = ACC.alloc(NA * nt_)) || (rc = ACCi.alloc(NIA * nt_)) || (rc = DOidx.alloc(nt_)) ||
            (rc = DOt.alloc((size_t)ndo * nt_)) || (rc = DOx.alloc((size_t)SD * ndo * nt_)) ||
            (rc = status.alloc(nt_)) || (rc = nev.alloc(nt_)) || (rc = nacc.alloc(nt_)) || (rc = nrej.alloc(nt_)) ||
            (rc = lam.alloc(SD * nt_)) || (rc = gp.alloc(NCP * nt_)))
            return rc;
        MPGOS_HIP_TRY(hipMalloc((void**)&queue, 256));
        rtol.assign(SD, 1e-8);
        atol.assign(SD, 1e-8);
        etol.assign(NE, 1e-6);
        edir.assign(NE, 0);
        // device buffers (inputs and per-instance outputs) start from the zeroed host mirrors
        if ((rc = copy(nacc, true)) || (rc = copy(nrej, true)) || (rc = copy(status, true)) || (rc = copy(nev, true)))
            return rc;
        return sync(MPGOS_PT_ALL, true);
    }

    int dims(mpgos_pt_dims* d) const override
    {
        d->nt = NT; d->sd = SD; d->ncp = NCP; d->nsp = NSP; d->nisp = NISP; d->ne = NE;
        d->na = NA; d->nia = NIA; d->ndo = NDO; d->algorithm = algorithm;
        d->precision = KAHAN ? MPGOS_FLOAT_KAHAN : (std::is_same<P, float>::value ? MPGOS_FLOAT : MPGOS_DOUBLE);
        d->has_vjp = Sys::HAS_VJP ? 1 : 0;
        return MPGOS_OK;
    }

    static int bound(const char* fn, const char* what, int v, int lo, int hi) { return bound_check(fn, what, v, lo, hi); }

    int option(int opt, int idx, double v) override
    {
        switch (opt) {
        case MPGOS_PT_THREADS_PER_BLOCK:
            if ((int)v < 1 || (int)v > kMaxBlock) { set_error("ThreadsPerBlock must be in 1..512"); return MPGOS_E_ARG; }
            blockSize = (int)v; blockSet = true; return MPGOS_OK;
        case MPGOS_PT_INITIAL_TIME_STEP: dt0 = v; return MPGOS_OK;
        case MPGOS_PT_ACTIVE_NUMBER_OF_THREADS:
            if ((int)v < 0 || (int)v > NT) { set_error("ActiveNumberOfThreads must be in 0..NT"); return MPGOS_E_ARG; }
            nActive = (int)v; orderDirty = true; return MPGOS_OK;
        case MPGOS_PT_MAXIMUM_TIME_STEP: dtmax = v; return MPGOS_OK;
        case MPGOS_PT_MINIMUM_TIME_STEP: dtmin = v; return MPGOS_OK;
        case MPGOS_PT_TIME_STEP_GROW_LIMIT: grow = v; return MPGOS_OK;
        case MPGOS_PT_TIME_STEP_SHRINK_LIMIT: shrink = v; return MPGOS_OK;
        case MPGOS_PT_RELATIVE_TOLERANCE:
            if (int rc = bound("SolverOption", "RelativeTolerance", idx, 0, SD - 1)) return rc;
            rtol[idx] = v; return MPGOS_OK;
        case MPGOS_PT_ABSOLUTE_TOLERANCE:
            if (int rc = bound("SolverOption", "AbsoluteTolerance", idx, 0, SD - 1)) return rc;
            atol[idx] = v; return MPGOS_OK;
        case MPGOS_PT_EVENT_TOLERANCE:
            if (int rc = bound("SolverOption", "EventTolerance", idx, 0, NE - 1)) return rc;
            etol[idx] = v; return MPGOS_OK;
        case MPGOS_PT_EVENT_DIRECTION:
            if (int rc = bound("SolverOption", "EventDirection", idx, 0, NE - 1)) return rc;
            edir[idx] = (int)v; return MPGOS_OK;
        case MPGOS_PT_DENSE_OUTPUT_MINIMUM_TIME_STEP: doMinDt = v; return MPGOS_OK;
        case MPGOS_PT_DENSE_OUTPUT_SAVE_FREQUENCY: doFreq = (int)v; return MPGOS_OK;
        case MPGOS_PT_PREFER_SHARED_MEMORY: preferShared = (int)v != 0; return MPGOS_OK;
        case MPGOS_PT_KERNEL_MODE:
            if ((int)v != MPGOS_KERNEL_PLAIN && (int)v != MPGOS_KERNEL_PERSISTENT && (int)v != MPGOS_KERNEL_RELAUNCH &&
                (int)v != MPGOS_KERNEL_SORTED) {
                set_error("KernelMode must be 0, 1, 2 or 3");
                return MPGOS_E_ARG;
            }
            kernelMode = (int)v; return MPGOS_OK;
        case MPGOS_PT_PERSISTENT_WAVES_PER_SIMD:
            if ((int)v < 0 || (int)v > 8) { set_error("PersistentWavesPerSimd must be in 0..8"); return MPGOS_E_ARG; }
            wavesPerSimd = (int)v; return MPGOS_OK;
        case MPGOS_PT_GRAPH_CAPTURE: graphCapture = (int)v != 0; return MPGOS_OK;
        case MPGOS_PT_PERSISTENT_ROUND_STEPS:
            if ((int)v < 1) { set_error("PersistentRoundSteps must be >= 1"); return MPGOS_E_ARG; }
            roundSteps = (int)v; return MPGOS_OK;
        case MPGOS_PT_QUEUE_ORDER:
            if ((int)v != MPGOS_QUEUE_INDEX && (int)v != MPGOS_QUEUE_LONGEST_FIRST) { set_error("QueueOrder must be 0 or 1"); return MPGOS_E_ARG; }
            queueOrder = (int)v; orderDirty = true; return MPGOS_OK;
        case MPGOS_PT_RELAUNCH_STEPS:
            if ((int)v < 1) { set_error("RelaunchSteps must be >= 1"); return MPGOS_E_ARG; }
            relaunchSteps = (int)v; return MPGOS_OK;
        case MPGOS_PT_RELAUNCH_LAUNCHES:
            if ((int)v < 0 || (int)v > 4096) { set_error("RelaunchLaunches must be in 0..4096"); return MPGOS_E_ARG; }
            relaunchLaunches = (int)v; return MPGOS_OK;
        case MPGOS_PT_PERSISTENT_SOLO_INSTANCES:
            if ((int)v < 0) { set_error("PersistentSoloInstances must be >= 0"); return MPGOS_E_ARG; }
            soloInstances = (int)v; return MPGOS_OK;
        case MPGOS_PT_LAUNCH_VARIANT:
            if ((int)v < 0 || (int)v > 3) { set_error("LaunchVariant must be in 0..3"); return MPGOS_E_ARG; }
            variant = (int)v; return MPGOS_OK;
        case MPGOS_PT_EVENT_LOG_CAPACITY: {
            if ((int)v < 0) { set_error("EventLogCapacity must be >= 0"); return MPGOS_E_ARG; }
            MPGOS_HIP_TRY(hipStreamSynchronize(strm));
            evlog.release(); evlogidx.release();
            evcap = (int)v;
            int rc;
            if ((rc = evlog.alloc((size_t)(SD + 1) * evcap * NT)) || (rc = evlogidx.alloc((size_t)evcap * NT))) return rc;
            // start from the zeroed host mirrors so unused slots read as zero
            if ((rc = copy(evlog, true)) || (rc = copy(evlogidx, true))) return rc;
            return MPGOS_OK;
        }
        default:
            set_error("SolverOption: unknown or not applicable option " + std::to_string(opt));
            return MPGOS_E_ARG;
        }
    }

    // (variable, problem, serial, step) -> pointer into the host mirror; exact
    // index rules and bound checks of SetHost/GetHost (PT-IF:705-830, 967-1080).
    template <class F>
    int access(const char* fn, int var, int problem, int serial, int step, F&& f) const
    {
        int rc;
        switch (var) {
        case MPGOS_PT_TIME_DOMAIN: case MPGOS_PT_ACTUAL_STATE: case MPGOS_PT_CONTROL_PARAMETERS:
        case MPGOS_PT_ACCESSORIES: case MPGOS_PT_INTEGER_ACCESSORIES: case MPGOS_PT_DENSE_TIME: {
            if ((rc = bound(fn, "ProblemNumber", problem, 0, NT - 1))) return rc;
            const size_t idx = (size_t)problem + (size_t)serial * NT;
            switch (var) {
            case MPGOS_PT_TIME_DOMAIN: if ((rc = bound(fn, "TimeDomain", serial, 0, 1))) return rc; f(&TD.h[idx]); break;
            case MPGOS_PT_ACTUAL_STATE: if ((rc = bound(fn, "ActualState", serial, 0, SD - 1))) return rc; f(&X.h[idx]); break;
            case MPGOS_PT_CONTROL_PARAMETERS: if ((rc = bound(fn, "ControlParameters", serial, 0, NCP - 1))) return rc; f(&cPAR.h[idx]); break;
            case MPGOS_PT_ACCESSORIES: if ((rc = bound(fn, "Accessories", serial, 0, NA - 1))) return rc; f(&ACC.h[idx]); break;
            case MPGOS_PT_INTEGER_ACCESSORIES: if ((rc = bound(fn, "IntegerAccessories", serial, 0, NIA - 1))) return rc; f(&ACCi.h[idx]); break;
            default: if ((rc = bound(fn, "DenseTime", serial, 0, NDO - 1))) return rc; f(&DOt.h[idx]); break;
            }
            return MPGOS_OK;
        }
        case MPGOS_PT_ACTUAL_TIME: case MPGOS_PT_DENSE_INDEX:
        case MPGOS_PT_ACCEPTED_STEPS: case MPGOS_PT_REJECTED_STEPS: case MPGOS_PT_STATUS: case MPGOS_PT_EVENT_COUNT:
            if ((rc = bound(fn, "ProblemNumber", problem, 0, NT - 1))) return rc;
            switch (var) {
            case MPGOS_PT_ACTUAL_TIME: f(&T.h[problem]); break;
            case MPGOS_PT_DENSE_INDEX: f(&DOidx.h[problem]); break;
            case MPGOS_PT_ACCEPTED_STEPS: f(&nacc.h[problem]); break;
            case MPGOS_PT_REJECTED_STEPS: f(&nrej.h[problem]); break;
            case MPGOS_PT_STATUS: f(&status.h[problem]); break;
            default: f(&nev.h[problem]); break;
            }
            return MPGOS_OK;
        case MPGOS_PT_SHARED_PARAMETERS:
            if ((rc = bound(fn, "SharedParameters", serial, 0, NSP - 1))) return rc;
            f(&sPAR.h[serial]); return MPGOS_OK;
        case MPGOS_PT_INTEGER_SHARED_PARAMETERS:
            if ((rc = bound(fn, "IntegerSharedParameters", serial, 0, NISP - 1))) return rc;
            f(&sPARi.h[serial]); return MPGOS_OK;
        case MPGOS_PT_DENSE_STATE: {
            if ((rc = bound(fn, "ProblemNumber", problem, 0, NT - 1))) return rc;
            if ((rc = bound(fn, "DenseState/ComponentNumber", serial, 0, SD - 1))) return rc;
            if ((rc = bound(fn, "DenseState/TimeStepNumber", step, 0, NDO - 1))) return rc;
            f(&DOx.h[(size_t)problem + (size_t)serial * NT + (size_t)step * NT * SD]);
            return MPGOS_OK;
        }
        case MPGOS_PT_EVENT_LOG: {
            if ((rc = bound(fn, "ProblemNumber", problem, 0, NT - 1))) return rc;
            if ((rc = bound(fn, "EventLog/Column", serial, 0, SD))) return rc;
            if ((rc = bound(fn, "EventLog/Hit", step, 0, evcap - 1))) return rc;
            f(&evlog.h[(size_t)problem + ((size_t)step * (SD + 1) + serial) * NT]);
            return MPGOS_OK;
        }
        case MPGOS_PT_EVENT_LOG_INDEX: {
            if ((rc = bound(fn, "ProblemNumber", problem, 0, NT - 1))) return rc;
            if ((rc = bound(fn, "EventLogIndex/Hit", step, 0, evcap - 1))) return rc;
            f(&evlogidx.h[(size_t)problem + (size_t)step * NT]);
            return MPGOS_OK;
        }
        case MPGOS_PT_ADJOINT_STATE:
            if ((rc = bound(fn, "ProblemNumber", problem, 0, NT - 1)) || (rc = bound(fn, "AdjointState", serial, 0, SD - 1))) return rc;
            f(&lam.h[(size_t)problem + (size_t)serial * NT]);
            return MPGOS_OK;
        case MPGOS_PT_PARAMETER_GRADIENT:
            if ((rc = bound(fn, "ProblemNumber", problem, 0, NT - 1)) || (rc = bound(fn, "ParameterGradient", serial, 0, NCP - 1))) return rc;
            f(&gp.h[(size_t)problem + (size_t)serial * NT]);
            return MPGOS_OK;
        default:
            set_error(std::string(fn) + ": variable not applicable with this argument configuration");
            return MPGOS_E_ARG;
        }
    }

    int set_host(int var, int problem, int serial, int step, double v) override
    {
        if (var == MPGOS_PT_DENSE_INDEX) {
            // PT-IF:767-770: the value itself is bound-checked
            if (int rc = bound("SetHost", "DenseIndex-Value", (int)v, 0, NDO - 1)) return rc;
        }
        if (var >= MPGOS_PT_ACCEPTED_STEPS && var != MPGOS_PT_ADJOINT_STATE) { set_error("SetHost: output-only variable"); return MPGOS_E_ARG; }
        return access("SetHost", var, problem, serial, step, [&](auto* p) {
            using T_ = typename std::remove_pointer<decltype(p)>::type;
            *p = std::is_integral<T_>::value ? (T_)(int)v : (T_)v;
        });
    }

    int get_host(int var, int problem, int serial, int step, double* v) const override
    {
        return access("GetHost", var, problem, serial, step, [&](auto* p) { *v = (double)*p; });
    }

    template <class T_> static void* ret(Buffer<T_>& b, bool host, size_t* nbytes)
    {
        if (nbytes) *nbytes = b.bytes();
        return host ? (void*)b.h : (void*)b.d;
    }

    void* buffer(int var, bool host, size_t* nbytes)
    {
        switch (var) {
        case MPGOS_PT_TIME_DOMAIN: return ret(TD, host, nbytes);
        case MPGOS_PT_ACTUAL_STATE: return ret(X, host, nbytes);
        case MPGOS_PT_ACTUAL_TIME: return ret(T, host, nbytes);
        case MPGOS_PT_CONTROL_PARAMETERS: return ret(cPAR, host, nbytes);
        case MPGOS_PT_SHARED_PARAMETERS: return ret(sPAR, host, nbytes);
        case MPGOS_PT_INTEGER_SHARED_PARAMETERS: return ret(sPARi, host, nbytes);
        case MPGOS_PT_ACCESSORIES: return ret(ACC, host, nbytes);
        case MPGOS_PT_INTEGER_ACCESSORIES: return ret(ACCi, host, nbytes);
        case MPGOS_PT_DENSE_INDEX: return ret(DOidx, host, nbytes);
        case MPGOS_PT_DENSE_TIME: return ret(DOt, host, nbytes);
        case MPGOS_PT_DENSE_STATE: return ret(DOx, host, nbytes);
        case MPGOS_PT_ACCEPTED_STEPS: return ret(nacc, host, nbytes);
        case MPGOS_PT_REJECTED_STEPS: return ret(nrej, host, nbytes);
        case MPGOS_PT_STATUS: return ret(status, host, nbytes);
        case MPGOS_PT_EVENT_COUNT: return ret(nev, host, nbytes);
        case MPGOS_PT_EVENT_LOG: return ret(evlog, host, nbytes);
        case MPGOS_PT_EVENT_LOG_INDEX: return ret(evlogidx, host, nbytes);
        case MPGOS_PT_ADJOINT_STATE: return ret(lam, host, nbytes);
        case MPGOS_PT_PARAMETER_GRADIENT: return ret(gp, host, nbytes);
        default:
            set_error("buffer: variable has no single buffer");
            if (nbytes) *nbytes = 0;
            return nullptr;
        }
    }
    void* host_buffer(int var, size_t* nbytes) override { return buffer(var, true, nbytes); }
    void* device_buffer(int var, size_t* nbytes) override { return buffer(var, false, nbytes); }

    template <class T_> int copy(Buffer<T_>& b, bool to_device)
    {
        if (b.n == 0) return MPGOS_OK;
        if (to_device) MPGOS_HIP_TRY(hipMemcpyAsync(b.d, b.h, b.bytes(), hipMemcpyHostToDevice, strm));
        else MPGOS_HIP_TRY(hipMemcpyAsync(b.h, b.d, b.bytes(), hipMemcpyDeviceToHost, strm));
        return MPGOS_OK;
    }

    // PT-IF:832-965
    int sync(int var, bool to_device) override
    {
        MPGOS_HIP_TRY(hipSetDevice(device));
        int rc = MPGOS_OK;
        if (to_device && (var == MPGOS_PT_TIME_DOMAIN || var == MPGOS_PT_ALL)) orderDirty = true;
        switch (var) {
        case MPGOS_PT_TIME_DOMAIN: return copy(TD, to_device);
        case MPGOS_PT_ACTUAL_STATE: return copy(X, to_device);
        case MPGOS_PT_ACTUAL_TIME: return copy(T, to_device);
        case MPGOS_PT_CONTROL_PARAMETERS: return copy(cPAR, to_device);
        case MPGOS_PT_SHARED_PARAMETERS: return copy(sPAR, to_device);
        case MPGOS_PT_INTEGER_SHARED_PARAMETERS: return copy(sPARi, to_device);
        case MPGOS_PT_ACCESSORIES: return copy(ACC, to_device);
        case MPGOS_PT_INTEGER_ACCESSORIES: return copy(ACCi, to_device);
        case MPGOS_PT_DENSE_OUTPUT:
            if ((rc = copy(DOidx, to_device)) || (rc = copy(DOt, to_device)) || (rc = copy(DOx, to_device))) return rc;
            return MPGOS_OK;
        case MPGOS_PT_ALL:
            if ((rc = copy(TD, to_device)) || (rc = copy(X, to_device)) || (rc = copy(T, to_device)) ||
                (rc = copy(cPAR, to_device)) || (rc = copy(sPAR, to_device)) || (rc = copy(ACC, to_device)) ||
                (rc = copy(sPARi, to_device)) || (rc = copy(ACCi, to_device)) || (rc = copy(DOidx, to_device)) ||
                (rc = copy(DOt, to_device)) || (rc = copy(DOx, to_device)))
                return rc;
            if (!to_device) {
                if ((rc = copy(nacc, false)) || (rc = copy(nrej, false)) || (rc = copy(status, false)) || (rc = copy(nev, false))) return rc;
            }
            return MPGOS_OK;
        case MPGOS_PT_ADJOINT_STATE: return copy(lam, to_device);
        case MPGOS_PT_PARAMETER_GRADIENT:
            if (to_device) { set_error("SynchroniseFromHostToDevice: output-only variable"); return MPGOS_E_ARG; }
            return copy(gp, false);
        case MPGOS_PT_ACCEPTED_STEPS: case MPGOS_PT_REJECTED_STEPS: case MPGOS_PT_STATUS: case MPGOS_PT_EVENT_COUNT:
        case MPGOS_PT_EVENT_LOG: case MPGOS_PT_EVENT_LOG_INDEX:
            if (to_device) { set_error("SynchroniseFromHostToDevice: output-only variable"); return MPGOS_E_ARG; }
            switch (var) {
            case MPGOS_PT_ACCEPTED_STEPS: return copy(nacc, false);
            case MPGOS_PT_REJECTED_STEPS: return copy(nrej, false);
            case MPGOS_PT_STATUS: return copy(status, false);
            case MPGOS_PT_EVENT_COUNT: return copy(nev, false);
            case MPGOS_PT_EVENT_LOG: return copy(evlog, false);
            default: return copy(evlogidx, false);
            }
        default:
            set_error("Synchronise: variable not applicable");
            return MPGOS_E_ARG;
        }
    }

    Args<Sys, P, CP> make_args() const
    {
        Args<Sys, P, CP> a;
        std::memset(&a, 0, sizeof a);
        a.TD = TD.d; a.X = X.d; a.T = T.d; a.cPAR = cPAR.d; a.sPAR = sPAR.d; a.sPARi = sPARi.d;
        a.ACC = ACC.d; a.ACCi = ACCi.d; a.DOidx = DOidx.d; a.DOt = DOt.d; a.DOx = DOx.d;
        a.nacc = nacc.d; a.nrej = nrej.d; a.status = status.d; a.nev = nev.d;
        a.evlog = evlog.d; a.evlogidx = evlogidx.d; a.queue = queue;
        a.NT = NT; a.nActive = nActive; a.NDO = NDO; a.evcap = evcap; a.preferShared = preferShared;
        a.list = nullptr; a.listLen = nActive; a.slice = 0; a.listLenDev = nullptr; a.outCount = nullptr; a.sortKey = nullptr;
        a.doFreq = doFreq;
        a.roundSteps = roundSteps;
        a.solo = 0;
        a.dt0 = (CP)(P)dt0;   // options are stored as Precision in the reference (PT-IF:619-650)
        a.dtmax = (CP)(P)dtmax; a.dtmin = (CP)(P)dtmin; a.grow = (CP)(P)grow; a.shrink = (CP)(P)shrink;
        if (KAHAN) { a.dt0 = (CP)dt0; a.dtmax = (CP)dtmax; a.dtmin = (CP)dtmin; a.grow = (CP)grow; a.shrink = (CP)shrink; }
        a.doMinDt = (P)doMinDt;
        for (int i = 0; i < SD; i++) {
            a.rtol[i] = KAHAN ? (CP)rtol[i] : (CP)(P)rtol[i];
            a.atol[i] = KAHAN ? (CP)atol[i] : (CP)(P)atol[i];
        }
        for (int i = 0; i < NE; i++) { a.etol[i] = (P)etol[i]; a.edir[i] = edir[i]; }
        return a;
    }

    using KernelFn = void (*)(Args<Sys, P, CP>);

    template <int ALG, int MODE>
    static KernelFn pick_variant(int v)
    {
        switch (v) {
        case 1: return solve_kernel<Sys, P, KAHAN, ALG, MODE, 2>;
        case 2: return solve_kernel<Sys, P, KAHAN, ALG, MODE, 4>;
        case 3: return solve_kernel<Sys, P, KAHAN, ALG, MODE, 8>;
        default: return solve_kernel<Sys, P, KAHAN, ALG, MODE, 1>;
        }
    }

    KernelFn kernel(bool sliced = false) const
    {
        const bool plain = kernelMode != MPGOS_KERNEL_PERSISTENT;
        sliced = sliced || kernelMode == MPGOS_KERNEL_RELAUNCH || kernelMode == MPGOS_KERNEL_SORTED;   // sliced lanes
        if (algorithm == MPGOS_RK4) {
            if (sliced) return plain ? pick_variant<MPGOS_RK4, MPGOS_KERNEL_PLAIN + 2>(variant)
                                     : pick_variant<MPGOS_RK4, MPGOS_KERNEL_PERSISTENT + 2>(variant);
            return plain ? pick_variant<MPGOS_RK4, MPGOS_KERNEL_PLAIN>(variant)
                         : pick_variant<MPGOS_RK4, MPGOS_KERNEL_PERSISTENT>(variant);
        }
        if (sliced) return plain ? pick_variant<MPGOS_RKCK45, MPGOS_KERNEL_PLAIN + 2>(variant)
                                 : pick_variant<MPGOS_RKCK45, MPGOS_KERNEL_PERSISTENT + 2>(variant);
        return plain ? pick_variant<MPGOS_RKCK45, MPGOS_KERNEL_PLAIN>(variant)
                     : pick_variant<MPGOS_RKCK45, MPGOS_KERNEL_PERSISTENT>(variant);
    }

    size_t dyn_lds() const
    {
        const size_t sharedBytes = preferShared ? (size_t)NSP * sizeof(P) + (size_t)NISP * sizeof(int) : 0;
        if (kernelMode != MPGOS_KERNEL_PERSISTENT) return sharedBytes;
        // persistent form: shared parameters (16-byte aligned), control words, exchange buffer
        return ((((size_t)NSP * sizeof(P) + (size_t)NISP * sizeof(int)) + 15) & ~(size_t)15) +
               persistent_lds_bytes<Sys, P, KAHAN, MPGOS_RK4>(block());
    }

    int launch_geometry(KernelFn fn, int& grid) const
    {
        const int B = block();
        if (kernelMode == MPGOS_KERNEL_PLAIN) {
            grid = NT / B + (NT % B == 0 ? 0 : 1);   // PT-IF:344
            return MPGOS_OK;
        }
        if (kernelMode == MPGOS_KERNEL_RELAUNCH || kernelMode == MPGOS_KERNEL_SORTED) {   // lists are at most nActive long
            grid = nActive / B + (nActive % B == 0 ? 0 : 1);
            return MPGOS_OK;
        }
        int perCU = 0;
        MPGOS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void*)fn, B, dyn_lds()));
        if (perCU < 1) perCU = 1;
        const int wavesPerBlock = (B + 63) / 64;
        if (wavesPerSimd > 0) {
            const int want = (4 * wavesPerSimd + wavesPerBlock - 1) / wavesPerBlock;
            if (want < perCU) perCU = want;
        }
        grid = perCU * numCU;
        const int needed = nActive / B + (nActive % B == 0 ? 0 : 1);
        if (grid > needed) grid = needed;
        if (grid < 1) grid = 1;
        return MPGOS_OK;
    }

    // Device memory of the relaunch form: the lane records of suspended
    // instances, their phases, the two instance lists of consecutive launches
    // and the per-launch suspended counts.
    struct RelaunchBuffers {
        uint32_t* saved; int* phase; int* listA; int* listB; unsigned* counts;
    };
    // Bytes of RelaunchBuffers carved from one block (records first, 16-byte
    // multiples; the rest 4-byte words).
    size_t relaunch_bytes() const
    {
        return lane_words() * 4 * (size_t)NT + 3 * 4 * (size_t)NT + 4 * (size_t)(relaunchLaunches + 1);
    }
    RelaunchBuffers carve_relaunch(void* base) const
    {
        RelaunchBuffers r;
        unsigned char* p = static_cast<unsigned char*>(base);
        r.saved = reinterpret_cast<uint32_t*>(p); p += lane_words() * 4 * (size_t)NT;
        r.phase = reinterpret_cast<int*>(p); p += 4 * (size_t)NT;
        r.listA = reinterpret_cast<int*>(p); p += 4 * (size_t)NT;
        r.listB = reinterpret_cast<int*>(p); p += 4 * (size_t)NT;
        r.counts = reinterpret_cast<unsigned*>(p);
        return r;
    }

    // Relaunch form: relaunchLaunches launches of relaunchSteps attempted steps
    // per instance, each working on the previous launch's suspended instances
    // packed into the lowest lanes (list length read on the device), then one
    // launch without a budget; queued together, no host round trip.
    int enqueue_relaunch(const Args<Sys, P, CP>& a, KernelFn fn, int grid, int B, size_t lds, const RelaunchBuffers& r)
    {
        const int S = relaunchLaunches;
        hipLaunchKernelGGL(reset_relaunch_kernel<0>, dim3((unsigned)((NT > S + 1 ? NT : S + 1) + 255) / 256), dim3(256), 0, strm,
                           r.phase, NT, r.counts, S + 1);
        for (int s = 0; s <= S; s++) {
            Args<Sys, P, CP> b = a;
            b.saved = r.saved; b.phase = r.phase;
            b.list = s == 0 ? a.list : (s % 2 == 1 ? r.listA : r.listB);
            b.listOut = s == S ? nullptr : (s % 2 == 0 ? r.listA : r.listB);
            b.listLen = nActive;
            b.listLenDev = s == 0 ? nullptr : r.counts + (s - 1);
            b.outCount = r.counts + s;
            b.slice = s < S ? relaunchSteps : 0x7fffffff;
            hipLaunchKernelGGL(fn, dim3(grid), dim3(B), lds, strm, b);
        }
        MPGOS_HIP_TRY(hipGetLastError());
        return MPGOS_OK;
    }

    // Sorted form: a pilot launch of relaunchSteps attempted steps per
    // instance that also writes each suspended instance's estimated remaining
    // steps, a counting sort of the suspended list by that estimate (largest
    // first), and one launch that resumes the sorted instances and runs them
    // to their ends; queued together, no host round trip.
    int enqueue_sorted(const Args<Sys, P, CP>& a, KernelFn fn, int grid, int B, size_t lds, const RelaunchBuffers& r)
    {
        hipLaunchKernelGGL(reset_relaunch_kernel<0>, dim3((unsigned)(NT + 255) / 256), dim3(256), 0, strm, r.phase, NT, r.counts, 1);
        hipLaunchKernelGGL(zero_kernel<0>, dim3(1), dim3(256), 0, strm, sortHist, 256);
        Args<Sys, P, CP> b = a;
        b.saved = r.saved; b.phase = r.phase;
        b.listOut = r.listA; b.listLen = nActive; b.listLenDev = nullptr; b.outCount = r.counts;
        b.slice = relaunchSteps; b.sortKey = sortKey;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(B), lds, strm, b);
        const unsigned sg = (unsigned)std::min(1024, (nActive + 255) / 256);
        hipLaunchKernelGGL(sort_histogram_kernel<0>, dim3(sg), dim3(256), 0, strm, sortKey, r.counts, sortHist);
        hipLaunchKernelGGL(sort_scan_kernel<0>, dim3(1), dim3(64), 0, strm, sortHist);
        hipLaunchKernelGGL(sort_scatter_kernel<0>, dim3((unsigned)(nActive + 255) / 256), dim3(256), 0, strm, sortKey, r.listA,
                           r.counts, sortHist, r.listB);
        b.list = r.listB; b.listOut = nullptr; b.listLenDev = r.counts; b.outCount = nullptr;
        b.slice = 0x7fffffff; b.sortKey = nullptr;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(B), lds, strm, b);
        MPGOS_HIP_TRY(hipGetLastError());
        return MPGOS_OK;
    }

    // The relaunch form over caller-provided device memory (`bytes` bytes at
    // `scratch`): nothing is allocated, whatever the KernelMode option says;
    // ThreadsPerBlock, LaunchVariant, RelaunchSteps and RelaunchLaunches apply.
    // GraphCapture is not used (a graph instantiation allocates).
    int solve_noalloc(void* scratch, size_t bytes) override
    {
        MPGOS_HIP_TRY(hipSetDevice(device));
        if (!scratch) { set_error("mpgos_pt_solve_noalloc: no scratch memory on this device"); return MPGOS_E_STATE; }
        if (relaunch_bytes() > bytes) {
            set_error("mpgos_pt_solve_noalloc: the relaunch form needs " + std::to_string(relaunch_bytes()) +
                      " bytes of scratch for NT = " + std::to_string(NT) + " (" + std::to_string(lane_words() * 4 + 12) +
                      " per instance), more than the " + std::to_string(bytes) + " available");
            return MPGOS_E_ARG;
        }
        if (nActive == 0) return MPGOS_OK;
        const RelaunchBuffers r = carve_relaunch(scratch);
        const Args<Sys, P, CP> a = make_args();
        const KernelFn fn = algorithm == MPGOS_RK4 ? pick_variant<MPGOS_RK4, MPGOS_KERNEL_PLAIN + 2>(variant)
                                                   : pick_variant<MPGOS_RKCK45, MPGOS_KERNEL_PLAIN + 2>(variant);
        const int B = blockSize;
        const int grid = nActive / B + (nActive % B == 0 ? 0 : 1);
        const size_t lds = preferShared ? (size_t)NSP * sizeof(P) + (size_t)NISP * sizeof(int) : 0;
        MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
        if (int rc = enqueue_relaunch(a, fn, grid, B, lds, r)) return rc;
        MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
        return MPGOS_OK;
    }

    int ensure_slice_buffers()
    {
        if (!saved) {
            if (hipMalloc((void**)&saved, lane_words() * 4 * (size_t)NT) != hipSuccess || hipMalloc((void**)&phase, 4 * (size_t)NT) != hipSuccess ||
                hipMalloc((void**)&listA, 4 * (size_t)NT) != hipSuccess || hipMalloc((void**)&listB, 4 * (size_t)NT) != hipSuccess) {
                set_error("time-sliced launches: allocation failed");
                return MPGOS_E_MEMORY;
            }
        }
        if (relaunchCountsN < relaunchLaunches + 1) {
            MPGOS_HIP_TRY(hipStreamSynchronize(strm));
            if (relaunchCounts) (void)hipFree(relaunchCounts);
            relaunchCounts = nullptr; relaunchCountsN = 0;
            if (hipMalloc((void**)&relaunchCounts, sizeof(unsigned) * (size_t)(relaunchLaunches + 1)) != hipSuccess) {
                set_error("relaunch form: allocation failed");
                return MPGOS_E_MEMORY;
            }
            relaunchCountsN = relaunchLaunches + 1;
        }
        return MPGOS_OK;
    }

    int enqueue(const Args<Sys, P, CP>& a, KernelFn fn, int grid)
    {
        if (kernelMode == MPGOS_KERNEL_RELAUNCH)
            return enqueue_relaunch(a, fn, grid, block(), dyn_lds(), RelaunchBuffers{saved, phase, listA, listB, relaunchCounts});
        if (kernelMode == MPGOS_KERNEL_SORTED)
            return enqueue_sorted(a, fn, grid, block(), dyn_lds(), RelaunchBuffers{saved, phase, listA, listB, relaunchCounts});
        if (kernelMode == MPGOS_KERNEL_PERSISTENT)
            hipLaunchKernelGGL(reset_queue_kernel<0>, dim3(1), dim3(64), 0, strm, queue);
        hipLaunchKernelGGL(fn, dim3(grid), dim3(block()), dyn_lds(), strm, a);
        MPGOS_HIP_TRY(hipGetLastError());
        return MPGOS_OK;
    }

    // list entries run one per workgroup: at most one per workgroup of the grid
    int solo_entries(int grid, int listLen) const
    {
        if (kernelMode != MPGOS_KERNEL_PERSISTENT) return 0;
        int d = soloInstances < grid ? soloInstances : grid;
        return d < listLen ? d : listLen;
    }

    // Work-queue order of the persistent form: the active instances sorted by
    // decreasing integration interval (ties by index), from the time domains
    // last uploaded.  Recomputed after the time domains, the active count or
    // the option change; the order only schedules work.
    bool ordered() const { return kernelMode == MPGOS_KERNEL_PERSISTENT && queueOrder == MPGOS_QUEUE_LONGEST_FIRST; }
    int ensure_order()
    {
        if (!orderDirty && order) return MPGOS_OK;
        if (!order) {
            if (hipMalloc((void**)&order, sizeof(int) * (size_t)NT) != hipSuccess) { set_error("QueueOrder: allocation failed"); return MPGOS_E_MEMORY; }
        }
        orderHost.resize((size_t)nActive);
        for (int i = 0; i < nActive; i++) orderHost[(size_t)i] = i;
        const P* td = TD.h;
        std::stable_sort(orderHost.begin(), orderHost.end(), [&](int x, int y) {
            return (double)(td[x + NT] - td[x]) > (double)(td[y + NT] - td[y]);
        });
        rankHost.assign((size_t)NT, 0);
        for (int r = 0; r < nActive; r++) rankHost[(size_t)orderHost[(size_t)r]] = r;
        MPGOS_HIP_TRY(hipStreamSynchronize(strm));      // a running Solve may still read the old order
        if (nActive > 0) MPGOS_HIP_TRY(hipMemcpy(order, orderHost.data(), sizeof(int) * (size_t)nActive, hipMemcpyHostToDevice));
        orderDirty = false;
        return MPGOS_OK;
    }

    // PT-IF:1279-1285
    int solve() override
    {
        MPGOS_HIP_TRY(hipSetDevice(device));
        if (nActive == 0) return MPGOS_OK;
        if (ordered()) {
            if (int rc = ensure_order()) return rc;
        }
        if (kernelMode == MPGOS_KERNEL_RELAUNCH || kernelMode == MPGOS_KERNEL_SORTED) {
            if (int rc = ensure_slice_buffers()) return rc;
        }
        if (kernelMode == MPGOS_KERNEL_SORTED && !sortKey) {
            if (hipMalloc((void**)&sortKey, sizeof(float) * (size_t)NT) != hipSuccess ||
                hipMalloc((void**)&sortHist, sizeof(unsigned) * 256) != hipSuccess) {
                set_error("sorted form: allocation failed");
                return MPGOS_E_MEMORY;
            }
        }
        Args<Sys, P, CP> a = make_args();
        if (ordered()) a.list = order;
        KernelFn fn = kernel();
        int grid = 0;
        if (int rc = launch_geometry(fn, grid)) return rc;
        a.solo = solo_entries(grid, a.listLen);
        MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
        if (!graphCapture) {
            if (int rc = enqueue(a, fn, grid)) return rc;
        } else {
            const long long key = ((long long)relaunchSteps << 32) | ((long long)relaunchLaunches << 16) | (algorithm << 8) |
                                  (kernelMode << 4) | variant;
            const unsigned char* ab = reinterpret_cast<const unsigned char*>(&a);
            const bool same = graphExec && key == graphKey && graphArgs.size() == sizeof a &&
                              std::memcmp(graphArgs.data(), ab, sizeof a) == 0 && graphGrid == grid &&
                              graphBlock == block();
            if (!same) {
                if (graphExec) { (void)hipGraphExecDestroy(graphExec); graphExec = nullptr; }
                if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
                MPGOS_HIP_TRY(hipStreamBeginCapture(strm, hipStreamCaptureModeThreadLocal));
                int rc = enqueue(a, fn, grid);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(strm, &g);
                if (rc) return rc;
                if (e != hipSuccess) { set_error(std::string("graph capture failed: ") + hipGetErrorString(e)); return MPGOS_E_HIP; }
                graph = g;
                MPGOS_HIP_TRY(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0));
                graphArgs.assign(ab, ab + sizeof a);
                graphKey = key; graphGrid = grid; graphBlock = block();
            }
            MPGOS_HIP_TRY(hipGraphLaunch(graphExec, strm));
        }
        MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
        return MPGOS_OK;
    }
    int graphGrid = 0, graphBlock = 0;

    // ------------------------------------------------------------ time-sliced solve
    uint32_t* saved = nullptr;       // lane states of suspended instances [words][NT]
    int* phase = nullptr;            // per instance: 0 not started, 1 suspended, 2 finished
    int* listA = nullptr;            // instance lists of consecutive slices
    int* listB = nullptr;
    const int* sliceList = nullptr;
    int sliceLen = 0;
    bool sliceActive = false;

    static constexpr size_t lane_words()
    {
        const size_t a = Lane<Sys, P, KAHAN, MPGOS_RK4>::kSavedWords, b = Lane<Sys, P, KAHAN, MPGOS_RKCK45>::kSavedWords;
        return a > b ? a : b;
    }

    // One slice of a Solve: every unfinished instance attempts at most
    // max_attempted_steps steps in this launch, then is suspended (lane state
    // kept in HBM) or finished.  The first call of a sequence starts the Solve
    // (initialisation), the call that returns *unfinished == 0 completes it;
    // the results are bit-identical to Solve().  Synchronous.
    int solve_slice(long long budget, int* unfinished) override
    {
        if (budget < 1 || budget > 0x7fffffffLL) { set_error("solve_slice: step budget must be in 1..2^31-1"); return MPGOS_E_ARG; }
        if (graphCapture) { set_error("solve_slice: not available with GraphCapture"); return MPGOS_E_STATE; }
        MPGOS_HIP_TRY(hipSetDevice(device));
        if (int rc = ensure_slice_buffers()) return rc;
        if (!sliceActive) {
            MPGOS_HIP_TRY(hipMemsetAsync(phase, 0, 4 * (size_t)NT, strm));
            sliceList = nullptr;
            if (ordered()) {
                if (int rc = ensure_order()) return rc;
                MPGOS_HIP_TRY(hipMemcpyAsync(listA, order, sizeof(int) * (size_t)nActive, hipMemcpyDeviceToDevice, strm));
                sliceList = listA;
            }
            sliceLen = nActive;
            sliceActive = true;
        }
        if (sliceLen == 0) { sliceActive = false; *unfinished = 0; return MPGOS_OK; }
        Args<Sys, P, CP> a = make_args();
        a.saved = saved; a.phase = phase; a.list = sliceList; a.listLen = sliceLen; a.slice = (int)budget;
        a.listOut = sliceList == listA ? listB : listA;
        KernelFn fn = kernel(true);
        int grid = 0;
        if (int rc = launch_geometry(fn, grid)) return rc;
        if (kernelMode == MPGOS_KERNEL_PLAIN) grid = (sliceLen + block() - 1) / block();
        a.solo = solo_entries(grid, sliceLen);
        MPGOS_HIP_TRY(hipMemsetAsync(queue, 0, 12, strm));
        MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
        hipLaunchKernelGGL(fn, dim3(grid), dim3(block()), dyn_lds(), strm, a);
        MPGOS_HIP_TRY(hipGetLastError());
        MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
        unsigned left = 0;
        MPGOS_HIP_TRY(hipMemcpyAsync(&left, queue + 1, 4, hipMemcpyDeviceToHost, strm));
        MPGOS_HIP_TRY(hipStreamSynchronize(strm));
        sliceList = a.listOut;
        sliceLen = (int)left;
        if (ordered() && sliceLen > 1) {
            // the kernel appends suspended instances in completion order: restore the queue order
            std::vector<int> l((size_t)sliceLen);
            MPGOS_HIP_TRY(hipMemcpy(l.data(), a.listOut, sizeof(int) * l.size(), hipMemcpyDeviceToHost));
            std::sort(l.begin(), l.end(), [&](int x, int y) { return rankHost[(size_t)x] < rankHost[(size_t)y]; });
            MPGOS_HIP_TRY(hipMemcpy(a.listOut, l.data(), sizeof(int) * l.size(), hipMemcpyHostToDevice));
        }
        if (sliceLen == 0) sliceActive = false;
        *unfinished = sliceLen;
        return MPGOS_OK;
    }

    // ------------------------------------------------------------ adjoint (pt_adjoint.h)
    static constexpr bool kAdjoint = Sys::HAS_VJP && std::is_same<P, double>::value && !KAHAN;

    int adjoint_args(AdjArgs<Sys>& A)
    {
        if constexpr (kAdjoint) {
            A.a = make_args();
            A.ckpt = ckpt; A.seg = seg; A.lam = lam.d; A.gp = gp.d; A.K = adjK; A.ckcap = adjCap;
        }
        return MPGOS_OK;
    }

    int adjoint_record(int K, int capacity) override
    {
        if constexpr (!kAdjoint) {
            set_error("adjoint: needs FP64 and a system with a right-hand-side VJP");
            return MPGOS_E_SYSTEM;
        } else {
            if (K < 1 || capacity < 1) { set_error("adjoint: checkpoint interval and capacity must be >= 1"); return MPGOS_E_ARG; }
            MPGOS_HIP_TRY(hipSetDevice(device));
            const size_t needC = (size_t)capacity * checkpoint_words<Sys>() * NT;
            const size_t needS = segment_in_lds(K) ? 0 : (size_t)K * (SD + 2) * NT;   // HBM segment buffer only when LDS is too small
            if (needC > ckptCount || needS > segCount) MPGOS_HIP_TRY(hipStreamSynchronize(strm));
            if (needC > ckptCount) {
                if (ckpt) (void)hipFree(ckpt);
                ckpt = nullptr; ckptCount = 0;
                if (hipMalloc((void**)&ckpt, needC * sizeof(double)) != hipSuccess) { set_error("adjoint: checkpoint allocation failed"); return MPGOS_E_MEMORY; }
                ckptCount = needC;
            }
            if (needS > segCount) {
                if (seg) (void)hipFree(seg);
                seg = nullptr; segCount = 0;
                if (hipMalloc((void**)&seg, needS * sizeof(double)) != hipSuccess) { set_error("adjoint: segment allocation failed"); return MPGOS_E_MEMORY; }
                segCount = needS;
            }
            adjK = K; adjCap = capacity;
            AdjArgs<Sys> A;
            adjoint_args(A);
            const int B = blockSize;
            const int grid = NT / B + (NT % B == 0 ? 0 : 1);
            const size_t lds = preferShared ? (size_t)NSP * sizeof(P) + (size_t)NISP * sizeof(int) : 0;
            MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
            if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((adjoint_record_kernel<Sys, MPGOS_RK4>), dim3(grid), dim3(B), lds, strm, A);
            else hipLaunchKernelGGL((adjoint_record_kernel<Sys, MPGOS_RKCK45>), dim3(grid), dim3(B), lds, strm, A);
            MPGOS_HIP_TRY(hipGetLastError());
            MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
            return MPGOS_OK;
        }
    }

    // Segment records of the reverse pass in LDS: K records of SD + 2 doubles
    // per lane of the workgroup (lanes counted in whole wavefronts), when they
    // fit in 64 KiB with the staged shared parameters.
    size_t segment_lds_bytes(int K) const { return (size_t)K * (SD + 2) * sizeof(double) * (size_t)((blockSize + 63) / 64 * 64); }
    bool segment_in_lds(int K) const
    {
        return adjoint_shared_bytes<Sys>(preferShared) + segment_lds_bytes(K) <= (size_t)64 * 1024;
    }

    int adjoint_reverse() override
    {
        if constexpr (!kAdjoint) {
            set_error("adjoint: needs FP64 and a system with a right-hand-side VJP");
            return MPGOS_E_SYSTEM;
        } else {
            if (!ckpt || adjK == 0) { set_error("adjoint: no recorded forward pass (call mpgos_pt_adjoint_record first)"); return MPGOS_E_STATE; }
            MPGOS_HIP_TRY(hipSetDevice(device));
            AdjArgs<Sys> A;
            adjoint_args(A);
            A.a.NDO = 0;       // the recomputation writes nothing but the segment buffer
            A.a.evcap = 0;
            const int B = blockSize;
            const int grid = NT / B + (NT % B == 0 ? 0 : 1);
            const bool ldsSeg = segment_in_lds(adjK);
            const size_t lds = adjoint_shared_bytes<Sys>(preferShared) + (ldsSeg ? segment_lds_bytes(adjK) : 0);
            if (!ldsSeg && segCount < (size_t)adjK * (SD + 2) * NT) {
                set_error("adjoint: the segment records no longer fit in LDS (ThreadsPerBlock changed since the record pass); record again");
                return MPGOS_E_STATE;
            }
            MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
            if (ldsSeg) {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((adjoint_reverse_kernel<Sys, MPGOS_RK4, true>), dim3(grid), dim3(B), lds, strm, A);
                else hipLaunchKernelGGL((adjoint_reverse_kernel<Sys, MPGOS_RKCK45, true>), dim3(grid), dim3(B), lds, strm, A);
            } else {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((adjoint_reverse_kernel<Sys, MPGOS_RK4, false>), dim3(grid), dim3(B), lds, strm, A);
                else hipLaunchKernelGGL((adjoint_reverse_kernel<Sys, MPGOS_RKCK45, false>), dim3(grid), dim3(B), lds, strm, A);
            }
            MPGOS_HIP_TRY(hipGetLastError());
            MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
            return MPGOS_OK;
        }
    }

    int synchronise_device() override { MPGOS_HIP_TRY(hipSetDevice(device)); MPGOS_HIP_TRY(hipDeviceSynchronize()); return MPGOS_OK; }
    int insert_sync_point() override { MPGOS_HIP_TRY(hipSetDevice(device)); MPGOS_HIP_TRY(hipEventRecord(syncEvent, strm)); return MPGOS_OK; }
    int synchronise_solver() override { MPGOS_HIP_TRY(hipSetDevice(device)); MPGOS_HIP_TRY(hipEventSynchronize(syncEvent)); return MPGOS_OK; }
    int last_solve_ms(double* ms) const override
    {
        float f = 0.f;
        MPGOS_HIP_TRY(hipEventElapsedTime(&f, evStart, evStop));
        *ms = f;
        return MPGOS_OK;
    }
    void* stream() override { return (void*)strm; }
    int device_index() const override { return device; }

    // ------------------------------------------------------------ Print (PT-IF:1082-1276)
    template <class T_>
    static void write_general(const std::string& file, int rows, int cols, const T_* data)
    {
        std::ofstream f(file);
        f.precision(10);
        f.flags(std::ios::scientific);
        for (int i = 0; i < rows; i++) {
            for (int j = 0; j < cols; j++) {
                f.width(18);
                f << data[i + (size_t)j * rows];
                if (j < cols - 1) f << ',';
            }
            f << '\n';
        }
    }

    int print(int var, int tid, const char* dir) override
    {
        const std::string pre = dir ? std::string(dir) + "/" : std::string();
        switch (var) {
        case MPGOS_PT_TIME_DOMAIN: write_general(pre + "TimeDomainInSolverObject.txt", NT, 2, TD.h); return MPGOS_OK;
        case MPGOS_PT_ACTUAL_STATE: write_general(pre + "ActualStateInSolverObject.txt", NT, SD, X.h); return MPGOS_OK;
        case MPGOS_PT_ACTUAL_TIME: write_general(pre + "ActualTimeInSolverObject.txt", NT, 1, T.h); return MPGOS_OK;
        case MPGOS_PT_CONTROL_PARAMETERS: write_general(pre + "ControlParametersInSolverObject.txt", NT, NCP, cPAR.h); return MPGOS_OK;
        case MPGOS_PT_SHARED_PARAMETERS: write_general(pre + "SharedParametersInSolverObject.txt", NSP, 1, sPAR.h); return MPGOS_OK;
        case MPGOS_PT_INTEGER_SHARED_PARAMETERS: write_general(pre + "IntegerSharedParametersInSolverObject.txt", NISP, 1, sPARi.h); return MPGOS_OK;
        case MPGOS_PT_ACCESSORIES: write_general(pre + "AccessoriesInSolverObject.txt", NT, NA, ACC.h); return MPGOS_OK;
        case MPGOS_PT_INTEGER_ACCESSORIES: write_general(pre + "IntegerAccessoriesInSolverObject.txt", NT, NIA, ACCi.h); return MPGOS_OK;
        case MPGOS_PT_DENSE_OUTPUT: break;
        default: set_error("Print: variable not applicable"); return MPGOS_E_ARG;
        }
        if (int rc = bound("Print", "Thread", tid, 0, NT - 1)) return rc;
        std::ofstream f(pre + "DenseOutput_" + std::to_string(tid) + ".txt");
        f.precision(10);
        f.flags(std::ios::scientific);
        auto row = [&](const char* title, int n, auto get) {
            f << title << ":\n";
            for (int i = 0; i < n; i++) { f.width(18); f << get(i); if (i < n - 1) f << ','; }
        };
        row("ControlParameters", NCP, [&](int i) { return cPAR.h[tid + (size_t)i * NT]; }); f << '\n';
        row("SharedParameters", NSP, [&](int i) { return sPAR.h[i]; }); f << '\n';
        row("IntegerSharedParameters", NISP, [&](int i) { return sPARi.h[i]; }); f << '\n';
        row("Accessories", NA, [&](int i) { return ACC.h[tid + (size_t)i * NT]; }); f << '\n';
        row("IntegerAccessories", NIA, [&](int i) { return ACCi.h[tid + (size_t)i * NT]; }); f << "\n\n";
        f << "Time series:\n";
        if (NDO > 0) {
            const int n = DOidx.h[tid] < NDO ? DOidx.h[tid] : NDO;
            for (int i = 0; i < n; i++) {
                f.width(18); f << DOt.h[tid + (size_t)i * NT] << ',';
                for (int j = 0; j < SD; j++) {
                    f.width(18); f << DOx.h[tid + (size_t)j * NT + (size_t)i * NT * SD];
                    if (j < SD - 1) f << ',';
                }
                f << '\n';
            }
        }
        return MPGOS_OK;
    }
};

} // namespace pt
} // namespace mpgos
