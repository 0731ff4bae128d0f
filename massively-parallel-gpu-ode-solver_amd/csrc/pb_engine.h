// Host side of the per-block module: buffers, options, Set/GetHost with the
// reference's scopes and coupling-matrix storage forms, transfers and launches
// of pb_device.h.  Implements the member functions of the reference's
// per-block ProblemSolver (SourceCodes/CoupledSystems_PerBlock_Interface.cuh:
// constructor 353-657, SolverOption 772-865, SetHost 868-1140 with the
// coupling-matrix mapping 994-1081, Synchronise* 1143-1337, GetHost 1339-1529,
// Solve 1949-1965) with return codes instead of exit().
//
// Unit-scope layout: (system * UPS + unit) + component * NS*UPS (the reference
// pads each block's logical threads; the padding is an artefact of its launch
// shape and is not visible through SetHost/GetHost, so it is not kept).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "mpgos.h"
#include "pb_adjoint.h"
#include "pb_device.h"
#include "pt_engine.h"   // Buffer, set_error, MPGOS_HIP_TRY

namespace mpgos {
namespace pb {

using pt::Buffer;

class SolverBase {
public:
    virtual ~SolverBase() {}
    virtual int init(int algorithm, int ns, int ups, int spb, int cbw, int cci, int ndo, int device) = 0;
    virtual int dims(mpgos_pb_dims* d) const = 0;
    virtual int option(int opt, int idx, double v) = 0;
    virtual int set_host(int var, int a0, int a1, int a2, int a3, double v) = 0;
    virtual int get_host(int var, int a0, int a1, int a2, int a3, double* v) const = 0;
    virtual void* buffer(int var, bool host, size_t* nbytes) = 0;
    virtual int sync(int var, bool to_device) = 0;
    virtual int solve() = 0;
    virtual int synchronise_device() = 0;
    virtual int insert_sync_point() = 0;
    virtual int synchronise_solver() = 0;
    virtual int last_solve_ms(double* ms) const = 0;
    virtual void* stream() = 0;
    virtual int adjoint_record(int K, int checkpoint_capacity, int step_capacity) = 0;
    virtual int adjoint_reverse() = 0;
};

// systems that provide OdeVJP (needed by the discrete adjoint)
template <class Sys, class = void> struct HasVjp : std::false_type {};
template <class Sys>
struct HasVjp<Sys, decltype((void)&Sys::template OdeVJP<double>)> : std::true_type {};

// KAHAN (P = float): the FLOAT_KAHAN mode (compensated FP32 state and time,
// FP64 step size, error norm, controller and event step control).
template <class Sys, class P, bool KAHAN = false>
class Solver : public SolverBase {
public:
    using CP = typename std::conditional<KAHAN, double, P>::type;
    static constexpr int UD = Sys::UD, NC = Sys::NC, NUP = Sys::NUP, NSP = Sys::NSP, NGP = Sys::NGP, NiGP = Sys::NiGP;
    static constexpr int NUA = Sys::NUA, NiUA = Sys::NiUA, NSA = Sys::NSA, NiSA = Sys::NiSA, NE = Sys::NE;

    int algorithm = MPGOS_RKCK45, NS = 0, UPS = 0, SPB = 1, CBW = 0, CCI = 0, NDO = 0, device = 0;
    int cMatSize = 0;   // elements per coupling matrix in the stored form
    hipStream_t strm = nullptr;
    hipEvent_t syncEvent = nullptr, evStart = nullptr, evStop = nullptr;
    size_t ldsMax = 0;

    int nActive = 0, sharedGlobals = 1, sharedMatrix = 0, doFreq = 1;
    int couplingKernel = 0;   // 0 automatic, 1 lane-per-pair, 2 dense contraction (MFMA)
    int lastForm = -1;        // coupling form of the last launch
    int mfmaPairs = 4;        // units per lane of the dense-contraction form (2 or 4)
    double dt0 = 1e-2, dtmax = 1.0e6, dtmin = 1.0e-12, grow = 5.0, shrink = 0.1, doMinDt = 0.0;
    std::vector<double> rtol, atol, etol;
    std::vector<int> edir;

    Buffer<P> X, uPAR, uACC, TD, T, sPAR, sACC, strength, DOt, DOx, gPAR, cMat;
    Buffer<int> uACCi, sACCi, DOidx, gPARi, cIndex, status;
    Buffer<long long> nacc, nrej;
    std::vector<P> cMatOriginal;   // the full matrices as set (GetHost), NC * UPS * UPS

    // discrete adjoint (FP64 systems with OdeVJP)
    static constexpr bool kAdjoint = std::is_same<P, double>::value && HasVjp<Sys>::value;
    Buffer<P> lamB, gUB, gSB;           // dL/dx, dL/d(unit parameters), dL/d(system parameters)
    double* recX = nullptr; double* recT = nullptr; double* recH = nullptr; double* seg = nullptr; double* cMatT = nullptr;
    double* stg = nullptr; double* gSp = nullptr;
    double* cTile = nullptr;                // pair-tiled copies (tile_pair_matrix) of the dense coupling matrices
    double* cTileT = nullptr;               // and of their transposes (adjoint sweep)
    size_t cTileN = 0, cTileTN = 0;
    size_t recXn = 0, recTn = 0, recHn = 0, segn = 0, cMatTn = 0, stgn = 0, gSpn = 0;
    int adjK = 0, adjCap = 0, adjStepCap = 0;

    ~Solver() override
    {
        if (strm) (void)hipStreamSynchronize(strm);
        lamB.release(); gUB.release(); gSB.release();
        for (double* q : {recX, recT, recH, seg, cMatT, stg, gSp, cTile, cTileT})
            if (q) (void)hipFree(q);
        for (auto* b : {&X, &uPAR, &uACC, &TD, &T, &sPAR, &sACC, &strength, &DOt, &DOx, &gPAR, &cMat}) b->release();
        for (auto* b : {&uACCi, &sACCi, &DOidx, &gPARi, &cIndex, &status}) b->release();
        nacc.release(); nrej.release();
        if (syncEvent) (void)hipEventDestroy(syncEvent);
        if (evStart) (void)hipEventDestroy(evStart);
        if (evStop) (void)hipEventDestroy(evStop);
        if (strm) (void)hipStreamDestroy(strm);
    }

    int init(int alg, int ns, int ups, int spb, int cbw, int cci, int ndo, int dev) override
    {
        if (alg != MPGOS_RK4 && alg != MPGOS_RKCK45) { set_error("unknown algorithm"); return MPGOS_E_ARG; }
        if (ns <= 0 || ups <= 0 || spb <= 0 || ndo < 0 || cbw < 0 || (cci != 0 && cci != 1)) {
            set_error("per-block: NS, UPS, SPB must be positive, NDO >= 0, CBW >= 0, CCI in {0, 1}");
            return MPGOS_E_ARG;
        }
        if ((long)spb * ups > kMaxBlock) { set_error("per-block: SPB * UPS must not exceed 1024"); return MPGOS_E_ARG; }
        if (cbw > 0 && 2 * cbw + 1 > ups) { set_error("per-block: 2 * CBW + 1 must not exceed UPS"); return MPGOS_E_ARG; }
        algorithm = alg; NS = ns; UPS = ups; SPB = spb; CBW = cbw; CCI = cci; NDO = ndo; device = dev; nActive = ns;
        if (CCI == 0 && CBW == 0) cMatSize = UPS * UPS;
        else if (CCI == 0) cMatSize = UPS * (2 * CBW + 1);
        else if (CBW > 0) cMatSize = 2 * CBW + 1;
        else cMatSize = UPS;
        MPGOS_HIP_TRY(hipSetDevice(device));
        MPGOS_HIP_TRY(hipStreamCreateWithFlags(&strm, hipStreamNonBlocking));
        MPGOS_HIP_TRY(hipEventCreateWithFlags(&syncEvent, hipEventDisableTiming));
        MPGOS_HIP_TRY(hipEventCreate(&evStart));
        MPGOS_HIP_TRY(hipEventCreate(&evStop));
        hipDeviceProp_t prop;
        MPGOS_HIP_TRY(hipGetDeviceProperties(&prop, device));
        ldsMax = prop.sharedMemPerBlock;
        const size_t NU = (size_t)NS * UPS;
        int rc;
        if ((rc = X.alloc(UD * NU)) || (rc = uPAR.alloc(NUP * NU)) || (rc = uACC.alloc(NUA * NU)) ||
            (rc = uACCi.alloc(NiUA * NU)) || (rc = TD.alloc(2 * (size_t)NS)) || (rc = T.alloc(NS)) ||
            (rc = sPAR.alloc((size_t)NSP * NS)) || (rc = sACC.alloc((size_t)NSA * NS)) || (rc = sACCi.alloc((size_t)NiSA * NS)) ||
            (rc = strength.alloc((size_t)NC * NS)) || (rc = DOidx.alloc(NS)) || (rc = DOt.alloc((size_t)NDO * NS)) ||
            (rc = DOx.alloc((size_t)NDO * UD * NU)) || (rc = gPAR.alloc(NGP)) || (rc = gPARi.alloc(NiGP)) ||
            (rc = cIndex.alloc(NC)) || (rc = cMat.alloc((size_t)NC * cMatSize)) || (rc = status.alloc(NS)) ||
            (rc = nacc.alloc(NS)) || (rc = nrej.alloc(NS)))
            return rc;
        if (kAdjoint && ((rc = lamB.alloc(UD * NU)) || (rc = gUB.alloc(NUP * NU)) || (rc = gSB.alloc((size_t)NSP * NS))))
            return rc;
        cMatOriginal.assign((size_t)NC * UPS * UPS, P(0));
        rtol.assign(UD, 1e-8);
        atol.assign(UD, 1e-8);
        etol.assign(NE, 1e-6);
        edir.assign(NE, 0);
        if ((rc = copy(status, true)) || (rc = copy(nacc, true)) || (rc = copy(nrej, true))) return rc;
        return sync(MPGOS_PB_ALL, true);
    }

    int dims(mpgos_pb_dims* d) const override
    {
        d->ns = NS; d->ups = UPS; d->ud = UD; d->spb = SPB; d->nc = NC; d->cbw = CBW; d->cci = CCI;
        d->nup = NUP; d->nsp = NSP; d->ngp = NGP; d->nigp = NiGP; d->nua = NUA; d->niua = NiUA; d->nsa = NSA;
        d->nisa = NiSA; d->ne = NE; d->ndo = NDO; d->algorithm = algorithm;
        d->precision = KAHAN ? MPGOS_FLOAT_KAHAN : (std::is_same<P, float>::value ? MPGOS_FLOAT : MPGOS_DOUBLE);
        return MPGOS_OK;
    }

    static int bound(const char* fn, const char* what, int v, int lo, int hi) { return bound_check(fn, what, v, lo, hi); }

    int option(int opt, int idx, double v) override
    {
        switch (opt) {
        case MPGOS_PB_INITIAL_TIME_STEP: dt0 = v; return MPGOS_OK;
        case MPGOS_PB_ACTIVE_SYSTEMS:
            if ((int)v < 0 || (int)v > NS) { set_error("ActiveSystems must be in 0..NS"); return MPGOS_E_ARG; }
            nActive = (int)v; return MPGOS_OK;
        case MPGOS_PB_MAXIMUM_TIME_STEP: dtmax = v; return MPGOS_OK;
        case MPGOS_PB_MINIMUM_TIME_STEP: dtmin = v; return MPGOS_OK;
        case MPGOS_PB_TIME_STEP_GROW_LIMIT: grow = v; return MPGOS_OK;
        case MPGOS_PB_TIME_STEP_SHRINK_LIMIT: shrink = v; return MPGOS_OK;
        case MPGOS_PB_RELATIVE_TOLERANCE:
            if (int rc = bound("SolverOption", "RelativeTolerance", idx, 0, UD - 1)) return rc;
            rtol[idx] = v; return MPGOS_OK;
        case MPGOS_PB_ABSOLUTE_TOLERANCE:
            if (int rc = bound("SolverOption", "AbsoluteTolerance", idx, 0, UD - 1)) return rc;
            atol[idx] = v; return MPGOS_OK;
        case MPGOS_PB_EVENT_TOLERANCE:
            if (int rc = bound("SolverOption", "EventTolerance", idx, 0, NE - 1)) return rc;
            etol[idx] = v; return MPGOS_OK;
        case MPGOS_PB_EVENT_DIRECTION:
            if (int rc = bound("SolverOption", "EventDirection", idx, 0, NE - 1)) return rc;
            edir[idx] = (int)v; return MPGOS_OK;
        case MPGOS_PB_DENSE_OUTPUT_SAVE_FREQUENCY: doFreq = (int)v; return MPGOS_OK;
        case MPGOS_PB_DENSE_OUTPUT_MINIMUM_TIME_STEP: doMinDt = v; return MPGOS_OK;
        case MPGOS_PB_SHARED_GLOBAL_VARIABLES: sharedGlobals = (int)v != 0; return MPGOS_OK;
        case MPGOS_PB_SHARED_COUPLING_MATRICES: sharedMatrix = (int)v != 0; return MPGOS_OK;
        case MPGOS_PB_MFMA_UNITS_PER_LANE:
            if ((int)v != 2 && (int)v != 4) { set_error("MfmaUnitsPerLane must be 2 or 4"); return MPGOS_E_ARG; }
            mfmaPairs = (int)v;
            return MPGOS_OK;
        case MPGOS_PB_COUPLING_KERNEL:
            if ((int)v < 0 || (int)v > 2) { set_error("CouplingKernel must be 0 (automatic), 1 (lane per pair) or 2 (dense contraction)"); return MPGOS_E_ARG; }
            couplingKernel = (int)v;
            return MPGOS_OK;
        default:
            set_error("SolverOption: unknown or not applicable option " + std::to_string(opt));
            return MPGOS_E_ARG;
        }
    }

    // Maps (variable, a0..a3) to an element of a host mirror; index rules and
    // bound checks of PB-IF:868-1140.
    template <class F>
    int access(const char* fn, int var, int a0, int a1, int a2, int a3, bool setting, double v, F&& f)
    {
        int rc;
        const size_t NU = (size_t)NS * UPS;
        switch (var) {
        case MPGOS_PB_ACTUAL_STATE: case MPGOS_PB_UNIT_PARAMETERS: case MPGOS_PB_UNIT_ACCESSORIES:
        case MPGOS_PB_INTEGER_UNIT_ACCESSORIES: {
            if ((rc = bound(fn, "SystemNumber", a0, 0, NS - 1)) || (rc = bound(fn, "UnitNumber", a1, 0, UPS - 1))) return rc;
            const size_t idx = (size_t)a0 * UPS + a1 + (size_t)a2 * NU;
            switch (var) {
            case MPGOS_PB_ACTUAL_STATE: if ((rc = bound(fn, "ActualState", a2, 0, UD - 1))) return rc; f(&X.h[idx]); break;
            case MPGOS_PB_UNIT_PARAMETERS: if ((rc = bound(fn, "UnitParameters", a2, 0, NUP - 1))) return rc; f(&uPAR.h[idx]); break;
            case MPGOS_PB_UNIT_ACCESSORIES: if ((rc = bound(fn, "UnitAccessories", a2, 0, NUA - 1))) return rc; f(&uACC.h[idx]); break;
            default: if ((rc = bound(fn, "IntegerUnitAccessories", a2, 0, NiUA - 1))) return rc; f(&uACCi.h[idx]); break;
            }
            return MPGOS_OK;
        }
        case MPGOS_PB_TIME_DOMAIN: case MPGOS_PB_SYSTEM_PARAMETERS: case MPGOS_PB_SYSTEM_ACCESSORIES:
        case MPGOS_PB_INTEGER_SYSTEM_ACCESSORIES: case MPGOS_PB_DENSE_TIME: case MPGOS_PB_COUPLING_STRENGTH: {
            if ((rc = bound(fn, "SystemNumber", a0, 0, NS - 1))) return rc;
            const size_t idx = (size_t)a0 + (size_t)a2 * NS;
            switch (var) {
            case MPGOS_PB_TIME_DOMAIN: if ((rc = bound(fn, "TimeDomain", a2, 0, 1))) return rc; f(&TD.h[idx]); break;
            case MPGOS_PB_SYSTEM_PARAMETERS: if ((rc = bound(fn, "SystemParameters", a2, 0, NSP - 1))) return rc; f(&sPAR.h[idx]); break;
            case MPGOS_PB_SYSTEM_ACCESSORIES: if ((rc = bound(fn, "SystemAccessories", a2, 0, NSA - 1))) return rc; f(&sACC.h[idx]); break;
            case MPGOS_PB_INTEGER_SYSTEM_ACCESSORIES: if ((rc = bound(fn, "IntegerSystemAccessories", a2, 0, NiSA - 1))) return rc; f(&sACCi.h[idx]); break;
            case MPGOS_PB_DENSE_TIME: if ((rc = bound(fn, "DenseTime", a2, 0, NDO - 1))) return rc; f(&DOt.h[idx]); break;
            default: if ((rc = bound(fn, "CouplingStrength", a2, 0, NC - 1))) return rc; f(&strength.h[idx]); break;
            }
            return MPGOS_OK;
        }
        case MPGOS_PB_GLOBAL_PARAMETERS:
            if ((rc = bound(fn, "GlobalParameters", a2, 0, NGP - 1))) return rc;
            f(&gPAR.h[a2]); return MPGOS_OK;
        case MPGOS_PB_INTEGER_GLOBAL_PARAMETERS:
            if ((rc = bound(fn, "IntegerGlobalParameters", a2, 0, NiGP - 1))) return rc;
            f(&gPARi.h[a2]); return MPGOS_OK;
        case MPGOS_PB_COUPLING_INDEX:
            if ((rc = bound(fn, "CouplingIndex", a2, 0, NC - 1))) return rc;
            if (setting && ((int)v < 0 || (int)v >= UD)) { set_error(std::string(fn) + ": CouplingIndex value must be a component 0..UD-1"); return MPGOS_E_ARG; }
            f(&cIndex.h[a2]); return MPGOS_OK;
        case MPGOS_PB_COUPLING_MATRIX: {
            // a0 = coupling, a1 = row, a2 = col (PB-IF:994-1081)
            if ((rc = bound(fn, "CouplingNumber", a0, 0, NC - 1)) || (rc = bound(fn, "Coupling Matrix Row", a1, 0, UPS - 1)) ||
                (rc = bound(fn, "Coupling Matrix Col", a2, 0, UPS - 1)))
                return rc;
            const size_t orig = (size_t)a1 + (size_t)a2 * UPS + (size_t)a0 * UPS * UPS;
            if (!setting) { f(&cMatOriginal[orig]); return MPGOS_OK; }
            size_t idx;
            if (CCI == 0 && CBW == 0) idx = orig;
            else {
                int mc = a2 - a1 + CBW;
                if (mc >= UPS) mc -= UPS;
                if (mc < 0) mc += UPS;
                const int lim = CBW > 0 ? 2 * CBW + 1 : UPS;
                if ((rc = bound(fn, "Coupling Matrix Shifted Diagonal", mc, 0, lim - 1))) return rc;
                if (CCI == 0) idx = (size_t)a1 + (size_t)mc * UPS + (size_t)a0 * UPS * lim;
                else idx = (size_t)mc + (size_t)a0 * lim;
            }
            cMatOriginal[orig] = (P)v;
            f(&cMat.h[idx]);
            return MPGOS_OK;
        }
        case MPGOS_PB_DENSE_INDEX: case MPGOS_PB_ACTUAL_TIME: case MPGOS_PB_ACCEPTED_STEPS: case MPGOS_PB_REJECTED_STEPS:
        case MPGOS_PB_STATUS:
            if ((rc = bound(fn, "SystemNumber", a0, 0, NS - 1))) return rc;
            switch (var) {
            case MPGOS_PB_DENSE_INDEX:
                if (setting && (rc = bound(fn, "DenseIndex-Value", (int)v, 0, NDO - 1))) return rc;
                f(&DOidx.h[a0]); break;
            case MPGOS_PB_ACTUAL_TIME: f(&T.h[a0]); break;
            case MPGOS_PB_ACCEPTED_STEPS: f(&nacc.h[a0]); break;
            case MPGOS_PB_REJECTED_STEPS: f(&nrej.h[a0]); break;
            default: f(&status.h[a0]); break;
            }
            return MPGOS_OK;
        case MPGOS_PB_DENSE_STATE: {
            if ((rc = bound(fn, "SystemNumber", a0, 0, NS - 1)) || (rc = bound(fn, "UnitNumber", a1, 0, UPS - 1)) ||
                (rc = bound(fn, "ComponentNumber in dense state", a2, 0, UD - 1)) || (rc = bound(fn, "DenseState", a3, 0, NDO - 1)))
                return rc;
            f(&DOx.h[(size_t)a0 * UPS + a1 + (size_t)a2 * NU + (size_t)a3 * NU * UD]);
            return MPGOS_OK;
        }
        default:
            set_error(std::string(fn) + ": variable not applicable with this argument configuration");
            return MPGOS_E_ARG;
        }
    }

    int set_host(int var, int a0, int a1, int a2, int a3, double v) override
    {
        if (var == MPGOS_PB_ACCEPTED_STEPS || var == MPGOS_PB_REJECTED_STEPS || var == MPGOS_PB_STATUS) {
            set_error("SetHost: output-only variable");
            return MPGOS_E_ARG;
        }
        return access("SetHost", var, a0, a1, a2, a3, true, v, [&](auto* p) {
            using T_ = typename std::remove_pointer<decltype(p)>::type;
            *p = std::is_integral<T_>::value ? (T_)(int)v : (T_)v;
        });
    }

    int get_host(int var, int a0, int a1, int a2, int a3, double* v) const override
    {
        return const_cast<Solver*>(this)->access("GetHost", var, a0, a1, a2, a3, false, 0.0, [&](auto* p) { *v = (double)*p; });
    }

    template <class T_> static void* ret(Buffer<T_>& b, bool host, size_t* nbytes)
    {
        if (nbytes) *nbytes = b.bytes();
        return host ? (void*)b.h : (void*)b.d;
    }

    void* buffer(int var, bool host, size_t* nbytes) override
    {
        switch (var) {
        case MPGOS_PB_TIME_DOMAIN: return ret(TD, host, nbytes);
        case MPGOS_PB_ACTUAL_TIME: return ret(T, host, nbytes);
        case MPGOS_PB_ACTUAL_STATE: return ret(X, host, nbytes);
        case MPGOS_PB_UNIT_PARAMETERS: return ret(uPAR, host, nbytes);
        case MPGOS_PB_SYSTEM_PARAMETERS: return ret(sPAR, host, nbytes);
        case MPGOS_PB_GLOBAL_PARAMETERS: return ret(gPAR, host, nbytes);
        case MPGOS_PB_INTEGER_GLOBAL_PARAMETERS: return ret(gPARi, host, nbytes);
        case MPGOS_PB_UNIT_ACCESSORIES: return ret(uACC, host, nbytes);
        case MPGOS_PB_INTEGER_UNIT_ACCESSORIES: return ret(uACCi, host, nbytes);
        case MPGOS_PB_SYSTEM_ACCESSORIES: return ret(sACC, host, nbytes);
        case MPGOS_PB_INTEGER_SYSTEM_ACCESSORIES: return ret(sACCi, host, nbytes);
        case MPGOS_PB_COUPLING_MATRIX: return ret(cMat, host, nbytes);
        case MPGOS_PB_COUPLING_STRENGTH: return ret(strength, host, nbytes);
        case MPGOS_PB_COUPLING_INDEX: return ret(cIndex, host, nbytes);
        case MPGOS_PB_DENSE_INDEX: return ret(DOidx, host, nbytes);
        case MPGOS_PB_DENSE_TIME: return ret(DOt, host, nbytes);
        case MPGOS_PB_DENSE_STATE: return ret(DOx, host, nbytes);
        case MPGOS_PB_ACCEPTED_STEPS: return ret(nacc, host, nbytes);
        case MPGOS_PB_REJECTED_STEPS: return ret(nrej, host, nbytes);
        case MPGOS_PB_STATUS: return ret(status, host, nbytes);
        case MPGOS_PB_ADJOINT_STATE: return ret(lamB, host, nbytes);
        case MPGOS_PB_UNIT_PARAMETER_GRADIENT: return ret(gUB, host, nbytes);
        case MPGOS_PB_SYSTEM_PARAMETER_GRADIENT: return ret(gSB, host, nbytes);
        default:
            set_error("buffer: variable has no single buffer");
            if (nbytes) *nbytes = 0;
            return nullptr;
        }
    }

    template <class T_> int copy(Buffer<T_>& b, bool to_device)
    {
        if (b.n == 0) return MPGOS_OK;
        if (to_device) MPGOS_HIP_TRY(hipMemcpyAsync(b.d, b.h, b.bytes(), hipMemcpyHostToDevice, strm));
        else MPGOS_HIP_TRY(hipMemcpyAsync(b.h, b.d, b.bytes(), hipMemcpyDeviceToHost, strm));
        return MPGOS_OK;
    }

    // PB-IF:1143-1337
    int sync(int var, bool to_device) override
    {
        MPGOS_HIP_TRY(hipSetDevice(device));
        int rc = MPGOS_OK;
        switch (var) {
        case MPGOS_PB_ALL:
            for (int v = MPGOS_PB_TIME_DOMAIN; v <= MPGOS_PB_DENSE_STATE; v++) {
                if (v == MPGOS_PB_DENSE_OUTPUT) continue;
                if ((rc = sync(v, to_device))) return rc;
            }
            if (!to_device && ((rc = copy(nacc, false)) || (rc = copy(nrej, false)) || (rc = copy(status, false)))) return rc;
            return MPGOS_OK;
        case MPGOS_PB_DENSE_OUTPUT:
            if ((rc = copy(DOidx, to_device)) || (rc = copy(DOt, to_device)) || (rc = copy(DOx, to_device))) return rc;
            return MPGOS_OK;
        case MPGOS_PB_ACCEPTED_STEPS: case MPGOS_PB_REJECTED_STEPS: case MPGOS_PB_STATUS:
            if (to_device) { set_error("SynchroniseFromHostToDevice: output-only variable"); return MPGOS_E_ARG; }
            return var == MPGOS_PB_ACCEPTED_STEPS ? copy(nacc, false) : var == MPGOS_PB_REJECTED_STEPS ? copy(nrej, false) : copy(status, false);
        default: {
            switch (var) {
            case MPGOS_PB_TIME_DOMAIN: return copy(TD, to_device);
            case MPGOS_PB_ACTUAL_TIME: return copy(T, to_device);
            case MPGOS_PB_ACTUAL_STATE: return copy(X, to_device);
            case MPGOS_PB_UNIT_PARAMETERS: return copy(uPAR, to_device);
            case MPGOS_PB_SYSTEM_PARAMETERS: return copy(sPAR, to_device);
            case MPGOS_PB_GLOBAL_PARAMETERS: return copy(gPAR, to_device);
            case MPGOS_PB_INTEGER_GLOBAL_PARAMETERS: return copy(gPARi, to_device);
            case MPGOS_PB_UNIT_ACCESSORIES: return copy(uACC, to_device);
            case MPGOS_PB_INTEGER_UNIT_ACCESSORIES: return copy(uACCi, to_device);
            case MPGOS_PB_SYSTEM_ACCESSORIES: return copy(sACC, to_device);
            case MPGOS_PB_INTEGER_SYSTEM_ACCESSORIES: return copy(sACCi, to_device);
            case MPGOS_PB_COUPLING_MATRIX: return copy(cMat, to_device);
            case MPGOS_PB_COUPLING_STRENGTH: return copy(strength, to_device);
            case MPGOS_PB_COUPLING_INDEX: return copy(cIndex, to_device);
            case MPGOS_PB_DENSE_INDEX: return copy(DOidx, to_device);
            case MPGOS_PB_DENSE_TIME: return copy(DOt, to_device);
            case MPGOS_PB_DENSE_STATE: return copy(DOx, to_device);
            case MPGOS_PB_ADJOINT_STATE: return copy(lamB, to_device);
            case MPGOS_PB_UNIT_PARAMETER_GRADIENT: return copy(gUB, to_device);
            case MPGOS_PB_SYSTEM_PARAMETER_GRADIENT: return copy(gSB, to_device);
            default: set_error("Synchronise: variable not applicable"); return MPGOS_E_ARG;
            }
        }
        }
    }

    Args<Sys, P, CP> make_args() const
    {
        Args<Sys, P, CP> a;
        std::memset(&a, 0, sizeof a);
        a.X = X.d; a.uPAR = uPAR.d; a.uACC = uACC.d; a.uACCi = uACCi.d;
        a.TD = TD.d; a.T = T.d; a.sPAR = sPAR.d; a.sACC = sACC.d; a.sACCi = sACCi.d; a.strength = strength.d;
        a.DOidx = DOidx.d; a.DOt = DOt.d; a.DOx = DOx.d;
        a.gPAR = gPAR.d; a.gPARi = gPARi.d; a.cIndex = cIndex.d; a.cMat = cMat.d;
        a.nacc = nacc.d; a.nrej = nrej.d; a.status = status.d;
        a.NS = NS; a.UPS = UPS; a.SPB = SPB; a.nActive = nActive; a.NDO = NDO; a.CBW = CBW; a.CCI = CCI;
        a.cMatSize = cMatSize; a.sharedGlobals = sharedGlobals; a.doFreq = doFreq;
        // options are stored in the solver's precision, as the reference does; FP64 in the Kahan mode
        a.dt0 = (CP)(P)dt0; a.dtmax = (CP)(P)dtmax; a.dtmin = (CP)(P)dtmin; a.grow = (CP)(P)grow; a.shrink = (CP)(P)shrink;
        if (KAHAN) { a.dt0 = (CP)dt0; a.dtmax = (CP)dtmax; a.dtmin = (CP)dtmin; a.grow = (CP)grow; a.shrink = (CP)shrink; }
        a.doMinDt = (P)doMinDt;
        for (int i = 0; i < UD; i++) {
            a.rtol[i] = KAHAN ? (CP)rtol[i] : (CP)(P)rtol[i];
            a.atol[i] = KAHAN ? (CP)atol[i] : (CP)(P)atol[i];
        }
        for (int i = 0; i < NE; i++) { a.etol[i] = (P)etol[i]; a.edir[i] = edir[i]; }
        return a;
    }

    // the dense-contraction (MFMA) form applies to full FP64 matrices with UPS a
    // multiple of 32 in [32, 256] (2 * UPS threads fill whole waves);
    // CouplingKernel 0 picks it when it applies
    bool mfma_applies() const
    {
        return std::is_same<P, double>::value && CCI == 0 && CBW == 0 && UPS % 32 == 0 && UPS >= 32 && UPS <= 256;
    }
    int coupling_form() const
    {
        if (couplingKernel == 1) return kLaneCoupling;
        if (couplingKernel == 2) return kMfmaCoupling;
        return mfma_applies() ? kMfmaCoupling : kLaneCoupling;
    }

    size_t lds_bytes(int cpl, int spb, int sm, int ldsPairs = 0) const
    {
        return lds_layout<Sys, P, KAHAN>(cpl, spb, UPS, sharedGlobals, sm, cMatSize, ldsPairs).total;
    }

    // Dense-contraction form with the tiled matrix: the leading column-block
    // pairs of every tile that fit into the LDS the workgroup leaves free (a
    // multiple of 2 pairs, one group of the two-tile contraction loop; at
    // most all).  At U = 256: 6 of 32 pairs, 96 KiB.  (Letting the results
    // overwrite the coupling terms after a barrier made room for 8 pairs but
    // ran 1.4 % slower.)
    int lds_pairs(int spb) const
    {
        const size_t perPair = (size_t)std::max(NC, 1) * (UPS / 16) * 64 * 16;
        const size_t base = lds_bytes(kMfmaCoupling, spb, 0, 1) - perPair;
        if (base >= ldsMax) return 0;
        int lp = (int)std::min<size_t>((ldsMax - base) / perPair, (size_t)(UPS / 8));
        lp -= lp % 2;
        return lp;
    }

    // PB-IF:1949-1965
    int solve() override
    {
        MPGOS_HIP_TRY(hipSetDevice(device));
        Args<Sys, P, CP> a = make_args();
        const int cpl = coupling_form();
        if (cpl == kMfmaCoupling && !mfma_applies()) {
            set_error("per-block: the dense-contraction coupling kernel needs a full FP64 coupling matrix and UPS a multiple of 32 up to 256");
            return MPGOS_E_ARG;
        }
        const int spb = cpl == kMfmaCoupling ? kMfmaSystems : SPB;
        a.SPB = spb;
        // the coupling matrix goes to LDS only when asked for and when it fits
        a.sharedMatrix = sharedMatrix && lds_bytes(cpl, spb, 1) <= ldsMax ? 1 : 0;
        const bool tiled = !KAHAN && cpl == kMfmaCoupling && !a.sharedMatrix && std::is_same<P, double>::value;
        if (tiled && (mfmaPairs == 2 ? mfma_lds_matrix<Sys, 2>() : mfma_lds_matrix<Sys, 4>())) a.ldsPairs = lds_pairs(spb);
        const size_t lds = lds_bytes(cpl, spb, a.sharedMatrix, a.ldsPairs);
        if (lds > ldsMax) { set_error("per-block: system scope and coupling terms exceed the LDS of a workgroup"); return MPGOS_E_ARG; }
        const int threads = cpl == kMfmaCoupling ? kMfmaSystems * UPS / mfmaPairs : ((SPB * UPS + 63) / 64) * 64;
        const int grid = NS / spb + (NS % spb == 0 ? 0 : 1);
        lastForm = cpl;
        MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
        if constexpr (KAHAN) {   // lane-per-pair form only
            if (threads <= 256) {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_kernel<Sys, P, MPGOS_RK4, 256, true>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_kernel<Sys, P, MPGOS_RKCK45, 256, true>), dim3(grid), dim3(threads), lds, strm, a);
            } else {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_kernel<Sys, P, MPGOS_RK4, kMaxBlock, true>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_kernel<Sys, P, MPGOS_RKCK45, kMaxBlock, true>), dim3(grid), dim3(threads), lds, strm, a);
            }
        } else {
            if (tiled) {
                // the matrices in column-block pairs per lane, refreshed every Solve
                // (a kernel on the solver's stream, so graph capture records it)
                if (int rc = tile_pairs((const double*)cMat.d, cTile, cTileN)) return rc;
                a.cTile = (const P*)cTile;
            }
            if (cpl == kMfmaCoupling && mfmaPairs == 2) {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_mfma_kernel<Sys, P, MPGOS_RK4, 2>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_mfma_kernel<Sys, P, MPGOS_RKCK45, 2>), dim3(grid), dim3(threads), lds, strm, a);
            } else if (cpl == kMfmaCoupling) {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_mfma_kernel<Sys, P, MPGOS_RK4, 4>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_mfma_kernel<Sys, P, MPGOS_RKCK45, 4>), dim3(grid), dim3(threads), lds, strm, a);
            } else if (threads <= 256) {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_kernel<Sys, P, MPGOS_RK4, 256>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_kernel<Sys, P, MPGOS_RKCK45, 256>), dim3(grid), dim3(threads), lds, strm, a);
            } else {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_kernel<Sys, P, MPGOS_RK4, kMaxBlock>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_kernel<Sys, P, MPGOS_RKCK45, kMaxBlock>), dim3(grid), dim3(threads), lds, strm, a);
            }
        }
        MPGOS_HIP_TRY(hipGetLastError());
        MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
        return MPGOS_OK;
    }

    // The pair-tiled copy of NC full UPS x UPS matrices at src (column-major),
    // written by a kernel on the solver's stream (allocated on first use; the
    // size never changes for a solver, so no buffer in use is freed).
    int tile_pairs(const double* src, double*& dst, size_t& have)
    {
        const size_t n = (size_t)NC * UPS * UPS;
        if (n > have) {
            if (dst) { MPGOS_HIP_TRY(hipStreamSynchronize(strm)); (void)hipFree(dst); }
            dst = nullptr; have = 0;
            if (hipMalloc((void**)&dst, n * sizeof(double)) != hipSuccess) {
                dst = nullptr;
                set_error("per-block: allocation of the tiled coupling matrix failed");
                return MPGOS_E_MEMORY;
            }
            have = n;
        }
        hipLaunchKernelGGL(tile_pair_matrix<0>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, strm, src, dst, UPS, NC);
        MPGOS_HIP_TRY(hipGetLastError());
        return MPGOS_OK;
    }

    static int grow_device(double*& p, size_t& have, size_t need, const char* what)
    {
        if (need <= have) return MPGOS_OK;
        if (p) (void)hipFree(p);
        p = nullptr; have = 0;
        if (hipMalloc((void**)&p, need * sizeof(double)) != hipSuccess) {
            set_error(std::string("per-block adjoint: allocation of ") + what + " failed");
            return MPGOS_E_MEMORY;
        }
        have = need;
        return MPGOS_OK;
    }

    // Record pass: a Solve() that also keeps the unit states every K accepted
    // steps (at most checkpoint_capacity per system) and (t, h) of every
    // accepted step (at most step_capacity per system).
    int adjoint_record(int K, int ckcap, int stepcap) override
    {
        if constexpr (!kAdjoint) {
            set_error("per-block adjoint: needs FP64 and a system with OdeVJP");
            return MPGOS_E_SYSTEM;
        } else {
            if (K < 1 || ckcap < 1 || stepcap < 1) { set_error("per-block adjoint: interval and capacities must be >= 1"); return MPGOS_E_ARG; }
            MPGOS_HIP_TRY(hipSetDevice(device));
            const size_t NU = (size_t)NS * UPS;
            MPGOS_HIP_TRY(hipStreamSynchronize(strm));
            int rc;
            if ((rc = grow_device(recX, recXn, (size_t)ckcap * UD * NU, "checkpoints")) || (rc = grow_device(recT, recTn, (size_t)stepcap * NS, "step times")) ||
                (rc = grow_device(recH, recHn, (size_t)stepcap * NS, "step sizes")))
                return rc;
            adjK = K; adjCap = ckcap; adjStepCap = stepcap;
            Args<Sys, double> a = make_args();
            a.recX = recX; a.recT = recT; a.recH = recH; a.recK = K; a.recCkCap = ckcap; a.recStepCap = stepcap;
            const int cpl = coupling_form();
            const int spb = cpl == kMfmaCoupling ? kMfmaSystems : SPB;
            a.SPB = spb;
            a.sharedMatrix = sharedMatrix && lds_bytes(cpl, spb, 1) <= ldsMax ? 1 : 0;
            if (cpl == kMfmaCoupling && !a.sharedMatrix && mfma_lds_matrix<Sys, 4>()) a.ldsPairs = lds_pairs(spb);
            const size_t lds = lds_bytes(cpl, spb, a.sharedMatrix, a.ldsPairs);
            if (lds > ldsMax) { set_error("per-block: system scope and coupling terms exceed the LDS of a workgroup"); return MPGOS_E_ARG; }
            const int threads = cpl == kMfmaCoupling ? kMfmaSystems * UPS / 4 : ((SPB * UPS + 63) / 64) * 64;
            const int grid = NS / spb + (NS % spb == 0 ? 0 : 1);
            MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
            if (cpl == kMfmaCoupling && !a.sharedMatrix) {
                if ((rc = tile_pairs((const double*)cMat.d, cTile, cTileN))) return rc;
                a.cTile = cTile;
            }
            if (cpl == kMfmaCoupling) {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_record_mfma_kernel<Sys, MPGOS_RK4>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_record_mfma_kernel<Sys, MPGOS_RKCK45>), dim3(grid), dim3(threads), lds, strm, a);
            } else if (threads <= 256) {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_record_kernel<Sys, MPGOS_RK4, 256>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_record_kernel<Sys, MPGOS_RKCK45, 256>), dim3(grid), dim3(threads), lds, strm, a);
            } else {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_record_kernel<Sys, MPGOS_RK4, kMaxBlock>), dim3(grid), dim3(threads), lds, strm, a);
                else hipLaunchKernelGGL((pb_record_kernel<Sys, MPGOS_RKCK45, kMaxBlock>), dim3(grid), dim3(threads), lds, strm, a);
            }
            MPGOS_HIP_TRY(hipGetLastError());
            MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
            return MPGOS_OK;
        }
    }

    // Reverse pass: AdjointState holds dL/dx(end) on entry (device buffer) and
    // dL/dx(start) on exit; UnitParameterGradient / SystemParameterGradient
    // receive the parameter gradients.  Uses the last record pass.
    int adjoint_reverse() override
    {
        if constexpr (!kAdjoint) {
            set_error("per-block adjoint: needs FP64 and a system with OdeVJP");
            return MPGOS_E_SYSTEM;
        } else {
            if (adjK == 0) { set_error("per-block adjoint: no record pass yet"); return MPGOS_E_STATE; }
            MPGOS_HIP_TRY(hipSetDevice(device));
            const size_t NU = (size_t)NS * UPS;
            MPGOS_HIP_TRY(hipStreamSynchronize(strm));
            const int cpl = coupling_form();
            if (cpl == kMfmaCoupling && !mfma_applies()) {
                set_error("per-block: the dense-contraction coupling kernel needs a full FP64 coupling matrix and UPS a multiple of 32 up to 256");
                return MPGOS_E_ARG;
            }
            const int ST = algorithm == MPGOS_RK4 ? 4 : 6;
            int rc;
            if (cpl == kMfmaCoupling) {
                if ((rc = grow_device(stg, stgn, (size_t)adjK * ST * (UD + NC) * NU, "stage buffer"))) return rc;
            } else if ((rc = grow_device(seg, segn, (size_t)adjK * UD * NU, "segment buffer"))) {
                return rc;
            }
            // transposed coupling matrices in the stored form (from the host mirror)
            std::vector<double> mt((size_t)NC * cMatSize, 0.0);
            for (int i = 0; i < NC; i++)
                for (int r = 0; r < UPS; r++)
                    for (int c = 0; c < UPS; c++) {
                        const long dst = stored_index(i, r, c), src = stored_index(i, c, r);
                        if (dst >= 0 && src >= 0) mt[dst] = (double)cMat.h[src];
                    }
            if ((rc = grow_device(cMatT, cMatTn, (size_t)NC * cMatSize, "transposed coupling matrix"))) return rc;
            MPGOS_HIP_TRY(hipMemcpy(cMatT, mt.data(), mt.size() * sizeof(double), hipMemcpyHostToDevice));
            Args<Sys, double> a = make_args();
            a.recX = recX; a.recT = recT; a.recH = recH; a.recK = adjK; a.recCkCap = adjCap; a.recStepCap = adjStepCap;
            AdjArgs<Sys> A;
            A.seg = seg; A.stg = stg; A.gSp = nullptr; A.lam = lamB.d; A.gU = gUB.d; A.gS = gSB.d; A.cMatT = cMatT;
            A.cTileT = nullptr;
            lastForm = cpl;
            if (cpl == kMfmaCoupling) {
                // one replay and one sweep launch per round; rounds = the largest
                // segment count of a system whose record did not overflow
                a.SPB = kMfmaSystems;
                const size_t lds = reverse_mfma_layout<Sys>(UPS).total;
                if (lds > ldsMax) { set_error("per-block adjoint: reverse-pass LDS exceeds a workgroup's"); return MPGOS_E_ARG; }
                if ((rc = grow_device(gSp, gSpn, (size_t)(NSP > 0 ? NSP : 1) * NU, "system-parameter partials"))) return rc;
                A.gSp = gSp;
                std::vector<long long> nh((size_t)NS);
                std::vector<int> sh((size_t)NS);
                MPGOS_HIP_TRY(hipMemcpy(nh.data(), nacc.d, nh.size() * sizeof(long long), hipMemcpyDeviceToHost));
                MPGOS_HIP_TRY(hipMemcpy(sh.data(), status.d, sh.size() * sizeof(int), hipMemcpyDeviceToHost));
                int rounds = 0;
                for (int i = 0; i < nActive && i < NS; i++)
                    if ((sh[i] & MPGOS_STATUS_CHECKPOINT_OVERFLOW) == 0) {
                        const long long nseg = (nh[i] + adjK - 1) / adjK;
                        if (nseg > rounds) rounds = (int)nseg;
                    }
                const int threads = kMfmaSystems * UPS / kMfmaPairs;
                const int grid = NS / kMfmaSystems + (NS % kMfmaSystems == 0 ? 0 : 1);
                const int fgrid = (int)((NU + 255) / 256);
                MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
                if ((rc = tile_pairs((const double*)cMat.d, cTile, cTileN)) || (rc = tile_pairs(cMatT, cTileT, cTileTN))) return rc;
                a.cTile = cTile;
                A.cTileT = cTileT;
                for (int r = 0; r < rounds; r++) {
                    if (algorithm == MPGOS_RK4) {
                        hipLaunchKernelGGL((pb_replay_mfma_kernel<Sys, MPGOS_RK4, kMfmaPairs>), dim3(grid), dim3(threads), lds, strm, a, A, r);
                        hipLaunchKernelGGL((pb_sweep_mfma_kernel<Sys, MPGOS_RK4, kMfmaPairs>), dim3(grid), dim3(threads), lds, strm, a, A, r);
                    } else {
                        hipLaunchKernelGGL((pb_replay_mfma_kernel<Sys, MPGOS_RKCK45, kMfmaPairs>), dim3(grid), dim3(threads), lds, strm, a, A, r);
                        hipLaunchKernelGGL((pb_sweep_mfma_kernel<Sys, MPGOS_RKCK45, kMfmaPairs>), dim3(grid), dim3(threads), lds, strm, a, A, r);
                    }
                }
                hipLaunchKernelGGL((pb_reverse_finish_kernel<Sys>), dim3(fgrid), dim3(256), 0, strm, a, A);
                MPGOS_HIP_TRY(hipGetLastError());
                MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
                return MPGOS_OK;
            }
            const size_t lds = reverse_layout<Sys>(SPB, UPS).total;
            if (lds > ldsMax) { set_error("per-block adjoint: reverse-pass LDS exceeds a workgroup's"); return MPGOS_E_ARG; }
            const int threads = ((SPB * UPS + 63) / 64) * 64;
            const int grid = NS / SPB + (NS % SPB == 0 ? 0 : 1);
            MPGOS_HIP_TRY(hipEventRecord(evStart, strm));
            if (threads <= 256) {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_reverse_kernel<Sys, MPGOS_RK4, 256>), dim3(grid), dim3(threads), lds, strm, a, A);
                else hipLaunchKernelGGL((pb_reverse_kernel<Sys, MPGOS_RKCK45, 256>), dim3(grid), dim3(threads), lds, strm, a, A);
            } else {
                if (algorithm == MPGOS_RK4) hipLaunchKernelGGL((pb_reverse_kernel<Sys, MPGOS_RK4, kMaxBlock>), dim3(grid), dim3(threads), lds, strm, a, A);
                else hipLaunchKernelGGL((pb_reverse_kernel<Sys, MPGOS_RKCK45, kMaxBlock>), dim3(grid), dim3(threads), lds, strm, a, A);
            }
            MPGOS_HIP_TRY(hipGetLastError());
            MPGOS_HIP_TRY(hipEventRecord(evStop, strm));
            return MPGOS_OK;
        }
    }

    // index of matrix element (row, col) of coupling i in the stored form, -1 outside the band (PB-IF:994-1081)
    long stored_index(int i, int row, int col) const
    {
        if (CCI == 0 && CBW == 0) return (long)row + (long)col * UPS + (long)i * UPS * UPS;
        int mc = col - row + CBW;
        if (mc >= UPS) mc -= UPS;
        if (mc < 0) mc += UPS;
        const int lim = CBW > 0 ? 2 * CBW + 1 : UPS;
        if (mc >= lim) return -1;
        return CCI == 0 ? (long)row + (long)mc * UPS + (long)i * UPS * lim : (long)mc + (long)i * lim;
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
};

} // namespace pb
} // namespace mpgos
