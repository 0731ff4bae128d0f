// ORACLE -- test infrastructure only (see pt_oracle.h).  C entry points for
// the Python tests and bench.py's cpu_baseline leg (ctypes).
#include <omp.h>

#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "dual.h"
#include "flops.h"
#include "pb_oracle.h"
#include "pb_sens.h"
#include "pt_oracle.h"
#include "systems/affine2.h"
#include "systems/duffing.h"
#include "systems/duffing_dw.h"
#include "systems/relief_valve.h"
#include "systems/keller_miksis.h"
#include "systems/lorenz.h"
#include "systems/pb_ring.h"

namespace oracle {
namespace flops {
thread_local long long g_count = 0;
F cos(F a) { g_count += 38; return F(mpgos::math::cos(a.v)); }
F sin(F a) { g_count += 38; return F(mpgos::math::sin(a.v)); }
F exp(F a) { g_count += 31; return F(mpgos::math::exp(a.v)); }
F log(F a) { g_count += 28; return F(mpgos::math::log(a.v)); }
F pow(F a, F b) { g_count += 70; return F(mpgos::math::pow(a.v, b.v)); }
F sqrt(F a) { g_count += 1; return F(mpgos::math::sqrt(a.v)); }
F ctrl_root(F a, bool fifth) { g_count += fifth ? 44 : 2; return F(mpgos::math::ctrl_root(a.v, fifth)); }
} // namespace flops

template <class Sys, class S>
static void run_all(const Job& j)
{
    std::vector<S> sp(Sys::NSP > 0 ? Sys::NSP : 1);
    std::vector<int> spi(Sys::NISP > 0 ? Sys::NISP : 1);
    for (int i = 0; i < Sys::NSP; i++) sp[i] = S(j.sPAR[i]);
    for (int i = 0; i < Sys::NISP; i++) spi[i] = j.sPARi[i];
    const int threads = j.nthreads > 0 ? j.nthreads : omp_get_max_threads();
    // chunks of up to 64 instances, small enough that every thread gets work
    int chunk = j.n_active / (threads * 8);
    chunk = chunk < 1 ? 1 : (chunk > 64 ? 64 : chunk);
#pragma omp parallel for schedule(dynamic, chunk) num_threads(threads)
    for (int tid = 0; tid < j.n_active; tid++) {
        std::vector<S> spl(sp);   // per-thread copy: callbacks may not write it, but keep lanes independent
        std::vector<int> spil(spi);
        Instance<Sys, S> inst(j, tid, spl.data(), spil.data());
        inst.run();
    }
}

template <class Sys>
static void trace_one(const Job& j, int tid, typename Instance<Sys, double>::Trace* tr)
{
    std::vector<double> sp(Sys::NSP > 0 ? Sys::NSP : 1);
    std::vector<int> spi(Sys::NISP > 0 ? Sys::NISP : 1);
    for (int i = 0; i < Sys::NSP; i++) sp[i] = j.sPAR[i];
    for (int i = 0; i < Sys::NISP; i++) spi[i] = j.sPARi[i];
    Instance<Sys, double> inst(j, tid, sp.data(), spi.data());
    inst.trace = tr;
    inst.run();
}

// Tangent-linear derivatives of the final state with respect to the initial
// state and the control parameters, on the frozen step sequence:
// jac[(i * (SD + NCP) + j) * NT + tid] = d x_i(T) / d z_j, z = (x(0), cPAR).
template <class Sys>
static void sensitivity_all(const Job& j, double* jac)
{
    constexpr int N = Sys::SD + Sys::NCP;
    using S = dual::D<N>;
    std::vector<S> sp(Sys::NSP > 0 ? Sys::NSP : 1);
    std::vector<int> spi(Sys::NISP > 0 ? Sys::NISP : 1);
    for (int i = 0; i < Sys::NSP; i++) sp[i] = S(j.sPAR[i]);
    for (int i = 0; i < Sys::NISP; i++) spi[i] = j.sPARi[i];
    const int threads = j.nthreads > 0 ? j.nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int tid = 0; tid < j.n_active; tid++) {
        std::vector<S> spl(sp);
        std::vector<int> spil(spi);
        Instance<Sys, S> inst(j, tid, spl.data(), spil.data());
        inst.run();
        for (int a = 0; a < Sys::SD; a++)
            for (int b = 0; b < N; b++) jac[((size_t)a * N + b) * j.nt + tid] = inst.x[a].d[b];
    }
}

typedef void (*Runner)(const Job&);
typedef void (*Sensitivity)(const Job&, double*);
struct TraceArgs { int cap; int n; double* t; double* h; int* acc; double* ratio; };
typedef void (*Tracer)(const Job&, int, TraceArgs*);
template <class Sys>
static void trace_entry(const Job& j, int tid, TraceArgs* a)
{
    static_assert(sizeof(TraceArgs) == sizeof(typename Instance<Sys, double>::Trace), "trace layout");
    trace_one<Sys>(j, tid, reinterpret_cast<typename Instance<Sys, double>::Trace*>(a));
}
struct Entry { const char* name; Runner run; Runner count; Tracer trace; Sensitivity sens; };

static const Entry kSystems[] = {
    {"duffing", run_all<mpgos::systems::Duffing, double>, run_all<mpgos::systems::Duffing, flops::F>, trace_entry<mpgos::systems::Duffing>, sensitivity_all<mpgos::systems::Duffing>},
    {"duffing_q", run_all<mpgos::systems::DuffingQ, double>, run_all<mpgos::systems::DuffingQ, flops::F>, trace_entry<mpgos::systems::DuffingQ>, sensitivity_all<mpgos::systems::DuffingQ>},
    {"duffing_eps", run_all<mpgos::systems::DuffingEps, double>, run_all<mpgos::systems::DuffingEps, flops::F>, trace_entry<mpgos::systems::DuffingEps>, sensitivity_all<mpgos::systems::DuffingEps>},
    {"duffing_dw", run_all<mpgos::systems::DuffingDW, double>, run_all<mpgos::systems::DuffingDW, flops::F>, trace_entry<mpgos::systems::DuffingDW>, sensitivity_all<mpgos::systems::DuffingDW>},
    {"relief_valve", run_all<mpgos::systems::ReliefValve, double>, run_all<mpgos::systems::ReliefValve, flops::F>, trace_entry<mpgos::systems::ReliefValve>, sensitivity_all<mpgos::systems::ReliefValve>},
    {"keller_miksis", run_all<mpgos::systems::KellerMiksis, double>, run_all<mpgos::systems::KellerMiksis, flops::F>, trace_entry<mpgos::systems::KellerMiksis>, sensitivity_all<mpgos::systems::KellerMiksis>},
    {"keller_miksis_max", run_all<mpgos::systems::KellerMiksisMax, double>, run_all<mpgos::systems::KellerMiksisMax, flops::F>, trace_entry<mpgos::systems::KellerMiksisMax>, sensitivity_all<mpgos::systems::KellerMiksisMax>},
    {"lorenz", run_all<mpgos::systems::Lorenz, double>, run_all<mpgos::systems::Lorenz, flops::F>, trace_entry<mpgos::systems::Lorenz>, sensitivity_all<mpgos::systems::Lorenz>},
    {"affine2", run_all<mpgos::systems::Affine2, double>, run_all<mpgos::systems::Affine2, flops::F>, trace_entry<mpgos::systems::Affine2>, sensitivity_all<mpgos::systems::Affine2>},
};

static const Entry* find(const char* name)
{
    for (const auto& e : kSystems)
        if (std::strcmp(e.name, name) == 0) return &e;
    return nullptr;
}

} // namespace oracle

namespace oracle {
// Per-block coupled systems (pb_oracle.h), OpenMP over systems.
template <class Sys>
static void run_pb(const oracle::PBJob& j)
{
    const int threads = j.nthreads > 0 ? j.nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int s = 0; s < j.n_active; s++) {
        oracle::PBSystem<Sys> sys(j, s);
        sys.run();
    }
}

} // namespace oracle

extern "C" {

int oracle_pt_solve(const char* system, oracle::Job* job, double* seconds)
{
    const oracle::Entry* e = oracle::find(system);
    if (!e || !job) return -4;
    const auto t0 = std::chrono::steady_clock::now();
    e->run(*job);
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

// Algorithmic FLOPs of the same integration (single thread).
int oracle_pt_count_flops(const char* system, oracle::Job* job, long long* flops)
{
    const oracle::Entry* e = oracle::find(system);
    if (!e || !job || !flops) return -4;
    oracle::Job j = *job;
    j.nthreads = 1;
    oracle::flops::g_count = 0;
    e->count(j);
    *flops = oracle::flops::g_count;
    return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

int oracle_pb_solve(const char* system, oracle::PBJob* job, double* seconds)
{
    if (!system || !job) return -4;
    void (*fn)(const oracle::PBJob&) = nullptr;
    if (std::strcmp(system, "ring_duffing") == 0) fn = oracle::run_pb<mpgos::systems::RingDuffing>;
    else if (std::strcmp(system, "ring_dw") == 0) fn = oracle::run_pb<mpgos::systems::RingDW>;
    else if (std::strcmp(system, "pb_linear") == 0) fn = oracle::run_pb<mpgos::systems::PBLinear>;
    else if (std::strcmp(system, "ring_duffing_q") == 0) fn = oracle::run_pb<mpgos::systems::RingDuffingQ>;
    if (!fn) return -4;
    const auto t0 = std::chrono::steady_clock::now();
    fn(*job);
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

// Frozen-step forward sensitivities of every system (pb_sens.h):
// jac[((s * N + i) * UD + c) * UPS + u], N = NSP + UD + NUP of the system.
int oracle_pb_sensitivity(const char* system, oracle::PBJob* job, double* jac, int* ndir)
{
    if (!system || !job || !jac) return -4;
    int n = -1;
    auto go = [&](auto tag) {
        using Sys = decltype(tag);
        using S = oracle::PBSens<Sys>;
        n = S::N;
        const size_t per = (size_t)S::N * Sys::UD * job->ups;
#pragma omp parallel for schedule(dynamic, 1) num_threads(job->nthreads > 0 ? job->nthreads : omp_get_max_threads())
        for (int sid = 0; sid < job->n_active; sid++) S::run(*job, sid, jac + per * sid);
    };
    if (std::strcmp(system, "ring_duffing_q") == 0) go(mpgos::systems::RingDuffingQ{});
    else if (std::strcmp(system, "pb_linear") == 0) go(mpgos::systems::PBLinear{});
    else return -4;
    if (ndir) *ndir = n;
    return 0;
}

int oracle_pt_sensitivity(const char* system, oracle::Job* job, double* jac)
{
    const oracle::Entry* e = oracle::find(system);
    if (!e || !job || !jac) return -4;
    e->sens(*job, jac);
    return 0;
}

// Per-attempt trace of instance `tid` (time before the step, step size,
// accepted flag, tol/err ratio of the RKCK45 controller).  Returns the number
// of recorded attempts in *n.
int oracle_pt_trace(const char* system, oracle::Job* job, int tid, int cap, double* t, double* h, int* acc,
                    double* ratio, int* n)
{
    const oracle::Entry* e = oracle::find(system);
    if (!e || !job || tid < 0 || tid >= job->nt) return -4;
    oracle::TraceArgs a{cap, 0, t, h, acc, ratio};
    e->trace(*job, tid, &a);
    *n = a.n;
    return 0;
}

} // extern "C"
