// extern "C" entry points of libmpgos.so (declared in include/mpgos.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "mpgos.h"
#include "mpgos_math.h"
#include "registry.h"

namespace mpgos {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

namespace pt {
static const Registration kSystems[] = {
    {"duffing", make_pt_duffing},
    {"duffing_q", make_pt_duffing_q},
    {"duffing_eps", make_pt_duffing_eps},
    {"duffing_dw", make_pt_duffing_dw},
    {"relief_valve", make_pt_relief_valve},
    {"keller_miksis", make_pt_keller_miksis},
    {"keller_miksis_max", make_pt_keller_miksis_max},
    {"lorenz", make_pt_lorenz},
    {"affine2", make_pt_affine2},
};
} // namespace pt

namespace pb {
static const Registration kSystems[] = {
    {"ring_duffing", make_pb_ring_duffing},
    {"ring_duffing_q", make_pb_ring_duffing_q},
    {"ring_dw", make_pb_ring_dw},
    {"pb_linear", make_pb_pb_linear},
};
} // namespace pb

// Elementary-function probe kernel: same source as the solver's functions.
__global__ void math_eval_kernel(int fn, const double* x, const double* y, double* out, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    switch (fn) {
    case 0: out[i] = math::cos(x[i]); break;
    case 1: out[i] = math::sin(x[i]); break;
    case 2: out[i] = math::exp(x[i]); break;
    case 3: out[i] = math::log(x[i]); break;
    case 5: out[i] = math::sqrt(x[i]); break;
    case 6: out[i] = math::pow_ctrl(x[i], y[i]); break;
    case 7: out[i] = math::ctrl_root(x[i], y[i] != 0.0); break;
    default: out[i] = math::pow(x[i], y[i]); break;
    }
}

// Scratch memory of mpgos_pt_solve_noalloc: the relaunch form's lane records,
// phases, instance lists and counts.  One copy per GPU, held by the loaded code
// object until the process ends.  Its host-side name is not a device address:
// host code uses only the address hipGetSymbolAddress returned for the device.
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
std::mutex g_scratch_mutex;
std::map<int, void*> g_scratch_addr;        // device -> address of `scratch` on that device
std::map<int, std::string> g_scratch_error; // device -> why the lookup failed

// Once per device, with that device current (called when a solver is created).
void scratch_setup(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_addr.count(device) || g_scratch_error.count(device)) return;
    hipError_t e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_scratch_error[device] = std::string("scratch memory lookup failed: ") + hipGetErrorString(e);
        return;
    }
    g_scratch_addr[device] = p;
}

void* scratch_address(int device, std::string& why)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    auto er = g_scratch_error.find(device);
    why = er != g_scratch_error.end() ? er->second : "scratch memory not set up for this device";
    return nullptr;
}
} // namespace

} // namespace mpgos

struct mpgos_solver {
    mpgos::pt::SolverBase* pt = nullptr;   // per-thread module
    mpgos::pb::SolverBase* pb = nullptr;   // per-block module
};

using mpgos::set_error;

extern "C" {

const char* mpgos_last_error(void) { return mpgos::g_last_error.c_str(); }

const char* mpgos_pt_list_systems(void)
{
    static std::string names;
    if (names.empty()) {
        for (const auto& r : mpgos::pt::kSystems) {
            if (!names.empty()) names += ",";
            names += r.name;
        }
    }
    return names.c_str();
}

int mpgos_pt_create(const char* system, int algorithm, int precision, int nt, int ndo, int device, mpgos_solver** out)
{
    if (!out || !system) { set_error("mpgos_pt_create: null argument"); return MPGOS_E_ARG; }
    *out = nullptr;
    for (const auto& r : mpgos::pt::kSystems) {
        if (std::strcmp(r.name, system) != 0) continue;
        mpgos::pt::SolverBase* s = r.make(precision);
        if (!s) { set_error("mpgos_pt_create: unsupported precision"); return MPGOS_E_SYSTEM; }
        int rc = s->init(algorithm, nt, ndo, device);
        if (rc != MPGOS_OK) { delete s; return rc; }
        mpgos::scratch_setup(device);
        mpgos_solver* h = new mpgos_solver;
        h->pt = s;
        *out = h;
        return MPGOS_OK;
    }
    set_error(std::string("mpgos_pt_create: unknown system '") + system + "' (registered: " + mpgos_pt_list_systems() + ")");
    return MPGOS_E_SYSTEM;
}

#define MPGOS_CHECK_HANDLE(h)                                     \
    do {                                                          \
        if (!(h) || !(h)->pt) { set_error("null solver handle"); return MPGOS_E_ARG; } \
    } while (0)

int mpgos_pt_destroy(mpgos_solver* h)
{
    MPGOS_CHECK_HANDLE(h);
    delete h->pt;
    delete h;
    return MPGOS_OK;
}

int mpgos_pt_get_dims(const mpgos_solver* h, mpgos_pt_dims* out)
{
    MPGOS_CHECK_HANDLE(h);
    if (!out) { set_error("null dims pointer"); return MPGOS_E_ARG; }
    return h->pt->dims(out);
}

int mpgos_pt_solver_option(mpgos_solver* h, int option, int index, double value)
{
    MPGOS_CHECK_HANDLE(h);
    return h->pt->option(option, index, value);
}

int mpgos_pt_set_host(mpgos_solver* h, int variable, int problem, int serial, int step, double value)
{
    MPGOS_CHECK_HANDLE(h);
    return h->pt->set_host(variable, problem, serial, step, value);
}

int mpgos_pt_get_host(const mpgos_solver* h, int variable, int problem, int serial, int step, double* value)
{
    MPGOS_CHECK_HANDLE(h);
    if (!value) { set_error("null value pointer"); return MPGOS_E_ARG; }
    return h->pt->get_host(variable, problem, serial, step, value);
}

void* mpgos_pt_host_buffer(mpgos_solver* h, int variable, size_t* nbytes)
{
    if (!h || !h->pt) { set_error("null solver handle"); return nullptr; }
    return h->pt->host_buffer(variable, nbytes);
}

void* mpgos_pt_device_buffer(mpgos_solver* h, int variable, size_t* nbytes)
{
    if (!h || !h->pt) { set_error("null solver handle"); return nullptr; }
    return h->pt->device_buffer(variable, nbytes);
}

int mpgos_pt_sync_h2d(mpgos_solver* h, int variable) { MPGOS_CHECK_HANDLE(h); return h->pt->sync(variable, true); }
int mpgos_pt_sync_d2h(mpgos_solver* h, int variable) { MPGOS_CHECK_HANDLE(h); return h->pt->sync(variable, false); }
int mpgos_pt_solve(mpgos_solver* h) { MPGOS_CHECK_HANDLE(h); return h->pt->solve(); }
int mpgos_pt_solve_slice(mpgos_solver* h, long long max_attempted_steps, int* unfinished)
{
    MPGOS_CHECK_HANDLE(h);
    if (!unfinished) { set_error("mpgos_pt_solve_slice: null pointer"); return MPGOS_E_ARG; }
    return h->pt->solve_slice(max_attempted_steps, unfinished);
}
int mpgos_pt_solve_noalloc(mpgos_solver* h)
{
    MPGOS_CHECK_HANDLE(h);
    std::string why;
    void* p = mpgos::scratch_address(h->pt->device_index(), why);
    if (!p) { set_error("mpgos_pt_solve_noalloc: " + why); return MPGOS_E_STATE; }
    return h->pt->solve_noalloc(p, mpgos::kNoallocScratchBytes);
}
int mpgos_pt_synchronise_device(mpgos_solver* h) { MPGOS_CHECK_HANDLE(h); return h->pt->synchronise_device(); }
int mpgos_pt_insert_synchronisation_point(mpgos_solver* h) { MPGOS_CHECK_HANDLE(h); return h->pt->insert_sync_point(); }
int mpgos_pt_synchronise_solver(mpgos_solver* h) { MPGOS_CHECK_HANDLE(h); return h->pt->synchronise_solver(); }
int mpgos_pt_print(mpgos_solver* h, int variable, int thread_id, const char* directory)
{
    MPGOS_CHECK_HANDLE(h);
    return h->pt->print(variable, thread_id, directory);
}
int mpgos_pt_last_solve_ms(const mpgos_solver* h, double* ms)
{
    MPGOS_CHECK_HANDLE(h);
    if (!ms) { set_error("null pointer"); return MPGOS_E_ARG; }
    return h->pt->last_solve_ms(ms);
}
void* mpgos_pt_stream(mpgos_solver* h)
{
    if (!h || !h->pt) { set_error("null solver handle"); return nullptr; }
    return h->pt->stream();
}

int mpgos_pt_adjoint_record(mpgos_solver* h, int checkpoint_interval, int checkpoint_capacity)
{
    MPGOS_CHECK_HANDLE(h);
    return h->pt->adjoint_record(checkpoint_interval, checkpoint_capacity);
}

int mpgos_pt_adjoint_reverse(mpgos_solver* h)
{
    MPGOS_CHECK_HANDLE(h);
    return h->pt->adjoint_reverse();
}

// ---------------------------------------------------------------- per-block
#define MPGOS_CHECK_PB(h)                                                                         \
    do {                                                                                          \
        if (!(h) || !(h)->pb) { set_error("null or non-per-block solver handle"); return MPGOS_E_ARG; } \
    } while (0)

const char* mpgos_pb_list_systems(void)
{
    static std::string names;
    if (names.empty()) {
        for (const auto& r : mpgos::pb::kSystems) {
            if (!names.empty()) names += ",";
            names += r.name;
        }
    }
    return names.c_str();
}

int mpgos_pb_create(const char* system, int algorithm, int precision, int ns, int ups, int spb, int cbw, int cci,
                    int ndo, int device, mpgos_solver** out)
{
    if (!out || !system) { set_error("mpgos_pb_create: null argument"); return MPGOS_E_ARG; }
    *out = nullptr;
    for (const auto& r : mpgos::pb::kSystems) {
        if (std::strcmp(r.name, system) != 0) continue;
        mpgos::pb::SolverBase* s = r.make(precision);
        if (!s) { set_error("mpgos_pb_create: unsupported precision"); return MPGOS_E_SYSTEM; }
        int rc = s->init(algorithm, ns, ups, spb, cbw, cci, ndo, device);
        if (rc != MPGOS_OK) { delete s; return rc; }
        mpgos_solver* h = new mpgos_solver;
        h->pb = s;
        *out = h;
        return MPGOS_OK;
    }
    set_error(std::string("mpgos_pb_create: unknown system '") + system + "' (registered: " + mpgos_pb_list_systems() + ")");
    return MPGOS_E_SYSTEM;
}

int mpgos_pb_destroy(mpgos_solver* h) { MPGOS_CHECK_PB(h); delete h->pb; delete h; return MPGOS_OK; }
int mpgos_pb_get_dims(const mpgos_solver* h, mpgos_pb_dims* out)
{
    MPGOS_CHECK_PB(h);
    if (!out) { set_error("null dims pointer"); return MPGOS_E_ARG; }
    return h->pb->dims(out);
}
int mpgos_pb_solver_option(mpgos_solver* h, int option, int index, double value) { MPGOS_CHECK_PB(h); return h->pb->option(option, index, value); }
int mpgos_pb_set_host(mpgos_solver* h, int variable, int a0, int a1, int a2, int a3, double value)
{
    MPGOS_CHECK_PB(h);
    return h->pb->set_host(variable, a0, a1, a2, a3, value);
}
int mpgos_pb_get_host(const mpgos_solver* h, int variable, int a0, int a1, int a2, int a3, double* value)
{
    MPGOS_CHECK_PB(h);
    if (!value) { set_error("null value pointer"); return MPGOS_E_ARG; }
    return h->pb->get_host(variable, a0, a1, a2, a3, value);
}
void* mpgos_pb_host_buffer(mpgos_solver* h, int variable, size_t* nbytes)
{
    if (!h || !h->pb) { set_error("null or non-per-block solver handle"); return nullptr; }
    return h->pb->buffer(variable, true, nbytes);
}
void* mpgos_pb_device_buffer(mpgos_solver* h, int variable, size_t* nbytes)
{
    if (!h || !h->pb) { set_error("null or non-per-block solver handle"); return nullptr; }
    return h->pb->buffer(variable, false, nbytes);
}
int mpgos_pb_sync_h2d(mpgos_solver* h, int variable) { MPGOS_CHECK_PB(h); return h->pb->sync(variable, true); }
int mpgos_pb_sync_d2h(mpgos_solver* h, int variable) { MPGOS_CHECK_PB(h); return h->pb->sync(variable, false); }
int mpgos_pb_solve(mpgos_solver* h) { MPGOS_CHECK_PB(h); return h->pb->solve(); }
int mpgos_pb_synchronise_device(mpgos_solver* h) { MPGOS_CHECK_PB(h); return h->pb->synchronise_device(); }
int mpgos_pb_insert_synchronisation_point(mpgos_solver* h) { MPGOS_CHECK_PB(h); return h->pb->insert_sync_point(); }
int mpgos_pb_synchronise_solver(mpgos_solver* h) { MPGOS_CHECK_PB(h); return h->pb->synchronise_solver(); }
int mpgos_pb_last_solve_ms(const mpgos_solver* h, double* ms)
{
    MPGOS_CHECK_PB(h);
    if (!ms) { set_error("null pointer"); return MPGOS_E_ARG; }
    return h->pb->last_solve_ms(ms);
}
int mpgos_pb_adjoint_record(mpgos_solver* h, int checkpoint_interval, int checkpoint_capacity, int step_capacity)
{
    MPGOS_CHECK_PB(h);
    return h->pb->adjoint_record(checkpoint_interval, checkpoint_capacity, step_capacity);
}
int mpgos_pb_adjoint_reverse(mpgos_solver* h) { MPGOS_CHECK_PB(h); return h->pb->adjoint_reverse(); }
void* mpgos_pb_stream(mpgos_solver* h)
{
    if (!h || !h->pb) { set_error("null or non-per-block solver handle"); return nullptr; }
    return h->pb->stream();
}

int mpgos_math_eval_host(int fn, const double* x, const double* y, double* out, long n)
{
    if (fn < 0 || fn > 7 || !x || !out || ((fn == 4 || fn >= 6) && !y)) { set_error("mpgos_math_eval_host: bad arguments"); return MPGOS_E_ARG; }
    for (long i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = mpgos::math::cos(x[i]); break;
        case 1: out[i] = mpgos::math::sin(x[i]); break;
        case 2: out[i] = mpgos::math::exp(x[i]); break;
        case 3: out[i] = mpgos::math::log(x[i]); break;
        case 5: out[i] = mpgos::math::sqrt(x[i]); break;
        case 6: out[i] = mpgos::math::pow_ctrl(x[i], y[i]); break;
        case 7: out[i] = mpgos::math::ctrl_root(x[i], y[i] != 0.0); break;
        default: out[i] = mpgos::math::pow(x[i], y[i]); break;
        }
    }
    return MPGOS_OK;
}

int mpgos_math_eval_device(int fn, const double* x_dev, const double* y_dev, double* out_dev, long n)
{
    if (fn < 0 || fn > 7 || !x_dev || !out_dev || ((fn == 4 || fn >= 6) && !y_dev) || n < 0) {
        set_error("mpgos_math_eval_device: bad arguments");
        return MPGOS_E_ARG;
    }
    if (n == 0) return MPGOS_OK;
    const int block = 256;
    const long grid = (n + block - 1) / block;
    hipLaunchKernelGGL(mpgos::math_eval_kernel, dim3((unsigned)grid), dim3(block), 0, 0, fn, x_dev, y_dev, out_dev, n);
    MPGOS_HIP_TRY(hipGetLastError());
    MPGOS_HIP_TRY(hipDeviceSynchronize());
    return MPGOS_OK;
}

} // extern "C"
