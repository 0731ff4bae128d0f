// Registry of the systems compiled into libmpgos.so.  Each system's kernels
// are instantiated in its own translation unit (sys_<name>.hip) so the build
// parallelises; the C ABI (capi.hip) selects a factory by name.
#pragma once

#include "pb_engine.h"
#include "pt_engine.h"

#define MPGOS_PT_FACTORY(NAME, SYS)                                                   \
    ::mpgos::pt::SolverBase* make_pt_##NAME(int precision)                            \
    {                                                                                 \
        switch (precision) {                                                          \
        case MPGOS_DOUBLE: return new ::mpgos::pt::Solver<SYS, double, false>();      \
        case MPGOS_FLOAT: return new ::mpgos::pt::Solver<SYS, float, false>();        \
        case MPGOS_FLOAT_KAHAN: return new ::mpgos::pt::Solver<SYS, float, true>();   \
        default: return nullptr;                                                      \
        }                                                                             \
    }

#define MPGOS_PB_FACTORY(NAME, SYS)                                                   \
    ::mpgos::pb::SolverBase* make_pb_##NAME(int precision)                            \
    {                                                                                 \
        switch (precision) {                                                          \
        case MPGOS_DOUBLE: return new ::mpgos::pb::Solver<SYS, double>();             \
        case MPGOS_FLOAT: return new ::mpgos::pb::Solver<SYS, float>();               \
        case MPGOS_FLOAT_KAHAN: return new ::mpgos::pb::Solver<SYS, float, true>();   \
        default: return nullptr;                                                      \
        }                                                                             \
    }

namespace mpgos {
namespace pb {
struct Registration {
    const char* name;
    SolverBase* (*make)(int precision);
};
SolverBase* make_pb_ring_duffing(int precision);
SolverBase* make_pb_ring_dw(int precision);
SolverBase* make_pb_pb_linear(int precision);
SolverBase* make_pb_ring_duffing_q(int precision);
} // namespace pb

namespace pt {

struct Registration {
    const char* name;
    SolverBase* (*make)(int precision);
};

SolverBase* make_pt_duffing(int precision);
SolverBase* make_pt_duffing_q(int precision);
SolverBase* make_pt_duffing_eps(int precision);
SolverBase* make_pt_duffing_dw(int precision);
SolverBase* make_pt_relief_valve(int precision);
SolverBase* make_pt_keller_miksis(int precision);
SolverBase* make_pt_keller_miksis_max(int precision);
SolverBase* make_pt_lorenz(int precision);
SolverBase* make_pt_affine2(int precision);

} // namespace pt
} // namespace mpgos
