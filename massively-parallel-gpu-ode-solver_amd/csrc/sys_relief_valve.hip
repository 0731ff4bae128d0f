// Kernel instantiations of the registered system "relief_valve" (all precisions,
// algorithms, kernel forms and register-budget variants).
#include "registry.h"
#include "systems/relief_valve.h"

namespace mpgos {
namespace pt {
MPGOS_PT_FACTORY(relief_valve, ::mpgos::systems::ReliefValve)
} // namespace pt
} // namespace mpgos
