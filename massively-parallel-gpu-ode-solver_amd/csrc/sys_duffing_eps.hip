// Kernel instantiations of the registered system "duffing_eps" (all precisions,
// algorithms, kernel forms and register-budget variants).
#include "registry.h"
#include "systems/duffing.h"

namespace mpgos {
namespace pt {
MPGOS_PT_FACTORY(duffing_eps, ::mpgos::systems::DuffingEps)
} // namespace pt
} // namespace mpgos
