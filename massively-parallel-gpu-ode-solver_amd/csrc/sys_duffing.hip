// Kernel instantiations of the registered system "duffing" (all precisions,
// algorithms, kernel forms and register-budget variants).
#include "registry.h"
#include "systems/duffing.h"

namespace mpgos {
namespace pt {
MPGOS_PT_FACTORY(duffing, ::mpgos::systems::Duffing)
} // namespace pt
} // namespace mpgos
