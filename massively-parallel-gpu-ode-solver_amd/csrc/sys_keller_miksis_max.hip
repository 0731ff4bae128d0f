// Kernel instantiations of the registered system "keller_miksis_max" (all
// precisions, algorithms, kernel forms and register-budget variants).
#include "registry.h"
#include "systems/keller_miksis.h"

namespace mpgos {
namespace pt {
MPGOS_PT_FACTORY(keller_miksis_max, ::mpgos::systems::KellerMiksisMax)
} // namespace pt
} // namespace mpgos
