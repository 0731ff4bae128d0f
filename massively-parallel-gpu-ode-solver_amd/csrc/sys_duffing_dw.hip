// Kernel instantiations of the registered system "duffing_dw" (all precisions,
// algorithms, kernel forms and register-budget variants).
#include "registry.h"
#include "systems/duffing_dw.h"

namespace mpgos {
namespace pt {
MPGOS_PT_FACTORY(duffing_dw, ::mpgos::systems::DuffingDW)
} // namespace pt
} // namespace mpgos
