// Kernel instantiations of the registered system "lorenz" (all precisions,
// algorithms, kernel forms and register-budget variants).
#include "registry.h"
#include "systems/lorenz.h"

namespace mpgos {
namespace pt {
MPGOS_PT_FACTORY(lorenz, ::mpgos::systems::Lorenz)
} // namespace pt
} // namespace mpgos
