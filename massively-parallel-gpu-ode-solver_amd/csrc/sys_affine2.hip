// Kernel instantiations of the registered system "affine2" (all precisions,
// algorithms, kernel forms and register-budget variants).
#include "registry.h"
#include "systems/affine2.h"

namespace mpgos {
namespace pt {
MPGOS_PT_FACTORY(affine2, ::mpgos::systems::Affine2)
} // namespace pt
} // namespace mpgos
