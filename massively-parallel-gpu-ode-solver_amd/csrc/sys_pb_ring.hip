// Kernel instantiations of the registered per-block systems.
#include "registry.h"
#include "systems/pb_ring.h"

namespace mpgos {
namespace pb {
MPGOS_PB_FACTORY(ring_duffing, ::mpgos::systems::RingDuffing)
MPGOS_PB_FACTORY(ring_dw, ::mpgos::systems::RingDW)
MPGOS_PB_FACTORY(pb_linear, ::mpgos::systems::PBLinear)
MPGOS_PB_FACTORY(ring_duffing_q, ::mpgos::systems::RingDuffingQ)
} // namespace pb
} // namespace mpgos
