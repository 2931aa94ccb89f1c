// Particle-mesh repulsion (placeholder until the PM kernels land; the session refuses PM).
#include "fa2_kernels.h"

namespace gr {
void pm_repulsion(FaDev &d) { (void)d; }
}  // namespace gr
