// Cooperative kernel instantiated for the iso5 shape (see pdq_shapes.h).
#include "pdq_coop.h"

namespace pdq {

int launch_coop_iso5(KArgs& a, int G, hipStream_t st) {
  return launch_fixed_shape<FixDims<5, 2, 2, 1, 5, 2, 0>>(a, G, st);
}

}  // namespace pdq
