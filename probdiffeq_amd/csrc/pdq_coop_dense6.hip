// Cooperative kernel instantiated for the dense6 shape (see pdq_shapes.h).
#include "pdq_coop.h"

namespace pdq {

int launch_coop_dense6(KArgs& a, int G, hipStream_t st) {
  return launch_fixed_shape<FixDims<12, 1, 1, 2, 6, 2, 1>>(a, G, st);
}

}  // namespace pdq
