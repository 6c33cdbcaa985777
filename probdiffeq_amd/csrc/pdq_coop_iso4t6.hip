// Cooperative kernel instantiated for the iso4t6 shape (see pdq_shapes.h).
#include "pdq_coop.h"

namespace pdq {

int launch_coop_iso4t6(KArgs& a, int G, hipStream_t st) {
  return launch_fixed_shape<FixDims<4, 14, 2, 1, 4, 2, 0>>(a, G, st);
}

}  // namespace pdq
