// Cooperative kernel instantiated for the dense5 shape (see pdq_shapes.h).
#include "pdq_coop.h"

namespace pdq {

int launch_coop_dense5(KArgs& a, int G, hipStream_t st) {
  return launch_fixed_shape<FixDims<10, 1, 1, 2, 5, 2, 1>>(a, G, st);
}

}  // namespace pdq
