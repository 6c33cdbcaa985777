// Dispatch of the block-diagonal kernel to its instantiations (compiled in
// pdq_bd_<strategy>.hip).
#include "pdq_bd_launch.h"

namespace pdq {

namespace {

template <int N>
int launch_bd_n(BdArgs& a, hipStream_t st) {
  switch (a.cfg.strategy) {
    case PDQ_FILTER: return launch_bd_ns<N, PDQ_FILTER>(a, st);
    case PDQ_FIXEDPOINT: return launch_bd_ns<N, PDQ_FIXEDPOINT>(a, st);
    case PDQ_FIXEDINTERVAL: return launch_bd_ns<N, PDQ_FIXEDINTERVAL>(a, st);
    default: return host_error(PDQ_ERR_INVALID, "unknown strategy %d", a.cfg.strategy);
  }
}

}  // namespace

int launch_bd(BdArgs& a, hipStream_t st) {
  switch (a.cfg.num_derivatives + 1) {
    case 2: return launch_bd_n<2>(a, st);
    case 3: return launch_bd_n<3>(a, st);
    case 4: return launch_bd_n<4>(a, st);
    case 5: return launch_bd_n<5>(a, st);
    case 6: return launch_bd_n<6>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "blockdiag supports num_derivatives in [1, 5] (got %d)", a.cfg.num_derivatives);
}

}  // namespace pdq
