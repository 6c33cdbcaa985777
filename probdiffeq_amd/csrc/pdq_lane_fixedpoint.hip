// Lane-kernel instantiations of the fixedpoint strategy.
#define PDQ_LANE_DEFINE_LAUNCH
#include "pdq_lane_launch.h"

namespace pdq {
PDQ_LANE_INSTANTIATE(PDQ_FIXEDPOINT)
}  // namespace pdq
