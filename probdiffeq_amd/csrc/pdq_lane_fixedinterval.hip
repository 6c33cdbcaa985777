// Lane-kernel instantiations of the fixedinterval strategy.
#define PDQ_LANE_DEFINE_LAUNCH
#include "pdq_lane_launch.h"

namespace pdq {
PDQ_LANE_INSTANTIATE(PDQ_FIXEDINTERVAL)
}  // namespace pdq
