// Lane-kernel instantiations of the filter strategy.
#define PDQ_LANE_DEFINE_LAUNCH
#include "pdq_lane_launch.h"

namespace pdq {
PDQ_LANE_INSTANTIATE(PDQ_FILTER)
}  // namespace pdq
