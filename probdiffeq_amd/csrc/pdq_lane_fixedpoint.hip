// Lane-kernel instantiations of the fixedpoint strategy, with the error
// estimator's and the calibration's options at their defaults.
#define PDQ_LANE_DEFINE_LAUNCH
#include "pdq_lane_launch.h"

namespace pdq {
PDQ_LANE_INSTANTIATE(PDQ_FIXEDPOINT, false)
}  // namespace pdq
