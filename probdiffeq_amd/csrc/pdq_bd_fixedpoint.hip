// Block-diagonal kernel instantiations for strategy_smoother_fixedpoint (nu = 1..5).
#define PDQ_BD_DEFINE_LAUNCH
#include "pdq_bd_launch.h"

namespace pdq {
PDQ_BD_INSTANTIATE(PDQ_FIXEDPOINT)
}  // namespace pdq
