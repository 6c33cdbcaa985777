// Dispatch of the lane-per-instance isotropic kernel to its instantiations
// (compiled in pdq_lane_<strategy>.hip with the option defaults folded in,
// and in pdq_lane_<strategy>_opts.hip with the options read at run time).
#include "pdq_lane_launch.h"

namespace pdq {

bool lane_supported(int N, int D) { return D == 2 && (N == 4 || N == 5); }

namespace {

template <int N, int D, int S, bool O>
int launch_lane_o(LArgs& a, hipStream_t st) {
  switch (a.cfg.solver) {
    case PDQ_SOLVER: return launch_lane_nd<N, D, S, PDQ_SOLVER, O>(a, st);
    case PDQ_SOLVER_DYNAMIC: return launch_lane_nd<N, D, S, PDQ_SOLVER_DYNAMIC, O>(a, st);
    case PDQ_SOLVER_MLE: return launch_lane_nd<N, D, S, PDQ_SOLVER_MLE, O>(a, st);
    default: return host_error(PDQ_ERR_INVALID, "unknown solver %d", a.cfg.solver);
  }
}

template <int N, int D, int S>
int launch_lane_s(LArgs& a, hipStream_t st) {
  return lane_default_options(a.cfg) ? launch_lane_o<N, D, S, false>(a, st) : launch_lane_o<N, D, S, true>(a, st);
}

template <int N, int D>
int launch_lane_shape(LArgs& a, hipStream_t st) {
  switch (a.cfg.strategy) {
    case PDQ_FILTER: return launch_lane_s<N, D, PDQ_FILTER>(a, st);
    case PDQ_FIXEDPOINT: return launch_lane_s<N, D, PDQ_FIXEDPOINT>(a, st);
    case PDQ_FIXEDINTERVAL: return launch_lane_s<N, D, PDQ_FIXEDINTERVAL>(a, st);
    default: return host_error(PDQ_ERR_INVALID, "unknown strategy %d", a.cfg.strategy);
  }
}

}  // namespace

int launch_lane(LArgs& a, int N, int D, hipStream_t st) {
  if (D == 2) {
    if (N == 4) return launch_lane_shape<4, 2>(a, st);
    if (N == 5) return launch_lane_shape<5, 2>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "no lane kernel for nu + 1 = %d, d = %d", N, D);
}

}  // namespace pdq
