// The step kernels in f64, coefficient mode 0.
#include "step_launch_impl.hpp"

namespace pxfdtd {
template int launch_mode<double, 0>(const StepArgs<double> &, int, int, hipStream_t, size_t, const SideStream *,
                                  const UnitCounts &, int);
template int launch_pair<double, 0>(const StepArgs<double> &, const PairTaps<double> &, int, hipStream_t, size_t,
                                  const PairStreams *, const PairCounts &);
}  // namespace pxfdtd
