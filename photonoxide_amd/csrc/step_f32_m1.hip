// The step kernels in f32, coefficient mode 1.
#include "step_launch_impl.hpp"

namespace pxfdtd {
template int launch_mode<float, 1>(const StepArgs<float> &, int, int, hipStream_t, size_t, const SideStream *,
                                  const UnitCounts &, int);
template int launch_pair<float, 1>(const StepArgs<float> &, const PairTaps<float> &, int, hipStream_t, size_t,
                                  const PairStreams *, const PairCounts &);
}  // namespace pxfdtd
