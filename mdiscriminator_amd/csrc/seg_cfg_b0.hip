// seg_cfg_b0.hip -- every segmented candidate configuration for 0 class
// bits (at most 1 classes), both discriminator kinds.
#include "segmented_launch.hpp"

namespace mdiscr {
template bool launch_nbits<0, true>(int, const SegArgs&, hipStream_t);
template bool launch_nbits<0, false>(int, const SegArgs&, hipStream_t);
}  // namespace mdiscr
