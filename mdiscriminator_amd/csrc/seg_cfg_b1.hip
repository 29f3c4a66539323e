// seg_cfg_b1.hip -- every segmented candidate configuration for 1 class
// bits (at most 2 classes), both discriminator kinds.
#include "segmented_launch.hpp"

namespace mdiscr {
template bool launch_nbits<1, true>(int, const SegArgs&, hipStream_t);
template bool launch_nbits<1, false>(int, const SegArgs&, hipStream_t);
}  // namespace mdiscr
