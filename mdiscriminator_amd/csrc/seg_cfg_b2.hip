// seg_cfg_b2.hip -- every segmented candidate configuration for 2 class
// bits (at most 4 classes), both discriminator kinds.
#include "segmented_launch.hpp"

namespace mdiscr {
template bool launch_nbits<2, true>(int, const SegArgs&, hipStream_t);
template bool launch_nbits<2, false>(int, const SegArgs&, hipStream_t);
}  // namespace mdiscr
