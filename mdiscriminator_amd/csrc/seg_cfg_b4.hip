// seg_cfg_b4.hip -- every segmented candidate configuration for 4 class
// bits (at most 16 classes), both discriminator kinds.
#include "segmented_launch.hpp"

namespace mdiscr {
template bool launch_nbits<4, true>(int, const SegArgs&, hipStream_t);
template bool launch_nbits<4, false>(int, const SegArgs&, hipStream_t);
}  // namespace mdiscr
