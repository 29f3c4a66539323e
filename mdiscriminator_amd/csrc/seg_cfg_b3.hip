// seg_cfg_b3.hip -- every segmented candidate configuration for 3 class
// bits (at most 8 classes), both discriminator kinds.
#include "segmented_launch.hpp"

namespace mdiscr {
template bool launch_nbits<3, true>(int, const SegArgs&, hipStream_t);
template bool launch_nbits<3, false>(int, const SegArgs&, hipStream_t);
}  // namespace mdiscr
