// seg_cfg_b6.hip -- every segmented candidate configuration for 6 class
// bits (at most 64 classes), both discriminator kinds.
#include "segmented_launch.hpp"

namespace mdiscr {
template bool launch_nbits<6, true>(int, const SegArgs&, hipStream_t);
template bool launch_nbits<6, false>(int, const SegArgs&, hipStream_t);
}  // namespace mdiscr
