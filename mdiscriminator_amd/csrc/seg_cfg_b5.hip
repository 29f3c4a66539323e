// seg_cfg_b5.hip -- every segmented candidate configuration for 5 class
// bits (at most 32 classes), both discriminator kinds.
#include "segmented_launch.hpp"

namespace mdiscr {
template bool launch_nbits<5, true>(int, const SegArgs&, hipStream_t);
template bool launch_nbits<5, false>(int, const SegArgs&, hipStream_t);
}  // namespace mdiscr
