// Instantiation unit: persistent mixed-width kernel, lane-group buckets, strong PRP.
#include "mixed_impl.h"
namespace pgap {
hipError_t launch_mixed_lg_sprp(const MixedLaunch &a, int device, hipStream_t stream) {
    return launch_mixed_lg<KIND_SPRP>(a, device, stream);
}
}  // namespace pgap
