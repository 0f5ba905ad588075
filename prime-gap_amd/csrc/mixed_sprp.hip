// Instantiation unit: persistent mixed-width kernel, one-lane buckets, strong PRP.
#include "mixed_impl.h"
namespace pgap {
hipError_t launch_mixed_sprp(const MixedLaunch &a, int device, hipStream_t stream) {
    return launch_mixed<KIND_SPRP>(a, device, stream);
}
}  // namespace pgap
