// Instantiation unit: persistent mixed-width kernel, lane-group buckets, Fermat base 2.
#include "mixed_impl.h"
namespace pgap {
hipError_t launch_mixed_lg_fermat(const MixedLaunch &a, int device, hipStream_t stream) {
    return launch_mixed_lg<KIND_FERMAT2>(a, device, stream);
}
}  // namespace pgap
