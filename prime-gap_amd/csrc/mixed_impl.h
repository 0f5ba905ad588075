// mixed_impl.h -- numbers of several widths in one persistent, work-queue launch.
//
// Reference counterpart: the reference compiles one container width per
// binary (BITS, gap_test_gpu.cu:48-55, suggested ladder :669-686), so a
// mixed workload either runs in the widest container or in one run per
// width.  Here the numbers are grouped into size buckets, each bucket is cut
// into tiles of 64 numbers (one per lane), tiles are queued widest first
// (longest-processing-time order), and a persistent grid of waves pulls
// tiles from an atomic counter; each wave runs the radix-2^28 variant of its
// tile's bucket.  Every wave leaves when the queue is empty.
#pragma once
#include <algorithm>
#include "kernels.h"
#include "prp28.h"

namespace pgap {

// Buckets of the persistent kernel (digits of radix 2^28): up to 332, 724,
// 1116, 1564 and 2180-bit numbers for Fermat base 2.
template <int... Ls>
struct Buckets {
    static constexpr int n = sizeof...(Ls);
    static constexpr int digits[n] = {Ls...};
};
using MixedBuckets = Buckets<12, 26, 40, 56, 78>;

// LDS park of the widest bucket (the last), one column per thread.
constexpr size_t kMixedParkBytes =
    (size_t)MixedBuckets::digits[MixedBuckets::n - 1] * r28::PARK_STRIDE * sizeof(uint32_t);

template <int KIND, int L>
__device__ __forceinline__ uint32_t test_one(const uint32_t *nw, uint32_t limbs, const Bases &bases,
                                             uint32_t *park) {
    // the persistent kernel runs one wavefront per SIMD (its widest bucket
    // sets the registers): three chains per column in the wide squaring
    if (KIND == KIND_FERMAT2) return r28::fermat2<L, 3>(nw, limbs, park);
    uint32_t v = 1;
    for (uint32_t k = 0; k < bases.n && v; ++k) v = r28::sprp<L, 3>(nw, limbs, bases.b[k], park);
    return v;
}

template <int KIND, int... Ls>
__device__ __forceinline__ uint32_t dispatch(int bucket, const uint32_t *nw, uint32_t limbs, const Bases &bases,
                                             uint32_t *park, Buckets<Ls...>) {
    uint32_t v = 0;
    int i = 0;
    ((bucket == i++ ? (void)(v = test_one<KIND, Ls>(nw, limbs, bases, park)) : (void)0), ...);
    return v;
}

template <int KIND>
__global__ void __launch_bounds__(256) mixed_kernel(const uint32_t *__restrict__ words,
                                                    const uint64_t *__restrict__ offset,
                                                    const uint32_t *__restrict__ limbs,
                                                    const uint32_t *__restrict__ order,
                                                    const uint8_t *__restrict__ tile_bucket, uint32_t ntiles,
                                                    uint32_t *__restrict__ next_tile, Bases bases,
                                                    uint8_t *__restrict__ out) {
    extern __shared__ uint32_t mixed_park[];
    uint32_t *park = mixed_park + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(next_tile, 1u);
        t = (uint32_t)__shfl((int)t, 0, 64);
        if (t >= ntiles) break;
        const int bucket = tile_bucket[t];
        const uint32_t item = order[(size_t)t * 64 + lane];
        if (item != 0xffffffffu) {
            const uint32_t v = dispatch<KIND>(bucket, words + offset[item], limbs[item], bases, park, MixedBuckets{});
            out[item] = (uint8_t)v;
        }
    }
}

// Launches the persistent kernel of one test kind (defined in mixed_fermat.hip
// and mixed_sprp.hip so the two compile in parallel).
struct MixedLaunch {
    const uint32_t *words;
    const uint64_t *offset;
    const uint32_t *limbs;
    const uint32_t *order;
    const uint8_t *tile_bucket;
    uint32_t ntiles;
    uint32_t *next_tile;
    Bases bases;
    uint8_t *out;
};
hipError_t launch_mixed_fermat(const MixedLaunch &a, int device, hipStream_t stream);
hipError_t launch_mixed_sprp(const MixedLaunch &a, int device, hipStream_t stream);

template <int KIND>
hipError_t launch_mixed(const MixedLaunch &a, int device, hipStream_t stream) {
    int per_cu = 0, cus = 0;
    (void)hipFuncSetAttribute((const void *)mixed_kernel<KIND>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kMixedParkBytes);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mixed_kernel<KIND>, 256, kMixedParkBytes) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 256;
    const uint32_t grid = (uint32_t)std::min<size_t>((size_t)per_cu * cus, ((size_t)a.ntiles + 3) / 4);
    hipLaunchKernelGGL(mixed_kernel<KIND>, dim3(std::max(grid, 1u)), dim3(256), kMixedParkBytes, stream, a.words, a.offset, a.limbs,
                       a.order, a.tile_bucket, a.ntiles, a.next_tile, a.bases, a.out);
    return hipGetLastError();
}

}  // namespace pgap
