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

// One test of one number.  A strong-PRP launch tests one base, bases.b[0]
// (the host runs later bases on the survivors), so no loop keeps a test live
// across bases.  The kernel is scheduled for two wavefronts per SIMD, the
// batch kernels' choice for the widest bucket, so the wide squaring uses one
// chain per column.
template <int KIND, int L>
__device__ __forceinline__ uint32_t test_one(const uint32_t *nw, uint32_t limbs, const Bases &bases,
                                             uint32_t *park) {
    if (KIND == KIND_FERMAT2) return r28::fermat2<L, 1>(nw, limbs, park);
    return r28::sprp<L, 1>(nw, limbs, bases.b[0], park);
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
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) mixed_kernel(const uint32_t *__restrict__ words,
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

// The kernel is scheduled for two wavefronts per SIMD, which pays when every
// SIMD works through several tiles (2^19 numbers, 8192 tiles: 160.8 ms against
// 182.4 ms for the one-wave build).  With few tiles the dispatcher may place
// the first (widest) tiles two to a SIMD while other SIMDs idle (163840
// numbers, 2560 tiles: 86.5 ms against 53.1 ms), so below two tiles per
// wave slot at two waves the launch reserves enough LDS to hold one workgroup
// (one wave per SIMD) per CU.
constexpr uint32_t kMixedTilesPerCuTwoWaves = 16;
constexpr size_t kMixedOneWaveLds = 96 * 1024;  // > 160 KiB / 2: one workgroup per CU

template <int KIND>
hipError_t launch_mixed(const MixedLaunch &a, int device, hipStream_t stream) {
    int per_cu = 0, cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        cus = 256;
    const bool two_waves = a.ntiles >= (uint32_t)cus * kMixedTilesPerCuTwoWaves;
    const size_t lds = two_waves ? kMixedParkBytes : std::max(kMixedParkBytes, kMixedOneWaveLds);
    (void)hipFuncSetAttribute((const void *)mixed_kernel<KIND>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mixed_kernel<KIND>, 256, lds) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    const uint32_t grid = (uint32_t)std::min<size_t>((size_t)per_cu * cus, ((size_t)a.ntiles + 3) / 4);
    hipLaunchKernelGGL(mixed_kernel<KIND>, dim3(std::max(grid, 1u)), dim3(256), lds, stream, a.words, a.offset,
                       a.limbs, a.order, a.tile_bucket, a.ntiles, a.next_tile, a.bases, a.out);
    return hipGetLastError();
}

}  // namespace pgap
