// mixed_impl.h -- numbers of several widths in one persistent, work-queue launch.
//
// Reference counterpart: the reference compiles one container width per
// binary (BITS, gap_test_gpu.cu:48-55, suggested ladder :669-686), so a
// mixed workload either runs in the widest container or in one run per
// width.  Here the numbers are grouped into size buckets, each bucket is cut
// into tiles of one wavefront (64 numbers of a one-lane bucket, 64 / G of a
// bucket with G lanes per number), tiles are queued widest first
// (longest-processing-time order), and a persistent grid of waves pulls
// tiles from an atomic counter; each wave runs the radix-2^28 variant of its
// tile's bucket.  Every wave leaves when the queue is empty.
#pragma once
#include <algorithm>
#include "kernels.h"
#include "prp28.h"
#include "prp_lg.h"

namespace pgap {

// A bucket: one number per lane with L digits (prp28.h), or one number per
// group of G lanes with S digits per lane (prp_lg.h).
template <int L>
struct OneLane {
    static constexpr int lanes = 1, digits = L;
};
template <int G, int S>
struct LaneGroup {
    static constexpr int lanes = G, digits = G * S, stage_digits = S;
};
template <class... Bs>
struct Buckets {
    static constexpr int n = sizeof...(Bs);
    static constexpr int digits[n] = {Bs::digits...};
    static constexpr int lanes[n] = {Bs::lanes...};
};
// Digits of radix 2^28.  One lane per number up to 332, 724, 1116, 1564 and
// 2180-bit numbers (Fermat base 2); groups of 8 lanes up to 2684, 3356, 3580,
// 4252 and 4476 bits.  The two kinds run as two persistent kernels: the
// one-lane buckets need up to 256 registers (two wavefronts per SIMD), the
// lane groups about 160 (three).
using MixedOneLane = Buckets<OneLane<12>, OneLane<26>, OneLane<40>, OneLane<56>, OneLane<78>>;
using MixedGroups = Buckets<LaneGroup<8, 12>, LaneGroup<8, 15>, LaneGroup<8, 16>, LaneGroup<8, 19>, LaneGroup<8, 20>>;

// LDS park of the widest one-lane bucket (the last), one column per thread;
// the lane-group kernel's stage for its widest bucket.
constexpr size_t kMixedParkBytes =
    (size_t)MixedOneLane::digits[MixedOneLane::n - 1] * r28::PARK_STRIDE * sizeof(uint32_t);
constexpr int kMixedStageStride = lg::stage_stride(20);
constexpr size_t kMixedStageBytes = lg::stage_bytes(20);

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

// One tile of a bucket in this wave: a one-lane bucket tests order[64 t +
// lane]; a G-lane bucket tests order[64 t + lane / G] with the group's lanes.
template <int KIND, int L>
__device__ __forceinline__ void run_tile(OneLane<L>, const uint32_t *words, const uint64_t *offset,
                                         const uint32_t *limbs, const uint32_t *order, uint32_t t, uint32_t lane,
                                         const Bases &bases, uint32_t *lds, uint8_t *out) {
    const uint32_t item = order[(size_t)t * 64 + lane];
    if (item == 0xffffffffu) return;
    out[item] = (uint8_t)test_one<KIND, L>(words + offset[item], limbs[item], bases, lds + threadIdx.x);
}
template <int KIND, int G, int S>
__device__ __forceinline__ void run_tile(LaneGroup<G, S>, const uint32_t *words, const uint64_t *offset,
                                         const uint32_t *limbs, const uint32_t *order, uint32_t t, uint32_t lane,
                                         const Bases &bases, uint32_t *lds, uint8_t *out) {
    static_assert(lg::stage_stride(S) <= kMixedStageStride && 64 % G == 0, "a wave's stage slice");
    const uint32_t item = order[(size_t)t * 64 + lane / G], g = lane % G;
    if (item == 0xffffffffu) return;  // the whole group
    // each wave owns a fixed slice of the stage, sized for the widest bucket:
    // the waves of a workgroup may run buckets of different S at once
    uint32_t *stage = lds + (threadIdx.x / 64) * 64 * kMixedStageStride + (lane / G) * G * lg::stage_stride(S);
    const uint32_t *nw = words + offset[item];
    uint32_t v;
    if constexpr (KIND == KIND_FERMAT2)
        v = lg::fermat2<S, G>(nw, limbs[item], g, stage);
    else
        v = lg::sprp<S, G>(nw, limbs[item], bases.b[0], g, stage);
    if (g == 0) out[item] = (uint8_t)v;
}

template <int KIND, class... Bs>
__device__ __forceinline__ void dispatch(int bucket, const uint32_t *words, const uint64_t *offset,
                                         const uint32_t *limbs, const uint32_t *order, uint32_t t, uint32_t lane,
                                         const Bases &bases, uint32_t *lds, uint8_t *out, Buckets<Bs...>) {
    int i = 0;
    ((bucket == i++ ? run_tile<KIND>(Bs{}, words, offset, limbs, order, t, lane, bases, lds, out) : (void)0), ...);
}

// Pull tiles until the queue is empty.
template <int KIND, class BL>
__device__ __forceinline__ void mixed_loop(const uint32_t *words, const uint64_t *offset, const uint32_t *limbs,
                                           const uint32_t *order, const uint8_t *tile_bucket, uint32_t ntiles,
                                           uint32_t *next_tile, const Bases &bases, uint32_t *lds, uint8_t *out) {
    const uint32_t lane = threadIdx.x & 63;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(next_tile, 1u);
        t = (uint32_t)__shfl((int)t, 0, 64);
        if (t >= ntiles) break;
        dispatch<KIND>(tile_bucket[t], words, offset, limbs, order, t, lane, bases, lds, out, BL{});
    }
}

template <int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) mixed_kernel(const uint32_t *__restrict__ words,
                                                    const uint64_t *__restrict__ offset,
                                                    const uint32_t *__restrict__ limbs,
                                                    const uint32_t *__restrict__ order,
                                                    const uint8_t *__restrict__ tile_bucket, uint32_t ntiles,
                                                    uint32_t *__restrict__ next_tile, Bases bases,
                                                    uint8_t *__restrict__ out) {
    extern __shared__ uint32_t mixed_lds[];
    mixed_loop<KIND, MixedOneLane>(words, offset, limbs, order, tile_bucket, ntiles, next_tile, bases, mixed_lds,
                                   out);
}

template <int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) mixed_lg_kernel(const uint32_t *__restrict__ words,
                                                       const uint64_t *__restrict__ offset,
                                                       const uint32_t *__restrict__ limbs,
                                                       const uint32_t *__restrict__ order,
                                                       const uint8_t *__restrict__ tile_bucket, uint32_t ntiles,
                                                       uint32_t *__restrict__ next_tile, Bases bases,
                                                       uint8_t *__restrict__ out) {
    extern __shared__ uint32_t mixed_lds[];
    mixed_loop<KIND, MixedGroups>(words, offset, limbs, order, tile_bucket, ntiles, next_tile, bases, mixed_lds,
                                  out);
}

// Launches a persistent kernel of one test kind: the one-lane buckets
// (mixed_fermat.hip, mixed_sprp.hip) or the lane-group buckets
// (mixed_lg_fermat.hip, mixed_lg_sprp.hip), four units that compile in
// parallel.
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
hipError_t launch_mixed_lg_fermat(const MixedLaunch &a, int device, hipStream_t stream);
hipError_t launch_mixed_lg_sprp(const MixedLaunch &a, int device, hipStream_t stream);

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

// The lane-group kernel: as many waves as its registers allow, no more
// waves than tiles.
template <int KIND>
hipError_t launch_mixed_lg(const MixedLaunch &a, int device, hipStream_t stream) {
    int per_cu = 0, cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mixed_lg_kernel<KIND>, 256, kMixedStageBytes) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    const uint32_t grid = (uint32_t)std::min<size_t>((size_t)per_cu * cus, ((size_t)a.ntiles + 3) / 4);
    hipLaunchKernelGGL(mixed_lg_kernel<KIND>, dim3(std::max(grid, 1u)), dim3(256), kMixedStageBytes, stream,
                       a.words, a.offset, a.limbs, a.order, a.tile_bucket, a.ntiles, a.next_tile, a.bases, a.out);
    return hipGetLastError();
}

}  // namespace pgap
