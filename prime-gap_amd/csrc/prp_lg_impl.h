// prp_lg_impl.h -- batch and scan kernels for the lane-group radix-2^28
// variants (G = 4, 8 or 16 lanes per number) and the strong Lucas kernels;
// included by the prp_lg_*.hip and lucas_lg_*.hip instantiation units.
#pragma once
#include <algorithm>
#include "kernels.h"
#include "lucas_lg.h"
#include "prp_lg.h"

namespace pgap {

// Dynamic LDS of a lane-group launch: the squaring's stage, or the occupancy
// cap's allocation when that is larger.
template <typename K>
size_t lg_lds_bytes(K kernel, int S, int blocks_cap) {
    const size_t bytes = std::max(lg::stage_bytes(S), occupancy_cap_bytes(kernel, blocks_cap));
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return bytes;
}

// This thread's group stage in the block's dynamic LDS.
template <int S, int G>
__device__ __forceinline__ uint32_t *group_stage(uint32_t *lds) {
    return lds + (threadIdx.x / G) * G * lg::stage_stride(S);
}

// G lanes per number: item = thread / G.  A group returns together (its
// lanes share the item), so DPP exchanges never read a finished lane of the
// same group.
template <int S, int G, int KIND>
__device__ __forceinline__ void prp_lg_item(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t count,
                                            const uint32_t *__restrict__ d_count,
                                            const uint32_t *__restrict__ d_idx, Bases bases,
                                            uint8_t *__restrict__ out) {
    extern __shared__ uint32_t lg_lds[];
    uint32_t *stage = group_stage<S, G>(lg_lds);
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t item = t / G, g = t % G;
    const uint32_t cnt = d_count ? min(*d_count, count) : count;
    if (item >= cnt) return;
    const uint32_t row = d_idx ? d_idx[item] : item;
    const uint32_t *nw = n + (size_t)row * limbs;
    uint32_t v;
    if constexpr (KIND == KIND_FERMAT2)
        v = lg::fermat2<S, G>(nw, limbs, g, stage);
    else if constexpr (KIND == KIND_SPRP)
        v = lg::sprp<S, G>(nw, limbs, bases.b[0], g, stage);
    else
        v = lg::strong_lucas<S, G>(nw, limbs, g, stage);
    if (g == 0) out[row] = (uint8_t)v;
}

template <int S, int G, int KIND>
__global__ void __launch_bounds__(256) prp_lg_kernel(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t count,
                                                     const uint32_t *__restrict__ d_count,
                                                     const uint32_t *__restrict__ d_idx, Bases bases,
                                                     uint8_t *__restrict__ out) {
    prp_lg_item<S, G, KIND>(n, limbs, count, d_count, d_idx, bases, out);
}

// Lanes of more than 32 digits (the quads of 4247-bit numbers): held to
// exactly two wavefronts per SIMD, where x, N and the accumulators take 152
// of a wavefront's 256 registers; left to choose, the compiler targets one
// wavefront and moves registers through AGPRs.
template <int S, int G, int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
prp_lg_wide_kernel(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t count,
                   const uint32_t *__restrict__ d_count, const uint32_t *__restrict__ d_idx, Bases bases,
                   uint8_t *__restrict__ out) {
    prp_lg_item<S, G, KIND>(n, limbs, count, d_count, d_idx, bases, out);
}

template <int S, int G, int KIND>
constexpr auto prp_lg_fn() {
    if constexpr (S > 32)
        return &prp_lg_wide_kernel<S, G, KIND>;
    else
        return &prp_lg_kernel<S, G, KIND>;
}

template <int S, int G, int KIND>
hipError_t launch_lg(const uint32_t *d_n, uint32_t limbs, uint32_t count, const uint32_t *d_count,
                     const uint32_t *d_idx, Bases bases, uint8_t *d_out, int blocks_cap, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    if (KIND == KIND_SPRP && bases.n != 1) return hipErrorInvalidValue;
    const dim3 grid((uint32_t)(((size_t)count * G + 255) / 256));
    constexpr auto kernel = prp_lg_fn<S, G, KIND>();
    const size_t lds = lg_lds_bytes(kernel, S, blocks_cap);
    hipLaunchKernelGGL(kernel, grid, dim3(256), lds, stream, d_n, limbs, count, d_count, d_idx, bases, d_out);
    return hipGetLastError();
}

// Streaming-scan round, a group per number: confirmation items padded to
// whole wavefronts (64 / G items) first, then Fermat items.
template <int S, int G>
__device__ __forceinline__ void scan_lg_item(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t conf_rows,
                                             uint32_t conf_max, uint32_t max_items,
                                             const uint32_t *__restrict__ ctr, const uint8_t *__restrict__ base,
                                             uint8_t *__restrict__ out) {
    extern __shared__ uint32_t lg_lds[];
    uint32_t *stage = group_stage<S, G>(lg_lds);
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t item = t / G, g = t % G;
    const uint32_t nc = min(ctr[3], conf_max), nf = min(ctr[0], max_items);
    constexpr uint32_t per_wave = 64 / G;
    const uint32_t ncw = (nc + per_wave - 1) / per_wave * per_wave;
    uint32_t row, v;
    if (item < ncw) {
        if (item >= nc) return;
        row = item;
        v = lg::sprp<S, G>(n + (size_t)row * limbs, limbs, base[item], g, stage);
    } else {
        if (item - ncw >= nf) return;
        row = conf_rows + (item - ncw);
        v = lg::fermat2<S, G>(n + (size_t)row * limbs, limbs, g, stage);
    }
    if (g == 0) out[row] = (uint8_t)v;
}

template <int S, int G>
__global__ void __launch_bounds__(256) scan_lg_kernel(const uint32_t *__restrict__ n, uint32_t limbs,
                                                      uint32_t conf_rows, uint32_t conf_max, uint32_t max_items,
                                                      const uint32_t *__restrict__ ctr,
                                                      const uint8_t *__restrict__ base, uint8_t *__restrict__ out) {
    scan_lg_item<S, G>(n, limbs, conf_rows, conf_max, max_items, ctr, base, out);
}

template <int S, int G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
scan_lg_wide_kernel(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t conf_rows, uint32_t conf_max,
                    uint32_t max_items, const uint32_t *__restrict__ ctr, const uint8_t *__restrict__ base,
                    uint8_t *__restrict__ out) {
    scan_lg_item<S, G>(n, limbs, conf_rows, conf_max, max_items, ctr, base, out);
}

template <int S, int G>
constexpr auto scan_lg_fn() {
    if constexpr (S > 32)
        return &scan_lg_wide_kernel<S, G>;
    else
        return &scan_lg_kernel<S, G>;
}

template <int S, int G>
hipError_t launch_scan_lg(const uint32_t *d_n, uint32_t limbs, uint32_t conf_rows, uint32_t conf_max,
                          uint32_t max_items, const uint32_t *d_ctr, const uint8_t *d_base, uint8_t *d_out,
                          int blocks_cap, hipStream_t stream) {
    constexpr size_t per_wave = 64 / G;
    const size_t items = ((size_t)conf_max + per_wave - 1) / per_wave * per_wave + max_items;
    if (items == 0) return hipSuccess;
    const dim3 grid((uint32_t)((items * G + 255) / 256));
    constexpr auto kernel = scan_lg_fn<S, G>();
    const size_t lds = lg_lds_bytes(kernel, S, blocks_cap);
    hipLaunchKernelGGL(kernel, grid, dim3(256), lds, stream, d_n, limbs, conf_rows, conf_max, max_items, d_ctr,
                       d_base, d_out);
    return hipGetLastError();
}

template <int S, int G>
int occupancy_scan_lg() {
    int n = 0;
    constexpr auto kernel = scan_lg_fn<S, G>();
    const size_t lds = lg_lds_bytes(kernel, S, 0);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, 256, lds) != hipSuccess || n <= 0)
        n = 1;
    return n;
}

template <int S, int G, int KIND>
int occupancy_lg() {
    int n = 0;
    constexpr auto kernel = prp_lg_fn<S, G, KIND>();
    const size_t lds = lg_lds_bytes(kernel, S, 0);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, 256, lds) != hipSuccess || n <= 0)
        n = 1;
    return n;
}

template <int G, int... Ss>
struct RegisterLG {
    RegisterLG() {
        (register_variant(Variant{lg::RB, G * Ss, G, KIND_FERMAT2, lg::RB * G * Ss - 4,
                                  &launch_lg<Ss, G, KIND_FERMAT2>, &occupancy_lg<Ss, G, KIND_FERMAT2>}),
         ...);
        (register_variant(Variant{lg::RB, G * Ss, G, KIND_SPRP, lg::RB * G * Ss - 8, &launch_lg<Ss, G, KIND_SPRP>,
                                  &occupancy_lg<Ss, G, KIND_SPRP>, &launch_scan_lg<Ss, G>,
                                  &occupancy_scan_lg<Ss, G>}),
         ...);
    }
};

// Strong Lucas with one number per wavefront (lucas_lg.h, WAVE): its four
// 16-lane rows compute a step's three Montgomery products at once, so a test
// takes about one product's latency per exponent bit.  LDS: the rows' stages
// (as for 16 lanes per number), then 64 S exchange words per wavefront.
template <int S>
__global__ void __launch_bounds__(256) lucas_wave_kernel(const uint32_t *__restrict__ n, uint32_t limbs,
                                                         uint32_t count, const uint32_t *__restrict__ d_count,
                                                         const uint32_t *__restrict__ d_idx,
                                                         uint8_t *__restrict__ out) {
    extern __shared__ uint32_t lg_lds[];
    uint32_t *stage = group_stage<S, 16>(lg_lds);
    uint32_t *xch = lg_lds + 256 * lg::stage_stride(S) + (threadIdx.x >> 6) * 64 * S;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t item = t >> 6, lane = t & 63u;
    const uint32_t cnt = d_count ? min(*d_count, count) : count;
    if (item >= cnt) return;
    const uint32_t row = d_idx ? d_idx[item] : item;
    const uint32_t v =
        lg::strong_lucas<S, 16, true>(n + (size_t)row * limbs, limbs, lane & 15u, stage, lane >> 4, xch);
    if (lane == 0) out[row] = (uint8_t)v;
}

template <int S>
size_t lucas_wave_lds(int blocks_cap) {
    const size_t bytes = std::max(lg::stage_bytes(S) + (size_t)256 * S * sizeof(uint32_t),
                                  occupancy_cap_bytes(lucas_wave_kernel<S>, blocks_cap));
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)lucas_wave_kernel<S>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
    return bytes;
}

template <int S>
hipError_t launch_lucas_wave(const uint32_t *d_n, uint32_t limbs, uint32_t count, const uint32_t *d_count,
                             const uint32_t *d_idx, Bases, uint8_t *d_out, int blocks_cap, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const dim3 grid((uint32_t)(((size_t)count * 64 + 255) / 256));
    hipLaunchKernelGGL(lucas_wave_kernel<S>, grid, dim3(256), lucas_wave_lds<S>(blocks_cap), stream, d_n, limbs,
                       count, d_count, d_idx, d_out);
    return hipGetLastError();
}

template <int S>
int occupancy_lucas_wave() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lucas_wave_kernel<S>, 256, lucas_wave_lds<S>(0)) !=
            hipSuccess ||
        n <= 0)
        n = 1;
    return n;
}

template <int... Ss>
struct RegisterLucasWave {
    RegisterLucasWave() {
        (register_variant(Variant{lg::RB, 16 * Ss, 64, KIND_LUCAS, lg::RB * 16 * Ss - lg::kLucasHeadroom,
                                  &launch_lucas_wave<Ss>, &occupancy_lucas_wave<Ss>}),
         ...);
    }
};

// Strong Lucas variants (batch launch only; the scan launches them for its
// Baillie-PSW confirmation rows through the same entry point).
template <int G, int... Ss>
struct RegisterLucas {
    RegisterLucas() {
        (register_variant(Variant{lg::RB, G * Ss, G, KIND_LUCAS, lg::RB * G * Ss - lg::kLucasHeadroom,
                                  &launch_lg<Ss, G, KIND_LUCAS>,
                                  &occupancy_lg<Ss, G, KIND_LUCAS>}),
         ...);
    }
};

}  // namespace pgap
