// prp_r28_impl.h -- batch kernels for the one-number-per-lane radix-2^28
// variants; included by the prp_r28_*.hip instantiation units.
#pragma once
#include <algorithm>
#include "kernels.h"
#include "prp28.h"

namespace pgap {

// Waves per SIMD the compiler is asked to schedule for.  Wide variants
// (L > 62): exactly two for batch kernels, whose parked exponentiation then
// fits 256 registers or nearly (78 digits, 2^18 Fermat tests: 166.7 ms), and
// exactly one for the streaming-scan kernel, whose rounds are sized to one
// resident wave per SIMD so that a round costs one test's latency (169.0 ms
// for the same batch; left to choose, the compiler's schedule took 194 ms).
// The narrow ones: the compiler's own choice (up to 8).
constexpr int r28_wmin(int L, bool scan) { return L > 62 ? (scan ? 1 : 2) : 1; }
constexpr int r28_wmax(int L, bool scan) { return L > 62 ? (scan ? 1 : 2) : 8; }

// Chains per column of the wide squaring in the one-wave scan kernels.
constexpr int kScanChains = 3;

// LDS park bytes of an r28 kernel (one column of L words per thread).
constexpr size_t r28_park_bytes(int L, bool parks) {
    return parks ? (size_t)L * r28::PARK_STRIDE * sizeof(uint32_t) : 0;
}

// Dynamic LDS of an r28 launch: the park, or the occupancy cap's allocation
// when that is larger.
template <typename K>
size_t r28_lds_bytes(K kernel, size_t park, int blocks_cap) {
    const size_t bytes = std::max(park, occupancy_cap_bytes(kernel, blocks_cap));
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return bytes;
}

// One lane per number; 256-thread blocks (4 waves).  Input words are read
// once per number plus one exponent word per 32 squarings, so the layout
// (row-major per number) does not bound the kernel: it is VALU-bound.
template <int L, int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(r28_wmin(L, false), r28_wmax(L, false))))
prp_r28_kernel(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t count, const uint32_t *__restrict__ d_count,
               const uint32_t *__restrict__ d_idx, Bases bases, uint8_t *__restrict__ out) {
    extern __shared__ uint32_t r28_park[];
    uint32_t *park = r28_park + threadIdx.x;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t cnt = d_count ? min(*d_count, count) : count;
    if (i >= cnt) return;
    const uint32_t row = d_idx ? d_idx[i] : i;
    const uint32_t *nw = n + (size_t)row * limbs;
    uint32_t v;
    if (KIND == KIND_FERMAT2)
        v = r28::fermat2<L>(nw, limbs, park);
    else
        v = r28::sprp<L>(nw, limbs, bases.b[0], park);
    out[row] = (uint8_t)v;
}

template <int L, int KIND>
hipError_t launch_r28(const uint32_t *d_n, uint32_t limbs, uint32_t count, const uint32_t *d_count,
                      const uint32_t *d_idx, Bases bases, uint8_t *d_out, int blocks_cap, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    if (KIND == KIND_SPRP && bases.n != 1) return hipErrorInvalidValue;
    const dim3 grid((count + 255u) / 256u);
    const size_t lds = r28_lds_bytes(prp_r28_kernel<L, KIND>,
                                     r28_park_bytes(L, KIND != KIND_FERMAT2 || r28::fermat_parks(L)), blocks_cap);
    hipLaunchKernelGGL((prp_r28_kernel<L, KIND>), grid, dim3(256), lds, stream, d_n, limbs, count, d_count, d_idx,
                       bases, d_out);
    return hipGetLastError();
}

// Streaming-scan round: confirmation items (one strong-PRP base each) in
// whole wavefronts first, then Fermat items.
template <int L>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(r28_wmin(L, true), r28_wmax(L, true))))
scan_r28_kernel(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t conf_rows, uint32_t conf_max,
                uint32_t max_items, const uint32_t *__restrict__ ctr, const uint8_t *__restrict__ base,
                uint8_t *__restrict__ out) {
    extern __shared__ uint32_t r28_park[];
    uint32_t *park = r28_park + threadIdx.x;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t nc = min(ctr[3], conf_max), nf = min(ctr[0], max_items);
    const uint32_t ncw = (nc + 63u) & ~63u;
    uint32_t row, v;
    if (t < ncw) {
        if (t >= nc) return;
        row = t;
        v = r28::sprp<L, kScanChains>(n + (size_t)row * limbs, limbs, base[t], park);
    } else {
        if (t - ncw >= nf) return;
        row = conf_rows + (t - ncw);
        v = r28::fermat2<L, kScanChains>(n + (size_t)row * limbs, limbs, park);
    }
    out[row] = (uint8_t)v;
}

template <int L>
hipError_t launch_scan_r28(const uint32_t *d_n, uint32_t limbs, uint32_t conf_rows, uint32_t conf_max,
                           uint32_t max_items, const uint32_t *d_ctr, const uint8_t *d_base, uint8_t *d_out,
                           int blocks_cap, hipStream_t stream) {
    const size_t threads = (((size_t)conf_max + 63) & ~(size_t)63) + max_items;
    if (threads == 0) return hipSuccess;
    const dim3 grid((uint32_t)((threads + 255) / 256));
    const size_t lds = r28_lds_bytes(scan_r28_kernel<L>, r28_park_bytes(L, true), blocks_cap);
    hipLaunchKernelGGL((scan_r28_kernel<L>), grid, dim3(256), lds, stream, d_n, limbs, conf_rows, conf_max,
                       max_items, d_ctr, d_base, d_out);
    return hipGetLastError();
}

template <int L>
int occupancy_scan_r28() {
    int n = 0;
    const size_t park = r28_lds_bytes(scan_r28_kernel<L>, r28_park_bytes(L, true), 0);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, scan_r28_kernel<L>, 256, park) != hipSuccess || n <= 0)
        n = 1;
    return n;
}

template <int L, int KIND>
int occupancy_r28() {
    int n = 0;
    const size_t park = r28_lds_bytes(prp_r28_kernel<L, KIND>,
                                      r28_park_bytes(L, KIND != KIND_FERMAT2 || r28::fermat_parks(L)), 0);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, prp_r28_kernel<L, KIND>, 256, park) != hipSuccess || n <= 0)
        n = 1;
    return n;
}

template <int... Ls>
struct RegisterR28 {
    RegisterR28() {
        (register_variant(Variant{28, Ls, 1, KIND_FERMAT2, 28 * Ls - 4, &launch_r28<Ls, KIND_FERMAT2>,
                                  &occupancy_r28<Ls, KIND_FERMAT2>}),
         ...);
        (register_variant(Variant{28, Ls, 1, KIND_SPRP, 28 * Ls - 8, &launch_r28<Ls, KIND_SPRP>,
                                  &occupancy_r28<Ls, KIND_SPRP>, &launch_scan_r28<Ls>, &occupancy_scan_r28<Ls>}),
         ...);
    }
};

}  // namespace pgap
