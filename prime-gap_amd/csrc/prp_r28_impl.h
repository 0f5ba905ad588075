// prp_r28_impl.h -- batch kernels for the one-number-per-lane radix-2^28
// variants; included by the prp_r28_*.hip instantiation units.
#pragma once
#include "kernels.h"
#include "prp28.h"

namespace pgap {

// One lane per number; 256-thread blocks (4 waves).  Input words are read
// once per number plus one exponent word per 32 squarings, so the layout
// (row-major per number) does not bound the kernel: it is VALU-bound.
template <int L, int KIND>
__global__ void __launch_bounds__(256) prp_r28_kernel(const uint32_t *__restrict__ n, uint32_t limbs,
                                                      uint32_t count, const uint32_t *__restrict__ d_count,
                                                      const uint32_t *__restrict__ d_idx, Bases bases,
                                                      uint8_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t cnt = d_count ? min(*d_count, count) : count;
    if (i >= cnt) return;
    const uint32_t row = d_idx ? d_idx[i] : i;
    const uint32_t *nw = n + (size_t)row * limbs;
    uint32_t v;
    if (KIND == KIND_FERMAT2) {
        v = r28::fermat2<L>(nw, limbs);
    } else {
        v = 1;
        for (uint32_t k = 0; k < bases.n && v; ++k) v = r28::sprp<L>(nw, limbs, bases.b[k]);
    }
    out[row] = (uint8_t)v;
}

template <int L, int KIND>
hipError_t launch_r28(const uint32_t *d_n, uint32_t limbs, uint32_t count, const uint32_t *d_count,
                      const uint32_t *d_idx, Bases bases, uint8_t *d_out, int /*waves_per_simd*/,
                      hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const dim3 grid((count + 255u) / 256u);
    hipLaunchKernelGGL((prp_r28_kernel<L, KIND>), grid, dim3(256), 0, stream, d_n, limbs, count, d_count, d_idx,
                       bases, d_out);
    return hipGetLastError();
}

template <int L, int KIND>
int occupancy_r28() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, prp_r28_kernel<L, KIND>, 256, 0) != hipSuccess || n <= 0)
        n = 1;
    return n;
}

template <int... Ls>
struct RegisterR28 {
    RegisterR28() {
        (register_variant(Variant{28, Ls, 1, KIND_FERMAT2, 28 * Ls - 4, &launch_r28<Ls, KIND_FERMAT2>,
                                  &occupancy_r28<Ls, KIND_FERMAT2>}),
         ...);
        (register_variant(
             Variant{28, Ls, 1, KIND_SPRP, 28 * Ls - 14, &launch_r28<Ls, KIND_SPRP>, &occupancy_r28<Ls, KIND_SPRP>}),
         ...);
    }
};

}  // namespace pgap
