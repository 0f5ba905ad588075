// prp_lg_impl.h -- batch kernels for the quad-per-number radix-2^26 variants;
// included by the prp_lg_*.hip instantiation units.
#pragma once
#include "kernels.h"
#include "prp_lg.h"

namespace pgap {

// Four lanes per number: item = thread / 4.  A quad returns together (its
// lanes share the item), so DPP exchanges never read a finished lane.
template <int S, int KIND>
__global__ void __launch_bounds__(256) prp_lg_kernel(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t count,
                                                     const uint32_t *__restrict__ d_count,
                                                     const uint32_t *__restrict__ d_idx, Bases bases,
                                                     uint8_t *__restrict__ out) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t item = t >> 2, g = t & 3u;
    const uint32_t cnt = d_count ? min(*d_count, count) : count;
    if (item >= cnt) return;
    const uint32_t row = d_idx ? d_idx[item] : item;
    const uint32_t *nw = n + (size_t)row * limbs;
    uint32_t v;
    if (KIND == KIND_FERMAT2) {
        v = lg::fermat2<S>(nw, limbs, g);
    } else {
        v = 1;
        for (uint32_t k = 0; k < bases.n && v; ++k) v = lg::sprp<S>(nw, limbs, bases.b[k], g);
    }
    if (g == 0) out[row] = (uint8_t)v;
}

template <int S, int KIND>
hipError_t launch_lg(const uint32_t *d_n, uint32_t limbs, uint32_t count, const uint32_t *d_count,
                     const uint32_t *d_idx, Bases bases, uint8_t *d_out, int /*waves_per_simd*/, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const dim3 grid((uint32_t)(((size_t)count * 4 + 255) / 256));
    hipLaunchKernelGGL((prp_lg_kernel<S, KIND>), grid, dim3(256), 0, stream, d_n, limbs, count, d_count, d_idx, bases,
                       d_out);
    return hipGetLastError();
}

template <int S, int KIND>
int occupancy_lg() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, prp_lg_kernel<S, KIND>, 256, 0) != hipSuccess || n <= 0)
        n = 1;
    return n;
}

template <int... Ss>
struct RegisterLG {
    RegisterLG() {
        (register_variant(Variant{lg::RB, 4 * Ss, 4, KIND_FERMAT2, lg::RB * 4 * Ss - 4, &launch_lg<Ss, KIND_FERMAT2>,
                                  &occupancy_lg<Ss, KIND_FERMAT2>}),
         ...);
        (register_variant(Variant{lg::RB, 4 * Ss, 4, KIND_SPRP, lg::RB * 4 * Ss - 14, &launch_lg<Ss, KIND_SPRP>,
                                  &occupancy_lg<Ss, KIND_SPRP>}),
         ...);
    }
};

}  // namespace pgap
