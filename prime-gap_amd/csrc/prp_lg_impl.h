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
                     const uint32_t *d_idx, Bases bases, uint8_t *d_out, int blocks_cap, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const dim3 grid((uint32_t)(((size_t)count * 4 + 255) / 256));
    const size_t lds = occupancy_cap_bytes(prp_lg_kernel<S, KIND>, blocks_cap);
    hipLaunchKernelGGL((prp_lg_kernel<S, KIND>), grid, dim3(256), lds, stream, d_n, limbs, count, d_count, d_idx,
                       bases, d_out);
    return hipGetLastError();
}

// Streaming-scan round, a quad per number: confirmation items padded to
// whole wavefronts (16 items) first, then Fermat items.
template <int S>
__global__ void __launch_bounds__(256) scan_lg_kernel(const uint32_t *__restrict__ n, uint32_t limbs,
                                                      uint32_t conf_rows, uint32_t conf_max, uint32_t max_items,
                                                      const uint32_t *__restrict__ ctr,
                                                      const uint8_t *__restrict__ base, uint8_t *__restrict__ out) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t item = t >> 2, g = t & 3u;
    const uint32_t nc = min(ctr[3], conf_max), nf = min(ctr[0], max_items);
    const uint32_t ncw = (nc + 15u) & ~15u;
    uint32_t row, v;
    if (item < ncw) {
        if (item >= nc) return;
        row = item;
        v = lg::sprp<S>(n + (size_t)row * limbs, limbs, base[item], g);
    } else {
        if (item - ncw >= nf) return;
        row = conf_rows + (item - ncw);
        v = lg::fermat2<S>(n + (size_t)row * limbs, limbs, g);
    }
    if (g == 0) out[row] = (uint8_t)v;
}

template <int S>
hipError_t launch_scan_lg(const uint32_t *d_n, uint32_t limbs, uint32_t conf_rows, uint32_t conf_max,
                          uint32_t max_items, const uint32_t *d_ctr, const uint8_t *d_base, uint8_t *d_out,
                          int blocks_cap, hipStream_t stream) {
    const size_t items = (((size_t)conf_max + 15) & ~(size_t)15) + max_items;
    if (items == 0) return hipSuccess;
    const dim3 grid((uint32_t)((items * 4 + 255) / 256));
    const size_t lds = occupancy_cap_bytes(scan_lg_kernel<S>, blocks_cap);
    hipLaunchKernelGGL((scan_lg_kernel<S>), grid, dim3(256), lds, stream, d_n, limbs, conf_rows, conf_max,
                       max_items, d_ctr, d_base, d_out);
    return hipGetLastError();
}

template <int S>
int occupancy_scan_lg() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, scan_lg_kernel<S>, 256, 0) != hipSuccess || n <= 0) n = 1;
    return n;
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
                                  &occupancy_lg<Ss, KIND_SPRP>, &launch_scan_lg<Ss>, &occupancy_scan_lg<Ss>}),
         ...);
    }
};

}  // namespace pgap
