// peak.hip -- the denominator of the PRP kernels' roofline: the sustained
// rate of v_mad_u64_u32 (32x32-bit products accumulated into 64 bits), the
// instruction the Montgomery squarings spend nearly all their issue slots on.
//
// Each thread runs kChains independent accumulators; each step of a chain is
// one multiply-add whose multiplicand is the chain's own low word, so the
// compiler can neither hoist nor fold it, and the chains give every wavefront
// kChains independent instructions in flight.  The launch keeps
// kBlocksPerCu 256-thread workgroups per compute unit (8 waves per SIMD),
// enough to hide the instruction's latency.  No counterpart in the reference.
#include "engine_impl.h"

namespace pgap {

constexpr int kChains = 8;
constexpr int kUnroll = 16;
constexpr int kBlocksPerCu = 8;

__global__ void __launch_bounds__(256) mad_peak_kernel(uint32_t iters, uint32_t seed, uint64_t *__restrict__ out) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint64_t a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = (uint64_t)(seed ^ t) * (2 * k + 1) + k;
    const uint32_t b = 0x9e3779b9u ^ blockIdx.x;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) a[k] = (uint64_t)(uint32_t)a[k] * b + a[k];
        }
    }
    uint64_t x = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) x ^= a[k];
    out[t] = x;
}

}  // namespace pgap

using namespace pgap;

extern "C" pgap_status pgap_mad_peak(pgap_engine *e, uint32_t iters, double *mac_per_s, double *kernel_ms,
                                     uint64_t *macs) {
    if (!e || iters == 0) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    (void)hipSetDevice(e->opt.device);
    const uint32_t blocks = (uint32_t)e->num_cus * kBlocksPerCu;
    pgap_status st;
    if ((st = e->ensure(e->d_small, (size_t)blocks * 256 * sizeof(uint64_t), "mad peak output"))) return st;
    hipError_t he;
    float ms = 0;
    // one untimed launch (clocks ramp, code object loads), then the timed one
    for (int rep = 0; rep < 2; ++rep) {
        if ((he = hipEventRecord(e->ev0, e->stream)) != hipSuccess) return e->hip_fail(he, "hipEventRecord");
        hipLaunchKernelGGL(mad_peak_kernel, dim3(blocks), dim3(256), 0, e->stream, iters, 12345u,
                           (uint64_t *)e->d_small.p);
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "mad peak launch");
        if ((he = hipEventRecord(e->ev1, e->stream)) != hipSuccess ||
            (he = hipEventSynchronize(e->ev1)) != hipSuccess || (he = hipEventElapsedTime(&ms, e->ev0, e->ev1)))
            return e->hip_fail(he, "mad peak timing");
    }
    const uint64_t n = (uint64_t)blocks * 256 * iters * kUnroll * kChains;
    if (macs) *macs = n;
    if (kernel_ms) *kernel_ms = ms;
    if (mac_per_s) *mac_per_s = ms > 0 ? n / (ms * 1e-3) : 0;
    return PGAP_OK;
}
