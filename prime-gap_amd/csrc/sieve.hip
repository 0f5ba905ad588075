// sieve.hip -- the unknown-file producer on the device: for every m of a
// batch, the offsets x in [-SL, SL] for which m*K + x has no prime factor
// <= max_prime.
//
// Reference counterpart: combined_sieve (combined_sieve.cpp:409-750 method1,
// :2014-2447 method2) and its writers (:1261-1445).  The reference sieves
// with primesieve on the host and finds, per prime, the m's it hits with a
// modular search (modulo_search.h:33-60); here every (prime, m) pair is
// evaluated directly on the GPU:
//   1. the primes <= max_prime come from a segmented sieve of Eratosthenes,
//      one 32 KiB LDS bitmap per workgroup, survivors appended unordered;
//   2. primes <= 2*SL + 1 ("small") mark a per-m LDS bitmap, one workgroup
//      per m, a thread per prime looping over its multiples;
//   3. larger primes hit each m's interval at most once: a thread per prime
//      loops over the batch's m's and sets one bit in the m's bitmap.
// The bitmaps come back to the host, which lists the survivors in outward
// order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine_impl.h"

namespace pgap {

// a mod p for any a < 2^64 and p < 2^62, mu = floor((2^64 - 1) / p): the
// estimated quotient is short by at most 2.
__device__ __forceinline__ uint64_t mod_b(uint64_t a, uint64_t p, uint64_t mu) {
    const uint64_t q = __umul64hi(a, mu);
    uint64_t r = a - q * p;
    r = r >= p ? r - p : r;
    return r >= p ? r - p : r;
}

// a * b mod p for a, b < p.  Below 2^32 the product fits 64 bits; above
// (p < 2^40), product = hi * 2^64 + lo with hi < 2^16, and 2^64 mod p = c64.
__device__ __forceinline__ uint64_t mulmod_b(uint64_t a, uint64_t b, uint64_t p, uint64_t mu, uint64_t c64) {
    if (p < (1ull << 32)) return mod_b(a * b, p, mu);
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
    return mod_b(mod_b(lo, p, mu) + mod_b(hi * c64, p, mu), p, mu);
}

constexpr uint32_t SEG_WORDS = 8192;              // 32 KiB of LDS
constexpr uint32_t SEG_BITS = SEG_WORDS * 32;     // odd numbers per segment

// Odd primes in [3, limit]: workgroup b covers odd n = 2k+1, k in
// [b*SEG_BITS, (b+1)*SEG_BITS).
__global__ void __launch_bounds__(256) segment_primes_kernel(const uint32_t *__restrict__ small, uint32_t nsmall,
                                                             uint64_t limit, uint64_t *__restrict__ out,
                                                             uint32_t *__restrict__ count, uint32_t cap) {
    __shared__ uint32_t bits[SEG_WORDS];
    for (uint32_t i = threadIdx.x; i < SEG_WORDS; i += 256) bits[i] = 0;
    __syncthreads();
    const uint64_t k0 = (uint64_t)blockIdx.x * SEG_BITS;
    const uint64_t lo = 2 * k0 + 1, hi = 2 * (k0 + SEG_BITS) + 1;  // odd numbers in [lo, hi)
    for (uint32_t s = threadIdx.x; s < nsmall; s += 256) {
        const uint64_t p = small[s];
        uint64_t start = p * p;
        if (start < lo) start = ((lo + p - 1) / p) * p;
        if ((start & 1) == 0) start += p;
        for (uint64_t n = start; n < hi; n += 2 * p) {
            const uint32_t k = (uint32_t)((n - 1) / 2 - k0);
            atomicOr(&bits[k >> 5], 1u << (k & 31));
        }
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < SEG_WORDS; w += 256) {
        uint32_t free_bits = ~bits[w];
        const uint64_t kw = k0 + 32ull * w;
        if (kw == 0) free_bits &= ~1u;  // n = 1
        // n = 2k+1 <= limit  <=>  k <= (limit-1)/2
        const uint64_t kmax = (limit - 1) / 2;
        if (kw > kmax) free_bits = 0;
        else if (kw + 31 > kmax) free_bits &= (uint32_t)((1ull << (kmax - kw + 1)) - 1);
        const uint32_t cnt = (uint32_t)__popc(free_bits);
        if (!cnt) continue;
        uint32_t base = atomicAdd(count, cnt);
        while (free_bits) {
            const uint32_t b = (uint32_t)__ffs(free_bits) - 1;
            free_bits &= free_bits - 1;
            if (base < cap) out[base] = 2 * (kw + b) + 1;
            ++base;
        }
    }
}

__global__ void mu_kernel(const uint64_t *__restrict__ p, size_t n, uint64_t *__restrict__ mu) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) mu[i] = ~0ull / p[i];
}

// K mod p, 16 bits at a time so that r * 2^16 + digit stays below 2^64 for
// p < 2^48.
__global__ void kmod_barrett_kernel(const uint32_t *__restrict__ K, uint32_t nk, const uint64_t *__restrict__ primes,
                                    const uint64_t *__restrict__ mu, size_t np, uint64_t *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const uint64_t p = primes[i], u = mu[i];
    uint64_t r = 0;
    for (int j = (int)nk - 1; j >= 0; --j) {
        r = mod_b((r << 16) | (K[j] >> 16), p, u);
        r = mod_b((r << 16) | (K[j] & 0xffffu), p, u);
    }
    out[i] = r;
}

// One workgroup per m: primes <= 2SL+1 mark the m's bitmap in LDS.
__global__ void __launch_bounds__(256) sieve_small_kernel(const uint64_t *__restrict__ m, uint32_t SL,
                                                          const uint32_t *__restrict__ primes,
                                                          const uint32_t *__restrict__ kmod,
                                                          const uint64_t *__restrict__ mu, uint32_t np,
                                                          uint32_t words, uint32_t *__restrict__ bitmap) {
    extern __shared__ uint32_t bm[];
    for (uint32_t i = threadIdx.x; i < words; i += 256) bm[i] = 0;
    __syncthreads();
    const uint64_t mi = m[blockIdx.x];
    const uint32_t W = 2 * SL + 1;
    for (uint32_t s = threadIdx.x; s < np; s += 256) {
        const uint64_t p = primes[s], u = mu[s];
        const uint64_t r = mod_b(mod_b(mi, p, u) * kmod[s], p, u);  // m*K mod p
        // index i = SL + x with x = -r mod p
        for (uint64_t i = (SL % p + p - r) % p; i < W; i += p) atomicOr(&bm[i >> 5], 1u << (i & 31));
    }
    __syncthreads();
    uint32_t *o = bitmap + (size_t)blockIdx.x * words;
    for (uint32_t i = threadIdx.x; i < words; i += 256) o[i] = bm[i];
}

// Primes > 2SL+1 (the rest of the table) hit each interval at most once.
__global__ void __launch_bounds__(256) sieve_large_kernel(const uint64_t *__restrict__ m, uint32_t nm, uint32_t SL,
                                                          const uint64_t *__restrict__ primes,
                                                          const uint64_t *__restrict__ kmod,
                                                          const uint64_t *__restrict__ mu, size_t np,
                                                          uint32_t words, uint32_t *__restrict__ bitmap) {
    const uint32_t W = 2 * SL + 1;
    for (size_t s = (size_t)blockIdx.x * 256 + threadIdx.x; s < np; s += (size_t)gridDim.x * 256) {
        const uint64_t p = primes[s];
        if (p <= W) continue;  // handled by sieve_small_kernel
        const uint64_t u = mu[s], kp = kmod[s];
        uint64_t c64 = mod_b(~0ull, p, u) + 1;  // 2^64 mod p
        if (c64 == p) c64 = 0;
        for (uint32_t j = 0; j < nm; ++j) {
            const uint64_t r = mulmod_b(mod_b(m[j], p, u), kp, p, u, c64);
            const uint64_t i = r <= SL ? SL - r : SL + p - r;
            if (i < W) atomicOr(&bitmap[(size_t)j * words + (i >> 5)], 1u << (i & 31));
        }
    }
}

}  // namespace pgap

using namespace pgap;

void pgap_engine::sieve_shutdown() {
    if (sieve_primes) (void)hipFree(sieve_primes);
    if (sieve_mu) (void)hipFree(sieve_mu);
    sieve_primes = nullptr;
    sieve_mu = nullptr;
    sieve_nprimes = 0;
    sieve_limit = 0;
    for (Family &f : families) {
        if (f.d_skmod) (void)hipFree(f.d_skmod);
        f.d_skmod = nullptr;
        f.skmod_limit = 0;
    }
}

static pgap_status sieve_tables(pgap_engine *e, Family &f, uint64_t limit) {
    hipError_t he;
    if (e->sieve_limit != limit) {
        e->sieve_shutdown();
        const uint64_t root = (uint64_t)std::sqrt((double)limit) + 2;
        std::vector<uint32_t> small;
        for (uint32_t p : primes_upto((uint32_t)root))
            if (p > 2) small.push_back(p);
        const double lg = std::log((double)std::max<uint64_t>(limit, 17));
        const uint32_t cap = (uint32_t)(1.26 * (double)limit / lg) + 64;
        uint32_t *d_small = nullptr, *d_count = nullptr;
        if ((he = hipMalloc(&e->sieve_primes, (size_t)cap * 8)) != hipSuccess ||
            (he = hipMalloc(&e->sieve_mu, (size_t)cap * 8)) != hipSuccess ||
            (he = hipMalloc(&d_small, std::max<size_t>(small.size(), 1) * 4)) != hipSuccess ||
            (he = hipMalloc(&d_count, 4)) != hipSuccess) {
            (void)hipGetLastError();
            if (d_small) (void)hipFree(d_small);
            e->sieve_shutdown();
            e->set_error("pgap_sieve: hipMalloc for primes <= %llu: %s", (unsigned long long)limit,
                         hipGetErrorString(he));
            return PGAP_E_OOM;
        }
        const uint64_t two = 2;
        const uint32_t one = 1;
        (void)hipMemcpyAsync(e->sieve_primes, &two, 8, hipMemcpyHostToDevice, e->stream);
        (void)hipMemcpyAsync(d_count, &one, 4, hipMemcpyHostToDevice, e->stream);
        (void)hipMemcpyAsync(d_small, small.data(), small.size() * 4, hipMemcpyHostToDevice, e->stream);
        const uint64_t nseg = ((limit - 1) / 2 + SEG_BITS) / SEG_BITS;
        hipLaunchKernelGGL(segment_primes_kernel, dim3((uint32_t)nseg), dim3(256), 0, e->stream, d_small,
                           (uint32_t)small.size(), limit, e->sieve_primes, d_count, cap);
        uint32_t n = 0;
        he = hipGetLastError();
        if (he == hipSuccess) he = hipMemcpyAsync(&n, d_count, 4, hipMemcpyDeviceToHost, e->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
        (void)hipFree(d_small);
        (void)hipFree(d_count);
        if (he != hipSuccess) {
            e->sieve_shutdown();
            return e->hip_fail(he, "prime generation");
        }
        if (n > cap) {
            e->sieve_shutdown();
            e->set_error("pgap_sieve: prime table overflow (%u > %u)", n, cap);
            return PGAP_E_STATE;
        }
        hipLaunchKernelGGL(mu_kernel, dim3((n + 255) / 256), dim3(256), 0, e->stream, e->sieve_primes, (size_t)n,
                           e->sieve_mu);
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "mu kernel");
        e->sieve_nprimes = n;
        e->sieve_limit = limit;
    }
    if (f.skmod_limit != limit) {
        if (f.d_skmod) (void)hipFree(f.d_skmod);
        f.d_skmod = nullptr;
        if ((he = hipMalloc(&f.d_skmod, std::max<size_t>(e->sieve_nprimes, 1) * 8)) != hipSuccess) {
            (void)hipGetLastError();
            e->set_error("pgap_sieve: hipMalloc K mod p: %s", hipGetErrorString(he));
            return PGAP_E_OOM;
        }
        hipLaunchKernelGGL(kmod_barrett_kernel, dim3((uint32_t)((e->sieve_nprimes + 255) / 256)), dim3(256), 0,
                           e->stream, f.d_K, (uint32_t)f.K.size(), e->sieve_primes, e->sieve_mu, e->sieve_nprimes,
                           f.d_skmod);
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "kmod kernel");
        f.skmod_limit = limit;
    }
    return PGAP_OK;
}

extern "C" pgap_status pgap_sieve(pgap_engine *e, int family, const uint64_t *m, size_t count, uint64_t max_prime,
                                  int32_t *prev, uint32_t *n_prev, int32_t *next, uint32_t *n_next) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    Family *f = e->family(family);
    if (!f || (count && (!m || !prev || !n_prev || !next || !n_next)) || max_prime < 2 ||
        max_prime > 20000000000ull) {
        e->set_error("pgap_sieve: bad arguments (max_prime in [2, 2e10])");
        return PGAP_E_ARG;
    }
    const uint32_t SL = f->SL;
    if (SL == 0 || SL > 200000) {
        e->set_error("pgap_sieve: family sieve length %u outside [1, 200000]", SL);
        return PGAP_E_ARG;
    }
    if (count == 0) return PGAP_OK;
    (void)hipSetDevice(e->opt.device);
    pgap_status st;
    if ((st = sieve_tables(e, *f, max_prime))) return st;

    // small primes (<= 2SL+1) with K mod p and Barrett constants, host-built
    const uint32_t W = 2 * SL + 1, words = (W + 31) / 32;
    std::vector<uint32_t> sp;
    for (uint32_t p : primes_upto((uint32_t)std::min<uint64_t>(W, max_prime))) sp.push_back(p);
    std::vector<uint32_t> skm(sp.size());
    std::vector<uint64_t> smu(sp.size());
    for (size_t i = 0; i < sp.size(); ++i) {
        skm[i] = mod_small(f->K, sp[i]);
        smu[i] = ~0ull / sp[i];
    }
    const size_t batch = std::min<size_t>(count, 4096);
    DevBuf &bm = e->d_n;  // reuse the batch number buffer for bitmaps
    if ((st = e->ensure(bm, batch * words * 4, "sieve bitmaps")) || (st = e->ensure(e->d_m, batch * 8, "m")) ||
        (st = e->ensure(e->d_idx0, sp.size() * 4 + 4, "small primes")) ||
        (st = e->ensure(e->d_idx1, sp.size() * 4 + 4, "small K mod p")) ||
        (st = e->ensure(e->d_x, sp.size() * 8 + 8, "small mu")))
        return st;
    hipError_t he;
    if ((he = hipMemcpyAsync(e->d_idx0.p, sp.data(), sp.size() * 4, hipMemcpyHostToDevice, e->stream)) !=
            hipSuccess ||
        (he = hipMemcpyAsync(e->d_idx1.p, skm.data(), skm.size() * 4, hipMemcpyHostToDevice, e->stream)) !=
            hipSuccess ||
        (he = hipMemcpyAsync(e->d_x.p, smu.data(), smu.size() * 8, hipMemcpyHostToDevice, e->stream)) != hipSuccess)
        return e->hip_fail(he, "sieve upload");
    std::vector<uint32_t> host_bm(batch * words);
    for (size_t off = 0; off < count; off += batch) {
        const uint32_t nb = (uint32_t)std::min(batch, count - off);
        if ((he = hipMemcpyAsync(e->d_m.p, m + off, nb * 8, hipMemcpyHostToDevice, e->stream)) != hipSuccess)
            return e->hip_fail(he, "sieve m upload");
        hipLaunchKernelGGL(sieve_small_kernel, dim3(nb), dim3(256), words * 4, e->stream, (const uint64_t *)e->d_m.p,
                           SL, (const uint32_t *)e->d_idx0.p, (const uint32_t *)e->d_idx1.p,
                           (const uint64_t *)e->d_x.p, (uint32_t)sp.size(), words, (uint32_t *)bm.p);
        if (e->sieve_nprimes > sp.size()) {
            const uint32_t grid = (uint32_t)std::min<size_t>((e->sieve_nprimes + 255) / 256, 65536);
            hipLaunchKernelGGL(sieve_large_kernel, dim3(grid), dim3(256), 0, e->stream, (const uint64_t *)e->d_m.p,
                               nb, SL, e->sieve_primes, f->d_skmod, e->sieve_mu, e->sieve_nprimes, words,
                               (uint32_t *)bm.p);
        }
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "sieve kernels");
        if ((he = hipMemcpyAsync(host_bm.data(), bm.p, (size_t)nb * words * 4, hipMemcpyDeviceToHost, e->stream)) !=
                hipSuccess ||
            (he = hipStreamSynchronize(e->stream)) != hipSuccess)
            return e->hip_fail(he, "sieve readback");
        for (uint32_t j = 0; j < nb; ++j) {
            const uint32_t *b = host_bm.data() + (size_t)j * words;
            auto composite = [&](uint32_t i) { return (b[i >> 5] >> (i & 31)) & 1u; };
            int32_t *pv = prev + (off + j) * SL, *nx = next + (off + j) * SL;
            uint32_t a = 0, c = 0;
            for (uint32_t d = 1; d <= SL; ++d) {
                if (!composite(SL - d)) pv[a++] = (int32_t)d;
                if (!composite(SL + d)) nx[c++] = (int32_t)d;
            }
            n_prev[off + j] = a;
            n_next[off + j] = c;
        }
    }
    return PGAP_OK;
}
