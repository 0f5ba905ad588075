// family.hip -- K = P#/d families and device-side candidate construction /
// trial-division screening of N = m*K + x.
//
// Reference counterparts: init_K / K_stats (gap_common.cpp:116-140) for K;
// the host batch fill mpz_mul_ui(center, K, m) + mpz_sub_ui/mpz_add_ui and
// from_mpz export (gap_test_gpu.cu:418-476, miller_rabin.h:37-52) for the
// construction; the per-prime offset test of combined_sieve for screening.
#include <algorithm>
#include <cstring>

#include "engine_impl.h"

namespace pgap {

// K mod p for every p in the table (Horner over K's words, high to low).
__global__ void kmod_kernel(const uint32_t *__restrict__ K, uint32_t nk, const uint32_t *__restrict__ primes,
                            uint32_t np, uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= np) return;
    const uint64_t p = primes[i];
    uint64_t r = 0;
    for (int j = (int)nk - 1; j >= 0; --j) r = ((r << 32) | K[j]) % p;
    out[i] = (uint32_t)r;
}

// a mod p with mu = floor((2^64-1)/p): the quotient estimate is at most 2
// short, so two conditional subtractions finish it.
__device__ __forceinline__ uint64_t mod_barrett(uint64_t a, uint64_t p, uint64_t mu) {
    const uint64_t q = __umul64hi(a, mu);
    uint64_t r = a - q * p;
    r = r >= p ? r - p : r;
    return r >= p ? r - p : r;
}

// One stage of the trial-division screen over primes [k0, k1).  Stage 0
// reads every candidate; later stages read the survivors the previous stage
// appended to idx_in (count in *n_in), so waves stay dense instead of
// idling behind the few lanes that survive the small primes.
__global__ void filter_stage_kernel(const uint64_t *__restrict__ m, const int64_t *__restrict__ x,
                                    const uint32_t *__restrict__ idx_in, const uint32_t *__restrict__ n_in,
                                    uint32_t count, const uint32_t *__restrict__ primes,
                                    const uint32_t *__restrict__ kmod, const uint64_t *__restrict__ mu, uint32_t k0,
                                    uint32_t k1, int last, uint8_t *__restrict__ keep, uint32_t *__restrict__ idx_out,
                                    uint32_t *__restrict__ n_out) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t n = idx_in ? *n_in : count;
    if (t >= n) return;
    const uint32_t i = idx_in ? idx_in[t] : t;
    const uint64_t mi = m[i];
    const int64_t xi = x[i];
    const uint64_t ax = xi < 0 ? (uint64_t)(-(xi + 1)) + 1u : (uint64_t)xi;
    for (uint32_t k = k0; k < k1; ++k) {
        const uint64_t p = primes[k], u = mu[k];
        const uint64_t mr = mod_barrett(mod_barrett(mi, p, u) * kmod[k], p, u);  // m*K mod p
        const uint64_t xr = mod_barrett(ax, p, u);
        const bool hit = xi < 0 ? (mr == xr) : ((mr + xr == p) || (mr == 0 && xr == 0));
        if (hit) return;  // keep[i] stays 0
    }
    if (last) {
        keep[i] = 1;
    } else {
        idx_out[atomicAdd(n_out, 1u)] = i;
    }
}

// out[i] = m_i*K + x_i as `limbs` words; *err |= 1 on overflow or negative.
__global__ void build_kernel(const uint32_t *__restrict__ K, uint32_t nk, const uint64_t *__restrict__ m,
                             const int64_t *__restrict__ x, uint32_t count, uint32_t limbs,
                             uint32_t *__restrict__ out, uint32_t *__restrict__ err) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    uint32_t *o = out + (size_t)i * limbs;
    const uint64_t mi = m[i];
    const uint64_t mlo = (uint32_t)mi, mhi = mi >> 32;
    // m*K: two passes of a 32x(nk) product
    uint64_t c0 = 0, c1 = 0;  // carries of the lo and hi partial rows
    uint32_t bad = 0;
    for (uint32_t j = 0; j < limbs; ++j) {
        const uint64_t kj = j < nk ? K[j] : 0u;
        const uint64_t kj1 = (j >= 1 && j - 1 < nk) ? K[j - 1] : 0u;
        const uint64_t t0 = kj * mlo + c0;           // < 2^64
        c0 = t0 >> 32;
        const uint64_t t1 = kj1 * mhi + (t0 & 0xffffffffu) + c1;
        c1 = t1 >> 32;
        o[j] = (uint32_t)t1;
    }
    if (c0 | c1) bad = 1;
    for (uint32_t j = limbs; j < nk + 2 && !bad; ++j) {
        // words beyond `limbs` must all be zero: recompute their top part
        const uint64_t kj = j < nk ? K[j] : 0u;
        const uint64_t kj1 = (j >= 1 && j - 1 < nk) ? K[j - 1] : 0u;
        if (kj * mlo || kj1 * mhi) bad = 1;
    }
    // + x
    const int64_t xi = x[i];
    if (xi >= 0) {
        uint64_t c = (uint64_t)xi;
        for (uint32_t j = 0; j < limbs && c; ++j) {
            const uint64_t t = (uint64_t)o[j] + (c & 0xffffffffu);
            o[j] = (uint32_t)t;
            c = (c >> 32) + (t >> 32);
        }
        if (c) bad = 1;
    } else {
        uint64_t b = (uint64_t)(-(xi + 1)) + 1u;
        for (uint32_t j = 0; j < limbs && b; ++j) {
            const uint64_t sub = b & 0xffffffffu;
            const uint64_t cur = o[j];
            o[j] = (uint32_t)(cur - sub);
            b = (b >> 32) + (cur < sub ? 1u : 0u);
        }
        if (b) bad = 1;
    }
    if (bad) atomicOr(err, 1u);
}

}  // namespace pgap

using namespace pgap;

static void free_tables(Family &f) {
    if (f.d_primes) (void)hipFree(f.d_primes);
    if (f.d_kmod) (void)hipFree(f.d_kmod);
    if (f.d_mu) (void)hipFree(f.d_mu);
    f.d_primes = f.d_kmod = nullptr;
    f.d_mu = nullptr;
    f.filter_limit = 0;
}

void pgap_engine::families_shutdown() {
    for (Family &f : families) {
        if (f.d_K) (void)hipFree(f.d_K);
        f.d_K = nullptr;
        free_tables(f);
    }
    families.clear();
}

static pgap_status family_tables(pgap_engine *e, Family &f, uint32_t limit) {
    if (f.filter_limit == limit && f.d_primes) return PGAP_OK;
    free_tables(f);
    std::vector<uint32_t> ps = primes_upto(limit);
    std::vector<uint64_t> mu(ps.size());
    for (size_t i = 0; i < ps.size(); ++i) mu[i] = UINT64_MAX / ps[i];
    f.stage_end[0] = (uint32_t)(std::lower_bound(ps.begin(), ps.end(), 1u << 10) - ps.begin());
    f.stage_end[1] = (uint32_t)(std::lower_bound(ps.begin(), ps.end(), 1u << 15) - ps.begin());
    const size_t bytes = std::max<size_t>(ps.size(), 1) * sizeof(uint32_t);
    hipError_t he;
    if ((he = hipMalloc(&f.d_primes, bytes)) != hipSuccess || (he = hipMalloc(&f.d_kmod, bytes)) != hipSuccess ||
        (he = hipMalloc(&f.d_mu, 2 * bytes)) != hipSuccess) {
        (void)hipGetLastError();
        e->set_error("hipMalloc for %zu filter primes: %s", ps.size(), hipGetErrorString(he));
        free_tables(f);
        return PGAP_E_OOM;
    }
    if (!ps.empty()) {
        he = hipMemcpyAsync(f.d_primes, ps.data(), ps.size() * sizeof(uint32_t), hipMemcpyHostToDevice, e->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(f.d_mu, mu.data(), mu.size() * sizeof(uint64_t), hipMemcpyHostToDevice, e->stream);
        if (he != hipSuccess) return e->hip_fail(he, "upload primes");
        hipLaunchKernelGGL(kmod_kernel, dim3((uint32_t)((ps.size() + 255) / 256)), dim3(256), 0, e->stream, f.d_K,
                           (uint32_t)f.K.size(), f.d_primes, (uint32_t)ps.size(), f.d_kmod);
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "kmod kernel");
        if ((he = hipStreamSynchronize(e->stream)) != hipSuccess) return e->hip_fail(he, "kmod sync");
    }
    f.nprimes = ps.size();
    f.filter_limit = limit;
    return PGAP_OK;
}

extern "C" {

pgap_status pgap_family(pgap_engine *e, uint32_t P, uint32_t D, uint32_t sieve_length, int *family_id) {
    if (!e || !family_id) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    if (P < 2 || P > 1000000 || D == 0) {
        e->set_error("pgap_family: bad P=%u D=%u", P, D);
        return PGAP_E_ARG;
    }
    for (size_t i = 0; i < e->families.size(); ++i) {
        const Family &f = e->families[i];
        if (f.P == P && f.D == D && f.SL == sieve_length) {
            *family_id = (int)i;
            return PGAP_OK;
        }
    }
    Family f;
    f.P = P;
    f.D = D;
    f.SL = sieve_length;
    if (!primorial_over(P, D, f.K)) {
        e->set_error("pgap_family: d=%u must divide %u# with P#/d > 1 (gap_common.cpp:117-120)", D, P);
        return PGAP_E_ARG;
    }
    f.K_log = ln_words(f.K);
    (void)hipSetDevice(e->opt.device);
    hipError_t he = hipMalloc(&f.d_K, f.K.size() * sizeof(uint32_t));
    if (he != hipSuccess) {
        (void)hipGetLastError();
        e->set_error("hipMalloc K: %s", hipGetErrorString(he));
        return PGAP_E_OOM;
    }
    he = hipMemcpy(f.d_K, f.K.data(), f.K.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        (void)hipFree(f.d_K);
        return e->hip_fail(he, "upload K");
    }
    e->families.push_back(std::move(f));
    *family_id = (int)e->families.size() - 1;
    return PGAP_OK;
}

pgap_status pgap_build_candidates_device(pgap_engine *e, int family, const uint64_t *m, const int64_t *x,
                                         size_t count, uint32_t limbs, uint32_t *d_out) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    Family *f = e->family(family);
    if (!f || (count && (!m || !x || !d_out)) || limbs == 0 || count > 0xffffffffu) {
        e->set_error("pgap_build_candidates_device: bad arguments");
        return PGAP_E_ARG;
    }
    if (count == 0) return PGAP_OK;
    (void)hipSetDevice(e->opt.device);
    pgap_status st;
    if ((st = e->ensure(e->d_m, count * 8, "m")) || (st = e->ensure(e->d_x, count * 8, "x")) ||
        (st = e->ensure(e->d_small, 64, "flags")))
        return st;
    hipError_t he;
    if ((he = hipMemcpyAsync(e->d_m.p, m, count * 8, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(e->d_x.p, x, count * 8, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemsetAsync(e->d_small.p, 0, 4, e->stream)) != hipSuccess)
        return e->hip_fail(he, "build upload");
    hipLaunchKernelGGL(build_kernel, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, e->stream, f->d_K,
                       (uint32_t)f->K.size(), (const uint64_t *)e->d_m.p, (const int64_t *)e->d_x.p, (uint32_t)count,
                       limbs, d_out, (uint32_t *)e->d_small.p);
    if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "build kernel");
    uint32_t err = 0;
    if ((he = hipMemcpyAsync(&err, e->d_small.p, 4, hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
        (he = hipStreamSynchronize(e->stream)) != hipSuccess)
        return e->hip_fail(he, "build sync");
    if (err) {
        e->set_error("pgap_build_candidates_device: some m*K+x does not fit %u words or is negative", limbs);
        return PGAP_E_TOO_LARGE;
    }
    return PGAP_OK;
}

pgap_status pgap_filter_candidates(pgap_engine *e, int family, const uint64_t *m, const int64_t *x, size_t count,
                                   uint32_t limit, uint8_t *keep) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    Family *f = e->family(family);
    if (!f || (count && (!m || !x || !keep)) || count > 0xffffffffu || limit > 2000000000u) {
        e->set_error("pgap_filter_candidates: bad arguments");
        return PGAP_E_ARG;
    }
    if (count == 0) return PGAP_OK;
    (void)hipSetDevice(e->opt.device);
    pgap_status st;
    if ((st = family_tables(e, *f, limit))) return st;
    if ((st = e->ensure(e->d_m, count * 8, "m")) || (st = e->ensure(e->d_x, count * 8, "x")) ||
        (st = e->ensure(e->d_keep, count, "keep")) || (st = e->ensure(e->d_idx0, count * 4, "index")) ||
        (st = e->ensure(e->d_idx1, count * 4, "index")) || (st = e->ensure(e->d_small, 64, "counters")))
        return st;
    hipError_t he;
    if ((he = hipMemcpyAsync(e->d_m.p, m, count * 8, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(e->d_x.p, x, count * 8, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemsetAsync(e->d_keep.p, 0, count, e->stream)) != hipSuccess ||
        (he = hipMemsetAsync(e->d_small.p, 0, 16, e->stream)) != hipSuccess)
        return e->hip_fail(he, "filter upload");
    // three stages: primes < 2^10, < 2^15, the rest; survivors are compacted
    // between stages (counters in d_small[0], d_small[1]).
    const uint32_t ends[3] = {std::min(f->stage_end[0], (uint32_t)f->nprimes),
                              std::min(f->stage_end[1], (uint32_t)f->nprimes), (uint32_t)f->nprimes};
    uint32_t *ctr = (uint32_t *)e->d_small.p;
    uint32_t *idx[2] = {(uint32_t *)e->d_idx0.p, (uint32_t *)e->d_idx1.p};
    uint32_t k0 = 0;
    for (int s = 0; s < 3; ++s) {
        const bool last = (s == 2);
        const uint32_t *in = s == 0 ? nullptr : idx[(s - 1) & 1];
        const uint32_t *n_in = s == 0 ? nullptr : ctr + (s - 1);
        hipLaunchKernelGGL(filter_stage_kernel, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, e->stream,
                           (const uint64_t *)e->d_m.p, (const int64_t *)e->d_x.p, in, n_in, (uint32_t)count,
                           f->d_primes, f->d_kmod, f->d_mu, k0, ends[s], last ? 1 : 0, (uint8_t *)e->d_keep.p,
                           last ? nullptr : idx[s & 1], last ? nullptr : ctr + s);
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "filter kernel");
        k0 = ends[s];
    }
    if ((he = hipMemcpyAsync(keep, e->d_keep.p, count, hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
        (he = hipStreamSynchronize(e->stream)) != hipSuccess)
        return e->hip_fail(he, "filter sync");
    return PGAP_OK;
}

}  // extern "C"
