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

// keep[i] = 1 iff m_i*K + x_i has no prime factor in the table.
__global__ void filter_kernel(const uint64_t *__restrict__ m, const int64_t *__restrict__ x, uint32_t count,
                              const uint32_t *__restrict__ primes, const uint32_t *__restrict__ kmod, uint32_t np,
                              uint8_t *__restrict__ keep) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    const uint64_t mi = m[i];
    const int64_t xi = x[i];
    const uint64_t ax = (uint64_t)(xi < 0 ? -xi : xi);
    uint8_t ok = 1;
    for (uint32_t k = 0; k < np; ++k) {
        const uint64_t p = primes[k];
        const uint64_t mr = (mi % p) * kmod[k] % p;   // m*K mod p
        const uint64_t xr = ax % p;
        const uint64_t r = xi < 0 ? (mr + p - xr) % p : (mr + xr) % p;
        if (r == 0) {
            ok = 0;
            break;
        }
    }
    keep[i] = ok;
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

void pgap_engine::families_shutdown() {
    for (Family &f : families) {
        if (f.d_K) (void)hipFree(f.d_K);
        if (f.d_primes) (void)hipFree(f.d_primes);
        if (f.d_kmod) (void)hipFree(f.d_kmod);
        f.d_K = nullptr;
        f.d_primes = f.d_kmod = nullptr;
    }
    families.clear();
}

static pgap_status family_tables(pgap_engine *e, Family &f, uint32_t limit) {
    if (f.filter_limit == limit && f.d_primes) return PGAP_OK;
    std::vector<uint32_t> ps = primes_upto(limit);
    if (f.d_primes) (void)hipFree(f.d_primes);
    if (f.d_kmod) (void)hipFree(f.d_kmod);
    f.d_primes = f.d_kmod = nullptr;
    f.filter_limit = 0;
    const size_t bytes = std::max<size_t>(ps.size(), 1) * sizeof(uint32_t);
    hipError_t he;
    if ((he = hipMalloc(&f.d_primes, bytes)) != hipSuccess || (he = hipMalloc(&f.d_kmod, bytes)) != hipSuccess) {
        (void)hipGetLastError();
        e->set_error("hipMalloc for %zu filter primes: %s", ps.size(), hipGetErrorString(he));
        return PGAP_E_OOM;
    }
    if (!ps.empty()) {
        he = hipMemcpyAsync(f.d_primes, ps.data(), ps.size() * sizeof(uint32_t), hipMemcpyHostToDevice, e->stream);
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
        e->set_error("pgap_family: d=%u does not divide %u# (gap_common.cpp:119)", D, P);
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
        (st = e->ensure(e->d_keep, count, "keep")))
        return st;
    hipError_t he;
    if ((he = hipMemcpyAsync(e->d_m.p, m, count * 8, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(e->d_x.p, x, count * 8, hipMemcpyHostToDevice, e->stream)) != hipSuccess)
        return e->hip_fail(he, "filter upload");
    hipLaunchKernelGGL(filter_kernel, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, e->stream,
                       (const uint64_t *)e->d_m.p, (const int64_t *)e->d_x.p, (uint32_t)count, f->d_primes, f->d_kmod,
                       (uint32_t)f->nprimes, (uint8_t *)e->d_keep.p);
    if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "filter kernel");
    if ((he = hipMemcpyAsync(keep, e->d_keep.p, count, hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
        (he = hipStreamSynchronize(e->stream)) != hipSuccess)
        return e->hip_fail(he, "filter sync");
    return PGAP_OK;
}

}  // extern "C"
