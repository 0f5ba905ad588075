// prp_lg.h -- Fermat base 2 and strong-PRP tests with one number per quad
// of lanes (radix 2^26, S digits per lane, L = 4S digits).  Same definitions
// and the same algorithm outline as the one-lane tests in prp28.h; every
// multi-digit step runs distributed over the quad (mont_lg.h).
#pragma once
#include "mont_lg.h"

namespace pgap {
namespace lg {

// This lane's S digits (digit index g*S + j) of N from 32-bit words.
template <int S>
__device__ __forceinline__ void load_digits(const uint32_t *__restrict__ nw, uint32_t limbs, uint32_t g,
                                            uint32_t (&d)[S]) {
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const uint32_t bit = (uint32_t)RB * (g * S + (uint32_t)j);
        const uint32_t wi = bit >> 5, sh = bit & 31;
        const uint32_t lo = wi < limbs ? nw[wi] : 0u;
        const uint32_t hi = wi + 1 < limbs ? nw[wi + 1] : 0u;
        const uint32_t v = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
        d[j] = v & MASK;
    }
}

template <int S>
__device__ __forceinline__ uint32_t bitlen(const uint32_t (&d)[S], uint32_t g) {
    uint32_t b = 0;
#pragma unroll
    for (int j = 0; j < S; ++j)
        if (d[j]) b = (uint32_t)RB * (g * S + (uint32_t)j) + (32u - (uint32_t)__builtin_clz(d[j]));
    return qmax(b);
}

// one = R mod N by modular doublings from 2^(b-1) (26*4S - (b-1) of them).
template <int S>
__device__ __forceinline__ void mont_one(const uint32_t (&n)[S], uint32_t b, uint32_t g, uint32_t (&one)[S]) {
    const uint32_t top = b - 1, td = top / RB, tb = top % RB;
#pragma unroll
    for (int j = 0; j < S; ++j) one[j] = (g * S + (uint32_t)j == td) ? (1u << tb) : 0u;
    const uint32_t cnt = (uint32_t)RB * 4 * S - top;
    for (uint32_t c = 0; c < cnt; ++c) mod_double<S>(one, n, g);
}

// one = R mod N, two = 2R mod N.
template <int S>
__device__ __forceinline__ void mont_one_two(const uint32_t (&n)[S], uint32_t b, uint32_t g, uint32_t (&one)[S],
                                             uint32_t (&two)[S]) {
    mont_one<S>(n, b, g, one);
#pragma unroll
    for (int j = 0; j < S; ++j) two[j] = one[j];
    mod_double<S>(two, n, g);
}

template <int S>
__device__ __forceinline__ uint32_t fermat2(const uint32_t *__restrict__ nw, uint32_t limbs, uint32_t g) {
    uint32_t n[S];
    load_digits<S>(nw, limbs, g, n);
    const uint32_t n0 = qbcast<0>(n[0]);
    if ((n0 & 1u) == 0) return 0;
    const uint32_t b = bitlen<S>(n, g);
    if (b < 2) return 0;
    const uint32_t np = neg_inv(n0);
    // x = Mont(2); Mont(1) is rebuilt after the loop instead of held
    uint32_t x[S];
    mont_one<S>(n, b, g, x);
    mod_double<S>(x, n, g);
    int i = (int)b - 2;
    uint32_t ew = nw[i >> 5];
    while (i >= 0) {
        const uint32_t s = (i == 0) ? 0u : ((ew >> (i & 31)) & 1u);
        uint32_t nxt = ew;
        if ((i & 31) == 0 && i > 0) nxt = nw[(i >> 5) - 1];
        sqr_redc<S>(x, n, np, g);
        shl<S>(x, s, g);
        ew = nxt;
        --i;
    }
    cond_sub<S>(x, n, g);
    uint32_t one[S];
    mont_one<S>(n, b, g, one);
    return equal<S>(x, one) ? 1u : 0u;
}

// x <- x * b (b < 64), digits renormalised.
template <int S>
__device__ __forceinline__ void mul_small(uint32_t (&x)[S], uint32_t b, uint32_t g) {
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) t[p] = (uint64_t)x[p] * b;
    normalize<S>(t, x, g);
}

// a <- (a + c) mod N for a, c < N.
template <int S>
__device__ __forceinline__ void mod_add(uint32_t (&a)[S], const uint32_t (&c)[S], const uint32_t (&n)[S],
                                        uint32_t g) {
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) t[p] = (uint64_t)a[p] + c[p];
    normalize<S>(t, a, g);
    cond_sub<S>(a, n, g);
}

template <int S>
__device__ __forceinline__ uint32_t sprp(const uint32_t *__restrict__ nw, uint32_t limbs, uint32_t base,
                                         uint32_t g) {
    uint32_t n[S];
    load_digits<S>(nw, limbs, g, n);
    const uint32_t n0 = qbcast<0>(n[0]);
    if ((n0 & 1u) == 0) return 0;
    const uint32_t b = bitlen<S>(n, g);
    if (b < 2) return 0;
    const uint32_t np = neg_inv(n0);
    uint32_t one[S], x[S];
    mont_one_two<S>(n, b, g, one, x);
    int tz = 1;
    while (((nw[tz >> 5] >> (tz & 31)) & 1u) == 0) ++tz;
    uint32_t bb = base;
    if (b <= 32) bb = base % nw[0];
    if (bb == 0) return 0;
    if (bb != 2) {
#pragma unroll
        for (int j = 0; j < S; ++j) x[j] = 0;
        for (int t = 5; t >= 0; --t) {
            mod_double<S>(x, n, g);
            if ((bb >> t) & 1u) mod_add<S>(x, one, n, g);
        }
    }
    uint32_t mone[S];
    (void)sub_n<S>(n, one, mone, g);  // N - Mont(1): no borrow since Mont(1) < N
    int i = (int)b - 2, checks = 0;
    uint32_t ew = i >= tz ? nw[i >> 5] : 0u;
    bool have = (i < tz);
    for (;;) {
        uint32_t s = 0, bit = 0, nxt = ew;
        const bool checking = have;
        if (checking) {
            for (int r = 0; r < 128; ++r) {  // x < 2*bb*N: reduce below N
                uint32_t u[S];
                if (sub_n<S>(x, n, u, g)) break;
#pragma unroll
                for (int j = 0; j < S; ++j) x[j] = u[j];
            }
            if (equal<S>(x, mone)) return 1;
            if (equal<S>(x, one)) return checks == 0 ? 1u : 0u;
            if (checks + 1 >= tz) return 0;
            ++checks;
        } else {
            bit = (ew >> (i & 31)) & 1u;
            if ((i & 31) == 0 && i > tz) nxt = nw[(i >> 5) - 1];
            s = bb == 2 ? bit : 0u;
        }
        sqr_redc<S>(x, n, np, g);
        shl<S>(x, s, g);
        if (!checking) {
            if (bb != 2 && bit) mul_small<S>(x, bb, g);
            ew = nxt;
            --i;
            have = (i < tz);
        }
    }
}

}  // namespace lg
}  // namespace pgap
