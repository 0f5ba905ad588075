// prp_lg.h -- Fermat base 2 and strong-PRP tests with one number per group
// of G lanes (radix 2^28, S digits per lane, L = G*S digits).  Same definitions
// and the same algorithm outline as the one-lane tests in prp28.h; every
// multi-digit step runs distributed over the group (mont_lg.h).
#pragma once
#include "mont_lg.h"

namespace pgap {
namespace lg {

// This lane's S digits (digit index g*S + j) of N from 32-bit words.
template <int S, int G>
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

template <int S, int G>
__device__ __forceinline__ uint32_t bitlen(const uint32_t (&d)[S], uint32_t g) {
    uint32_t b = 0;
#pragma unroll
    for (int j = 0; j < S; ++j)
        if (d[j]) b = (uint32_t)RB * (g * S + (uint32_t)j) + (32u - (uint32_t)__builtin_clz(d[j]));
    return gmax<G>(b);
}

// one = R mod N by modular doublings from 2^(b-1) (28*G*S - (b-1) of them).
template <int S, int G>
__device__ __forceinline__ void mont_one(const uint32_t (&n)[S], uint32_t b, uint32_t g, uint32_t (&one)[S]) {
    const uint32_t top = b - 1, td = top / RB, tb = top % RB;
#pragma unroll
    for (int j = 0; j < S; ++j) one[j] = (g * S + (uint32_t)j == td) ? (1u << tb) : 0u;
    const uint32_t cnt = (uint32_t)RB * G * S - top;
    for (uint32_t c = 0; c < cnt; ++c) mod_double<S, G>(one, n, g);
}

// one = R mod N, two = 2R mod N.
template <int S, int G>
__device__ __forceinline__ void mont_one_two(const uint32_t (&n)[S], uint32_t b, uint32_t g, uint32_t (&one)[S],
                                             uint32_t (&two)[S]) {
    mont_one<S, G>(n, b, g, one);
#pragma unroll
    for (int j = 0; j < S; ++j) two[j] = one[j];
    mod_double<S, G>(two, n, g);
}

template <int S, int G>
__device__ __forceinline__ uint32_t fermat2(const uint32_t *__restrict__ nw, uint32_t limbs, uint32_t g,
                                            uint32_t *stage) {
    uint32_t n[S];
    load_digits<S, G>(nw, limbs, g, n);
    const uint32_t n0 = gbcast0<G>(n[0]);
    if ((n0 & 1u) == 0) return 0;
    const uint32_t b = bitlen<S, G>(n, g);
    if (b < 2) return 0;
    const uint32_t np = neg_inv(n0);
    // x = Mont(2); Mont(1) is rebuilt after the loop instead of held
    uint32_t x[S];
    mont_one<S, G>(n, b, g, x);
    mod_double<S, G>(x, n, g);
    int i = (int)b - 2;
    uint32_t ew = nw[i >> 5];
    while (i >= 0) {
        const uint32_t s = (i == 0) ? 0u : ((ew >> (i & 31)) & 1u);
        uint32_t nxt = ew;
        if ((i & 31) == 0 && i > 0) nxt = nw[(i >> 5) - 1];
        sqr_redc<S, G, true>(x, n, np, g, stage, s);
        ew = nxt;
        --i;
    }
    canon<S, G>(x, g);
    cond_sub<S, G>(x, n, g);
    uint32_t one[S];
    mont_one<S, G>(n, b, g, one);
    return equal<S, G>(x, one) ? 1u : 0u;
}

// x <- x * b (b < 64).  LAZY: lazy_carry digits out (the exponent loop's
// form; lazy digits in, each product < 2^35); otherwise exact digits.
template <int S, int G, bool LAZY = false>
__device__ __forceinline__ void mul_small(uint32_t (&x)[S], uint32_t b, uint32_t g) {
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) t[p] = (uint64_t)x[p] * b;
    if constexpr (LAZY)
        lazy_carry<S, G>(t, x, g, 0u);
    else
        normalize<S, G>(t, x, g);
}

// a <- (a + c) mod N for a, c < N.
template <int S, int G>
__device__ __forceinline__ void mod_add(uint32_t (&a)[S], const uint32_t (&c)[S], const uint32_t (&n)[S],
                                        uint32_t g) {
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) t[p] = (uint64_t)a[p] + c[p];
    normalize<S, G>(t, a, g);
    cond_sub<S, G>(a, n, g);
}

template <int S, int G>
__device__ __forceinline__ uint32_t sprp(const uint32_t *__restrict__ nw, uint32_t limbs, uint32_t base,
                                         uint32_t g, uint32_t *stage) {
    uint32_t n[S];
    load_digits<S, G>(nw, limbs, g, n);
    const uint32_t n0 = gbcast0<G>(n[0]);
    if ((n0 & 1u) == 0) return 0;
    const uint32_t b = bitlen<S, G>(n, g);
    if (b < 2) return 0;
    const uint32_t np = neg_inv(n0);
    int tz = 1;
    while (((nw[tz >> 5] >> (tz & 31)) & 1u) == 0) ++tz;
    uint32_t bb = base;
    if (b <= 32) bb = base % nw[0];
    if (bb == 0) return 0;
    uint32_t x[S];
    {
        // x = Mont(bb); Mont(1) is needed again only in phase 2
        uint32_t one[S];
        mont_one_two<S, G>(n, b, g, one, x);
        if (bb != 2) {
#pragma unroll
            for (int j = 0; j < S; ++j) x[j] = 0;
            for (int t = 5; t >= 0; --t) {
                mod_double<S, G>(x, n, g);
                if ((bb >> t) & 1u) mod_add<S, G>(x, one, n, g);
            }
        }
    }
    // Phase 1: exponent bits of d (positions b-2 .. tz of N - 1)
    int i = (int)b - 2;
    uint32_t ew = i >= tz ? nw[i >> 5] : 0u;
    while (i >= tz) {
        const uint32_t bit = (ew >> (i & 31)) & 1u;
        uint32_t nxt = ew;
        if ((i & 31) == 0 && i > tz) nxt = nw[(i >> 5) - 1];
        sqr_redc<S, G, true>(x, n, np, g, stage, bb == 2 ? bit : 0u);
        if (bb != 2 && bit) mul_small<S, G, true>(x, bb, g);
        ew = nxt;
        --i;
    }
    canon<S, G>(x, g);
    // Phase 2: x = Mont(b^d) < 2*bb*N; reduce, compare with Mont(+-1), and
    // square at most tz - 1 times looking for Mont(-1)
    for (int r = 0; r < 128; ++r) {
        uint32_t u[S];
        if (sub_n<S, G>(x, n, u, g)) break;
#pragma unroll
        for (int j = 0; j < S; ++j) x[j] = u[j];
    }
    uint32_t one[S], mone[S];
    mont_one<S, G>(n, b, g, one);
    (void)sub_n<S, G>(n, one, mone, g);  // N - Mont(1): no borrow since Mont(1) < N
    if (equal<S, G>(x, one) || equal<S, G>(x, mone)) return 1;
    for (int c = 1; c < tz; ++c) {
        sqr_redc<S, G>(x, n, np, g, stage);
        cond_sub<S, G>(x, n, g);  // x < 2N -> x < N
        if (equal<S, G>(x, mone)) return 1;
        if (equal<S, G>(x, one)) return 0;
    }
    return 0;
}

}  // namespace lg
}  // namespace pgap
