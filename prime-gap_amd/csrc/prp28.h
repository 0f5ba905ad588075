// prp28.h -- Fermat base 2 and strong-PRP (Miller-Rabin) tests of one number
// per lane, radix 2^28, L digits.
//
// Reference counterpart: miller_rabin_t::miller_rabin / powm
// (miller_rabin.h:157-245).  Verdicts follow the textbook definitions that the
// oracle states on GMP:
//   fermat2(N) = [2^(N-1) mod N == 1]                (0 for even N and N <= 1)
//   sprp(N, b) = strong probable prime to base b     (0 for even N and N < 3)
// The reference kernel's extra squaring (miller_rabin.h:233-241) is not kept.
#pragma once
#include "mont28.h"

namespace pgap {
namespace r28 {

// Loads N from `limbs` little-endian 32-bit words into L digits.
template <int L>
__device__ __forceinline__ void load_n(const uint32_t* __restrict__ nw, uint32_t limbs, uint32_t (&n)[L]) {
    constexpr int W = (28 * L + 31) / 32;
    uint32_t w[W];
#pragma unroll
    for (int j = 0; j < W; ++j) w[j] = (uint32_t)j < limbs ? nw[j] : 0u;
    from_words<L, W>(w, n);
}

// Shared set-up: Mont(1) = R mod N and y = Mont(2) = 2R mod N, built by
// modular doublings from 2^(b-1) < N (no division needed).
template <int L>
__device__ __forceinline__ void mont_one_two(const uint32_t (&n)[L], int b, uint32_t (&one)[L],
                                             uint32_t (&two)[L]) {
    const int top = b - 1;
    const int td = top / 28, tb = top - 28 * td;
#pragma unroll
    for (int j = 0; j < L; ++j) two[j] = (j == td) ? (1u << tb) : 0u;
    // 2^(b-1) -> 2^(28L) mod N takes 28L - (b-1) doublings.
    const int cnt = 28 * L - top;
    for (int c = 0; c < cnt; ++c) mod_double<L>(two, n);
#pragma unroll
    for (int j = 0; j < L; ++j) one[j] = two[j];
    mod_double<L>(two, n);
}

// Fermat base 2.  The top bit of E = N-1 is consumed by starting at Mont(2);
// every further bit is one squaring with the doubling folded in.
template <int L>
__device__ __forceinline__ uint32_t fermat2(const uint32_t* __restrict__ nw, uint32_t limbs) {
    uint32_t n[L];
    load_n<L>(nw, limbs, n);
    if ((n[0] & 1u) == 0) return 0;
    const int b = bitlen<L>(n);
    if (b < 2) return 0;
    const uint32_t np = neg_inv28(n[0]);
    uint32_t one[L], x[L];
    mont_one_two<L>(n, b, one, x);

    int i = b - 2;
    uint32_t ew = nw[i >> 5];
    while (i >= 0) {
        const uint32_t s = (i == 0) ? 0u : ((ew >> (i & 31)) & 1u);
        uint32_t nxt = ew;
        if ((i & 31) == 0 && i > 0) nxt = nw[(i >> 5) - 1];
        sqr_redc<L>(x, n, np, s);
        ew = nxt;
        --i;
    }
    cond_sub<L>(x, n);
    return equal<L>(x, one) ? 1u : 0u;
}

// c = (a + b) mod N for a, b < N.
template <int L>
__device__ __forceinline__ void mod_add(uint32_t (&a)[L], const uint32_t (&b)[L], const uint32_t (&n)[L]) {
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        uint32_t v = a[j] + b[j] + c;
        c = v >> 28;
        a[j] = v & MASK;
    }
    cond_sub<L>(a, n);
}

// Strong probable prime test to one base.
template <int L>
__device__ __forceinline__ uint32_t sprp(const uint32_t* __restrict__ nw, uint32_t limbs, uint32_t base) {
    uint32_t n[L];
    load_n<L>(nw, limbs, n);
    if ((n[0] & 1u) == 0) return 0;
    const int b = bitlen<L>(n);
    if (b < 2 || (b == 2 && n[0] < 3)) return 0;
    const uint32_t np = neg_inv28(n[0]);
    uint32_t one[L], two[L];
    mont_one_two<L>(n, b, one, two);

    // N - 1 = d * 2^tz
    int tz = 1;
    while (((nw[tz >> 5] >> (tz & 31)) & 1u) == 0) ++tz;

    // Reduce the base below N when N fits one word.
    uint32_t bb = base;
    if (b <= 32) {
        const uint32_t n32 = nw[0];
        bb = base % n32;
    }
    uint32_t x[L];
    if (bb == 2) {
#pragma unroll
        for (int j = 0; j < L; ++j) x[j] = two[j];
    } else {
        // Mont(bb) = bb * R mod N by double-and-add over the bits of bb.
#pragma unroll
        for (int j = 0; j < L; ++j) x[j] = 0;
        for (int t = 31; t >= 0; --t) {
            mod_double<L>(x, n);
            if ((bb >> t) & 1u) mod_add<L>(x, one, n);
        }
    }
    if (bb == 0) {
        return 0;  // base == 0 mod N: x stays 0, never 1 or N-1
    }
    uint32_t bm[L];
#pragma unroll
    for (int j = 0; j < L; ++j) bm[j] = x[j];

    // Left-to-right over the bits of d = (N-1) >> tz, top bit consumed.
    int i = b - 2;
    uint32_t ew = nw[i >> 5];
    while (i >= tz) {
        const uint32_t bit = (ew >> (i & 31)) & 1u;
        uint32_t nxt = ew;
        if ((i & 31) == 0 && i > 0) nxt = nw[(i >> 5) - 1];
        if (bb == 2) {
            sqr_redc<L>(x, n, np, bit);
        } else {
            sqr_redc<L>(x, n, np, 0u);
            if (bit) mul_redc<L>(x, x, bm, n, np);
        }
        ew = nxt;
        --i;
    }
    // x = Mont(base^d) < 4N (the last bit-folded squaring may double):
    // three conditional subtractions make it canonical.
    cond_sub<L>(x, n);
    cond_sub<L>(x, n);
    cond_sub<L>(x, n);
    uint32_t mone[L];  // Mont(-1) = N - Mont(1)
    {
        uint32_t br = 0;
#pragma unroll
        for (int j = 0; j < L; ++j) {
            uint32_t v = n[j] - one[j] - br;
            br = v >> 31;
            mone[j] = v & MASK;
        }
    }
    if (equal<L>(x, one) || equal<L>(x, mone)) return 1;
    for (int r = 1; r < tz; ++r) {
        sqr_redc<L>(x, n, np, 0u);
        cond_sub<L>(x, n);
        if (equal<L>(x, mone)) return 1;
        if (equal<L>(x, one)) return 0;
    }
    return 0;
}

}  // namespace r28
}  // namespace pgap
