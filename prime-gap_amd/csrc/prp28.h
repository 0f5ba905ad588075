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

// Mont(1) = R mod N for a b-bit N, built by modular doublings from
// 2^(b-1) < N (no division needed): 28L - (b-1) doublings.
template <int L>
__device__ __forceinline__ void mont_one(const uint32_t (&n)[L], int b, uint32_t (&one)[L]) {
    const int top = b - 1;
    const int td = top / 28, tb = top - 28 * td;
#pragma unroll
    for (int j = 0; j < L; ++j) one[j] = (j == td) ? (1u << tb) : 0u;
    const int cnt = 28 * L - top;
    for (int c = 0; c < cnt; ++c) mod_double<L>(one, n);
}

// Parking: a lane's column of L words in LDS, stride PARK_STRIDE (one column
// per thread of a 256-thread block; consecutive lanes hit consecutive banks).
// In the wide variants (L > 62) the result of an exponentiation loop leaves
// it through the park and the code after the loop reads x and N again (N
// from global memory).  With x and N live out of the fully unrolled loop
// into the check, the register allocator spreads the loop itself over AGPRs
// (78 digits: 439 registers, 182 ms for 2^18 Fermat tests; parked: 166 ms).
// The narrow variants keep the check in registers: parked, the 40-digit
// Fermat kernel measured 1.3% slower.
constexpr int PARK_STRIDE = 256;

// Whether fermat2<L> uses the park (sprp always does).
constexpr bool fermat_parks(int L) { return L > 62; }

template <int L>
__device__ __forceinline__ void park_store(uint32_t *park, const uint32_t (&x)[L]) {
#pragma unroll
    for (int j = 0; j < L; ++j) park[j * PARK_STRIDE] = x[j];
}

// x and N from the park and the number's words; the memory clobber keeps the
// compiler from forwarding the stored values or reusing earlier loads.
template <int L>
__device__ __forceinline__ void unpark(const uint32_t *park, const uint32_t *__restrict__ nw, uint32_t limbs,
                                       uint32_t (&x)[L], uint32_t (&n)[L]) {
    asm volatile("" ::: "memory");
#pragma unroll
    for (int j = 0; j < L; ++j) x[j] = park[j * PARK_STRIDE];
    load_n<L>(nw, limbs, n);
}

// redc_class of the parked value.
template <int L>
__device__ __forceinline__ uint32_t parked_class(const uint32_t *park, const uint32_t *__restrict__ nw,
                                                 uint32_t limbs) {
    uint32_t x[L], n[L];
    unpark<L>(park, nw, limbs, x, n);
    return redc_class<L>(x, n, neg_inv28(n[0]));
}

// One exponent step x <- x^2 R^{-1} 2^s: the column-folded doubling up to 62
// digits, the digit-pass doubling beyond (the column bound of the folded form).
// C: chains per column of the wide form (1, or 3 for kernels that run one
// wavefront per SIMD).
template <int L, int C = 1>
__device__ __forceinline__ void sqr_step(uint32_t (&x)[L], const uint32_t (&n)[L], uint32_t np, uint32_t s) {
    if constexpr (L <= 62)
        sqr_redc<L>(x, n, np, s);
    else
        sqr_redc_wide<L, C>(x, n, np, s);
}

// x <- x * b for a small base b (b < 64): digits times b with the carry
// propagated.  Lazy bound: x < 2N in gives x < 2bN out, and a squaring of an
// input < 2bN returns < 2N as long as N < R / (4 b^2): the strong-PRP
// variants' max_bits (28L - 8) covers b <= 8, and the engine selects with six
// more bits for larger bases (N < R/2^14 covers every b < 64).
template <int L>
__device__ __forceinline__ void mul_small(uint32_t (&x)[L], uint32_t b) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const uint64_t v = mad(x[j], b, c);  // < 2^34
        c = v >> 28;
        x[j] = (uint32_t)v & MASK;
    }
}

// Fermat base 2.  The top bit of E = N-1 is consumed by starting at Mont(2);
// every further bit is one squaring with the doubling folded in.
template <int L, int C = 1>
__device__ __forceinline__ uint32_t fermat2(const uint32_t* __restrict__ nw, uint32_t limbs, uint32_t* park) {
    uint32_t n[L];
    load_n<L>(nw, limbs, n);
    if ((n[0] & 1u) == 0) return 0;
    const int b = bitlen<L>(n);
    if (b < 2) return 0;
    const uint32_t np = neg_inv28(n[0]);
    // x = Mont(2)
    uint32_t x[L];
    mont_one<L>(n, b, x);
    mod_double<L>(x, n);

    int i = b - 2;
    uint32_t ew = nw[i >> 5];
    while (i >= 0) {
        const uint32_t s = (i == 0) ? 0u : ((ew >> (i & 31)) & 1u);
        uint32_t nxt = ew;
        if ((i & 31) == 0 && i > 0) nxt = nw[(i >> 5) - 1];
        sqr_step<L, C>(x, n, np, s);
        ew = nxt;
        --i;
    }
    if constexpr (fermat_parks(L)) {
        // x = Mont(1) is read from the residue itself (redc_class)
        park_store<L>(park, x);
        return parked_class<L>(park, nw, limbs) == 1u ? 1u : 0u;
    } else {
        // Mont(1) is rebuilt after the loop rather than held through it
        cond_sub<L>(x, n);
        uint32_t one[L];
        mont_one<L>(n, b, one);
        return equal<L>(x, one) ? 1u : 0u;
    }
}

// Strong probable prime test to one base b < 64 (N - 1 = d * 2^tz):
// x = b^d, pass if x = +-1 or x^(2^r) = -1 for some 0 < r < tz.  Two
// phases, each with its own squaring loop: the exponent bits of d, which
// holds only x, N and the squaring's temporaries (so it runs as fast as the
// Fermat loop), then at most tz - 1 checking squarings, the only place that
// needs Mont(1) and Mont(-1).
template <int L, int C = 1>
__device__ __forceinline__ uint32_t sprp(const uint32_t* __restrict__ nw, uint32_t limbs, uint32_t base,
                                         uint32_t* park) {
    uint32_t n[L];
    load_n<L>(nw, limbs, n);
    if ((n[0] & 1u) == 0) return 0;
    const int b = bitlen<L>(n);
    if (b < 2) return 0;  // N = 1
    const uint32_t np = neg_inv28(n[0]);

    int tz = 1;
    while (((nw[tz >> 5] >> (tz & 31)) & 1u) == 0) ++tz;

    uint32_t bb = base;
    if (b <= 32) bb = base % nw[0];  // N < 2^32: reduce the base
    if (bb == 0) return 0;           // x = 0 is never +-1
    // x = bb * Mont(1), congruent to Mont(bb) = bb * R mod N and < bb * N,
    // which the first squaring absorbs (see mul_small)
    uint32_t x[L];
    mont_one<L>(n, b, x);
    mul_small<L>(x, bb);

    // Phase 1: exponent bits of d below its top bit, positions b-2 .. tz of
    // N - 1.  Base 2 folds the multiply into the squaring; other bases
    // multiply by bb (x < 2*bb*N afterwards, which the next squaring
    // absorbs: see mul_small).
    int i = b - 2;
    uint32_t ew = i >= tz ? nw[i >> 5] : 0u;
    while (i >= tz) {
        const uint32_t bit = (ew >> (i & 31)) & 1u;
        uint32_t nxt = ew;
        if ((i & 31) == 0 && i > tz) nxt = nw[(i >> 5) - 1];
        sqr_step<L, C>(x, n, np, bb == 2 ? bit : 0u);
        if (bb != 2 && bit) mul_small<L>(x, bb);
        ew = nxt;
        --i;
    }

    // Phase 2: x = Mont(b^d) < 2*bb*N.  Pass if b^d = +-1, else square at
    // most tz - 1 times looking for -1 (redc_class reads the residue of the
    // lazy Montgomery form directly).  x lives in the park between steps.
    park_store<L>(park, x);
    if (parked_class<L>(park, nw, limbs) != 0) return 1;
    for (int c = 1; c < tz; ++c) {
        {
            uint32_t y[L], m[L];
            unpark<L>(park, nw, limbs, y, m);
            sqr_step<L, C>(y, m, neg_inv28(m[0]), 0u);
            park_store<L>(park, y);
        }
        const uint32_t k = parked_class<L>(park, nw, limbs);
        if (k == 2) return 1;
        if (k == 1) return 0;  // 1 without passing -1: composite
    }
    return 0;
}

}  // namespace r28
}  // namespace pgap
