// lucas_lg.h -- strong Lucas probable-prime test (Selfridge's method A) of
// one number per group of G lanes, radix 2^28: the second half of the
// Baillie-PSW test that the reference's CPU tester runs for every candidate
// through GMP (mpz_probab_prime_p(N, 25), gap_test_common.cpp:144,154).
// The engine runs it after a strong base-2 test, so a side's first PRP is
// the first candidate that passes Baillie-PSW, as in test_interval_cpu.
//
//   D = first of 5, -7, 9, -11, 13, ... with Jacobi(D/N) = -1 (a D that N
//       divides is skipped; any other D with Jacobi 0 shares a factor with N:
//       composite); P = 1, Q = (1 - D)/4; N + 1 = d * 2^s, d odd;
//   N passes if U_d = 0 or V_{d*2^r} = 0 (mod N) for some 0 <= r < s.
// 2 passes; other even N, N < 2 and perfect squares fail.
//
// Arithmetic (Montgomery form, exact 28-bit digits throughout).  The chain
// runs on (V_k, V_k+1, Q^k) with branch-free selects (the numbers of one
// wavefront take different exponent bits):
//   bit 0:  V_2k   = V_k^2 - 2 Q^k,      V_2k+1 = V_k V_k+1 - Q^k,   Q^2k   = Q^k (Q^k)
//   bit 1:  V_2k+1 = V_k V_k+1 - Q^k,    V_2k+2 = V_k+1^2 - 2 Q^k+1, Q^2k+1 = Q^k (Q Q^k)
// Q is small (|Q| < 2^23), so Q^(k+bit) = c Q^k with c = 1 or Q is a scalar
// multiple that is never reduced: Q^k stays below N, c Q^k below 2^24 N, and
// the V values are kept below 2^26 N by adding 2^25 N before subtracting
// 2 c Q^k.  With N < R / 2^54 (the variants' width bound) a Montgomery
// product of two such values is below 2N, so each exponent bit costs three
// products (V_k V_k+1, X^2, Q^k (c Q^k)) and linear work, and no V is ever
// reduced.  The final tests read residues through one Montgomery reduction
// of the value itself ((v + mN)/R lies in [0, N] for v < R: zero iff v = 0
// mod N).  U_d is never formed: D U_d = 2 V_d+1 - P V_d and gcd(D, N) = 1,
// so U_d = 0 iff 2 V_d+1 = V_d.
//
// The three products of a step depend only on the state before it.  With
// WAVE set, one number takes a whole wavefront: its four 16-lane DPP rows
// each hold the full state, rows 0, 1 and 2 compute V_k V_k+1, X^2 and
// Q^k (c Q^k) in one Montgomery product with per-row operands (row 3
// repeats X^2), and the rows trade their results through LDS (`xch`: 64 S
// words per wavefront, in order within the wavefront, so no barrier).  A
// step then costs one product's latency instead of three, for the strong
// Lucas tests a streaming scan waits on.
#pragma once
#include "prp_lg.h"

namespace pgap {
namespace lg {

// N mod a (a odd, a < 2^25) from the number's 32-bit words, most significant
// first; every lane of a group computes it (same loads, no divergence).
__device__ __forceinline__ uint32_t words_mod(const uint32_t *__restrict__ nw, uint32_t nwords, uint32_t a) {
    if (a < 65536u) {
        uint32_t r = 0;
        for (int i = (int)nwords - 1; i >= 0; --i) {
            const uint32_t w = nw[i];
            r = ((r << 16) | (w >> 16)) % a;
            r = ((r << 16) | (w & 0xffffu)) % a;
        }
        return r;
    }
    uint64_t r = 0;
    for (int i = (int)nwords - 1; i >= 0; --i) r = ((r << 32) | nw[i]) % a;
    return (uint32_t)r;
}

// Jacobi symbol (x/a) for odd a > 0.
__device__ __forceinline__ int jacobi(uint32_t x, uint32_t a) {
    int t = 1;
    x %= a;
    while (x) {
        const int z = __builtin_ctz(x);
        x >>= z;
        if ((z & 1) && ((a & 7u) == 3u || (a & 7u) == 5u)) t = -t;
        if ((x & 3u) == 3u && (a & 3u) == 3u) t = -t;
        const uint32_t r = a % x;
        a = x;
        x = r;
    }
    return a == 1 ? t : 0;
}

// Whether N (b bits, exact digits) is a perfect square: the bitwise integer
// square root, two bits of N per step; N is a square iff nothing remains.
template <int S, int G>
__device__ bool is_square(const uint32_t (&n)[S], uint32_t b, uint32_t g) {
    uint32_t x[S], c[S];
#pragma unroll
    for (int p = 0; p < S; ++p) {
        x[p] = n[p];
        c[p] = 0;
    }
    for (int pos = (int)((b - 1) & ~1u); pos >= 0; pos -= 2) {
        uint32_t t[S], u[S];
#pragma unroll
        for (int p = 0; p < S; ++p) t[p] = c[p];
        add_pow2<S, G>(t, (uint32_t)pos, g);  // c + 4^k
        const uint32_t borrow = sub_n<S, G>(x, t, u, g);
        shr1<S, G>(c, g);
        if (!borrow) {
#pragma unroll
            for (int p = 0; p < S; ++p) x[p] = u[p];
            add_pow2<S, G>(c, (uint32_t)pos, g);
        }
    }
    return is_zero<S, G>(x);
}

// |D| bound of the parameter search: beyond every non-square N's first
// suitable D under GRH (the least prime q with Jacobi(N/q) = -1 is below
// 2 ln(4N)^2, 2.0e7 at 4480 bits); an N reaching it is reported composite.
constexpr uint32_t kLucasMaxD = 1u << 25;
// Bits of headroom below R = 2^(28 G S) the chain needs: N < R / 2^54.
constexpr int kLucasHeadroom = 54;
// Candidates tried before the perfect-square check (a square has no D).
constexpr uint32_t kLucasSquareCheck = 32;

// p1 = V_k V_k+1, x2 = X^2, q3 = Q^k (c Q^k), each R^-1 mod N and below 2N.
template <int S, int G, bool WAVE>
__device__ __forceinline__ void lucas_products(const uint32_t (&vk)[S], const uint32_t (&vk1)[S],
                                               const uint32_t (&x)[S], const uint32_t (&qk)[S],
                                               const uint32_t (&qc)[S], uint32_t (&p1)[S], uint32_t (&x2)[S],
                                               uint32_t (&q3)[S], const uint32_t (&n)[S], uint32_t np, uint32_t g,
                                               uint32_t row, uint32_t *stage, uint32_t *xch) {
    if constexpr (WAVE) {
        uint32_t a[S], y[S];
#pragma unroll
        for (int p = 0; p < S; ++p) {
            a[p] = row == 0 ? vk[p] : (row == 2 ? qk[p] : x[p]);
            y[p] = row == 0 ? vk1[p] : (row == 2 ? qc[p] : x[p]);
        }
        mul_redc<S, G>(a, y, n, np, g, stage);
#pragma unroll
        for (int p = 0; p < S; ++p) xch[(row * S + p) * 16 + g] = a[p];
#pragma unroll
        for (int p = 0; p < S; ++p) {
            p1[p] = xch[(0 * S + p) * 16 + g];
            x2[p] = xch[(1 * S + p) * 16 + g];
            q3[p] = xch[(2 * S + p) * 16 + g];
        }
    } else {
#pragma unroll
        for (int p = 0; p < S; ++p) {
            p1[p] = vk[p];
            x2[p] = x[p];
            q3[p] = qk[p];
        }
        mul_redc<S, G>(p1, vk1, n, np, g, stage);
        sqr_redc<S, G>(x2, n, np, g, stage);
        mul_redc<S, G>(q3, qc, n, np, g, stage);
    }
}

// WAVE: `row` is the lane's 16-lane row of its wavefront and `xch` the
// wavefront's exchange area (G must be 16).
template <int S, int G, bool WAVE = false>
__device__ __forceinline__ uint32_t strong_lucas(const uint32_t *__restrict__ nw, uint32_t limbs, uint32_t g,
                                                 uint32_t *stage, uint32_t row = 0, uint32_t *xch = nullptr) {
    static_assert(!WAVE || G == 16, "a wavefront's rows are 16-lane groups");
    uint32_t n[S];
    load_digits<S, G>(nw, limbs, g, n);
    const uint32_t b = bitlen<S, G>(n, g);
    if (b < 2) return 0;
    const uint32_t n0 = gbcast0<G>(n[0]);
    if ((n0 & 1u) == 0) return b == 2 ? 1u : 0u;  // 2 passes, other even N fail
    const uint32_t nwords = (b + 31) / 32;

    // Selfridge's method A: D* = a for a = 1 mod 4, -a for a = 3 mod 4 (so
    // D* = 1 mod 4), and by reciprocity Jacobi(D*/N) = Jacobi(N/a).
    int32_t D = 0;
    for (uint32_t a = 5, k = 0; a < kLucasMaxD; a += 2, ++k) {
        if (k == kLucasSquareCheck && is_square<S, G>(n, b, g)) return 0;
        const int j = jacobi(words_mod(nw, nwords, a), a);
        if (j < 0) {
            D = (a & 2u) ? -(int32_t)a : (int32_t)a;
            break;
        }
        // Jacobi 0: N shares a factor with a; only N dividing a (N <= a <
        // 2^25, so N is its low word) is not a proof of compositeness
        if (j == 0 && !(b <= 25 && a % nw[0] == 0)) return 0;
    }
    if (D == 0) return 0;
    const int32_t Q = (1 - D) / 4;
    const uint32_t qa = Q < 0 ? (uint32_t)(-Q) : (uint32_t)Q;

    const uint32_t np = neg_inv(n0);
    uint32_t vk[S], vk1[S], qk[S];
    mont_one_two<S, G>(n, b, g, vk1, vk);  // V_1 = P = 1, V_0 = 2
#pragma unroll
    for (int p = 0; p < S; ++p) qk[p] = vk1[p];  // Q^0 = 1
    uint32_t w25[S];  // 2^25 N
    {
        uint64_t t[S];
#pragma unroll
        for (int p = 0; p < S; ++p) t[p] = (uint64_t)n[p] << 25;
        normalize<S, G>(t, w25, g);
    }

    // s = trailing ones of N; the bits of N + 1 from the top down to s are
    // N's own bits above s and a one at s (N = 2^b - 1: N + 1 = 2^b, d = 1)
    uint32_t s = 0;
    while (s < b && ((nw[s >> 5] >> (s & 31)) & 1u)) ++s;
    const int top = s == b ? (int)b : (int)b - 1;
    for (int i = top; i >= (int)s; --i) {
        const uint32_t bit = i == (int)s ? 1u : ((nw[i >> 5] >> (i & 31)) & 1u);
        const uint32_t c = bit ? qa : 1u;
        const bool neg = bit && Q < 0;  // Q^(k+bit) = -c Q^k
        uint32_t x[S], p1[S], qc[S], q3[S];
        uint64_t t[S];
#pragma unroll
        for (int p = 0; p < S; ++p) {
            x[p] = bit ? vk1[p] : vk[p];
            t[p] = (uint64_t)qk[p] * c;
        }
        normalize<S, G>(t, qc, g);                 // c Q^k < 2^24 N
        lucas_products<S, G, WAVE>(vk, vk1, x, qk, qc, p1, x, q3, n, np, g, row, stage, xch);
        // P1 = V_k V_k+1 - Q^k, as (V_k V_k+1) + N - Q^k in (0, 3N)
#pragma unroll
        for (int p = 0; p < S; ++p) t[p] = (uint64_t)p1[p] + n[p];
        normalize<S, G>(t, p1, g);
        (void)sub_n<S, G>(p1, qk, p1, g);
        // P2 = X^2 - 2 Q^(k+bit), as X^2 + 2^25 N - 2 c Q^k (or X^2 + 2 c Q^k)
        {
            uint32_t z[S], a[S];
#pragma unroll
            for (int p = 0; p < S; ++p) z[p] = qc[p];
            shl<S, G>(z, 1u, g);                   // 2 c Q^k < 2^25 N
            (void)sub_n<S, G>(w25, z, a, g);       // 2^25 N - 2 c Q^k >= 0
#pragma unroll
            for (int p = 0; p < S; ++p) t[p] = (uint64_t)x[p] + (neg ? z[p] : a[p]);
            normalize<S, G>(t, x, g);              // < 2N + 2^25 N < 2^26 N
        }
        // Q^(2k+bit) = Q^k (c Q^k), sign applied
#pragma unroll
        for (int p = 0; p < S; ++p) qk[p] = q3[p];
        cond_sub<S, G>(qk, n, g);
        if (neg) {
            uint32_t z[S];
#pragma unroll
            for (int p = 0; p < S; ++p) z[p] = 0;
            mod_sub<S, G>(z, qk, n, g);
#pragma unroll
            for (int p = 0; p < S; ++p) qk[p] = z[p];
        }
#pragma unroll
        for (int p = 0; p < S; ++p) {
            vk[p] = bit ? p1[p] : x[p];
            vk1[p] = bit ? x[p] : p1[p];
        }
    }

    // U_d = 0 iff 2 V_d+1 - V_d = 0; then V_d, V_2d, ..., V_{d 2^(s-1)}
    uint32_t e1[S];  // the integer 1: a Montgomery product with it reduces
#pragma unroll
    for (int p = 0; p < S; ++p) e1[p] = (g == 0 && p == 0) ? 1u : 0u;
    const auto zero_mod_n = [&](const uint32_t (&v)[S]) {
        uint32_t r[S];
#pragma unroll
        for (int p = 0; p < S; ++p) r[p] = v[p];
        mul_redc<S, G>(r, e1, n, np, g, stage);  // v R^-1 mod N, in [0, N]
        return is_zero<S, G>(r) || equal<S, G>(r, n);
    };
    if (zero_mod_n(vk)) return 1;
    {
        uint32_t w[S], a[S];
        uint64_t t[S];
#pragma unroll
        for (int p = 0; p < S; ++p) w[p] = w25[p];
        shl<S, G>(w, 1u, g);                       // 2^26 N > V_d
        (void)sub_n<S, G>(w, vk, a, g);
#pragma unroll
        for (int p = 0; p < S; ++p) t[p] = ((uint64_t)vk1[p] << 1) + a[p];
        normalize<S, G>(t, w, g);                  // 2 V_d+1 + 2^26 N - V_d < 2^28 N
        if (zero_mod_n(w)) return 1;
    }
    for (uint32_t r = 1; r < s; ++r) {
        // V <- V^2 - 2 Q^k, Q^k <- (Q^k)^2
        uint32_t z[S], a[S];
        uint64_t t[S];
#pragma unroll
        for (int p = 0; p < S; ++p) z[p] = qk[p];
        shl<S, G>(z, 1u, g);
        (void)sub_n<S, G>(w25, z, a, g);
        sqr_redc<S, G>(vk, n, np, g, stage);
#pragma unroll
        for (int p = 0; p < S; ++p) t[p] = (uint64_t)vk[p] + a[p];
        normalize<S, G>(t, vk, g);
        if (zero_mod_n(vk)) return 1;
        sqr_redc<S, G>(qk, n, np, g, stage);
        cond_sub<S, G>(qk, n, g);
    }
    return 0;
}

}  // namespace lg
}  // namespace pgap
