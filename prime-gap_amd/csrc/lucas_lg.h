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
// Arithmetic: every value is held fully reduced (exact digits, < N) in
// Montgomery form, so the products use mul_redc / sqr_redc followed by one
// conditional subtraction, and the additive steps are modular.  The chain
// runs on (V_k, V_k+1, Q^k) with branch-free selects (numbers of one
// wavefront take different exponent bits):
//   bit 0:  V_2k   = V_k^2 - 2 Q^k,         V_2k+1 = V_k V_k+1 - Q^k,   Q^2k   = (Q^k)^2
//   bit 1:  V_2k+1 = V_k V_k+1 - Q^k,       V_2k+2 = V_k+1^2 - 2 Q^k+1, Q^2k+1 = Q^k Q^k+1
// i.e. four Montgomery products per exponent bit.  U_d is never formed:
// D U_d = 2 V_d+1 - P V_d and gcd(D, N) = 1, so U_d = 0 iff 2 V_d+1 = V_d.
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
// Candidates tried before the perfect-square check (a square has no D).
constexpr uint32_t kLucasSquareCheck = 32;

template <int S, int G>
__device__ __forceinline__ uint32_t strong_lucas(const uint32_t *__restrict__ nw, uint32_t limbs, uint32_t g,
                                                 uint32_t *stage) {
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

    const uint32_t np = neg_inv(n0);
    uint32_t one[S], vk[S];
    mont_one_two<S, G>(n, b, g, one, vk);  // V_0 = 2
    uint32_t qm[S];                        // Mont(Q)
    {
        const uint32_t qa = Q < 0 ? (uint32_t)(-Q) : (uint32_t)Q;
#pragma unroll
        for (int p = 0; p < S; ++p) qm[p] = 0;
        for (int t = 31 - __builtin_clz(qa); t >= 0; --t) {
            mod_double<S, G>(qm, n, g);
            if ((qa >> t) & 1u) mod_add<S, G>(qm, one, n, g);
        }
        if (Q < 0) {
            uint32_t z[S];
#pragma unroll
            for (int p = 0; p < S; ++p) z[p] = 0;
            mod_sub<S, G>(z, qm, n, g);
#pragma unroll
            for (int p = 0; p < S; ++p) qm[p] = z[p];
        }
    }
    uint32_t vk1[S], qk[S];  // V_1 = P = 1, Q^0 = 1
#pragma unroll
    for (int p = 0; p < S; ++p) {
        vk1[p] = one[p];
        qk[p] = one[p];
    }

    // s = trailing ones of N; the bits of N + 1 from the top down to s are
    // N's own bits above s and a one at s (N = 2^b - 1: N + 1 = 2^b, d = 1)
    uint32_t s = 0;
    while (s < b && ((nw[s >> 5] >> (s & 31)) & 1u)) ++s;
    const int top = s == b ? (int)b : (int)b - 1;
    for (int i = top; i >= (int)s; --i) {
        const uint32_t bit = i == (int)s ? 1u : ((nw[i >> 5] >> (i & 31)) & 1u);
        uint32_t x[S], p1[S], qx[S], q2[S];
#pragma unroll
        for (int p = 0; p < S; ++p) {
            x[p] = bit ? vk1[p] : vk[p];
            p1[p] = vk[p];
            qx[p] = qk[p];
        }
        mul_redc<S, G>(p1, vk1, n, np, g, stage);  // V_k V_k+1 - Q^k
        cond_sub<S, G>(p1, n, g);
        mod_sub<S, G>(p1, qk, n, g);
        mul_redc<S, G>(qx, qm, n, np, g, stage);   // Q^k+1 when the bit is set
        cond_sub<S, G>(qx, n, g);
#pragma unroll
        for (int p = 0; p < S; ++p) {
            qx[p] = bit ? qx[p] : qk[p];
            q2[p] = qx[p];
        }
        sqr_redc<S, G>(x, n, np, g, stage);        // X^2 - 2 Q^(k+bit)
        cond_sub<S, G>(x, n, g);
        mod_double<S, G>(q2, n, g);
        mod_sub<S, G>(x, q2, n, g);
        mul_redc<S, G>(qk, qx, n, np, g, stage);   // Q^(2k+bit)
        cond_sub<S, G>(qk, n, g);
#pragma unroll
        for (int p = 0; p < S; ++p) {
            vk[p] = bit ? p1[p] : x[p];
            vk1[p] = bit ? x[p] : p1[p];
        }
    }

    // U_d = 0 iff 2 V_d+1 = V_d; then V_d, V_2d, ..., V_{d 2^(s-1)}
    uint32_t w[S];
#pragma unroll
    for (int p = 0; p < S; ++p) w[p] = vk1[p];
    mod_double<S, G>(w, n, g);
    if (is_zero<S, G>(vk) || equal<S, G>(w, vk)) return 1;
    for (uint32_t r = 1; r < s; ++r) {
        uint32_t q2[S];
#pragma unroll
        for (int p = 0; p < S; ++p) q2[p] = qk[p];
        mod_double<S, G>(q2, n, g);
        sqr_redc<S, G>(vk, n, np, g, stage);
        cond_sub<S, G>(vk, n, g);
        mod_sub<S, G>(vk, q2, n, g);
        if (is_zero<S, G>(vk)) return 1;
        sqr_redc<S, G>(qk, n, np, g, stage);
        cond_sub<S, G>(qk, n, g);
    }
    return 0;
}

}  // namespace lg
}  // namespace pgap
