// mont28.h -- one-number-per-lane Montgomery arithmetic in radix 2^28.
//
// A number of up to 28*L bits is held as L digits of 28 bits, each in a
// 32-bit VGPR.  A 28x28-bit product is < 2^56, so a column of up to 255
// products accumulates in a 64-bit register with plain v_mad_u64_u32 (the
// 64-bit addend folded in) and never needs a carry-out: one VALU slot per
// product.  With 32-bit digits the same column needs v_mad_u64_u32's carry-out
// plus a v_addc per product (two slots), and on gfx950 v_mad_u64_u32 issues at
// full rate, so the carry add costs as much as the multiply.
//
// Squaring is finely integrated product scanning (FIPS): column k of a^2 (the
// cross products once, doubled, plus the diagonal) and of m*N are summed in the
// same accumulator, m_k = (col mod 2^28) * n' is formed while the low columns
// run, and the high columns are the result.  Lazy reduction: with N < R/16,
// R = 2^(28L), an input < 4N gives an output < 2N, so no conditional
// subtraction runs inside the exponentiation.
//
// Reference counterpart: CGBN's cgbn_mont_sqr / cgbn_mont_mul as called by
// miller_rabin_t::powm (miller_rabin.h:157-211); the arithmetic here is
// written from the Montgomery/FIPS method, not from CGBN.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgap {
namespace r28 {

constexpr uint32_t MASK = (1u << 28) - 1;

__device__ __forceinline__ uint64_t mad(uint32_t a, uint32_t b, uint64_t c) {
    return (uint64_t)a * b + c;
}

// -N^{-1} mod 2^28 for odd n0 (Newton iteration, each step doubles the
// number of correct low bits: 3 -> 6 -> 12 -> 24 -> 48).
__device__ __forceinline__ uint32_t neg_inv28(uint32_t n0) {
    uint32_t x = n0;                 // n0 * n0 == 1 mod 8
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    return (0u - x) & MASK;
}

// Convert W little-endian 32-bit words to L radix-2^28 digits (bits above
// 28*L must be zero; the caller guarantees it by bucket choice).
template <int L, int W>
__device__ __forceinline__ void from_words(const uint32_t (&w)[W], uint32_t (&d)[L]) {
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const int bit = 28 * j;
        const int wi = bit >> 5, sh = bit & 31;
        uint32_t lo = wi < W ? w[wi] : 0u;
        uint32_t hi = (wi + 1) < W ? w[wi + 1] : 0u;
        uint32_t v = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
        d[j] = v & MASK;
    }
}

// Bit length of the value held in digits (0 for zero).
template <int L>
__device__ __forceinline__ int bitlen(const uint32_t (&d)[L]) {
    int b = 0;
#pragma unroll
    for (int j = 0; j < L; ++j)
        if (d[j]) b = 28 * j + (32 - __builtin_clz(d[j]));
    return b;
}

// y = 2y mod N for y < N (digits normalised).
template <int L>
__device__ __forceinline__ void mod_double(uint32_t (&y)[L], const uint32_t (&n)[L]) {
    uint32_t t[L];
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        uint32_t v = (y[j] << 1) | c;
        c = v >> 28;
        t[j] = v & MASK;
    }
    // t - N; keep it when no borrow (or when 2y overflowed the top digit).
    uint32_t br = 0;
    uint32_t u[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        uint32_t v = t[j] - n[j] - br;
        br = v >> 31;
        u[j] = v & MASK;
    }
    const bool take = (br == 0) || (c != 0);
#pragma unroll
    for (int j = 0; j < L; ++j) y[j] = take ? u[j] : t[j];
}

// y = y mod N for y < 2N.
template <int L>
__device__ __forceinline__ void cond_sub(uint32_t (&y)[L], const uint32_t (&n)[L]) {
    uint32_t br = 0;
    uint32_t u[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        uint32_t v = y[j] - n[j] - br;
        br = v >> 31;
        u[j] = v & MASK;
    }
#pragma unroll
    for (int j = 0; j < L; ++j) y[j] = br ? y[j] : u[j];
}

template <int L>
__device__ __forceinline__ bool equal(const uint32_t (&a)[L], const uint32_t (&b)[L]) {
    uint32_t x = 0;
#pragma unroll
    for (int j = 0; j < L; ++j) x |= a[j] ^ b[j];
    return x == 0;
}

// Classifies x * R^{-1} mod N for a Montgomery-form x < R (every lazy bound
// used here qualifies: x < 4N with N < R/16, x < 2bN with N < R/(4b^2)): 1 if
// it is 1, 2 if it is N - 1, 0 otherwise.  This is the Montgomery reduction
// of x itself, (x + mN)/R, which is congruent to x R^{-1} and lies in [0, N]
// because x < R and m < R, so the two residues are recognised by their
// digits alone: no Mont(1) or Mont(-1) array is held or rebuilt, and nothing
// beyond x, N and one working copy is live.
//
// Row-wise: each row adds m*N with m chosen to clear the lowest position,
// then drops that position (the window w[] moves down one digit per row, so
// after L rows it holds the result).  A carry digit may reach 2^28 until the
// next row normalises it.  The row loop is not unrolled: L^2 multiply-adds
// once per test (or per checking squaring of a strong PRP) cost nothing
// measurable, and one copy of a row keeps the unrolled exponentiation's
// register allocation undisturbed.
template <int L>
__device__ __forceinline__ uint32_t redc_class(const uint32_t (&x)[L], const uint32_t (&n)[L], uint32_t np) {
    uint32_t w[L];
#pragma unroll
    for (int j = 0; j < L; ++j) w[j] = x[j];
#pragma unroll 1
    for (int i = 0; i < L; ++i) {
        const uint32_t m = (w[0] * np) & MASK;
        uint64_t c = mad(m, n[0], w[0]) >> 28;  // the low 28 bits cancel
#pragma unroll
        for (int j = 1; j < L; ++j) {
            const uint64_t v = mad(m, n[j], (uint64_t)w[j] + c);
            // the digit is cut off before the carry moves on, through an
            // opaque value: otherwise the cut is deferred to the end of the
            // row and every 64-bit sum of the row stays live until then
            uint32_t d = (uint32_t)v & MASK;
            asm("" : "+v"(d));
            w[j - 1] = d;
            c = (v - d) >> 28;
        }
        w[L - 1] = (uint32_t)c;
    }
    uint32_t d1 = w[0] ^ 1u, dm = w[0] ^ (n[0] - 1u);  // N odd: N - 1 borrows nothing
#pragma unroll
    for (int j = 1; j < L; ++j) {
        d1 |= w[j];
        dm |= w[j] ^ n[j];
    }
    return d1 == 0 ? 1u : (dm == 0 ? 2u : 0u);
}

// a <- a^2 * R^{-1} * 2^s mod N (s in {0,1}), lazy: a < 4N in, a < 4N out.
// The 2^s factor is folded into the high (result) columns, so the doubling
// step of a base-2 exponentiation costs two VALU slots per digit of output.
// One dependent chain per column: the narrow variants keep two or more
// wavefronts per SIMD, which hide the multiply-add latency (splitting the
// column into independent chains measured 1% slower at 40 digits).
template <int L>
__device__ __forceinline__ void sqr_redc(uint32_t (&a)[L], const uint32_t (&n)[L], uint32_t np,
                                         uint32_t s) {
    static_assert(L >= 2 && L <= 63, "radix-2^28 column bound");
    uint32_t m[L];
    uint64_t T = 0;
#pragma unroll
    for (int k = 0; k < 2 * L - 1; ++k) {
        const int ilo = k - (L - 1) > 0 ? k - (L - 1) : 0;
        // cross products a_i * a_{k-i}, i < k-i
        uint64_t X = 0;
#pragma unroll
        for (int i = ilo; 2 * i < k; ++i) X = mad(a[i], a[k - i], X);
        if (k < L) {
            T = (X << 1) + T;
            if ((k & 1) == 0) T = mad(a[k >> 1], a[k >> 1], T);
#pragma unroll
            for (int i = ilo; i < k; ++i) T = mad(m[i], n[k - i], T);
            m[k] = ((uint32_t)T * np) & MASK;
            T = mad(m[k], n[0], T);
            T >>= 28;
        } else {
            uint64_t Z = 0;
            if ((k & 1) == 0) Z = mad(a[k >> 1], a[k >> 1], Z);
#pragma unroll
            for (int i = ilo; i < L; ++i) Z = mad(m[i], n[k - i], Z);
            Z = (X << 1) + Z;
            if (k == L) T <<= s;
            T = (Z << s) + T;
            // digit k-L of the result; a[k-L] and m[k-L] are dead here
            a[k - L] = (uint32_t)T & MASK;
            T >>= 28;
        }
    }
    a[L - 1] = (uint32_t)T & MASK;
}

// Wide variant (L up to 127): the 2^s factor is applied to the finished
// digits (a one-bit shift with carry) instead of to the column sums, whose
// extra bit would break the 64-bit column bound above 63 digits.
//
// C = chains per column.  C = 1: one dependent chain, as in the narrow form,
// for kernels that keep two wavefronts per SIMD (78 digits, parked check,
// 2^18 Fermat tests: 161.7 ms against 166.7 ms with three chains).  C = 3:
// the cross products and the m*N products in two halves summed in
// independent chains, issued round-robin and merged once per column, for
// kernels that run one wavefront per SIMD, where nothing else hides the
// latency of a dependent multiply-add (1511# streaming scan: 11.3 k m/s
// against 10.0 k with one chain; persistent mixed-width batch 2.88 M tests/s
// against 2.80 M).  Exact integer sums: the result does not depend on C.
template <int L, int C = 1>
__device__ __forceinline__ void sqr_redc_wide(uint32_t (&a)[L], const uint32_t (&n)[L], uint32_t np,
                                              uint32_t s) {
    static_assert(L >= 2 && L <= 127, "radix-2^28 column bound");
    static_assert(C == 1 || C == 3, "one or three chains per column");
    uint32_t m[L];
    uint64_t T = 0;
#pragma unroll
    for (int k = 0; k < 2 * L - 1; ++k) {
        const int ilo = k - (L - 1) > 0 ? k - (L - 1) : 0;
        const int ye = k < L ? k : L;  // m_i n_{k-i}, i in [ilo, ye) (m_k joins below)
        if constexpr (C == 1) {
            // cross products a_i * a_{k-i}, i < k-i
            uint64_t X = 0;
#pragma unroll
            for (int i = ilo; 2 * i < k; ++i) X = mad(a[i], a[k - i], X);
            T = (X << 1) + T;
            if ((k & 1) == 0) T = mad(a[k >> 1], a[k >> 1], T);
#pragma unroll
            for (int i = ilo; i < ye; ++i) T = mad(m[i], n[k - i], T);
        } else {
            const int xe = (k + 1) / 2;         // cross products a_i a_{k-i}, i in [ilo, xe)
            const int ym = (ilo + ye + 1) / 2;  // split point of the m*N chain
            const int nx = xe - ilo, n0 = ym - ilo, n1 = ye - ym;
            const int rounds = nx > n0 ? (nx > n1 ? nx : n1) : (n0 > n1 ? n0 : n1);
            uint64_t X = 0, Y0 = T, Y1 = 0;
#pragma unroll
            for (int r = 0; r < rounds; ++r) {
                if (r < nx) X = mad(a[ilo + r], a[k - ilo - r], X);
                if (r < n0) Y0 = mad(m[ilo + r], n[k - ilo - r], Y0);
                if (r < n1) Y1 = mad(m[ym + r], n[k - ym - r], Y1);
            }
            // opaque to the optimiser: otherwise reassociation folds the three
            // sums back into one dependent chain
            asm("" : "+v"(X), "+v"(Y0), "+v"(Y1));
            T = (X << 1) + Y0;
            T += Y1;
            if ((k & 1) == 0) T = mad(a[k >> 1], a[k >> 1], T);
        }
        if (k < L) {
            m[k] = ((uint32_t)T * np) & MASK;
            T = mad(m[k], n[0], T);
            T >>= 28;
        } else {
            // digit k-L of the result; a[k-L] and m[k-L] are dead here
            a[k - L] = (uint32_t)T & MASK;
            T >>= 28;
        }
    }
    a[L - 1] = (uint32_t)T & MASK;
    // a <- a * 2^s (a < 2N < R/8, so the shifted value still fits L digits)
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const uint32_t v = (a[j] << s) | c;
        c = v >> 28;
        a[j] = v & MASK;
    }
}

}  // namespace r28
}  // namespace pgap
