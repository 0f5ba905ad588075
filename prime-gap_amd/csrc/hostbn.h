// hostbn.h -- the little host-side multi-precision arithmetic the engine
// needs: K = P#/d, m*K, m*K +- x, ln K, residues mod small primes.  Numbers
// are little-endian vectors of 32-bit words (the layout the kernels load).
//
// Reference counterparts (GMP there): init_K (gap_common.cpp:116-121),
// _log (gap_common.cpp:100-104), mpz_mul_ui / mpz_add_ui / mpz_sub_ui in
// the batch fill (gap_test_gpu.cu:431-476).
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace pgap {

using Words = std::vector<uint32_t>;

inline void trim(Words &a) {
    while (!a.empty() && a.back() == 0) a.pop_back();
}

inline uint32_t bitlen(const Words &a) {
    for (size_t i = a.size(); i-- > 0;)
        if (a[i]) return 32u * (uint32_t)i + (32u - (uint32_t)__builtin_clz(a[i]));
    return 0;
}

inline void mul_small(Words &a, uint64_t m) {
    // a *= m for m < 2^64, via two 32-bit halves
    const uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
    Words r(a.size() + 2, 0);
    for (int half = 0; half < 2; ++half) {
        const uint64_t f = half ? hi : lo;
        if (!f) continue;
        uint64_t carry = 0;
        for (size_t i = 0; i < a.size(); ++i) {
            unsigned __int128 t = (unsigned __int128)a[i] * f + r[i + half] + carry;
            r[i + half] = (uint32_t)t;
            carry = (uint64_t)(t >> 32);
        }
        for (size_t i = a.size() + half; carry; ++i) {
            uint64_t t = (uint64_t)r[i] + carry;
            r[i] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    trim(r);
    a.swap(r);
}

// a /= d; returns the remainder.
inline uint32_t div_small(Words &a, uint32_t d) {
    uint64_t rem = 0;
    for (size_t i = a.size(); i-- > 0;) {
        uint64_t cur = (rem << 32) | a[i];
        a[i] = (uint32_t)(cur / d);
        rem = cur % d;
    }
    trim(a);
    return (uint32_t)rem;
}

inline uint32_t mod_small(const Words &a, uint32_t p) {
    uint64_t rem = 0;
    for (size_t i = a.size(); i-- > 0;) rem = ((rem << 32) | a[i]) % p;
    return (uint32_t)rem;
}

// a += x (x may be negative); returns false if the result would be negative.
inline bool add_signed(Words &a, int64_t x) {
    if (x >= 0) {
        uint64_t c = (uint64_t)x;
        for (size_t i = 0; c; ++i) {
            if (i == a.size()) a.push_back(0);
            uint64_t t = (uint64_t)a[i] + (c & 0xffffffffu);
            a[i] = (uint32_t)t;
            c = (c >> 32) + (t >> 32);
        }
        return true;
    }
    uint64_t b = (uint64_t)(-(x + 1)) + 1u;
    for (size_t i = 0; b; ++i) {
        if (i == a.size()) return false;
        uint64_t sub = b & 0xffffffffu;
        uint64_t cur = a[i];
        a[i] = (uint32_t)(cur - sub);
        b = (b >> 32) + (cur < sub ? 1 : 0);
    }
    trim(a);
    return true;
}

inline std::vector<uint32_t> primes_upto(uint32_t n) {
    std::vector<uint32_t> ps;
    if (n < 2) return ps;
    std::vector<uint8_t> comp(n / 2 + 1, 0);
    ps.push_back(2);
    for (uint64_t i = 1; 2 * i + 1 <= n; ++i) {
        if (comp[i]) continue;
        const uint64_t p = 2 * i + 1;
        ps.push_back((uint32_t)p);
        for (uint64_t j = p * p / 2; j <= n / 2; j += p) comp[j] = 1;
    }
    return ps;
}

// K = P# / d; false if d does not divide P# or K <= 1 (init_K asserts both,
// gap_common.cpp:117-120).
inline bool primorial_over(uint32_t P, uint32_t D, Words &K) {
    K.assign(1, 1);
    for (uint32_t p : primes_upto(P)) mul_small(K, p);
    if (D == 0) return false;
    if (div_small(K, D) != 0) return false;
    return bitlen(K) > 1;
}

// ln K the way gap_common.cpp:100-104 computes it: mpz_get_d_2exp gives the
// mantissa truncated to 53 bits in [0.5, 1) and the exponent.
inline double ln_words(const Words &K) {
    const uint32_t b = bitlen(K);
    if (b == 0) return -INFINITY;
    // top 64 bits, left-aligned
    uint64_t top = 0;
    for (int k = 0; k < 64; ++k) {
        const int64_t bit = (int64_t)b - 1 - k;
        if (bit < 0) break;
        const uint32_t w = K[(size_t)bit >> 5];
        top |= (uint64_t)((w >> (bit & 31)) & 1u) << (63 - k);
    }
    const double mant = std::ldexp((double)(top >> 11), -53);
    return std::log(mant) + std::log(2) * (double)b;
}

}  // namespace pgap
