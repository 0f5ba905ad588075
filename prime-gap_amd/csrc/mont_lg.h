// mont_lg.h -- Montgomery arithmetic on numbers spread over a quad of lanes
// (4 lanes per number), radix 2^26.
//
// Used where one lane cannot hold a number (above ~2240 bits the three
// L-digit arrays of the one-lane kernel no longer fit the register file).
// Lane g of a quad owns digits [g*S, (g+1)*S) of x and N; L = 4*S.
//
// Squaring is row-wise Montgomery (CIOS): for i = 0..L-1
//     T += x_i * x;   m = T_0 * n' mod 2^26;   T += m * N;   T /= 2^26
// x_i and m are broadcast inside the quad with DPP quad_perm (a modifier on
// a v_mov, no LDS).  T is kept as 64-bit unnormalised digit accumulators; a
// 26x26-bit product is < 2^52 and each accumulator takes two per row for at
// most L = 164 rows, < 2^61, so no carry moves until the row loop ends.  The
// division by 2^26 moves every accumulator down one position; positions are
// mapped to registers with a rotating offset (position p of row i lives in
// register (i + p) mod S), so the shift costs one 64-bit DPP move from the
// next lane, not S register moves.  The multiplier digits x_i come from a
// copy of x that rotates one lane down after every S rows, so the broadcast
// source is always lane 0 of the quad and the S-row body is unrolled once.
//
// Reference counterpart: CGBN's warp-cooperative cgbn_mont_sqr with
// TPI threads per instance (miller_rabin.h:157-211, TPI = 8 in
// gap_test_gpu.cu:80); written here from the CIOS method for AMD quads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgap {
namespace lg {

constexpr int RB = 26;
constexpr uint32_t MASK = (1u << RB) - 1;

__device__ __forceinline__ uint64_t mad(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }

// Lane g of quad receives lane SRC's value.
template <int P0, int P1, int P2, int P3>
__device__ __forceinline__ uint32_t qperm(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, P0 | (P1 << 2) | (P2 << 4) | (P3 << 6), 0xF, 0xF, false);
}
template <int SRC>
__device__ __forceinline__ uint32_t qbcast(uint32_t v) { return qperm<SRC, SRC, SRC, SRC>(v); }
// lane g <- lane g+1 (lane 3 keeps its own value; callers mask it)
__device__ __forceinline__ uint32_t qdown(uint32_t v) { return qperm<1, 2, 3, 3>(v); }
// lane g <- lane g-1 (lane 0 keeps its own value; callers mask it)
__device__ __forceinline__ uint32_t qup(uint32_t v) { return qperm<0, 0, 1, 2>(v); }
__device__ __forceinline__ uint64_t qdown64(uint64_t v) {
    return (uint64_t)qdown((uint32_t)v) | ((uint64_t)qdown((uint32_t)(v >> 32)) << 32);
}
__device__ __forceinline__ uint64_t qup64(uint64_t v) {
    return (uint64_t)qup((uint32_t)v) | ((uint64_t)qup((uint32_t)(v >> 32)) << 32);
}
__device__ __forceinline__ uint32_t qmax(uint32_t v) {
    v = max(v, qperm<1, 0, 3, 2>(v));
    return max(v, qperm<2, 3, 0, 1>(v));
}
__device__ __forceinline__ uint32_t qor(uint32_t v) {
    v |= qperm<1, 0, 3, 2>(v);
    return v | qperm<2, 3, 0, 1>(v);
}

__device__ __forceinline__ uint32_t quad_lane() { return threadIdx.x & 3u; }

__device__ __forceinline__ uint32_t neg_inv(uint32_t n0) {
    uint32_t x = n0;
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    return (0u - x) & MASK;
}

// Carry-normalise unnormalised accumulators t (this lane's S positions) into
// digits; the carry crosses lanes in three rounds.
template <int S>
__device__ __forceinline__ void normalize(const uint64_t (&t)[S], uint32_t (&x)[S], uint32_t g) {
    uint64_t c = 0;
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const uint64_t v = t[p] + c;
        x[p] = (uint32_t)v & MASK;
        c = v >> RB;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        uint64_t cin = qup64(c);
        if (g == 0) cin = 0;
        c = 0;
        uint64_t carry = cin;
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const uint64_t v = (uint64_t)x[p] + carry;
            x[p] = (uint32_t)v & MASK;
            carry = v >> RB;
        }
        c = carry;
    }
}

// y <- y - n over the quad; returns the final borrow (1 if y < n), valid in
// every lane.  Digits of y are left as y - n mod R (caller selects).
template <int S>
__device__ __forceinline__ uint32_t sub_n(const uint32_t (&y)[S], const uint32_t (&n)[S], uint32_t (&u)[S],
                                          uint32_t g) {
    uint32_t br = 0;
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const uint32_t v = y[p] - n[p] - br;
        br = v >> 31;
        u[p] = v & MASK;
    }
    // Borrows ripple up one lane per round.  A lane that already borrowed
    // locally holds u >= 1 in its own digits, so an incoming borrow cannot
    // make it borrow twice: only newly caused borrows travel on.
    uint32_t total = br, pending = br;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        uint32_t bin = qup(pending);
        if (g == 0) bin = 0;
        uint32_t b2 = bin;
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const uint32_t v = u[p] - b2;
            b2 = v >> 31;
            u[p] = v & MASK;
        }
        pending = b2;
        total |= b2;
    }
    return qbcast<3>(total);
}

// y <- y mod N for y < 2N.
template <int S>
__device__ __forceinline__ void cond_sub(uint32_t (&y)[S], const uint32_t (&n)[S], uint32_t g) {
    uint32_t u[S];
    const uint32_t borrow = sub_n<S>(y, n, u, g);
#pragma unroll
    for (int p = 0; p < S; ++p) y[p] = borrow ? y[p] : u[p];
}

// y <- y * 2^s (s in {0,1}), value stays below R.
template <int S>
__device__ __forceinline__ void shl(uint32_t (&y)[S], uint32_t s, uint32_t g) {
    uint32_t top = y[S - 1] >> (RB - 1);
    uint32_t cin = qup(top);
    if (g == 0) cin = 0;
#pragma unroll
    for (int p = S - 1; p >= 1; --p) y[p] = ((y[p] << s) | ((y[p - 1] >> (RB - 1)) & s)) & MASK;
    y[0] = ((y[0] << s) | (cin & s)) & MASK;
}

// y <- 2y mod N for y < N.
template <int S>
__device__ __forceinline__ void mod_double(uint32_t (&y)[S], const uint32_t (&n)[S], uint32_t g) {
    shl<S>(y, 1u, g);
    cond_sub<S>(y, n, g);
}

template <int S>
__device__ __forceinline__ bool equal(const uint32_t (&a)[S], const uint32_t (&b)[S]) {
    uint32_t d = 0;
#pragma unroll
    for (int p = 0; p < S; ++p) d |= a[p] ^ b[p];
    return qor(d) == 0;
}

// x <- x^2 R^{-1} mod N (lazy: x < 4N in, < 2N out), R = 2^(26*4S).
template <int S>
__device__ __forceinline__ void sqr_redc(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t np, uint32_t g) {
    uint64_t t[S];
    uint32_t xs[S];
#pragma unroll
    for (int p = 0; p < S; ++p) {
        t[p] = 0;
        xs[p] = x[p];
    }
    const bool lane0 = (g == 0), lane3 = (g == 3);
    for (int o = 0; o < 4; ++o) {
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const uint32_t xi = qbcast<0>(xs[u]);
#pragma unroll
            for (int j = 0; j < S; ++j) t[(u + j) % S] = mad(xi, x[j], t[(u + j) % S]);
            const uint32_t m = qbcast<0>(((uint32_t)t[u] * np) & MASK);
#pragma unroll
            for (int j = 0; j < S; ++j) t[(u + j) % S] = mad(m, n[j], t[(u + j) % S]);
            const uint64_t in = qdown64(t[u]);
            const uint64_t c = lane0 ? (t[u] >> RB) : 0ull;
            t[(u + 1) % S] += c;
            t[u] = lane3 ? 0ull : in;
        }
        // the next S rows broadcast the next lane's digits
#pragma unroll
        for (int u = 0; u < S; ++u) xs[u] = qdown(xs[u]);
    }
    normalize<S>(t, x, g);
}

}  // namespace lg
}  // namespace pgap
