// mont_lg.h -- Montgomery arithmetic on numbers spread over a group of G
// lanes (G = 4, 8 or 16 lanes per number), radix 2^28.
//
// Used where one lane cannot hold a number (above ~2240 bits the three
// L-digit arrays of the one-lane kernel no longer fit the register file).
// Lane g of a group owns digits [g*S, (g+1)*S) of x and N; L = G*S.
//
// Squaring is row-wise Montgomery (CIOS): for i = 0..L-1
//     T += x_i * x;   m = T_0 * n' mod 2^28;   T += m * N;   T /= 2^28
// m is broadcast inside the group with DPP (a modifier on a v_mov): quad_perm
// within a quad, plus a row shift under a bank mask for groups of 8; the row
// multipliers x_i are read from an LDS stage (see sqr_redc).  T is kept as
// 64-bit unnormalised digit accumulators.  The division by 2^28 moves every
// accumulator down one position; positions are mapped to registers with a
// rotating offset (position p of row i lives in register (i + p) mod S), so
// the shift moves only each lane's bottom accumulator to the lane below, not
// S registers.  Before it moves, every lane adds that accumulator's bits
// above 28 to its own next position, so what crosses the lane boundary is a
// 28-bit digit: one 32-bit DPP move per row (see sqr_redc for the bound).
// The S rows that multiply by one lane's digits share one unrolled body.
//
// A 4247-bit number (152 digits) runs on G = 4 lanes of S = 38 digits at
// two wavefronts per SIMD (x, N and the accumulators take 152 registers per
// lane; sqr_redc keeps its rows in order, see there), or on G = 8 lanes of
// 19 digits at three.  The quad issues 17% fewer instructions per test (one
// m broadcast per 76 products instead of per 38) and is the faster of the
// two.  G = 16 (a whole DPP row) cuts a test's latency at a lower
// throughput; the streaming scan uses it while few m's are in flight.
//
// Reference counterpart: CGBN's warp-cooperative cgbn_mont_sqr with
// TPI threads per instance (miller_rabin.h:157-211, TPI = 8 in
// gap_test_gpu.cu:80); written here from the CIOS method for AMD quads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgap {
namespace lg {

constexpr int RB = 28;
constexpr uint32_t MASK = (1u << RB) - 1;

__device__ __forceinline__ uint64_t mad(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }

// DPP control codes (GFX9 DPP16; row_newbcast from gfx90a on).
constexpr int DPP_ROW_SHL1 = 0x101, DPP_ROW_SHL4 = 0x104, DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR4 = 0x114,
              DPP_ROW_MIRROR = 0x140, DPP_ROW_HALF_MIRROR = 0x141, DPP_ROW_NEWBCAST0 = 0x150,
              DPP_ROW_NEWBCAST15 = 0x15F;

// Lane g of quad receives lane SRC's value.
template <int P0, int P1, int P2, int P3>
__device__ __forceinline__ uint32_t qperm(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, P0 | (P1 << 2) | (P2 << 4) | (P3 << 6), 0xF, 0xF, false);
}
template <int SRC>
__device__ __forceinline__ uint32_t qbcast(uint32_t v) { return qperm<SRC, SRC, SRC, SRC>(v); }

// Group lane 0's value to every lane of the group.  For G = 8 the quad
// broadcast is followed by a 4-lane shift written only to the upper quad of
// each group (banks 1 and 3 of the 16-lane row); a group of 16 is the whole
// DPP row, one row_newbcast.
template <int G>
__device__ __forceinline__ uint32_t gbcast0(uint32_t v) {
    if constexpr (G == 16) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, DPP_ROW_NEWBCAST0, 0xF, 0xF, false);
    const uint32_t q = qbcast<0>(v);
    if constexpr (G == 4) return q;
    return (uint32_t)__builtin_amdgcn_update_dpp((int)q, (int)q, DPP_ROW_SHR4, 0xF, 0xA, false);
}
// Group lane G-1's value to every lane of the group.
template <int G>
__device__ __forceinline__ uint32_t gbcast_top(uint32_t v) {
    if constexpr (G == 16) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, DPP_ROW_NEWBCAST15, 0xF, 0xF, false);
    const uint32_t q = qbcast<3>(v);
    if constexpr (G == 4) return q;
    return (uint32_t)__builtin_amdgcn_update_dpp((int)q, (int)q, DPP_ROW_SHL4, 0xF, 0x5, false);
}
// lane g <- lane g-1 (lane 0's value is undefined; callers mask it)
template <int G>
__device__ __forceinline__ uint32_t gup(uint32_t v) {
    if constexpr (G == 4) return qperm<0, 0, 1, 2>(v);
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_SHR1, 0xF, 0xF, false);
}
template <int G>
__device__ __forceinline__ uint64_t gup64(uint64_t v) {
    return (uint64_t)gup<G>((uint32_t)v) | ((uint64_t)gup<G>((uint32_t)(v >> 32)) << 32);
}
// Reductions over the group: butterflies within the quad, then (G >= 8) the
// mirror of the 8-lane half-row, which pairs every lane with the other quad,
// then (G = 16) the mirror of the row, which pairs the two halves.
template <int G>
__device__ __forceinline__ uint32_t gmax(uint32_t v) {
    v = max(v, qperm<1, 0, 3, 2>(v));
    v = max(v, qperm<2, 3, 0, 1>(v));
    if constexpr (G >= 8)
        v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_HALF_MIRROR, 0xF, 0xF, false));
    if constexpr (G == 16)
        v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_MIRROR, 0xF, 0xF, false));
    return v;
}
template <int G>
__device__ __forceinline__ uint32_t gor(uint32_t v) {
    v |= qperm<1, 0, 3, 2>(v);
    v |= qperm<2, 3, 0, 1>(v);
    if constexpr (G >= 8) v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_HALF_MIRROR, 0xF, 0xF, false);
    if constexpr (G == 16) v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_MIRROR, 0xF, 0xF, false);
    return v;
}

__device__ __forceinline__ uint32_t neg_inv(uint32_t n0) {
    uint32_t x = n0;
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    x *= 2u - n0 * x;
    return (0u - x) & MASK;
}

// lane g <- lane g - SH of the same 16-lane DPP row (lanes below SH read 0).
template <int SH>
__device__ __forceinline__ uint32_t row_shr(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 + SH, 0xF, 0xF, true);
}

// One-bit carries (or borrows) across the group by carry lookahead: lane g
// generates `gen` (a carry out of its digits) and propagates (`prop`: a
// carry into it would pass through); a Kogge-Stone scan over the lanes
// (log2 G DPP steps instead of G - 1 rounds over S digits) gives each lane
// its carry in.
template <int G, int SH>
__device__ __forceinline__ void lookahead_step(uint32_t &gp, uint32_t g) {
    if constexpr (SH < G) {
        uint32_t b = row_shr<SH>(gp);
        if (g < (uint32_t)SH) b = 2u;  // lanes of the group below lane 0: gen 0, prop 1
        gp = ((gp | ((gp >> 1) & b)) & 1u) | (gp & b & 2u);
        lookahead_step<G, SH * 2>(gp, g);
    }
}
template <int G>
__device__ __forceinline__ uint32_t carry_in(uint32_t gen, uint32_t prop, uint32_t g) {
    uint32_t gp = gen | (prop << 1);
    lookahead_step<G, 1>(gp, g);  // gp & 1: a carry leaves lanes [0, g]
    const uint32_t cin = gup<G>(gp & 1u);
    return g == 0 ? 0u : cin;
}

// Carry-normalise unnormalised accumulators t (this lane's S positions) into
// digits: each lane carries through its own positions, adds the carry out
// of the lane below (after which at most one bit leaves any lane, S >= 2),
// and the one-bit carries cross the group by lookahead.
template <int S, int G>
__device__ __forceinline__ void normalize(const uint64_t (&t)[S], uint32_t (&x)[S], uint32_t g) {
    static_assert(S >= 2, "one-bit carries after the first exchange");
    uint64_t c = 0;
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const uint64_t v = t[p] + c;
        x[p] = (uint32_t)v & MASK;
        c = v >> RB;
    }
    uint64_t carry = gup64<G>(c);
    if (g == 0) carry = 0;
    uint32_t all = MASK;
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const uint64_t v = (uint64_t)x[p] + carry;
        x[p] = (uint32_t)v & MASK;
        carry = v >> RB;
        all &= x[p];
    }
    uint32_t cin = carry_in<G>((uint32_t)carry, all == MASK ? 1u : 0u, g);
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const uint32_t v = x[p] + cin;
        x[p] = v & MASK;
        cin = v >> RB;
    }
}

// y <- y - n over the group; returns the final borrow (1 if y < n), valid in
// every lane.  Digits of y are left as y - n mod R (caller selects).  Each
// lane subtracts its own digits (borrow out 0 or 1); a borrow into a lane
// passes through it only when its difference digits are all zero; the
// borrows cross the group by lookahead.
template <int S, int G>
__device__ __forceinline__ uint32_t sub_n(const uint32_t (&y)[S], const uint32_t (&n)[S], uint32_t (&u)[S],
                                          uint32_t g) {
    uint32_t br = 0, any = 0;
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const uint32_t v = y[p] - n[p] - br;
        br = v >> 31;
        u[p] = v & MASK;
        any |= u[p];
    }
    const uint32_t prop = any == 0 ? 1u : 0u;
    uint32_t bin = carry_in<G>(br, prop, g);
    const uint32_t out = br | (prop & bin);  // borrow out of this lane
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const uint32_t v = u[p] - bin;
        bin = v >> 31;
        u[p] = v & MASK;
    }
    return gbcast_top<G>(out);
}

// y <- y mod N for y < 2N.
template <int S, int G>
__device__ __forceinline__ void cond_sub(uint32_t (&y)[S], const uint32_t (&n)[S], uint32_t g) {
    uint32_t u[S];
    const uint32_t borrow = sub_n<S, G>(y, n, u, g);
#pragma unroll
    for (int p = 0; p < S; ++p) y[p] = borrow ? y[p] : u[p];
}

// y <- y * 2^s (s in {0,1}), value stays below R.
template <int S, int G>
__device__ __forceinline__ void shl(uint32_t (&y)[S], uint32_t s, uint32_t g) {
    uint32_t top = y[S - 1] >> (RB - 1);
    uint32_t cin = gup<G>(top);
    if (g == 0) cin = 0;
#pragma unroll
    for (int p = S - 1; p >= 1; --p) y[p] = ((y[p] << s) | ((y[p - 1] >> (RB - 1)) & s)) & MASK;
    y[0] = ((y[0] << s) | (cin & s)) & MASK;
}

// y <- 2y mod N for y < N.
template <int S, int G>
__device__ __forceinline__ void mod_double(uint32_t (&y)[S], const uint32_t (&n)[S], uint32_t g) {
    shl<S, G>(y, 1u, g);
    cond_sub<S, G>(y, n, g);
}

template <int S, int G>
__device__ __forceinline__ bool equal(const uint32_t (&a)[S], const uint32_t (&b)[S]) {
    uint32_t d = 0;
#pragma unroll
    for (int p = 0; p < S; ++p) d |= a[p] ^ b[p];
    return gor<G>(d) == 0;
}

// Row multipliers through LDS.  Each squaring, lane g of a group writes its
// S digits to the group's stage (row g, stride stage_stride(S) words); the S
// rows that use lane o's digits then read them with one broadcast load each
// (every lane of the group reads the same word), in place of a two-step DPP
// broadcast plus its operand copy per row on the VALU.  The stage is private
// to one group of one wavefront, whose LDS operations complete in order, so
// no barrier is needed.  The odd stride puts the groups of a wavefront in
// different banks.
constexpr int stage_stride(int S) { return S | 1; }
constexpr size_t stage_bytes(int S) { return (size_t)256 * stage_stride(S) * sizeof(uint32_t); }

// lane g <- lane g+1 across the whole 16-lane DPP row (the row's last lane
// reads zero); a group's top lane reads the next group's lane 0, which the
// caller replaces.
__device__ __forceinline__ uint32_t row_down(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, DPP_ROW_SHL1, 0xF, 0xF, true);
}

// Digits of a lane-local accumulator array t times 2^s (s in {0,1}) with one
// cross-lane step instead of G - 1: each lane carries through its own
// positions, doubles its (now exact) digits, and adds the carry out of the
// lane below, split into two digits, to its two bottom digits.  Digits 0 and
// 1 of a lane may then exceed 28 bits (< 2^29 and < 2^28 + 2^10); every
// other digit is exact.  The value is exact: the top lane's own carry is zero
// whenever the value is below R.  sqr_redc accepts such digits (see its
// bound); everything else wants canon() first.
template <int S, int G>
__device__ __forceinline__ void lazy_carry(const uint64_t (&t)[S], uint32_t (&x)[S], uint32_t g, uint32_t s) {
    uint64_t c = 0;
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const uint64_t v = t[p] + c;
        x[p] = (uint32_t)v & MASK;
        c = v >> RB;
    }
    const uint32_t out = (x[S - 1] >> (RB - 1)) & s;  // the doubling's bit out of the top digit
#pragma unroll
    for (int p = S - 1; p >= 1; --p) x[p] = ((x[p] << s) | ((x[p - 1] >> (RB - 1)) & s)) & MASK;
    x[0] = (x[0] << s) & MASK;
    c = (c << s) + out;  // the lane's carry out, < 2^38
    uint32_t c0 = gup<G>((uint32_t)c & MASK), c1 = gup<G>((uint32_t)(c >> RB));
    if (g == 0) c0 = c1 = 0;
    x[0] += c0;
    x[1] += c1;
}

// Exact 28-bit digits from lazy ones (lazy_carry's output or any digits whose
// value is below R).
template <int S, int G>
__device__ __forceinline__ void canon(uint32_t (&x)[S], uint32_t g) {
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) t[p] = x[p];
    normalize<S, G>(t, x, g);
}

// One row of the squaring at rotation u: T += x_i * x; m from the bottom
// digit (lane 0); T += m * N; then every lane adds its bottom accumulator's
// bits above 28 to its next position and hands the low 28 bits to the lane
// below; the group's top lane takes zero.
template <int S, int G>
__device__ __forceinline__ void cios_row(uint64_t (&t)[S], const uint32_t (&x)[S], const uint32_t (&n)[S],
                                         uint32_t xi, uint32_t np, bool top, int u) {
#pragma unroll
    for (int j = 0; j < S; ++j) t[(u + j) % S] = mad(xi, x[j], t[(u + j) % S]);
    const uint32_t m = gbcast0<G>(((uint32_t)t[u] * np) & MASK);
#pragma unroll
    for (int j = 0; j < S; ++j) t[(u + j) % S] = mad(m, n[j], t[(u + j) % S]);
    t[(u + 1) % S] += t[u] >> RB;
    const uint32_t lo = row_down((uint32_t)t[u] & MASK);
    // a group of 16 is a whole DPP row, whose last lane the move already
    // zeroes; a smaller group's top lane reads the next group's lane 0
    t[u] = (G == 16 || !top) ? lo : 0u;
}

// cios_row for the wide lanes (S > 32): m is broadcast by a DPP move with
// no old value when the group is a quad (every lane has a source), and the
// digit handed down is masked after the move by a per-lane mask that is
// zero in the group's top lane, in place of a mask and a select: 7 other
// instructions per row instead of 9 (1% of the 4247-bit kernel).
template <int S, int G>
__device__ __forceinline__ void cios_row_wide(uint64_t (&t)[S], const uint32_t (&x)[S], const uint32_t (&n)[S],
                                              uint32_t xi, uint32_t np, uint32_t lane_mask, int u) {
#pragma unroll
    for (int j = 0; j < S; ++j) t[(u + j) % S] = mad(xi, x[j], t[(u + j) % S]);
    const uint32_t mu = ((uint32_t)t[u] * np) & MASK;
    uint32_t m;
    if constexpr (G == 4)
        m = (uint32_t)__builtin_amdgcn_mov_dpp((int)mu, 0 /* quad_perm [0,0,0,0] */, 0xF, 0xF, true);
    else
        m = gbcast0<G>(mu);
#pragma unroll
    for (int j = 0; j < S; ++j) t[(u + j) % S] = mad(m, n[j], t[(u + j) % S]);
    t[(u + 1) % S] += t[u] >> RB;
    t[u] = row_down((uint32_t)t[u]) & lane_mask;
}

// x <- x^2 R^{-1} mod N (lazy: x < 4N in, < 2N out), R = 2^(28*G*S).
// `stage` is the group's LDS stage (G rows of stage_stride(S) words).
// LAZY: the result comes as lazy_carry digits times 2^s (< 4N); otherwise as
// exact digits (s must be 0).
//
// Rows as in cios_row.  Lane 0's low bits are zero after each row (that is
// what m is for), so T /= 2^28 is exact.  Accumulator bound: a value enters
// a lane below 2^28, takes two products (< 2^56 each) per row for at most S
// rows and one carry below 2^36, and leaves the lane at 28 bits again:
// 2S(2^28 - 1)^2 + 2^36 + 2^28 < 2^64 for S <= 127, whatever G, with exact
// input digits.  With lazy_carry digits (x_i * x_j < 2^58, m * n_j < 2^56)
// the same count gives S(2^58 + 2^56) + 2^37 < 2^64 for S <= 51.
template <int S, int G, bool LAZY = false>
__device__ __forceinline__ void sqr_redc(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t np, uint32_t g,
                                         uint32_t *stage, uint32_t s = 0) {
    static_assert(S <= 51 && (G == 4 || G == 8 || G == 16), "accumulator bound; groups tile a 16-lane DPP row");
    constexpr int ST = stage_stride(S);
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) {
        t[p] = 0;
        stage[g * ST + p] = x[p];
    }
    const bool top = (g == G - 1);
    if constexpr (S > 32) {
        // Wide lanes (quads of 4247-bit numbers at two wavefronts per SIMD).
        // Left to itself the compiler reassociates a block's rows into one
        // dependent chain of multiply-adds per accumulator, which keeps every
        // row's multiplier and m live and spills.  Here each row's multiplier
        // is read one row ahead, a scheduling barrier keeps the rows in
        // order, and every fourth row the accumulators pass through an empty
        // asm, which bounds each chain at 8 products: no spills, and 2^15
        // tests of 4247 bits take 204 ms against 218-220 without the pin.
        const uint32_t lane_mask = top ? 0u : MASK;
        uint32_t xnext = stage[0];
        for (int o = 0; o < G; ++o) {
#pragma unroll
            for (int u = 0; u < S; ++u) {
                const uint32_t xi = xnext;
                if (u + 1 < S)
                    xnext = stage[o * ST + u + 1];
                else if (o + 1 < G)
                    xnext = stage[(o + 1) * ST];
                cios_row_wide<S, G>(t, x, n, xi, np, lane_mask, u);
                __builtin_amdgcn_sched_barrier(0);
                if (u % 4 == 3) {
#pragma unroll
                    for (int p = 0; p < S; ++p) asm volatile("" : "+v"(t[p]));
                }
            }
        }
    } else {
        for (int o = 0; o < G; ++o) {
            uint32_t xb[S];
#pragma unroll
            for (int u = 0; u < S; ++u) xb[u] = stage[o * ST + u];
#pragma unroll
            for (int u = 0; u < S; ++u) cios_row<S, G>(t, x, n, xb[u], np, top, u);
        }
    }
    if constexpr (LAZY)
        lazy_carry<S, G>(t, x, g, s);
    else
        normalize<S, G>(t, x, g);
}

// x <- x * y * R^{-1} mod N, exact digits in and out (x, y < N in, < 2N out
// for N < R/2: (xy + mN)/R < N^2/R + N).  The same row-wise CIOS as
// sqr_redc with the rows x_i * y: x goes through the stage, y stays in
// registers.  Accumulator bound as for sqr_redc with exact digits.
template <int S, int G>
__device__ __forceinline__ void mul_redc(uint32_t (&x)[S], const uint32_t (&y)[S], const uint32_t (&n)[S],
                                         uint32_t np, uint32_t g, uint32_t *stage) {
    static_assert(S <= 32 && (G == 4 || G == 8 || G == 16), "accumulator bound; groups tile a 16-lane DPP row");
    constexpr int ST = stage_stride(S);
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) {
        t[p] = 0;
        stage[g * ST + p] = x[p];
    }
    const bool top = (g == G - 1);
    for (int o = 0; o < G; ++o) {
        uint32_t xb[S];
#pragma unroll
        for (int u = 0; u < S; ++u) xb[u] = stage[o * ST + u];
#pragma unroll
        for (int u = 0; u < S; ++u) cios_row<S, G>(t, y, n, xb[u], np, top, u);
    }
    normalize<S, G>(t, x, g);
}

// a <- (a - c) mod N for a, c < N (exact digits).  a - c wraps to a - c + R
// when it borrows; adding N then carries out of the top lane, which drops
// the R.
template <int S, int G>
__device__ __forceinline__ void mod_sub(uint32_t (&a)[S], const uint32_t (&c)[S], const uint32_t (&n)[S],
                                        uint32_t g) {
    uint32_t u[S];
    const uint32_t borrow = sub_n<S, G>(a, c, u, g);
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) t[p] = (uint64_t)u[p] + (borrow ? n[p] : 0u);
    normalize<S, G>(t, a, g);
}

template <int S, int G>
__device__ __forceinline__ bool is_zero(const uint32_t (&a)[S]) {
    uint32_t d = 0;
#pragma unroll
    for (int p = 0; p < S; ++p) d |= a[p];
    return gor<G>(d) == 0;
}

// y <- y >> 1 over the group (the top lane shifts in zero).
template <int S, int G>
__device__ __forceinline__ void shr1(uint32_t (&y)[S], uint32_t g) {
    uint32_t above = row_down(y[0] & 1u);  // lane g <- lane g+1's lowest bit
    if (g == G - 1) above = 0;
#pragma unroll
    for (int p = 0; p < S - 1; ++p) y[p] = (y[p] >> 1) | ((y[p + 1] & 1u) << (RB - 1));
    y[S - 1] = (y[S - 1] >> 1) | (above << (RB - 1));
}

// y <- y + 2^pos (exact digits out; the value stays below R).
template <int S, int G>
__device__ __forceinline__ void add_pow2(uint32_t (&y)[S], uint32_t pos, uint32_t g) {
    const uint32_t d = pos / RB, v = 1u << (pos % RB);
#pragma unroll
    for (int p = 0; p < S; ++p)
        if (g * S + (uint32_t)p == d) y[p] += v;
    uint64_t t[S];
#pragma unroll
    for (int p = 0; p < S; ++p) t[p] = y[p];
    normalize<S, G>(t, y, g);
}

}  // namespace lg
}  // namespace pgap
