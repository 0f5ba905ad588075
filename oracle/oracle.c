/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * GMP 6.2.1 restatement of the prime-gap tester path.  Never linked into the
 * product library; loaded by tests/, smoke() and bench.py's cpu_baseline.
 */
#include "oracle.h"

#include <gmp.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static void import_limbs(mpz_t z, const uint32_t *n, int limbs) {
    /* 32-bit little-endian words, as miller_rabin.h:37-52 exports them. */
    mpz_import(z, (size_t)limbs, -1, sizeof(uint32_t), 0, 0, n);
}

static int export_limbs(const mpz_t z, uint32_t *out, int cap) {
    size_t cnt = (mpz_sizeinbase(z, 2) + 31) / 32;
    if (mpz_sgn(z) == 0) cnt = 0;
    if ((int)cnt > cap) return -1;
    memset(out, 0, sizeof(uint32_t) * (size_t)cap);
    size_t written = 0;
    mpz_export(out, &written, -1, sizeof(uint32_t), 0, 0, z);
    return (int)cnt;
}

/* Fermat base 2, mpz_powm(2, N-1, N) == 1: the engine's scan filter.  (The
 * reference GPU kernel with ROUNDS=1, gap_test_gpu.cu:81, runs the stronger
 * strong base-2 test, sprp_z below with base 2: miller_rabin.h:213-245.) */
static int fermat2_z(const mpz_t N, mpz_t e, mpz_t r) {
    if (mpz_cmp_ui(N, 1) <= 0) return 0;
    mpz_sub_ui(e, N, 1);
    mpz_set_ui(r, 2);
    mpz_powm(r, r, e, N);
    return mpz_cmp_ui(r, 1) == 0;
}

/* Strong probable prime to `base`, textbook Miller-Rabin: N-1 = d*2^s,
 * x = base^d; pass if x == 1 or x == N-1, or x^(2^j) == N-1 for some
 * 0 < j < s.  miller_rabin.h:213-245 allows one squaring more than this; the
 * engine follows the textbook form, pinned here. */
static int sprp_z(const mpz_t N, uint32_t base, mpz_t d, mpz_t x, mpz_t nm1) {
    if (mpz_cmp_ui(N, 3) < 0 || mpz_even_p(N)) return 0;
    mpz_sub_ui(nm1, N, 1);
    mp_bitcnt_t s = mpz_scan1(nm1, 0);
    mpz_tdiv_q_2exp(d, nm1, s);
    mpz_set_ui(x, base);
    mpz_powm(x, x, d, N);
    if (mpz_cmp_ui(x, 1) == 0 || mpz_cmp(x, nm1) == 0) return 1;
    for (mp_bitcnt_t j = 1; j < s; j++) {
        mpz_mul(x, x, x);
        mpz_mod(x, x, N);
        if (mpz_cmp(x, nm1) == 0) return 1;
        if (mpz_cmp_ui(x, 1) == 0) return 0;
    }
    return 0;
}

int or_fermat2(const uint32_t *n, int limbs) {
    mpz_t N, e, r;
    mpz_inits(N, e, r, NULL);
    import_limbs(N, n, limbs);
    int v = fermat2_z(N, e, r);
    mpz_clears(N, e, r, NULL);
    return v;
}

int or_sprp(const uint32_t *n, int limbs, uint32_t base) {
    mpz_t N, d, x, nm1;
    mpz_inits(N, d, x, nm1, NULL);
    import_limbs(N, n, limbs);
    int v = sprp_z(N, base, d, x, nm1);
    mpz_clears(N, d, x, nm1, NULL);
    return v;
}

/* Strong Lucas probable-prime test with Selfridge's method A parameters, the
 * second half of the Baillie-PSW test that GMP 6.2's mpz_probab_prime_p runs
 * (gap_test_common.cpp:144,154 call it with reps = 25: trial division,
 * strong base 2, strong Lucas, then reps - 24 Miller-Rabin rounds).
 *   D = first of 5, -7, 9, -11, 13, ... with Jacobi(D/N) = -1 (skipping a D
 *       that N divides: N = |D| or a proper divisor of it; any other D with
 *       Jacobi 0 shares a factor with N, so N is composite);
 *   P = 1, Q = (1 - D)/4;  N + 1 = d * 2^s, d odd;
 *   N passes if U_d == 0 or V_{d*2^r} == 0 (mod N) for some 0 <= r < s.
 * 2 passes, other even N, N < 2 and perfect squares fail.  The D search is
 * bounded at |D| < 2^25, beyond every non-square N's first suitable D under
 * GRH (the least prime q with Jacobi(N/q) = -1 is below 2 ln(4N)^2); an N
 * reaching the bound is reported composite.
 * Computed with the textbook binary chain on (U_k, V_k, Q^k): doubling
 * U_2k = U_k V_k, V_2k = V_k^2 - 2Q^k; increment U_k+1 = (P U_k + V_k)/2,
 * V_k+1 = (D U_k + P V_k)/2 (halving mod N). */
static int strong_lucas_z(const mpz_t N) {
    if (mpz_cmp_ui(N, 2) < 0) return 0;
    if (mpz_cmp_ui(N, 2) == 0) return 1;
    if (mpz_even_p(N)) return 0;
    if (mpz_perfect_square_p(N)) return 0;
    long D = 0;
    for (unsigned long a = 5; a < (1ul << 25); a += 2) {
        const long Ds = (a & 2) ? -(long)a : (long)a;
        const int j = mpz_si_kronecker(Ds, N);
        if (j == -1) { D = Ds; break; }
        if (j == 0 && !(mpz_cmp_ui(N, a) <= 0 && a % mpz_get_ui(N) == 0)) return 0;
    }
    if (D == 0) return 0;
    const long Q = (1 - D) / 4;
    mpz_t d, U, V, Qk, t, u;
    mpz_inits(d, U, V, Qk, t, u, NULL);
    mpz_add_ui(d, N, 1);
    const mp_bitcnt_t s = mpz_scan1(d, 0);
    mpz_tdiv_q_2exp(d, d, s);
    /* k = 1: U_1 = 1, V_1 = P = 1, Q^1 = Q */
    mpz_set_ui(U, 1);
    mpz_set_ui(V, 1);
    mpz_set_si(Qk, Q);
    mpz_mod(Qk, Qk, N);
    for (long i = (long)mpz_sizeinbase(d, 2) - 2; i >= 0; i--) {
        mpz_mul(U, U, V);                 /* U_2k = U_k V_k */
        mpz_mod(U, U, N);
        mpz_mul(V, V, V);                 /* V_2k = V_k^2 - 2 Q^k */
        mpz_submul_ui(V, Qk, 2);
        mpz_mod(V, V, N);
        mpz_mul(Qk, Qk, Qk);
        mpz_mod(Qk, Qk, N);
        if (mpz_tstbit(d, (mp_bitcnt_t)i)) {
            mpz_add(t, U, V);             /* U_k+1 = (U_k + V_k)/2 (P = 1) */
            mpz_mul_si(u, U, D);          /* V_k+1 = (D U_k + V_k)/2 */
            mpz_add(u, u, V);
            mpz_mod(t, t, N);
            mpz_mod(u, u, N);
            if (mpz_odd_p(t)) mpz_add(t, t, N);
            if (mpz_odd_p(u)) mpz_add(u, u, N);
            mpz_tdiv_q_2exp(U, t, 1);
            mpz_tdiv_q_2exp(V, u, 1);
            mpz_mul_si(Qk, Qk, Q);
            mpz_mod(Qk, Qk, N);
        }
    }
    int pass = mpz_sgn(U) == 0 || mpz_sgn(V) == 0;
    for (mp_bitcnt_t r = 1; r < s && !pass; r++) {
        mpz_mul(V, V, V);                 /* V_{2k} = V_k^2 - 2 Q^k */
        mpz_submul_ui(V, Qk, 2);
        mpz_mod(V, V, N);
        mpz_mul(Qk, Qk, Qk);
        mpz_mod(Qk, Qk, N);
        pass = mpz_sgn(V) == 0;
    }
    mpz_clears(d, U, V, Qk, t, u, NULL);
    return pass;
}

/* Baillie-PSW: strong base 2, then strong Lucas (2 itself passes).  For every
 * N this equals mpz_probab_prime_p(N, reps) != 0 unless N is a BPSW
 * pseudoprime (none is known; none below 2^64). */
static int bpsw_z(const mpz_t N, mpz_t a, mpz_t b, mpz_t c) {
    if (mpz_cmp_ui(N, 2) == 0) return 1;
    return sprp_z(N, 2, a, b, c) && strong_lucas_z(N);
}

int or_strong_lucas(const uint32_t *n, int limbs) {
    mpz_t N;
    mpz_init(N);
    import_limbs(N, n, limbs);
    int v = strong_lucas_z(N);
    mpz_clear(N);
    return v;
}

int or_bpsw(const uint32_t *n, int limbs) {
    mpz_t N, a, b, c;
    mpz_inits(N, a, b, c, NULL);
    import_limbs(N, n, limbs);
    int v = bpsw_z(N, a, b, c);
    mpz_clears(N, a, b, c, NULL);
    return v;
}

/* gap_test_common.cpp:144,154: mpz_probab_prime_p(N, 25). */
int or_probab_prime(const uint32_t *n, int limbs, int reps) {
    mpz_t N;
    mpz_init(N);
    import_limbs(N, n, limbs);
    int v = mpz_probab_prime_p(N, reps);
    mpz_clear(N);
    return v;
}

typedef struct {
    const uint32_t *n;
    size_t lo, hi;
    int limbs, kind, nbases;
    const uint32_t *bases;
    uint8_t *out;
} verdict_job;

static void *verdict_worker(void *arg) {
    verdict_job *j = (verdict_job *)arg;
    mpz_t N, a, b, c;
    mpz_inits(N, a, b, c, NULL);
    for (size_t i = j->lo; i < j->hi; i++) {
        import_limbs(N, j->n + i * (size_t)j->limbs, j->limbs);
        int v;
        if (j->kind == 0) {
            v = fermat2_z(N, a, b);
        } else if (j->kind == 2) {
            v = strong_lucas_z(N);
        } else if (j->kind == 3) {
            v = bpsw_z(N, a, b, c);
        } else {
            v = 1;
            for (int k = 0; k < j->nbases && v; k++) v = sprp_z(N, j->bases[k], a, b, c);
        }
        if (j->out) j->out[i] = (uint8_t)v;
    }
    mpz_clears(N, a, b, c, NULL);
    return NULL;
}

void or_verdicts(const uint32_t *n, size_t count, int limbs, int kind,
                 const uint32_t *bases, int nbases, uint8_t *out, int threads) {
    if (threads <= 0) threads = 1;
    if ((size_t)threads > count) threads = count ? (int)count : 1;
    pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    verdict_job *jobs = (verdict_job *)calloc((size_t)threads, sizeof(verdict_job));
    for (int t = 0; t < threads; t++) {
        jobs[t].n = n; jobs[t].limbs = limbs; jobs[t].kind = kind;
        jobs[t].bases = bases; jobs[t].nbases = nbases; jobs[t].out = out;
        jobs[t].lo = count * (size_t)t / (size_t)threads;
        jobs[t].hi = count * (size_t)(t + 1) / (size_t)threads;
        pthread_create(&tid[t], NULL, verdict_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    free(tid);
    free(jobs);
}

/* gap_common.cpp:116-121 */
static int init_K(mpz_t K, uint32_t P, uint32_t D) {
    mpz_primorial_ui(K, P);
    if (D == 0 || !mpz_divisible_ui_p(K, D)) return -1;
    mpz_divexact_ui(K, K, D);
    return 0;
}

int or_K(uint32_t P, uint32_t D, uint32_t *out, int cap) {
    mpz_t K;
    mpz_init(K);
    int r = init_K(K, P, D);
    if (r == 0) r = export_limbs(K, out, cap);
    mpz_clear(K);
    return r;
}

/* gap_common.cpp:100-104 */
double or_K_log(uint32_t P, uint32_t D) {
    mpz_t K;
    mpz_init(K);
    if (init_K(K, P, D) != 0) { mpz_clear(K); return -1.0; }
    long exp;
    double mantis = mpz_get_d_2exp(&exp, K);
    double v = log(mantis) + log(2) * exp;
    mpz_clear(K);
    return v;
}

static void center_of(mpz_t c, uint32_t P, uint32_t D, uint64_t m) {
    init_K(c, P, D);
    mpz_t mm;
    mpz_init(mm);
    mpz_import(mm, 1, -1, sizeof(uint64_t), 0, 0, &m);
    mpz_mul(c, c, mm);
    mpz_clear(mm);
}

static void add_si64(mpz_t z, int64_t x) {
    if (x >= 0) {
        mpz_t t; mpz_init(t);
        uint64_t u = (uint64_t)x;
        mpz_import(t, 1, -1, sizeof(uint64_t), 0, 0, &u);
        mpz_add(z, z, t);
        mpz_clear(t);
    } else {
        mpz_t t; mpz_init(t);
        uint64_t u = (uint64_t)(-(x + 1)) + 1u;
        mpz_import(t, 1, -1, sizeof(uint64_t), 0, 0, &u);
        mpz_sub(z, z, t);
        mpz_clear(t);
    }
}

/* Largest prime strictly below z (z > 2); GMP 6.2.1 has no mpz_prevprime. */
static void prevprime_z(mpz_t r, const mpz_t z) {
    mpz_sub_ui(r, z, 1);
    if (mpz_cmp_ui(r, 2) <= 0) return;
    if (mpz_even_p(r)) mpz_sub_ui(r, r, 1);
    while (!mpz_probab_prime_p(r, 25)) mpz_sub_ui(r, r, 2);
}

int64_t or_prev_prime_dist(uint32_t P, uint32_t D, uint64_t m, int64_t start) {
    mpz_t c, t;
    mpz_inits(c, t, NULL);
    center_of(c, P, D, m);
    mpz_set(t, c);
    add_si64(t, -start);
    prevprime_z(t, t);
    mpz_sub(t, c, t);
    int64_t d = (int64_t)mpz_get_ui(t);
    mpz_clears(c, t, NULL);
    return d;
}

int64_t or_next_prime_dist(uint32_t P, uint32_t D, uint64_t m, int64_t start) {
    mpz_t c, t;
    mpz_inits(c, t, NULL);
    center_of(c, P, D, m);
    mpz_set(t, c);
    add_si64(t, start);
    mpz_nextprime(t, t);
    mpz_sub(t, t, c);
    int64_t d = (int64_t)mpz_get_ui(t);
    mpz_clears(c, t, NULL);
    return d;
}

/* The in-sieve test of or_test_interval by kind. */
static int interval_test(const mpz_t t, int kind, mpz_t e, mpz_t r, mpz_t q) {
    switch (kind) {
    case 0: return mpz_probab_prime_p(t, 25) != 0;
    case 1: return fermat2_z(t, e, r);
    case 2: return sprp_z(t, 2, e, r, q);
    default: return bpsw_z(t, e, r, q);
    }
}

/* gap_test_common.cpp:128-181 */
void or_test_interval(uint32_t P, uint32_t D, uint64_t m, uint32_t SL,
                      const int32_t *prev, int np, const int32_t *next, int nn,
                      int kind, int64_t *out) {
    mpz_t center, t, e, r, q;
    mpz_inits(center, t, e, r, q, NULL);
    center_of(center, P, D, m);
    int64_t prev_p = 0, next_p = 0, tests = 0, ptests = 0, ntests = 0;
    for (int i = 0; i < np; i++) {
        tests++; ptests++;
        mpz_sub_ui(t, center, (unsigned long)prev[i]);
        int v = interval_test(t, kind, e, r, q);
        if (v) { prev_p = prev[i]; break; }
    }
    for (int i = 0; i < nn; i++) {
        tests++; ntests++;
        mpz_add_ui(t, center, (unsigned long)next[i]);
        int v = interval_test(t, kind, e, r, q);
        if (v) { next_p = next[i]; break; }
    }
    int oos_prev = 0, oos_next = 0;
    if (prev_p == 0) {
        oos_prev = 1;
        mpz_sub_ui(t, center, SL);
        prevprime_z(t, t);
        mpz_sub(t, center, t);
        prev_p = (int64_t)mpz_get_ui(t);
    }
    if (next_p == 0) {
        oos_next = 1;
        mpz_add_ui(t, center, SL);
        mpz_nextprime(t, t);
        mpz_sub(t, t, center);
        next_p = (int64_t)mpz_get_ui(t);
    }
    out[0] = prev_p; out[1] = next_p; out[2] = tests;
    out[3] = oos_prev; out[4] = oos_next; out[5] = ptests; out[6] = ntests;
    mpz_clears(center, t, e, r, q, NULL);
}

/* ---- sieve: the unknown set combined_sieve writes ---- */
struct or_sieve {
    uint32_t P, D, SL;
    uint64_t max_prime;
    size_t nprimes;
    uint64_t *primes;
    uint64_t *kmod;   /* K mod p */
};

/* a * b mod p for a, b < p < 2^63 (128-bit product) */
static uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t p) {
    return (uint64_t)(((unsigned __int128)a * b) % p);
}

/* Primes in (lo, max_prime] with K mod p.  lo > 0 keeps only the upper
 * range (a test of the primes above 2^32 need not pay for K mod p below). */
or_sieve *or_sieve_new_range(uint32_t P, uint32_t D, uint32_t SL, uint64_t lo, uint64_t max_prime) {
    if (max_prime > 20000000000ull || max_prime < 2) return NULL;
    mpz_t K;
    mpz_init(K);
    if (init_K(K, P, D) != 0) { mpz_clear(K); return NULL; }
    uint64_t half = max_prime / 2 + 1;
    uint8_t *comp = (uint8_t *)calloc(half / 8 + 1, 1);  /* bit i <-> odd 2i+1 */
    size_t cnt = 1;
    for (uint64_t i = 1; i < half; i++) {
        if (comp[i >> 3] & (1u << (i & 7))) continue;
        uint64_t p = 2 * i + 1;
        if (p > max_prime) break;
        cnt++;
        if (p <= max_prime / p)
            for (uint64_t j = (p * p) / 2; j < half; j += p) comp[j >> 3] |= (uint8_t)(1u << (j & 7));
    }
    or_sieve *s = (or_sieve *)calloc(1, sizeof(or_sieve));
    s->P = P; s->D = D; s->SL = SL; s->max_prime = max_prime;
    s->primes = (uint64_t *)malloc(sizeof(uint64_t) * cnt);
    s->kmod = (uint64_t *)malloc(sizeof(uint64_t) * cnt);
    size_t k = 0;
    if (lo < 2) s->primes[k++] = 2;
    for (uint64_t i = 1; i < half; i++) {
        uint64_t p = 2 * i + 1;
        if (p > max_prime) break;
        if (p > lo && !(comp[i >> 3] & (1u << (i & 7)))) s->primes[k++] = p;
    }
    s->nprimes = k;
    /* K mod p for any p < 2^63: mpz_fdiv_ui takes an unsigned long (64-bit here) */
    for (size_t i = 0; i < k; i++) s->kmod[i] = (uint64_t)mpz_fdiv_ui(K, (unsigned long)s->primes[i]);
    free(comp);
    mpz_clear(K);
    return s;
}

or_sieve *or_sieve_new(uint32_t P, uint32_t D, uint32_t SL, uint64_t max_prime) {
    return or_sieve_new_range(P, D, SL, 0, max_prime);
}

void or_sieve_free(or_sieve *s) {
    if (!s) return;
    free(s->primes);
    free(s->kmod);
    free(s);
}

/* For the offsets x in [-SL, SL] that some prime in (lo, max_prime] removes:
 * x and the first such prime (in increasing order).  Returns the count, or
 * -1 when more than cap.  Lets a test check the 64-bit arithmetic exactly:
 * p must divide m*K + x. */
int or_sieve_removed(const or_sieve *s, uint64_t m, uint64_t lo, int64_t *xs, uint64_t *ps, int cap) {
    const uint32_t SL = s->SL;
    const uint32_t W = 2 * SL + 1;
    uint64_t *by = (uint64_t *)calloc(W, sizeof(uint64_t));
    for (size_t i = 0; i < s->nprimes; i++) {
        uint64_t p = s->primes[i];
        if (p <= lo) continue;
        uint64_t r = mulmod64(m % p, s->kmod[i], p);
        uint64_t first = ((SL % p) + p - r) % p;
        for (uint64_t j = first; j < W; j += p)
            if (!by[j] || p < by[j]) by[j] = p;
    }
    int n = 0;
    for (uint32_t j = 0; j < W; j++) {
        if (!by[j]) continue;
        if (n >= cap) { free(by); return -1; }
        xs[n] = (int64_t)j - (int64_t)SL;
        ps[n] = by[j];
        n++;
    }
    free(by);
    return n;
}

int or_sieve_m(const or_sieve *s, uint64_t m, int32_t *prev, int *np, int32_t *next, int *nn, int cap) {
    const uint32_t SL = s->SL;
    const uint32_t W = 2 * SL + 1;
    uint8_t *comp = (uint8_t *)calloc(W, 1);
    for (size_t i = 0; i < s->nprimes; i++) {
        uint64_t p = s->primes[i];
        uint64_t r = mulmod64(m % p, s->kmod[i], p);                 /* m*K mod p */
        uint64_t first = ((SL % p) + p - r) % p;                     /* index of x = -r */
        for (uint64_t j = first; j < W; j += p) comp[j] = 1;
    }
    comp[SL] = 1; /* x = 0 is never an unknown */
    int a = 0, b = 0;
    for (uint32_t d = 1; d <= SL; d++) {
        if (!comp[SL - d]) { if (a >= cap) { free(comp); return -1; } prev[a++] = (int32_t)d; }
        if (!comp[SL + d]) { if (b >= cap) { free(comp); return -1; } next[b++] = (int32_t)d; }
    }
    *np = a;
    *nn = b;
    free(comp);
    return 0;
}

/* ---- threaded gap_test_simple baseline ---- */
typedef struct {
    uint32_t P, D, SL;
    const uint64_t *m;
    const int32_t *prev, *next;
    const uint64_t *prev_off, *next_off;
    const uint32_t *n_prev, *n_next;
    int kind;
    size_t lo, hi;
    int64_t *rows;
} scan_job;

static void *scan_worker(void *arg) {
    scan_job *j = (scan_job *)arg;
    int64_t out[7];
    for (size_t i = j->lo; i < j->hi; i++) {
        or_test_interval(j->P, j->D, j->m[i], j->SL, j->prev + j->prev_off[i], (int)j->n_prev[i],
                         j->next + j->next_off[i], (int)j->n_next[i], j->kind, out);
        if (j->rows) {
            j->rows[3 * i] = out[0];
            j->rows[3 * i + 1] = out[1];
            j->rows[3 * i + 2] = out[2];
        }
    }
    return NULL;
}

size_t or_bench_scan(uint32_t P, uint32_t D, uint32_t SL, const uint64_t *m, size_t count, const int32_t *prev,
                     const uint64_t *prev_off, const uint32_t *n_prev, const int32_t *next,
                     const uint64_t *next_off, const uint32_t *n_next, int kind, int threads, int64_t *rows) {
    if (threads <= 0) threads = 1;
    if ((size_t)threads > count) threads = count ? (int)count : 1;
    pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    scan_job *jobs = (scan_job *)calloc((size_t)threads, sizeof(scan_job));
    for (int t = 0; t < threads; t++) {
        scan_job *j = &jobs[t];
        j->P = P; j->D = D; j->SL = SL; j->m = m; j->prev = prev; j->next = next;
        j->prev_off = prev_off; j->next_off = next_off; j->n_prev = n_prev; j->n_next = n_next;
        j->kind = kind; j->rows = rows;
        j->lo = count * (size_t)t / (size_t)threads;
        j->hi = count * (size_t)(t + 1) / (size_t)threads;
        pthread_create(&tid[t], NULL, scan_worker, j);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    free(tid);
    free(jobs);
    return count;
}

/* ---- gap_test.py's probabilistic one-side skip ---------------------------- */

double or_prob_prime_after_sieve(double prob_prime, uint64_t max_prime) {
    /* gap_test_stats.py:241-250: unknowns_after_sieve = 1 / (ln(max_prime) e^gamma)
     * (Mertens' third theorem, gamma written as 0.577215665 there) */
    const double unknowns_after_sieve = 1.0 / (log((double)max_prime) * exp(0.577215665));
    return prob_prime / unknowns_after_sieve;
}

static int small_is_prime(uint32_t n) {
    if (n < 2) return 0;
    for (uint32_t t = 2; t * t <= n; t++)
        if (n % t == 0) return 0;
    return 1;
}

double or_setup_extended_gap(uint32_t SL, uint32_t P, uint32_t D, double prob_prime, int32_t *xs, int cap,
                             int *nx) {
    /* gap_test_stats.py:386-411 */
    double ppc = prob_prime;
    for (uint32_t p = 2; p <= P; p++)
        if (small_is_prime(p)) ppc /= (1 - 1.0 / p);
    int n = 0;
    for (uint32_t x = SL + 1; x <= 2 * SL; x++) {
        int coprime = 1;
        for (uint32_t p = 2; p <= P && coprime; p++)
            if (D % p != 0 && x % p == 0 && small_is_prime(p)) coprime = 0;
        if (coprime) {
            if (n < cap) xs[n] = (int32_t)x;
            n++;
        }
    }
    *nx = n;
    return ppc;
}

static int is_record(const int64_t *records, size_t n, int64_t megagap, int64_t gap) {
    if (megagap) return gap > megagap;
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (records[mid] < gap) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && records[lo] == gap;
}

static uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

double or_prob_record_one_sided(const int64_t *records, size_t n_records, int64_t megagap, int64_t other_side,
                                const int32_t *unknowns_next, size_t n_unknowns, double pps, uint64_t m,
                                uint64_t K_mod_d, uint64_t d, const int32_t *ext, size_t n_ext, double ppc) {
    /* gap_test_stats.py:428-453, with its condition on the extension kept as
     * written: an offset x counts when gcd(m * (K mod d) + x, d) > 1 */
    double prob_record = 0, prob_nth = 1;
    for (size_t i = 0; i < n_unknowns; i++) {
        if (is_record(records, n_records, megagap, unknowns_next[i] + other_side)) prob_record += prob_nth * pps;
        prob_nth *= 1 - pps;
    }
    for (size_t i = 0; i < n_ext; i++) {
        if (gcd_u64(m * K_mod_d + (uint64_t)ext[i], d) > 1) {
            if (is_record(records, n_records, megagap, ext[i] + other_side)) prob_record += prob_nth * ppc;
            prob_nth *= 1 - ppc;
        }
    }
    return prob_record + prob_nth;
}

void or_side_decisions(size_t n, const double *prob_record, const double *new_prob_record, double prob_threshold,
                       uint8_t *test_next) {
    /* gap_test.py:274-275 (start: 5 * prob_threshold over 10 sides) and
     * :196-202, one m after another */
    double total = 5 * prob_threshold;
    long sides = 10;
    for (size_t i = 0; i < n; i++) {
        const double threshold = total / sides;
        const int t = new_prob_record[i] > threshold;
        total += prob_record[i] - (t ? 0 : new_prob_record[i]);
        sides += 1 + t;
        test_next[i] = (uint8_t)t;
    }
}
