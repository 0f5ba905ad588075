/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, on GMP 6.2.1, of the prime-gap tester path that the HIP
 * engine replaces.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so, and only as the checker or as the
 * timed CPU baseline; the product library (libpgap.so) never links it.
 *
 * Each function cites the reference file:line whose behaviour it restates.
 * How each function is pinned independently of itself:
 *   or_fermat2 / or_sprp   : Python's own big-integer pow() (an independent
 *                            bignum implementation) at every size, brute force
 *                            factoring for N < 2^32, the published base-2
 *                            Fermat / strong pseudoprime lists, Mersenne and
 *                            Fermat numbers (tests/test_oracle.py).
 *   or_probab_prime        : brute force below 2^32, known primes/composites.
 *   or_strong_lucas        : the strong Lucas pseudoprimes (Selfridge method A)
 *                            below 10^5 of OEIS A217255 are exactly the
 *                            composites it passes; every prime passes; an
 *                            independent linear-recurrence evaluation of
 *                            U_n, V_n in Python below 3000.
 *   or_bpsw                : equals mpz_probab_prime_p != 0 on every N below
 *                            2*10^5 and on the pseudoprime lists.
 *   or_prevprime/nextprime : brute force on small ranges; the record gap
 *                            11244911*809#/1841070 (-5788, +17354) from the
 *                            reference README.md:40-46.
 *   or_test_interval       : README.md:265-288 rows for 907#/2190 (prev/next
 *                            per m) and the prp-test totals printed in
 *                            README.md:241-262 (144, 643, 1550, 2590).
 *   or_sieve_*             : md5 15a5cbff... of the reference's own unknown
 *                            file (misc/tests.sh:62-67) and the per-m unknown
 *                            counts in README.md:245-262.
 *   or_K / or_K_log        : mpz_primorial_ui against a Python product of
 *                            primes; K_log against math.log of the integer.
 */
#ifndef PGAP_ORACLE_H
#define PGAP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Verdicts on one number given as `limbs` little-endian 32-bit words. */
int or_fermat2(const uint32_t *n, int limbs);                 /* 2^(N-1) mod N == 1 */
int or_sprp(const uint32_t *n, int limbs, uint32_t base);     /* strong PRP, textbook */
int or_probab_prime(const uint32_t *n, int limbs, int reps);  /* mpz_probab_prime_p */
int or_strong_lucas(const uint32_t *n, int limbs);            /* strong Lucas, Selfridge A */
int or_bpsw(const uint32_t *n, int limbs);                    /* strong base 2 + strong Lucas */

/* Batched verdicts: n is row-major [count][limbs].  kind 0 = Fermat base 2,
 * kind 1 = strong PRP to each base in bases[0..nbases), kind 2 = strong
 * Lucas, kind 3 = Baillie-PSW. threads <= 0 means 1. */
void or_verdicts(const uint32_t *n, size_t count, int limbs, int kind,
                 const uint32_t *bases, int nbases, uint8_t *out, int threads);

/* K = P#/d (gap_common.cpp:116-121).  Writes up to cap limbs, returns the
 * number of limbs of K, or -1 if d does not divide P#. */
int or_K(uint32_t P, uint32_t D, uint32_t *out, int cap);
/* ln(K) computed exactly as gap_common.cpp:100-104 (mpz_get_d_2exp). */
double or_K_log(uint32_t P, uint32_t D);

/* Distances from centre = m*K to the previous prime below centre - start and
 * the next prime above centre + start, restating the fallback of
 * gap_test_common.cpp:160-178 (mpz_prevprime / mpz_nextprime).  The GMP here
 * (6.2.1) lacks mpz_prevprime, so the previous prime is a downward scan over
 * mpz_probab_prime_p(., 25). */
int64_t or_prev_prime_dist(uint32_t P, uint32_t D, uint64_t m, int64_t start);
int64_t or_next_prime_dist(uint32_t P, uint32_t D, uint64_t m, int64_t start);

/* One m of gap_test_simple (gap_test_common.cpp:128-181): scan prev then next
 * unknown distances with mpz_probab_prime_p(N, 25), first PRP wins, fall back
 * beyond SL.  out[0]=prev_p out[1]=next_p out[2]=prp tests (including hits)
 * out[3]=prev fell back out[4]=next fell back out[5]=prev tests out[6]=next
 * tests.  The in-sieve test by kind: 0 mpz_probab_prime_p(N,25) as
 * gap_test_simple does (gap_test_common.cpp:144,154); 1 Fermat base 2; 2 the
 * strong base-2 test of the reference GPU tester (gap_test_gpu.cu:81 ROUNDS=1,
 * miller_rabin.h:213-245 in its textbook form); 3 Baillie-PSW. */
void or_test_interval(uint32_t P, uint32_t D, uint64_t m, uint32_t SL,
                      const int32_t *prev, int np, const int32_t *next, int nn,
                      int kind, int64_t *out);

/* Sieve: unknowns of m*K + x for x in [-SL, SL] \ {0} free of every prime
 * <= max_prime, the set combined_sieve writes (combined_sieve.cpp:1261-1445).
 * prev gets positive distances in increasing order, next likewise. */
typedef struct or_sieve or_sieve;
or_sieve *or_sieve_new(uint32_t P, uint32_t D, uint32_t SL, uint64_t max_prime);
or_sieve *or_sieve_new_range(uint32_t P, uint32_t D, uint32_t SL, uint64_t lo, uint64_t max_prime);
void or_sieve_free(or_sieve *s);
int or_sieve_m(const or_sieve *s, uint64_t m, int32_t *prev, int *np, int32_t *next, int *nn, int cap);
int or_sieve_removed(const or_sieve *s, uint64_t m, uint64_t lo, int64_t *xs, uint64_t *ps, int cap);

/* CPU baseline, timed by the caller: gap_test_simple over count m's (gap_test_simple.cpp:152-195 calling
 * test_interval_cpu), m's split statically over `threads` pthreads.  List i
 * is prev[prev_off[i] .. prev_off[i] + n_prev[i]) (likewise next).  kind as
 * in or_test_interval.  rows[3*i..] = prev_p, next_p, prp tests. */
size_t or_bench_scan(uint32_t P, uint32_t D, uint32_t SL, const uint64_t *m, size_t count, const int32_t *prev,
                     const uint64_t *prev_off, const uint32_t *n_prev, const int32_t *next,
                     const uint64_t *next_off, const uint32_t *n_next, int kind, int threads, int64_t *rows);

/* gap_test.py's probabilistic one-side skip (host policy; the product's
 * side_prob.h is checked against these):
 *   or_prob_prime_after_sieve   prob_prime_sieve_length's value (gap_test_stats.py:241-284);
 *   or_setup_extended_gap       prob_prime_coprime, and the x in (SL, 2SL] coprime to
 *                               the primes <= P not dividing D (gap_test_stats.py:386-411);
 *   or_prob_record_one_sided    gap_test_stats.py:428-453 (records sorted ascending;
 *                               megagap > 0: a gap counts when above it);
 *   or_side_decisions           the running threshold, gap_test.py:274-275, 196-202. */
double or_prob_prime_after_sieve(double prob_prime, uint64_t max_prime);
double or_setup_extended_gap(uint32_t SL, uint32_t P, uint32_t D, double prob_prime, int32_t *xs, int cap,
                             int *nx);
double or_prob_record_one_sided(const int64_t *records, size_t n_records, int64_t megagap, int64_t other_side,
                                const int32_t *unknowns_next, size_t n_unknowns, double pps, uint64_t m,
                                uint64_t K_mod_d, uint64_t d, const int32_t *ext, size_t n_ext, double ppc);
void or_side_decisions(size_t n, const double *prob_record, const double *new_prob_record, double prob_threshold,
                       uint8_t *test_next);

#ifdef __cplusplus
}
#endif


#endif
