/*
 * pgap.h -- C ABI of the MI355X PRP engine for prime-gap searches.
 *
 * The engine replaces the tester stage of sethtroisi/prime-gap:
 *   - the in-process seam test_runner_t<params>(n, rounds) / run_test(tests,
 *     results) (miller_rabin.h:317-383), called from run_gpu_thread
 *     (gap_test_gpu.cu:232-268), by pgap_create + pgap_prp_batch;
 *   - the per-m scan state machine of load_batch_thread (gap_test_gpu.cu:
 *     323-655) and test_interval_cpu (gap_test_common.cpp:128-181), by the
 *     streaming calls pgap_scan_*.
 *
 * Conventions:
 *   - numbers are little-endian 32-bit words, the layout miller_rabin.h:37-52
 *     exports from mpz_t (mpz_export(..., -1, 4, 0, 0)); row-major
 *     [count][limbs] arrays;
 *   - every call returns a pgap_status; the engine never calls exit() (the
 *     reference's CUDA_CHECK/CGBN_CHECK exit(1), miller_rabin.h:55-94);
 *     pgap_last_error() gives the message of the last failure;
 *   - inputs are copied before a call returns; the caller keeps ownership;
 *   - one engine serves one submitting thread plus one polling thread.
 */
#ifndef PGAP_H
#define PGAP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgap_engine pgap_engine;

typedef enum {
    PGAP_OK = 0,
    PGAP_E_ARG = -1,        /* bad argument (null pointer, zero limbs, ...) */
    PGAP_E_DEVICE = -2,     /* HIP runtime error, or the HIP code object is missing */
    PGAP_E_OOM = -3,        /* device allocation failed / admission refused by the HBM budget */
    PGAP_E_TOO_LARGE = -4,  /* number wider than the largest kernel variant */
    PGAP_E_STATE = -5,      /* call not valid in the engine's current state */
    PGAP_E_TIMEOUT = -6     /* pgap_scan_poll: no row within the time allowed */
} pgap_status;

/* The test a batch runs.
 *   FERMAT2  2^(N-1) == 1 (mod N): the scan's filter.
 *   SPRP     a strong-probable-prime (Miller-Rabin) test for each base in
 *            order; a number passes when every base passes.  The reference
 *            GPU tester's test_runner_t(n, ROUNDS) runs this with the bases
 *            generate_primes(ROUNDS) = 2, 3, 5, ... (miller_rabin.h:269-286,
 *            317-319, results[i] = passed == rounds at :381); its build uses
 *            ROUNDS = 1 (gap_test_gpu.cu:81), i.e. SPRP with bases {2}.
 *   LUCAS    the strong Lucas test with Selfridge's method A parameters
 *            (bases ignored).
 *   BPSW     Baillie-PSW: SPRP base 2, then LUCAS (bases ignored); the
 *            verdict of mpz_probab_prime_p(N, 25) != 0, the test of the
 *            reference CPU tester (gap_test_common.cpp:144,154), for every N
 *            that is not a BPSW pseudoprime (none is known). */
typedef enum { PGAP_FERMAT2 = 0, PGAP_SPRP = 1, PGAP_LUCAS = 2, PGAP_BPSW = 3 } pgap_test;

typedef struct {
    int device;               /* HIP device ordinal (the reference hard-codes 0, miller_rabin.h:326) */
    size_t hbm_budget;        /* bytes the streaming scan may keep resident; 0 = 1/4 of free HBM */
    const char *tuning_path;  /* autotuner table written by pgap_autotune(); NULL = built-in defaults */
    int radix;                /* force the digit radix in bits (every variant is 28; 0 = tuned) */
    int lanes_per_number;     /* force lanes per number (0 = tuned) */
    int waves_per_simd;       /* force an occupancy cap: resident 256-thread workgroups per CU, i.e.
                                 waves per SIMD (0 = tuned, else the kernel's natural occupancy) */
} pgap_options;

/* Engine lifetime: owns the stream, device buffers and streaming state.
 * Mirrors test_runner_t's constructor/destructor (miller_rabin.h:317-339). */
pgap_status pgap_create(const pgap_options *opt, pgap_engine **out);
void pgap_destroy(pgap_engine *e);
/* Message of the engine's last failure, copied for the calling thread: valid
 * until that thread calls pgap_last_error again. */
const char *pgap_last_error(const pgap_engine *e);
/* Library version string, e.g. "pgap 0.1 gfx950". */
const char *pgap_version(void);

/* ---- batch mode ------------------------------------------------------- */

/* Test count numbers of `limbs` words each (host memory, row-major) and
 * write verdict[i] in {0,1}.  Replaces test_runner_t::run_test
 * (miller_rabin.h:341-383), without its n-per-batch ceiling (:345-348). */
pgap_status pgap_prp_batch(pgap_engine *e, uint32_t limbs, const uint32_t *n, size_t count,
                           pgap_test kind, const uint32_t *bases, uint32_t nbases, uint8_t *verdict);

/* Same on device-resident buffers (d_n, d_verdict are device pointers),
 * asynchronous on the engine's stream; `max_bits` is an upper bound of the
 * numbers' bit length (0 = let the engine measure it).  pgap_sync() waits. */
pgap_status pgap_prp_batch_device(pgap_engine *e, uint32_t limbs, const uint32_t *d_n, size_t count,
                                  pgap_test kind, const uint32_t *bases, uint32_t nbases,
                                  uint8_t *d_verdict, uint32_t max_bits);

/* pgap_prp_batch_device for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  It calls no hipMalloc or
 * hipFree: the survivor scratch that several SPRP bases and BPSW need is a
 * fixed array of the library, one per GPU (4 MiB, held until the process
 * ends), so count <= PGAP_NOALLOC_MAX_COUNT, and max_bits must be given
 * (> 0).  Every call on one GPU shares that array: calls that overlap on a
 * GPU (several engines or streams at once) race and must be serialised.
 * Asynchronous on the engine's stream; pgap_sync() waits.  No timing is
 * recorded for pgap_last_launch. */
#define PGAP_NOALLOC_MAX_COUNT (1u << 20)
pgap_status pgap_prp_batch_device_noalloc(pgap_engine *e, uint32_t limbs, const uint32_t *d_n, size_t count,
                                          pgap_test kind, const uint32_t *bases, uint32_t nbases,
                                          uint8_t *d_verdict, uint32_t max_bits);

/* Mixed widths in one persistent launch.  Number i has limbs[i] words
 * starting at words + offset[i].  Numbers are grouped into size buckets (one
 * lane per number up to 2180 bits, groups of 4 or 8 lanes up to 4476 bits
 * for FERMAT2, 4472 for SPRP) and one work-queue kernel serves every bucket.
 * FERMAT2 and SPRP only.  pgap_last_launch reports lanes_per_number = 0 (it
 * varies by bucket). */
pgap_status pgap_prp_batch_mixed(pgap_engine *e, const uint32_t *words, const uint64_t *offset,
                                 const uint32_t *limbs, size_t count, pgap_test kind,
                                 const uint32_t *bases, uint32_t nbases, uint8_t *verdict);

pgap_status pgap_sync(pgap_engine *e);

/* Timing of the last PRP launch: device time measured with HIP events on
 * the engine's stream, and the kernel variant used. */
typedef struct {
    double kernel_ms;
    uint32_t digits;            /* digits per number of the variant */
    uint32_t radix_bits;        /* digit radix in bits (28) */
    uint32_t lanes_per_number;
    uint32_t launches;
    uint64_t tests;
    uint32_t graph;             /* 1 when the launch replayed a captured hipGraph */
    uint32_t reserved;
} pgap_launch_info;
/* Timing and shape of the last batch launch.  A batch call repeated with
 * exactly the same arguments (variant, buffers, count, bases) is captured
 * into a hipGraph on its second occurrence and replayed from then on, so a
 * steady-state loop costs one graph launch per call. */
pgap_status pgap_last_launch(const pgap_engine *e, pgap_launch_info *info);

/* ---- streaming scan (serving) ----------------------------------------- */

/* A family is K = P#/d (gap_common.cpp:116-121); registering it once lets
 * every request name only m. */
pgap_status pgap_family(pgap_engine *e, uint32_t P, uint32_t D, uint32_t sieve_length, int *family_id);

/* Candidate construction and filtering for a family, on the device.
 *
 * pgap_build_candidates_device writes N_i = m_i*K + x_i as `limbs` words per
 * row into d_out (device memory, [count][limbs]); it replaces the host-side
 * mpz_add_ui/mpz_sub_ui + mpz_export batch fill of gap_test_gpu.cu:431-476
 * and miller_rabin.h:352-354.  m and x are host arrays.  PGAP_E_TOO_LARGE if
 * some N does not fit `limbs` words or is negative.
 *
 * pgap_filter_candidates sets keep[i] = 1 when m_i*K + x_i has no prime
 * factor <= limit (trial division by every prime <= limit on the device, the
 * test combined_sieve applies to each offset); used to draw sieved synthetic
 * candidates and to screen offsets beyond the sieve length.  N itself equal
 * to a prime <= limit is reported as 0. */
pgap_status pgap_build_candidates_device(pgap_engine *e, int family, const uint64_t *m, const int64_t *x,
                                         size_t count, uint32_t limbs, uint32_t *d_out);
pgap_status pgap_filter_candidates(pgap_engine *e, int family, const uint64_t *m, const int64_t *x,
                                   size_t count, uint32_t limit, uint8_t *keep);

/* Device sieve: the unknowns combined_sieve writes (combined_sieve.cpp:
 * 1261-1445), i.e. the offsets x in [-SL, SL], x != 0, for which m*K + x has
 * no prime factor <= max_prime (SL = the family's sieve length).  For each
 * m_i the prev distances go to prev + i*SL (increasing), the next distances
 * to next + i*SL, and their counts to n_prev[i], n_next[i] (host arrays).
 * The primes <= max_prime are generated on the device and kept for later
 * calls with the same limit.  max_prime <= 2e10 (primes above 2^32 take
 * 64-bit modular arithmetic; the table costs 24 bytes per prime of HBM). */
pgap_status pgap_sieve(pgap_engine *e, int family, const uint64_t *m, size_t count, uint64_t max_prime,
                       int32_t *prev, uint32_t *n_prev, int32_t *next, uint32_t *n_next);

/* Which sides of m*K a request scans.  gap_test.py's probabilistic
 * side-skip (gap_test.py:169-202) finds the previous prime first and then
 * decides, from it, whether the next side is worth scanning: the caller
 * submits the m with PGAP_SIDES_PREV, decides, and submits it again with
 * PGAP_SIDES_NEXT when it scans on. */
enum {
    PGAP_SIDES_BOTH = 0,
    PGAP_SIDES_PREV = 1,  /* prev side only: row with PGAP_ROW_NEXT_SKIPPED */
    PGAP_SIDES_NEXT = 2,  /* next side only: row with PGAP_ROW_PREV_SKIPPED */
};

typedef struct {
    int family;
    uint64_t m;
    const int32_t *prev;  /* positive distances below m*K, increasing (outward) */
    uint32_t n_prev;
    const int32_t *next;  /* positive distances above m*K, increasing (outward) */
    uint32_t n_next;
    uint64_t tag;         /* caller cookie, echoed in the row */
    uint32_t sides;       /* PGAP_SIDES_*; 0 (a zeroed request) scans both */
} pgap_scan_req;

enum {
    PGAP_ROW_PREV_FALLBACK = 1,  /* no PRP among prev unknowns: searched beyond SL */
    PGAP_ROW_NEXT_FALLBACK = 2,
    PGAP_ROW_NEXT_SKIPPED = 4,  /* next side not scanned (side-skip or PGAP_SIDES_PREV): next_p = n_tests = 0 */
    PGAP_ROW_PREV_SKIPPED = 8,  /* prev side not scanned (PGAP_SIDES_NEXT): prev_p = p_tests = 0 */
};

/* One finished m: the row test_interval_cpu produces (gap_test_common.cpp:
 * 128-181).  p_tests / n_tests count tests up to and including the first PRP
 * on each side, as s_total_prp_tests does; speculative tests beyond it are
 * not counted. */
typedef struct {
    uint64_t m, tag;
    int32_t prev_p, next_p;
    uint32_t p_tests, n_tests;
    uint32_t flags;
    uint32_t family;
    double t_submit, t_done;  /* seconds on the engine's steady clock */
} pgap_scan_row;

/* Submit one m.  PGAP_E_OOM means the HBM budget has no room right now:
 * poll, then retry (admission control).  The first submit starts the
 * engine's scan worker (a host thread with its own stream) that runs rounds
 * while anything is in flight; m's submitted while others are on the device
 * join the next round.  Submit and poll may be called from one client thread
 * (not concurrently from several). */
pgap_status pgap_scan_submit(pgap_engine *e, const pgap_scan_req *req);
/* Collect finished rows (up to cap).  timeout_ms = 0 returns at once;
 * otherwise the call waits until at least one row is ready or nothing is in
 * flight, without a limit when timeout_ms < 0, else for at most timeout_ms
 * milliseconds, after which it returns PGAP_E_TIMEOUT with *n_out = 0 (the
 * scan goes on; poll again).  An error inside the worker is returned by the
 * next poll or submit. */
pgap_status pgap_scan_poll(pgap_engine *e, pgap_scan_row *rows, size_t cap, size_t *n_out, int timeout_ms);
/* Number of m's submitted and not yet returned. */
size_t pgap_scan_in_flight(const pgap_engine *e);

/* Scan telemetry since the first submit: PRP tests issued (including the
 * speculative ones past a side's first PRP), rounds run, and the number of
 * slots the HBM budget allows. */
pgap_status pgap_scan_stats(const pgap_engine *e, uint64_t *tests_issued, uint64_t *rounds, uint32_t *slots);

/* Autotuner.  For each width in bits[0..n), times every kernel family that
 * covers it (one lane per number; a group of 4, 8 or 16 lanes per number;
 * radix 2^28) at its smallest covering size and under a few occupancy caps,
 * for each test (Fermat, strong PRP, strong Lucas), and keeps the fastest.
 * The table replaces the engine's
 * current one (later batches and scans use it) and, when path is non-NULL,
 * is written there for pgap_options.tuning_path.  Text format, one line per
 * (test, width): "<test> <max_bits> <radix> <digits> <lanes> <cap> <tests/s>".
 * Replaces the reference's compile-time choice of BITS / WINDOW_BITS /
 * THREADS_PER_INSTANCE (gap_test_gpu.cu:48-81). */
pgap_status pgap_autotune(pgap_engine *e, const uint32_t *bits, size_t n, const char *path);

/* The denominator of the PRP kernels' roofline, measured: `iters` rounds of
 * independent v_mad_u64_u32 chains (32x32-bit products accumulated into 64
 * bits, the instruction the Montgomery squarings issue) on every compute
 * unit at 8 waves per SIMD.  Reports the sustained multiply-add rate, the
 * timed launch's kernel time and its multiply-add count.  No counterpart in
 * the reference (measurement only). */
pgap_status pgap_mad_peak(pgap_engine *e, uint32_t iters, double *mac_per_s, double *kernel_ms, uint64_t *macs);

/* Side-skip, gap_test_gpu's order (gap_test_gpu.cu:591-617): each m scans
 * its prev side first and starts the next side only after the prev PRP; when
 * that PRP's merit, (float)(prev_p / (K_log + ln m)), is below
 * min_merit / 2 - 2 the m finishes without the next side (row flag
 * PGAP_ROW_NEXT_SKIPPED, next_p = n_tests = 0).  Off by default (both sides
 * scan together, test_interval_cpu's result for every m).  Must be called
 * before the first submit. */
pgap_status pgap_scan_side_skip(pgap_engine *e, int enable, float min_merit);

/* What decides a side's prime.  Every scan tests candidates with FERMAT2;
 * a side's first Fermat PRP is then confirmed, and one that fails the
 * confirmation is treated as composite (the scan continues past it):
 *   NONE   no confirmation: the first Fermat base-2 PRP;
 *   SPRP2  a strong base-2 test: the reference GPU tester's verdict
 *          (gap_test_gpu.cu:81 ROUNDS = 1, miller_rabin.h:213-245; the
 *          textbook test, without that kernel's one extra squaring);
 *   BPSW   strong base 2 plus the strong Lucas test: the verdict of
 *          mpz_probab_prime_p(N, 25), gap_test_simple's row
 *          (gap_test_common.cpp:144,154).  The default. */
enum { PGAP_CONFIRM_NONE = 0, PGAP_CONFIRM_SPRP2 = 1, PGAP_CONFIRM_BPSW = 2 };

/* Streaming knobs: the cap on speculative depth per side (a round gives
 * each unfinished side about resident-lanes / sides tests, never more than
 * this; default 256, at most 4096), and the confirmation (PGAP_CONFIRM_*,
 * default PGAP_CONFIRM_BPSW).  Must be called before the first submit. */
pgap_status pgap_scan_config(pgap_engine *e, uint32_t depth, uint32_t confirm);

#ifdef __cplusplus
}
#endif

#endif
