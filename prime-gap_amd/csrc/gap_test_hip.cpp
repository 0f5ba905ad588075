// gap_test_hip -- drop-in for the reference testers gap_test_simple /
// gap_test_gpu: reads a combined_sieve unknown file, finds the previous and
// next prime around every m*P#/d on the GPU (streaming scan of libpgap) and
// prints the same per-m rows and stats lines.
//
// Reference counterparts: main + prime_gap_test of gap_test_simple.cpp:43-200
// (rows "%-5d %.4f  %ld * %ld#/%ld -%d to +%d", merit = gap / (K_log +
// log m) as float), StatsCounters (gap_test_common.cpp:29-124), argument
// handling of Args::argparse for the testers (gap_common.cpp:805-1099).
//
// Differences by design: PRP tests run on the GPU: Fermat base 2 on every
// candidate, and each side's first hit confirmed by Baillie-PSW (strong base
// 2 plus strong Lucas), the verdict mpz_probab_prime_p(N, 25) gives; m's are
// scanned concurrently and rows are printed in m order as they complete;
// malformed input is an error message and exit status 1, not an assert.
//
// --side-skip selects gap_test_gpu's order instead (gap_test_gpu.cu:591-617):
// prev side first, and an m whose prev gap has merit below min_merit/2 - 2
// ends without its next side (counted as "only prev_prime", no row).
// --side-skip=prob selects gap_test.py's probabilistic one-side skip
// (gap_test.py:184-202): every m scans its prev side first; then, in m
// order, the chance of a record given that side alone is set against the
// running average (side_prob.h), and the next side is scanned only when it
// is above it.  It reads m_stats.prob_record from --db and the record gaps
// from --prime-gaps-db (gap_test.py:465-486); a skipped m is saved with
// next_p = 0, as gap_test.py saves it.
//
// Extra flags: --depth N, --confirm bpsw|sprp2|none (scan knobs; sprp2 is
// the reference GPU tester's strong base-2 verdict), --hbm-budget-mb N,
// --results FILE (tab-separated P D m next_p prev_p merit, the columns of the
// reference's result table, schema.sql:142-154), --dump-unknowns (parse and
// print the lists, no GPU).
//
// --db FILE records every finished m in an SQLite file the way gap_test.py
// does through gap_test_stats.save (gap_test_stats.py:217-238): a `result`
// row (next_p 0 for a side-skipped m) and the m_stats columns next_p, prev_p,
// merit, prp_next, prp_prev, test_time (a row is added when gap_stats did not
// create one).  m's whose `result` row already has next_p >= 0 are not tested
// again (load_existing, gap_test_stats.py:145-150, and the resume rule of
// gap_test.py:341-349).  test_time is the m's time in flight in the scan.
// --prp-top-percent X (with --db) tests only the m's whose m_stats
// prob_record is at least the X% threshold (determine_test_threshold,
// gap_test_stats.py:456-486); partial rows are tested whatever it says.
#include <getopt.h>

#include <algorithm>
#include <chrono>
#include <clocale>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/pgap.h"
#include "hostbn.h"
#include "resultdb.h"
#include "side_prob.h"
#include "unknowns.h"

using namespace pgap;

namespace {

struct Args {
    UnknownConfig u;
    float min_merit = 18;
    uint64_t mskip = 0;
    int verbose = 2;
    uint32_t depth = 256, confirm = PGAP_CONFIRM_BPSW;
    int poll_timeout_s = 600;
    size_t hbm_budget = 0;
    std::string results, db;
    int top_percent = 0;
    bool dump = false;
    bool side_skip = false;       // gap_test_gpu's rule
    bool side_skip_prob = false;  // gap_test.py's probabilistic rule
    std::string gaps_db;          // --prime-gaps-db: the record gaps
    int64_t megagap = 0;
    bool valid = true;
};

void usage(const char *name) {
    printf("Usage: %s\n", name);
    printf("[REQUIRED]\n  -u, --unknown-filename <P>_<D>_<mstart>_<minc>_s<SL>_l<MP>M.txt\n");
    printf("    (or -p <p> -d <d> --mstart <m> --minc <n> --sieve-length <SL> --max-prime <MP>)\n\n");
    printf("[OPTIONALLY]\n  --min-merit <x>   only display gaps with merit > x (default 18)\n");
    printf("  --mskip <m>       skip m < mskip (partial resume)\n");
    printf("  --depth <n>       cap on speculative tests per side per round (default 256)\n");
    printf("  --confirm <test>  confirmation of each side's first Fermat PRP: bpsw (default, the\n"
           "                    mpz_probab_prime_p verdict), sprp2 (strong base 2, gap_test_gpu's), none\n");
    printf("  --hbm-budget-mb <n>  device memory the scan may keep resident\n");
    printf("  --poll-timeout <s>  give up when no row arrives for this long (default 600)\n");
    printf("  --side-skip       gap_test_gpu order: prev side first, skip next if prev merit < min/2-2\n");
    printf("  --side-skip=prob  gap_test.py's one-side skip: prev side first, next side only when\n"
           "                    prob(record | prev) beats the running average (needs --db with\n"
           "                    m_stats.prob_record and --prime-gaps-db)\n");
    printf("  --prime-gaps-db <file>  record gaps (table gaps: gapsize, merit) for --side-skip=prob\n");
    printf("  --megagap <n>     with --side-skip=prob: count gaps above n instead of records\n");
    printf("  --results <file>  append result rows (P D m next_p prev_p merit, tab separated)\n");
    printf("  --db <file>       record result / m_stats rows in an SQLite file; skip m's it holds\n");
    printf("  --prp-top-percent <x>  only test the top x%% of m's by m_stats.prob_record (needs --db)\n");
    printf("  --dump-unknowns   print the parsed unknown lists and exit (no GPU)\n");
    printf("  -q, --quiet       less output (twice for less)\n  -h, --help\n");
}

bool is_prime_small(uint64_t n) {
    if (n < 2) return false;
    for (uint64_t t = 2; t * t <= n; ++t)
        if (n % t == 0) return false;
    return true;
}

Args parse_args(int argc, char **argv) {
    static struct option opts[] = {
        {"unknown-filename", required_argument, 0, 'u'}, {"min-merit", required_argument, 0, 3},
        {"mskip", required_argument, 0, 16},             {"p", required_argument, 0, 'p'},
        {"d", required_argument, 0, 'd'},                {"mstart", required_argument, 0, 1},
        {"minc", required_argument, 0, 2},               {"sieve-length", required_argument, 0, 4},
        {"max-prime", required_argument, 0, 5},          {"depth", required_argument, 0, 20},
        {"confirm", required_argument, 0, 21},           {"hbm-budget-mb", required_argument, 0, 22},
        {"results", required_argument, 0, 23},           {"dump-unknowns", no_argument, 0, 24},
        {"side-skip", optional_argument, 0, 25},               {"db", required_argument, 0, 26},
        {"prp-top-percent", required_argument, 0, 27},  {"poll-timeout", required_argument, 0, 28},
        {"prime-gaps-db", required_argument, 0, 29},     {"megagap", required_argument, 0, 30},
        {"method1", no_argument, 0, 8},                  {"quiet", no_argument, 0, 'q'},
        {"help", no_argument, 0, 'h'},                   {0, 0, 0, 0}};
    Args a;
    std::string err;
    int c, idx = 0;
    while ((c = getopt_long(argc, argv, "qhp:d:u:", opts, &idx)) >= 0) {
        switch (c) {
            case 'h': usage(argv[0]); exit(0);
            case 'q': a.verbose--; break;
            case 'p': a.u.p = (uint32_t)atoi(optarg); break;
            case 'd': a.u.d = (uint32_t)atoi(optarg); break;
            case 1: a.u.mstart = strtoull(optarg, nullptr, 10); break;
            case 2: a.u.minc = strtoull(optarg, nullptr, 10); break;
            case 4: a.u.sieve_length = (uint32_t)atoi(optarg); break;
            case 5: a.u.max_prime = strtoull(optarg, nullptr, 10) * 1000000ull; break;
            case 8: a.u.method1 = true; break;
            case 16: a.mskip = strtoull(optarg, nullptr, 10); break;
            case 3: a.min_merit = (float)atof(optarg); break;
            case 20: a.depth = (uint32_t)atoi(optarg); break;
            case 21:
                if (!strcmp(optarg, "bpsw")) a.confirm = PGAP_CONFIRM_BPSW;
                else if (!strcmp(optarg, "sprp2")) a.confirm = PGAP_CONFIRM_SPRP2;
                else if (!strcmp(optarg, "none")) a.confirm = PGAP_CONFIRM_NONE;
                else {
                    printf("--confirm takes bpsw, sprp2 or none: %s\n", optarg);
                    a.valid = false;
                }
                break;
            case 22: a.hbm_budget = (size_t)strtoull(optarg, nullptr, 10) << 20; break;
            case 23: a.results = optarg; break;
            case 24: a.dump = true; break;
            case 25:
                if (!optarg)
                    a.side_skip = true;
                else if (!strcmp(optarg, "prob"))
                    a.side_skip_prob = true;
                else {
                    printf("--side-skip takes no value or =prob: %s\n", optarg);
                    a.valid = false;
                }
                break;
            case 29: a.gaps_db = optarg; break;
            case 30: a.megagap = strtoll(optarg, nullptr, 10); break;
            case 26: a.db = optarg; break;
            case 27: a.top_percent = atoi(optarg); break;
            case 28: a.poll_timeout_s = std::max(1, atoi(optarg)); break;
            case 'u':
                if (!parse_unknown_filename(optarg, a.u, err)) {
                    fprintf(stderr, "%s\n", err.c_str());
                    a.valid = false;
                }
                break;
            default: a.valid = false;
        }
    }
    if (optind < argc) {
        printf("unknown positional argument: %s\n", argv[optind]);
        a.valid = false;
    }
    if (a.u.path.empty() && a.u.p) a.u.path = gen_unknown_fn(a.u);
    if (a.u.mstart == 0) { printf("mstart must be greater than 0: %lu\n", (unsigned long)a.u.mstart); a.valid = false; }
    if (a.u.minc == 0) { printf("minc must be greater than 0: %lu\n", (unsigned long)a.u.minc); a.valid = false; }
    if (a.u.d == 0) { printf("d must be greater than 0: %u\n", a.u.d); a.valid = false; }
    if (!is_prime_small(a.u.p)) { printf("p# not prime (p=%u)\n", a.u.p); a.valid = false; }
    if (a.u.sieve_length == 0) { printf("Must set sieve-length\n"); a.valid = false; }
    if (a.top_percent && (a.top_percent < 1 || a.top_percent > 100 || a.db.empty())) {
        printf("--prp-top-percent takes 1..100 and needs --db (m_stats rows from gap_stats)\n");
        a.valid = false;
    }
    if (a.side_skip_prob && (a.db.empty() || a.gaps_db.empty() || a.u.max_prime < 1000000)) {
        printf("--side-skip=prob needs --db (m_stats rows from gap_stats), --prime-gaps-db and max_prime >= 1M\n");
        a.valid = false;
    }
    if (a.megagap < 0 || (a.megagap && !a.side_skip_prob)) {
        printf("--megagap takes a positive gap size and goes with --side-skip=prob\n");
        a.valid = false;
    }
    if (a.mskip && (a.mskip < a.u.mstart || a.mskip >= a.u.mstart + a.u.minc)) {
        printf("mskip must lie in [mstart, mstart + minc)\n");
        a.valid = false;
    }
    return a;
}

// StatsCounters (gap_test_common.cpp:29-124), restated.
struct Stats {
    std::chrono::steady_clock::time_point start = std::chrono::steady_clock::now();
    uint32_t tests = 0;
    size_t total_unknown = 0, unk_prev = 0, unk_next = 0, prp_tests = 0;
    size_t fallback_prev = 0, fallback_next = 0, skips = 0;
    float best_merit = 0;
    long best_m = 0;
    double tests_per_second = 0;

    void process(int verbose, double SL, long m, bool last, size_t ul, size_t uu, int prev_p, int next_p,
                 size_t pt, size_t nt, float merit) {
        tests += 1;
        total_unknown += ul + uu;
        unk_prev += ul;
        unk_next += uu;
        prp_tests += pt + nt;
        if (merit > best_merit) {
            best_merit = merit;
            best_m = m;
        }
        size_t interval = (size_t)(1800 * tests_per_second);
        if (interval > 10000) interval -= interval % 10000;
        bool power = tests == 1;
        for (long p = 10; p <= (long)tests; p *= 10) power |= tests == p || tests == 3 * p || tests == 5 * p;
        if (!(last || power || (interval > 0 && tests % interval == 0))) return;
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        tests_per_second = tests / secs;
        if (verbose + (int)last < 1) return;
        printf("\tm=%ld %4ld <- unknowns -> %-4ld\t%4d <- gap -> %-4d\n", m, (long)ul, (long)uu, prev_p, next_p);
        if (tests > secs)
            printf("\t    tests     %-10d (%.2f/sec)  %.0f seconds elapsed\n", tests, tests / secs, secs);
        else
            printf("\t    tests     %-10d (%.2f secs/test)  %.0f seconds elapsed\n", tests, secs / tests, secs);
        printf("\t    unknowns  %-10ld (avg: %.2f), %.2f%% composite  %.2f%% <- %% -> %.2f%%\n", (long)total_unknown,
               total_unknown / (double)tests, 100.0 * (1 - total_unknown / ((2.0 * SL + 1) * tests)),
               100.0 * unk_prev / total_unknown, 100.0 * unk_next / total_unknown);
        printf("\t    prp tests %-10ld (avg: %.2f) (%.1f tests/sec)\n", (long)prp_tests, prp_tests / (float)tests,
               prp_tests / secs);
        if (verbose >= 2) {
            if (skips)
                printf("\t    only prev_prime %ld (%.2f%%)\n", (long)skips, 100.0 * skips / tests);
            if (fallback_prev + fallback_next > 0)
                printf("\t    fallback prev_gap %ld (%.2f%%), next_gap %ld (%.2f%%)\n", (long)fallback_prev,
                       100.0 * fallback_prev / tests, (long)fallback_next, 100.0 * fallback_next / tests);
            printf("\t    best merit this interval: %.3f (at m=%ld)\n", best_merit, best_m);
        }
        best_merit = 0;
        best_m = 0;
    }
};

struct Pending {
    size_t ul = 0, uu = 0;
};

}  // namespace

int main(int argc, char **argv) {
    Args a = parse_args(argc, argv);
    if (!a.valid) {
        usage(argv[0]);
        return 1;
    }
    Words K;
    if (!primorial_over(a.u.p, a.u.d, K)) {
        fprintf(stderr, "d=%u must divide %u# with P#/d > 1\n", a.u.d, a.u.p);
        return 1;
    }
    const double K_log = ln_words(K);

    setlocale(LC_NUMERIC, "");
    if (a.verbose >= 2) printf("Compiled with %s\n", pgap_version());
    if (a.verbose >= 0) {
        printf("\n");
        printf("Testing m * %u#/%u, m = %lu + [0, %'lu)\n", a.u.p, a.u.d, (unsigned long)a.u.mstart,
               (unsigned long)a.u.minc);
    }
    if (a.mskip > 0) printf("\tskipping m < %'lu\n", (unsigned long)a.mskip);
    setlocale(LC_NUMERIC, "C");

    std::ifstream in(a.u.path, std::ios::in | std::ios::binary);
    if (!in.is_open()) {
        // a bare name resolves under unknowns/ first; accept it where it lies too
        const std::string alt = a.u.path.substr(a.u.path.find('/') + 1);
        in.open(alt, std::ios::in | std::ios::binary);
        if (in.is_open()) a.u.path = alt;
    }
    if (!in.is_open()) {
        fprintf(stderr, "can't open unknown file '%s'\n", a.u.path.c_str());
        return 1;
    }
    if (a.verbose >= 1) printf("\nReading unknowns from '%s'\n", a.u.path.c_str());

    std::string first;
    if (!std::getline(in, first)) {
        fprintf(stderr, "empty unknown file\n");
        return 1;
    }
    const int format = detect_format(first);
    uint64_t first_m = a.u.mstart;
    while (gcd64(first_m, a.u.d) > 1) ++first_m;
    if (format < 0 || strtoull(first.c_str(), nullptr, 10) != first_m) {
        fprintf(stderr, "unknown file must start with the line of m=%lu (lines start with <mstart + mi>)\n",
                (unsigned long)first_m);
        return 1;
    }
    in.clear();
    in.seekg(0);

    BitLayout bl;
    if (format == 2) bl.build(K, a.u.p, a.u.d, a.u.sieve_length);

    uint64_t valid_ms = 0, last_mi = 0;
    for (uint64_t mi = 0; mi < a.u.minc; ++mi)
        if (gcd64(a.u.mstart + mi, a.u.d) == 1) {
            ++valid_ms;
            last_mi = mi;
        }

    // m's this run leaves out: finished rows of --db, and with
    // --prp-top-percent those below the probability threshold
    ResultDb db;
    std::set<uint64_t> skip_m, partial;
    const uint64_t m_hi = a.u.mstart + a.u.minc - 1;
    if (!a.db.empty()) {
        if (!db.open(a.db) || !db.existing(a.u.p, a.u.d, a.u.mstart, m_hi, skip_m, partial)) {
            fprintf(stderr, "database '%s': %s\n", a.db.c_str(), db.error().c_str());
            return 1;
        }
        if (a.verbose >= 0)
            printf("Found %lu results in '%s'\n", (unsigned long)(skip_m.size() + partial.size()), a.db.c_str());
    }
    // gap_test.py's m_probs[m][1]: m_stats.prob_record, or prob_merit for
    // --megagap with --prp-top-percent below 100 (gap_test_stats.py:465-466)
    std::map<uint64_t, double> m_prob;
    double prob_threshold = 0;
    if (a.top_percent > 0 || a.side_skip_prob) {
        // determine_test_threshold (gap_test_stats.py:456-486)
        std::vector<std::pair<uint64_t, float>> probs;
        const bool by_merit = a.megagap && a.top_percent > 0 && a.top_percent < 100;
        if (!db.probs(a.u.p, a.u.d, a.u.mstart, m_hi, probs, by_merit)) {
            fprintf(stderr, "database '%s': %s (run gap_stats first)\n", a.db.c_str(), db.error().c_str());
            return 1;
        }
        if (probs.size() != valid_ms) {
            fprintf(stderr, "m_stats holds %lu rows for this range, expected %lu (run gap_stats first)\n",
                    (unsigned long)probs.size(), (unsigned long)valid_ms);
            return 1;
        }
        for (auto &pr : probs) m_prob[pr.first] = pr.second;
        if (a.top_percent > 0 && a.top_percent < 100) {
            std::vector<float> best;
            for (auto &pr : probs) best.push_back(pr.second);
            std::sort(best.begin(), best.end(), std::greater<float>());
            // Python's round(): halves go to the even neighbour
            size_t at = (size_t)std::nearbyint(best.size() * (double)a.top_percent / 100.0);
            const float thr = best[std::min(at, best.size() - 1)];
            size_t n_sel = 0;
            for (auto &pr : probs) {
                if (pr.second >= thr)
                    ++n_sel;
                else if (!partial.count(pr.first))  // partial rows run whatever their probability
                    skip_m.insert(pr.first);
            }
            if (a.verbose >= 0) printf("Testing %lu m where prob(record) >= %.3g\n", (unsigned long)n_sel, thr);
            prob_threshold = thr;
            // a partial row runs whatever its probability, counted at 1.01 x
            // the threshold (gap_test.py:332-336)
            for (uint64_t m : partial)
                if (m_prob.count(m) && m_prob[m] < prob_threshold) m_prob[m] = 1.01 * prob_threshold;
        }
    }
    std::unique_ptr<OneSidedSkip> one_side;
    if (a.side_skip_prob) {
        // gap_test.py:465-486, 514-516: records near ln N at mstart, and the
        // chance that an unknown left by the sieve is prime
        const double M_log = K_log + log((double)a.u.mstart);
        std::vector<int64_t> records;
        std::string err;
        if (!load_record_gaps(a.gaps_db, M_log, records, err)) {
            fprintf(stderr, "prime gaps database '%s': %s\n", a.gaps_db.c_str(), err.c_str());
            return 1;
        }
        if (records.empty() && !a.megagap) {
            fprintf(stderr, "prime gaps database '%s' holds no record gap for ln N = %.2f\n", a.gaps_db.c_str(),
                    M_log);
            return 1;
        }
        const double prob_prime = side_prob_prime(M_log);
        const double pps = side_prob_after_sieve(prob_prime, a.u.max_prime);
        one_side.reset(new OneSidedSkip(records, a.megagap, a.u.p, a.u.d, a.u.sieve_length,
                                        mod_small(K, a.u.d), prob_prime, pps, prob_threshold));
        if (a.verbose >= 1) {
            if (!records.empty())
                printf("\tLoaded %lu records (%ld to %ld) from '%s'\n", (unsigned long)records.size(),
                       (long)records.front(), (long)records.back(), a.gaps_db.c_str());
            if (a.megagap) printf("\tMin merit for megagap(%ld) %.2f\n", (long)a.megagap, a.megagap / M_log);
            printf("\t~2x%.1f = %.1f PRP tests per m\n", 1 / pps, 2 / pps);
        }
    }
    // the last m this run tests: its row closes the stats
    uint64_t stats_last_mi = last_mi;
    for (uint64_t mi = last_mi + 1; mi-- > 0;)
        if (gcd64(a.u.mstart + mi, a.u.d) == 1 && !skip_m.count(a.u.mstart + mi)) {
            stats_last_mi = mi;
            break;
        }
    const float min_merit = a.min_merit;
    if (a.verbose >= 1) {
        const float m_log = (float)log((double)a.u.mstart);
        printf("Min Gap ~= %d (for merit > %.1f)\n", (int)(min_merit * (K_log + m_log)), min_merit);
        printf("\n%lu tests M_start(%lu) + mi(%d to %lu)\n\n", (unsigned long)valid_ms, (unsigned long)a.u.mstart, 0,
               (unsigned long)last_mi);
    }

    pgap_engine *e = nullptr;
    int fam = -1;
    if (!a.dump) {
        pgap_options opt{};
        opt.hbm_budget = a.hbm_budget;
        if (pgap_create(&opt, &e) != PGAP_OK) {
            fprintf(stderr, "pgap_create: %s\n", pgap_last_error(e));
            pgap_destroy(e);
            return 1;
        }
        if (pgap_scan_config(e, a.depth, a.confirm) != PGAP_OK ||
            (a.side_skip && pgap_scan_side_skip(e, 1, a.min_merit) != PGAP_OK) ||
            pgap_family(e, a.u.p, a.u.d, a.u.sieve_length, &fam) != PGAP_OK) {
            fprintf(stderr, "scan set-up: %s\n", pgap_last_error(e));
            pgap_destroy(e);
            return 1;
        }
    }
    FILE *results = nullptr;
    if (!a.results.empty() && !(results = fopen(a.results.c_str(), "a"))) {
        fprintf(stderr, "can't open %s\n", a.results.c_str());
        return 1;
    }

    Stats stats;
    std::map<uint64_t, Pending> pending;  // submitted, in m order
    std::map<uint64_t, pgap_scan_row> done;
    std::vector<pgap_scan_row> rows(4096);
    int status = 0;

    // --side-skip=prob: m's whose prev side is submitted and whose decision
    // is pending (decided in m order), prev rows back, next sides decided on
    // and waiting for admission
    struct Held {
        std::vector<int32_t> next;
        double prob_record;
    };
    std::map<uint64_t, Held> undecided;
    std::map<uint64_t, pgap_scan_row> prev_rows;
    std::map<uint64_t, std::vector<int32_t>> next_lists;
    std::deque<uint64_t> outbox;

    auto decide = [&]() {
        while (!undecided.empty()) {
            auto it = undecided.begin();
            auto pr = prev_rows.find(it->first);
            if (pr == prev_rows.end()) break;
            const Held &h = it->second;
            const double p1 = one_side->one_sided(pr->second.prev_p, h.next.data(), h.next.size(), it->first);
            if (one_side->decide(h.prob_record, p1)) {
                outbox.push_back(it->first);
                next_lists[it->first] = std::move(it->second.next);
            } else {
                done[it->first] = pr->second;  // flagged PGAP_ROW_NEXT_SKIPPED, next_p = 0
                prev_rows.erase(pr);
            }
            undecided.erase(it);
        }
    };
    // submit decided next sides until admission says the HBM budget is full
    auto submit_next_sides = [&]() -> bool {
        while (!outbox.empty()) {
            const uint64_t m = outbox.front();
            const std::vector<int32_t> &nx = next_lists[m];
            pgap_scan_req req{};
            req.family = fam;
            req.m = m;
            req.next = nx.data();
            req.n_next = (uint32_t)nx.size();
            req.tag = m;
            req.sides = PGAP_SIDES_NEXT;
            const pgap_status st = pgap_scan_submit(e, &req);
            if (st == PGAP_E_OOM) return true;
            if (st != PGAP_OK) {
                fprintf(stderr, "scan submit m=%lu: %s\n", (unsigned long)m, pgap_last_error(e));
                return false;
            }
            next_lists.erase(m);
            outbox.pop_front();
        }
        return true;
    };

    // print finished rows in m order
    auto flush = [&](bool all) {
        while (!pending.empty()) {
            auto it = pending.begin();
            auto d = done.find(it->first);
            if (d == done.end()) break;
            const pgap_scan_row &r = d->second;
            const long m = (long)r.m;
            const int prev_p = r.prev_p, next_p = r.next_p, gap = prev_p + next_p;
            const float merit = (float)(gap / (K_log + log((double)m)));
            const bool skipped = (r.flags & PGAP_ROW_NEXT_SKIPPED) != 0;
            if (skipped) stats.skips += 1;
            if (one_side) {
                // gap_test_stats.py:614-619: a skipped m's row has next_p = 0
                // and is printed like any other when its merit or record says so
                const bool record = one_side->is_record(gap);
                if (record || merit > min_merit)
                    printf("%-5d %.4f  %ld * %u#/%u -%d to +%d%s\n", gap, merit, m, a.u.p, a.u.d, prev_p, next_p,
                           record ? "\tRECORD!" : "");
                if (results)
                    fprintf(results, "%u\t%u\t%ld\t%d\t%d\t%.4f\n", a.u.p, a.u.d, m, next_p, prev_p, merit);
            } else if (!skipped) {  // a gap_test_gpu skip prints nothing (gap_test_gpu.cu:607-613)
                if (merit > min_merit)
                    printf("%-5d %.4f  %ld * %u#/%u -%d to +%d\n", gap, merit, m, a.u.p, a.u.d, prev_p, next_p);
                if (results)
                    fprintf(results, "%u\t%u\t%ld\t%d\t%d\t%.4f\n", a.u.p, a.u.d, m, next_p, prev_p, merit);
            }
            if (!a.db.empty() && !db.save(a.u.p, a.u.d, r.m, next_p, prev_p, gap / (K_log + log((double)m)),
                                          r.n_tests, r.p_tests, r.t_done - r.t_submit)) {
                fprintf(stderr, "database '%s': %s\n", a.db.c_str(), db.error().c_str());
                status = 1;
            }
            stats.fallback_prev += (r.flags & PGAP_ROW_PREV_FALLBACK) ? 1 : 0;
            stats.fallback_next += (r.flags & PGAP_ROW_NEXT_FALLBACK) ? 1 : 0;
            const bool is_last = (r.m - a.u.mstart) == stats_last_mi;
            stats.process(a.verbose, a.u.sieve_length, m, is_last, it->second.ul, it->second.uu, prev_p, next_p,
                          r.p_tests, r.n_tests, merit);
            done.erase(d);
            pending.erase(it);
        }
        (void)all;
    };
    auto drain = [&](bool wait) -> bool {
        size_t n = 0;
        // a scan that returns no row for --poll-timeout seconds is an error
        // (the worker stalled), not a wait without end
        if (pgap_scan_poll(e, rows.data(), rows.size(), &n, wait ? a.poll_timeout_s * 1000 : 0) != PGAP_OK) {
            fprintf(stderr, "scan: %s\n", pgap_last_error(e));
            return false;
        }
        for (size_t i = 0; i < n; ++i) {
            const pgap_scan_row &r = rows[i];
            if (!one_side) {
                done[r.m] = r;
            } else if (r.flags & PGAP_ROW_PREV_SKIPPED) {
                // the next side of an m whose prev side came back earlier
                auto pr = prev_rows.find(r.m);
                pgap_scan_row w = pr->second;
                w.next_p = r.next_p;
                w.n_tests = r.n_tests;
                w.flags = (pr->second.flags & PGAP_ROW_PREV_FALLBACK) | (r.flags & PGAP_ROW_NEXT_FALLBACK);
                w.t_done = r.t_done;
                done[r.m] = w;
                prev_rows.erase(pr);
            } else {
                prev_rows[r.m] = r;
            }
        }
        if (one_side) {
            decide();
            if (!submit_next_sides()) return false;
        }
        flush(false);
        return true;
    };

    std::string line, err;
    std::vector<int32_t> prev, next;
    for (uint64_t mi = 0; mi < a.u.minc && status == 0; ++mi) {
        const uint64_t m = a.u.mstart + mi;
        if (gcd64(m, a.u.d) > 1) continue;
        if (!std::getline(in, line)) {
            fprintf(stderr, "unknown file ended before m=%lu\n", (unsigned long)m);
            status = 1;
            break;
        }
        if (m < a.mskip || skip_m.count(m)) continue;
        uint64_t mp = 0;
        if (!parse_unknown_line(line, format, format == 2 ? &bl : nullptr, mp, prev, next, err) || mp != m) {
            fprintf(stderr, "bad line for m=%lu: %s\n", (unsigned long)m, mp != m ? "m mismatch" : err.c_str());
            status = 1;
            break;
        }
        if (a.dump) {
            printf("%s", plain_line(m, prev, next).c_str());
            continue;
        }
        pgap_scan_req req{};
        req.family = fam;
        req.m = m;
        req.prev = prev.data();
        req.n_prev = (uint32_t)prev.size();
        req.next = next.data();
        req.n_next = (uint32_t)next.size();
        req.tag = m;
        if (one_side) {
            req.sides = PGAP_SIDES_PREV;
            if (!submit_next_sides()) {  // decided next sides go first
                status = 1;
                break;
            }
        }
        for (;;) {
            pgap_status st = pgap_scan_submit(e, &req);
            if (st == PGAP_OK) break;
            if (st != PGAP_E_OOM) {
                fprintf(stderr, "scan submit m=%lu: %s\n", (unsigned long)m, pgap_last_error(e));
                status = 1;
                break;
            }
            if (!drain(true)) {
                status = 1;
                break;
            }
        }
        if (status) break;
        pending[m] = Pending{prev.size(), next.size()};
        if (one_side) {
            auto pm = m_prob.find(m);
            undecided[m] = Held{next, pm == m_prob.end() ? 0.0 : pm->second};
        }
    }
    while (status == 0 && e && (pgap_scan_in_flight(e) > 0 || !outbox.empty()))
        if (!drain(true)) status = 1;
    if (status == 0) flush(true);
    if (results) fclose(results);
    db.close();
    pgap_destroy(e);
    return status;
}
