/*
 * pgap_test_runner.h -- C++ drop-in for the reference GPU tester's
 * test_runner_t<params> (miller_rabin.h:293-383), over the C ABI of pgap.h.
 *
 * A caller of the reference (run_gpu_thread, gap_test_gpu.cu:232-268) keeps
 * its code: `test_runner_t<params> runner(n, ROUNDS); runner.run_test(tests,
 * results);`.  Semantics kept:
 *   - rounds = the number of strong-probable-prime bases, the first `rounds`
 *     primes 2, 3, 5, ... (generate_primes, miller_rabin.h:269-286): ROUNDS
 *     = 1 (gap_test_gpu.cu:81) is a strong test to base 2, not a Fermat test;
 *   - results[i] = 1 when tests[i] passes every base (passed == rounds,
 *     miller_rabin.h:381), else 0; the caller sizes results (>= tests.size());
 *   - numbers cross as mpz_export(..., -1, 4, 0, 0) words (miller_rabin.h:
 *     37-52), zero-padded to the widest number of the batch.
 * Differences: errors throw std::runtime_error (the reference calls
 * exit(1)); a batch larger than n is accepted (the reference exits,
 * miller_rabin.h:345-348); rounds is limited to 1..8 (bases below 64); the
 * strong test is the textbook one (the reference kernel allows one squaring
 * more, miller_rabin.h:233-241); `params` (BITS / TPI) is accepted and
 * ignored: the engine picks its kernel per batch width.
 *
 * Needs <gmp.h>; link with -lpgap -lgmp.
 */
#ifndef PGAP_TEST_RUNNER_H
#define PGAP_TEST_RUNNER_H

#include <gmp.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgap.h"

template <class params = void>
class test_runner_t {
  public:
    const size_t n;       /* batch size the caller plans for (not a ceiling here) */
    const size_t rounds;  /* strong-probable-prime bases: the first `rounds` primes */

    test_runner_t(size_t n, size_t rounds) : n(n), rounds(rounds) {
        if (rounds < 1 || rounds > 8) throw std::invalid_argument("test_runner_t: rounds must be in [1, 8]");
        // generate_primes(rounds): 2, then odd numbers free of the odd primes so far
        bases_.push_back(2);
        for (uint32_t c = 3; bases_.size() < rounds; c += 2) {
            bool prime = true;
            for (size_t k = 1; k < bases_.size() && prime; ++k) prime = c % bases_[k] != 0;
            if (prime) bases_.push_back(c);
        }
        pgap_options opt{};
        pgap_status st = pgap_create(&opt, &engine_);
        if (st != PGAP_OK) {
            const std::string msg = engine_ ? pgap_last_error(engine_) : "pgap_create failed";
            pgap_destroy(engine_);
            engine_ = nullptr;
            throw std::runtime_error("test_runner_t: " + msg);
        }
    }
    ~test_runner_t() { pgap_destroy(engine_); }
    test_runner_t(const test_runner_t &) = delete;
    test_runner_t &operator=(const test_runner_t &) = delete;

    void run_test(const std::vector<mpz_t *> &tests, std::vector<int> &results) {
        if (tests.empty()) return;
        if (results.size() < tests.size()) throw std::invalid_argument("test_runner_t: results too small");
        size_t limbs = 1;
        for (mpz_t *z : tests) limbs = std::max(limbs, (mpz_sizeinbase(*z, 2) + 31) / 32);
        words_.assign(tests.size() * limbs, 0);
        for (size_t i = 0; i < tests.size(); ++i)
            mpz_export(&words_[i * limbs], nullptr, -1, sizeof(uint32_t), 0, 0, *tests[i]);
        verdict_.assign(tests.size(), 0);
        const pgap_status st = pgap_prp_batch(engine_, (uint32_t)limbs, words_.data(), tests.size(), PGAP_SPRP,
                                              bases_.data(), (uint32_t)bases_.size(), verdict_.data());
        if (st != PGAP_OK) throw std::runtime_error(std::string("test_runner_t: ") + pgap_last_error(engine_));
        for (size_t i = 0; i < tests.size(); ++i) results[i] = verdict_[i];
    }

    const std::vector<uint32_t> &bases() const { return bases_; }

  private:
    pgap_engine *engine_ = nullptr;
    std::vector<uint32_t> bases_, words_;
    std::vector<uint8_t> verdict_;
};

#endif
