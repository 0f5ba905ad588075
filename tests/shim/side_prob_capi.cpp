// C entry points over prime-gap_amd/csrc/side_prob.h so the CPU tests can
// call the one-side skip policy through ctypes and compare it with the
// oracle's restatement.  Test-only.
#include <cstdint>
#include <vector>

#include "side_prob.h"

using pgap::OneSidedSkip;

extern "C" {

void *sp_new(const int64_t *records, size_t n, int64_t megagap, uint32_t P, uint32_t D, uint32_t SL,
             uint64_t K_mod_d, double prob_prime, double prob_after_sieve, double prob_threshold) {
    std::vector<int64_t> r(records, records + n);
    return new OneSidedSkip(r, megagap, P, D, SL, K_mod_d, prob_prime, prob_after_sieve, prob_threshold);
}

void sp_free(void *h) { delete static_cast<OneSidedSkip *>(h); }

double sp_one_sided(const void *h, int64_t prev_p, const int32_t *next, size_t n, uint64_t m) {
    return static_cast<const OneSidedSkip *>(h)->one_sided(prev_p, next, n, m);
}

int sp_decide(void *h, double prob_record, double new_prob_record) {
    return static_cast<OneSidedSkip *>(h)->decide(prob_record, new_prob_record) ? 1 : 0;
}

double sp_prob_prime_coprime(const void *h) { return static_cast<const OneSidedSkip *>(h)->prob_prime_coprime(); }

size_t sp_extended(const void *h, int32_t *out, size_t cap) {
    const auto &x = static_cast<const OneSidedSkip *>(h)->extended();
    for (size_t i = 0; i < x.size() && i < cap; ++i) out[i] = x[i];
    return x.size();
}

double sp_after_sieve(double prob_prime, uint64_t max_prime) { return pgap::side_prob_after_sieve(prob_prime, max_prime); }

double sp_prob_prime(double log_n) { return pgap::side_prob_prime(log_n); }
}
