// side_prob.h -- gap_test.py's probabilistic one-side skip, host C++ (no HIP).
//
// After an m's previous prime is known, the chance that its gap turns out a
// record is recomputed from that side alone (prob_record_one_sided,
// gap_test_stats.py:428-453); when that chance is not above a running
// threshold -- the average prob(record) per side tested so far -- the next
// side is not scanned (gap_test.py:184-202).  The caller feeds m's in m order
// (the reference's order with one worker thread).
//
// Arithmetic follows the reference's double-precision steps one for one
// (same operations, same order), so the decisions repeat the reference's on
// the same inputs.  Build without FMA contraction (-ffp-contract=off).
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace pgap {

// The Euler-Mascheroni constant as gap_test_stats.py:245 writes it.
constexpr double kSideSkipGamma = 0.577215665;

// prob_prime = 1/ln N - 1/ln^2 N with ln N taken at mstart (gap_test.py:515).
inline double side_prob_prime(double log_n) { return 1 / log_n - 1 / (log_n * log_n); }

// The return value of prob_prime_sieve_length (gap_test_stats.py:241-284):
// a number left by a sieve to max_prime is prime with probability
// prob_prime * ln(max_prime) * e^gamma (Mertens' third theorem).
inline double side_prob_after_sieve(double prob_prime, uint64_t max_prime) {
    const double left = 1.0 / (std::log((double)max_prime) * std::exp(kSideSkipGamma));
    return prob_prime / left;
}

class OneSidedSkip {
  public:
    // records: the record gap sizes (load_records, gap_test_stats.py:116-142);
    // megagap > 0 replaces them with "any gap above megagap".
    // P, D, SL: the family; K_mod_d = (P#/D) mod D; prob_prime as above.
    OneSidedSkip(const std::vector<int64_t> &records, int64_t megagap, uint32_t P, uint32_t D, uint32_t SL,
                 uint64_t K_mod_d, double prob_prime, double prob_after_sieve, double prob_threshold)
        : megagap_(megagap), D_(D), K_mod_d_(K_mod_d), pps_(prob_after_sieve) {
        for (int64_t g : records)
            if (g > 0) {
                if ((size_t)g >= record_.size()) record_.resize((size_t)g + 1, 0);
                record_[(size_t)g] = 1;
            }
        // setup_extended_gap (gap_test_stats.py:386-411): beyond the sieve,
        // an offset coprime to every prime of P# is prime with probability
        // prob_prime / prod_{p <= P} (1 - 1/p); the offsets in (SL, 2SL]
        // coprime to K's primes (those not dividing D) are the candidates.
        std::vector<uint8_t> composite(P + 1, 0);
        std::vector<uint8_t> off(2 * (size_t)SL + 1, 1);
        ppc_ = prob_prime;
        for (uint32_t p = 2; p <= P; ++p) {
            if (composite[p]) continue;
            for (uint64_t q = (uint64_t)p * p; q <= P; q += p) composite[q] = 1;
            ppc_ /= (1 - 1.0 / p);
            if (D % p == 0) continue;
            for (size_t x = 0; x < off.size(); x += p) off[x] = 0;
        }
        for (size_t x = (size_t)SL + 1; x < off.size(); ++x)
            if (off[x]) extended_.push_back((int32_t)x);
        // gap_test.py:274-275: the threshold starts as 5 * prob_threshold
        // spread over 10 sides
        total_ = 5 * prob_threshold;
        sides_ = 10;
    }

    // prob_record_one_sided: the chance that prev_p + (the next prime's
    // distance) is a record, given the next side's unknowns (outward).  The
    // extension beyond SL counts an offset x when gcd(m * (K mod D) + x, D)
    // > 1, the reference's condition as written (gap_test_stats.py:446).
    double one_sided(int64_t prev_p, const int32_t *next, size_t n_next, uint64_t m) const {
        double prob = 0, nth = 1;
        for (size_t i = 0; i < n_next; ++i) {
            if (counts(prev_p + next[i])) prob += nth * pps_;
            nth *= 1 - pps_;
        }
        const uint64_t base = m * K_mod_d_;
        for (int32_t x : extended_) {
            if (gcd(base + (uint64_t)x, D_) <= 1) continue;
            if (counts(prev_p + x)) prob += nth * ppc_;
            nth *= 1 - ppc_;
        }
        return prob + nth;
    }

    // gap_test.py:196-202: scan the next side when the one-sided chance is
    // above the running average; the m's a-priori prob_record (m_stats) joins
    // the average, minus the one-sided chance given up when skipping.
    bool decide(double prob_record, double new_prob_record) {
        const double threshold = total_ / (double)sides_;
        const bool test_next = new_prob_record > threshold;
        total_ += prob_record - (test_next ? 0 : new_prob_record);
        sides_ += test_next ? 2 : 1;
        return test_next;
    }

    bool is_record(int64_t gap) const { return gap > 0 && (size_t)gap < record_.size() && record_[(size_t)gap]; }
    double threshold() const { return total_ / (double)sides_; }
    double prob_prime_coprime() const { return ppc_; }
    const std::vector<int32_t> &extended() const { return extended_; }

  private:
    bool counts(int64_t gap) const { return megagap_ ? gap > megagap_ : is_record(gap); }
    static uint64_t gcd(uint64_t a, uint64_t b) {
        while (b) {
            const uint64_t t = a % b;
            a = b;
            b = t;
        }
        return a;
    }

    int64_t megagap_;
    uint64_t D_, K_mod_d_;
    double pps_, ppc_ = 0;
    std::vector<uint8_t> record_;
    std::vector<int32_t> extended_;
    double total_ = 0;
    int64_t sides_ = 10;
};

}  // namespace pgap
