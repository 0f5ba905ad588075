// engine_impl.h -- the engine object behind the opaque pgap_engine handle.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>

#include "../../include/pgap.h"
#include "kernels.h"

#include <vector>

#include "hostbn.h"

namespace pgap {
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};
class Scanner;

// One autotuner decision: numbers of up to `bits` bits of test `kind` run on
// the variant (radix, digits, lanes) with an occupancy cap.
struct TuneEntry {
    int kind = 0;
    uint32_t bits = 0;
    int radix = 0, digits = 0, lanes = 0, blocks_cap = 0;
    double rate = 0;
};

// Rows per batch launch: grid dimensions are 32-bit.
constexpr size_t kBatchChunk = (size_t)1 << 30;

// A variant and the occupancy cap to launch it with.
struct Choice {
    const Variant *v = nullptr;
    int cap = 0;
};

// The last batch launch's arguments and, once the same launch has come
// twice in a row, its captured graph.
struct BatchGraph {
    const Variant *v = nullptr;
    const uint32_t *d_n = nullptr;
    uint32_t limbs = 0;
    size_t count = 0;
    Bases bases{};
    uint8_t *d_out = nullptr;
    int cap = 0;
    void *sel = nullptr, *nsel = nullptr;
    hipGraphExec_t exec = nullptr;
    bool same(const BatchGraph &o) const {
        if (v != o.v || d_n != o.d_n || limbs != o.limbs || count != o.count || d_out != o.d_out || cap != o.cap ||
            sel != o.sel || nsel != o.nsel || bases.n != o.bases.n)
            return false;
        for (uint32_t i = 0; i < bases.n; ++i)
            if (bases.b[i] != o.bases.b[i]) return false;
        return true;
    }
};

// K = P#/d with what the device needs to build and screen m*K + x.
struct Family {
    uint32_t P = 0, D = 0, SL = 0;
    Words K;                 // host copy
    double K_log = 0;        // ln K (gap_common.cpp:100-104)
    uint32_t *d_K = nullptr; // device copy, K.size() words
    // trial-division tables for pgap_filter_candidates (built per limit)
    uint32_t filter_limit = 0;
    size_t nprimes = 0;
    uint32_t *d_primes = nullptr;
    uint32_t *d_kmod = nullptr;  // K mod p
    uint64_t *d_mu = nullptr;    // floor((2^64-1)/p) for Barrett reduction
    uint32_t stage_end[2] = {0, 0};  // prime-index ends of the first two filter stages
    // K mod p over the engine's sieve prime table (pgap_sieve)
    uint64_t *d_skmod = nullptr;
    uint64_t skmod_limit = 0;
};
}  // namespace pgap

struct pgap_engine {
    pgap_options opt{};
    bool ready = false;
    int num_cus = 256;  // compute units of the device (queried at create)
    // autotuner table (sorted by kind, bits); tune_mu guards it against a
    // pgap_autotune call while the scan worker picks variants
    std::vector<pgap::TuneEntry> tune;
    mutable std::mutex tune_mu;
    bool load_tuning(const char *path);
    const pgap::TuneEntry *tuned(int kind, uint32_t bits) const;  // caller holds tune_mu
    pgap::Choice select(int kind, uint32_t bits) const;             // variant + occupancy cap, or none
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing_pending = false;
    pgap_launch_info last{};

    mutable std::mutex err_mu;
    std::string err;

    pgap::DevBuf d_n, d_out, d_small, d_m, d_x, d_keep, d_idx0, d_idx1;
    pgap::DevBuf d_sel, d_nsel;  // survivors between strong-PRP bases
    pgap::BatchGraph graph;       // steady-state replay of a repeated batch launch
    uint32_t graph_launches = 0;  // kernel launches inside the captured graph
    hipError_t enqueue_batch(const pgap::Choice &c, const uint32_t *d_n, uint32_t limbs, size_t count,
                             const pgap::Bases &bases, uint8_t *d_out, uint32_t *sel, uint32_t *nsel,
                             uint32_t *launches);
    hipError_t enqueue_bpsw(const pgap::Choice &cs, const pgap::Choice &cl, const uint32_t *d_n, uint32_t limbs,
                            size_t count, uint8_t *d_out, uint32_t *sel, uint32_t *nsel, uint32_t *launches);
    std::vector<pgap::Family> families;  // appended only by pgap_family (client thread)
    pgap::Scanner *scanner = nullptr;  // owned; released by scan_shutdown()

    // primes <= sieve_limit (unordered) with Barrett constants, for pgap_sieve
    uint64_t *sieve_primes = nullptr;
    uint64_t *sieve_mu = nullptr;
    size_t sieve_nprimes = 0;
    uint64_t sieve_limit = 0;
    void sieve_shutdown();

    pgap::Family *family(int id) {
        return (id >= 0 && (size_t)id < families.size()) ? &families[(size_t)id] : nullptr;
    }
    void families_shutdown();

    void set_error(const char *fmt, ...) __attribute__((format(printf, 2, 3)));
    pgap_status hip_fail(hipError_t e, const char *what);
    pgap_status ensure(pgap::DevBuf &b, size_t bytes, const char *what);
    pgap_status pick(int kind, uint32_t bits, pgap::Choice *out);  // select(), error when none covers
    pgap_status launch_batch(const pgap::Choice &c, const uint32_t *d_n, uint32_t limbs, size_t count,
                             const pgap::Bases &bases, uint8_t *d_out);
    pgap_status launch_bpsw(const pgap::Choice &cs, const pgap::Choice &cl, const uint32_t *d_n, uint32_t limbs,
                            size_t count, uint8_t *d_out);
    pgap_status run_batch(pgap_test kind, const pgap::Bases &bs, uint32_t bits, const uint32_t *d_n,
                          uint32_t limbs, size_t count, uint8_t *d_out);
    void collect_timing();
    void scan_shutdown();
};
