// tune.hip -- the autotuner: for each number width, time every kernel family
// (one lane per number, a group of 4, 8 or 16 lanes per number; radix 2^28) at its
// smallest covering digit count and under a few occupancy caps, and keep the
// fastest.  The table is plain text and is loaded by pgap_create through
// pgap_options.tuning_path.
//
// Reference counterpart: none in code -- the reference fixes BITS,
// WINDOW_BITS and THREADS_PER_INSTANCE at compile time (gap_test_gpu.cu:
// 48-81) and lists "auto optimize BATCH_SIZE, TPI" as an open item
// (EXTENDED.md:527-529).  There is no window width to tune here: base 2 and
// the small strong-PRP bases never use a table.
//
// Table format, one entry per line:
//   <kind> <max_bits> <radix> <digits> <lanes> <blocks_cap> <tests_per_s>
// kind 0 = Fermat base 2, 1 = strong PRP, 2 = strong Lucas; blocks_cap 0 =
// natural occupancy.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "engine_impl.h"

using namespace pgap;

bool pgap_engine::load_tuning(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return false;
    char line[256];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        TuneEntry t{};
        if (sscanf(line, "%d %u %d %d %d %d %lf", &t.kind, &t.bits, &t.radix, &t.digits, &t.lanes, &t.blocks_cap,
                   &t.rate) == 7)
            tune.push_back(t);
    }
    fclose(f);
    std::sort(tune.begin(), tune.end(),
              [](const TuneEntry &a, const TuneEntry &b) { return a.kind != b.kind ? a.kind < b.kind : a.bits < b.bits; });
    return true;
}

const TuneEntry *pgap_engine::tuned(int kind, uint32_t bits) const {
    for (const TuneEntry &t : tune)
        if (t.kind == kind && t.bits >= bits) return &t;
    return nullptr;
}

static void fill_random(std::vector<uint32_t> &w, size_t count, uint32_t limbs, uint32_t bits, uint64_t seed) {
    std::mt19937_64 rng(seed);
    w.assign(count * limbs, 0);
    for (size_t i = 0; i < count; ++i) {
        uint32_t *r = w.data() + i * limbs;
        for (uint32_t j = 0; j * 32 < bits; ++j) r[j] = (uint32_t)rng();
        const uint32_t top = (bits - 1) / 32, tb = (bits - 1) % 32;
        r[top] &= (tb == 31) ? 0xffffffffu : ((1u << (tb + 1)) - 1);
        r[top] |= 1u << tb;
        r[0] |= 1u;
    }
}

extern "C" pgap_status pgap_autotune(pgap_engine *e, const uint32_t *bits, size_t n, const char *path) {
    if (!e || (n && !bits)) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    (void)hipSetDevice(e->opt.device);
    std::vector<TuneEntry> found;
    std::vector<uint32_t> host;
    for (size_t q = 0; q < n; ++q) {
        const uint32_t b = bits[q];
        if (b < 8 || b > 1u << 16) {
            e->set_error("pgap_autotune: width %u outside [8, 65536]", b);
            return PGAP_E_ARG;
        }
        for (int kind = KIND_FERMAT2; kind <= KIND_LUCAS; ++kind) {
            // smallest covering variant of each (radix, lanes) family
            std::vector<const Variant *> cand;
            for (size_t i = 0; i < variant_count(); ++i) {
                const Variant *v = variant_at(i);
                if (v->kind != kind || v->max_bits < (int)b) continue;
                bool replaced = false;
                for (const Variant *&c : cand)
                    if (c->radix == v->radix && c->lanes == v->lanes) {
                        if (v->max_bits < c->max_bits) c = v;
                        replaced = true;
                    }
                if (!replaced) cand.push_back(v);
            }
            if (cand.empty()) continue;
            const uint32_t limbs = (b + 31) / 32;
            TuneEntry best{};
            best.rate = -1;
            Bases bs{};
            if (kind == KIND_SPRP) {  // one base per launch (strong PRP to base 3)
                bs.n = 1;
                bs.b[0] = 3;
            }
            for (const Variant *v : cand) {
                const int natural = v->blocks_per_cu ? v->blocks_per_cu() : 1;
                const size_t count = (size_t)std::max(1, natural) * e->num_cus * 256 / (size_t)v->lanes * 2;
                fill_random(host, count, limbs, b, 0x9e3779b97f4a7c15ull ^ b);
                pgap_status st;
                if ((st = e->ensure(e->d_n, host.size() * 4, "tuning numbers")) ||
                    (st = e->ensure(e->d_out, count, "tuning verdicts")))
                    return st;
                hipError_t he = hipMemcpy(e->d_n.p, host.data(), host.size() * 4, hipMemcpyHostToDevice);
                if (he != hipSuccess) return e->hip_fail(he, "tuning upload");
                std::vector<int> caps = {0};
                if (natural > 1) caps.push_back(1);
                if (natural > 2) caps.push_back(natural - 1);
                for (int cap : caps) {
                    float best_ms = 1e30f;
                    for (int rep = 0; rep < 2; ++rep) {
                        if ((he = hipEventRecord(e->ev0, e->stream)) != hipSuccess ||
                            (he = v->launch((const uint32_t *)e->d_n.p, limbs, (uint32_t)count, nullptr, nullptr, bs,
                                            (uint8_t *)e->d_out.p, cap, e->stream)) != hipSuccess ||
                            (he = hipEventRecord(e->ev1, e->stream)) != hipSuccess ||
                            (he = hipEventSynchronize(e->ev1)) != hipSuccess)
                            return e->hip_fail(he, "tuning launch");
                        float ms = 0;
                        (void)hipEventElapsedTime(&ms, e->ev0, e->ev1);
                        if (rep) best_ms = std::min(best_ms, ms);
                    }
                    const double rate = count / (best_ms * 1e-3);
                    if (rate > best.rate) {
                        best.kind = kind;
                        best.bits = b;
                        best.radix = v->radix;
                        best.digits = v->digits;
                        best.lanes = v->lanes;
                        best.blocks_cap = cap;
                        best.rate = rate;
                    }
                }
            }
            if (best.rate > 0) found.push_back(best);
        }
    }
    if (path && *path) {
        FILE *f = fopen(path, "w");
        if (!f) {
            e->set_error("pgap_autotune: cannot write %s", path);
            return PGAP_E_ARG;
        }
        fprintf(f, "# pgap tuning table: kind max_bits radix digits lanes blocks_cap tests_per_s\n");
        for (const TuneEntry &t : found)
            fprintf(f, "%d %u %d %d %d %d %.1f\n", t.kind, t.bits, t.radix, t.digits, t.lanes, t.blocks_cap, t.rate);
        fclose(f);
    }
    std::sort(found.begin(), found.end(), [](const TuneEntry &a, const TuneEntry &b) {
        return a.kind != b.kind ? a.kind < b.kind : a.bits < b.bits;
    });
    std::lock_guard<std::mutex> g(e->tune_mu);
    e->tune = found;
    return PGAP_OK;
}
