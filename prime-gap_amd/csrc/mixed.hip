// mixed.hip -- pgap_prp_batch_mixed: bucket the numbers, queue tiles widest
// first and run the persistent kernel of mixed_impl.h.
#include <algorithm>
#include <vector>

#include "engine_impl.h"
#include "mixed_impl.h"

using namespace pgap;

static pgap_status mixed_once(pgap_engine *e, const uint32_t *words, const uint64_t *offset, const uint32_t *limbs,
                              size_t count, pgap_test kind, const uint32_t *bases, uint32_t nbases,
                              uint8_t *verdict);

// Several strong-PRP bases: the first base runs on every number, each later
// base only on the numbers that passed so far (a few percent of sieved
// candidates), so the widest lane of a wavefront never carries the others
// through bases they have already failed.
extern "C" pgap_status pgap_prp_batch_mixed(pgap_engine *e, const uint32_t *words, const uint64_t *offset,
                                            const uint32_t *limbs, size_t count, pgap_test kind,
                                            const uint32_t *bases, uint32_t nbases, uint8_t *verdict) {
    if (!e || kind != PGAP_SPRP || nbases <= 1 || !bases || nbases > (uint32_t)kMaxBases || !verdict || !offset ||
        !limbs)
        return mixed_once(e, words, offset, limbs, count, kind, bases, nbases, verdict);
    pgap_status st = mixed_once(e, words, offset, limbs, count, kind, bases, 1, verdict);
    for (uint32_t k = 1; k < nbases && st == PGAP_OK; ++k) {
        std::vector<uint64_t> off;
        std::vector<uint32_t> lim;
        std::vector<size_t> at;
        for (size_t i = 0; i < count; ++i)
            if (verdict[i]) {
                off.push_back(offset[i]);
                lim.push_back(limbs[i]);
                at.push_back(i);
            }
        if (at.empty()) break;
        std::vector<uint8_t> v(at.size());
        st = mixed_once(e, words, off.data(), lim.data(), at.size(), kind, bases + k, 1, v.data());
        for (size_t j = 0; j < at.size() && st == PGAP_OK; ++j) verdict[at[j]] = v[j];
    }
    return st;
}

static pgap_status mixed_once(pgap_engine *e, const uint32_t *words, const uint64_t *offset, const uint32_t *limbs,
                              size_t count, pgap_test kind, const uint32_t *bases, uint32_t nbases,
                              uint8_t *verdict) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    if (count == 0) return PGAP_OK;
    if (!words || !offset || !limbs || !verdict || count > 0xfffffff0u) {
        e->set_error("pgap_prp_batch_mixed: bad arguments");
        return PGAP_E_ARG;
    }
    Bases bs{};
    if (kind == PGAP_SPRP) {
        if (nbases == 0 || nbases > (uint32_t)kMaxBases || !bases) {
            e->set_error("SPRP needs 1..%d bases", kMaxBases);
            return PGAP_E_ARG;
        }
        bs.n = nbases;
        for (uint32_t i = 0; i < nbases; ++i) {
            if (bases[i] < 2 || bases[i] > 63) {
                e->set_error("SPRP base %u outside [2, 63]", bases[i]);
                return PGAP_E_ARG;
            }
            bs.b[i] = bases[i];
        }
        if (nbases > 1) {  // the kernel tests bases.b[0]; pgap_prp_batch_mixed stages the others
            e->set_error("pgap_prp_batch_mixed: one base per persistent launch");
            return PGAP_E_ARG;
        }
    } else if (kind != PGAP_FERMAT2) {
        e->set_error("pgap_prp_batch_mixed runs FERMAT2 or SPRP (kind %d)", (int)kind);
        return PGAP_E_ARG;
    }
    // bucket per number: the smallest digit count whose bound covers it
    constexpr int NB = MixedBuckets::n;
    // headroom below R = 2^(28 digits): the variants' lazy-reduction bounds
    const int slack = kind == PGAP_FERMAT2 ? 4 : 8 + (int)(select_bits(KIND_SPRP, 0, bs));
    std::vector<std::vector<uint32_t>> by_bucket(NB);
    size_t total_words = 0;
    for (size_t i = 0; i < count; ++i) {
        const uint32_t *w = words + offset[i];
        uint32_t bits = 0;
        for (int j = (int)limbs[i] - 1; j >= 0; --j)
            if (w[j]) {
                bits = 32u * (uint32_t)j + (32u - (uint32_t)__builtin_clz(w[j]));
                break;
            }
        int b = 0;
        while (b < NB && 28 * MixedBuckets::digits[b] - slack < (int)std::max(bits, 2u)) ++b;
        if (b == NB) {
            e->set_error("pgap_prp_batch_mixed: number %zu has %u bits, above the widest bucket (%d)", i, bits,
                         28 * MixedBuckets::digits[NB - 1] - slack);
            return PGAP_E_TOO_LARGE;
        }
        by_bucket[(size_t)b].push_back((uint32_t)i);
        total_words = std::max<size_t>(total_words, offset[i] + limbs[i]);
    }
    // tiles of 64, widest bucket first
    std::vector<uint32_t> order;
    std::vector<uint8_t> tile_bucket;
    for (int b = NB - 1; b >= 0; --b) {
        const auto &v = by_bucket[(size_t)b];
        for (size_t k = 0; k < v.size(); k += 64) {
            tile_bucket.push_back((uint8_t)b);
            for (size_t l = 0; l < 64; ++l) order.push_back(k + l < v.size() ? v[k + l] : 0xffffffffu);
        }
    }
    const uint32_t ntiles = (uint32_t)tile_bucket.size();
    (void)hipSetDevice(e->opt.device);
    pgap_status st;
    const size_t b_words = total_words * 4, b_off = count * 8, b_limbs = count * 4, b_order = order.size() * 4;
    if ((st = e->ensure(e->d_n, b_words + b_off + b_limbs + b_order + ntiles + 64, "mixed inputs")) ||
        (st = e->ensure(e->d_out, count, "verdicts")) || (st = e->ensure(e->d_small, 64, "tile counter")))
        return st;
    char *base = (char *)e->d_n.p;
    uint32_t *d_words = (uint32_t *)base;
    uint64_t *d_off = (uint64_t *)(base + ((b_words + 7) & ~(size_t)7));
    uint32_t *d_limbs = (uint32_t *)((char *)d_off + b_off);
    uint32_t *d_order = (uint32_t *)((char *)d_limbs + b_limbs);
    uint8_t *d_tb = (uint8_t *)((char *)d_order + b_order);
    hipError_t he;
    if ((he = hipMemcpyAsync(d_words, words, b_words, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(d_off, offset, b_off, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(d_limbs, limbs, b_limbs, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(d_order, order.data(), b_order, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(d_tb, tile_bucket.data(), ntiles, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemsetAsync(e->d_small.p, 0, 4, e->stream)) != hipSuccess)
        return e->hip_fail(he, "mixed upload");
    MixedLaunch ml{d_words, d_off, d_limbs, d_order, d_tb, ntiles, (uint32_t *)e->d_small.p, bs,
                   (uint8_t *)e->d_out.p};
    if ((he = hipEventRecord(e->ev0, e->stream)) != hipSuccess) return e->hip_fail(he, "event");
    he = kind == PGAP_FERMAT2 ? launch_mixed_fermat(ml, e->opt.device, e->stream)
                              : launch_mixed_sprp(ml, e->opt.device, e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "mixed kernel");
    if ((he = hipEventRecord(e->ev1, e->stream)) != hipSuccess) return e->hip_fail(he, "event");
    e->last.digits = 0;
    e->last.radix_bits = 28;
    e->last.lanes_per_number = 1;
    e->last.launches = 1;
    e->last.tests = count;
    e->timing_pending = true;
    if ((he = hipMemcpyAsync(verdict, e->d_out.p, count, hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
        (he = hipStreamSynchronize(e->stream)) != hipSuccess)
        return e->hip_fail(he, "mixed readback");
    e->collect_timing();
    return PGAP_OK;
}
