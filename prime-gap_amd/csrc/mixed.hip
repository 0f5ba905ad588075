// mixed.hip -- pgap_prp_batch_mixed: bucket the numbers, queue tiles widest
// first and run the persistent kernels of mixed_impl.h (lane-group buckets
// and one-lane buckets side by side on two streams).
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
    // bucket per number: the smallest digit count whose bound covers it, the
    // one-lane buckets first, then the lane-group ones; headroom below
    // R = 2^(28 digits): the variants' lazy-reduction bounds (one lane: the
    // window's; a lane group: 4 or 8 bits)
    constexpr int N1 = MixedOneLane::n, NB = N1 + MixedGroups::n;
    const int slack1 = kind == PGAP_FERMAT2 ? 4 : 8 + (int)(select_bits(KIND_SPRP, 0, bs));
    const int slack_g = kind == PGAP_FERMAT2 ? 4 : 8;
    auto max_bits = [&](int b) {
        return b < N1 ? 28 * MixedOneLane::digits[b] - slack1 : 28 * MixedGroups::digits[b - N1] - slack_g;
    };
    auto lanes_of = [&](int b) { return b < N1 ? 1 : MixedGroups::lanes[b - N1]; };
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
        while (b < NB && max_bits(b) < (int)std::max(bits, 2u)) ++b;
        if (b == NB) {
            e->set_error("pgap_prp_batch_mixed: number %zu has %u bits, above the widest bucket (%d)", i, bits,
                         max_bits(NB - 1));
            return PGAP_E_TOO_LARGE;
        }
        by_bucket[(size_t)b].push_back((uint32_t)i);
        total_words = std::max<size_t>(total_words, offset[i] + limbs[i]);
    }
    // tiles of one wavefront (64 / lanes numbers), widest bucket first, in
    // two queues: [0] the lane-group kernel's, [1] the one-lane kernel's
    std::vector<uint32_t> order[2];
    std::vector<uint8_t> tile_bucket[2];
    for (int b = NB - 1; b >= 0; --b) {
        const int q = b < N1 ? 1 : 0;
        const auto &v = by_bucket[(size_t)b];
        const size_t per = 64 / (size_t)lanes_of(b);
        for (size_t k = 0; k < v.size(); k += per) {
            tile_bucket[q].push_back((uint8_t)(b < N1 ? b : b - N1));
            for (size_t l = 0; l < 64; ++l)
                order[q].push_back(l < per && k + l < v.size() ? v[k + l] : 0xffffffffu);
        }
    }
    const uint32_t ntiles[2] = {(uint32_t)tile_bucket[0].size(), (uint32_t)tile_bucket[1].size()};
    (void)hipSetDevice(e->opt.device);
    pgap_status st;
    hipError_t he;
    // The one-lane kernel runs beside the lane-group kernel on a stream of
    // its own, created for this call and destroyed at its end: a process
    // gets a few hardware queues (GPU_MAX_HW_QUEUES, 4 by default), and a
    // stream kept alive would make the streaming scan's two streams share
    // one with another (measured: 9.8 k -> 6.7 k m/s at 1511#).
    struct AuxStream {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        ~AuxStream() {
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
            if (ev) (void)hipEventDestroy(ev);
        }
    } aux;
    if (ntiles[0] && ntiles[1]) {
        if ((he = hipStreamCreateWithFlags(&aux.s, hipStreamNonBlocking)) != hipSuccess ||
            (he = hipEventCreateWithFlags(&aux.ev, hipEventDisableTiming)) != hipSuccess)
            return e->hip_fail(he, "mixed stream");
    }
    const size_t b_words = total_words * 4, b_off = count * 8, b_limbs = count * 4;
    const size_t b_order = (order[0].size() + order[1].size()) * 4;
    if ((st = e->ensure(e->d_n, b_words + b_off + b_limbs + b_order + ntiles[0] + ntiles[1] + 64, "mixed inputs")) ||
        (st = e->ensure(e->d_out, count, "verdicts")) || (st = e->ensure(e->d_small, 64, "tile counters")))
        return st;
    char *base = (char *)e->d_n.p;
    uint32_t *d_words = (uint32_t *)base;
    uint64_t *d_off = (uint64_t *)(base + ((b_words + 7) & ~(size_t)7));
    uint32_t *d_limbs = (uint32_t *)((char *)d_off + b_off);
    uint32_t *d_order0 = (uint32_t *)((char *)d_limbs + b_limbs);
    uint32_t *d_order1 = d_order0 + order[0].size();
    uint8_t *d_tb0 = (uint8_t *)(d_order1 + order[1].size());
    uint8_t *d_tb1 = d_tb0 + ntiles[0];
    uint32_t *d_next = (uint32_t *)e->d_small.p;  // [0]: lane-group queue, [1]: one-lane queue
    if ((he = hipMemcpyAsync(d_words, words, b_words, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(d_off, offset, b_off, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(d_limbs, limbs, b_limbs, hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (order[0].size() && (he = hipMemcpyAsync(d_order0, order[0].data(), order[0].size() * 4,
                                                 hipMemcpyHostToDevice, e->stream)) != hipSuccess) ||
        (order[1].size() && (he = hipMemcpyAsync(d_order1, order[1].data(), order[1].size() * 4,
                                                 hipMemcpyHostToDevice, e->stream)) != hipSuccess) ||
        (ntiles[0] && (he = hipMemcpyAsync(d_tb0, tile_bucket[0].data(), ntiles[0], hipMemcpyHostToDevice,
                                           e->stream)) != hipSuccess) ||
        (ntiles[1] && (he = hipMemcpyAsync(d_tb1, tile_bucket[1].data(), ntiles[1], hipMemcpyHostToDevice,
                                           e->stream)) != hipSuccess) ||
        (he = hipMemsetAsync(d_next, 0, 8, e->stream)) != hipSuccess)
        return e->hip_fail(he, "mixed upload");
    const MixedLaunch lg_launch{d_words, d_off, d_limbs, d_order0, d_tb0, ntiles[0], d_next, bs, (uint8_t *)e->d_out.p};
    const MixedLaunch one_launch{d_words, d_off, d_limbs, d_order1, d_tb1, ntiles[1], d_next + 1, bs,
                                 (uint8_t *)e->d_out.p};
    // the lane-group kernel (the widest tiles) on the engine's stream, the
    // one-lane kernel beside it on the second stream; the second joins the
    // first before the end event
    if ((he = hipEventRecord(e->ev0, e->stream)) != hipSuccess) return e->hip_fail(he, "event");
    const bool both = ntiles[0] && ntiles[1];
    if (ntiles[0]) {
        he = kind == PGAP_FERMAT2 ? launch_mixed_lg_fermat(lg_launch, e->opt.device, e->stream)
                                  : launch_mixed_lg_sprp(lg_launch, e->opt.device, e->stream);
        if (he != hipSuccess) return e->hip_fail(he, "mixed lane-group kernel");
    }
    if (ntiles[1]) {
        hipStream_t s1 = both ? aux.s : e->stream;
        if (both && (he = hipStreamWaitEvent(s1, e->ev0, 0)) != hipSuccess) return e->hip_fail(he, "mixed stream");
        he = kind == PGAP_FERMAT2 ? launch_mixed_fermat(one_launch, e->opt.device, s1)
                                  : launch_mixed_sprp(one_launch, e->opt.device, s1);
        if (he != hipSuccess) return e->hip_fail(he, "mixed kernel");
        if (both && ((he = hipEventRecord(aux.ev, s1)) != hipSuccess ||
                     (he = hipStreamWaitEvent(e->stream, aux.ev, 0)) != hipSuccess))
            return e->hip_fail(he, "mixed stream");
    }
    if ((he = hipEventRecord(e->ev1, e->stream)) != hipSuccess) return e->hip_fail(he, "event");
    e->last.digits = 0;
    e->last.radix_bits = 28;
    e->last.lanes_per_number = 0;  // per bucket
    e->last.launches = (ntiles[0] ? 1u : 0u) + (ntiles[1] ? 1u : 0u);
    e->last.tests = count;
    e->timing_pending = true;
    if ((he = hipMemcpyAsync(verdict, e->d_out.p, count, hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
        (he = hipStreamSynchronize(e->stream)) != hipSuccess)
        return e->hip_fail(he, "mixed readback");
    e->collect_timing();
    return PGAP_OK;
}
