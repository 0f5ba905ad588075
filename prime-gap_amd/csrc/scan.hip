// scan.hip -- streaming prev/next-prime scan: device-resident per-m state,
// continuous batching and HBM-budget admission.
//
// Reference counterparts: the loader/scan state machine load_batch_thread
// (gap_test_gpu.cu:323-655) with its DataM / GPUBatch records
// (gap_test_gpu.cu:139-215), the GPU batch thread run_gpu_thread (:232-268),
// the beyond-SL fallback run_overflow_thread (:270-321), and the per-m scan
// semantics of test_interval_cpu (gap_test_common.cpp:128-181).
//
// Design.  Every admitted m owns a slot: its centre m*K, its prev and next
// unknown lists (outward order) and per-side state (next index to test,
// index of the first confirmed PRP) stay in HBM until the m finishes.  A
// round tests, for every unfinished side, either the next `depth` unknowns
// (Fermat base 2) or, when the side's first Fermat hit is awaiting
// confirmation, that hit's confirmation tests (a strong base-2 test and, for
// Baillie-PSW, a strong Lucas test, all in one round):
//   issue -> materialise N = centre -+ x -> one grid running the round's
//   confirmations and Fermat tests -> retire (first Fermat hit per side,
//   confirmation stage) -> per-slot update, emitting finished m's and
//   exhausted sides.
// Strong-PRP confirmations ride in the next round's grid instead of a launch
// of their own, so a round costs one test's latency; strong Lucas rows (lane
// groups only) run in a launch of their own on a second stream, beside the
// grid, and are joined before the retire step.  A hit
// that fails a confirmation (a base-2 pseudoprime) restarts its side just
// past it.
// Every index below a side's first PRP has been tested in this or an earlier
// round, so with Baillie-PSW confirmation "first confirmed PRP" is the
// reference's first mpz_probab_prime_p hit; tests issued past it in the same round are
// speculative and are not counted (p_tests / n_tests count up to the hit).
// A side whose unknowns are exhausted continues past SL with candidates that
// have no factor below 2^14 -- the prevprime(centre - SL) / nextprime(centre
// + SL) fallback of gap_test_common.cpp:160-178, here on the GPU.
// New requests are admitted between rounds as slots and budget allow, so m's
// that arrive while others are on the device join the next round.  Rounds run
// on an engine-owned worker thread with its own stream from the first submit
// on; submit and poll only exchange requests and rows with it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine_impl.h"
#include "waitq.h"

namespace pgap {

constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t FALLBACK_CHUNK = 4096;  // candidates appended per exhausted side
constexpr uint32_t FALLBACK_SPAN = 65536;  // offsets screened per append
constexpr uint32_t MAX_SLOTS = 1u << 16;  // m's in flight (the HBM budget may allow fewer)

// Latency-first rounds.  A round costs one test's latency whatever it
// carries, and a group of 8 lanes finishes a wide test sooner than one lane
// (2128 bits, 2048 numbers: 11.7 against 40.2 ms) at 59% of its throughput.
// While at most LATENCY_SIDES sides scan, rounds run on the 8-lane variants
// (when the one-lane variant is wide): many short rounds, each side a
// shallower window; above the threshold the one-lane rounds take over
// again.  1511#/2190 with Baillie-PSW confirmation (p50/p99 per m, ms):
// threshold 1024: 78/138 at 0.3 of capacity, 118/332 at 4900 m/s, 169/265
// at half of capacity (5.9 k m/s); 2048: 64/124, 123/364, 193/559; 640:
// 80/139, 164/229, 168/238.
constexpr uint32_t LATENCY_SIDES = 1024;     // scanning sides at most
constexpr uint32_t LATENCY_SIDES_ROW = 256;  // ... and at most this many for 16 lanes instead of 8
constexpr int LATENCY_MIN_DIGITS = 52;       // one-lane variant at least this wide
constexpr uint32_t LUCAS_POOLS = 4;  // Lucas batches that may run at once (a ring)

enum : uint32_t {
    F_PREV_DONE = 1,
    F_NEXT_DONE = 2,
    F_PREV_WAIT = 4,  // exhausted, waiting for the host to append candidates
    F_NEXT_WAIT = 8,
    F_SEQ = 16,        // side-skip mode: the next side starts after the prev side's PRP
    F_NEXT_SKIP = 32,  // side-skip taken: the next side was not scanned
    F_PREV_CONF = 64,  // the pending hit's confirmations are issued and not all back
    F_NEXT_CONF = 128,
    F_FINISHED = 256,  // the finished event was written (a second update pass skips the slot)
    F_PREV_SKIP = 512, // a next-only request: the prev side was not scanned
};

struct SlotState {
    uint32_t cursor[2];
    uint32_t found[2];  // index of the side's confirmed PRP (NONE until then)
    uint32_t len[2];
    uint32_t flags;
    uint32_t cwords;
    uint32_t cand[2];   // first Fermat hit awaiting confirmation (NONE if none)
    uint32_t stage[2];  // confirmations it passed so far
    uint32_t done[2];   // confirmations whose verdict is back
    uint32_t skip_below;  // F_SEQ: a prev distance below this skips the next side
    uint32_t pad;
};

// Event record written for a slot that finished or needs fallback candidates.
struct Event {
    uint32_t slot, code;  // code bit 0: finished; bit 1/2: prev/next exhausted
    uint32_t found[2];
    int32_t x[2];
    uint32_t flags;       // slot flags when the event was written
    uint32_t pad;
};

// Admission / append descriptors (host -> device, one per request).
struct Upload {
    uint32_t slot;
    uint32_t kind;        // 0 = new m, 1 = append to side `side`
    uint32_t side;
    uint32_t cwords;
    uint32_t n0, n1;      // list entries (kind 0: prev, next; kind 1: appended)
    uint32_t off_center;  // word offsets into the data part
    uint32_t off_l0, off_l1;
    uint32_t init_flags;  // kind 0: initial slot flags (F_SEQ)
    uint32_t skip_below;  // kind 0: see SlotState
    uint32_t pad;
};

// ---- kernels --------------------------------------------------------------

__global__ void scatter_kernel(const Upload *__restrict__ up, const uint32_t *__restrict__ data, uint32_t nup,
                               SlotState *__restrict__ st, uint32_t *__restrict__ center, uint32_t cw,
                               int32_t *__restrict__ lists, uint32_t cap) {
    const uint32_t u = blockIdx.x;
    if (u >= nup) return;
    const Upload d = up[u];
    int32_t *l0 = lists + (size_t)d.slot * 2 * cap;
    int32_t *l1 = l0 + cap;
    if (d.kind == 0) {
        uint32_t *c = center + (size_t)d.slot * cw;
        for (uint32_t j = threadIdx.x; j < cw; j += blockDim.x) c[j] = j < d.cwords ? data[d.off_center + j] : 0u;
        // the host bounds n0, n1 by cap; the guard keeps a slot's lists in place regardless
        for (uint32_t j = threadIdx.x; j < min(d.n0, cap); j += blockDim.x) l0[j] = (int32_t)data[d.off_l0 + j];
        for (uint32_t j = threadIdx.x; j < min(d.n1, cap); j += blockDim.x) l1[j] = (int32_t)data[d.off_l1 + j];
        if (threadIdx.x == 0) {
            SlotState s;
            s.cursor[0] = s.cursor[1] = 0;
            s.found[0] = s.found[1] = NONE;
            s.cand[0] = s.cand[1] = NONE;
            s.stage[0] = s.stage[1] = 0;
            s.done[0] = s.done[1] = 0;
            s.len[0] = min(d.n0, cap);
            s.len[1] = min(d.n1, cap);
            s.flags = d.init_flags;
            s.skip_below = d.skip_below;
            s.pad = 0;
            s.cwords = d.cwords;
            st[d.slot] = s;
        }
    } else {
        int32_t *l = d.side ? l1 : l0;
        const uint32_t base = st[d.slot].len[d.side];  // read before the barrier below
        __syncthreads();
        const uint32_t n0 = base < cap ? min(d.n0, cap - base) : 0u;
        for (uint32_t j = threadIdx.x; j < n0; j += blockDim.x) l[base + j] = (int32_t)data[d.off_l0 + j];
        __syncthreads();
        if (threadIdx.x == 0) {
            st[d.slot].len[d.side] = base + n0;
            // both sides of one slot can be appended in the same launch (two
            // blocks): clear this side's bit without losing the other's
            atomicAnd(&st[d.slot].flags, ~(d.side ? F_NEXT_WAIT : F_PREV_WAIT));
        }
    }
}

// N = centre + x (side 1) or centre - x (side 0) as `w` words at o; the
// centre has cw words (zero above).
__device__ __forceinline__ void write_n(const uint32_t *__restrict__ c, uint32_t cw, uint32_t side, uint32_t x,
                                       uint32_t *__restrict__ o, uint32_t w) {
    if (side) {
        uint64_t carry = x;
        for (uint32_t j = 0; j < w; ++j) {
            const uint64_t v = (uint64_t)(j < cw ? c[j] : 0u) + carry;
            o[j] = (uint32_t)v;
            carry = v >> 32;
        }
    } else {
        uint32_t borrow = x;
        for (uint32_t j = 0; j < w; ++j) {
            const uint32_t cj = j < cw ? c[j] : 0u;
            o[j] = cj - borrow;
            borrow = cj < borrow ? 1u : 0u;
        }
    }
}

// What a side does this round: 0 nothing, 1 the confirmation tests of its
// pending hit (every base at once, one row each), 2 Fermat tests of its next
// unknowns.
__device__ __forceinline__ uint32_t side_role(const SlotState &s, uint32_t side) {
    const uint32_t done = side ? F_NEXT_DONE : F_PREV_DONE;
    const uint32_t wait = side ? F_NEXT_WAIT : F_PREV_WAIT;
    if (s.flags & (done | wait)) return 0;
    if (side == 1 && (s.flags & (F_SEQ | F_PREV_DONE)) == F_SEQ) return 0;  // next waits for prev
    if (s.flags & (side ? F_NEXT_CONF : F_PREV_CONF)) return 0;              // verdicts outstanding
    if (s.cand[side] != NONE) return 1;
    return s.cursor[side] < s.len[side] ? 2 : 0;
}

// Share weight of a scanning side in a round: a side that has issued 128
// candidates without a confirmed PRP (the tail of the gap distribution) gets
// twice the share of the others, which cuts the rounds the longest gaps
// take.  1511#/2190 at half of capacity: p99 134 -> 120 ms, p50 unchanged
// (49 ms); 3001#/2190: p99 298 -> 261 ms, p50 133 -> 141 ms.
__device__ __forceinline__ uint32_t side_weight(uint32_t cursor) { return 1u + min(cursor / 128u, 1u); }

// ctr[1] = sides that run Fermat tests this round, ctr[4] = confirmation
// rows, ctr[5] = the scanning sides' total weight, ctr[6] = confirming sides
// (with Baillie-PSW, each starts a strong Lucas test that takes a whole
// wavefront beside the grid; the Fermat rows divide the resident lanes the
// confirmations leave among the scanning sides).
__global__ void count_kernel(const uint32_t *__restrict__ active, uint32_t na, const SlotState *__restrict__ st,
                             uint32_t nconf, uint32_t *__restrict__ ctr) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t side = t & 1u;
    const SlotState *s = t < 2 * na ? &st[active[t >> 1]] : nullptr;
    const uint32_t role = s ? side_role(*s, side) : 0u;
    const uint64_t bs = __ballot(role == 2), bc = __ballot(role == 1);
    uint32_t w = role == 2 ? side_weight(s->cursor[side]) : 0u;
    for (int off = 32; off > 0; off >>= 1) w += (uint32_t)__shfl_xor((int)w, off, 64);
    if ((threadIdx.x & 63u) == 0) {
        if (bs) atomicAdd(&ctr[1], (uint32_t)__popcll(bs));
        if (bc) {
            atomicAdd(&ctr[4], (uint32_t)__popcll(bc) * nconf);
            atomicAdd(&ctr[6], (uint32_t)__popcll(bc));
        }
        if (w) atomicAdd(&ctr[5], w);
    }
}

// One thread per (active slot, side): for a pending Fermat hit, `conf.n`
// strong-PRP confirmation items, one per base (rows [0, cr), counter ctr[3]),
// plus, when `lpool` is set, a strong Lucas item appended to the Lucas pool
// (item, and its number N = centre -+ x written here: `lw` words at
// lnum + index * lw, count *lcount); or Fermat items (rows cr + [0, ctr[0])).  Each scanning side takes its weighted share
// (side_weight, total ctr[5]) of the resident lanes the confirmations (padded
// to whole wavefronts of `cpad` items) leave, at least one, capped by `depth`
// and by the `maxi` rows the round holds: a grid larger than the resident
// lanes would run a second, mostly empty wave and double the round's time.
__global__ void issue_kernel(const uint32_t *__restrict__ active, uint32_t na, SlotState *__restrict__ st,
                             uint32_t cr, uint32_t depth, uint32_t resident, uint32_t cpad, uint32_t maxi, Bases conf,
                             uint32_t *__restrict__ lpool, uint32_t *__restrict__ lcount, uint32_t lres,
                             uint32_t *__restrict__ lnum, uint32_t lw, const uint32_t *__restrict__ center,
                             uint32_t cw, const int32_t *__restrict__ lists, uint32_t cap,
                             uint32_t *__restrict__ items, uint32_t *__restrict__ item_idx,
                             uint8_t *__restrict__ item_base, uint32_t *__restrict__ ctr) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= 2 * na) return;
    const uint32_t slot = active[t >> 1], side = t & 1u;
    SlotState &s = st[slot];
    const uint32_t role = side_role(s, side);
    if (role == 0) return;
    if (role == 1) {
        const uint32_t r = atomicAdd(&ctr[3], conf.n);
        for (uint32_t j = 0; j < conf.n; ++j) {
            items[r + j] = (slot << 1) | side;
            item_idx[r + j] = s.cand[side];
            item_base[r + j] = (uint8_t)conf.b[j];
        }
        if (lpool) {
            const uint32_t l = atomicAdd(lcount, 1u);
            lpool[l] = (slot << 1) | side;
            const uint32_t x = (uint32_t)lists[((size_t)slot * 2 + side) * cap + s.cand[side]];
            write_n(center + (size_t)slot * cw, cw, side, x, lnum + (size_t)l * lw, lw);
        }
        s.stage[side] = 0;
        s.done[side] = 0;
        atomicOr(&s.flags, side ? F_NEXT_CONF : F_PREV_CONF);  // the other side's thread writes flags too
        return;
    }
    const uint32_t c = s.cursor[side], l = s.len[side];
    const uint32_t ns = max(ctr[1], 1u), wsum = max(ctr[5], 1u), w = side_weight(c);
    // confirmation rows in whole wavefronts, and the lanes of the strong Lucas
    // tests that run beside the grid (`lres` lanes each, in units of this
    // grid's items of 64 / cpad lanes)
    const uint32_t used = (ctr[4] + cpad - 1) / cpad * cpad + (lpool ? (ctr[6] * lres * cpad + 63) / 64 : 0u);
    const uint32_t room = resident > used ? resident - used : 0u;
    // weighted shares, at least one item each: the sum stays within room
    // (when room >= ns) and within maxi (maxi >= 2 na >= ns)
    const uint32_t share_room = room > ns ? (uint32_t)((uint64_t)(room - ns) * w / wsum) + 1u : 1u;
    const uint32_t share_max = (uint32_t)((uint64_t)(maxi - ns) * w / wsum) + 1u;
    const uint32_t d = min(min(depth, share_room), share_max);
    const uint32_t k = min(d, l - c);
    const uint32_t base = cr + atomicAdd(&ctr[0], k);
    for (uint32_t j = 0; j < k; ++j) {
        items[base + j] = (slot << 1) | side;
        item_idx[base + j] = c + j;
    }
    s.cursor[side] = c + k;
}

// Rows in use this round: strong-PRP confirmations [0, ctr[3]), Fermat
// cr + [0, ctr[0]).
__device__ __forceinline__ bool row_live(uint32_t t, uint32_t cr, const uint32_t *ctr) {
    return t < cr ? t < ctr[3] : t - cr < ctr[0];
}

// N = centre -+ x for every item, `w` words per row.
__global__ void materialize_kernel(const uint32_t *__restrict__ items, const uint32_t *__restrict__ item_idx,
                                   uint32_t cr, uint32_t nrows, const uint32_t *__restrict__ ctr,
                                   const uint32_t *__restrict__ center, uint32_t cw,
                                   const int32_t *__restrict__ lists, uint32_t cap, uint32_t *__restrict__ nbuf,
                                   uint32_t w) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= nrows || !row_live(t, cr, ctr)) return;
    const uint32_t slot = items[t] >> 1, side = items[t] & 1u;
    const uint32_t x = (uint32_t)lists[((size_t)slot * 2 + side) * cap + item_idx[t]];
    write_n(center + (size_t)slot * cw, cw, side, x, nbuf + (size_t)t * w, w);
}

// Confirmation rows count the tests their side's pending hit passed and the
// verdicts back (the update decides); Fermat hits become the side's pending
// hit (or, with no confirmation, its PRP).  A side has either confirmation
// rows or Fermat rows in a round, never both.
__global__ void retire_kernel(const uint32_t *__restrict__ items, const uint32_t *__restrict__ item_idx,
                              uint32_t cr, uint32_t nrows, const uint8_t *__restrict__ verdict,
                              const uint32_t *__restrict__ ctr, uint32_t confirm, SlotState *__restrict__ st) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= nrows || !row_live(t, cr, ctr)) return;
    const uint32_t slot = items[t] >> 1, side = items[t] & 1u;
    SlotState &s = st[slot];
    if (t < cr) {
        if (verdict[t]) atomicAdd(&s.stage[side], 1u);
        atomicAdd(&s.done[side], 1u);
        return;
    }
    if (!verdict[t]) return;
    atomicMin(confirm ? &s.cand[side] : &s.found[side], item_idx[t]);
}

// A finished strong Lucas batch: each item's verdict counts toward its
// side's pending confirmations.
__global__ void lucas_retire_kernel(const uint32_t *__restrict__ pool, const uint8_t *__restrict__ verdict,
                                    uint32_t count, SlotState *__restrict__ st) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= count) return;
    const uint32_t slot = pool[t] >> 1, side = pool[t] & 1u;
    if (verdict[t]) atomicAdd(&st[slot].stage[side], 1u);
    atomicAdd(&st[slot].done[side], 1u);
}

// Counter reset as a kernel (the round sequence carries no memset nodes).
__global__ void zero_word_kernel(uint32_t *__restrict__ p) {
    if (threadIdx.x == 0) *p = 0;
}

__global__ void update_kernel(const uint32_t *__restrict__ active, uint32_t na, SlotState *__restrict__ st,
                              const int32_t *__restrict__ lists, uint32_t cap, uint32_t nconf,
                              uint32_t *__restrict__ ctr, Event *__restrict__ events) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= na) return;
    const uint32_t slot = active[t];
    SlotState &s = st[slot];
    uint32_t flags = s.flags, code = 0;
    if (flags & F_FINISHED) return;
    // every confirmation verdict back: all passed -> the side's PRP; any
    // failed (a base-2 pseudoprime) -> resume just past it
    for (uint32_t side = 0; side < 2; ++side) {
        const uint32_t conf = side ? F_NEXT_CONF : F_PREV_CONF;
        if (!(flags & conf) || s.done[side] < nconf) continue;
        flags &= ~conf;
        if (s.stage[side] >= nconf) {
            s.found[side] = s.cand[side];
        } else {
            s.cursor[side] = s.cand[side] + 1;
            s.cand[side] = NONE;
        }
        s.stage[side] = 0;
        s.done[side] = 0;
    }
    for (uint32_t side = 0; side < 2; ++side) {
        const uint32_t done = side ? F_NEXT_DONE : F_PREV_DONE;
        const uint32_t wait = side ? F_NEXT_WAIT : F_PREV_WAIT;
        if (flags & (done | wait)) continue;
        if (side == 1 && (flags & (F_SEQ | F_PREV_DONE)) == F_SEQ) continue;  // not started
        if (s.found[side] != NONE) {
            flags |= done;
            // side-skip (gap_test_gpu.cu:591-617): a small prev gap ends the m
            if (side == 0 && (flags & F_SEQ) &&
                (uint32_t)lists[(size_t)slot * 2 * cap + s.found[0]] < s.skip_below)
                flags |= F_NEXT_DONE | F_NEXT_SKIP;
        } else if (s.cand[side] == NONE && s.cursor[side] >= s.len[side]) {
            flags |= wait;
            code |= 2u << side;
        }
    }
    if ((flags & (F_PREV_DONE | F_NEXT_DONE)) == (F_PREV_DONE | F_NEXT_DONE)) {
        code |= 1u;
        flags |= F_FINISHED;
    }
    s.flags = flags;
    if (!code) return;
    Event ev;
    ev.slot = slot;
    ev.code = code;
    for (int side = 0; side < 2; ++side) {
        ev.found[side] = s.found[side];
        ev.x[side] = s.found[side] != NONE ? lists[((size_t)slot * 2 + side) * cap + s.found[side]] : 0;
    }
    ev.flags = flags;
    ev.pad = 0;
    events[atomicAdd(&ctr[2], 1u)] = ev;
}

// ---- host side --------------------------------------------------------------

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// What the worker needs of a family, copied at the family's first submit so
// the worker never reads the engine's family table (the client thread may
// append to it at any time).
struct FamInfo {
    int id = 0;
    Words K;
    uint32_t SL = 0;
    double K_log = 0;
    std::vector<uint32_t> kmod;  // K mod p over the fallback primes (worker, lazily)
};

struct SlotHost {
    bool used = false;
    FamInfo *fam = nullptr;
    uint64_t m = 0, tag = 0;
    uint32_t orig_len[2] = {0, 0};
    uint32_t len[2] = {0, 0};
    uint32_t fb_from[2] = {0, 0};  // next offset the fallback screen starts at
    bool fb[2] = {false, false};
    double t_submit = 0;
    uint32_t bits = 0;             // bit length of the centre
    uint64_t room = 0;             // 2^bits - centre, saturated at 2^40
    uint64_t small = 0;            // the centre when below 2^63, else 0
};

// The centre's value when it is below 2^63, else 0.
static uint64_t small_value(const Words &c) {
    if (bitlen(c) > 63) return 0;
    return (uint64_t)(c.size() > 0 ? c[0] : 0) | ((uint64_t)(c.size() > 1 ? c[1] : 0) << 32);
}

// 2^bitlen(c) - c, saturated at 2^40: next-side candidates centre + x with
// x < room keep the centre's bit length.
static uint64_t room_above(const Words &c) {
    constexpr uint64_t kCap = 1ull << 40;
    Words d = c;
    add_signed(d, (int64_t)kCap);
    const uint32_t b = bitlen(c);
    if (bitlen(d) == b) return kCap;
    const uint64_t lo = (uint64_t)(c.size() > 0 ? c[0] : 0) | ((uint64_t)(c.size() > 1 ? c[1] : 0) << 32);
    return b >= 64 ? 0 - lo : (1ull << b) - lo;  // < 2^40: the low 64 bits decide it
}

struct Staged {
    FamInfo *fam;
    uint64_t m, tag;
    uint32_t sides;  // PGAP_SIDES_*
    std::vector<int32_t> prev, next;
    double t_submit;
};

class Scanner {
  public:
    pgap_engine *e;
    uint32_t depth = 256, confirm = PGAP_CONFIRM_BPSW;
    bool side_skip = false;       // gap_test_gpu's prev-first order with side-skip
    float skip_min_merit = 0;     // min_merit / 2 - 2 (float, gap_test_gpu.cu:604)
    uint32_t S = 0, CW = 0, CAP = 0, W = 0, MAXI = 0;
    uint32_t SL_MAX = 0;  // largest sieve length the slots' lists were sized for
    bool inited = false;

    // Shared with the client thread, under mu: staged requests, ready rows,
    // the admitted count, the family copies, telemetry and the worker's
    // status.  Everything else belongs to the worker alone.
    std::mutex mu;
    std::condition_variable cv_work;   // worker: requests arrived / stop
    std::condition_variable cv_ready;  // client: rows ready / worker stopped
    std::deque<Staged> staged;
    std::vector<pgap_scan_row> ready;
    size_t admitted = 0;               // m's in slots
    std::map<int, FamInfo> fams;       // node addresses stay valid across inserts
    pgap_status worker_status = PGAP_OK;
    std::string worker_err;
    bool stop = false;
    std::thread worker;
    hipStream_t stream = nullptr;      // the worker's stream
    // Strong Lucas confirmations (Baillie-PSW) are not part of a round's
    // grid: a Lucas test costs more than a Fermat test, and a round is sized
    // to one.  A round's Lucas tests (one per confirming side) go to a pool
    // of a ring of LUCAS_POOLS (the item and its number), which starts as a
    // batch right after the issue step on a stream of its own (default
    // priority; batches run there in order); finished batches are retired,
    // oldest first, at the end of a round, before a second update pass, so
    // the m's they decide finish in that round.  A side waits for all its
    // confirmations while the other sides keep scanning.
    hipStream_t lstream = nullptr;
    hipEvent_t ev_rows = nullptr, ev_lucas[LUCAS_POOLS] = {};
    uint32_t *d_lpool = nullptr, *d_lnum = nullptr, *d_lctr = nullptr;
    uint8_t *d_lverd = nullptr;
    uint32_t LP = 0;                       // items per pool (two per slot)
    uint32_t lhead = 0, lbusy = 0;         // oldest running pool, pools running
    uint32_t lcount[LUCAS_POOLS] = {};
    // latency-first rounds are chosen by the sides that scan: the last
    // round's count plus two per m admitted since
    uint32_t last_scanning = 0, admitted_last = 0;

    std::vector<SlotHost> slots;
    std::vector<uint32_t> free_slots, active;
    std::vector<Upload> pending_append;      // fallback appends for the next round
    std::vector<std::vector<uint32_t>> pending_append_data;

    // small primes for the fallback screen
    std::vector<uint32_t> fb_primes;

    SlotState *d_state = nullptr;
    uint32_t *d_center = nullptr, *d_active = nullptr, *d_items = nullptr, *d_idx = nullptr, *d_nbuf = nullptr,
             *d_ctr = nullptr, *d_updata = nullptr;
    int32_t *d_lists = nullptr;
    uint8_t *d_verdict = nullptr, *d_base = nullptr;
    Event *d_events = nullptr;
    Upload *d_up = nullptr;
    size_t up_cap = 0, updata_cap = 0;

    // pinned host staging
    uint32_t *h_active = nullptr, *h_ctr = nullptr;  // h_ctr mirrors the 8 round counters
    Event *h_events = nullptr;
    Upload *h_up = nullptr;
    uint32_t *h_updata = nullptr;

    uint64_t tests_issued = 0, rounds = 0;  // under mu

    explicit Scanner(pgap_engine *eng) : e(eng) {}
    ~Scanner() {
        shutdown();
        release();
    }

    // Stop the worker (it finishes the round in progress) and wait for it.
    void shutdown() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
    }

    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : {(void *)d_state, (void *)d_center, (void *)d_active, (void *)d_items, (void *)d_idx,
                        (void *)d_nbuf, (void *)d_base, (void *)d_ctr, (void *)d_lists, (void *)d_verdict,
                        (void *)d_events, (void *)d_up, (void *)d_updata})
            if (p) (void)hipFree(p);
        for (void *p : {(void *)h_active, (void *)h_ctr, (void *)h_events, (void *)h_up, (void *)h_updata})
            if (p) (void)hipHostFree(p);
        d_state = nullptr;
        d_center = d_active = d_items = d_idx = d_nbuf = d_ctr = d_updata = nullptr;
        d_lists = nullptr;
        d_verdict = d_base = nullptr;
        d_events = nullptr;
        d_up = nullptr;
        h_active = h_ctr = h_updata = nullptr;
        h_events = nullptr;
        h_up = nullptr;
        if (lstream) (void)hipStreamSynchronize(lstream);
        for (void *p : {(void *)d_lpool, (void *)d_lnum, (void *)d_lctr, (void *)d_lverd})
            if (p) (void)hipFree(p);
        d_lpool = d_lnum = d_lctr = nullptr;
        d_lverd = nullptr;
        lhead = lbusy = 0;
        last_scanning = admitted_last = 0;
        for (uint32_t p = 0; p < LUCAS_POOLS; ++p) {
            lcount[p] = 0;
            if (ev_lucas[p]) (void)hipEventDestroy(ev_lucas[p]);
            ev_lucas[p] = nullptr;
        }
        if (ev_rows) (void)hipEventDestroy(ev_rows);
        if (lstream) (void)hipStreamDestroy(lstream);
        if (stream) (void)hipStreamDestroy(stream);
        ev_rows = nullptr;
        lstream = nullptr;
        stream = nullptr;
        inited = false;
    }

    size_t in_flight_locked() const { return staged.size() + admitted + ready.size(); }
    size_t in_flight() {
        std::lock_guard<std::mutex> g(mu);
        return in_flight_locked();
    }

    void worker_main() {
        (void)hipSetDevice(e->opt.device);
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !staged.empty() || admitted > 0; });
                if (stop) return;
            }
            const pgap_status st = round();
            if (st != PGAP_OK) {
                std::lock_guard<std::mutex> g(mu);
                worker_status = st;
                worker_err = pgap_last_error(e);
                cv_ready.notify_all();
                return;
            }
        }
    }

    void start_worker() {
        if (!worker.joinable()) worker = std::thread([this] { worker_main(); });
    }

    template <class T>
    pgap_status dalloc(T **p, size_t n, const char *what) {
        hipError_t he = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
        if (he != hipSuccess) {
            (void)hipGetLastError();
            e->set_error("scan: hipMalloc %s (%zu bytes): %s", what, n * sizeof(T), hipGetErrorString(he));
            return PGAP_E_OOM;
        }
        return PGAP_OK;
    }
    template <class T>
    pgap_status halloc(T **p, size_t n, const char *what) {
        hipError_t he = hipHostMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T), 0);
        if (he != hipSuccess) {
            (void)hipGetLastError();
            e->set_error("scan: hipHostMalloc %s: %s", what, hipGetErrorString(he));
            return PGAP_E_OOM;
        }
        return PGAP_OK;
    }

    // Size the slot arena from the families registered so far and the budget.
    pgap_status init() {
        if (inited) return PGAP_OK;
        if (e->families.empty()) {
            e->set_error("scan: register a family first (pgap_family)");
            return PGAP_E_STATE;
        }
        uint32_t kw = 0, sl = 0;
        for (const Family &f : e->families) {
            kw = std::max<uint32_t>(kw, (uint32_t)f.K.size());
            sl = std::max(sl, f.SL);
        }
        CW = kw + 2;   // m < 2^64
        W = CW + 1;    // room for the carry of + x
        SL_MAX = std::max<uint32_t>(sl, 1);
        CAP = SL_MAX + 4 * FALLBACK_CHUNK;
        // A round carries about as many Fermat tests as the chosen variant
        // keeps resident (the depth rule in round()), plus at most one per
        // side when depth 1 already exceeds that: the row buffer is sized by
        // the largest resident count over the variants plus two rows per
        // slot, not by slots x depth.
        size_t resident_max = 0;
        for (size_t i = 0; i < variant_count(); ++i) {
            const Variant *v = variant_at(i);
            const size_t bpc = (size_t)std::max(1, v->blocks_per_cu ? v->blocks_per_cu() : 1);
            resident_max = std::max(resident_max, bpc * (size_t)e->num_cus * 256 / (size_t)std::max(1, v->lanes));
        }
        const size_t row_bytes = 4ull * W + 4 + 4 + 1;  // number, item, index, verdict
        // per slot: state, centre, lists, active entry, event, and per side
        // one strong-PRP confirmation row (+ its base), one Fermat row and
        // an item in each strong Lucas pool
        const size_t nconf = 1, nlucas = confirm == PGAP_CONFIRM_BPSW ? 1 : 0;
        const size_t per_slot = sizeof(SlotState) + 4ull * CW + 8ull * CAP + 4 + sizeof(Event) +
                                2ull * (nconf * (row_bytes + 1) + row_bytes) + nlucas * 2ull * LUCAS_POOLS * row_bytes;
        size_t budget = e->opt.hbm_budget;
        if (budget == 0) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)8 << 30;
            budget = fr / 4;
        }
        // speculation rows take at most half of the budget
        const size_t extra = std::min(resident_max, budget / 2 / row_bytes);
        S = (uint32_t)std::min<size_t>(MAX_SLOTS, (budget - extra * row_bytes) / per_slot);
        if (S == 0) {
            e->set_error("scan: HBM budget %zu bytes cannot hold one m (%zu bytes per slot)", budget, per_slot);
            return PGAP_E_TOO_LARGE;
        }
        MAXI = (uint32_t)std::min<size_t>(extra + 2ull * S, 0x7fffffffu);
        // strong-PRP confirmation rows, then Fermat rows
        const size_t rows = 2ull * S * nconf + MAXI;
        LP = 2 * S;
        pgap_status st;
        if ((st = dalloc(&d_state, S, "state")) || (st = dalloc(&d_center, (size_t)S * CW, "centres")) ||
            (st = dalloc(&d_lists, (size_t)S * 2 * CAP, "lists")) || (st = dalloc(&d_active, S, "active")) ||
            (st = dalloc(&d_items, rows, "items")) || (st = dalloc(&d_idx, rows, "item index")) ||
            (st = dalloc(&d_nbuf, rows * W, "numbers")) || (st = dalloc(&d_verdict, rows, "verdicts")) ||
            (st = dalloc(&d_base, 2ull * S * nconf, "confirmation bases")) || (st = dalloc(&d_ctr, 8, "counters")) ||
            (st = dalloc(&d_events, S, "events")) || (st = halloc(&h_active, S, "active")) ||
            (st = halloc(&h_ctr, 16, "counters")) || (st = halloc(&h_events, S, "events")) ||
            (nlucas && ((st = dalloc(&d_lpool, (size_t)LUCAS_POOLS * LP, "Lucas pools")) ||
                        (st = dalloc(&d_lnum, (size_t)LUCAS_POOLS * LP * W, "Lucas pool numbers")) ||
                        (st = dalloc(&d_lverd, (size_t)LUCAS_POOLS * LP, "Lucas pool verdicts")) ||
                        (st = dalloc(&d_lctr, 64, "Lucas pool counters"))))) {
            release();
            return st;
        }
        hipError_t he;
        if ((he = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess ||
            (he = hipStreamCreateWithFlags(&lstream, hipStreamNonBlocking)) != hipSuccess ||
            (he = hipEventCreateWithFlags(&ev_rows, hipEventDisableTiming)) != hipSuccess ||
            (d_lctr && ((he = hipMemsetAsync(d_lctr, 0, 64 * sizeof(uint32_t), stream)) != hipSuccess ||
                        (he = hipStreamSynchronize(stream)) != hipSuccess))) {
            release();
            return e->hip_fail(he, "scan streams");
        }
        for (uint32_t p = 0; p < LUCAS_POOLS; ++p) {
            if ((he = hipEventCreateWithFlags(&ev_lucas[p], hipEventDisableTiming)) != hipSuccess) {
                release();
                return e->hip_fail(he, "scan streams");
            }
        }
        slots.assign(S, SlotHost());
        free_slots.clear();
        for (uint32_t s = S; s-- > 0;) free_slots.push_back(s);
        // small primes for the beyond-SL screen (K mod p per family on demand)
        fb_primes = primes_upto(1u << 14);
        inited = true;
        return PGAP_OK;
    }

    // Smallest prev distance d whose merit, computed as gap_test_gpu.cu:595
    // does (float)(d / (K_log + log m)), is not below skip_min_merit: a prev
    // PRP nearer than this skips the next side.
    uint32_t skip_distance(double K_log, uint64_t m) const {
        const double den = K_log + std::log((double)m);
        const auto below = [&](double d) { return (float)(d / den) < skip_min_merit; };
        double d = std::floor((double)skip_min_merit * den) - 2;
        if (d < 0) d = 0;
        while (d > 0 && !below(d - 1)) d -= 1;
        while (below(d)) d += 1;
        return d > 4294967295.0 ? 0xffffffffu : (uint32_t)d;
    }

    const std::vector<uint32_t> &kmod_small(FamInfo *f) {
        if (f->kmod.empty()) {
            f->kmod.resize(fb_primes.size());
            for (size_t i = 0; i < fb_primes.size(); ++i) f->kmod[i] = mod_small(f->K, fb_primes[i]);
        }
        return f->kmod;
    }

    uint32_t depth_max() const { return std::max<uint32_t>(depth, 1); }

    pgap_status ensure_upload(size_t nup, size_t nwords) {
        if (nup > up_cap) {
            if (d_up) (void)hipFree(d_up);
            if (h_up) (void)hipHostFree(h_up);
            d_up = nullptr;
            h_up = nullptr;
            up_cap = 0;
            const size_t c = std::max<size_t>(nup, 256);
            pgap_status st;
            if ((st = dalloc(&d_up, c, "upload descriptors")) || (st = halloc(&h_up, c, "upload descriptors")))
                return st;
            up_cap = c;
        }
        if (nwords > updata_cap) {
            if (d_updata) (void)hipFree(d_updata);
            if (h_updata) (void)hipHostFree(h_updata);
            d_updata = nullptr;
            h_updata = nullptr;
            updata_cap = 0;
            const size_t c = std::max<size_t>(nwords, (size_t)1 << 20);
            pgap_status st;
            if ((st = dalloc(&d_updata, c, "upload data")) || (st = halloc(&h_updata, c, "upload data")))
                return st;
            updata_cap = c;
        }
        return PGAP_OK;
    }

    // Screen offsets beyond SL for one exhausted side: no factor below 2^14
    // (N itself a prime below 2^14 passes).  The prev side stops at N = 3:
    // the scan's tests take odd N, and a centre below SL + 4 is refused at
    // submit, so a prime >= 3 always lies in that range.  `end_of_range` is
    // set when the prev side has nothing left.
    std::vector<uint32_t> fallback_candidates(SlotHost &sh, int side, bool *end_of_range) {
        const std::vector<uint32_t> &km = kmod_small(sh.fam);
        std::vector<uint32_t> out;
        uint32_t x = sh.fb_from[side];
        uint64_t end = (uint64_t)x + FALLBACK_SPAN;
        if (side == 0 && sh.small) end = std::min<uint64_t>(end, sh.small - 2);  // x <= centre - 3
        *end_of_range = side == 0 && sh.small && x >= end;
        if (x >= end) return out;
        // sieve [x, end): N = m*K - x (prev) or m*K + x (next) is divisible by p
        // iff x == m*K (prev) or x == -m*K (next) mod p; N itself a small
        // prime (small centres only) stays
        const uint32_t span = (uint32_t)(end - x);
        std::vector<uint8_t> comp(span, 0);
        for (size_t i = 0; i < fb_primes.size(); ++i) {
            const uint32_t p = fb_primes[i];
            const uint32_t r = (uint32_t)((sh.m % p) * km[i] % p);
            const uint32_t hit = side ? (p - r) % p : r;                  // x mod p of the multiples
            uint64_t j = (uint64_t)((hit + p - x % p) % p);               // first index with x + j == hit
            for (; j < span; j += p) {
                const uint64_t xx = x + j;
                const uint64_t N = sh.small ? (side ? sh.small + xx : sh.small - xx) : ~0ull;
                if (N != p) comp[j] = 1;
            }
        }
        uint32_t j = 0;
        for (; j < span && out.size() < FALLBACK_CHUNK; ++j)
            if (!comp[j]) out.push_back(x + j);
        sh.fb_from[side] = x + j;
        return out;
    }

    pgap_status admit() {
        std::vector<Staged> take;
        {
            std::lock_guard<std::mutex> g(mu);
            while (!staged.empty() && take.size() < free_slots.size()) {
                take.push_back(std::move(staged.front()));
                staged.pop_front();
            }
            admitted += take.size();  // no window where an m is neither staged nor admitted
        }
        admitted_last = (uint32_t)take.size();
        const size_t nup = take.size() + pending_append.size();
        if (nup == 0) return PGAP_OK;
        size_t words = 0;
        for (const Staged &r : take) words += CW + r.prev.size() + r.next.size();
        for (const auto &d : pending_append_data) words += d.size();
        pgap_status st;
        if ((st = ensure_upload(nup, words))) return st;
        size_t u = 0, off = 0;
        for (Staged &r : take) {
            const uint32_t slot = free_slots.back();
            free_slots.pop_back();
            const FamInfo &f = *r.fam;
            Words c = f.K;
            mul_small(c, r.m);
            SlotHost &sh = slots[slot];
            sh = SlotHost();
            sh.used = true;
            sh.fam = r.fam;
            sh.m = r.m;
            sh.tag = r.tag;
            sh.orig_len[0] = sh.len[0] = (uint32_t)r.prev.size();
            sh.orig_len[1] = sh.len[1] = (uint32_t)r.next.size();
            sh.fb_from[0] = sh.fb_from[1] = f.SL + 1;
            sh.t_submit = r.t_submit;
            sh.bits = bitlen(c);
            sh.room = room_above(c);
            sh.small = small_value(c);
            Upload d{};
            d.slot = slot;
            d.kind = 0;
            d.cwords = (uint32_t)c.size();
            if (side_skip) {
                d.init_flags = F_SEQ;
                d.skip_below = skip_distance(f.K_log, r.m);
            }
            // a one-sided request starts with the other side finished
            if (r.sides == PGAP_SIDES_PREV) d.init_flags |= F_NEXT_DONE | F_NEXT_SKIP;
            if (r.sides == PGAP_SIDES_NEXT) d.init_flags |= F_PREV_DONE | F_PREV_SKIP;
            d.n0 = sh.len[0];
            d.n1 = sh.len[1];
            d.off_center = (uint32_t)off;
            memcpy(h_updata + off, c.data(), c.size() * 4);
            off += CW;
            d.off_l0 = (uint32_t)off;
            memcpy(h_updata + off, r.prev.data(), r.prev.size() * 4);
            off += r.prev.size();
            d.off_l1 = (uint32_t)off;
            memcpy(h_updata + off, r.next.data(), r.next.size() * 4);
            off += r.next.size();
            h_up[u++] = d;
            active.push_back(slot);
        }
        for (size_t i = 0; i < pending_append.size(); ++i) {
            Upload d = pending_append[i];
            d.off_l0 = (uint32_t)off;
            const auto &v = pending_append_data[i];
            memcpy(h_updata + off, v.data(), v.size() * 4);
            off += v.size();
            h_up[u++] = d;
        }
        pending_append.clear();
        pending_append_data.clear();
        hipError_t he;
        if ((he = hipMemcpyAsync(d_up, h_up, nup * sizeof(Upload), hipMemcpyHostToDevice, stream)) !=
                hipSuccess ||
            (he = hipMemcpyAsync(d_updata, h_updata, std::max<size_t>(off, 1) * 4, hipMemcpyHostToDevice,
                                 stream)) != hipSuccess)
            return e->hip_fail(he, "scan upload");
        hipLaunchKernelGGL(scatter_kernel, dim3((uint32_t)nup), dim3(256), 0, stream, d_up, d_updata,
                           (uint32_t)nup, d_state, d_center, CW, d_lists, CAP);
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "scatter kernel");
        return PGAP_OK;
    }

    // The variant for a Lucas batch of tests of up to `bits` bits: one
    // number per wavefront (a step's three products at once), whatever the
    // batch size.
    // Its tests finish soonest, and beside the rounds' grids it also gives
    // the larger capacity: 1511#/2190 11.3-11.7 k m/s against 9.7-9.8 k
    // with the 8-lane variant for batches above 256 tests (3001#/2190: 453-467
    // against 430-439 m/s), at the same latency for the same arrival rate
    // (p50/p99 125/367 against 122/358 ms at 4900 m/s).
    pgap_status lucas_choice(uint32_t bits, Choice *c) {
        const Variant *vw =
            e->opt.lanes_per_number || e->opt.radix ? nullptr : find_variant(KIND_LUCAS, (int)bits, 28, 64);
        if (vw) {
            *c = Choice{vw, 0};
            return PGAP_OK;
        }
        return e->pick(KIND_LUCAS, bits, c);
    }

    // Start pool p (n items, written by the issue step already queued on the
    // round's stream) as a batch on the Lucas stream.
    pgap_status lucas_launch(uint32_t p, const Choice &c, uint32_t n) {
        Bases none{};
        hipError_t he;
        if ((he = hipEventRecord(ev_rows, stream)) != hipSuccess ||
            (he = hipStreamWaitEvent(lstream, ev_rows, 0)) != hipSuccess ||
            (he = c.v->launch(d_lnum + (size_t)p * LP * W, W, n, nullptr, nullptr, none, d_lverd + (size_t)p * LP,
                              c.cap, lstream)) != hipSuccess ||
            (he = hipEventRecord(ev_lucas[p], lstream)) != hipSuccess)
            return e->hip_fail(he, "strong Lucas batch launch");
        lcount[p] = n;
        ++lbusy;
        return PGAP_OK;
    }

    // Retire the finished Lucas batches, oldest first (`wait`: wait for the
    // oldest one): their verdicts count toward their sides' confirmations.
    // With `update`, queue a second update pass on the round's stream when
    // any was retired (at the end of a round; at its start the round's own
    // update decides).  `any`: whether one was.
    pgap_status lucas_retire(uint32_t na, uint32_t nconf, bool wait, bool update, bool *any, uint64_t *issued) {
        *any = false;
        while (lbusy) {
            const uint32_t p = lhead;
            hipError_t he;
            if (wait && !*any) {
                if ((he = hipEventSynchronize(ev_lucas[p])) != hipSuccess)
                    return e->hip_fail(he, "strong Lucas batch");
            } else if ((he = hipEventQuery(ev_lucas[p])) == hipErrorNotReady) {
                break;
            } else if (he != hipSuccess) {
                return e->hip_fail(he, "strong Lucas batch");
            }
            const uint32_t n = lcount[p];
            const dim3 b256(256);
            hipLaunchKernelGGL(lucas_retire_kernel, dim3((n + 255) / 256), b256, 0, stream,
                               d_lpool + (size_t)p * LP, d_lverd + (size_t)p * LP, n, d_state);
            hipLaunchKernelGGL(zero_word_kernel, dim3(1), dim3(64), 0, stream, d_lctr + p);
            *issued += n;
            lcount[p] = 0;
            lhead = (lhead + 1) % LUCAS_POOLS;
            --lbusy;
            *any = true;
        }
        if (*any && update) {
            hipLaunchKernelGGL(update_kernel, dim3((na + 255) / 256), dim3(256), 0, stream, d_active, na, d_state,
                               d_lists, CAP, nconf, d_ctr, d_events);
            hipError_t he = hipGetLastError();
            if (he != hipSuccess) return e->hip_fail(he, "strong Lucas retire");
        }
        return PGAP_OK;
    }

    pgap_status round() {
        pgap_status st;
        if ((st = admit())) return st;
        if (active.empty()) return PGAP_OK;
        const uint32_t na = (uint32_t)active.size();
        // widest N this round: a next-side candidate centre + x gains a bit
        // only when x reaches the room above the centre (x stays below the
        // sieve length or the fallback screen's next offset)
        uint32_t bits = 0;
        for (uint32_t s : active) {
            const SlotHost &sh = slots[s];
            const uint64_t xmax = std::max<uint64_t>(sh.fam->SL, std::max(sh.fb_from[0], sh.fb_from[1]));
            bits = std::max(bits, sh.bits + (xmax >= sh.room ? 1u : 0u));
        }
        // confirmation of a side's first Fermat PRP: a strong base-2 row in
        // the round's grid, plus a strong Lucas row for Baillie-PSW
        Bases none{}, conf{};
        if (confirm != PGAP_CONFIRM_NONE) {
            conf.n = 1;
            conf.b[0] = 2;
        }
        const bool lucas = confirm == PGAP_CONFIRM_BPSW;
        Choice cf, cc;
        if ((st = e->pick(KIND_FERMAT2, bits, &cf))) return st;
        if (conf.n && (st = e->pick(KIND_SPRP, bits, &cc))) return st;
        // sides that will scan this round (m's waiting for confirmations do not)
        const uint32_t sides = (uint32_t)std::min<uint64_t>(2ull * na, (uint64_t)last_scanning + 2ull * admitted_last);
        if (sides <= LATENCY_SIDES && !e->opt.lanes_per_number && !e->opt.radix) {
            // wide one-lane numbers move to 8 lanes (16 while very few sides
            // scan), 4- and 8-lane numbers to 16
            const bool row = sides <= LATENCY_SIDES_ROW;
            const int lanes = cf.v->lanes == 1 && cf.v->digits >= LATENCY_MIN_DIGITS ? (row ? 16 : 8)
                              : cf.v->lanes == 4 || cf.v->lanes == 8                   ? 16
                                                                                       : 0;
            const Variant *lf = lanes ? find_variant(KIND_FERMAT2, (int)bits, 28, lanes) : nullptr;
            const Variant *lc = lanes && conf.n ? find_variant(KIND_SPRP, (int)bits, 28, lanes) : nullptr;
            if (lf && (!conf.n || (lc && lc->scan))) {
                cf = Choice{lf, 0};
                cc = Choice{lc, 0};
            }
        }
        const Variant *vf = cf.v, *vc = cc.v;
        if (vc && !vc->scan) {
            e->set_error("scan: variant radix %d x %d has no scan kernel", vc->radix, vc->digits);
            return PGAP_E_STATE;
        }
        const int cap_f = cf.cap, cap_c = cc.cap;
        // Speculative depth.  A round costs about one test's latency whether
        // it carries one item or a full device's worth (the kernel is VALU
        // bound per lane), so every round is sized to the numbers the chosen
        // variant keeps resident at once: few active sides get a deep
        // speculative window, many get one or two tests each.
        //
        // When the variant keeps more than one wavefront per SIMD resident,
        // a full round takes that many tests' time; while the scanning sides
        // are few (at most a quarter of one wave per SIMD's lanes) the round
        // is held to one wave per SIMD, so it costs one test's latency and
        // each side still gets a deep window.
        const Variant *vr = vc ? vc : vf;
        int (*occ)() = vc ? vc->scan_blocks_per_cu : vf->blocks_per_cu;
        const uint32_t per_wave = e->num_cus * 256u / (uint32_t)std::max(1, vr->lanes);
        const uint32_t resident_full = (uint32_t)std::max(1, occ ? occ() : 1) * per_wave;
        const uint32_t resident = 4ull * sides >= per_wave ? resident_full : std::min(resident_full, per_wave);
        const uint32_t cpad = 64u / (uint32_t)std::max(1, vr->lanes);  // items per wavefront

        Choice lc;
        uint32_t lres = 0, lpool_now = 0, nluc = 0;
        uint64_t issued = 0;
        if (lucas && lbusy == LUCAS_POOLS) {
            // every pool still running: wait for the oldest (the round's own
            // update decides the m's it confirms)
            bool any = false;
            if ((st = lucas_retire(na, conf.n + 1, true, false, &any, &issued))) return st;
        }
        if (lucas) lpool_now = (lhead + lbusy) % LUCAS_POOLS;

        memcpy(h_active, active.data(), na * 4);
        hipError_t he;
        if ((he = hipMemcpyAsync(d_active, h_active, na * 4, hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (he = hipMemsetAsync(d_ctr, 0, 8 * sizeof(uint32_t), stream)) != hipSuccess)
            return e->hip_fail(he, "scan round set-up");
        const uint32_t cr = 2 * na * std::max<uint32_t>(conf.n, 1);  // confirmation rows: every base of a side
        // Fermat rows: the per-side depth (set on the device from the number
        // of scanning sides) gives at most resident + one row per side
        const uint32_t maxi = (uint32_t)std::min<uint64_t>(MAXI, (uint64_t)resident + 2ull * na);
        const uint32_t nrows = cr + maxi;
        const dim3 b256(256);
        hipLaunchKernelGGL(count_kernel, dim3((2 * na + 255) / 256), b256, 0, stream, d_active, na, d_state,
                           conf.n, d_ctr);
        if (lucas) {
            // this round's strong Lucas batch: one test per confirming side
            // (just counted) and, in rounds held to one wave per SIMD, the
            // lanes it takes beside the grid, reserved from the Fermat rows
            if ((he = hipMemcpyAsync(h_ctr, d_ctr, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) !=
                    hipSuccess ||
                (he = hipStreamSynchronize(stream)) != hipSuccess)
                return e->hip_fail(he, "scan count");
            nluc = h_ctr[6];
            if (nluc) {
                if ((st = lucas_choice(bits, &lc))) return st;
                lres = resident < resident_full ? (uint32_t)lc.v->lanes : 0u;
            }
        }
        hipLaunchKernelGGL(issue_kernel, dim3((2 * na + 255) / 256), b256, 0, stream, d_active, na, d_state, cr,
                           depth_max(), resident, cpad, maxi, conf,
                           lucas ? d_lpool + (size_t)lpool_now * LP : nullptr, lucas ? d_lctr + lpool_now : nullptr,
                           lres, lucas ? d_lnum + (size_t)lpool_now * LP * W : nullptr, W,
                           d_center, CW, d_lists, CAP, d_items, d_idx, d_base, d_ctr);
        hipLaunchKernelGGL(materialize_kernel, dim3((nrows + 255) / 256), b256, 0, stream, d_items, d_idx, cr,
                           nrows, d_ctr, d_center, CW, d_lists, CAP, d_nbuf, W);
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "scan issue");
        // this round's confirmations start their strong Lucas tests now,
        // beside the grid
        if (nluc && (st = lucas_launch(lpool_now, lc, nluc))) return st;
        if (vc)
            he = vc->scan(d_nbuf, W, cr, cr, maxi, d_ctr, d_base, d_verdict, cap_c, stream);
        else
            he = vf->launch(d_nbuf + (size_t)cr * W, W, maxi, d_ctr, nullptr, none, d_verdict + cr, cap_f,
                            stream);
        if (he != hipSuccess) return e->hip_fail(he, "scan test launch");
        hipLaunchKernelGGL(retire_kernel, dim3((nrows + 255) / 256), b256, 0, stream, d_items, d_idx, cr, nrows,
                           d_verdict, d_ctr, confirm, d_state);
        hipLaunchKernelGGL(update_kernel, dim3((na + 255) / 256), b256, 0, stream, d_active, na, d_state, d_lists,
                           CAP, conf.n + (lucas ? 1u : 0u), d_ctr, d_events);
        if ((he = hipGetLastError()) != hipSuccess) return e->hip_fail(he, "scan retire");
        if ((he = hipMemcpyAsync(h_ctr, d_ctr, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
            (he = hipStreamSynchronize(stream)) != hipSuccess)
            return e->hip_fail(he, "scan round");
        uint32_t nev = h_ctr[2];
        issued += (uint64_t)h_ctr[0] + h_ctr[3];
        last_scanning = h_ctr[1];
        if (lucas) {
            // finished Lucas batches decide their m's in this round (when the
            // round tested nothing, wait for the oldest instead of spinning)
            bool any = false;
            if ((st = lucas_retire(na, conf.n + 1, h_ctr[0] + h_ctr[3] == 0, true, &any, &issued))) return st;
            if (any) {
                if ((he = hipMemcpyAsync(h_ctr, d_ctr, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) !=
                        hipSuccess ||
                    (he = hipStreamSynchronize(stream)) != hipSuccess)
                    return e->hip_fail(he, "strong Lucas retire");
                nev = h_ctr[2];
            }
        }
        if (nev) {
            if ((he = hipMemcpyAsync(h_events, d_events, nev * sizeof(Event), hipMemcpyDeviceToHost, stream)) !=
                    hipSuccess ||
                (he = hipStreamSynchronize(stream)) != hipSuccess)
                return e->hip_fail(he, "scan events");
        }
        const double t = now_s();
        std::vector<uint32_t> finished;
        std::vector<pgap_scan_row> rows;
        for (uint32_t i = 0; i < nev; ++i) {
            const Event &ev = h_events[i];
            SlotHost &sh = slots[ev.slot];
            for (int side = 0; side < 2; ++side) {
                if (!(ev.code & (2u << side))) continue;
                // exhausted: append screened candidates beyond SL (or past the last chunk)
                sh.fb[side] = true;
                bool exhausted = false;
                std::vector<uint32_t> c = fallback_candidates(sh, side, &exhausted);
                while (c.empty() && !exhausted) c = fallback_candidates(sh, side, &exhausted);
                if (c.empty()) {
                    e->set_error("scan: m=%llu has no probable prime below its centre down to 3",
                                 (unsigned long long)sh.m);
                    return PGAP_E_STATE;
                }
                Upload d{};
                d.slot = ev.slot;
                d.kind = 1;
                d.side = (uint32_t)side;
                d.n0 = (uint32_t)c.size();
                if (sh.len[side] + d.n0 > CAP) {
                    e->set_error("scan: m=%llu side %d ran past %u fallback candidates",
                                 (unsigned long long)sh.m, side, CAP);
                    return PGAP_E_STATE;
                }
                sh.len[side] += d.n0;
                pending_append.push_back(d);
                pending_append_data.push_back(std::move(c));
            }
            if (ev.code & 1u) {
                pgap_scan_row r{};
                r.m = sh.m;
                r.tag = sh.tag;
                r.family = (uint32_t)sh.fam->id;
                r.prev_p = ev.x[0];
                r.next_p = ev.x[1];
                r.p_tests = ev.found[0] == NONE ? 0 : std::min(ev.found[0] + 1, sh.orig_len[0]);
                r.n_tests = ev.found[1] == NONE ? 0 : std::min(ev.found[1] + 1, sh.orig_len[1]);
                r.flags = (sh.fb[0] ? PGAP_ROW_PREV_FALLBACK : 0) | (sh.fb[1] ? PGAP_ROW_NEXT_FALLBACK : 0) |
                          ((ev.flags & F_NEXT_SKIP) ? PGAP_ROW_NEXT_SKIPPED : 0) |
                          ((ev.flags & F_PREV_SKIP) ? PGAP_ROW_PREV_SKIPPED : 0);
                r.t_submit = sh.t_submit;
                r.t_done = t;
                rows.push_back(r);
                finished.push_back(ev.slot);
            }
        }
        if (!finished.empty()) {
            std::sort(finished.begin(), finished.end());
            std::vector<uint32_t> keep;
            keep.reserve(active.size());
            for (uint32_t s : active)
                if (!std::binary_search(finished.begin(), finished.end(), s)) keep.push_back(s);
            active.swap(keep);
            for (uint32_t s : finished) {
                slots[s].used = false;
                free_slots.push_back(s);
            }
        }
        {
            std::lock_guard<std::mutex> g(mu);
            ready.insert(ready.end(), rows.begin(), rows.end());
            admitted -= finished.size();
            tests_issued += issued;
            ++rounds;
        }
        if (!rows.empty()) cv_ready.notify_all();
        return PGAP_OK;
    }
};

}  // namespace pgap

using namespace pgap;

void pgap_engine::scan_shutdown() {
    delete scanner;  // stops and joins the worker first
    scanner = nullptr;
}

static Scanner *get_scanner(pgap_engine *e) {
    if (!e->scanner) e->scanner = new Scanner(e);
    return e->scanner;
}

extern "C" {

pgap_status pgap_scan_side_skip(pgap_engine *e, int enable, float min_merit) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    Scanner *s = get_scanner(e);
    if (s->inited) {
        e->set_error("pgap_scan_side_skip: must be called before the first submit");
        return PGAP_E_STATE;
    }
    s->side_skip = enable != 0;
    s->skip_min_merit = min_merit / 2 - 2;  // gap_test_gpu.cu:604, in float
    return PGAP_OK;
}

pgap_status pgap_scan_config(pgap_engine *e, uint32_t depth, uint32_t confirm) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    Scanner *s = get_scanner(e);
    if (s->inited) {
        e->set_error("pgap_scan_config: must be called before the first submit");
        return PGAP_E_STATE;
    }
    if (depth == 0 || depth > 4096 || confirm > PGAP_CONFIRM_BPSW) {
        e->set_error("pgap_scan_config: depth in [1,4096], confirm one of PGAP_CONFIRM_NONE/SPRP2/BPSW");
        return PGAP_E_ARG;
    }
    s->depth = depth;
    s->confirm = confirm;
    return PGAP_OK;
}

pgap_status pgap_scan_submit(pgap_engine *e, const pgap_scan_req *req) {
    if (!e || !req) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    Family *f = e->family(req->family);
    if (!f || req->m == 0 || (req->n_prev && !req->prev) || (req->n_next && !req->next) ||
        req->sides > PGAP_SIDES_NEXT) {
        e->set_error("pgap_scan_submit: bad family, m, list or sides");
        return PGAP_E_ARG;
    }
    Scanner *s = get_scanner(e);
    (void)hipSetDevice(e->opt.device);
    pgap_status st = s->init();
    if (st) return st;
    if ((size_t)f->K.size() + 2 > s->CW) {
        e->set_error("pgap_scan_submit: family registered after the scan started is wider than its slots");
        return PGAP_E_STATE;
    }
    if (f->SL > s->SL_MAX) {
        e->set_error("pgap_scan_submit: family sieve length %u exceeds the %u the scan's slots were sized for "
                     "(register every family before the first submit)",
                     f->SL, s->SL_MAX);
        return PGAP_E_STATE;
    }
    {
        // the prev side's numbers m*K - x stay >= 4 (and its beyond-SL
        // fallback finds a prime >= 3) only when m*K >= SL + 4
        Words c = f->K;
        mul_small(c, req->m);
        const uint64_t cen = small_value(c);
        if (cen && cen < (uint64_t)f->SL + 4) {
            e->set_error("pgap_scan_submit: centre m*K = %llu below sieve length + 4", (unsigned long long)cen);
            return PGAP_E_ARG;
        }
    }
    if (req->n_prev > f->SL || req->n_next > f->SL) {
        e->set_error("pgap_scan_submit: more unknowns than the sieve length allows");
        return PGAP_E_TOO_LARGE;
    }
    for (uint32_t i = 0; i < req->n_prev; ++i)
        if (req->prev[i] <= 0 || (uint32_t)req->prev[i] > f->SL) {
            e->set_error("pgap_scan_submit: prev distance %d outside [1, SL]", req->prev[i]);
            return PGAP_E_ARG;
        }
    for (uint32_t i = 0; i < req->n_next; ++i)
        if (req->next[i] <= 0 || (uint32_t)req->next[i] > f->SL) {
            e->set_error("pgap_scan_submit: next distance %d outside [1, SL]", req->next[i]);
            return PGAP_E_ARG;
        }
    Staged r;
    r.m = req->m;
    r.tag = req->tag;
    r.sides = req->sides;
    // the side a one-sided request leaves out is neither uploaded nor scanned
    if (req->sides != PGAP_SIDES_NEXT) r.prev.assign(req->prev, req->prev + req->n_prev);
    if (req->sides != PGAP_SIDES_PREV) r.next.assign(req->next, req->next + req->n_next);
    r.t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->worker_status != PGAP_OK) {
            e->set_error("%s", s->worker_err.c_str());
            return s->worker_status;
        }
        // admission control: staged requests wait for a free slot; refuse
        // when the backlog already fills every slot once more
        if (s->staged.size() >= s->S) {
            e->set_error("pgap_scan_submit: HBM budget full (%u slots busy, %zu waiting)", s->S,
                         s->staged.size());
            return PGAP_E_OOM;
        }
        auto it = s->fams.find(req->family);
        if (it == s->fams.end()) {
            FamInfo fi;
            fi.id = req->family;
            fi.K = f->K;
            fi.SL = f->SL;
            fi.K_log = f->K_log;
            it = s->fams.emplace(req->family, std::move(fi)).first;
        }
        r.fam = &it->second;
        s->staged.push_back(std::move(r));
        s->start_worker();
    }
    s->cv_work.notify_one();
    return PGAP_OK;
}

pgap_status pgap_scan_poll(pgap_engine *e, pgap_scan_row *rows, size_t cap, size_t *n_out, int timeout_ms) {
    if (!e || !n_out || (cap && !rows)) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    *n_out = 0;
    Scanner *s = get_scanner(e);
    if (!s->inited) return PGAP_OK;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool woke = wait_bounded(s->cv_ready, lk, timeout_ms, [&] {
        return !s->ready.empty() || s->worker_status != PGAP_OK || s->in_flight_locked() == 0;
    });
    if (!woke && timeout_ms > 0) {
        e->set_error("pgap_scan_poll: no row within %d ms (%zu m's in flight)", timeout_ms, s->in_flight_locked());
        return PGAP_E_TIMEOUT;
    }
    if (s->ready.empty() && s->worker_status != PGAP_OK) {
        e->set_error("%s", s->worker_err.c_str());
        return s->worker_status;
    }
    const size_t n = std::min(cap, s->ready.size());
    for (size_t i = 0; i < n; ++i) rows[i] = s->ready[i];
    s->ready.erase(s->ready.begin(), s->ready.begin() + (long)n);
    *n_out = n;
    return PGAP_OK;
}

size_t pgap_scan_in_flight(const pgap_engine *e) {
    if (!e || !e->scanner) return 0;
    return e->scanner->in_flight();
}

pgap_status pgap_scan_stats(const pgap_engine *e, uint64_t *tests_issued, uint64_t *rounds, uint32_t *slots) {
    if (!e) return PGAP_E_ARG;
    Scanner *s = e->scanner;
    uint64_t t = 0, r = 0;
    uint32_t n = 0;
    if (s) {
        std::lock_guard<std::mutex> g(s->mu);
        t = s->tests_issued;
        r = s->rounds;
        n = s->S;
    }
    if (tests_issued) *tests_issued = t;
    if (rounds) *rounds = r;
    if (slots) *slots = n;
    return PGAP_OK;
}

}  // extern "C"
