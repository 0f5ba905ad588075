// engine.hip -- the C ABI of include/pgap.h: engine lifetime, batch PRP
// tests and launch bookkeeping.  The streaming scan lives in scan.hip.
//
// Reference seam replaced: test_runner_t (miller_rabin.h:293-383).  Unlike
// the reference, errors are returned as pgap_status (never exit()), device
// buffers are reused across calls, and timing comes from HIP events on the
// engine's own stream.
#include "engine_impl.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace pgap {

// ---- variant registry ---------------------------------------------------

static std::vector<Variant> &registry() {
    static std::vector<Variant> v;
    return v;
}

void register_variant(const Variant &v) { registry().push_back(v); }
size_t variant_count() { return registry().size(); }
const Variant *variant_at(size_t i) { return i < registry().size() ? &registry()[i] : nullptr; }

const Variant *find_variant(int kind, int bits, int radix, int lanes) {
    const Variant *best = nullptr;
    for (const Variant &v : registry()) {
        if (v.kind != kind || v.max_bits < bits) continue;
        if (radix && v.radix != radix) continue;
        if (lanes && v.lanes != lanes) continue;
        if (!best || v.max_bits < best->max_bits ||
            (v.max_bits == best->max_bits && v.lanes < best->lanes))
            best = &v;
    }
    return best;
}

// ---- small device helpers ----------------------------------------------

// Max bit length over a row-major batch (atomicMax into *out).
__global__ void max_bits_kernel(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t count,
                                uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    uint32_t b = 0;
    if (i < count) {
        const uint32_t *w = n + (size_t)i * limbs;
        for (int j = (int)limbs - 1; j >= 0; --j) {
            if (w[j]) {
                b = 32u * (uint32_t)j + (32u - (uint32_t)__builtin_clz(w[j]));
                break;
            }
        }
    }
    // wave-level max first, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = (uint32_t)__shfl_down((int)b, off, 64);
        b = b > o ? b : o;
    }
    if ((threadIdx.x & 63) == 0 && b) atomicMax(out, b);
}

// ---- scratch of the allocation-free batch entry ---------------------------
//
// pgap_prp_batch_device_noalloc keeps its survivor index list and counter
// (several strong-PRP bases, Baillie-PSW) in this array instead of engine
// allocations.  Every GPU the code object loads on holds its own copy until
// the process ends; pgap_create looks up that copy's address once per device
// (hipGetSymbolAddress, with the device current) and every later use goes
// through that address.  The host-side name `scratch` is not a device address
// and is never passed to a kernel.
constexpr size_t kNoAllocWords = (size_t)PGAP_NOALLOC_MAX_COUNT + 64;  // index list, then the counter
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

static std::mutex scratch_mu;
static std::vector<uint32_t *> scratch_dev;  // per device ordinal

// Called with `device` current.
static hipError_t scratch_setup(int device) {
    std::lock_guard<std::mutex> g(scratch_mu);
    if ((size_t)device >= scratch_dev.size()) scratch_dev.resize((size_t)device + 1, nullptr);
    if (scratch_dev[(size_t)device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (he != hipSuccess) return he;
    scratch_dev[(size_t)device] = (uint32_t *)p;
    return hipSuccess;
}

static uint32_t *scratch_of(int device) {
    std::lock_guard<std::mutex> g(scratch_mu);
    return (size_t)device < scratch_dev.size() ? scratch_dev[(size_t)device] : nullptr;
}

}  // namespace pgap

using namespace pgap;

// ---- engine -------------------------------------------------------------

void pgap_engine::set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> g(err_mu);
    err = buf;
}

pgap_status pgap_engine::hip_fail(hipError_t e, const char *what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return PGAP_E_DEVICE;
}

pgap_status pgap_engine::ensure(DevBuf &b, size_t bytes, const char *what) {
    if (b.cap >= bytes) return PGAP_OK;
    if (b.p) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = std::max(bytes, (size_t)4096);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc(%zu) for %s: %s", want, what, hipGetErrorString(e));
        return PGAP_E_OOM;
    }
    b.cap = want;
    return PGAP_OK;
}

Choice pgap_engine::select(int kind, uint32_t bits) const {
    Choice c;
    std::lock_guard<std::mutex> g(tune_mu);
    const TuneEntry *t = (opt.radix || opt.lanes_per_number) ? nullptr : tuned(kind, bits);
    if (t) c.v = find_variant(kind, (int)bits, t->radix, t->lanes);
    if (!c.v) c.v = find_variant(kind, (int)bits, opt.radix, opt.lanes_per_number);
    if (!c.v && (opt.radix || opt.lanes_per_number)) c.v = find_variant(kind, (int)bits, 0, 0);
    if (!c.v) return c;
    if (opt.waves_per_simd > 0)
        c.cap = opt.waves_per_simd;
    else if (t && t->radix == c.v->radix && t->lanes == c.v->lanes && t->digits == c.v->digits)
        c.cap = t->blocks_cap;
    return c;
}

pgap_status pgap_engine::pick(int kind, uint32_t bits, Choice *out) {
    *out = select(kind, bits);
    if (!out->v) {
        set_error("no kernel variant for %u-bit numbers (kind %d)", bits, kind);
        return PGAP_E_TOO_LARGE;
    }
    return PGAP_OK;
}

// Rows whose verdict is 1: the numbers a later strong-PRP base still tests.
__global__ void select_kernel(const uint8_t *__restrict__ verdict, uint32_t count, uint32_t *__restrict__ sel,
                              uint32_t *__restrict__ nsel) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < count && verdict[t]) sel[atomicAdd(nsel, 1u)] = t;
}

// A counter reset as a kernel, not a memset node, so the sequence can sit in
// a captured graph.
__global__ void zero_kernel(uint32_t *__restrict__ p) {
    if (threadIdx.x == 0) *p = 0;
}

// Several strong-PRP bases: the first on every row, each later one only on
// the rows that passed so far (selected on the device), so lanes that
// already failed do not ride along through the later bases.  Every launch
// tests one base (the kernels' contract).
// `sel` (a row index per number of a chunk) and `nsel` (one counter) are the
// survivor scratch, used when several bases are asked for.
hipError_t pgap_engine::enqueue_batch(const Choice &c, const uint32_t *d_n, uint32_t limbs, size_t count,
                                      const Bases &bases, uint8_t *d_out, uint32_t *sel, uint32_t *nsel,
                                      uint32_t *launches) {
    const Variant *v = c.v;
    const size_t chunk = kBatchChunk;  // grid dimensions are 32-bit: split very large batches
    const uint32_t nb = v->kind == KIND_SPRP ? bases.n : 1;
    Bases first = bases;
    if (v->kind == KIND_SPRP) first.n = 1;
    hipError_t e;
    for (size_t off = 0; off < count; off += chunk) {
        const uint32_t n = (uint32_t)std::min(chunk, count - off);
        const uint32_t *dn = d_n + off * limbs;
        uint8_t *dout = d_out + off;
        if ((e = v->launch(dn, limbs, n, nullptr, nullptr, first, dout, c.cap, stream)) != hipSuccess) return e;
        ++*launches;
        for (uint32_t k = 1; k < nb; ++k) {
            Bases one{};
            one.n = 1;
            one.b[0] = bases.b[k];
            hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, stream, nsel);
            hipLaunchKernelGGL(select_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, dout, n, sel, nsel);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if ((e = v->launch(dn, limbs, n, nsel, sel, one, dout, c.cap, stream)) != hipSuccess) return e;
            ++*launches;
        }
    }
    return hipSuccess;
}

pgap_status pgap_engine::launch_batch(const Choice &c, const uint32_t *d_n, uint32_t limbs, size_t count,
                                      const Bases &bases, uint8_t *d_out) {
    const Variant *v = c.v;
    if (v->kind == KIND_SPRP && bases.n > 1) {
        pgap_status st;
        if ((st = ensure(d_sel, std::min(count, kBatchChunk) * 4, "survivor index")) ||
            (st = ensure(d_nsel, 64, "survivor count")))
            return st;
    }
    BatchGraph key;
    key.v = v;
    key.d_n = d_n;
    key.limbs = limbs;
    key.count = count;
    key.bases = bases;
    key.d_out = d_out;
    key.cap = c.cap;
    key.sel = d_sel.p;
    key.nsel = d_nsel.p;
    const bool repeat = graph.same(key);
    if (!repeat && graph.exec) {
        (void)hipGraphExecDestroy(graph.exec);
        graph.exec = nullptr;
    }
    hipError_t e = hipEventRecord(ev0, stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    uint32_t launches = 0;
    bool replayed = false;
    if (repeat) {
        // the same launch as last time: capture it once, then replay
        if (!graph.exec) {
            hipGraph_t g = nullptr;
            if ((e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return hip_fail(e, "graph capture");
            const hipError_t le = enqueue_batch(c, d_n, limbs, count, bases, d_out, (uint32_t *)d_sel.p,
                                                (uint32_t *)d_nsel.p, &launches);
            e = hipStreamEndCapture(stream, &g);
            if (le != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hip_fail(le, "kernel launch (capture)");
            }
            if (e != hipSuccess) return hip_fail(e, "graph capture end");
            e = hipGraphInstantiate(&graph.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                graph.exec = nullptr;
                return hip_fail(e, "graph instantiate");
            }
        } else {
            launches = graph_launches;
        }
        if ((e = hipGraphLaunch(graph.exec, stream)) != hipSuccess) return hip_fail(e, "graph launch");
        replayed = true;
        graph_launches = launches;
    } else {
        if ((e = enqueue_batch(c, d_n, limbs, count, bases, d_out, (uint32_t *)d_sel.p, (uint32_t *)d_nsel.p,
                               &launches)) != hipSuccess)
            return hip_fail(e, "kernel launch");
        graph = key;
    }
    e = hipEventRecord(ev1, stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    last.digits = (uint32_t)v->digits;
    last.radix_bits = (uint32_t)v->radix;
    last.lanes_per_number = (uint32_t)v->lanes;
    last.launches = launches;
    last.tests = count;
    last.graph = replayed ? 1u : 0u;
    last.kernel_ms = -1.0;
    timing_pending = true;
    return PGAP_OK;
}

// Rows whose number is 2: Baillie-PSW passes it, the strong base-2 test
// (odd N only) does not.
__global__ void two_kernel(const uint32_t *__restrict__ n, uint32_t limbs, uint32_t count,
                           uint8_t *__restrict__ out) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= count) return;
    const uint32_t *w = n + (size_t)t * limbs;
    uint32_t hi = 0;
    for (uint32_t j = 1; j < limbs; ++j) hi |= w[j];
    if (w[0] == 2u && hi == 0) out[t] = 1;
}

// Baillie-PSW: the strong base-2 test on every row, the strong Lucas test on
// the rows that passed (selected on the device into `sel` / `nsel`), then 2
// itself.
hipError_t pgap_engine::enqueue_bpsw(const Choice &cs, const Choice &cl, const uint32_t *d_n, uint32_t limbs,
                                     size_t count, uint8_t *d_out, uint32_t *sel, uint32_t *nsel,
                                     uint32_t *launches) {
    Bases two{}, none{};
    two.n = 1;
    two.b[0] = 2;
    hipError_t e;
    for (size_t off = 0; off < count; off += kBatchChunk) {
        const uint32_t n = (uint32_t)std::min(kBatchChunk, count - off);
        const uint32_t *dn = d_n + off * limbs;
        uint8_t *dout = d_out + off;
        if ((e = cs.v->launch(dn, limbs, n, nullptr, nullptr, two, dout, cs.cap, stream)) != hipSuccess) return e;
        hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, stream, nsel);
        hipLaunchKernelGGL(select_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, dout, n, sel, nsel);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = cl.v->launch(dn, limbs, n, nsel, sel, none, dout, cl.cap, stream)) != hipSuccess) return e;
        hipLaunchKernelGGL(two_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, dn, limbs, n, dout);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        *launches += 2;
    }
    return hipSuccess;
}

pgap_status pgap_engine::launch_bpsw(const Choice &cs, const Choice &cl, const uint32_t *d_n, uint32_t limbs,
                                     size_t count, uint8_t *d_out) {
    pgap_status st;
    if ((st = ensure(d_sel, std::min(count, kBatchChunk) * 4, "survivor index")) ||
        (st = ensure(d_nsel, 64, "survivor count")))
        return st;
    hipError_t e = hipEventRecord(ev0, stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    uint32_t launches = 0;
    if ((e = enqueue_bpsw(cs, cl, d_n, limbs, count, d_out, (uint32_t *)d_sel.p, (uint32_t *)d_nsel.p,
                          &launches)) != hipSuccess)
        return hip_fail(e, "Baillie-PSW launch");
    e = hipEventRecord(ev1, stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    last.digits = (uint32_t)cl.v->digits;
    last.radix_bits = (uint32_t)cl.v->radix;
    last.lanes_per_number = (uint32_t)cl.v->lanes;
    last.launches = launches;
    last.tests = count;
    last.graph = 0;
    last.kernel_ms = -1.0;
    timing_pending = true;
    return PGAP_OK;
}

// Variant(s) for a batch of `kind` over numbers of up to `bits` bits, and the
// launch: BPSW runs a strong base-2 variant and a strong Lucas variant.
pgap_status pgap_engine::run_batch(pgap_test kind, const Bases &bs, uint32_t bits, const uint32_t *d_n,
                                   uint32_t limbs, size_t count, uint8_t *d_out) {
    bits = std::max(bits, 2u);
    pgap_status st;
    if (kind == PGAP_BPSW) {
        Choice cs, cl;
        if ((st = pick(KIND_SPRP, bits, &cs)) || (st = pick(KIND_LUCAS, bits, &cl))) return st;
        return launch_bpsw(cs, cl, d_n, limbs, count, d_out);
    }
    const int k = kind == PGAP_LUCAS ? KIND_LUCAS : (int)kind;
    Choice v;
    if ((st = pick(k, select_bits(k, bits, bs), &v))) return st;
    return launch_batch(v, d_n, limbs, count, bs, d_out);
}

void pgap_engine::collect_timing() {
    if (!timing_pending) return;
    float ms = 0.f;
    if (hipEventSynchronize(ev1) == hipSuccess && hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess)
        last.kernel_ms = ms;
    timing_pending = false;
}

static pgap_status check_bases(pgap_engine *e, pgap_test kind, const uint32_t *bases, uint32_t nbases,
                               Bases *out) {
    memset(out, 0, sizeof *out);
    if (kind == PGAP_FERMAT2 || kind == PGAP_LUCAS || kind == PGAP_BPSW) return PGAP_OK;
    if (kind != PGAP_SPRP) {
        e->set_error("unknown test kind %d", (int)kind);
        return PGAP_E_ARG;
    }
    if (nbases == 0 || nbases > (uint32_t)kMaxBases || !bases) {
        e->set_error("SPRP needs 1..%d bases (got %u)", kMaxBases, nbases);
        return PGAP_E_ARG;
    }
    out->n = nbases;
    for (uint32_t i = 0; i < nbases; ++i) {
        // bases are small primes in practice (generate_primes, miller_rabin.h:
        // 269-286); the kernels multiply by the base as a scalar, which the
        // lazy-reduction bound allows for bases below 64 (select_bits).
        if (bases[i] < 2 || bases[i] > 63) {
            e->set_error("SPRP base %u outside [2, 63]", bases[i]);
            return PGAP_E_ARG;
        }
        out->b[i] = bases[i];
    }
    return PGAP_OK;
}

static uint32_t host_max_bits(const uint32_t *n, uint32_t limbs, size_t count) {
    uint32_t best = 0;
    for (size_t i = 0; i < count; ++i) {
        const uint32_t *w = n + i * limbs;
        for (int j = (int)limbs - 1; j >= 0; --j) {
            if (w[j]) {
                uint32_t b = 32u * (uint32_t)j + (32u - (uint32_t)__builtin_clz(w[j]));
                best = std::max(best, b);
                break;
            }
        }
    }
    return best;
}

extern "C" {

const char *pgap_version(void) { return "pgap 0.1 gfx950"; }

pgap_status pgap_create(const pgap_options *opt, pgap_engine **out) {
    if (!out) return PGAP_E_ARG;
    *out = nullptr;
    pgap_engine *e = new pgap_engine();
    if (opt) e->opt = *opt;
    if (variant_count() == 0) {
        e->set_error("no HIP kernel variants registered (code objects missing)");
        *out = e;
        return PGAP_E_DEVICE;
    }
    int ndev = 0;
    hipError_t he = hipGetDeviceCount(&ndev);
    if (he != hipSuccess || ndev == 0) {
        e->set_error("no HIP device: %s", hipGetErrorString(he));
        *out = e;
        return PGAP_E_DEVICE;
    }
    if (e->opt.device < 0 || e->opt.device >= ndev) {
        e->set_error("device %d out of range (have %d)", e->opt.device, ndev);
        *out = e;
        return PGAP_E_ARG;
    }
    if ((he = hipSetDevice(e->opt.device)) != hipSuccess ||
        (he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess ||
        (he = hipEventCreate(&e->ev0)) != hipSuccess || (he = hipEventCreate(&e->ev1)) != hipSuccess) {
        e->hip_fail(he, "engine set-up");
        *out = e;
        return PGAP_E_DEVICE;
    }
    if ((he = scratch_setup(e->opt.device)) != hipSuccess) {
        e->hip_fail(he, "scratch symbol address");
        *out = e;
        return PGAP_E_DEVICE;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->opt.device) == hipSuccess && cus > 0)
        e->num_cus = cus;
    if (e->opt.tuning_path && !e->load_tuning(e->opt.tuning_path)) {
        e->set_error("cannot read tuning table %s", e->opt.tuning_path);
        e->opt.tuning_path = nullptr;
        *out = e;
        return PGAP_E_ARG;
    }
    e->opt.tuning_path = nullptr;  // the caller's string need not outlive create
    e->ready = true;
    *out = e;
    return PGAP_OK;
}

void pgap_destroy(pgap_engine *e) {
    if (!e) return;
    if (e->ready) {
        (void)hipSetDevice(e->opt.device);
        (void)hipStreamSynchronize(e->stream);
        e->scan_shutdown();
        e->sieve_shutdown();
        e->families_shutdown();
        if (e->graph.exec) (void)hipGraphExecDestroy(e->graph.exec);
        e->graph.exec = nullptr;
        for (DevBuf *b : {&e->d_n, &e->d_out, &e->d_small, &e->d_m, &e->d_x, &e->d_keep, &e->d_idx0, &e->d_idx1,
                          &e->d_sel, &e->d_nsel}) {
            if (b->p) (void)hipFree(b->p);
        }
        (void)hipEventDestroy(e->ev0);
        (void)hipEventDestroy(e->ev1);
        (void)hipStreamDestroy(e->stream);
    }
    delete e;
}

// The message is copied per calling thread: the scan worker may set a new
// one while the caller reads this one.  Valid until this thread's next call.
const char *pgap_last_error(const pgap_engine *e) {
    if (!e) return "null engine";
    static thread_local std::string copy;
    {
        std::lock_guard<std::mutex> g(e->err_mu);
        copy = e->err;
    }
    return copy.c_str();
}

pgap_status pgap_prp_batch(pgap_engine *e, uint32_t limbs, const uint32_t *n, size_t count, pgap_test kind,
                           const uint32_t *bases, uint32_t nbases, uint8_t *verdict) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    if (count == 0) return PGAP_OK;
    if (!n || !verdict || limbs == 0) {
        e->set_error("pgap_prp_batch: null buffer or zero limbs");
        return PGAP_E_ARG;
    }
    Bases bs;
    pgap_status st = check_bases(e, kind, bases, nbases, &bs);
    if (st) return st;
    (void)hipSetDevice(e->opt.device);
    const uint32_t bits = host_max_bits(n, limbs, count);
    const size_t in_bytes = count * limbs * sizeof(uint32_t);
    if ((st = e->ensure(e->d_n, in_bytes, "numbers")) || (st = e->ensure(e->d_out, count, "verdicts")))
        return st;
    hipError_t he = hipMemcpyAsync(e->d_n.p, n, in_bytes, hipMemcpyHostToDevice, e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "hipMemcpyAsync H2D");
    if ((st = e->run_batch(kind, bs, bits, (const uint32_t *)e->d_n.p, limbs, count, (uint8_t *)e->d_out.p)))
        return st;
    he = hipMemcpyAsync(verdict, e->d_out.p, count, hipMemcpyDeviceToHost, e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "hipMemcpyAsync D2H");
    he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "hipStreamSynchronize");
    e->collect_timing();
    return PGAP_OK;
}

pgap_status pgap_prp_batch_device(pgap_engine *e, uint32_t limbs, const uint32_t *d_n, size_t count,
                                  pgap_test kind, const uint32_t *bases, uint32_t nbases, uint8_t *d_verdict,
                                  uint32_t max_bits) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    if (count == 0) return PGAP_OK;
    if (!d_n || !d_verdict || limbs == 0) {
        e->set_error("pgap_prp_batch_device: null buffer or zero limbs");
        return PGAP_E_ARG;
    }
    Bases bs;
    pgap_status st = check_bases(e, kind, bases, nbases, &bs);
    if (st) return st;
    (void)hipSetDevice(e->opt.device);
    if (max_bits == 0) {
        if ((st = e->ensure(e->d_small, 64, "scratch"))) return st;
        hipError_t he = hipMemsetAsync(e->d_small.p, 0, sizeof(uint32_t), e->stream);
        if (he != hipSuccess) return e->hip_fail(he, "hipMemsetAsync");
        for (size_t off = 0; off < count; off += ((size_t)1 << 30)) {
            const uint32_t c = (uint32_t)std::min((size_t)1 << 30, count - off);
            hipLaunchKernelGGL(max_bits_kernel, dim3((c + 255) / 256), dim3(256), 0, e->stream, d_n + off * limbs,
                               limbs, c, (uint32_t *)e->d_small.p);
        }
        he = hipMemcpyAsync(&max_bits, e->d_small.p, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
        if (he != hipSuccess) return e->hip_fail(he, "max bits");
    }
    return e->run_batch(kind, bs, max_bits, d_n, limbs, count, d_verdict);
}

// The device batch without engine allocations: the same variants and
// launches as pgap_prp_batch_device, with the survivor scratch in the
// library's per-GPU array (no graph capture, no timing events).
pgap_status pgap_prp_batch_device_noalloc(pgap_engine *e, uint32_t limbs, const uint32_t *d_n, size_t count,
                                          pgap_test kind, const uint32_t *bases, uint32_t nbases,
                                          uint8_t *d_verdict, uint32_t max_bits) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    if (count == 0) return PGAP_OK;
    if (!d_n || !d_verdict || limbs == 0 || max_bits == 0) {
        e->set_error("pgap_prp_batch_device_noalloc: null buffer, zero limbs or max_bits 0 (the bound is required)");
        return PGAP_E_ARG;
    }
    if (count > (size_t)PGAP_NOALLOC_MAX_COUNT) {
        e->set_error("pgap_prp_batch_device_noalloc: %zu numbers exceed the %u its scratch holds (split the batch)",
                     count, (unsigned)PGAP_NOALLOC_MAX_COUNT);
        return PGAP_E_TOO_LARGE;
    }
    Bases bs;
    pgap_status st = check_bases(e, kind, bases, nbases, &bs);
    if (st) return st;
    (void)hipSetDevice(e->opt.device);
    uint32_t *scr = scratch_of(e->opt.device);
    if (!scr) {
        e->set_error("pgap_prp_batch_device_noalloc: no scratch address for device %d", e->opt.device);
        return PGAP_E_STATE;
    }
    uint32_t *sel = scr, *nsel = scr + PGAP_NOALLOC_MAX_COUNT;
    const uint32_t bits = std::max(max_bits, 2u);
    uint32_t launches = 0;
    hipError_t he;
    if (kind == PGAP_BPSW) {
        Choice cs, cl;
        if ((st = e->pick(KIND_SPRP, bits, &cs)) || (st = e->pick(KIND_LUCAS, bits, &cl))) return st;
        he = e->enqueue_bpsw(cs, cl, d_n, limbs, count, d_verdict, sel, nsel, &launches);
    } else {
        const int k = kind == PGAP_LUCAS ? KIND_LUCAS : (int)kind;
        Choice v;
        if ((st = e->pick(k, select_bits(k, bits, bs), &v))) return st;
        he = e->enqueue_batch(v, d_n, limbs, count, bs, d_verdict, sel, nsel, &launches);
    }
    if (he != hipSuccess) return e->hip_fail(he, "pgap_prp_batch_device_noalloc launch");
    return PGAP_OK;
}

pgap_status pgap_sync(pgap_engine *e) {
    if (!e) return PGAP_E_ARG;
    if (!e->ready) return PGAP_E_STATE;
    hipError_t he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "hipStreamSynchronize");
    e->collect_timing();
    return PGAP_OK;
}

pgap_status pgap_last_launch(const pgap_engine *e, pgap_launch_info *info) {
    if (!e || !info) return PGAP_E_ARG;
    *info = e->last;
    return PGAP_OK;
}

}  // extern "C"
