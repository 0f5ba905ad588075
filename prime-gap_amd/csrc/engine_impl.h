// engine_impl.h -- the engine object behind the opaque pgap_engine handle.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>

#include "../../include/pgap.h"
#include "kernels.h"

namespace pgap {
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};
class Scanner;
}  // namespace pgap

struct pgap_engine {
    pgap_options opt{};
    bool ready = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing_pending = false;
    pgap_launch_info last{};

    mutable std::mutex err_mu;
    std::string err;

    pgap::DevBuf d_n, d_out, d_small;
    pgap::Scanner *scanner = nullptr;  // owned; released by scan_shutdown()

    void set_error(const char *fmt, ...) __attribute__((format(printf, 2, 3)));
    pgap_status hip_fail(hipError_t e, const char *what);
    pgap_status ensure(pgap::DevBuf &b, size_t bytes, const char *what);
    pgap_status pick(int kind, uint32_t bits, const pgap::Variant **out);
    pgap_status launch_batch(const pgap::Variant *v, const uint32_t *d_n, uint32_t limbs, size_t count,
                             const pgap::Bases &bases, uint8_t *d_out);
    void collect_timing();
    void scan_shutdown();
};
