// scan.hip -- streaming scan (filled in below) and mixed-width batches.
#include "engine_impl.h"

namespace pgap {
class Scanner {};
}  // namespace pgap

void pgap_engine::scan_shutdown() {
    delete scanner;
    scanner = nullptr;
}

extern "C" {
pgap_status pgap_prp_batch_mixed(pgap_engine *e, const uint32_t *, const uint64_t *, const uint32_t *, size_t,
                                 pgap_test, const uint32_t *, uint32_t, uint8_t *) {
    if (!e) return PGAP_E_ARG;
    e->set_error("mixed batches: not built yet");
    return PGAP_E_STATE;
}
pgap_status pgap_scan_submit(pgap_engine *e, const pgap_scan_req *) {
    if (!e) return PGAP_E_ARG;
    e->set_error("scan: not built yet");
    return PGAP_E_STATE;
}
pgap_status pgap_scan_poll(pgap_engine *e, pgap_scan_row *, size_t, size_t *n_out, int) {
    if (!e) return PGAP_E_ARG;
    if (n_out) *n_out = 0;
    e->set_error("scan: not built yet");
    return PGAP_E_STATE;
}
size_t pgap_scan_in_flight(const pgap_engine *) { return 0; }
pgap_status pgap_scan_config(pgap_engine *e, uint32_t, uint32_t) {
    if (!e) return PGAP_E_ARG;
    return PGAP_E_STATE;
}
}
