// waitq.h -- the bounded wait of pgap_scan_poll, host-only C++ (no HIP), so a
// CPU test can drive it with a producer that stalls.
#pragma once
#include <chrono>
#include <condition_variable>
#include <mutex>

namespace pgap {

// Waits on cv (lk held) until ready() holds: timeout_ms == 0 checks once,
// timeout_ms < 0 waits without a limit, timeout_ms > 0 gives up after that
// many milliseconds.  Returns ready() at the end.
template <class Ready>
bool wait_bounded(std::condition_variable &cv, std::unique_lock<std::mutex> &lk, int timeout_ms, Ready ready) {
    if (timeout_ms == 0) return ready();
    if (timeout_ms < 0) {
        cv.wait(lk, ready);
        return true;
    }
    return cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
}

}  // namespace pgap
