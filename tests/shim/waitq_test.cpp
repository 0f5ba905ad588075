// The bounded wait of pgap_scan_poll (prime-gap_amd/csrc/waitq.h) with a
// producer thread: a row posted in time wakes the waiter; a producer that
// stalls makes the waiter give up after its timeout; no timeout checks once.
// Prints "ok" and exits 0 when every case behaves.
#include <chrono>
#include <cstdio>
#include <thread>

#include "../../prime-gap_amd/csrc/waitq.h"

int main() {
    using clk = std::chrono::steady_clock;
    std::mutex mu;
    std::condition_variable cv;
    int rows = 0;
    auto ready = [&] { return rows > 0; };
    // 1. a row arrives after 20 ms; a 2000 ms wait returns with it
    std::thread p([&] {
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
        std::lock_guard<std::mutex> g(mu);
        rows = 1;
        cv.notify_all();
    });
    {
        std::unique_lock<std::mutex> lk(mu);
        const auto t0 = clk::now();
        const bool got = pgap::wait_bounded(cv, lk, 2000, ready);
        const double ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        if (!got || ms > 1500) { printf("case 1: got %d after %.1f ms\n", got, ms); return 1; }
    }
    p.join();
    // 2. the producer stalls: a 50 ms wait gives up, reporting no row
    rows = 0;
    {
        std::unique_lock<std::mutex> lk(mu);
        const auto t0 = clk::now();
        const bool got = pgap::wait_bounded(cv, lk, 50, ready);
        const double ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        if (got || ms < 45 || ms > 2000) { printf("case 2: got %d after %.1f ms\n", got, ms); return 1; }
    }
    // 3. no wait: one check
    {
        std::unique_lock<std::mutex> lk(mu);
        if (pgap::wait_bounded(cv, lk, 0, ready)) { printf("case 3\n"); return 1; }
        rows = 2;
        if (!pgap::wait_bounded(cv, lk, 0, ready)) { printf("case 3b\n"); return 1; }
    }
    printf("ok\n");
    return 0;
}
