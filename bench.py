#!/usr/bin/env python3
"""Benchmark of the MI355X PRP engine on its headline workload.

Workload (the fixed-width throughput batch): N = m*809#/1841070 + x with m
uniform in [1e7, 1.2e7) coprime to 1841070, x uniform in [-15000, 15000],
every N free of prime factors <= 1e6 (screened on the device), 2^20
candidates (N ~ 1108 bits), seeded.  One step = one Fermat base-2 test of
every candidate, inputs resident in HBM.  Reported: tests/s over exactly
--steps steps, the dominant kernel's roofline against the int32
multiply-add rate measured in the same run (pgap_mad_peak) and against the
nominal one, and the GMP oracle timed on the host on a bounded sample of the
same candidates (cpu_baseline).  `extra` holds the other widths, the
mixed-width batch and the serving scans (503#/210 plain file, 1511#/2190,
3001#/2190).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --mad-peak      (only the multiply-add rate)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "prime-gap_amd"))

import numpy as np  # noqa: E402

METRIC = "Fermat PRP tests/sec at 1170-bit & 2180-bit; streamed per-m p50/p99 latency"
P, D = 809, 1841070
M_LO, M_HI = 10_000_000, 12_000_000
X_RANGE = 15000
SIEVE_LIMIT = 1_000_000
# One v_mad_u64_u32 per lane per clock: 256 CUs x 64 lanes x 2.4 GHz.
PEAK_INT32_MAC = 256 * 64 * 2.4e9
MAD_PEAK_ITERS = 16384


def measured_peak(engine, reps=3):
    """Sustained v_mad_u64_u32 rate of this device (best of `reps` launches)."""
    runs = [engine.mad_peak(MAD_PEAK_ITERS) for _ in range(reps)]
    best = max(runs, key=lambda r: r["mac_per_s"])
    return {"mac_per_s": best["mac_per_s"], "kernel_ms": best["kernel_ms"], "macs_per_launch": best["macs"],
            "runs_mac_per_s": [r["mac_per_s"] for r in runs]}


def host_cpus():
    """Threads the CPU baseline uses: the CPUs this process may run on, capped
    by OMP_NUM_THREADS (the host's share per GPU); plus what the host is."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    threads = min(allowed, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else allowed
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"threads_used": threads, "host_nproc": os.cpu_count(), "affinity_cpus": allowed,
                     "omp_num_threads": cap, "cpu_model": model}


def draw_candidates(engine, fam, count, seed):
    """Seeded (m, x) pairs whose N = m*K + x survive trial division to 1e6."""
    rng = np.random.default_rng(seed)
    ms, xs = [], []
    have = 0
    while have < count:
        m = rng.integers(M_LO, M_HI, size=1 << 22, dtype=np.int64)
        m = m[np.gcd(m, D) == 1]
        x = rng.integers(-X_RANGE, X_RANGE + 1, size=len(m), dtype=np.int64)
        keep = engine.filter_candidates(fam, m.astype(np.uint64), x, SIEVE_LIMIT).astype(bool)
        ms.append(m[keep])
        xs.append(x[keep])
        have += int(keep.sum())
    return np.concatenate(ms)[:count].astype(np.uint64), np.concatenate(xs)[:count]


def load_traffic(kernel_tag):
    """HBM bytes per dispatch measured with rocprofv3 --pmc (profiles/)."""
    path = os.path.join(REPO, "profiles", f"{kernel_tag}.traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_dispatch")
    except OSError:
        return None


def cpu_baseline(numbers_host, sample, threads):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib  # the GMP oracle: timed as the CPU baseline, never as the product
    arr = np.ascontiguousarray(numbers_host[:sample], dtype=np.uint32)
    t0 = time.perf_counter()
    verdicts = oracle_lib.verdicts(arr, 0, threads=threads)
    dt = time.perf_counter() - t0
    return arr.shape[0] / dt, dt, verdicts


def primorial_over(P, D):
    return math.prod(p for p in range(2, P + 1) if all(p % q for q in range(2, int(p ** 0.5) + 1))) // D


def bench_batch_2180(engine, hipmem, pgap, count, steps, seed, peak_mac):
    """The 2180-bit width of the metric: N = m*1511#/2190 + x, m ~ U[1, 1e6)
    coprime to 2190, x ~ U[-20000, 20000], no prime factor <= 1e6."""
    P2, D2 = 1511, 2190
    fam = engine.family(P2, D2, 0)
    rng = np.random.default_rng(seed)
    ms, xs, have = [], [], 0
    while have < count:
        m = rng.integers(1, 1_000_000, size=1 << 20, dtype=np.int64)
        m = m[np.gcd(m, D2) == 1]
        x = rng.integers(-20000, 20001, size=len(m), dtype=np.int64)
        keep = engine.filter_candidates(fam, m.astype(np.uint64), x, SIEVE_LIMIT).astype(bool)
        ms.append(m[keep])
        xs.append(x[keep])
        have += int(keep.sum())
    m = np.concatenate(ms)[:count].astype(np.uint64)
    x = np.concatenate(xs)[:count]
    K = primorial_over(P2, D2)
    bits_hi = (K * 1_000_000 + 20000).bit_length()
    limbs = (bits_hi + 31) // 32
    d_n = hipmem.DeviceArray((count, limbs), np.uint32)
    d_out = hipmem.DeviceArray(count, np.uint8)
    engine.build_candidates_device(fam, m, x, limbs, d_n.ptr)
    e_mean = float(np.mean([(int(a) * K + int(b)).bit_length() for a, b in zip(m[:2048], x[:2048])]))
    ks = []
    for i in range(steps + 1):
        engine.prp_device(d_n.ptr, limbs, count, d_out.ptr, pgap.FERMAT2, (), bits_hi)
        engine.sync()
        if i:
            ks.append(engine.last_launch().kernel_ms)
    info = engine.last_launch()
    avg = float(np.mean(ks)) * 1e-3
    L32 = (int(e_mean + 0.999) + 31) // 32
    return {
        "workload": f"N = m*{P2}#/{D2} + x, m~U[1,1e6) coprime to d, x~U[-20000,20000], no factor <= 1e6, "
                    f"{count} candidates, {e_mean:.0f} bits",
        "tests_per_s": count / avg,
        "kernel_ms_avg": avg * 1e3,
        "kernel": f"radix-2^{info.radix_bits} x {info.digits} digits",
        "roofline_frac": 2 * L32 * L32 * e_mean * count / avg / peak_mac,
        "roofline_frac_of_nominal": 2 * L32 * L32 * e_mean * count / avg / PEAK_INT32_MAC,
        "roofline_frac_container_72_limbs": 2 * 72 * 72 * e_mean * count / avg / peak_mac,
        "work_per_test": f"2*L^2*e, L={L32}, e={e_mean:.1f}",
    }


def sieved_family_rows(engine, hipmem, P2, D2, mlo, mhi, xr, count, rng):
    """count rows N = m*K + x (no factor <= 1e6) as a host [count][limbs] array."""
    fam = engine.family(P2, D2, 0)
    ms, xs, have = [], [], 0
    while have < count:
        m = rng.integers(mlo, mhi, size=1 << 20, dtype=np.int64)
        m = m[np.gcd(m, D2) == 1]
        x = rng.integers(-xr, xr + 1, size=len(m), dtype=np.int64)
        keep = engine.filter_candidates(fam, m.astype(np.uint64), x, SIEVE_LIMIT).astype(bool)
        ms.append(m[keep])
        xs.append(x[keep])
        have += int(keep.sum())
    m = np.concatenate(ms)[:count].astype(np.uint64)
    x = np.concatenate(xs)[:count]
    K = primorial_over(P2, D2)
    bits_hi = (K * (mhi - 1) + xr).bit_length()
    limbs = (bits_hi + 31) // 32
    d = hipmem.DeviceArray((count, limbs), np.uint32)
    engine.build_candidates_device(fam, m, x, limbs, d.ptr)
    rows = d.to_host()
    d.free()
    e_mean = float(np.mean([(int(a) * K + int(b)).bit_length() for a, b in zip(m[:1024], x[:1024])]))
    return rows, bits_hi, e_mean


MIX_UNEVEN = ((503, 210, 0.4, 1_000_000, 20000), (1511, 2190, 0.4, 1_000_000, 20000),
              (809, 1841070, 0.2, 1_000_000, 20000))
# the uneven mix widened to the large-operand family: 3001#/2190 with m < 1e5
# and x ~ U[-40000, 40000] (4231-4247 bits, eight lanes per number)
MIX_WIDE = ((503, 210, 0.4, 1_000_000, 20000), (1511, 2190, 0.3, 1_000_000, 20000),
            (3001, 2190, 0.3, 100_000, 40000))


def bench_mixed(engine, hipmem, pgap, count, seed, fams=MIX_UNEVEN):
    """Uneven batching: by default 40% 503#/210, 40% 1511#/2190, 20%
    809#/1841070, m ~ U[1, 1e6) coprime to d, x ~ U[-20000, 20000], no factor
    <= 1e6, in a seeded shuffle.  Compared: one persistent launch over all
    widths, one launch per width, and every number in the widest width (one
    container)."""
    rng = np.random.default_rng(seed)
    parts = []
    for (P2, D2, frac, mhi, xr) in fams:
        n = int(round(count * frac))
        rows, bits_hi, e_mean = sieved_family_rows(engine, hipmem, P2, D2, 1, mhi, xr, n, rng)
        parts.append((P2, rows, bits_hi, e_mean))
    total = sum(len(r) for _, r, _, _ in parts)
    # packed, shuffled
    fam_of = np.concatenate([np.full(len(r), i) for i, (_, r, _, _) in enumerate(parts)])
    idx_in = np.concatenate([np.arange(len(r)) for _, r, _, _ in parts])
    perm = rng.permutation(total)
    fam_of, idx_in = fam_of[perm], idx_in[perm]
    limbs = np.array([parts[f][1].shape[1] for f in fam_of], dtype=np.uint32)
    offset = np.zeros(total, dtype=np.uint64)
    offset[1:] = np.cumsum(limbs[:-1], dtype=np.uint64)
    words = np.zeros(int(limbs.sum()), dtype=np.uint32)
    for f, (_, rows, _, _) in enumerate(parts):
        sel = np.nonzero(fam_of == f)[0]
        w = rows.shape[1]
        pos = (offset[sel][:, None] + np.arange(w, dtype=np.uint64)[None, :]).ravel()
        words[pos] = rows[idx_in[sel]].ravel()
    engine.prp_mixed_packed(words, offset, limbs)          # warm-up
    v_mixed = engine.prp_mixed_packed(words, offset, limbs)
    t_mixed = engine.last_launch().kernel_ms
    # one launch per width (device-resident rows)
    t_per = 0.0
    v_per = np.zeros(total, dtype=np.uint8)
    dev = []
    for f, (_, rows, bits_hi, _) in enumerate(parts):
        d_n = hipmem.DeviceArray.from_host(rows)
        d_o = hipmem.DeviceArray(len(rows), np.uint8)
        engine.prp_device(d_n.ptr, rows.shape[1], len(rows), d_o.ptr, pgap.FERMAT2, (), bits_hi)
        engine.sync()
        engine.prp_device(d_n.ptr, rows.shape[1], len(rows), d_o.ptr, pgap.FERMAT2, (), bits_hi)
        engine.sync()
        t_per += engine.last_launch().kernel_ms
        vv = d_o.to_host()
        sel = np.nonzero(fam_of == f)[0]
        v_per[sel] = vv[idx_in[sel]]
        dev.append((d_n, d_o))
    # one container: everything at the widest width
    wmax = max(r.shape[1] for _, r, _, _ in parts)
    allrows = np.zeros((total, wmax), dtype=np.uint32)
    for f, (_, rows, _, _) in enumerate(parts):
        sel = np.nonzero(fam_of == f)[0]
        allrows[sel, : rows.shape[1]] = rows[idx_in[sel]]
    d_all = hipmem.DeviceArray.from_host(allrows)
    d_out = hipmem.DeviceArray(total, np.uint8)
    bits_all = max(b for _, _, b, _ in parts)
    engine.prp_device(d_all.ptr, wmax, total, d_out.ptr, pgap.FERMAT2, (), bits_all)
    engine.sync()
    engine.prp_device(d_all.ptr, wmax, total, d_out.ptr, pgap.FERMAT2, (), bits_all)
    engine.sync()
    t_one = engine.last_launch().kernel_ms
    v_one = d_out.to_host()
    desc = ", ".join(f"{100 * fr:.0f}% {P2}#/{D2} ({parts[i][3]:.0f} bits, m~U[1,{mhi:.0e}), x~U[-{xr},{xr}])"
                     for i, (P2, D2, fr, mhi, xr) in enumerate(fams))
    return {
        "workload": f"{total} candidates: {desc}, no factor <= 1e6, seeded shuffle",
        "persistent_tests_per_s": total / (t_mixed * 1e-3),
        "per_width_tests_per_s": total / (t_per * 1e-3),
        "one_container_tests_per_s": total / (t_one * 1e-3),
        "kernel_ms": {"persistent": t_mixed, "per_width_sum": t_per, "one_container": t_one},
        "verdicts_identical": bool(np.array_equal(v_mixed, v_per) and np.array_equal(v_mixed, v_one)),
    }


def bench_batch_family(engine, hipmem, pgap, P2, D2, m_hi, xr, count, seed, peak_mac, container_limbs):
    """A fixed-width Fermat batch of one family: N = m*P2#/D2 + x, m ~ U[1,
    m_hi), x ~ U[-xr, xr], no prime factor <= 1e6; one warm-up launch and
    two timed ones."""
    rng = np.random.default_rng(seed)
    rows, bits_hi, e_mean = sieved_family_rows(engine, hipmem, P2, D2, 1, m_hi, xr, count, rng)
    d_n = hipmem.DeviceArray.from_host(rows)
    d_o = hipmem.DeviceArray(count, np.uint8)
    ks = []
    for i in range(3):
        engine.prp_device(d_n.ptr, rows.shape[1], count, d_o.ptr, pgap.FERMAT2, (), bits_hi)
        engine.sync()
        if i:
            ks.append(engine.last_launch().kernel_ms)
    info = engine.last_launch()
    avg = float(np.mean(ks)) * 1e-3
    L32 = (int(e_mean + 0.999) + 31) // 32
    return {
        "workload": f"N = m*{P2}#/{D2} + x, m~U[1,{m_hi:.0e}), x~U[-{xr},{xr}], no factor <= 1e6, {count} "
                    f"candidates, {e_mean:.0f} bits",
        "tests_per_s": count / avg,
        "kernel_ms_avg": avg * 1e3,
        "kernel": f"radix-2^{info.radix_bits} x {info.digits} digits, {info.lanes_per_number} lanes/number",
        "roofline_frac": 2 * L32 * L32 * e_mean * count / avg / peak_mac,
        "roofline_frac_of_nominal": 2 * L32 * L32 * e_mean * count / avg / PEAK_INT32_MAC,
        f"roofline_frac_container_{container_limbs}_limbs":
            2 * container_limbs * container_limbs * e_mean * count / avg / peak_mac,
        "work_per_test": f"2*L^2*e, L={L32}, e={e_mean:.1f}",
    }


def bench_batch_4300(engine, hipmem, pgap, count, seed, peak_mac):
    """Large operands: N = m*3001#/2190 + x, m ~ U[1, 1e5), x ~ U[-40000,
    40000], no factor <= 1e6 (4231-4248 bits, eight lanes per number)."""
    return bench_batch_family(engine, hipmem, pgap, 3001, 2190, 100_000, 40000, count, seed, peak_mac, 136)


def bench_batch_700(engine, hipmem, pgap, count, seed, peak_mac):
    """Narrow operands, the 503#/210 family: N = m*503#/210 + x, m ~ U[1,
    1e6), x ~ U[-20000, 20000], no factor <= 1e6 (about 705 bits; a
    container of 24 32-bit limbs)."""
    return bench_batch_family(engine, hipmem, pgap, 503, 210, 1_000_000, 20000, count, seed, peak_mac, 24)


def bench_stream(engine, pgap, n_m, load, seed, cpu_sample, threads, P4=1511, D4=2190, SL4=40000, m_hi=200000,
                 sieve_limit=1_000_000_000, all_m=False, cpu_info=None):
    """Serving: P=1511, d=2190, SL=40000, m's in seeded random order from
    [1, 200000) coprime to d, each with its unknowns sieved to 1e9 (device
    sieve, untimed).  Phase 1 saturates the engine (capacity); phase 2 feeds
    Poisson arrivals at `load` x capacity and records arrival-to-row latency.
    The large-operand serving case passes P=3001, m < 1e5.  all_m: every m
    coprime to d below m_hi, in both phases (the plain 503#/210 file, whose
    m's are few), with every row checked against the CPU baseline's."""
    fam = engine.family(P4, D4, SL4)
    rng = np.random.default_rng(seed)
    pool = np.array([m for m in range(1, m_hi) if math.gcd(m, D4) == 1], dtype=np.uint64)
    if all_m:
        n_m = len(pool)
        order = rng.permutation(pool)
    else:
        order = rng.permutation(pool)[: 2 * n_m]
    t0 = time.perf_counter()
    lists = engine.sieve(fam, order, sieve_limit, SL4)
    sieve_s = time.perf_counter() - t0
    if all_m:
        cap_set = lat_set = list(range(n_m))
    else:
        cap_set, lat_set = list(range(n_m)), list(range(n_m, 2 * n_m))

    def run(idx, arrivals):
        rows = {}
        pending = list(idx)
        pos = 0
        start = time.monotonic()
        arr_abs = {}
        while pos < len(pending) or engine.scan_in_flight():
            now = time.monotonic()
            while pos < len(pending):
                i = pending[pos]
                if arrivals is not None and start + arrivals[pos] > now:
                    break
                p, q = lists[i]
                if engine.scan_submit(fam, int(order[i]), p, q, tag=i) == pgap.PGAP_E_OOM:
                    break
                arr_abs[i] = start + arrivals[pos] if arrivals is not None else now
                pos += 1
            if engine.scan_in_flight():
                for r in engine.scan_poll(wait=False):
                    rows[r.tag] = (r, r.t_done - arr_abs[r.tag])
            elif pos < len(pending):
                time.sleep(max(0.0, start + arrivals[pos] - time.monotonic()))
        return rows, time.monotonic() - start

    s0 = engine.scan_stats()
    rows_cap, cap_s = run(cap_set, None)
    s1 = engine.scan_stats()
    useful = sum(r.p_tests + r.n_tests for r, _ in rows_cap.values())
    cap_mps = len(rows_cap) / cap_s
    gaps = np.cumsum(rng.exponential(1.0 / (load * cap_mps), size=len(lat_set)))
    rows_lat, lat_s = run(lat_set, gaps)
    lat_ms = np.array([lat for _, lat in rows_lat.values()]) * 1e3
    # one m at a time (no other work in flight): the lightly loaded latency
    single = []
    for i in cap_set[:64]:
        p, q = lists[i]
        t_sub = time.monotonic()
        assert engine.scan_submit(fam, int(order[i]), p, q, tag=i) == 0
        got = []
        while not got:
            got = engine.scan_poll(wait=True)
        single.append((got[0].t_done - t_sub) * 1e3)
        r0 = rows_cap[i][0]
        assert (got[0].prev_p, got[0].next_p) == (r0.prev_p, r0.next_p)
    out = {
        "workload": f"P={P4} d={D4} SL={SL4}, {'every' if all_m else ''} m in seeded random order from [1,{m_hi}) coprime to d, unknowns "
                    f"sieved to {sieve_limit:.0e} ({np.mean([len(p) + len(q) for p, q in lists]):.0f} per m), Fermat-2 scan "
                    f"with Baillie-PSW confirmation of each side's first PRP (strong base 2 + strong Lucas), per-round depth sized to the resident lanes (cap 256)",
        "m": n_m,
        "capacity_m_per_s": cap_mps,
        "useful_tests_per_s": useful / cap_s,
        "issued_tests_per_s": (s1["tests_issued"] - s0["tests_issued"]) / cap_s,
        "rounds": s1["rounds"] - s0["rounds"],
        "offered_load": load,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "single_m_p50_ms": float(np.percentile(single, 50)),
        "single_m_p99_ms": float(np.percentile(single, 99)),
        "latency_phase_s": lat_s,
        "sieve_s_untimed": sieve_s,
        "slots": s1["slots"],
    }
    if cpu_sample:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_lib
        ids = cap_set[:cpu_sample]
        t0 = time.perf_counter()
        rows = oracle_lib.scan_many(P4, D4, SL4, [int(order[i]) for i in ids], [lists[i] for i in ids], 0, threads)
        dt = time.perf_counter() - t0
        mism = sum(1 for k, i in enumerate(ids)
                   if (rows_cap[i][0].prev_p, rows_cap[i][0].next_p) != (rows[k][0], rows[k][1]))
        out["cpu_baseline"] = {"m_per_s": len(ids) / dt, "tests_per_s": float(rows[:, 2].sum()) / dt,
                               "cores": threads, "kind": "port",
                               "sample": f"{len(ids)} of the same m's, gap_test_simple semantics "
                                         f"(mpz_probab_prime_p(N,25) scan), {threads} threads, {dt:.1f} s",
                               "host": cpu_info}
        out["parity_sample_mismatches"] = mism
        out["parity_sample_m"] = len(ids)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=20261202 + 1)
    ap.add_argument("--cpu-sample", type=int, default=393216,
                    help="candidates the CPU baseline tests (about 10 s on 16 host threads)")
    ap.add_argument("--stream-cpu-sample", type=int, default=768,
                    help="m's the CPU scan baseline runs (about 10 s on 16 host threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-2180", action="store_true", help="skip the 2180-bit batch")
    ap.add_argument("--no-stream", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-stream-4300", action="store_true", help="skip the large-operand serving measurement")
    ap.add_argument("--stream-4300-m", type=int, default=2048)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-width batch")
    ap.add_argument("--no-plain-503", action="store_true", help="skip the plain 503#/210 file scan")
    ap.add_argument("--mad-peak", action="store_true", help="only measure the v_mad_u64_u32 rate")
    ap.add_argument("--no-4300", action="store_true", help="skip the 4300-bit batch")
    ap.add_argument("--batch-4300-lanes", type=int, default=0,
                    help="lanes per number for the 4300-bit batch (0: the engine's choice, 4 lanes x 38 digits; "
                         "8: the 8 x 19 variant)")
    ap.add_argument("--no-700", action="store_true", help="skip the 700-bit batch")
    ap.add_argument("--stream-m", type=int, default=16384)
    ap.add_argument("--stream-load", type=float, default=0.5)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()

    if args.copy_calibration:
        import hipmem
        n = args.copy_calibration << 20
        a = hipmem.DeviceArray(n, np.uint8).zero()
        b = hipmem.DeviceArray(n, np.uint8)
        for _ in range(3):
            b.copy_from(a)
        hipmem.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": n}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(m's and candidates are independent, so a multi-GPU run is a host-side partition "
              "of this workload).", file=sys.stderr)
        return 2

    import hipmem
    import pgap

    engine = pgap.Engine(0)   # raises PgapError when no GPU / no HIP code object
    if args.mad_peak:
        pk = measured_peak(engine)
        print(json.dumps({"mad_peak": pk, "nominal_mac_per_s": PEAK_INT32_MAC,
                          "frac_of_nominal": pk["mac_per_s"] / PEAK_INT32_MAC}), flush=True)
        engine.close()
        return 0
    peak = measured_peak(engine)
    fam = engine.family(P, D, 0)
    t_setup = time.perf_counter()
    m, x = draw_candidates(engine, fam, args.count, args.seed)
    K = math.prod(p for p in range(2, P + 1) if all(p % q for q in range(2, int(p ** 0.5) + 1))) // D
    max_n = K * M_HI + X_RANGE
    bits_hi = max_n.bit_length()
    limbs = (bits_hi + 31) // 32
    d_n = hipmem.DeviceArray((args.count, limbs), np.uint32)
    d_out = hipmem.DeviceArray(args.count, np.uint8)
    engine.build_candidates_device(fam, m, x, limbs, d_n.ptr)
    hipmem.synchronize()
    # exponent bits per test: bitlen(N - 1) = bitlen(N) (N odd > 2)
    e_bits = np.array([(int(mi) * K + int(xi)).bit_length() for mi, xi in zip(m[:4096], x[:4096])])
    e_mean = float(e_bits.mean())
    setup_s = time.perf_counter() - t_setup

    def step():
        engine.prp_device(d_n.ptr, limbs, args.count, d_out.ptr, pgap.FERMAT2, (), bits_hi)
        engine.sync()
        return engine.last_launch().kernel_ms

    for _ in range(args.warmup):
        step()
    hipmem.synchronize()   # device-wide barrier (the process holds no torch context)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(step())
    hipmem.synchronize()
    elapsed = time.perf_counter() - t0
    info = engine.last_launch()

    tests = args.count * args.steps
    value = tests / elapsed
    avg_kernel_s = float(np.mean(kernel_ms)) * 1e-3
    # algorithmic work per test W = 2 L^2 e int32 MACs, L = 32-bit limbs of N
    L32 = (int(e_mean + 0.999) + 31) // 32
    w_test = 2 * L32 * L32 * e_mean
    achieved = w_test * args.count / avg_kernel_s
    w_container = 2 * 40 * 40 * e_mean   # the 1280-bit (40-limb) container

    verd = d_out.to_host()
    # the same batch through pgap_prp_batch (host buffers in and out, PCIe
    # included) -- reported beside `value`, never as it
    host_n = d_n.to_host()
    engine.prp(host_n, pgap.FERMAT2)
    t_h = time.perf_counter()
    host_verd = engine.prp(host_n, pgap.FERMAT2)
    pcie_rate = args.count / (time.perf_counter() - t_h)
    del host_n
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"fixed-width Fermat base-2 batch: N = m*{P}#/{D} + x, m~U[1e7,1.2e7) coprime to d, "
                        f"x~U[-15000,15000], no prime factor <= 1e6, {args.count} candidates "
                        f"({int(e_bits.min())}-{int(e_bits.max())} bits), seed {args.seed}",
            "global_batch": args.count,
            "bits": round(e_mean, 1),
            "kernel": f"radix-2^{info.radix_bits} x {info.digits} digits, {info.lanes_per_number} lane(s)/number",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": peak["mac_per_s"] / 1e12,
            "unit": "TOP/s (int32 multiply-add)",
            "frac": achieved / peak["mac_per_s"],
            "traffic": load_traffic(f"fermat2_r{info.radix_bits}_d{info.digits}"),
            "peak_source": f"measured in this run: pgap_mad_peak, {peak['macs_per_launch']} v_mad_u64_u32 in "
                           f"{peak['kernel_ms']:.2f} ms (best of 3)",
            "peak_nominal": PEAK_INT32_MAC / 1e12,
            "frac_of_nominal": achieved / PEAK_INT32_MAC,
            "work_per_test": f"2*L^2*e int32 MACs, L={L32} limbs (minimal), e={e_mean:.1f}",
            "work_per_launch": w_test * args.count,
            "frac_container_40_limbs": w_container * args.count / avg_kernel_s / peak["mac_per_s"],
            "frac_container_40_limbs_of_nominal": w_container * args.count / avg_kernel_s / PEAK_INT32_MAC,
            "kernel_ms_avg": avg_kernel_s * 1e3,
        },
        "prp_found": int(verd.sum()),
        "host_buffer_tests_per_s": pcie_rate,
        "host_buffer_verdicts_identical": bool((host_verd == verd).all()),
        "setup_s": round(setup_s, 2),
    }
    threads, cpu_info = host_cpus()
    if not args.no_cpu:
        host = d_n.to_host(rows=args.cpu_sample)
        rate, dt, cv = cpu_baseline(host, args.cpu_sample, threads)
        mism = int((cv != verd[: args.cpu_sample]).sum())
        result["cpu_baseline"] = {
            "value": rate,
            "unit": "tests/s",
            "cores": threads,
            "kind": "port",
            "sample": f"first {args.cpu_sample} of the same candidates, Fermat base 2 via GMP mpz_powm, "
                      f"{threads} threads, {dt:.1f} s",
            "host": cpu_info,
        }
        result["parity_sample_mismatches"] = mism
    extra = {}
    if not args.no_2180:
        extra["batch_2180_bit"] = bench_batch_2180(engine, hipmem, pgap, 1 << 18, 3, args.seed + 1, peak["mac_per_s"])
    if not args.no_mixed:
        extra["mixed_widths"] = bench_mixed(engine, hipmem, pgap, 1 << 19, args.seed + 3)
        extra["mixed_widths_with_4247_bit"] = bench_mixed(engine, hipmem, pgap, 1 << 17, args.seed + 4, MIX_WIDE)
    if not args.no_700:
        extra["batch_700_bit"] = bench_batch_700(engine, hipmem, pgap, 1 << 20, args.seed + 7, peak["mac_per_s"])
    if not args.no_4300:
        e4 = pgap.Engine(0, lanes=args.batch_4300_lanes) if args.batch_4300_lanes else engine
        extra["batch_4300_bit"] = bench_batch_4300(e4, hipmem, pgap, 1 << 15, args.seed + 4, peak["mac_per_s"])
        if e4 is not engine:
            e4.close()
    if not args.no_plain_503:
        # the reference's CPU-runnable case: 503#/210, SL=4000, every m < 2000
        # coprime to 210, unknowns sieved to 1e8 (a fresh engine: a scan's
        # slots are sized by the families it starts with)
        engine3 = pgap.Engine(0)
        extra["plain_503"] = bench_stream(engine3, pgap, 0, args.stream_load, args.seed + 6,
                                          0 if args.no_cpu else 10 ** 9, threads, P4=503, D4=210, SL4=4000,
                                          m_hi=2000, sieve_limit=100_000_000, all_m=True, cpu_info=cpu_info)
        engine3.close()
    if not args.no_stream:
        extra["stream"] = bench_stream(engine, pgap, args.stream_m, args.stream_load, args.seed + 2,
                                       0 if args.no_cpu else args.stream_cpu_sample, threads, cpu_info=cpu_info)
    if not args.no_stream_4300:
        # P=3001, d=2190, SL=40000, m < 1e5, unknowns sieved to 1e10
        # (a fresh engine: a scan's slots are sized by the families it starts with)
        engine2 = pgap.Engine(0)
        extra["stream_4300_bit"] = bench_stream(engine2, pgap, args.stream_4300_m, args.stream_load, args.seed + 5,
                                                0 if args.no_cpu else 16, threads, P4=3001, D4=2190, SL4=40000,
                                                m_hi=100000, sieve_limit=10_000_000_000, cpu_info=cpu_info)
        engine2.close()
    if extra:
        result["extra"] = extra
    print(json.dumps(result), flush=True)
    engine.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
