"""gap_test_hip end to end on the GPU: rows and stats lines equal what
gap_test_simple prints (gap_test_simple.cpp:101-200, gap_test_common.cpp:59-124)
for the oracle's per-m results, after the reference's own normalisation of
timing figures (misc/tests.sh:119-120)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
import unknown_format as uf

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "prime-gap_amd", "gap_test_hip")
FIXTURE = os.path.join(REPO, "tests", "golden", "907_2190_1_200_s11000_l100M.txt")


def normalise(text: str) -> str:
    # misc/tests.sh:119-120
    text = re.sub(r"\([0-9.]+ secs/test\)", "(XX.YY/sec)", text)  # a slow first m prints secs/test
    text = re.sub(r"[0-9.]+( ?)(tests)?/sec\)", r"XX.YY\1\2/sec)", text)
    text = re.sub(r"[0-9]+ sec", "Z sec", text)
    return re.sub(r"Reading unknowns from '[^']*'", "Reading unknowns from 'FILE'", text)


def expected_stdout(P, D, ms, mi, SL, lines, min_merit, verbose):
    K_log = ol.K_log(P, D)
    out = []
    w = out.append
    mm = np.float32(min_merit)
    w("\n")
    w(f"Testing m * {P}#/{D}, m = {ms} + [0, {mi})\n")
    w("\nReading unknowns from 'FILE'\n")
    w(f"Min Gap ~= {int(float(mm) * (K_log + float(np.float32(math.log(ms)))))} (for merit > {float(mm):.1f})\n")
    valid = [m for m in range(ms, ms + mi) if math.gcd(m, D) == 1]
    last_mi = max(m - ms for m in valid)
    w(f"\n{len(valid)} tests M_start({ms}) + mi(0 to {last_mi})\n\n")
    s_tests = tu = tp = tn = prp = 0
    for ln in lines:
        m, p, n = uf.parse_line(ln)
        r = ol.test_interval(P, D, m, SL, p, n, kind=0)
        gap = r["prev_p"] + r["next_p"]
        merit = np.float32(gap / (K_log + math.log(m)))
        if merit > mm:
            w("%-5d %.4f  %d * %d#/%d -%d to +%d\n" % (gap, merit, m, P, D, r["prev_p"], r["next_p"]))
        s_tests += 1
        tu += len(p) + len(n)
        tp += len(p)
        tn += len(n)
        prp += r["tests"]
        is_last = (m - ms) == last_mi
        power = s_tests == 1
        q = 10
        while q <= s_tests:
            power |= s_tests in (q, 3 * q, 5 * q)
            q *= 10
        if (is_last or power) and verbose + is_last >= 1:
            w("\tm=%d %4d <- unknowns -> %-4d\t%4d <- gap -> %-4d\n" % (m, len(p), len(n), r["prev_p"], r["next_p"]))
            w("\t    tests     %-10d (XX.YY/sec)  Z seconds elapsed\n" % s_tests)
            w("\t    unknowns  %-10d (avg: %.2f), %.2f%% composite  %.2f%% <- %% -> %.2f%%\n" % (
                tu, tu / s_tests, 100.0 * (1 - tu / ((2.0 * SL + 1) * s_tests)), 100.0 * tp / tu, 100.0 * tn / tu))
            w("\t    prp tests %-10d (avg: %.2f) (XX.YY tests/sec)\n" % (prp, np.float32(prp / np.float32(s_tests))))
    return "".join(out)


def test_cli_matches_gap_test_simple_output():
    r = subprocess.run([CLI, "-u", FIXTURE, "--min-merit", "8", "-q"], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    lines = open(FIXTURE, "rb").read().splitlines(keepends=True)
    want = expected_stdout(907, 2190, 1, 200, 11000, lines, 8.0, 1)
    assert normalise(r.stdout.decode()) == want


def test_cli_results_file_and_mskip(tmp_path):
    res = tmp_path / "results.tsv"
    r = subprocess.run([CLI, "-u", FIXTURE, "--min-merit", "8", "-qq", "--mskip", "100", "--depth", "2",
                        "--results", str(res)], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    rows = [l.split("\t") for l in res.read_text().splitlines()]
    lines = {uf.parse_line(ln)[0]: uf.parse_line(ln) for ln in open(FIXTURE, "rb").read().splitlines(keepends=True)}
    assert [int(x[2]) for x in rows] == [m for m in sorted(lines) if m >= 100]
    for P, D, m, next_p, prev_p, merit in rows:
        _, p, n = lines[int(m)]
        want = ol.test_interval(907, 2190, int(m), 11000, p, n, kind=0)
        assert (int(prev_p), int(next_p)) == (want["prev_p"], want["next_p"])


def test_cli_side_skip_rows_and_count():
    # --side-skip: gap_test_gpu.cu:591-617 -- an m whose prev gap's merit is
    # below min_merit/2 - 2 prints no row and counts as "only prev_prime".
    min_merit = 6.0
    r = subprocess.run([CLI, "-u", FIXTURE, "--min-merit", str(min_merit), "--side-skip"], capture_output=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    out = r.stdout.decode()
    K_log = ol.K_log(907, 2190)
    thr = np.float32(np.float32(min_merit) / np.float32(2) - np.float32(2))
    want_rows, skips = [], 0
    for ln in open(FIXTURE, "rb").read().splitlines(keepends=True):
        m, p, n = uf.parse_line(ln)
        w = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
        if np.float32(w["prev_p"] / (K_log + math.log(m))) < thr:
            skips += 1
            continue
        gap = w["prev_p"] + w["next_p"]
        merit = np.float32(gap / (K_log + math.log(m)))
        if merit > np.float32(min_merit):
            want_rows.append("%-5d %.4f  %d * 907#/2190 -%d to +%d" % (gap, merit, m, w["prev_p"], w["next_p"]))
    got_rows = [l for l in out.splitlines() if re.match(r"^\d+\s+\d+\.\d{4}  \d+ \* ", l)]
    assert got_rows == want_rows
    assert 0 < skips
    assert re.findall(r"only prev_prime (\d+) ", out)[-1] == str(skips)


def _fixture_rows():
    return {uf.parse_line(ln)[0]: uf.parse_line(ln) for ln in open(FIXTURE, "rb").read().splitlines(keepends=True)}


def test_cli_db_rows_m_stats_and_resume(tmp_path):
    # gap_test_stats.save (gap_test_stats.py:217-238): REPLACE INTO result,
    # m_stats next_p / prev_p / merit set and prp_next / prp_prev / test_time
    # added to; load_existing (gap_test_stats.py:145-150): a second run over
    # the range tests only the m's without a finished row.
    import sqlite3

    db = tmp_path / "gaps.db"
    lines = _fixture_rows()
    K_log = ol.K_log(907, 2190)
    con = sqlite3.connect(db)
    # m_stats as gap_stats leaves it: one row per m, with columns of its own
    con.execute("CREATE TABLE m_stats (P INTEGER, D INTEGER, m INTEGER, next_p INTEGER DEFAULT 0, "
                "prev_p INTEGER DEFAULT 0, merit REAL DEFAULT 0, prob_record REAL, prp_next INTEGER DEFAULT 0, "
                "prp_prev INTEGER DEFAULT 0, test_time REAL DEFAULT 0, PRIMARY KEY(P, D, m))")
    con.executemany("INSERT INTO m_stats(P, D, m, prob_record, prp_next, prp_prev, test_time) "
                    "VALUES(907, 2190, ?, ?, 1, 2, 0.5)", [(m, m / 1000) for m in lines])
    con.commit()
    con.close()

    r = subprocess.run([CLI, "-u", FIXTURE, "-qq", "--mskip", "150", "--db", str(db)], capture_output=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    con = sqlite3.connect(db)
    res = {row[0]: row[1:] for row in con.execute("SELECT m, next_p, prev_p, merit FROM result WHERE P = 907 AND D = 2190")}
    assert sorted(res) == [m for m in sorted(lines) if m >= 150]
    for m, (next_p, prev_p, merit) in res.items():
        _, p, n = lines[m]
        w = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
        assert (prev_p, next_p) == (w["prev_p"], w["next_p"]), m
        assert abs(merit - round((prev_p + next_p) / (K_log + math.log(m)), 4)) < 1.5e-4, m
        st = con.execute("SELECT next_p, prev_p, merit, prp_next, prp_prev, test_time, prob_record FROM m_stats "
                         "WHERE P = 907 AND D = 2190 AND m = ?", (m,)).fetchone()
        assert st[:3] == (next_p, prev_p, merit)
        assert (st[3], st[4]) == (1 + w["n_tests"], 2 + w["p_tests"]), m
        assert st[5] > 0.5 and st[6] == m / 1000
    untouched = con.execute("SELECT COUNT(*) FROM m_stats WHERE m < 150 AND prp_next = 1 AND next_p = 0").fetchone()
    assert untouched[0] == sum(1 for m in lines if m < 150)
    con.close()

    # resume: the whole range again -- only m < 150 are tested now
    r = subprocess.run([CLI, "-u", FIXTURE, "-q", "--db", str(db)], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    assert f"Found {len(res)} results" in r.stdout.decode()
    con = sqlite3.connect(db)
    assert con.execute("SELECT COUNT(*) FROM result").fetchone()[0] == len(lines)
    # rows of the first run were not tested twice: their prp counts stand
    for m in res:
        _, p, n = lines[m]
        w = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
        st = con.execute("SELECT prp_next, prp_prev FROM m_stats WHERE m = ?", (m,)).fetchone()
        assert st == (1 + w["n_tests"], 2 + w["p_tests"])
    con.close()


def test_cli_db_side_skip_rows(tmp_path):
    # a side-skipped m is saved with next_p = 0 and the prev side's merit
    # (gap_test.py:164-165, gap_test_stats.py:574-577); its m_stats row is
    # added when gap_stats made none.
    import sqlite3

    db = tmp_path / "gaps.db"
    min_merit = 6.0
    r = subprocess.run([CLI, "-u", FIXTURE, "-qq", "--min-merit", str(min_merit), "--side-skip", "--db", str(db)],
                       capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    K_log = ol.K_log(907, 2190)
    thr = np.float32(np.float32(min_merit) / np.float32(2) - np.float32(2))
    con = sqlite3.connect(db)
    skipped = 0
    for m, (_, p, n) in _fixture_rows().items():
        w = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
        row = con.execute("SELECT next_p, prev_p FROM result WHERE m = ?", (m,)).fetchone()
        st = con.execute("SELECT next_p, prev_p, prp_next FROM m_stats WHERE m = ?", (m,)).fetchone()
        if np.float32(w["prev_p"] / (K_log + math.log(m))) < thr:
            skipped += 1
            assert row == (0, w["prev_p"]) and st == (0, w["prev_p"], 0), m
        else:
            assert row == (w["next_p"], w["prev_p"]) and st == (w["next_p"], w["prev_p"], w["n_tests"]), m
    assert skipped > 0
    con.close()


@pytest.mark.parametrize("megagap,top", [(0, 0), (3000, 0), (0, 60)])
def test_cli_probabilistic_side_skip(tmp_path, megagap, top):
    # --side-skip=prob: gap_test.py's one-side skip (gap_test.py:184-202,
    # gap_test_stats.py:428-453) -- prev side first, the next side only when
    # prob(record | prev gap) beats the running average; skipped m's saved
    # with next_p = 0.  Decisions from the oracle's restatement, in m order.
    import sqlite3
    from test_side_prob import expected_decisions

    P, D, SL = 907, 2190, 11000
    fixture = _fixture_rows()
    ms = sorted(fixture)
    K_log = ol.K_log(P, D)
    db = tmp_path / "search.db"
    con = sqlite3.connect(db)
    con.execute("CREATE TABLE m_stats (P INTEGER, D INTEGER, m INTEGER, next_p INTEGER DEFAULT 0, "
                "prev_p INTEGER DEFAULT 0, merit REAL DEFAULT 0, prob_record REAL, prob_merit REAL, "
                "prp_next INTEGER DEFAULT 0, prp_prev INTEGER DEFAULT 0, test_time REAL DEFAULT 0, "
                "PRIMARY KEY(P, D, m))")
    probs = [float(np.float32(0.8 + 0.02 * (m % 5))) for m in ms]
    con.executemany("INSERT INTO m_stats(P, D, m, prob_record, prob_merit) VALUES(?, ?, ?, ?, 0.5)",
                    [(P, D, m, pr) for m, pr in zip(ms, probs)])
    con.commit()
    con.close()
    gaps = tmp_path / "gaps.db"
    con = sqlite3.connect(gaps)
    con.execute("CREATE TABLE gaps (gapsize INTEGER, merit REAL, primedigits INTEGER)")
    records = list(range(2500, int(35 * K_log), 2))
    con.executemany("INSERT INTO gaps VALUES(?, 1.0, 300)", [(g,) for g in records])
    # not records at this size: a merit above what the gap reaches here, and
    # a gap beyond merit 35
    con.executemany("INSERT INTO gaps VALUES(?, ?, 300)", [(2000, 5.0), (int(36 * K_log) // 2 * 2, 1.0)])
    con.commit()
    con.close()

    cmd = [CLI, "-u", FIXTURE, "--min-merit", "3", "--side-skip=prob", "--db", str(db),
           "--prime-gaps-db", str(gaps)] + (["--megagap", str(megagap)] if megagap else []) + \
        (["--prp-top-percent", str(top)] if top else [])
    r = subprocess.run(cmd, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    out = r.stdout.decode()
    # --prp-top-percent (determine_test_threshold, gap_test_stats.py:456-486):
    # the m's at or above the prob_record at index round(n * top / 100) of
    # the descending list; the one-side threshold starts at 5 x that value
    thr = 0.0
    if top:
        thr = sorted(probs, reverse=True)[round(len(probs) * top / 100)]
        keep = [i for i, pr in enumerate(probs) if pr >= thr]
        ms, probs = [ms[i] for i in keep], [probs[i] for i in keep]
        assert 0 < len(ms) < len(fixture)
    want = {m: ol.test_interval(P, D, m, SL, fixture[m][1], fixture[m][2], kind=0) for m in ms}
    dec = expected_decisions(P, D, SL, 100_000_000, 1, [(m, want[m]["prev_p"], fixture[m][2]) for m in ms],
                             [] if megagap else records, megagap, probs, thr)
    assert 0 < sum(dec) < len(dec)
    con = sqlite3.connect(db)
    res = {row[0]: row[1:] for row in con.execute("SELECT m, next_p, prev_p FROM result")}
    assert sorted(res) == ms
    st = {row[0]: row[1:] for row in con.execute("SELECT m, prp_next, prp_prev FROM m_stats")}
    con.close()
    rows_out = []
    for m, t in zip(ms, dec):
        w = want[m]
        next_p = w["next_p"] if t else 0
        assert res[m] == (next_p, w["prev_p"]), m
        assert st[m] == ((w["n_tests"] if t else 0), w["p_tests"]), m
        gap = w["prev_p"] + next_p
        merit = np.float32(gap / (K_log + math.log(m)))
        rec = gap in set(records)  # the RECORD! tag reads the record list also with --megagap
        if rec or merit > np.float32(3):
            rows_out.append("%-5d %.4f  %d * %d#/%d -%d to +%d%s" % (gap, merit, m, P, D, w["prev_p"], next_p,
                                                                     "\tRECORD!" if rec else ""))
    printed = [ln for ln in out.splitlines() if re.match(r"^\d+ +\d+\.\d{4}  \d+ \* ", ln)]
    assert printed == rows_out
    assert f"only prev_prime {len(ms) - sum(dec)} (" in out
