"""gap_test_hip input handling on the CPU (--dump-unknowns runs no GPU code):
the C++ reader of all three unknown-file formats, file-name grammar, and the
error paths the reference turns into asserts (gap_common.cpp:1186-1359)."""
import hashlib
import os
import subprocess

import pytest

import oracle_lib as ol
import unknown_format as uf

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "prime-gap_amd", "gap_test_hip")
FN = "907_2190_1_200_s11000_l100M.txt"
PLAIN_MD5 = "15a5cbff7301262caf047028c05f0525"  # misc/tests.sh:66-67


def run(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, cwd=cwd, timeout=120)


@pytest.fixture(scope="module")
def formats(tmp_path_factory, unknowns_907):
    K = ol.K(907, 2190)
    cop = uf.coprime_x(K, 11000)
    out = {}
    for name in ("plain", "rle", "bitcompressed"):
        d = tmp_path_factory.mktemp(name)
        data = b""
        for ln in unknowns_907.splitlines(keepends=True):
            m, p, n = uf.parse_line(ln)
            if name == "plain":
                data += uf.plain_line(m, p, n)
            elif name == "rle":
                data += uf.rle_line(m, p, n)
            else:
                data += uf.bitcompressed_line(m, p, n, K, 2190, 11000, cop)
        (d / FN).write_bytes(data)
        out[name] = str(d / FN)
    return out


@pytest.mark.parametrize("name", ["plain", "rle", "bitcompressed"])
def test_reader_all_formats(formats, name):
    r = run("-u", formats[name], "--dump-unknowns", "-qqq")
    assert r.returncode == 0, r.stderr
    assert hashlib.md5(r.stdout).hexdigest() == PLAIN_MD5


def test_bare_name_resolves_under_unknowns_dir(formats, tmp_path):
    (tmp_path / "unknowns").mkdir()
    os.symlink(formats["rle"], tmp_path / "unknowns" / FN)
    r = run("-u", FN, "--dump-unknowns", "-qqq", cwd=tmp_path)
    assert r.returncode == 0 and hashlib.md5(r.stdout).hexdigest() == PLAIN_MD5


def test_mskip_skips_lines(formats):
    r = run("-u", formats["plain"], "--dump-unknowns", "-qqq", "--mskip", "100")
    # gap_test_simple.cpp:74-78 prints the skip notice at every verbosity
    assert r.stdout.startswith(b"\tskipping m < 100\n")
    ms = [int(l.split(b" ")[0]) for l in r.stdout.splitlines()[1:]]
    assert r.returncode == 0 and ms and min(ms) >= 100 and 199 in ms


def test_header_lines(formats):
    r = run("-u", formats["plain"], "--dump-unknowns", "-q")
    out = r.stdout.decode()
    assert "Testing m * 907#/2190, m = 1 + [0, 200)" in out
    assert "53 tests M_start(1) + mi(0 to 198)" in out   # README.md:241-250


@pytest.mark.parametrize("bad", ["907_2190_1_200_s11000.txt", "907_2190_1_200_s11000_l100M.dat",
                                 "908_2190_1_200_s11000_l100M.txt"])
def test_bad_names_fail(bad, tmp_path):
    r = run("-u", str(tmp_path / bad), "--dump-unknowns")
    assert r.returncode != 0


def test_malformed_and_truncated_lines_fail(formats, tmp_path):
    lines = open(formats["plain"], "rb").read().splitlines(keepends=True)
    # truncated list (header promises more entries than the line holds)
    bad = lines[:3] + [lines[3][: len(lines[3]) // 2] + b"\n"] + lines[4:]
    p = tmp_path / FN
    p.write_bytes(b"".join(bad))
    r = run("-u", str(p), "--dump-unknowns", "-qqq")
    assert r.returncode != 0 and b"bad line" in r.stderr
    # file ends early
    p.write_bytes(b"".join(lines[:10]))
    r = run("-u", str(p), "--dump-unknowns", "-qqq")
    assert r.returncode != 0 and b"ended before" in r.stderr
    # first line is not the first coprime m
    p.write_bytes(b"".join(lines[1:]))
    r = run("-u", str(p), "--dump-unknowns", "-qqq")
    assert r.returncode != 0
    # missing file
    r = run("-u", str(tmp_path / "x" / FN), "--dump-unknowns")
    assert r.returncode != 0


def test_db_tables_and_resume(formats, tmp_path):
    # --db: the result / m_stats tables are created when absent, and an m whose
    # result row has next_p >= 0 is not read again (gap_test.py:341-349: 0 is a
    # side-skipped m, -1 a partial row that is tested again); rows of another
    # D are not this range's.
    import sqlite3

    db = tmp_path / "gaps.db"
    r = run("-u", formats["plain"], "--dump-unknowns", "-qqq", "--db", str(db))
    assert r.returncode == 0 and hashlib.md5(r.stdout).hexdigest() == PLAIN_MD5
    con = sqlite3.connect(db)
    cols = {t: [c[1] for c in con.execute(f"PRAGMA table_info({t})")] for t in ("result", "m_stats")}
    assert cols["result"] == ["P", "D", "m", "next_p", "prev_p", "merit"]
    assert {"next_p", "prev_p", "merit", "prp_next", "prp_prev", "test_time"} <= set(cols["m_stats"])
    con.executemany("INSERT INTO result VALUES(?, ?, ?, ?, ?, ?)",
                    [(907, 2190, 7, 10, 20, 1.0), (907, 2190, 11, 0, 30, 0.5),
                     (907, 2190, 13, -1, 40, 0.0), (907, 2310, 17, 8, 8, 0.1)])
    con.commit()
    con.close()
    r = run("-u", formats["plain"], "--dump-unknowns", "-q", "--db", str(db))
    assert r.returncode == 0, r.stderr
    out = r.stdout.decode()
    assert "Found 3 results" in out
    ms = {int(l.split(" ")[0]) for l in out.splitlines() if l[:1].isdigit() and " : " in l}
    assert 7 not in ms and 11 not in ms and {1, 13, 17, 199} <= ms
    assert len(ms) == 53 - 2


def test_db_unwritable_path_fails(formats, tmp_path):
    r = run("-u", formats["plain"], "--dump-unknowns", "-qqq", "--db", str(tmp_path / "no" / "dir" / "x.db"))
    assert r.returncode != 0 and b"database" in r.stderr


def test_db_prp_top_percent(formats, tmp_path):
    # determine_test_threshold (gap_test_stats.py:456-486): the threshold is
    # the prob_record at index round(n * percent / 100) of the descending
    # list; an m at or above it is tested, and so is a partial row.
    import sqlite3

    import numpy as np

    lines = open(formats["plain"], "rb").read().splitlines()
    ms = [int(l.split(b" ")[0]) for l in lines]
    db = tmp_path / "gaps.db"
    con = sqlite3.connect(db)
    con.execute("CREATE TABLE m_stats (P INTEGER, D INTEGER, m INTEGER, next_p INTEGER DEFAULT 0, "
                "prev_p INTEGER DEFAULT 0, merit REAL DEFAULT 0, prob_record REAL, prp_next INTEGER DEFAULT 0, "
                "prp_prev INTEGER DEFAULT 0, test_time REAL DEFAULT 0, PRIMARY KEY(P, D, m))")
    probs = {m: ((m * 7919) % 101) / 1e4 for m in ms}
    con.executemany("INSERT INTO m_stats(P, D, m, prob_record) VALUES(907, 2190, ?, ?)", probs.items())
    con.execute("CREATE TABLE result (P INTEGER, D INTEGER, m INTEGER, next_p INTEGER, prev_p INTEGER, "
                "merit REAL, PRIMARY KEY(P, D, m))")
    low = min(ms, key=lambda m: probs[m])
    con.execute("INSERT INTO result VALUES(907, 2190, ?, -1, 50, 0)", (low,))
    con.commit()
    con.close()
    for percent in (5, 10, 37, 100):
        r = run("-u", formats["plain"], "--dump-unknowns", "-q", "--db", str(db), "--prp-top-percent", str(percent))
        assert r.returncode == 0, r.stderr
        got = {int(l.split(" ")[0]) for l in r.stdout.decode().splitlines() if l[:1].isdigit() and " : " in l}
        f = {m: np.float32(v) for m, v in probs.items()}
        if percent == 100:
            want = set(ms)
        else:
            best = sorted(f.values(), reverse=True)
            thr = best[round(len(best) * percent / 100)]
            want = {m for m in ms if f[m] >= thr} | {low}
        assert got == want, percent


def test_db_prp_top_percent_needs_stats(formats, tmp_path):
    r = run("-u", formats["plain"], "--dump-unknowns", "--prp-top-percent", "10")
    assert r.returncode != 0
    # no m_stats rows from gap_stats in a fresh database
    r = run("-u", formats["plain"], "--dump-unknowns", "--db", str(tmp_path / "x.db"), "--prp-top-percent", "10")
    assert r.returncode != 0 and b"gap_stats" in r.stderr


def _search_db(path, P, D, ms, prob=0.5):
    import sqlite3
    con = sqlite3.connect(path)
    con.execute("CREATE TABLE m_stats (P INTEGER, D INTEGER, m INTEGER, next_p INTEGER DEFAULT 0, "
                "prev_p INTEGER DEFAULT 0, merit REAL DEFAULT 0, prob_record REAL, prob_merit REAL, "
                "prp_next INTEGER DEFAULT 0, prp_prev INTEGER DEFAULT 0, test_time REAL DEFAULT 0, "
                "PRIMARY KEY(P, D, m))")
    con.executemany("INSERT INTO m_stats(P, D, m, prob_record, prob_merit) VALUES(?, ?, ?, ?, ?)",
                    [(P, D, m, prob, prob) for m in ms])
    con.commit()
    con.close()


def test_prob_side_skip_arguments(formats, tmp_path):
    r = run("-u", formats["plain"], "--dump-unknowns", "--side-skip=prob")
    assert r.returncode != 0 and b"needs --db" in r.stdout
    r = run("-u", formats["plain"], "--dump-unknowns", "--side-skip=sometimes")
    assert r.returncode != 0 and b"--side-skip takes" in r.stdout
    r = run("-u", formats["plain"], "--dump-unknowns", "--megagap", "5000")
    assert r.returncode != 0 and b"--megagap" in r.stdout
    db = tmp_path / "search.db"
    _search_db(db, 907, 2190, [m for m in range(1, 200) if m % 2 and m % 3 and m % 5 and m % 73])
    r = run("-u", formats["plain"], "--dump-unknowns", "--side-skip=prob", "--db", str(db),
            "--prime-gaps-db", str(tmp_path / "missing.db"))
    assert r.returncode != 0 and b"prime gaps database" in r.stderr


def test_record_gaps_query_and_missing_gaps(tmp_path):
    # load_records (gap_test_stats.py:116-142): gaps whose merit at ln N
    # would beat their record, below merit 35; and, once 30 ln N passes
    # 113326, every even size from there to 30 ln N absent from the list.
    # 4001#/2190 with m = 1: ln N = 3911.7, so 30 ln N = 117352.
    import sqlite3

    P, D = 4001, 2190
    log_n = ol.K_log(P, D)
    assert 113326 < int(30 * log_n) < 120000
    unk = tmp_path / f"{P}_{D}_1_3_s100_l100M.txt"
    unk.write_bytes(uf.plain_line(1, [1, 3], [5, 7]))
    db = tmp_path / "search.db"
    _search_db(db, P, D, [1])
    gaps = tmp_path / "gaps.db"
    con = sqlite3.connect(gaps)
    con.execute("CREATE TABLE gaps (gapsize INTEGER, merit REAL)")
    table = [(10000, 2.0), (10002, 2.6), (20000, 5.0), (20002, 5.2), (136000, 34.0), (137600, 35.1),
             (113326, 29.0), (113400, 28.9), (117000, 30.0)]
    con.executemany("INSERT INTO gaps VALUES(?, ?)", table)
    con.commit()
    con.close()
    listed = [g for g, mr in table if g > mr * log_n - 0.001 and g < 35 * log_n]
    present = {g for g, _ in table if 113326 <= g <= int(30 * log_n)}
    missing = [g for g in range(113326, int(30 * log_n) + 1, 2) if g not in present]
    want = sorted(listed + missing)
    r = run("-u", str(unk), "--dump-unknowns", "--side-skip=prob", "--db", str(db), "--prime-gaps-db", str(gaps))
    assert r.returncode == 0, r.stderr
    assert f"Loaded {len(want)} records ({want[0]} to {want[-1]})".encode() in r.stdout
    # merit 2.0 / 5.0 at ln N 3911.7 beats 10000 / 20000; 2.6 / 5.2 do not;
    # 137600 is past merit 35; a listed size in the missing range is a
    # record only by its merit (113400 yes, 113326 and 117000 no)
    assert {10000, 20000, 136000, 113328, 113400} <= set(want)
    assert not {10002, 20002, 137600, 113326, 117000} & set(want)
