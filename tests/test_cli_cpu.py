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
