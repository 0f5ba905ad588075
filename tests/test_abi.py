"""The C-ABI library loads and exports every entry point include/pgap.h
declares.  No compute call runs here (this container has no GPU)."""
import ctypes
import os
import re

import pytest

import pgap

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "pgap.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pgap_[a-z_0-9]+)\s*\(", src)))


def test_library_built():
    assert os.path.exists(pgap.LIB_PATH), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(pgap.LIB_PATH)
    names = declared_functions()
    assert set(names) == set(pgap.EXPORTED)
    for name in names:
        assert hasattr(lib, name), name


def test_version_and_null_handling():
    lib = pgap.load()
    assert lib.pgap_version().startswith(b"pgap ")
    assert lib.pgap_prp_batch(None, 4, None, 1, 0, None, 0, None) == pgap.PGAP_E_ARG
    assert lib.pgap_last_error(None) == b"null engine"
    lib.pgap_destroy(None)


def test_no_gpu_fails_loudly():
    try:
        e = pgap.Engine(0)
    except pgap.PgapError as err:
        assert err.status == pgap.PGAP_E_DEVICE
        return
    e.close()
    pytest.skip("a GPU is present")


def test_struct_layout_and_constants_match_the_header(tmp_path):
    # the ctypes mirrors in pgap.py against what a C++ compiler makes of
    # include/pgap.h (sizes, field offsets, enum values)
    import subprocess

    src = tmp_path / "layout.cpp"
    src.write_text(
        '#include <cstddef>\n#include <cstdio>\n#include "pgap.h"\n'
        "#define F(t, f) std::printf(#t \".\" #f \" %zu\\n\", offsetof(t, f))\n"
        "int main() {\n"
        '  std::printf("pgap_scan_req %zu\\npgap_scan_row %zu\\npgap_launch_info %zu\\n",\n'
        "              sizeof(pgap_scan_req), sizeof(pgap_scan_row), sizeof(pgap_launch_info));\n"
        "  F(pgap_scan_req, family); F(pgap_scan_req, m); F(pgap_scan_req, prev); F(pgap_scan_req, n_prev);\n"
        "  F(pgap_scan_req, next); F(pgap_scan_req, n_next); F(pgap_scan_req, tag); F(pgap_scan_req, sides);\n"
        "  F(pgap_scan_row, m); F(pgap_scan_row, tag); F(pgap_scan_row, prev_p); F(pgap_scan_row, next_p);\n"
        "  F(pgap_scan_row, p_tests); F(pgap_scan_row, n_tests); F(pgap_scan_row, flags);\n"
        "  F(pgap_scan_row, family); F(pgap_scan_row, t_submit); F(pgap_scan_row, t_done);\n"
        '  std::printf("PGAP_ROW_NEXT_SKIPPED %d\\nPGAP_ROW_PREV_SKIPPED %d\\nPGAP_SIDES_PREV %d\\n'
        'PGAP_SIDES_NEXT %d\\nPGAP_E_TIMEOUT %d\\nPGAP_BPSW %d\\nPGAP_CONFIRM_BPSW %d\\n",\n'
        "              PGAP_ROW_NEXT_SKIPPED, PGAP_ROW_PREV_SKIPPED, PGAP_SIDES_PREV, PGAP_SIDES_NEXT,\n"
        "              PGAP_E_TIMEOUT, PGAP_BPSW, PGAP_CONFIRM_BPSW);\n"
        "}\n")
    exe = tmp_path / "layout"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(REPO, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines())
    got = {k: int(v) for k, v in got.items()}
    assert got["pgap_scan_req"] == ctypes.sizeof(pgap.ScanReq)
    assert got["pgap_scan_row"] == ctypes.sizeof(pgap.ScanRow)
    for cls, name in ((pgap.ScanReq, "pgap_scan_req"), (pgap.ScanRow, "pgap_scan_row")):
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == got[f"{name}.{field}"], (name, field)
    assert (got["PGAP_ROW_NEXT_SKIPPED"], got["PGAP_ROW_PREV_SKIPPED"]) == (pgap.ROW_NEXT_SKIPPED,
                                                                            pgap.ROW_PREV_SKIPPED)
    assert (got["PGAP_SIDES_PREV"], got["PGAP_SIDES_NEXT"]) == (pgap.SIDES_PREV, pgap.SIDES_NEXT)
    assert got["PGAP_E_TIMEOUT"] == pgap.PGAP_E_TIMEOUT
    assert got["PGAP_BPSW"] == pgap.BPSW and got["PGAP_CONFIRM_BPSW"] == pgap.CONFIRM_BPSW
