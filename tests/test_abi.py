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
