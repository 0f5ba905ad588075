"""The test_runner_t drop-in (include/pgap_test_runner.h) builds against the C
ABI and GMP as a caller of the reference would build it, and without a GPU
its constructor fails with a message (status 1), never exit() inside the
library.  Its results on a GPU are checked in test_gpu_shim.py."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "tests", "shim", "run_test_main")


def test_driver_built():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"


def test_rounds_outside_range_refused():
    r = subprocess.run([DRIVER, "9"], input=b"341\n", capture_output=True, timeout=60)
    assert r.returncode == 1 and b"rounds must be in [1, 8]" in r.stderr


def test_no_gpu_reports_and_fails():
    r = subprocess.run([DRIVER, "1"], input=b"341\n561\n", capture_output=True, timeout=120)
    if r.returncode == 0:
        pytest.skip("a GPU is present")
    assert r.returncode == 1 and b"test_runner_t:" in r.stderr


def test_poll_wait_gives_up_when_the_worker_stalls():
    # the bounded wait pgap_scan_poll uses, driven by a producer thread that
    # posts in time, then stalls (tests/shim/waitq_test.cpp)
    exe = os.path.join(REPO, "tests", "shim", "waitq_test")
    r = subprocess.run([exe], capture_output=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == b"ok", r.stdout
