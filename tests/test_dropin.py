"""The C++ drop-in classes (include/vren) through their test driver, tests/cpp/drop_in_test.cpp.

CPU: the driver's host-only mode checks that preconditions surface as the reference's exceptions
(std::invalid_argument for radix_sort.cpp:158-161, build_bvh.cpp:45 and the camera/config checks).
GPU: every primitive against std:: on the host, and cluster_and_shade against the batch chain, byte for byte."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "drop_in_test")


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/drop_in_test is not built (run __graft_entry__.build() or tests/cpp/build.py)")
    return DRIVER


def test_drop_in_host_only():
    r = subprocess.run([_driver(), "--host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout


@pytest.mark.gpu
def test_drop_in_on_device():
    print("[gpu] tests/cpp/drop_in_test", flush=True)
    r = subprocess.run([_driver()], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
