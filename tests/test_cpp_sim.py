"""The reference's C++ regression checks (tests/sim_tests.cpp), restated in
tests/cpp/sim_tests.cpp against the C++ ``Sim`` shim (stardust_amd/cpp), run as
a child process."""
import os
import subprocess

import pytest

import stardust_amd

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "sim_tests")


def _binary():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/sim_tests is not built (run __graft_entry__.build())")
    return BIN


def test_cpp_sim_without_device_exits_2():
    """Like the reference's main (sim_tests.cpp:547-551): no device -> exit code 2."""
    if stardust_amd.lib().sd_device_available():
        pytest.skip("a device is visible")
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_sim_regressions(gpu):
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 FAIL" in r.stdout
