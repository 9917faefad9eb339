"""Runs the C++ adapter test (reference-shaped rsfm:: API over the C ABI) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "cpp", "test_adapters")


def test_cpp_adapters(dev):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_adapters is not built (run __graft_entry__.build())")
    print("[case] tests/cpp/test_adapters", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout
