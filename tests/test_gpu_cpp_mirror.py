"""Runs the C++ drop-in test binary (include/paulib_hip.hpp on a host sum with
PauliSum<N,double>'s fields, compared bit for bit with the oracle)."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "_build" / "mirror_test"


@pytest.mark.gpu
def test_cpp_mirror():
    assert BIN.exists(), "build first (__graft_entry__.build())"
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
