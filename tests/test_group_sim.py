"""The parallel formulation of the child grouping (binary-searched group
ends plus pointer doubling, as the column-scan kernel computes it) against
the sequential greedy walk it replaces, on the CPU (tests/cpp/group_sim.cpp)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_parallel_grouping_matches_greedy_walk(tmp_path):
    exe = tmp_path / "group_sim"
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", str(ROOT / "tests" / "cpp" / "group_sim.cpp"), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0/200000 mismatches" in r.stdout
