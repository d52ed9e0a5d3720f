"""The context heap's range and chunk bookkeeping, on the CPU: tests/cpp/
heap_sim.cpp compiles paulib_amd/csrc/plh_heap.hip against a simulated
virtual-memory device (a fake reservation and a cap on physical chunks) and
checks, after every step of a seeded allocate/release sequence, that live
ranges are disjoint and fully backed, that counts add up, that trim keeps
exactly the chunks under live ranges, and that a request beyond the cap
fails with an out-of-memory status and leaves the heap usable."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_heap_bookkeeping_against_simulated_device(tmp_path):
    exe = tmp_path / "heap_sim"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(ROOT / "tests" / "cpp" / "heap_sim.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
