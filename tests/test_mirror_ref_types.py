"""include/paulib_hip.hpp compiles against the reference's own types: the
drop-in takes paulib::PauliSum<N,double> (include/PauliSumCore.h:20-134) and
paulib::PauliString<N> (include/PauliStringCore.h:16-27) as they are.  A
compile-only check (g++ -fsyntax-only) of tests/cpp/mirror_ref_types.cpp with
the reference headers; skipped where the reference tree is absent."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF_INCLUDE = Path("/root/reference/include")


@pytest.mark.skipif(not (REF_INCLUDE / "PauliSum.h").exists() or shutil.which("g++") is None,
                    reason="reference headers or g++ absent")
def test_mirror_header_accepts_reference_types():
    cmd = ["g++", "-std=c++20", "-fsyntax-only", "-fopenmp", "-Wall", "-Wextra", "-Werror",
           f"-I{REF_INCLUDE}", str(ROOT / "tests" / "cpp" / "mirror_ref_types.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
