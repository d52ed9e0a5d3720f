"""Builds libpaulib_hip.so (the HIP engine) in-tree with hipcc for gfx950.

Each .hip translation unit under csrc/ is compiled to an object in
paulib_amd/_build/ (in parallel), then linked into
paulib_amd/_build/libpaulib_hip.so.  Every call recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "_build"
LIB = OUT / "libpaulib_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _compile(src: Path) -> Path:
    obj = OUT / (src.stem + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-Wl,--no-undefined", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", file=sys.stderr)
    return LIB


ROOT = HERE.parent
CPP_TEST = ROOT / "tests" / "cpp" / "mirror_test.cpp"
CPP_TEST_BIN = ROOT / "tests" / "cpp" / "_build" / "mirror_test"


def build_cpp_tests() -> Path:
    """The C++ drop-in test (include/paulib_hip.hpp against the oracle)."""
    CPP_TEST_BIN.parent.mkdir(parents=True, exist_ok=True)
    oracle_dir = ROOT / "oracle" / "_build"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", str(CPP_TEST), "-o", str(CPP_TEST_BIN),
           f"-L{OUT}", "-lpaulib_hip", f"-L{oracle_dir}", "-lpauli_oracle",
           f"-Wl,-rpath,{OUT}", f"-Wl,-rpath,{oracle_dir}",
           "-Wl,-rpath,$ORIGIN/../../../paulib_amd/_build", "-Wl,-rpath,$ORIGIN/../../../oracle/_build"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ mirror test build failed:\n{r.stdout}\n{r.stderr}")
    return CPP_TEST_BIN


if __name__ == "__main__":
    build()
