"""The engine library loads on a CPU-only host and exports every entry point
that include/paulib_hip.h declares (no compute call is made here)."""
import ctypes as C
import re
from pathlib import Path

from paulib_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "paulib_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(plh_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_the_expected_surface():
    syms = declared_symbols()
    for must in ["plh_launch_mul_soa", "plh_multiply", "plh_multiply_canonical", "plh_commutator",
                 "plh_canonicalize", "plh_sort_and_combine", "plh_evolve", "plh_propagate",
                 "plh_apply_H", "plh_apply_S", "plh_apply_CNOT", "plh_zero_state_expectation",
                 "plh_apply_clifford", "plh_apply_iSWAP", "plh_apply_YC_Y", "plh_apply_sqrt_X",
                 "plh_apply_RX", "plh_apply_RY", "plh_apply_RZ", "plh_elementwise_multiply"]:
        assert must in syms


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_error_text_without_gpu_call():
    lib = _lib.lib()
    assert lib.plh_version() == 1
    assert b"gfx950" in lib.plh_build_info()
