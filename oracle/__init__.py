"""CPU checkers for the PauliSum engine -- TEST INFRASTRUCTURE ONLY.

Two libraries with one ABI (prefix orc_ / ref_):
  "oracle"    _build/libpauli_oracle.so, the C restatement (pauli_oracle.c);
  "reference" _ref/libpaulib_ref.so, the reference headers themselves
              instantiated by ref_bridge.cpp (built only where the
              reference tree exists; the built file travels with the repo).
Only tests, the smoke check and the benchmark's CPU-baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "_build" / "libpauli_oracle.so"
REF_LIB = HERE / "_ref" / "libpaulib_ref.so"

_p = C.c_void_p
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_i64 = C.c_int64

_SIG = {
    "sum_new": (_p, [C.c_int, _i64]),
    "sum_free": (None, [_p]),
    "sum_len": (_i64, [_p]),
    "sum_nq": (_i64, [_p]),
    "sum_set_nq": (None, [_p, _i64]),
    "sum_set": (C.c_int, [_p, _i64, _u64p, _u64p, _u8p, _f64p]),
    "sum_get": (None, [_p, _u64p, _u64p, _u8p, _f64p]),
    "multiply": (C.c_int, [_p, _p]),
    "commutator": (_p, [_p, _p, C.c_int]),
    "sort_and_combine": (C.c_int, [_p]),
    "canonicalize": (C.c_int, [_p]),
    "add": (C.c_int, [_p, _p]),
    "zero_state_expectation": (C.c_double, [_p]),
    "truncate_weight": (C.c_int, [_p, _i64]),
    "cutoff": (C.c_int, [_p, C.c_double]),
    "trim": (C.c_int, [_p, _i64]),
    "apply_H": (C.c_int, [_p, _i64]),
    "apply_S": (C.c_int, [_p, _i64]),
    "apply_CNOT": (C.c_int, [_p, _i64, _i64]),
    "evolve": (C.c_int, [_p, _p]),
    "apply_gate": (C.c_int, [_p, C.c_int, _i64, _i64]),
    "apply_rot_axis": (C.c_int, [_p, C.c_int, _i64, C.c_double]),
    "elementwise_multiply": (C.c_int, [_p, _p]),
    "apply_truncation": (C.c_int, [_p, _i64, C.c_double, _i64]),
    "propagate": (C.c_int, [_p, _p, _i64, C.c_double, _i64]),
    "trim_keep": (C.c_int, [_p, _i64, _p]),
    "scale": (C.c_int, [_p, C.c_double]),
    "sub": (C.c_int, [_p, _p]),
    "lookup": (C.c_double, [_p, _u64p, _u64p, C.c_uint8]),
    "trace": (C.c_double, [_p]),
    "frobenius_norm2": (C.c_double, [_p]),
    "apply_truncation_keep": (C.c_int, [_p, _i64, C.c_double, _i64, _p]),
    "run_program": (C.c_int, [_p, _i64, _i32p, _i64p, _i64p, _u64p, _u64p, _f64p, _i64, C.c_double, _i64]),
}

_libs: dict = {}


def build(quiet: bool = True) -> None:
    """Compile the oracle (and the reference bridge where the tree exists)."""
    r = subprocess.run(["make", "-C", str(HERE), "all"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if not quiet:
        print(r.stdout)


def available(kind: str) -> bool:
    return (ORACLE_LIB if kind == "oracle" else REF_LIB).exists()


def _load(kind: str):
    if kind in _libs:
        return _libs[kind]
    path, prefix = (ORACLE_LIB, "orc_") if kind == "oracle" else (REF_LIB, "ref_")
    if not path.exists():
        raise FileNotFoundError(path)
    h = C.CDLL(str(path))
    for name, (res, args) in _SIG.items():
        fn = getattr(h, prefix + name)
        fn.restype = res
        fn.argtypes = args
    _libs[kind] = (h, prefix)
    return _libs[kind]


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class CpuSum:
    """A sum held by one of the CPU checkers."""

    def __init__(self, kind: str, array_size: int, num_qubits: int, handle=None):
        self.kind = kind
        self._h, self._pre = _load(kind)
        self.array_size = array_size
        self.handle = handle if handle is not None else self._fn("sum_new")(array_size, num_qubits)
        if not self.handle:
            raise ValueError("unsupported array size for this checker")

    def _fn(self, name):
        return getattr(self._h, self._pre + name)

    @classmethod
    def from_host(cls, kind: str, host) -> "CpuSum":
        s = cls(kind, host.array_size, host.num_qubits)
        x = np.ascontiguousarray(host.x, np.uint64)
        z = np.ascontiguousarray(host.z, np.uint64)
        f = np.ascontiguousarray(host.flags, np.uint8)
        v = np.ascontiguousarray(host.value, np.float64)
        rc = s._fn("sum_set")(s.handle, len(host), _ptr(x, C.c_uint64), _ptr(z, C.c_uint64),
                              _ptr(f, C.c_uint8), _ptr(v, C.c_double))
        if rc:
            raise RuntimeError(f"sum_set failed ({rc})")
        return s

    def to_host(self):
        from paulib_amd.host import HostSum

        n = int(self._fn("sum_len")(self.handle))
        a = self.array_size
        x = np.empty((a, n), np.uint64)
        z = np.empty((a, n), np.uint64)
        f = np.empty(n, np.uint8)
        v = np.empty(n, np.float64)
        self._fn("sum_get")(self.handle, _ptr(x, C.c_uint64), _ptr(z, C.c_uint64), _ptr(f, C.c_uint8),
                            _ptr(v, C.c_double))
        return HostSum(a, int(self._fn("sum_nq")(self.handle)), x, z, f, v)

    def __len__(self):
        return int(self._fn("sum_len")(self.handle))

    def __del__(self):
        try:
            if self.handle:
                self._fn("sum_free")(self.handle)
                self.handle = None
        except Exception:
            pass

    def _rc(self, rc, name):
        if rc:
            raise RuntimeError(f"{self.kind} {name} failed ({rc})")
        return self

    # operations (in place unless noted)
    def multiply(self, rhs):
        return self._rc(self._fn("multiply")(self.handle, rhs.handle), "multiply")

    def commutator(self, b, by_i=False) -> "CpuSum":
        h = self._fn("commutator")(self.handle, b.handle, 1 if by_i else 0)
        return CpuSum(self.kind, self.array_size, 0, handle=h)

    def sort_and_combine(self):
        return self._rc(self._fn("sort_and_combine")(self.handle), "sort_and_combine")

    def canonicalize(self):
        return self._rc(self._fn("canonicalize")(self.handle), "canonicalize")

    def add(self, rhs):
        return self._rc(self._fn("add")(self.handle, rhs.handle), "add")

    def zero_state_expectation(self) -> float:
        return float(self._fn("zero_state_expectation")(self.handle))

    def truncate_weight(self, w):
        return self._rc(self._fn("truncate_weight")(self.handle, int(w)), "truncate_weight")

    def cutoff(self, eps):
        return self._rc(self._fn("cutoff")(self.handle, float(eps)), "cutoff")

    def trim(self, m):
        return self._rc(self._fn("trim")(self.handle, int(m)), "trim")

    def trim_keep(self, m, keep):
        return self._rc(self._fn("trim_keep")(self.handle, int(m), keep.handle), "trim_keep")

    def scale(self, c):
        return self._rc(self._fn("scale")(self.handle, float(c)), "scale")

    def sub(self, rhs):
        return self._rc(self._fn("sub")(self.handle, rhs.handle), "sub")

    def lookup(self, x, z, flags) -> float:
        xa = np.ascontiguousarray(np.asarray(x, np.uint64).reshape(self.array_size))
        za = np.ascontiguousarray(np.asarray(z, np.uint64).reshape(self.array_size))
        return float(self._fn("lookup")(self.handle, _ptr(xa, C.c_uint64), _ptr(za, C.c_uint64), int(flags) & 3))

    def trace(self) -> float:
        return float(self._fn("trace")(self.handle))

    def frobenius_norm2(self) -> float:
        return float(self._fn("frobenius_norm2")(self.handle))

    def apply_truncation_keep(self, keep, max_weight=None, cutoff=0.0, max_terms=None):
        return self._rc(self._fn("apply_truncation_keep")(
            self.handle, -1 if max_weight is None else max_weight, cutoff,
            -1 if max_terms is None else max_terms, keep.handle), "apply_truncation_keep")

    def apply_H(self, c):
        return self._rc(self._fn("apply_H")(self.handle, int(c)), "apply_H")

    def apply_S(self, c):
        return self._rc(self._fn("apply_S")(self.handle, int(c)), "apply_S")

    def apply_CNOT(self, c, t):
        return self._rc(self._fn("apply_CNOT")(self.handle, int(c), int(t)), "apply_CNOT")

    def apply_gate(self, gate, c, t=0):
        """Any gate of the extended Clifford set by name ('X', 'CZ', 'iSWAP', ...)."""
        from paulib_amd.gates import gate_id

        return self._rc(self._fn("apply_gate")(self.handle, gate_id(gate), int(c), int(t)), f"apply_{gate}")

    def apply_rot(self, axis, c, angle):
        """apply_RX / apply_RY / apply_RZ (axis 'X', 'Y' or 'Z')."""
        a = {"X": 1, "Y": 2, "Z": 3}[axis]
        return self._rc(self._fn("apply_rot_axis")(self.handle, a, int(c), float(angle)), f"apply_R{axis}")

    def elementwise_multiply(self, rhs):
        return self._rc(self._fn("elementwise_multiply")(self.handle, rhs.handle), "elementwise_multiply")

    def evolve(self, rot_host):
        r = CpuSum.from_host(self.kind, rot_host)
        return self._rc(self._fn("evolve")(self.handle, r.handle), "evolve")

    def apply_truncation(self, max_weight=None, cutoff=0.0, max_terms=None):
        return self._rc(self._fn("apply_truncation")(
            self.handle, -1 if max_weight is None else max_weight, cutoff,
            -1 if max_terms is None else max_terms), "apply_truncation")

    def propagate(self, rot_host, max_weight=None, cutoff=0.0, max_terms=None):
        r = CpuSum.from_host(self.kind, rot_host)
        return self._rc(self._fn("propagate")(
            self.handle, r.handle, -1 if max_weight is None else max_weight, cutoff,
            -1 if max_terms is None else max_terms), "propagate")

    def run_program(self, prog, max_weight=None, cutoff=0.0, max_terms=None):
        return self._rc(self._fn("run_program")(
            self.handle, len(prog.kind), _ptr(prog.kind, C.c_int32), _ptr(prog.q0, C.c_int64),
            _ptr(prog.q1, C.c_int64), _ptr(prog.px, C.c_uint64), _ptr(prog.pz, C.c_uint64),
            _ptr(prog.theta, C.c_double), -1 if max_weight is None else max_weight, cutoff,
            -1 if max_terms is None else max_terms), "run_program")
