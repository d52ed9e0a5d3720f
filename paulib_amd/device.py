"""Device-resident PauliSum over the C ABI -- the Python mirror of the
reference's PauliSum<N,double> hot-path API.

Method and function names follow the reference (include/PauliSumCore.h:58-133,
include/PauliSumOperations.h:348-356, include/PauliPropagation.h:26-71); every
call runs the HIP engine (libpaulib_hip.so) on the context's stream.  Errors
from the engine raise EngineError (the reference throws std::runtime_error).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .host import HostSum


def _u64(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _i32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Context:
    """One device and one HIP stream (plh_ctx)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        h = C.c_void_p()
        _lib.call("plh_ctx_create", device, C.c_void_p(stream) if stream else None, C.byref(h))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return _lib.lib().plh_ctx_stream(self._h) or 0

    def synchronize(self):
        _lib.call("plh_ctx_synchronize", self._h)

    def scratch_bytes(self) -> int:
        return int(_lib.lib().plh_ctx_scratch_bytes(self._h))

    def autotune(self, cache_path: str | None = None, force: bool = False) -> dict:
        """Tune the engine for this device (tile caps, batch threshold); see
        plh_ctx_autotune.  Returns the tuned values."""
        path = None if cache_path is None else str(cache_path).encode()
        _lib.call("plh_ctx_autotune", self._h, path, 1 if force else 0)
        return self.tuning()

    def tuning(self) -> dict:
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _lib.call("plh_ctx_tuning", self._h, C.byref(a), C.byref(b), C.byref(c))
        return {"tile_root_max": a.value, "tile_pass_max": b.value, "solo_pairs": c.value}

    def trim(self):
        _lib.call("plh_ctx_trim", self._h)

    def close(self):
        if self._h:
            _lib.lib().plh_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class TruncationOptions:
    """TruncationOptions<double> (include/PauliPropagation.h:26-31); None = unset."""

    max_weight: int | None = None
    coefficient_cutoff: float = 0.0
    max_terms: int | None = None

    def args(self):
        return (-1 if self.max_weight is None else int(self.max_weight),
                float(self.coefficient_cutoff),
                -1 if self.max_terms is None else int(self.max_terms))


class PauliSum:
    """Device-resident PauliSum<64*array_size, double> (plh_sum)."""

    def __init__(self, ctx: Context, array_size: int, num_qubits: int | None = None, capacity: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        nq = 64 * array_size if num_qubits is None else num_qubits
        _lib.call("plh_sum_create", ctx.handle, array_size, nq, capacity, C.byref(h))
        self._h = h

    # -------------------------------------------------------- lifetime
    @classmethod
    def from_host(cls, ctx: Context, host: HostSum) -> "PauliSum":
        s = cls(ctx, host.array_size, host.num_qubits, max(len(host), 16))
        s.upload(host)
        return s

    def upload(self, host: HostSum):
        x = np.ascontiguousarray(host.x)
        z = np.ascontiguousarray(host.z)
        f = np.ascontiguousarray(host.flags)
        v = np.ascontiguousarray(host.value)
        _lib.call("plh_sum_upload", self._h, len(host), _u64(x), _u64(z), _u8(f), _f64(v))
        _lib.call("plh_sum_set_num_qubits", self._h, host.num_qubits)

    def to_host(self) -> HostSum:
        n = len(self)
        a = self.array_size
        x = np.empty((a, n), np.uint64)
        z = np.empty((a, n), np.uint64)
        f = np.empty(n, np.uint8)
        v = np.empty(n, np.float64)
        _lib.call("plh_sum_download", self._h, _u64(x), _u64(z), _u8(f), _f64(v))
        return HostSum(a, self.num_qubits(), x, z, f, v)

    download = to_host

    def range_to_host(self, offset: int, count: int) -> HostSum:
        """Terms [offset, offset + count) (plh_sum_download_range)."""
        a = self.array_size
        x = np.empty((a, count), np.uint64)
        z = np.empty((a, count), np.uint64)
        f = np.empty(count, np.uint8)
        v = np.empty(count, np.float64)
        _lib.call("plh_sum_download_range", self._h, int(offset), int(count), _u64(x), _u64(z), _u8(f), _f64(v))
        return HostSum(a, self.num_qubits(), x, z, f, v)

    def word_at(self, t: int, plane: str = "x", k: int = 0) -> int:
        """One key word of term t (x or z plane, word k)."""
        h = self.range_to_host(t, 1)
        return int((h.x if plane == "x" else h.z)[k, 0])

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().plh_sum_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def __len__(self) -> int:
        return int(_lib.lib().plh_sum_length(self._h))

    get_length = __len__

    def num_qubits(self) -> int:
        return int(_lib.lib().plh_sum_num_qubits(self._h))

    @property
    def array_size(self) -> int:
        return int(_lib.lib().plh_sum_array_size(self._h))

    def device_planes(self):
        """(x, z, flags, value) device addresses and the plane stride."""
        ptrs = [C.c_void_p() for _ in range(4)]
        stride = C.c_int64()
        _lib.call("plh_sum_device_planes", self._h, *[C.byref(p) for p in ptrs], C.byref(stride))
        return tuple(p.value for p in ptrs), stride.value

    def copy(self) -> "PauliSum":
        out = PauliSum(self.ctx, self.array_size, self.num_qubits(), max(len(self), 16))
        _lib.call("plh_sum_copy", self._h, out._h)
        return out

    def _new(self) -> "PauliSum":
        return PauliSum(self.ctx, self.array_size, self.num_qubits())

    # ------------------------------------------------------ bulk ops
    def multiply(self, rhs: "PauliSum"):
        """PauliSum::multiply: self <- self * rhs, uncombined, out[i*n + j]."""
        _lib.call("plh_multiply", self._h, rhs._h)
        return self

    def __imul__(self, rhs):
        if isinstance(rhs, PauliSum):
            return self.multiply(rhs)
        return self.scale(float(rhs))

    def __mul__(self, rhs):
        """A * B (uncombined product) or A * c."""
        out = self.copy()
        out *= rhs
        return out

    def __rmul__(self, c):
        if isinstance(c, PauliSum):
            return NotImplemented
        return self.copy().scale(float(c))

    def __neg__(self):
        return self.copy().scale(-1.0)

    def multiply_canonical(self, rhs: "PauliSum") -> "PauliSum":
        """canonicalize(self * rhs) via the fused product->collation path."""
        out = self._new()
        _lib.call("plh_multiply_canonical", self._h, rhs._h, out._h)
        return out

    def sort_and_combine(self):
        _lib.call("plh_sort_and_combine", self._h)
        return self

    def canonicalize(self):
        _lib.call("plh_canonicalize", self._h)
        return self

    def __iadd__(self, rhs: "PauliSum"):
        _lib.call("plh_add", self._h, rhs._h)
        return self

    def __add__(self, rhs: "PauliSum") -> "PauliSum":
        out = self.copy()
        out += rhs
        return out

    def __isub__(self, rhs: "PauliSum"):
        """operator-=: append -rhs, then sort_and_combine."""
        _lib.call("plh_sub", self._h, rhs._h)
        return self

    def __sub__(self, rhs: "PauliSum") -> "PauliSum":
        out = self.copy()
        out -= rhs
        return out

    def zero_state_expectation(self) -> float:
        out = C.c_double()
        _lib.call("plh_zero_state_expectation", self._h, C.byref(out))
        return out.value

    def truncate_weight(self, max_weight: int):
        _lib.call("plh_truncate_weight", self._h, int(max_weight))
        return self

    def cutoff(self, eps: float):
        _lib.call("plh_cutoff", self._h, float(eps))
        return self

    def trim(self, max_terms: int, keep: "PauliSum | None" = None):
        """PauliSum::trim: keep the max_terms largest |c| in term order (ties:
        lower index); with keep, terms whose (x, z) pattern is in keep are
        never dropped and the rest fill max(0, max_terms - protected) places."""
        if keep is None:
            _lib.call("plh_trim", self._h, int(max_terms))
        else:
            _lib.call("plh_trim_keep", self._h, int(max_terms), keep._h)
        return self

    def scale(self, c: float):
        """PauliSum::scale: every coefficient times c."""
        _lib.call("plh_scale", self._h, float(c))
        return self

    def collate(self):
        """sort_and_combine under its Python name (python/bindings.cpp:170)."""
        return self.sort_and_combine()

    def lookup(self, pauli_string) -> float:
        """Coefficient of pauli_string (a string such as '-XZ', or a tuple
        (x words, z words, flags)); phase relative to it, +-i terms skipped."""
        a = self.array_size
        if isinstance(pauli_string, str):
            xw, zw, f, q = HostSum.parse_string(pauli_string, a)
        else:
            xw, zw, f = pauli_string
        x = np.ascontiguousarray(np.asarray(xw, np.uint64).reshape(a))
        z = np.ascontiguousarray(np.asarray(zw, np.uint64).reshape(a))
        out = C.c_double()
        _lib.call("plh_lookup", self._h, a, _u64(x), _u64(z), int(f) & 3, C.byref(out))
        return out.value

    def trace(self) -> float:
        """2^num_qubits times the sum of identity-term coefficients (real phase)."""
        out = C.c_double()
        _lib.call("plh_trace", self._h, C.byref(out))
        return out.value

    def frobenius_norm2(self) -> float:
        """Sum of squared coefficients (tr(O^dagger O) / 2^n of a canonical sum)."""
        out = C.c_double()
        _lib.call("plh_frobenius_norm2", self._h, C.byref(out))
        return out.value

    def term_weight(self, i: int) -> int:
        """Non-identity sites of term i."""
        self._check_term(i)
        h = self.range_to_host(int(i), 1)
        return int(sum(int(a | b).bit_count() for a, b in zip(h.x[:, 0], h.z[:, 0])))

    def _check_term(self, i: int):
        if not 0 <= int(i) < len(self):
            raise IndexError(f"term index {i} out of range for a sum of {len(self)} terms")

    @property
    def coefficients(self):
        """The coefficients in term order (a host copy)."""
        return self.to_host().value

    def get_pauli_string(self, i: int) -> str:
        self._check_term(i)
        return self.range_to_host(int(i), 1).term_string(0)

    def __getitem__(self, i: int):
        """(Pauli string with its phase prefix, coefficient) of term i."""
        self._check_term(i)
        h = self.range_to_host(int(i), 1)
        return h.term_string(0), float(h.value[0])

    def get_str(self) -> str:
        return self.to_host().get_str()

    __str__ = get_str
    __repr__ = get_str

    def apply_H(self, c: int):
        _lib.call("plh_apply_H", self._h, int(c))
        return self

    def apply_S(self, c: int):
        _lib.call("plh_apply_S", self._h, int(c))
        return self

    def apply_CNOT(self, c: int, t: int):
        _lib.call("plh_apply_CNOT", self._h, int(c), int(t))
        return self

    def apply_gate(self, gate, c: int, t: int = 0):
        """Any gate of include/PauliSumClifford.h by name ('X', 'CZ', 'iSWAP', ...) or id."""
        from .gates import gate_id

        _lib.call("plh_apply_clifford", self._h, gate_id(gate), int(c), int(t))
        return self

    def apply_RX(self, c: int, angle: float):
        _lib.call("plh_apply_RX", self._h, int(c), float(angle))
        return self

    def apply_RY(self, c: int, angle: float):
        _lib.call("plh_apply_RY", self._h, int(c), float(angle))
        return self

    def apply_RZ(self, c: int, angle: float):
        _lib.call("plh_apply_RZ", self._h, int(c), float(angle))
        return self

    def elementwise_multiply(self, rhs: "PauliSum"):
        _lib.call("plh_elementwise_multiply", self._h, rhs._h)
        return self

    def evolve(self, rotations: HostSum):
        """PauliSum::evolve: each rotation term in order; its coefficient is the angle."""
        px = np.ascontiguousarray(rotations.x.T)
        pz = np.ascontiguousarray(rotations.z.T)
        pf = np.ascontiguousarray(rotations.flags)
        th = np.ascontiguousarray(rotations.value)
        _lib.call("plh_evolve", self._h, len(rotations), _u64(px), _u64(pz), _u8(pf), _f64(th))
        return self


def commutator(a: PauliSum, b: PauliSum) -> PauliSum:
    """[A, B] uncombined, anticommuting pairs in (i over a, j over b) order."""
    out = a._new()
    _lib.call("plh_commutator", a._h, b._h, 0, out._h)
    return out


def i_commutator(a: PauliSum, b: PauliSum) -> PauliSum:
    out = a._new()
    _lib.call("plh_commutator", a._h, b._h, 1, out._h)
    return out


def commutator_canonical(a: PauliSum, b: PauliSum, multiply_by_i: bool = False) -> PauliSum:
    """canonicalize([A, B]) (or of i[A, B]) without the uncombined list."""
    out = a._new()
    _lib.call("plh_commutator_canonical", a._h, b._h, 1 if multiply_by_i else 0, out._h)
    return out


def apply_truncation(O: PauliSum, opts: TruncationOptions, keep: PauliSum | None = None):
    """apply_truncation(O, opts[, keep]): canonicalize, weight, cutoff, trim."""
    if keep is None:
        _lib.call("plh_apply_truncation", O._h, *opts.args())
    else:
        _lib.call("plh_apply_truncation_keep", O._h, *opts.args(), keep._h)
    return O


def propagate(O: PauliSum, rotations: HostSum, opts: TruncationOptions | None = None):
    """propagate(O, rotations, opts): gates in reverse, evolve + apply_truncation each."""
    opts = opts or TruncationOptions()
    px = np.ascontiguousarray(rotations.x.T)
    pz = np.ascontiguousarray(rotations.z.T)
    pf = np.ascontiguousarray(rotations.flags)
    th = np.ascontiguousarray(rotations.value)
    _lib.call("plh_propagate", O._h, len(rotations), _u64(px), _u64(pz), _u8(pf), _f64(th), *opts.args())
    return O


def propagate_expectation(O: PauliSum, rotations: HostSum, opts: TruncationOptions | None = None) -> float:
    work = O.copy()
    propagate(work, rotations, opts)
    r = work.zero_state_expectation()
    work.close()
    return r


def batch_multiply_canonical(lhs: list, rhs: list, out: PauliSum | None = None):
    """canonicalize(lhs[k] * rhs[k]) for every k in one collation.

    Returns (out, offsets): out holds the jobs one after another, job k is
    terms offsets[k] .. offsets[k + 1] - 1 (see job_terms)."""
    if len(lhs) != len(rhs) or not lhs:
        raise ValueError("lhs and rhs must be non-empty lists of equal length")
    ctx = lhs[0].ctx
    out = out or PauliSum(ctx, lhs[0].array_size, lhs[0].num_qubits())
    k = len(lhs)
    la = (C.c_void_p * k)(*[a._h.value for a in lhs])
    ra = (C.c_void_p * k)(*[b._h.value for b in rhs])
    off = np.zeros(k + 1, np.int64)
    _lib.call("plh_batch_multiply_canonical", k, la, ra, out._h, _i64(off))
    return out, off


def batch_multiply_canonical_each(lhs: list, rhs: list, outs: list | None = None, solo_pairs: int = 0) -> list:
    """canonicalize(lhs[k] * rhs[k]) into outs[k] for every k: products of at
    least solo_pairs pairs (0: the engine's default) run alone, the rest are
    collated together and split.  Returns the output sums."""
    if len(lhs) != len(rhs) or not lhs:
        raise ValueError("lhs and rhs must be non-empty lists of equal length")
    k = len(lhs)
    if outs is None:
        outs = [PauliSum(a.ctx, a.array_size, a.num_qubits()) for a in lhs]
    if len(outs) != k:
        raise ValueError("one output per job")
    la = (C.c_void_p * k)(*[a._h.value for a in lhs])
    ra = (C.c_void_p * k)(*[b._h.value for b in rhs])
    oa = (C.c_void_p * k)(*[o._h.value for o in outs])
    _lib.call("plh_batch_multiply_canonical_each", k, la, ra, oa, int(solo_pairs))
    return outs


def job_terms(out: PauliSum, offsets, k: int) -> HostSum:
    """Job k's terms of a batched result."""
    return out.range_to_host(int(offsets[k]), int(offsets[k + 1] - offsets[k]))


def run_program(O: PauliSum, prog, opts: TruncationOptions | None = None):
    """Heisenberg run of a Clifford + rotation program (see workloads.Program)."""
    opts = opts or TruncationOptions()
    _lib.call("plh_run_program", O._h, len(prog.kind), _i32(prog.kind), _i64(prog.q0), _i64(prog.q1),
              _u64(prog.px), _u64(prog.pz), _f64(prog.theta), *opts.args())
    return O


def _named_gate(name, two):
    if two:
        def f(self, c: int, t: int):
            _lib.call(f"plh_apply_{name}", self._h, int(c), int(t))
            return self
    else:
        def f(self, c: int):
            _lib.call(f"plh_apply_{name}", self._h, int(c))
            return self
    f.__name__ = f"apply_{name}"
    f.__doc__ = f"PauliSum::apply_{name} (include/PauliSumClifford.h)."
    return f


for _g in ("I", "X", "Y", "Z", "H_YZ", "H_XY", "sqrt_X", "sqrt_Y"):
    setattr(PauliSum, f"apply_{_g}", _named_gate(_g, False))
PauliSum.apply_SX = PauliSum.apply_sqrt_X  # python/bindings.cpp names
PauliSum.apply_SY = PauliSum.apply_sqrt_Y
for _g in ("CZ", "CY", "XC_X", "XC_Z", "XC_Y", "YC_X", "YC_Z", "YC_Y", "SWAP", "iSWAP"):
    setattr(PauliSum, f"apply_{_g}", _named_gate(_g, True))
