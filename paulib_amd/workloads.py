"""Seeded synthetic inputs of the benchmark workloads.

Random sums: each site is the identity with probability p_identity, else
X, Y or Z uniformly; coefficients are N(0, 1) or U[-1, 1] (as 2*k/2^53 - 1,
k uniform on [0, 2^53), exact in FP64); flags are 0.

Workloads (named by what they are):
  product_1000x1000_q32   A*B and [A,B] of 1,000-term sums on 32 qubits,
                          sites I w.p. 0.5, N(0,1) coefficients.
  product_65536x65536_q64 two 65,536-term sums on 64 qubits, sites non-I
                          w.p. 0.25, U[-1,1] coefficients (4.3e9 pairs).
  brickwork_q64_d20       Z_0 + Z_31 through a depth-20 brickwork circuit:
                          per layer a random H/S/none per qubit, brickwork
                          CNOTs, an X/Y/Z rotation per qubit and brickwork ZZ
                          rotations, angles U[0, 2pi); truncation |c| < 1e-4
                          or weight > 8.
  batch_512_products      512 products, terms log-uniform in [64, 65536],
                          qubits in {16, 32, 64, 128}.
  serving_stream          Poisson stream of propagation jobs, 32-128 qubits,
                          depth 5-40, weight-1..3 observables, cutoff
                          log-uniform in [1e-5, 1e-3].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .host import HostSum, words_for


def _pack_sites(sites: np.ndarray, array_size: int):
    """sites (n, nq) in {0=I, 1=X, 2=Y, 3=Z} -> x, z planes (AS, n)."""
    n, nq = sites.shape
    xb = (sites == 1) | (sites == 2)
    zb = (sites == 2) | (sites == 3)
    x = np.zeros((array_size, n), np.uint64)
    z = np.zeros((array_size, n), np.uint64)
    for w in range(array_size):
        lo, hi = 64 * w, min(nq, 64 * w + 64)
        if lo >= hi:
            continue
        weights = (np.uint64(1) << np.arange(hi - lo, dtype=np.uint64))
        x[w] = (xb[:, lo:hi].astype(np.uint64) * weights).sum(axis=1, dtype=np.uint64)
        z[w] = (zb[:, lo:hi].astype(np.uint64) * weights).sum(axis=1, dtype=np.uint64)
    return x, z


def random_sum(n: int, num_qubits: int, p_identity: float, coeff: str, seed: int,
               array_size: int | None = None) -> HostSum:
    rng = np.random.Generator(np.random.PCG64(seed))
    a = array_size or words_for(num_qubits)
    sites = np.zeros((n, num_qubits), np.int8)
    if n and num_qubits:
        nonid = rng.random((n, num_qubits)) >= p_identity
        which = rng.integers(1, 4, size=(n, num_qubits), dtype=np.int8)
        sites = np.where(nonid, which, 0).astype(np.int8)
    x, z = _pack_sites(sites, a)
    if coeff == "normal":
        v = rng.standard_normal(n)
    elif coeff == "uniform":
        k = rng.integers(0, 1 << 53, size=n, dtype=np.int64)
        v = 2.0 * (k.astype(np.float64) / float(1 << 53)) - 1.0
    elif coeff == "ones":
        v = np.ones(n)
    else:
        raise ValueError(coeff)
    return HostSum(a, num_qubits, x, z, np.zeros(n, np.uint8), v.astype(np.float64))


def random_words_sum(n: int, num_qubits: int, seed: int, with_flags: bool = True,
                     array_size: int | None = None) -> HostSum:
    """Uniform random words (padding cleared), random sign/phase flags and
    varied coefficients -- the shape of random_pauli_sum plus
    random_sum_with_coeffs in the reference tests (PauliSumUtils.h:113-154,
    tests/PauliSumTests.cpp:33-39)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = array_size or words_for(num_qubits)
    x = rng.integers(0, 2**64, size=(a, n), dtype=np.uint64)
    z = rng.integers(0, 2**64, size=(a, n), dtype=np.uint64)
    for w in range(a):
        bits = max(0, min(64, num_qubits - 64 * w))
        mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
        x[w] &= mask
        z[w] &= mask
    f = rng.integers(0, 4, size=n, dtype=np.uint8) if with_flags else np.zeros(n, np.uint8)
    v = np.array([(-1.0 if i % 2 else 1.0) * (0.5 + i) for i in range(n)], np.float64)
    return HostSum(a, num_qubits, x, z, f, v)


def product_1000x1000_q32(seed: int = 1):
    a = random_sum(1000, 32, 0.5, "normal", seed)
    b = random_sum(1000, 32, 0.5, "normal", seed + 1)
    return a, b


def product_65536x65536_q64(seed: int = 7, n: int = 65536, m: int | None = None):
    a = random_sum(n, 64, 0.75, "uniform", seed)
    b = random_sum(m if m is not None else n, 64, 0.75, "uniform", seed + 1)
    return a, b


# --------------------------------------------------------------- programs
@dataclass
class Program:
    """Ops in application order: kind 0 = H(q0), 1 = S(q0), 2 = CNOT(q0, q1),
    3 = rotation by theta about the Pauli word px/pz[l*AS : l*AS + AS],
    16 + g = gate id g of the extended set (paulib_amd.gates) on (q0, q1)."""

    array_size: int
    num_qubits: int
    kind: np.ndarray
    q0: np.ndarray
    q1: np.ndarray
    px: np.ndarray
    pz: np.ndarray
    theta: np.ndarray

    def __len__(self):
        return int(self.kind.shape[0])

    def num_rotations(self) -> int:
        return int((self.kind == 3).sum())


class ProgramBuilder:
    def __init__(self, nq: int):
        self.nq = nq
        self.a = words_for(nq)
        self.ops = []

    def clifford(self, kind: int, q0: int, q1: int = 0):
        self.ops.append((kind, q0, q1, [0] * self.a, [0] * self.a, 0.0))

    def gate(self, name, c: int, t: int = 0):
        """Any gate of the extended set by name ('sqrt_X', 'YC_Z', 'iSWAP', ...): kind 16 + id."""
        from .gates import gate_id

        self.clifford(16 + gate_id(name), c, t)

    def rotation(self, paulis: dict, theta: float):
        xw = [0] * self.a
        zw = [0] * self.a
        for q, p in paulis.items():
            w, b = divmod(q, 64)
            if p in "XY":
                xw[w] |= 1 << b
            if p in "ZY":
                zw[w] |= 1 << b
        self.ops.append((3, 0, 0, xw, zw, float(theta)))

    def build(self) -> Program:
        L = len(self.ops)
        return Program(
            self.a, self.nq,
            np.array([o[0] for o in self.ops], np.int32),
            np.array([o[1] for o in self.ops], np.int64),
            np.array([o[2] for o in self.ops], np.int64),
            np.array([w for o in self.ops for w in o[3]], np.uint64).reshape(L * self.a),
            np.array([w for o in self.ops for w in o[4]], np.uint64).reshape(L * self.a),
            np.array([o[5] for o in self.ops], np.float64))


def brickwork_program(num_qubits: int, depth: int, seed: int) -> Program:
    rng = np.random.Generator(np.random.PCG64(seed))
    b = ProgramBuilder(num_qubits)
    for layer in range(depth):
        cl = rng.integers(0, 3, size=num_qubits)  # 0 = H, 1 = S, 2 = none
        for q in range(num_qubits):
            if cl[q] < 2:
                b.clifford(int(cl[q]), q)
        off = layer % 2
        for q in range(off, num_qubits - 1, 2):
            b.clifford(2, q, q + 1)
        axes = rng.integers(0, 3, size=num_qubits)
        th = rng.random(num_qubits) * 2 * np.pi
        for q in range(num_qubits):
            b.rotation({q: "XYZ"[axes[q]]}, th[q])
        th2 = rng.random(num_qubits) * 2 * np.pi
        for q in range(off, num_qubits - 1, 2):
            b.rotation({q: "Z", q + 1: "Z"}, th2[q])
    return b.build()


def mixed_gate_program(num_qubits: int, depth: int, seed: int) -> Program:
    """Brickwork layers whose Cliffords are drawn from the whole gate set
    (single-qubit gates on random qubits, two-qubit gates on neighbour pairs
    in either order), each layer followed by single-qubit and ZZ rotations."""
    from .gates import SINGLE, TWO

    rng = np.random.Generator(np.random.PCG64(seed))
    b = ProgramBuilder(num_qubits)
    for layer in range(depth):
        for q in range(num_qubits):
            if rng.random() < 0.5:
                b.clifford(16 + int(rng.integers(len(SINGLE))), q)
        off = layer % 2
        for q in range(off, num_qubits - 1, 2):
            g = 16 + len(SINGLE) + int(rng.integers(len(TWO)))
            c, t = (q, q + 1) if rng.random() < 0.5 else (q + 1, q)
            b.clifford(g, c, t)
        axes = rng.integers(0, 3, size=num_qubits)
        th = rng.random(num_qubits) * 2 * np.pi
        for q in range(num_qubits):
            b.rotation({q: "XYZ"[axes[q]]}, th[q])
        th2 = rng.random(num_qubits) * 2 * np.pi
        for q in range(off, num_qubits - 1, 2):
            b.rotation({q: "Z", q + 1: "Z"}, th2[q])
    return b.build()


def mixed_field_ising(n: int = 40, layers: int = 10, seed: int = 7):
    """Trotterised mixed-field Ising circuit of the reference's
    bench/PropagationBench.cpp:30-42, as a rotation list (the angle is the
    coefficient): per layer ZZ on (q, q+1), then X on every qubit, then Z on
    every qubit, angles uniform on [-0.4, 0.4) (drawn here with numpy's PCG64,
    not the reference's mt19937), and the observable Z_0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    terms = []
    for _ in range(layers):
        for q in range(n - 1):
            s = ["I"] * n
            s[q] = s[q + 1] = "Z"
            terms.append(("".join(s), float(rng.uniform(-0.4, 0.4))))
        for a in "XZ":
            for q in range(n):
                s = ["I"] * n
                s[q] = a
                terms.append(("".join(s), float(rng.uniform(-0.4, 0.4))))
    circ = HostSum.from_terms(terms, array_size=words_for(n), num_qubits=n)
    O = observable(n, [{0: "Z"}])
    return O, circ


def observable(num_qubits: int, paulis: list[dict], coeffs=None) -> HostSum:
    a = words_for(num_qubits)
    terms = []
    for k, pd in enumerate(paulis):
        s = ["I"] * num_qubits
        for q, p in pd.items():
            s[q] = p
        terms.append(("".join(s), 1.0 if coeffs is None else coeffs[k]))
    return HostSum.from_terms(terms, array_size=a, num_qubits=num_qubits)


def brickwork_q64_d20(seed: int = 11):
    prog = brickwork_program(64, 20, seed)
    O = observable(64, [{0: "Z"}, {31: "Z"}])
    return O, prog, dict(max_weight=8, cutoff=1e-4)


def batch_products(count: int = 512, seed: int = 5, lo: int = 64, hi: int = 65536):
    rng = np.random.Generator(np.random.PCG64(seed))
    qubits = rng.choice([16, 32, 64, 128], size=count)
    ns = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(int)
    ms = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(int)
    jobs = []
    for k in range(count):
        jobs.append((int(qubits[k]), int(ns[k]), int(ms[k]), seed * 100003 + k))
    return jobs


def serving_jobs(count: int, rate_hz: float, seed: int = 9):
    """(arrival_s, observable, program, cutoff) tuples of a Poisson stream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate_hz)
        nq = int(rng.integers(32, 129))
        depth = int(rng.integers(5, 41))
        w = int(rng.integers(1, 4))
        qs = rng.choice(nq, size=w, replace=False)
        O = observable(nq, [{int(q): "XYZ"[int(rng.integers(0, 3))] for q in qs}])
        prog = brickwork_program(nq, depth, seed * 7919 + k)
        cut = float(np.exp(rng.uniform(np.log(1e-5), np.log(1e-3))))
        out.append((t, O, prog, cut))
    return out
