"""Pins the oracle with what the mathematics fixes, independently of the
reference code: dense Pauli matrices for products, phases and Clifford
conjugations; matrix exponentials for evolve; a statevector simulation for
propagate without truncation; and the commutator identity
canon([A,B]) = canon(A*B - B*A).  A dropped phase term, a swapped operand
or a wrong sign in the oracle fails one of these."""
import itertools
import math

import numpy as np
import pytest

from helpers import assert_terms_equal, cpu
from paulib_amd import workloads as W
from paulib_amd.host import HostSum

I2 = np.eye(2, dtype=complex)
PX = np.array([[0, 1], [1, 0]], dtype=complex)
PY = np.array([[0, -1j], [1j, 0]], dtype=complex)
PZ = np.array([[1, 0], [0, -1]], dtype=complex)
MAT = {"I": I2, "X": PX, "Y": PY, "Z": PZ}
PHASE = [1, 1j, -1, -1j]


def term_matrix(h: HostSum, t: int) -> np.ndarray:
    """Dense matrix of term t: i^flags * (P_0 (x) P_1 ...), qubit 0 leftmost."""
    s = h.term_string(t).lstrip("-i")
    m = np.array([[1.0 + 0j]])
    for ch in s:
        m = np.kron(m, MAT[ch])
    return PHASE[int(h.flags[t]) & 3] * m


def sum_matrix(h: HostSum) -> np.ndarray:
    d = 2 ** h.num_qubits
    m = np.zeros((d, d), complex)
    for t in range(len(h)):
        m += h.value[t] * term_matrix(h, t)
    return m


def all_terms(nq):
    return ["".join(p) for p in itertools.product("IXYZ", repeat=nq)]


def test_products_of_all_two_qubit_paulis_with_phases():
    strs = all_terms(2)
    for a, b in itertools.product(strs, strs):
        for fa, fb in [(0, 0), (1, 2), (3, 1)]:
            A = HostSum.from_terms([(a, 1.0)], array_size=1, num_qubits=2)
            B = HostSum.from_terms([(b, 1.0)], array_size=1, num_qubits=2)
            A.flags[:] = fa
            B.flags[:] = fb
            s = cpu("oracle", A)
            s.multiply(cpu("oracle", B))
            out = s.to_host()
            np.testing.assert_allclose(term_matrix(out, 0), term_matrix(A, 0) @ term_matrix(B, 0), atol=0)


def test_multiword_products_against_matrices():
    # 66 qubits -> two words; only a few sites are non-identity so the dense
    # check runs on the sites used.
    rng = np.random.default_rng(1)
    for _ in range(40):
        sites = sorted(int(q) for q in rng.choice(66, size=4, replace=False))
        pa = {int(q): "IXYZ"[rng.integers(0, 4)] for q in sites}
        pb = {int(q): "IXYZ"[rng.integers(0, 4)] for q in sites}
        A = W.observable(66, [pa])
        B = W.observable(66, [pb])
        s = cpu("oracle", A)
        s.multiply(cpu("oracle", B))
        out = s.to_host()
        ma = mb = mo = np.array([[1.0 + 0j]])
        for q in sites:
            ma = np.kron(ma, MAT[pa[q]])
            mb = np.kron(mb, MAT[pb[q]])
            xb = (int(out.x[q // 64, 0]) >> (q % 64)) & 1
            zb = (int(out.z[q // 64, 0]) >> (q % 64)) & 1
            mo = np.kron(mo, MAT["Y" if xb and zb else "X" if xb else "Z" if zb else "I"])
        np.testing.assert_allclose(PHASE[int(out.flags[0])] * mo, ma @ mb, atol=0)


CLIFF = {
    "H": np.array([[1, 1], [1, -1]], dtype=complex) / math.sqrt(2),
    "S": np.array([[1, 0], [0, 1j]], dtype=complex),
}
CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)


def test_clifford_conjugations():
    # apply_H: H P H; apply_S: S P S^dag; apply_CNOT(0,1): CNOT P CNOT
    for p in all_terms(2):
        for f in range(4):
            h = HostSum.from_terms([(p, 1.0)], array_size=1, num_qubits=2)
            h.flags[:] = f
            m = term_matrix(h, 0)
            for gate, U in [("H", np.kron(CLIFF["H"], I2)), ("S", np.kron(I2, CLIFF["S"]))]:
                q = 0 if gate == "H" else 1
                out = cpu("oracle", h)
                getattr(out, f"apply_{gate}")(q)
                np.testing.assert_allclose(term_matrix(out.to_host(), 0), U @ m @ U.conj().T, atol=1e-12)
            out = cpu("oracle", h)
            out.apply_CNOT(0, 1)
            np.testing.assert_allclose(term_matrix(out.to_host(), 0), CNOT @ m @ CNOT, atol=0)


def expm_pauli(P: np.ndarray, theta: float) -> np.ndarray:
    """exp(-i theta/2 P) for a Pauli matrix P (P^2 = 1)."""
    return math.cos(theta / 2) * np.eye(P.shape[0]) - 1j * math.sin(theta / 2) * P


@pytest.mark.parametrize("theta", [0.3, -1.1, 2.9, math.pi / 2, math.pi, -math.pi / 2])
def test_evolve_is_heisenberg_conjugation(theta):
    O = W.random_words_sum(12, 3, 5)
    O.flags[:] = 0
    for p in ["XZY", "ZZI", "IYI"]:
        R = HostSum.from_terms([(p, theta)], array_size=1, num_qubits=3)
        U = expm_pauli(term_matrix(HostSum.from_terms([(p, 1.0)], array_size=1, num_qubits=3), 0), theta)
        s = cpu("oracle", O)
        s.evolve(R)
        np.testing.assert_allclose(sum_matrix(s.to_host()), U.conj().T @ sum_matrix(O) @ U, atol=1e-10)


def statevector_expectation(O: HostSum, prog) -> float:
    """<0|U^dag O U|0> with U the program in application order; the program's
    S is the gate whose Heisenberg action apply_S computes (S P S^dag), i.e.
    S^dag on the state."""
    n = O.num_qubits
    psi = np.zeros(2 ** n, complex)
    psi[0] = 1.0

    def one(U, q):
        ops = [I2] * n
        ops[q] = U
        m = ops[0]
        for o in ops[1:]:
            m = np.kron(m, o)
        return m

    for l in range(len(prog)):
        k = int(prog.kind[l])
        if k == 0:
            psi = one(CLIFF["H"], int(prog.q0[l])) @ psi
        elif k == 1:
            psi = one(CLIFF["S"].conj().T, int(prog.q0[l])) @ psi
        elif k == 2:
            c, t = int(prog.q0[l]), int(prog.q1[l])
            new = psi.copy()
            for b in range(2 ** n):
                if (b >> (n - 1 - c)) & 1:
                    new[b ^ (1 << (n - 1 - t))] = psi[b]
            psi = new
        else:
            P = HostSum(1, n, prog.px[l:l + 1].reshape(1, 1), prog.pz[l:l + 1].reshape(1, 1),
                        np.zeros(1, np.uint8), np.ones(1))
            psi = expm_pauli(term_matrix(P, 0), float(prog.theta[l])) @ psi
    return float(np.real(psi.conj() @ sum_matrix(O) @ psi))


@pytest.mark.parametrize("n,depth,seed", [(4, 3, 1), (6, 2, 2), (7, 3, 3)])
def test_propagation_matches_statevector(n, depth, seed):
    prog = W.brickwork_program(n, depth, seed)
    O = W.observable(n, [{0: "Z"}, {n - 1: "X"}, {1: "Y", 2: "Z"}], coeffs=[1.0, 0.5, -0.25])
    s = cpu("oracle", O)
    s.run_program(prog)  # no truncation
    assert abs(s.zero_state_expectation() - statevector_expectation(O, prog)) <= 1e-10


def test_commutator_identity():
    A = W.random_words_sum(30, 5, 1)
    B = W.random_words_sum(25, 5, 2)
    c = cpu("oracle", A).commutator(cpu("oracle", B))
    c.canonicalize()
    ab = cpu("oracle", A)
    ab.multiply(cpu("oracle", B))
    ba = cpu("oracle", B)
    ba.multiply(cpu("oracle", A))
    neg = ba.to_host()
    neg.value = -neg.value
    ab.add(cpu("oracle", neg))
    ab.canonicalize()
    want = ab.to_host()
    got = c.to_host()
    # Same strings; coefficients agree to rounding (different addition order).
    assert len(got) == len(want)
    assert (got.x == want.x).all() and (got.z == want.z).all() and (got.flags == want.flags).all()
    np.testing.assert_allclose(got.value, want.value, rtol=1e-12)


def test_i_commutator_is_hermitian_real():
    # i[A, B] of Hermitian real-coefficient sums has even phases (flags & 1 == 0)
    A = W.random_sum(40, 6, 0.4, "normal", 3)
    B = W.random_sum(40, 6, 0.4, "normal", 4)
    c = cpu("oracle", A).commutator(cpu("oracle", B), by_i=True)
    c.canonicalize()
    assert (c.to_host().flags & 1 == 0).all()


def test_canonicalize_against_brute_force_map():
    S = W.random_words_sum(500, 3, 8)
    acc = {}
    for t in range(len(S)):
        f = int(S.flags[t])
        v = -S.value[t] if f & 2 else S.value[t]
        k = (int(S.x[0, t]), int(S.z[0, t]), f & 1)
        acc[k] = acc.get(k, 0.0) + v
    want = sorted((k, v) for k, v in acc.items() if v != 0.0)
    s = cpu("oracle", S)
    s.canonicalize()
    got = s.to_host()
    assert [((int(got.x[0, t]), int(got.z[0, t]), int(got.flags[t])), got.value[t])
            for t in range(len(got))] == want
