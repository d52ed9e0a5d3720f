"""The oracle's C restatement against the reference's own headers compiled as
oracle/_ref (ref_bridge.cpp), on seeded random inputs.  Strings and phases
must agree bit for bit; coefficients bit for bit except where a run sums
three or more addends (the reference is built with -ffast-math, so its
in-order loop may be re-associated), where 1e-12 relative applies."""
import math

import numpy as np
import pytest

import oracle
from helpers import assert_terms_equal, cpu
from paulib_amd import workloads as W
from paulib_amd.host import HostSum

pytestmark = pytest.mark.skipif(not oracle.available("reference"),
                                reason="oracle/_ref not built (needs the reference tree)")


def both(host, *ops):
    res = []
    for kind in ("oracle", "reference"):
        s = cpu(kind, host)
        for name, *args in ops:
            getattr(s, name)(*args)
        res.append(s.to_host())
    return res


@pytest.mark.parametrize("as_,nq,n,m", [(1, 64, 13, 7), (2, 128, 13, 7), (1, 64, 3, 2), (8, 500, 10, 10),
                                         (4, 200, 17, 9), (16, 1024, 4, 5), (1, 40, 300, 200)])
def test_multiply_and_canonicalize(as_, nq, n, m):
    A = W.random_words_sum(n, nq, 42, array_size=as_)
    B = W.random_words_sum(m, nq, 99, array_size=as_)
    out = []
    for kind in ("oracle", "reference"):
        s = cpu(kind, A)
        s.multiply(cpu(kind, B))
        raw = s.to_host()
        s.canonicalize()
        out.append((raw, s.to_host()))
    assert_terms_equal(out[0][0], out[1][0])
    assert_terms_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("nq", [3, 5, 8])
def test_collisions_sort_and_combine(nq):
    S = W.random_sum(20000, nq, 0.3, "normal", 4)
    S.flags[:] = np.arange(len(S)) % 4
    for op in ("sort_and_combine", "canonicalize"):
        o, r = both(S, (op,))
        assert_terms_equal(o, r, values="rtol")


@pytest.mark.parametrize("by_i", [False, True])
def test_commutator(by_i):
    A = W.random_words_sum(200, 40, 11)
    B = W.random_words_sum(150, 40, 22)
    res = []
    for kind in ("oracle", "reference"):
        c = cpu(kind, A).commutator(cpu(kind, B), by_i)
        c.canonicalize()
        res.append(c.to_host())
    assert_terms_equal(res[0], res[1])


# 1.7152103698830625: the C library's sincos and cos differ in the last bit of
# cos there (the reference's build calls sincos for its std::cos/std::sin pair)
@pytest.mark.parametrize("theta", [0.0, math.pi, math.pi / 2, -math.pi / 2, 0.4, 2.2, 1.7152103698830625])
def test_evolve(theta):
    S = W.random_words_sum(400, 30, 6)
    rot = HostSum.from_terms([("XYZ" + "I" * 27, theta), ("I" * 10 + "ZZ" + "I" * 18, 1.1)], array_size=1)
    o, r = both(S, ("evolve", rot))
    assert_terms_equal(o, r)


def test_gates_on_random_sum():
    S = W.random_words_sum(1000, 128, 3, array_size=2)
    o, r = both(S, ("apply_H", 5), ("apply_S", 70), ("apply_CNOT", 3, 100), ("apply_CNOT", 64, 2))
    assert_terms_equal(o, r)


def test_truncation():
    S = W.random_sum(5000, 12, 0.5, "normal", 2)
    o, r = both(S, ("apply_truncation", 4, 0.3, None))
    assert_terms_equal(o, r, values="rtol")


@pytest.mark.parametrize("nq,depth,seed,mw,cut", [(20, 5, 1, 8, 1e-4), (32, 4, 2, 6, 1e-3), (48, 3, 3, 8, 1e-5)])
def test_brickwork_programs(nq, depth, seed, mw, cut):
    prog = W.brickwork_program(nq, depth, seed)
    O = W.observable(nq, [{0: "Z"}, {nq // 2: "Z"}])
    res = []
    for kind in ("oracle", "reference"):
        s = cpu(kind, O)
        s.run_program(prog, mw, cut)
        res.append((s.to_host(), s.zero_state_expectation()))
    assert_terms_equal(res[0][0], res[1][0])
    assert abs(res[0][1] - res[1][1]) <= 1e-12 * max(1.0, np.abs(res[0][0].value).sum())


def test_propagate_rotation_list():
    rng = np.random.default_rng(5)
    terms = []
    for _ in range(60):
        s = ["I"] * 16
        for q in rng.choice(16, size=rng.integers(1, 3), replace=False):
            s[q] = "XYZ"[rng.integers(0, 3)]
        terms.append(("".join(s), float(rng.uniform(-math.pi, math.pi))))
    rot = HostSum.from_terms(terms, array_size=1)
    O = W.observable(16, [{3: "Z"}])
    res = []
    for kind in ("oracle", "reference"):
        s = cpu(kind, O)
        s.propagate(rot, 6, 1e-4)
        res.append(s.to_host())
    assert_terms_equal(res[0], res[1])


@pytest.mark.parametrize("n,layers,log2m", [(12, 4, 6), (20, 5, 8), (16, 6, 10), (24, 4, 9)])
def test_propagate_mixed_field_ising_trim_matches_reference(n, layers, log2m):
    """The reference's bench/PropagationBench.cpp circuit (smaller) with only
    max_terms set: the oracle's propagate equals the reference's bit for bit."""
    O, circ = W.mixed_field_ising(n, layers)
    a, b = cpu("oracle", O), cpu("reference", O)
    a.propagate(circ, None, 0.0, 1 << log2m)
    b.propagate(circ, None, 0.0, 1 << log2m)
    assert_terms_equal(a.to_host(), b.to_host())


def _distinct_magnitudes(host, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    host.value[:] = rng.permutation(np.arange(1, len(host) + 1)) * rng.choice([-1.0, 1.0], len(host)) / 7.0


@pytest.mark.parametrize("nq,n,m,nkeep", [(12, 400, 50, 20), (70, 300, 10, 40), (30, 200, 150, 199)])
def test_trim_keep_and_truncation_keep(nq, n, m, nkeep):
    """trim(max_terms, keep) with distinct magnitudes (nth_element's tie
    choice is left open), keep drawn partly from the sum itself."""
    S = W.random_sum(n, nq, 0.5, "normal", n + nq)
    _distinct_magnitudes(S, 3)
    S.flags[:] = np.arange(n) % 4
    pick = np.random.Generator(np.random.PCG64(nkeep)).choice(n, nkeep // 2, replace=False)
    K = W.random_sum(nkeep - len(pick), nq, 0.5, "normal", 77)
    K = HostSum(S.array_size, nq, np.concatenate([S.x[:, pick], K.x], axis=1),
                np.concatenate([S.z[:, pick], K.z], axis=1),
                np.concatenate([(S.flags[pick] + 1) % 4, K.flags]), np.ones(nkeep))
    res = []
    for kind in ("oracle", "reference"):
        s = cpu(kind, S)
        s.trim_keep(m, cpu(kind, K))
        res.append(s.to_host())
    assert_terms_equal(res[0], res[1])
    res = []
    for kind in ("oracle", "reference"):
        s = cpu(kind, S)
        s.apply_truncation_keep(cpu(kind, K), 6, 1e-3, m)
        res.append(s.to_host())
    assert_terms_equal(res[0], res[1])


def test_sub_scale_lookup_trace_norm():
    A = W.random_words_sum(300, 100, 5, array_size=2)
    A.x[:, 10:14] = 0  # identity terms for trace
    A.z[:, 10:14] = 0
    B = W.random_words_sum(200, 100, 6, array_size=2)
    B = HostSum(2, 100, np.concatenate([A.x[:, :50], B.x], axis=1), np.concatenate([A.z[:, :50], B.z], axis=1),
                np.concatenate([A.flags[:50], B.flags]), np.concatenate([A.value[:50] * 0.5, B.value]))
    res = []
    for kind in ("oracle", "reference"):
        s = cpu(kind, A)
        s.sub(cpu(kind, B))
        s.scale(-0.375)
        res.append(s.to_host())
    assert_terms_equal(res[0], res[1])
    o, r = cpu("oracle", A), cpu("reference", A)
    for t in [0, 7, 10, 49, 299]:
        x, z = A.x[:, t], A.z[:, t]
        for f in range(4):
            assert o.lookup(x, z, f) == r.lookup(x, z, f)
    assert o.trace() == r.trace() and o.trace() != 0.0
    o.canonicalize()
    r.canonicalize()
    assert math.isclose(o.frobenius_norm2(), r.frobenius_norm2(), rel_tol=1e-12)
