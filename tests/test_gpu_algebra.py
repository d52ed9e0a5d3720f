"""The algebra and query surface of the device PauliSum against the oracle:
trim(max_terms, keep) (include/PauliSumUtils.h:345-375), apply_truncation
with keep (include/PauliPropagation.h:35-47), operator- / unary minus /
scalar products (include/PauliSumCore.h:191-262), scale, lookup, trace and
frobenius_norm2 (include/PauliSumUtils.h:229-268, :395-410).

Strings, phases and kept-term sets compare bit for bit.  trim's choice
among equal magnitudes is the lower index in both the engine and the
oracle.  lookup, trace and frobenius_norm2 are sums whose order differs
(the engine reduces per block, the oracle in term order; the reference's
frobenius_norm2 is an OpenMP reduction), so they are held to 1e-12
relative to the sum of the absolute addends."""
import math

import numpy as np
import pytest

from helpers import assert_terms_equal, cpu
from paulib_amd import PauliSum, TruncationOptions, apply_truncation
from paulib_amd import workloads as W
from paulib_amd.host import HostSum

pytestmark = pytest.mark.gpu


def S(terms, nq=None, as_=1):
    return HostSum.from_terms(terms, array_size=as_, num_qubits=nq)


def _keep_from(Sx: HostSum, nkeep: int, seed: int) -> HostSum:
    """A keep sum: half its patterns taken from Sx (phases shifted), half random."""
    n = len(Sx)
    pick = np.random.Generator(np.random.PCG64(seed)).choice(n, min(n, nkeep // 2), replace=False)
    R = W.random_sum(nkeep - len(pick), Sx.num_qubits, 0.5, "normal", seed + 1, array_size=Sx.array_size)
    return HostSum(Sx.array_size, Sx.num_qubits, np.concatenate([Sx.x[:, pick], R.x], axis=1),
                   np.concatenate([Sx.z[:, pick], R.z], axis=1),
                   np.concatenate([(Sx.flags[pick] + 1) % 4, R.flags]).astype(np.uint8), np.ones(nkeep))


def test_reference_kats_on_device(ctx):
    # tests/PauliSumTests.cpp:1125-1137, :608-683, :768-817, :1042-1055
    d = PauliSum.from_host(ctx, S([("XI", 0.001), ("ZI", -5.0), ("IZ", 0.3), ("YI", 2.0)]))
    d.trim(2, PauliSum.from_host(ctx, S([("XI", 1.0)])))
    out = d.to_host()
    assert {out.term_string(t): out.value[t] for t in range(len(out))} == {"XI": 0.001, "ZI": -5.0}
    a = PauliSum.from_host(ctx, S([("XX", 1.0)]))
    a -= PauliSum.from_host(ctx, S([("ZZ", 2.0)]))
    assert a.get_str() == "-2.000000 * ZZ\n1.000000 * XX\n"
    a = PauliSum.from_host(ctx, S([("XX", 1.0), ("ZZ", 2.0)]))
    b = PauliSum.from_host(ctx, S([("XX", 3.0)]))
    c = a - b
    assert c.get_str() == "2.000000 * ZZ\n-2.000000 * XX\n"
    assert len(a) == 2 and a.to_host().value[0] == 1.0 and b.to_host().value[0] == 3.0
    z = a - a
    assert len(z) == 2 and list(z.to_host().value) == [0.0, 0.0]
    n = -PauliSum.from_host(ctx, S([("XX", 1.0), ("ZZ", -2.5)]))
    assert n.get_str() == "-1.000000 * XX\n2.500000 * ZZ\n"
    m = PauliSum.from_host(ctx, S([("XX", 1.0), ("ZZ", -2.0)]))
    assert (m * 3.0).get_str() == "3.000000 * XX\n-6.000000 * ZZ\n" == (3.0 * m).get_str()
    m *= 4.0
    assert m.get_str() == "4.000000 * XX\n-8.000000 * ZZ\n"
    s = PauliSum.from_host(ctx, S([("X", 1.0), ("-X", 0.25), ("iY", 2.0)]))
    assert [s.lookup(q) for q in ("X", "-X", "iY", "-iY", "Z")] == [0.75, -0.75, 2.0, -2.0, 0.0]
    s = PauliSum.from_host(ctx, S([("iZ", 3.0), ("Z", 1.0)]))
    assert s.lookup("Z") == 1.0 and s.lookup("iZ") == 3.0
    assert PauliSum(ctx, 1, 2).lookup("XX") == 0.0
    s = PauliSum.from_host(ctx, S([("IIII", 2.5), ("XIZI", 7.0), ("-IIII", 1.0), ("iIIII", 3.0)]))
    assert s.trace() == 16.0 * 1.5
    s = PauliSum.from_host(ctx, S([("XII", 0.5), ("ZZI", -0.2), ("XYZ", 1.5)])).canonicalize()
    assert abs(s.frobenius_norm2() - 2.54) <= 1e-14
    w = PauliSum.from_host(ctx, S([("IIII", 1.0), ("IXYZ", 1.0), ("ZIIZ", 1.0)]))
    assert [w.term_weight(i) for i in range(3)] == [0, 3, 2]
    assert w[1] == ("IXYZ", 1.0) and w.get_pauli_string(2) == "ZIIZ"
    with pytest.raises(IndexError):
        w[3]


@pytest.mark.parametrize("nq,n,nkeep", [(20, 3000, 40), (64, 5000, 1), (100, 2000, 300), (300, 700, 64)])
def test_trim_keep_matches_oracle(ctx, nq, n, nkeep):
    Sx = W.random_sum(n, nq, 0.6, "normal", n + nq)
    Sx.flags[:] = np.arange(n) % 4
    Sx.value[::5] = Sx.value[1]  # ties across the threshold
    K = _keep_from(Sx, nkeep, 7)
    for m in sorted({0, 1, nkeep // 3, n // 3, n - 1, n, n + 5}):
        d = PauliSum.from_host(ctx, Sx)
        d.trim(m, PauliSum.from_host(ctx, K))
        o = cpu("oracle", Sx)
        o.trim_keep(m, cpu("oracle", K))
        assert_terms_equal(d.to_host(), o.to_host())


def test_trim_keep_empty_keep_equals_trim(ctx):
    Sx = W.random_sum(4000, 40, 0.5, "uniform", 3)
    for m in [0, 17, 3999]:
        a = PauliSum.from_host(ctx, Sx)
        a.trim(m, PauliSum(ctx, 1, 40))
        b = PauliSum.from_host(ctx, Sx)
        b.trim(m)
        assert_terms_equal(a.to_host(), b.to_host())


def test_trim_keep_large_against_numpy(ctx):
    n, m, nk = 2_000_000, 300_000, 5000
    rng = np.random.Generator(np.random.PCG64(21))
    x = rng.integers(0, 2**63, size=(1, n), dtype=np.uint64)
    z = rng.integers(0, 2**63, size=(1, n), dtype=np.uint64)
    v = rng.standard_normal(n)
    Sx = HostSum(1, 64, x, z, np.zeros(n, np.uint8), v)
    kidx = rng.choice(n, nk, replace=False)
    K = HostSum(1, 64, x[:, kidx], z[:, kidx], np.zeros(nk, np.uint8), np.ones(nk))
    d = PauliSum.from_host(ctx, Sx)
    d.trim(m, PauliSum.from_host(ctx, K))
    got = d.to_host()
    prot = np.zeros(n, bool)
    prot[kidx] = True
    rest = np.nonzero(~prot)[0]
    top = rest[np.argsort(-np.abs(v[rest]), kind="stable")[: m - nk]]
    keep = np.sort(np.concatenate([kidx, top]))
    assert len(got) == m
    assert np.array_equal(got.x[0], x[0, keep]) and np.array_equal(got.z[0], z[0, keep])
    assert np.array_equal(got.value.view(np.uint64), v[keep].view(np.uint64))


def test_apply_truncation_with_keep(ctx):
    Sx = W.random_sum(6000, 30, 0.6, "normal", 11)
    K = _keep_from(Sx, 100, 2)
    d = PauliSum.from_host(ctx, Sx)
    apply_truncation(d, TruncationOptions(max_weight=12, coefficient_cutoff=1e-3, max_terms=500),
                     PauliSum.from_host(ctx, K))
    o = cpu("oracle", Sx)
    o.apply_truncation_keep(cpu("oracle", K), 12, 1e-3, 500)
    assert_terms_equal(d.to_host(), o.to_host())


@pytest.mark.parametrize("as_,nq", [(1, 50), (2, 100), (4, 200)])
def test_sub_scale_and_queries_match_oracle(ctx, as_, nq):
    A = W.random_words_sum(3000, nq, 5, array_size=as_)
    A.x[:, 10:20] = 0
    A.z[:, 10:20] = 0  # identity terms with every phase
    B0 = W.random_words_sum(2000, nq, 6, array_size=as_)
    B = HostSum(as_, nq, np.concatenate([A.x[:, :500], B0.x], axis=1), np.concatenate([A.z[:, :500], B0.z], axis=1),
                np.concatenate([A.flags[:500], B0.flags]), np.concatenate([A.value[:500] * 0.5, B0.value]))
    d = PauliSum.from_host(ctx, A)
    d -= PauliSum.from_host(ctx, B)
    d.scale(-0.375)
    o = cpu("oracle", A)
    o.sub(cpu("oracle", B))
    o.scale(-0.375)
    assert_terms_equal(d.to_host(), o.to_host())
    # self-subtraction through one handle
    e = PauliSum.from_host(ctx, A)
    e -= e
    oe = cpu("oracle", A)
    oe.sub(cpu("oracle", A))
    assert_terms_equal(e.to_host(), oe.to_host())
    # queries on the raw (uncombined, all-phase) sum
    dA, oA = PauliSum.from_host(ctx, A), cpu("oracle", A)
    tol = 1e-12 * float(np.abs(A.value).sum())
    for t in [0, 10, 15, 499, 2999]:
        for f in range(4):
            q = (A.x[:, t], A.z[:, t], f)
            assert abs(dA.lookup(q) - oA.lookup(*q)) <= tol
    assert math.isclose(dA.trace(), oA.trace(), rel_tol=1e-12, abs_tol=math.ldexp(tol, nq))
    assert abs(dA.frobenius_norm2() - oA.frobenius_norm2()) <= 1e-12 * oA.frobenius_norm2()
    assert PauliSum(ctx, as_, nq).trace() == 0.0 and PauliSum(ctx, as_, nq).frobenius_norm2() == 0.0


def test_frobenius_norm2_is_trace_of_square(ctx):
    """For a canonical Hermitian O (real coefficients, flags 0):
    tr(O O) / 2^n = sum of c^2, with O O formed by the product engine."""
    O = W.random_sum(300, 12, 0.5, "normal", 4)
    d = PauliSum.from_host(ctx, O).canonicalize()
    sq = d.multiply_canonical(d)
    want = d.frobenius_norm2()
    assert math.isclose(math.ldexp(sq.trace(), -12), want, rel_tol=1e-12)


def test_operators_compose(ctx):
    A = W.random_sum(500, 20, 0.5, "normal", 1)
    B = W.random_sum(400, 20, 0.5, "normal", 2)
    a, b = PauliSum.from_host(ctx, A), PauliSum.from_host(ctx, B)
    c = (a + b) - b  # = canonical-order A with B's keys at zero
    o = cpu("oracle", A)
    o.add(cpu("oracle", B))
    o.sub(cpu("oracle", B))
    assert_terms_equal(c.to_host(), o.to_host())
    p = a * b  # uncombined product, a unchanged
    op = cpu("oracle", A)
    op.multiply(cpu("oracle", B))
    assert_terms_equal(p.to_host(), op.to_host())
    assert len(a) == 500
