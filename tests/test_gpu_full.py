"""Full-size parity: the 65,536 x 65,536 product and commutator on 64 qubits.

The collated output (4.3e9 / 2.1e9 terms) is far too large for the oracle,
so parity is bucket-sampled: the output is sorted by key, hence the terms
whose x word 0 has a given top-16-bit prefix form one contiguous range.  On
the CPU the pairs whose product lands in that prefix are found directly (the
prefix of x_l ^ x_r is prefix(x_l) ^ prefix(x_r)), multiplied by the oracle
row by row in the reference's pair order and canonicalized; the device range
is located by binary search and compared bit for bit.  Size-independent
properties cover the rest: the term count, key order across the sampled
boundaries, flags in {0, 1} and no zero coefficients.
"""
import numpy as np
import pytest

import oracle
from helpers import assert_terms_equal
from paulib_amd import PauliSum, commutator_canonical
from paulib_amd import workloads as W
from paulib_amd.host import HostSum

pytestmark = pytest.mark.gpu

N = 65536
PREFIX_BITS = 16


def _top(x0):
    return (np.asarray(x0, np.uint64) >> np.uint64(64 - PREFIX_BITS)).astype(np.int64)


def _rows(h: HostSum, idx) -> HostSum:
    idx = np.asarray(idx, np.int64)
    return HostSum(h.array_size, h.num_qubits, h.x[:, idx], h.z[:, idx], h.flags[idx], h.value[idx])


def _concat(parts, like: HostSum) -> HostSum:
    if not parts:
        return HostSum(like.array_size, like.num_qubits)
    return HostSum(like.array_size, like.num_qubits, np.concatenate([p.x for p in parts], axis=1),
                   np.concatenate([p.z for p in parts], axis=1), np.concatenate([p.flags for p in parts]),
                   np.concatenate([p.value for p in parts]))


def _pick_prefixes(outer_top, inner_top, rng, k, lo=2000, hi=400_000):
    """Prefixes whose bucket holds between lo and hi pairs."""
    counts = np.bincount(inner_top, minlength=1 << PREFIX_BITS)
    picked = []
    for v in rng.permutation(1 << PREFIX_BITS):
        size = int(counts[np.bitwise_xor(outer_top, int(v))].sum())
        if lo <= size <= hi:
            picked.append(int(v))
        if len(picked) == k:
            break
    assert len(picked) == k
    return picked


def _device_range(d: PauliSum, v: int):
    """[first, last) of the terms whose x word 0 has prefix v."""
    n = len(d)

    def first_at_least(p):
        lo, hi = 0, n
        while lo < hi:
            mid = (lo + hi) // 2
            if int(_top(d.word_at(mid))) < p:
                lo = mid + 1
            else:
                hi = mid
        return lo

    return first_at_least(v), first_at_least(v + 1)


def _check_canonical(h: HostSum):
    assert np.all(h.flags <= 1)
    assert np.all(h.value != 0.0)
    keys = np.concatenate([h.x, h.z, h.flags[None, :].astype(np.uint64)]).T
    if len(keys) > 1:
        a, b = keys[:-1], keys[1:]
        neq = a != b
        first = np.argmax(neq, axis=1)
        rows = np.arange(len(a))
        assert neq.any(axis=1).all(), "repeated key"
        assert np.all(a[rows, first] < b[rows, first]), "keys out of order"


def test_product_65536x65536_q64_buckets(ctx):
    A, B = W.product_65536x65536_q64(seed=7)
    dA, dB = PauliSum.from_host(ctx, A), PauliSum.from_host(ctx, B)
    out = dA.multiply_canonical(dB)
    assert len(out) <= N * N
    # expected key collisions among 4.3e9 random 129-bit products: ~2e-7
    assert len(out) >= N * N - 2
    rng = np.random.Generator(np.random.PCG64(11))
    ta, tb = _top(A.x[0]), _top(B.x[0])
    by_prefix = {}
    order = np.argsort(ta, kind="stable")
    bounds = np.searchsorted(ta[order], np.arange((1 << PREFIX_BITS) + 1))
    for v in _pick_prefixes(tb, ta, rng, 3):
        parts = []
        a_cpu = {}
        for i in range(N):  # rhs rows in order; lhs j ascending inside each row
            need = int(tb[i]) ^ v
            js = order[bounds[need]:bounds[need + 1]]
            if js.size == 0:
                continue
            if need not in a_cpu:
                a_cpu[need] = _rows(A, js)
            o = oracle.CpuSum.from_host("oracle", a_cpu[need])
            o.multiply(oracle.CpuSum.from_host("oracle", _rows(B, [i])))
            parts.append(o.to_host())
        want = oracle.CpuSum.from_host("oracle", _concat(parts, A))
        want.canonicalize()
        want = want.to_host()
        lo, hi = _device_range(out, v)
        got = out.range_to_host(lo, hi - lo)
        by_prefix[v] = (lo, hi)
        assert_terms_equal(got, want)
        _check_canonical(got)
    # order across bucket boundaries: the term before each range is smaller
    for v, (lo, hi) in by_prefix.items():
        if lo > 0:
            assert int(_top(out.word_at(lo - 1))) < v


@pytest.mark.parametrize("by_i", [False])
def test_commutator_65536x65536_q64_buckets(ctx, by_i):
    A, B = W.product_65536x65536_q64(seed=17)
    dA, dB = PauliSum.from_host(ctx, A), PauliSum.from_host(ctx, B)
    out = commutator_canonical(dA, dB, multiply_by_i=by_i)
    rng = np.random.Generator(np.random.PCG64(12))
    ta, tb = _top(A.x[0]), _top(B.x[0])
    order = np.argsort(tb, kind="stable")
    bounds = np.searchsorted(tb[order], np.arange((1 << PREFIX_BITS) + 1))
    for v in _pick_prefixes(ta, tb, rng, 2):
        parts = []
        b_cpu = {}
        for i in range(N):  # a rows in order; b index j ascending inside each row
            need = int(ta[i]) ^ v
            js = order[bounds[need]:bounds[need + 1]]
            if js.size == 0:
                continue
            if need not in b_cpu:
                b_cpu[need] = oracle.CpuSum.from_host("oracle", _rows(B, js))
            c = oracle.CpuSum.from_host("oracle", _rows(A, [i])).commutator(b_cpu[need], by_i)
            parts.append(c.to_host())
        want = oracle.CpuSum.from_host("oracle", _concat(parts, A))
        want.canonicalize()
        want = want.to_host()
        lo, hi = _device_range(out, v)
        got = out.range_to_host(lo, hi - lo)
        assert_terms_equal(got, want)
        _check_canonical(got)
