"""trim (PauliSum::trim, include/PauliSumUtils.h:330-343) and max_terms in
apply_truncation / propagate / run_program, against the oracle.

trim keeps the max_terms terms of largest |c| in term order.  The reference
picks among equal magnitudes with std::nth_element (library-defined); the
engine and the oracle both keep the lower index, so ties compare exactly
too (that choice itself is not pinned to the reference: see the oracle's
header).  Large sums are checked against numpy's stable top-k.
"""
import numpy as np
import pytest

from helpers import assert_terms_equal, cpu
from paulib_amd import PauliSum, TruncationOptions, run_program
from paulib_amd import workloads as W
from paulib_amd.host import HostSum

pytestmark = pytest.mark.gpu


def _trim_oracle(S, m):
    o = cpu("oracle", S)
    o.trim(m)
    return o.to_host()


@pytest.mark.parametrize("nq,n", [(20, 3000), (64, 1), (100, 777), (300, 500)])
def test_trim_matches_oracle(ctx, nq, n):
    S = W.random_sum(n, nq, 0.6, "normal", n)
    S.flags[:] = np.arange(len(S)) % 4
    for m in sorted({0, 1, n // 3, n - 1, n, n + 5}):
        if m < 0:
            continue
        d = PauliSum.from_host(ctx, S)
        d.trim(m)
        assert_terms_equal(d.to_host(), _trim_oracle(S, m))


def test_trim_ties_keep_lower_index(ctx):
    # magnitudes from a set of four: long runs of equal |c| across the threshold
    S = W.random_sum(20000, 40, 0.5, "uniform", 3)
    rng = np.random.Generator(np.random.PCG64(5))
    S.value[:] = rng.choice([1.0, -1.0, 0.5, -0.5, 0.25, -0.25, 2.0, -2.0], size=len(S))
    for m in [1, 2500, 7777, 10000, 19999]:
        d = PauliSum.from_host(ctx, S)
        d.trim(m)
        assert_terms_equal(d.to_host(), _trim_oracle(S, m))


def test_trim_large_against_numpy(ctx):
    n, m = 3_000_000, 1_234_567
    rng = np.random.Generator(np.random.PCG64(9))
    x = rng.integers(0, 2**63, size=(1, n), dtype=np.uint64)
    z = rng.integers(0, 2**63, size=(1, n), dtype=np.uint64)
    f = rng.integers(0, 2, size=n, dtype=np.uint8)
    v = rng.standard_normal(n)
    v[::7] = v[3]  # a run of equal magnitudes
    S = HostSum(1, 64, x, z, f, v)
    d = PauliSum.from_host(ctx, S)
    d.trim(m)
    got = d.to_host()
    # stable order by descending |v|: ties keep the lower index; then back to term order
    keep = np.sort(np.argsort(-np.abs(v), kind="stable")[:m])
    assert len(got) == m
    assert np.array_equal(got.x[0], x[0, keep]) and np.array_equal(got.z[0], z[0, keep])
    assert np.array_equal(got.flags, f[keep])
    assert np.array_equal(got.value.view(np.uint64), v[keep].view(np.uint64))


def test_apply_truncation_and_programs_with_max_terms(ctx):
    from paulib_amd import apply_truncation

    S = W.random_sum(5000, 30, 0.6, "normal", 11)
    d = PauliSum.from_host(ctx, S)
    apply_truncation(d, TruncationOptions(max_weight=12, coefficient_cutoff=1e-3, max_terms=700))
    o = cpu("oracle", S)
    o.apply_truncation(12, 1e-3, 700)
    assert_terms_equal(d.to_host(), o.to_host())
    for nq, depth, seed, mt in [(16, 4, 1, 40), (24, 6, 3, 300), (40, 5, 4, 1000)]:
        prog = W.brickwork_program(nq, depth, seed)
        O = W.observable(nq, [{0: "Z"}, {nq // 2: "Z"}])
        d = PauliSum.from_host(ctx, O)
        run_program(d, prog, TruncationOptions(max_weight=8, coefficient_cutoff=1e-5, max_terms=mt))
        o = cpu("oracle", O)
        o.run_program(prog, 8, 1e-5, mt)
        assert_terms_equal(d.to_host(), o.to_host())


@pytest.mark.parametrize("n,layers,log2m", [(12, 4, 6), (20, 5, 8), (16, 6, 10)])
def test_propagate_mixed_field_ising_with_max_terms(ctx, n, layers, log2m):
    """propagate with only max_terms set (the reference's
    bench/PropagationBench.cpp workload, smaller), against the oracle."""
    from paulib_amd import propagate

    O, circ = W.mixed_field_ising(n, layers)
    d = PauliSum.from_host(ctx, O)
    propagate(d, circ, TruncationOptions(max_terms=1 << log2m))
    o = cpu("oracle", O)
    o.propagate(circ, None, 0.0, 1 << log2m)
    assert_terms_equal(d.to_host(), o.to_host())
    e_got, e_want = d.zero_state_expectation(), o.zero_state_expectation()
    assert abs(e_got - e_want) <= 1e-12 * max(1.0, np.abs(o.to_host().value).sum())
