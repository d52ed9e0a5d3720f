"""Product batches (plh_batch_multiply_canonical) against the oracle.

Every job of a batch must be exactly canonicalize(lhs * rhs) as the oracle
computes it (strings, phases and coefficients bit for bit) and identical to
the single-product path, whatever else shares its batch: empty operands,
one-term sums, heavy key collisions and uneven sizes in one collation.
"""
import numpy as np
import pytest

from helpers import assert_terms_equal, cpu
from paulib_amd import EngineError, PauliSum, batch_multiply_canonical, batch_multiply_canonical_each, job_terms
from paulib_amd import workloads as W

pytestmark = pytest.mark.gpu


def _jobs(seed, count, qubits, lo=1, hi=3000):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        nq = int(rng.choice(qubits))
        n = int(np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        m = int(np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        pid = float(rng.choice([0.5, 0.75]))
        out.append((W.random_sum(n, nq, pid, "uniform", seed * 1000 + 2 * k),
                    W.random_sum(m, nq, pid, "uniform", seed * 1000 + 2 * k + 1)))
    return out


def _oracle(A, B):
    o = cpu("oracle", A)
    o.multiply(cpu("oracle", B))
    o.canonicalize()
    return o.to_host()


def _check(ctx, jobs):
    L = [PauliSum.from_host(ctx, a) for a, _ in jobs]
    R = [PauliSum.from_host(ctx, b) for _, b in jobs]
    out, off = batch_multiply_canonical(L, R)
    assert off[0] == 0 and off[-1] == len(out) and np.all(np.diff(off) >= 0)
    for k, (a, b) in enumerate(jobs):
        got = job_terms(out, off, k)
        assert_terms_equal(got, _oracle(a, b))
        alone = L[k].multiply_canonical(R[k]).to_host()
        assert_terms_equal(got, alone)


def test_batch_one_word(ctx):
    jobs = _jobs(3, 30, [16, 32, 64])
    # edge jobs: an empty operand, one-term sums, all-identity collisions, 4 qubits (heavy collisions)
    jobs.insert(5, (W.random_sum(0, 32, 0.5, "uniform", 1), W.random_sum(40, 32, 0.5, "uniform", 2)))
    jobs.insert(9, (W.random_sum(1, 64, 0.5, "uniform", 3), W.random_sum(1, 64, 0.5, "uniform", 4)))
    jobs.insert(12, (W.random_sum(200, 8, 1.0, "uniform", 5), W.random_sum(300, 8, 1.0, "uniform", 6)))
    jobs.insert(20, (W.random_sum(2000, 4, 0.5, "uniform", 7), W.random_sum(1500, 4, 0.5, "uniform", 8)))
    _check(ctx, jobs)


def test_batch_two_words(ctx):
    _check(ctx, _jobs(4, 10, [100, 128], hi=1500))


def test_batch_single_job_matches_plain(ctx):
    _check(ctx, _jobs(5, 1, [64], lo=2000, hi=4000))


def test_batch_rejects_mixed_array_sizes(ctx):
    a = PauliSum.from_host(ctx, W.random_sum(10, 64, 0.5, "uniform", 1))
    b = PauliSum.from_host(ctx, W.random_sum(10, 128, 0.5, "uniform", 2))
    with pytest.raises(EngineError):
        batch_multiply_canonical([a, a], [a, b])


def _check_each(ctx, jobs, solo_pairs):
    L = [PauliSum.from_host(ctx, a) for a, _ in jobs]
    R = [PauliSum.from_host(ctx, b) for _, b in jobs]
    outs = batch_multiply_canonical_each(L, R, solo_pairs=solo_pairs)
    for k, (a, b) in enumerate(jobs):
        got = outs[k].to_host()
        assert_terms_equal(got, _oracle(a, b))
        assert_terms_equal(got, L[k].multiply_canonical(R[k]).to_host())
    return outs


@pytest.mark.parametrize("solo_pairs", [1, 50_000, 1 << 40])
def test_batch_each_matches_plain(ctx, solo_pairs):
    # mixed array sizes, uneven sizes, an empty operand and heavy collisions;
    # solo_pairs 1: every job alone; 1 << 40: every job batched; else a mix
    jobs = _jobs(6, 16, [16, 64, 100, 128], hi=1200)
    jobs.insert(3, (W.random_sum(0, 32, 0.5, "uniform", 1), W.random_sum(40, 32, 0.5, "uniform", 2)))
    jobs.insert(7, (W.random_sum(900, 6, 1.0, "uniform", 5), W.random_sum(700, 6, 1.0, "uniform", 6)))
    jobs.insert(9, (W.random_sum(60, 200, 0.5, "uniform", 7), W.random_sum(50, 200, 0.5, "uniform", 8)))
    _check_each(ctx, jobs, solo_pairs)


def test_batch_each_reuses_outputs(ctx):
    # the same outputs twice, the second batch with smaller jobs
    jobs = _jobs(9, 6, [32, 64], lo=500, hi=2000)
    outs = _check_each(ctx, jobs, 200_000)
    small = _jobs(10, 6, [32, 64], lo=1, hi=50)
    L = [PauliSum.from_host(ctx, a) for a, _ in small]
    R = [PauliSum.from_host(ctx, b) for _, b in small]
    batch_multiply_canonical_each(L, R, outs, solo_pairs=200_000)
    for k, (a, b) in enumerate(small):
        assert_terms_equal(outs[k].to_host(), _oracle(a, b))


def test_batch_each_rejects_aliasing(ctx):
    a = PauliSum.from_host(ctx, W.random_sum(10, 64, 0.5, "uniform", 1))
    b = PauliSum.from_host(ctx, W.random_sum(10, 64, 0.5, "uniform", 2))
    o = PauliSum(ctx, 1, 64)
    with pytest.raises(EngineError):
        batch_multiply_canonical_each([a, a], [b, b], [o, o])
    with pytest.raises(EngineError):
        batch_multiply_canonical_each([a], [b], [a])
