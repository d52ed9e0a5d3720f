"""Collated products whose root pass is built from lhs digit groups.

Products of at least 2^22 pairs, with 2^11..2^22 lhs terms and 512 rhs rows
or more, take their root bucket contents from the operands' digit groups
(a product key word is the XOR of its operands' words) instead of a
scatter.  Each case is compared with the oracle's multiply + canonicalize
(include/PauliSumOperations.h:255-306, include/PauliSumUtils.h:314-328)
bit for bit.  The cases place the root digit and the 32 bits carried after
it in every way the key layout allows: both inside x word 0, the digit
across the x/z word boundary, the carried bits running into the flags, and
the digit itself reaching the flags (which keeps the scatter).
"""
import numpy as np
import pytest

from helpers import assert_terms_equal, cpu
from paulib_amd import PauliSum
from paulib_amd import workloads as W

pytestmark = pytest.mark.gpu


def oracle_canonical_product(A, B):
    s = cpu("oracle", A)
    s.multiply(cpu("oracle", B))
    s.canonicalize()
    return s.to_host()


def check(ctx, A, B):
    print(f"group root: {A.x.shape[1]} x {B.x.shape[1]}, AS {A.array_size}, {A.num_qubits} qubits", flush=True)
    got = PauliSum.from_host(ctx, A).multiply_canonical(PauliSum.from_host(ctx, B)).to_host()
    assert_terms_equal(got, oracle_canonical_product(A, B))


def masked(h, xmask, zmask):
    """The sum with its x and z words ANDed with per-word masks."""
    h.x &= np.array(xmask, np.uint64)[:, None]
    h.z &= np.array(zmask, np.uint64)[:, None]
    return h


def test_group_root_config_distribution_64q(ctx):
    """Sites non-I w.p. 0.25 on 64 qubits: digit and carried bits in x word 0."""
    A, B = W.product_65536x65536_q64(seed=3, n=4096, m=1024)
    check(ctx, A, B)


def test_group_root_two_words(ctx):
    """100 qubits in two words per plane, uniform random words and flags."""
    A = W.random_words_sum(2048, 100, 5, array_size=2)
    B = W.random_words_sum(2048, 100, 6, array_size=2)
    check(ctx, A, B)


def test_group_root_ragged_sizes(ctx):
    """Lhs count not a power of two, rhs rows not a multiple of a row block."""
    A = W.random_words_sum(3001, 50, 7)
    B = W.random_words_sum(1499, 50, 8)
    check(ctx, A, B)


def test_group_root_digit_across_words(ctx):
    """X/Y only on qubits 0..3: the root digit starts 4 bits before z word 0."""
    A = masked(W.random_words_sum(2048, 64, 9), [0xF], [0xFFFFFFFFFFFFFFFF])
    B = masked(W.random_words_sum(2048, 64, 10), [0xF], [0xFFFFFFFFFFFFFFFF])
    check(ctx, A, B)


def test_group_root_carried_bits_reach_flags(ctx):
    """Only Z on qubits 0..20: the 32 carried bits run into the phase flags,
    and 2^21 keys collect long runs of addends."""
    A = masked(W.random_words_sum(2048, 64, 11), [0], [(1 << 21) - 1])
    B = masked(W.random_words_sum(2048, 64, 12), [0], [(1 << 21) - 1])
    check(ctx, A, B)


def test_digit_in_flags_keeps_scatter(ctx):
    """Only Z on qubits 0..3: the root digit reaches the flags, so the scatter
    builds the root (runs of ~2.6e5 addends)."""
    A = masked(W.random_words_sum(2048, 64, 13), [0], [0xF])
    B = masked(W.random_words_sum(2048, 64, 14), [0], [0xF])
    check(ctx, A, B)


def test_group_root_few_qubits_heavy_collisions(ctx):
    """8 qubits: 2^16 keys, every key summed over ~64 addends in order."""
    A = W.random_sum(4096, 8, 0.3, "normal", 21)
    B = W.random_sum(1024, 8, 0.3, "normal", 22)
    A.flags[:] = np.arange(4096) % 4
    check(ctx, A, B)
