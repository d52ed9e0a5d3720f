"""Parity of the extended gate set, apply_RX/RY/RZ and elementwise_multiply
(HIP engine through the C ABI) with the oracle, bit for bit: strings, phases
and coefficients.  A rotation's sort_and_combine gives each key at most the
old term and one new term, and both sides add a run in original-index order,
so its coefficients are exact too."""
import math

import numpy as np
import pytest

from helpers import assert_terms_equal, cpu
from paulib_amd import PauliSum, _lib
from paulib_amd import workloads as W
from paulib_amd.gates import GATES, SINGLE, TWO
from paulib_amd.host import HostSum

pytestmark = pytest.mark.gpu


def dev(ctx, host):
    return PauliSum.from_host(ctx, host)


@pytest.mark.parametrize("as_,nq,n", [(1, 64, 1000), (2, 128, 777), (1, 64, 70001), (4, 200, 300)])
def test_every_gate_matches_oracle(ctx, as_, nq, n):
    S = W.random_words_sum(n, nq, 5 + as_, array_size=as_)
    d = dev(ctx, S)
    o = cpu("oracle", S)
    last = nq - 1
    pairs = [(3, 9), (9, 3), (0, last), (last, 0), (62, 63)]
    if as_ > 1:
        pairs += [(5, 70), (70, 5), (63, 64)]
    for name in GATES:
        qs = pairs if name in TWO else [(0, 0), (17, 0), (last, 0)]
        for c, t in qs:
            print(f"gate {name} c={c} t={t} as={as_} n={n}", flush=True)
            d.apply_gate(name, c, t)
            o.apply_gate(name, c, t)
            assert_terms_equal(d.to_host(), o.to_host())


def test_named_entry_points(ctx):
    S = W.random_words_sum(500, 64, 3)
    d = dev(ctx, S)
    o = cpu("oracle", S)
    for name in SINGLE:
        if name in ("H", "S"):
            continue
        getattr(d, f"apply_{name}")(11)
        o.apply_gate(name, 11)
    for name in TWO:
        if name == "CNOT":
            continue
        getattr(d, f"apply_{name}")(11, 40)
        o.apply_gate(name, 11, 40)
    assert_terms_equal(d.to_host(), o.to_host())


def test_gate_errors_and_empty_sum(ctx):
    d = dev(ctx, W.random_words_sum(8, 64, 1))
    with pytest.raises(_lib.EngineError):
        d.apply_CZ(4, 4)
    with pytest.raises(_lib.EngineError):
        d.apply_X(64)
    with pytest.raises(_lib.EngineError):
        d.apply_gate(21, 0, 1)
    empty = HostSum(1, 64, np.zeros((1, 0), np.uint64), np.zeros((1, 0), np.uint64),
                    np.zeros(0, np.uint8), np.zeros(0))
    e = dev(ctx, empty)
    e.apply_iSWAP(0, 1)
    e.apply_RY(3, 0.5)
    e.elementwise_multiply(e)
    assert len(e.to_host()) == 0


# 3.430420739766125 = 2 * 1.7152103698830625, where sincos and a separate cos
# give different last bits of cos(angle / 2)
@pytest.mark.parametrize("angle", [0.0, 0.37, -1.9, math.pi, 2 * math.pi, 5.5, 3.430420739766125])
@pytest.mark.parametrize("axis", ["X", "Y", "Z"])
def test_half_angle_rotations(ctx, axis, angle):
    for as_, nq, q in [(1, 64, 5), (1, 64, 63), (2, 128, 100)]:
        S = W.random_words_sum(3000, nq, 17 + q, array_size=as_)
        d = dev(ctx, S)
        o = cpu("oracle", S)
        print(f"R{axis} angle={angle} q={q} as={as_}", flush=True)
        getattr(d, f"apply_R{axis}")(q, angle)
        o.apply_rot(axis, q, angle)
        assert_terms_equal(d.to_host(), o.to_host())


def test_rotation_sequence_with_collisions(ctx):
    """Repeated rotations on a few qubits of a 6-qubit sum: keys collide and
    combine, zeros and +/- pairs stay separate (sort_and_combine)."""
    S = W.random_words_sum(200, 6, 8)
    d = dev(ctx, S)
    o = cpu("oracle", S)
    rng = np.random.default_rng(4)
    for step in range(12):
        axis = "XYZ"[int(rng.integers(3))]
        q = int(rng.integers(6))
        angle = float(rng.uniform(-4, 4)) if step % 4 else 0.0
        getattr(d, f"apply_R{axis}")(q, angle)
        o.apply_rot(axis, q, angle)
        assert_terms_equal(d.to_host(), o.to_host())


@pytest.mark.parametrize("as_,nq,n", [(1, 64, 999), (2, 128, 4097), (16, 1000, 50)])
def test_elementwise_multiply(ctx, as_, nq, n):
    A = W.random_words_sum(n, nq, 31, array_size=as_)
    B = W.random_words_sum(n, nq, 32, array_size=as_)
    d = dev(ctx, A)
    d.elementwise_multiply(dev(ctx, B))
    o = cpu("oracle", A)
    o.elementwise_multiply(cpu("oracle", B))
    assert_terms_equal(d.to_host(), o.to_host())
    # rhs is the sum itself
    d2 = dev(ctx, A)
    d2.elementwise_multiply(d2)
    o2 = cpu("oracle", A)
    o2.elementwise_multiply(cpu("oracle", A))
    assert_terms_equal(d2.to_host(), o2.to_host())


def test_elementwise_length_mismatch(ctx):
    d = dev(ctx, W.random_words_sum(10, 64, 1))
    with pytest.raises(_lib.EngineError):
        d.elementwise_multiply(dev(ctx, W.random_words_sum(9, 64, 2)))


@pytest.mark.parametrize("nq,depth,seed", [(12, 5, 1), (64, 3, 2), (100, 2, 3), (300, 2, 4)])
def test_run_program_with_gate_set(ctx, nq, depth, seed):
    """run_program with the whole gate set (kinds 16 + id) and truncated
    rotations against the oracle's run_program.  A layer's gates go through
    the fused Clifford-run kernel in runs of at most 64 ops; 300 qubits take
    several runs per layer at five words per plane."""
    from paulib_amd import TruncationOptions, run_program

    prog = W.mixed_gate_program(nq, depth, seed)
    O = W.observable(nq, [{0: "Z"}, {nq // 2: "X", nq // 2 + 1: "Y"}])
    d = dev(ctx, O)
    run_program(d, prog, TruncationOptions(max_weight=6, coefficient_cutoff=1e-4))
    o = cpu("oracle", O)
    o.run_program(prog, 6, 1e-4)
    assert_terms_equal(d.to_host(), o.to_host())
    e_got, e_want = d.zero_state_expectation(), o.zero_state_expectation()
    assert abs(e_got - e_want) <= 1e-12 * max(1.0, np.abs(o.to_host().value).sum())
