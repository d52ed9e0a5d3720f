"""The benchmark configurations at their real shapes, HIP path against the
reference's results.

- canonicalize([A, B]) of the 1,000 x 1,000 product on 32 qubits (sites I
  w.p. 1/2, N(0,1) coefficients), against the oracle (commutator_impl,
  include/PauliSumOperations.h:315-356, then canonicalize,
  include/PauliSumUtils.h:314-328): strings, phases and coefficients bit
  for bit (every coefficient is a single product (2a)*b or a two-addend sum).
- The depth-20 brickwork propagation of Z_0 + Z_31 on 64 qubits (weight <= 8,
  |c| >= 1e-4; propagate, include/PauliPropagation.h:53-71, with the Clifford
  steps of include/PauliSumClifford.h), against the fixture written from the
  reference's own headers (tests/golden/make_c2_fixture.py): term count and
  SHA-256 of the canonical planes and of the coefficient bits, bit for bit;
  the expectation to 1e-12 of the sum of |coefficients| (the GPU reduces in a
  tree, the reference in term order).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from helpers import assert_terms_equal, cpu, digests, terms_of
from paulib_amd import PauliSum, TruncationOptions, commutator_canonical, run_program
from paulib_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "c2_brickwork_q64_d20.json"


@pytest.mark.parametrize("by_i", [False, True])
def test_commutator_canonical_config_1000x1000_q32(ctx, by_i):
    A, B = W.product_1000x1000_q32(seed=1)
    c = cpu("oracle", A).commutator(cpu("oracle", B), by_i)
    c.canonicalize()
    want = c.to_host()
    got = commutator_canonical(PauliSum.from_host(ctx, A), PauliSum.from_host(ctx, B), by_i).to_host()
    assert len(want) > 400000  # about half of the 1e6 pairs anticommute
    assert_terms_equal(got, want)


def test_propagation_config_brickwork_q64_d20(ctx):
    doc = json.loads(GOLDEN.read_text())
    O, prog, tr = W.brickwork_q64_d20()
    d = PauliSum.from_host(ctx, O)
    run_program(d, prog, TruncationOptions(max_weight=tr["max_weight"], coefficient_cutoff=tr["cutoff"]))
    e = d.zero_state_expectation()
    d.canonicalize()
    got = d.to_host()
    assert len(got) == doc["terms"]
    dg = digests(got)
    assert terms_of(got, range(min(32, len(got)))) == doc["head"]
    assert dg["planes_sha256"] == doc["planes_sha256"]
    assert dg["values_sha256"] == doc["values_sha256"]
    assert abs(e - doc["expectation"]) <= 1e-12 * max(1.0, doc["abs_value_sum"])
