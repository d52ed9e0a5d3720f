"""The oracle's propagation of the depth-20 brickwork configuration (Z_0 +
Z_31 on 64 qubits, weight <= 8, |c| >= 1e-4) against the fixture written
from the reference's own headers (tests/golden/make_c2_fixture.py): the
same terms bit for bit (planes and coefficient digests) and the same
expectation.  Runs the oracle's single-threaded C restatement (~90 s)."""
import json
from pathlib import Path

from helpers import cpu, digests, terms_of
from paulib_amd import workloads as W

GOLDEN = Path(__file__).resolve().parent / "golden" / "c2_brickwork_q64_d20.json"


def test_oracle_reproduces_c2_fixture():
    doc = json.loads(GOLDEN.read_text())
    O, prog, tr = W.brickwork_q64_d20()
    s = cpu("oracle", O)
    s.run_program(prog, tr["max_weight"], tr["cutoff"])
    e = s.zero_state_expectation()
    s.canonicalize()
    h = s.to_host()
    assert len(h) == doc["terms"]
    assert terms_of(h, range(32)) == doc["head"]
    assert terms_of(h, range(len(h) - 32, len(h))) == doc["tail"]
    assert digests(h) == {"planes_sha256": doc["planes_sha256"], "values_sha256": doc["values_sha256"]}
    assert float(e).hex() == doc["expectation_hex"]
