"""The engine's Clifford tables (built on the host from each gate's images
of X and Z, no GPU call) against the oracle's restatement of the
reference's per-term updates, on every local input: all 4 (single-qubit) or
16 (two-qubit) bit patterns, with the sign flip.  The oracle itself is
pinned by tests/test_oracle_gates.py."""
import ctypes as C

import numpy as np
import pytest

from helpers import cpu
from paulib_amd import _lib
from paulib_amd.gates import GATES, is_two_qubit
from paulib_amd.host import HostSum


def engine_table(gid):
    t, s = C.c_uint64(), C.c_uint32()
    rc = _lib.lib().plh_clifford_table(gid, C.byref(t), C.byref(s))
    assert rc == 0
    return t.value, s.value


@pytest.mark.parametrize("name", GATES)
def test_engine_table_matches_oracle(name):
    gid = GATES.index(name)
    tbl, sgn = engine_table(gid)
    two = is_two_qubit(name)
    # qubit c = 3, t = 9 of one word; one term per local pattern
    nb = 16 if two else 4
    x = np.zeros((1, nb), np.uint64)
    z = np.zeros((1, nb), np.uint64)
    for b in range(nb):
        x[0, b] = ((b & 1) << 3) | (((b >> 2) & 1) << 9)
        z[0, b] = (((b >> 1) & 1) << 3) | (((b >> 3) & 1) << 9)
    h = HostSum(1, 16, x, z, np.zeros(nb, np.uint8), np.ones(nb))
    s = cpu("oracle", h)
    s.apply_gate(name, 3, 9)
    out = s.to_host()
    for b in range(nb):
        ox, oz = int(out.x[0, b]), int(out.z[0, b])
        got = (ox >> 3 & 1) | ((oz >> 3 & 1) << 1) | ((ox >> 9 & 1) << 2) | ((oz >> 9 & 1) << 3)
        assert (tbl >> (4 * b)) & 15 == got, (name, b)
        assert (sgn >> b) & 1 == int(out.flags[b]) >> 1, (name, b)


def test_unknown_gate_is_refused():
    t, s = C.c_uint64(), C.c_uint32()
    assert _lib.lib().plh_clifford_table(21, C.byref(t), C.byref(s)) != 0
    assert _lib.lib().plh_clifford_table(-1, C.byref(t), C.byref(s)) != 0
