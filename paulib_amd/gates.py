"""The extended Clifford gate set of include/PauliSumClifford.h by id.

The ids are the ones the engine's plh_apply_clifford takes; the named C entry
points (plh_apply_X, plh_apply_CZ, ...) are thin wrappers over it.
"""
SINGLE = ("I", "X", "Y", "Z", "H", "S", "H_YZ", "H_XY", "sqrt_X", "sqrt_Y")
TWO = ("CNOT", "CZ", "CY", "XC_X", "XC_Z", "XC_Y", "YC_X", "YC_Z", "YC_Y", "SWAP", "iSWAP")
GATES = SINGLE + TWO


def gate_id(name) -> int:
    if isinstance(name, int):
        return name
    return GATES.index(name)


def is_two_qubit(name) -> bool:
    return gate_id(name) >= len(SINGLE)
