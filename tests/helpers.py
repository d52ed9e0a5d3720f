"""Comparison helpers shared by the tests."""
import hashlib

import numpy as np

import oracle


def oracle_kinds():
    """The CPU checkers present: the C restatement always, the compiled
    reference when oracle/_ref was built."""
    return ["oracle"] + (["reference"] if oracle.available("reference") else [])


def assert_terms_equal(got, want, values="exact", rtol=1e-12):
    """Strings and phases bit-exact, in order; values bit-exact or to rtol."""
    assert len(got) == len(want), f"length {len(got)} != {len(want)}"
    assert got.array_size == want.array_size
    if len(want) == 0:
        return
    bad = np.nonzero((got.x != want.x).any(axis=0) | (got.z != want.z).any(axis=0)
                     | (got.flags != want.flags))[0]
    assert bad.size == 0, f"{bad.size} terms differ in string/phase, first at {bad[0]}"
    if values == "exact":
        vb = np.nonzero(got.value.view(np.uint64) != want.value.view(np.uint64))[0]
        assert vb.size == 0, (f"{vb.size} values differ, first at {vb[0]}: "
                              f"{got.value[vb[0]]!r} vs {want.value[vb[0]]!r}")
    else:
        scale = np.maximum(np.abs(got.value), np.abs(want.value))
        err = np.abs(got.value - want.value)
        ok = err <= rtol * scale + 1e-300
        assert ok.all(), f"value mismatch beyond {rtol} at {np.nonzero(~ok)[0][:5]}"


def cpu(kind, host):
    return oracle.CpuSum.from_host(kind, host)


def digests(h) -> dict:
    """SHA-256 of a sum's planes (x words, z words, flags bytes) and of its
    coefficient bits, little-endian: a size-independent fingerprint of every
    bit of a result."""
    planes = hashlib.sha256()
    planes.update(np.ascontiguousarray(h.x, "<u8").tobytes())
    planes.update(np.ascontiguousarray(h.z, "<u8").tobytes())
    planes.update(np.ascontiguousarray(h.flags, "u1").tobytes())
    vals = hashlib.sha256(np.ascontiguousarray(h.value, "<f8").view("<u8").tobytes())
    return {"planes_sha256": planes.hexdigest(), "values_sha256": vals.hexdigest()}


def terms_of(h, idx) -> list:
    """Terms of a one-word sum as [x hex, z hex, flags, value hex] rows."""
    return [[f"{int(h.x[0, t]):016x}", f"{int(h.z[0, t]):016x}", int(h.flags[t]), float(h.value[t]).hex()]
            for t in idx]
