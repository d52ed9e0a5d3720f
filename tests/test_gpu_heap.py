"""The context's device heap: ranges carved from shared regions must never
overlap.  A seeded sequence of sums of mixed sizes (small ranges that share
regions, and ranges of 256 MiB and more that get regions of their own) is
created, filled, dropped and re-created in a fresh context; after every
step each live sum must still read back exactly what was written into it,
and the context's accounting must add up.
"""
import numpy as np
import pytest

from paulib_amd import Context, PauliSum
from paulib_amd.host import HostSum

pytestmark = pytest.mark.gpu

LARGE_TERMS = 11_000_000  # 25 B a term at one word per plane: over 256 MiB


def _fill(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.integers(0, 2**63, size=(1, n), dtype=np.uint64)
    z = rng.integers(0, 2**63, size=(1, n), dtype=np.uint64)
    f = rng.integers(0, 2, size=n, dtype=np.uint8)
    v = rng.standard_normal(n)
    return HostSum(1, 64, x, z, f, v)


def _same(d: PauliSum, h: HostSum):
    g = d.to_host()
    assert len(g) == len(h)
    assert np.array_equal(g.x, h.x) and np.array_equal(g.z, h.z)
    assert np.array_equal(g.flags, h.flags)
    assert np.array_equal(g.value.view(np.uint64), h.value.view(np.uint64))


def test_heap_ranges_stay_disjoint():
    ctx = Context(0)
    rng = np.random.Generator(np.random.PCG64(2024))
    live = {}
    try:
        for step in range(40):
            if live and (len(live) >= 6 or rng.random() < 0.4):
                k = list(live)[int(rng.integers(len(live)))]
                d, _ = live.pop(k)
                d.close()
            else:
                n = LARGE_TERMS if rng.random() < 0.15 else int(rng.integers(1, 300_000))
                h = _fill(n, step)
                live[step] = (PauliSum.from_host(ctx, h), h)
            for d, h in live.values():
                _same(d, h)
            assert ctx.scratch_bytes() >= 0
        for d, _ in live.values():
            d.close()
        live.clear()
        ctx.trim()
        assert ctx.scratch_bytes() == 0
    finally:
        for d, _ in live.values():
            d.close()
        ctx.close()


def test_heap_reuse_after_release_matches_fresh(ctx):
    # a product whose id lists (81M pairs, 324 MB each) are large scratch,
    # three times, with other work between:
    # the second and third results come from reused ranges
    from paulib_amd import workloads as W

    A = PauliSum.from_host(ctx, W.random_sum(9000, 64, 0.75, "uniform", 1))
    B = PauliSum.from_host(ctx, W.random_sum(9000, 64, 0.75, "uniform", 2))
    first = A.multiply_canonical(B).to_host()
    small = PauliSum.from_host(ctx, _fill(1000, 3))
    again = A.multiply_canonical(B).to_host()
    small.close()
    third = A.multiply_canonical(B).to_host()
    for got in (again, third):
        assert np.array_equal(got.x, first.x) and np.array_equal(got.z, first.z)
        assert np.array_equal(got.flags, first.flags)
        assert np.array_equal(got.value.view(np.uint64), first.value.view(np.uint64))
    A.close()
    B.close()
