"""Per-device autotuning (plh_ctx_autotune): the sweep picks tile caps from
its candidate sets and a power-of-two batch threshold, writes them to the
cache, a second context loads the same values from the cache without a
sweep, and collated products stay identical to the oracle under the tuned
caps."""
import time

import numpy as np
import pytest

from helpers import assert_terms_equal, cpu
from paulib_amd import Context, PauliSum
from paulib_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_autotune_sweep_cache_and_parity(ctx, tmp_path):
    import gc

    gc.collect()
    ctx.trim()  # the sweep needs ~150 GB: the session context gives its unused memory back first
    cache = tmp_path / "tune.txt"
    c1 = Context(0)
    try:
        t = c1.autotune(cache)
        assert t["tile_root_max"] in (65536, 131072, 262144, 524288)
        assert t["tile_pass_max"] in (16384, 32768, 65536)
        sp = t["solo_pairs"]
        assert (1 << 16) <= sp <= (1 << 26) and sp & (sp - 1) == 0
        line = cache.read_text().strip().split()
        assert [int(v) for v in line[1:]] == [t["tile_root_max"], t["tile_pass_max"], sp]
        # products under the tuned caps
        A = W.random_sum(3000, 64, 0.75, "uniform", 1)
        B = W.random_sum(2500, 64, 0.75, "uniform", 2)
        got = PauliSum.from_host(c1, A).multiply_canonical(PauliSum.from_host(c1, B)).to_host()
        o = cpu("oracle", A)
        o.multiply(cpu("oracle", B))
        o.canonicalize()
        assert_terms_equal(got, o.to_host())
    finally:
        c1.close()
    c2 = Context(0)
    try:
        t0 = time.perf_counter()
        t2 = c2.autotune(cache)
        assert time.perf_counter() - t0 < 1.0  # from the cache, no sweep
        assert t2 == t
    finally:
        c2.close()
