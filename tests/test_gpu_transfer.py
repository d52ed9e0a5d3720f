"""Downloads into pageable host memory: copies of 1 MiB and more go through
the context's two pinned staging chunks (64 MiB each, the DMA of one chunk
overlapping the host copy of the other); the bytes must arrive unchanged,
for whole sums and for ranges at unaligned offsets."""
import numpy as np
import pytest

from paulib_amd import Context, PauliSum
from paulib_amd.host import HostSum

pytestmark = pytest.mark.gpu


def _big(n, as_, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 1 << 63, size=(as_, n), dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    z = rng.integers(0, 1 << 63, size=(as_, n), dtype=np.uint64)
    f = rng.integers(0, 4, size=n, dtype=np.uint8)
    v = rng.uniform(-1, 1, size=n)
    return HostSum(as_, 64 * as_, x, z, f, v)


@pytest.mark.parametrize("n,as_", [(9_000_001, 1), (4_194_305, 2), (40_000_000, 1), (200_003, 1),
                                   (131_071, 1)])
def test_staged_download_is_exact(ctx, n, as_):
    h = _big(n, as_, n)
    print(f"staged download n={n} as={as_}", flush=True)
    d = PauliSum.from_host(ctx, h)
    got = d.to_host()
    assert (got.x == h.x).all() and (got.z == h.z).all()
    assert (got.flags == h.flags).all() and (got.value.view(np.uint64) == h.value.view(np.uint64)).all()
    off, cnt = 12_345, n - 12_345 - 777
    r = d.range_to_host(off, cnt)
    assert (r.x == h.x[:, off:off + cnt]).all() and (r.z == h.z[:, off:off + cnt]).all()
    assert (r.flags == h.flags[off:off + cnt]).all()
    assert (r.value.view(np.uint64) == h.value[off:off + cnt].view(np.uint64)).all()
    d.close()


def test_staged_download_in_a_second_context():
    """Each context owns its staging chunks; two contexts download in turn."""
    h = _big(5_000_003, 1, 3)
    c1, c2 = Context(0), Context(0)
    try:
        for c in (c1, c2, c1):
            d = PauliSum.from_host(c, h)
            got = d.to_host()
            assert (got.x == h.x).all() and (got.value.view(np.uint64) == h.value.view(np.uint64)).all()
            d.close()
    finally:
        c1.close()
        c2.close()
