"""plh_launch_commutator_soa_noalloc (the commutator entry point that
allocates nothing: its scratch is a fixed per-device array) against
plh_commutator, the entry point whose kernels it shares: the same terms,
phases and coefficients bit for bit and the same count, on two calls in a
row (the second smaller, so it runs over the first call's scratch), with
the output cut at a capacity below the count, and the size limit."""
import ctypes as C

import numpy as np
import pytest

from paulib_amd import PauliSum, _lib, commutator, i_commutator
from paulib_amd import workloads as W

pytestmark = pytest.mark.gpu


class _Hip:
    """The caller's own device memory, from the HIP runtime (hipMalloc)."""

    def __init__(self):
        try:
            self.h = C.CDLL("libamdhip64.so")
        except OSError:
            self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        if a.nbytes:
            assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        if out.nbytes:
            assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _call(hip, A, B, by_i, capacity):
    """Packed planes of A and B in caller-allocated device memory, output
    planes of `capacity` terms and a device count (not zeroed); returns the
    count and the first min(count, capacity) output terms."""
    as_ = A.array_size
    cap = max(capacity, 1)
    ax, az, af, av = hip.put(A.x), hip.put(A.z), hip.put(A.flags), hip.put(A.value)
    bx, bz, bf, bv = hip.put(B.x), hip.put(B.z), hip.put(B.flags), hip.put(B.value)
    ox = hip.put(np.full((as_, cap), ~np.uint64(0), np.uint64))
    oz = hip.put(np.full((as_, cap), ~np.uint64(0), np.uint64))
    of = hip.put(np.full(cap, 255, np.uint8))
    ov = hip.put(np.full(cap, np.nan))
    cnt = hip.put(np.array([12345], np.uint64))
    hip.sync()
    _lib.call("plh_launch_commutator_soa_noalloc", as_, 1 if by_i else 0, ax, az, af, av, bx, bz, bf, bv,
              ox, oz, of, ov, cnt, len(A), len(B), capacity, None)
    hip.sync()
    k = int(hip.get(cnt, np.zeros(1, np.uint64))[0])
    w = min(k, capacity)
    x = hip.get(ox, np.zeros((as_, cap), np.uint64))[:, :w]
    z = hip.get(oz, np.zeros((as_, cap), np.uint64))[:, :w]
    f = hip.get(of, np.zeros(cap, np.uint8))[:w]
    v = hip.get(ov, np.zeros(cap))[:w]
    hip.free()
    return k, x, z, f, v


def _want(ctx, A, B, by_i):
    a, b = PauliSum.from_host(ctx, A), PauliSum.from_host(ctx, B)
    return (i_commutator(a, b) if by_i else commutator(a, b)).to_host()


def _check(got, want, upto):
    k, x, z, f, v = got
    assert k == len(want)
    assert (x == want.x[:, :upto]).all() and (z == want.z[:, :upto]).all()
    assert (f == want.flags[:upto]).all()
    assert (v.view(np.uint64) == want.value[:upto].view(np.uint64)).all()


def test_noalloc_commutator_matches_plh_commutator(ctx):
    hip = _Hip()
    cases = [(2, 128, 700, 530, False), (1, 64, 90, 41, True)]  # second call smaller
    for as_, nq, n, m, by_i in cases:
        print(f"noalloc commutator as={as_} {n}x{m} by_i={by_i}", flush=True)
        A = W.random_words_sum(n, nq, 31 + n, array_size=as_)
        B = W.random_words_sum(m, nq, 57 + m, array_size=as_)
        want = _want(ctx, A, B, by_i)
        assert len(want) > 100
        _check(_call(hip, A, B, by_i, len(want) + 3), want, len(want))
    # output cut at a capacity below the count: the first terms, the full count
    A = W.random_words_sum(300, 200, 5, array_size=4)
    B = W.random_words_sum(120, 200, 6, array_size=4)
    want = _want(ctx, A, B, False)
    _check(_call(hip, A, B, False, len(want) // 3), want, len(want) // 3)
    # an empty operand: count 0
    E = W.random_words_sum(0, 64, 1)
    assert _call(hip, E, W.random_words_sum(10, 64, 2), False, 4)[0] == 0


def test_noalloc_commutator_rejects_oversize():
    with pytest.raises(_lib.EngineError) as e:
        _lib.call("plh_launch_commutator_soa_noalloc", 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                  70000, 70000, 16, None)
    assert e.value.code == -3
    with pytest.raises(_lib.EngineError) as e:
        _lib.call("plh_launch_commutator_soa_noalloc", 17, 0, None, None, None, None, None, None, None, None,
                  None, None, None, None, None, 1, 1, 1, None)
    assert e.value.code == -2
