"""CPU model of the batched sort-free rotation's searches (paulib_amd/csrc/
plh_rotate.hip), checked against brute force on random sorted batches.

It restates, step for step, the device logic that places the new terms of
a rotation without a sort:
- k_rj_first's common-prefix index: l0[t] = the key bits that term t shares
  with term t - 1, and its minima per 64 terms, per 1024-term tile and per
  16 tiles;
- k_rj_new's wave-at-a-time group bounds: ballots over 64 consecutive
  terms, and wave-wide searches past the wave's edges;
- the rank correction at each set bit of M (the anticommuting count of the
  other half of the term's group);
- k_rj_lb's lower-bound range, the other half of the first set bit's group.

The ranks must equal the new keys' positions in a full sort, and the lower
bounds must equal bisection over all old keys.  The batches hold several
jobs, some not rotated this step, and groups that span more than 16 tiles.
This pins the algorithm; the GPU tests pin the kernels against the oracle.
"""
import bisect

import numpy as np
import pytest

WORD, TILE, BLOCK = 64, 1024, 16  # terms per wave, terms per tile, tiles per l3 entry


def _batch(rng, njobs, per_job):
    """Sorted distinct keys (job, x, z, imag << 56) of one 64-bit word each."""
    keys = set()
    for j in range(njobs):
        dens = float(rng.choice([0.02, 0.1, 0.3]))
        nq = int(rng.integers(8, 65))
        for _ in range(int(rng.integers(1, per_job))):
            x = z = 0
            bits = rng.random((nq, 2)) < dens
            for q in range(nq):
                if bits[q, 0]:
                    x |= 1 << (63 - q)
                if bits[q, 1]:
                    z |= 1 << (63 - q)
            keys.add((j, x, z, int(rng.random() < 0.2) << 56))
    return sorted(keys)


def _lcp(a, b):
    for w in range(3):  # job, x, z words
        d = a[w] ^ b[w]
        if d:
            return 64 * w + (64 - d.bit_length())
    return 64 * 3


class _Index:
    def __init__(self, keys):
        self.n = n = len(keys)
        self.tiles = (n + TILE - 1) // TILE
        l0 = np.zeros(self.tiles * TILE, dtype=np.int64)
        for t in range(1, n):
            l0[t] = _lcp(keys[t], keys[t - 1])
        self.l0 = l0
        self.l1 = l0.reshape(-1, WORD).min(1)
        self.l2 = l0.reshape(-1, TILE).min(1)
        self.l3 = np.array([self.l2[u:u + BLOCK].min() for u in range(0, self.tiles, BLOCK)])

    @staticmethod
    def _mask(pred):
        return sum(1 << k for k, p in enumerate(pred) if p)

    def _word_last(self, w, thr):
        b = self._mask(self.l0[WORD * w + k] < thr for k in range(WORD))
        return WORD * w + (b.bit_length() - 1 if b else 0)

    def _tile_last(self, T, thr):
        b = self._mask(self.l1[16 * T + k] < thr for k in range(16))
        return self._word_last(16 * T + (b.bit_length() - 1 if b else 0), thr)

    def _block_last(self, u, thr):
        b = self._mask(16 * u + k < self.tiles and self.l2[16 * u + k] < thr for k in range(16))
        return self._tile_last(16 * u + (b.bit_length() - 1 if b else 0), thr)

    def prev_lt(self, p, thr):
        """wv_prev_lt: the last t < p with l0[t] < thr."""
        q = p - 1
        w, T = q >> 6, (q >> 6) >> 4
        b = self._mask(WORD * w + k <= q and self.l0[WORD * w + k] < thr for k in range(WORD))
        if b:
            return WORD * w + b.bit_length() - 1
        b = self._mask((w & ~15) + k < w and self.l1[(w & ~15) + k] < thr for k in range(16))
        if b:
            return self._word_last((w & ~15) + b.bit_length() - 1, thr)
        b = self._mask((T & ~15) + k < T and self.l2[(T & ~15) + k] < thr for k in range(16))
        if b:
            return self._tile_last((T & ~15) + b.bit_length() - 1, thr)
        U = T >> 4
        while U > 0:
            u0 = U - 64 if U >= 64 else 0
            b = self._mask(u0 + k < U and self.l3[u0 + k] < thr for k in range(64))
            if b:
                return self._block_last(u0 + b.bit_length() - 1, thr)
            U = u0
        return 0

    def _low(self, b):
        return (b & -b).bit_length() - 1

    def _word_first(self, w, thr):
        b = self._mask(self.l0[WORD * w + k] < thr for k in range(WORD))
        return WORD * w + (self._low(b) if b else 63)

    def _tile_first(self, T, thr):
        b = self._mask(self.l1[16 * T + k] < thr for k in range(16))
        return self._word_first(16 * T + (self._low(b) if b else 15), thr)

    def _block_first(self, u, thr):
        b = self._mask(16 * u + k < self.tiles and self.l2[16 * u + k] < thr for k in range(16))
        return self._tile_first(16 * u + self._low(b) if b else self.tiles - 1, thr)

    def next_lt(self, p, thr):
        """wv_next_lt: the first t > p with l0[t] < thr, n when none."""
        s, n = p + 1, self.n
        if s >= n:
            return n
        w, T = s >> 6, (s >> 6) >> 4
        b = self._mask(WORD * w + k >= s and self.l0[WORD * w + k] < thr for k in range(WORD))
        if b:
            return min(WORD * w + self._low(b), n)
        b = self._mask((w & ~15) + k > w and self.l1[(w & ~15) + k] < thr for k in range(16))
        if b:
            return min(self._word_first((w & ~15) + self._low(b), thr), n)
        b = self._mask((T & ~15) + k > T and (T & ~15) + k < self.tiles and self.l2[(T & ~15) + k] < thr
                       for k in range(16))
        if b:
            return min(self._tile_first((T & ~15) + self._low(b), thr), n)
        nu = (self.tiles + 15) >> 4
        u0 = (T >> 4) + 1
        while u0 < nu:
            b = self._mask(u0 + k < nu and self.l3[u0 + k] < thr for k in range(64))
            if b:
                return min(self._block_first(u0 + self._low(b), thr), n)
            u0 += 64
        return n


def _set_bits(px, pz):
    """M's set bits as key positions g, most significant first (x word, then z)."""
    gs = []
    for h, word in enumerate((px, pz)):
        while word:
            b = word.bit_length() - 1
            word &= ~(1 << b)
            gs.append(64 * (1 + h) + (63 - b))
    return gs


def _bit(key, g):
    return (key[g >> 6] >> (63 - (g & 63))) & 1


@pytest.mark.parametrize("seed,per_job", [(1, 4000), (2, 9000), (3, 60000)])
def test_ranks_and_lower_bounds_match_brute_force(seed, per_job):
    rng = np.random.default_rng(seed)
    keys = _batch(rng, int(rng.integers(2, 5)), per_job)
    n = len(keys)
    idx = _Index(keys)
    jobs = sorted({k[0] for k in keys})
    gen = {}
    for j in jobs:
        px = pz = 0
        if rng.random() < 0.8:  # the others apply no rotation this step
            for _ in range(int(rng.integers(1, 4))):
                q = int(rng.integers(0, 64))
                if rng.random() < 0.5:
                    px |= 1 << (63 - q)
                else:
                    pz |= 1 << (63 - q)
        gen[j] = (px, pz)
    anti = np.array([bin((k[1] & gen[k[0]][1]) ^ (k[2] & gen[k[0]][0])).count("1") & 1 for k in keys], np.int64)
    A = np.concatenate([[0], np.cumsum(anti)])
    newkeys = sorted(((k[0], k[1] ^ gen[k[0]][0], k[2] ^ gen[k[0]][1], k[3]), i)
                     for i, k in enumerate(keys) if anti[i])
    want_rank = {i: r for r, (_, i) in enumerate(newkeys)}
    checked = 0
    for i0 in range(0, n, WORD):
        lanes = range(WORD)
        inn = [i0 + L < n for L in lanes]
        act = [inn[L] and anti[i0 + L] == 1 for L in lanes]
        if not any(act):
            continue
        lz = [int(idx.l0[i0 + L]) if inn[L] else 0 for L in lanes]
        G = [_set_bits(*gen[keys[i0 + L][0]]) if inn[L] else [] for L in lanes]
        r = [int(A[i0 + L]) if act[L] else 0 for L in lanes]
        h = [(0, 0)] * WORD
        for k in range(max(len(g) for g in G)):
            has = [k < len(G[L]) for L in lanes]
            g = [G[L][k] if has[L] else 0 for L in lanes]
            tlt = [g[L] if has[L] else 64 for L in lanes]
            tle = [g[L] + 1 if has[L] else 64 for L in lanes]
            s_lt = _Index._mask(lz[L] < tlt[L] for L in lanes)
            s_le = _Index._mask(lz[L] < tle[L] for L in lanes)
            below = [(1 << (L + 1)) - 1 for L in lanes]
            full = (1 << WORD) - 1
            p_le = p_lt = 0
            n_le = n_lt = n
            if any(act[L] and has[L] and (s_lt & below[L]) == 0 for L in lanes):
                p_le = idx.prev_lt(i0, tle[0])
                p_lt = p_le if idx.l0[p_le] < tlt[0] else idx.prev_lt(p_le, tlt[0])
            if any(act[L] and has[L] and (s_lt & ~below[L] & full) == 0 for L in lanes):
                n_le = idx.next_lt(i0 + WORD - 1, tle[-1])
                n_lt = n_le if (n_le >= n or idx.l0[n_le] < tlt[-1]) else idx.next_lt(n_le, tlt[-1])
            for L in lanes:
                if not (act[L] and has[L]):
                    continue
                if _bit(keys[i0 + L], g[L]):
                    bm = s_le & below[L]
                    mid = i0 + bm.bit_length() - 1 if bm else p_le
                    if bm:
                        ml = bm.bit_length() - 1
                        bl = s_lt & ((1 << ml) - 1)
                        lo = (i0 + bl.bit_length() - 1 if bl else p_lt) if not (s_lt >> ml) & 1 else mid
                    else:
                        lo = p_lt
                    r[L] -= int(A[mid] - A[lo])
                    if k == 0:
                        h[L] = (lo, mid)
                else:
                    bm = s_le & ~below[L] & full
                    mid = i0 + (bm & -bm).bit_length() - 1 if bm else n_le
                    if bm:
                        ml = (bm & -bm).bit_length() - 1
                        split = not (s_lt >> ml) & 1 and mid < n
                        bh = 0 if ml == WORD - 1 else s_lt & ~((2 << ml) - 1) & full
                        hi = (i0 + (bh & -bh).bit_length() - 1 if bh else n_lt) if split else mid
                    else:
                        hi = n_lt
                    r[L] += int(A[hi] - A[mid])
                    if k == 0:
                        h[L] = (mid, hi)
        for L in lanes:
            if not act[L]:
                continue
            i = i0 + L
            key = keys[i]
            px, pz = gen[key[0]]
            nk = (key[0], key[1] ^ px, key[2] ^ pz, key[3])
            lo, hi = h[L]
            assert r[L] == want_rank[i], f"rank of term {i}"
            assert lo + bisect.bisect_left(keys[lo:hi], nk) == bisect.bisect_left(keys, nk), f"lower bound of term {i}"
            checked += 1
    assert checked == len(newkeys)
    if per_job > 20000:
        assert idx.tiles > BLOCK  # the 16-tile level is exercised
