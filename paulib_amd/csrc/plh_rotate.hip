// plh_rotate.hip -- a rotation of a sorted canonical sum without a sort.
//
// What it computes: evolve's general path (include/PauliSumOperations.h:
// 148-180) for one generator P followed by apply_truncation's canonicalize
// and filters (include/PauliPropagation.h:35-47, include/PauliSumUtils.h:
// 314-328, 377-393): every term Q that anticommutes with P keeps cos * v
// and gains Q * ps (ps = P with the i bit flipped) with (-sin) * v; equal
// (x, z, imag) keys combine as acc = 0; acc += old; acc += new; exact zeros,
// overweight terms and |c| < eps are dropped.  It requires the input to be
// sorted by (x, z, imag) with distinct keys (the output of any folding
// collation, kept by filters and trim), so every key has at most one old and
// one new addend and the order of combining is the reference's.
//
// How (no radix pass, no leaf):
//   1. anticommuting flags a_i, and their exclusive scan A.
//   2. a new term's key is k_i ^ M (M: P's bits; a real term times iP stays
//      real, so the imaginary bit is kept).  Its rank among the new terms is
//      A_i corrected at each set bit b of M: inside the group of keys that
//      share k_i's bits above b (a contiguous range of the sorted input, with
//      bit b = 0 first), XOR with M swaps the two halves, so the rank gains
//      the anticommuting count of the bit-1 half (k_i has bit b = 0) or loses
//      that of the bit-0 half (k_i has bit b = 1).  The halves come from
//      binary searches on the sorted input; the new terms are written in
//      sorted order.
//   3. old and new terms are merged: each finds its lower bound in the other
//      list (binary search); an old term whose key is also a new one takes
//      the sum; per-term keep flags (filters) are scanned once over both
//      lists, and each kept term's output position is its kept rank in its
//      own list plus the kept rank of its lower bound in the other.
// About seven launches and one read-back (the output count) per rotation.
//
// Batched sums (a job plane, the job id leading the key) rotate each term by
// its own job's generator when that job's step is a general rotation; the
// groups of step 2 never cross jobs because the job word lies above every
// bit of M.  Their form (rotate_sorted_jobs_as) keeps the flags as bits with
// per-tile counts and lets only the new terms search, so a step over a large
// batch moves a fraction of the bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "plh_block.hpp"
#include "plh_engine.hpp"
#include "plh_sources.hpp"

namespace plh {

namespace {

inline unsigned blocks_of(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// A sorted sum's keys on its planes: words W[0..KW) = (job,) x[0..AS),
// z[0..AS), W[KW] = (flags & 1) << 56; global bit g counts from the most
// significant bit of W[0].  Values and flags are read with the sign bit
// folded in (a sign flip of a fast-path rotation keeps the order).
template <int AS, bool JOB>
struct Sorted {
    static constexpr int XO = JOB ? 1 : 0;
    static constexpr int KW = 2 * AS + XO;
    const uint64_t* x;
    const uint64_t* z;
    const uint8_t* f;
    const double* v;
    const uint32_t* j;
    int64_t stride;
    int as;

    PLH_DEV uint64_t word(uint32_t i, int w) const {
        if (JOB && w == 0) return j ? j[i] : 0u;
        const int k = w - XO;
        if (k < AS) return k < as ? x[(int64_t)k * stride + i] : 0ull;
        if (k < 2 * AS) return (k - AS) < as ? z[(int64_t)(k - AS) * stride + i] : 0ull;
        return (uint64_t)(f[i] & 1u) << 56;
    }
    PLH_DEV void key(uint32_t i, uint64_t* kw) const {
#pragma unroll
        for (int w = 0; w <= KW; ++w) kw[w] = word(i, w);
    }
    PLH_DEV double value(uint32_t i) const { return (f[i] & 2u) ? -v[i] : v[i]; }
    PLH_DEV uint32_t job(uint32_t i) const { return JOB && j ? j[i] : 0u; }
};

template <int KW>
PLH_DEV bool key_less(const uint64_t* a, const uint64_t* b) {
#pragma unroll
    for (int w = 0; w <= KW; ++w)
        if (a[w] != b[w]) return a[w] < b[w];
    return false;
}

template <int KW>
PLH_DEV bool key_eq(const uint64_t* a, const uint64_t* b) {
#pragma unroll
    for (int w = 0; w <= KW; ++w)
        if (a[w] != b[w]) return false;
    return true;
}

PLH_DEV uint32_t bit_at(const uint64_t* kw, int g) { return (uint32_t)(kw[g >> 6] >> (63 - (g & 63))) & 1u; }

// Compares bits [0, g) of the key of item t with those of kw: -1, 0, 1.
template <int AS, bool JOB>
PLH_DEV int prefix_cmp(const Sorted<AS, JOB>& S, uint32_t t, const uint64_t* kw, int g) {
    const int wg = g >> 6, o = g & 63;
    for (int w = 0; w < wg; ++w) {
        const uint64_t a = S.word(t, w);
        if (a != kw[w]) return a < kw[w] ? -1 : 1;
    }
    if (o) {
        const uint64_t a = S.word(t, wg) >> (64 - o), b = kw[wg] >> (64 - o);
        if (a != b) return a < b ? -1 : 1;
    }
    return 0;
}

// The rotation of item i: false when none applies (a batched job whose step
// is not a general rotation).
template <int AS, bool JOB>
PLH_DEV bool rot_of(const RotParams<AS>& rp, const JobRotTables& jt, uint32_t job, const uint64_t*& px,
                    const uint64_t*& pz, uint32_t& psf, double& c, double& sn) {
    if (!JOB) {
        px = rp.px;
        pz = rp.pz;
        psf = rp.psf;
        c = rp.c;
        sn = rp.s;
        return true;
    }
    if (jt.kind[job] != 4) return false;
    px = jt.p + (size_t)job * 2 * AS;
    pz = px + AS;
    psf = jt.psf[job];
    c = jt.c[job];
    sn = jt.s[job];
    return true;
}

template <int AS>
PLH_DEV bool anticommutes(const uint64_t* kw, int xo, const uint64_t* px, const uint64_t* pz) {
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k) acc ^= (kw[xo + k] & pz[k]) ^ (kw[xo + AS + k] & px[k]);
    return popc(acc) & 1;
}

// Filters of an output term (canonicalize's zero drop, truncate_weight,
// cutoff; per job for batched sums, whose drained jobs keep nothing).
struct Filters {
    uint32_t mode;
    int64_t max_weight;
    double eps;
    const double* job_eps;
    const int64_t* job_maxw;
    const uint8_t* job_alive;
};

template <int AS, int XO>
PLH_DEV bool keep_of(const Filters& fl, const uint64_t* kw, double acc) {
    if ((fl.mode & kDropZero) && acc == 0.0) return false;
    const uint32_t job = XO ? (uint32_t)kw[0] : 0u;
    if (XO && fl.job_alive && !fl.job_alive[job]) return false;
    if (fl.mode & kMaxWeight) {
        int64_t wgt = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) wgt += popc(kw[XO + k] | kw[XO + AS + k]);
        const int64_t mw = (XO && fl.job_maxw) ? fl.job_maxw[job] : fl.max_weight;
        if (mw >= 0 && wgt > mw) return false;
    }
    if (fl.mode & kCutoff) {
        const double eps = (XO && fl.job_eps) ? fl.job_eps[job] : fl.eps;
        if (eps > 0.0 && !(fabs(acc) >= eps)) return false;
    }
    return true;
}

// New terms, sorted: planes of stride `cap`.
struct NewTerms {
    uint64_t* x;
    uint64_t* z;
    uint8_t* f;
    double* v;
    uint32_t* j;
    int64_t cap;
};

// Job ranges of a batched sum sorted by job: off[k] = first term of job k
// (k <= njobs).  Searches of the rotation stay inside an item's own job.
__global__ void k_rs_bounds(const uint32_t* __restrict__ J, uint32_t n, uint32_t njobs, uint32_t* __restrict__ off) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > njobs) return;
    uint32_t lo = 0, cnt = n;
    while (cnt > 0) {
        const uint32_t half = cnt / 2, t = lo + half;
        if (J[t] < k) {
            lo = t + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    off[k] = lo;
}

// ---- 1. anticommuting flags
template <int AS, bool JOB>
__global__ __launch_bounds__(kBlock) void k_rs_anti(Sorted<AS, JOB> S, uint32_t n, RotParams<AS> rp, JobRotTables jt,
                                                    uint32_t* __restrict__ a) {
    constexpr int KW = Sorted<AS, JOB>::KW;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t kw[KW + 1];
    S.key(i, kw);
    const uint64_t *px, *pz;
    uint32_t psf;
    double c, sn;
    const bool r = rot_of<AS, JOB>(rp, jt, S.job(i), px, pz, psf, c, sn);
    a[i] = (r && anticommutes<AS>(kw, S.XO, px, pz)) ? 1u : 0u;
}

// Galloping searches over a sorted range: the cost grows with the log of
// the distance to the answer, not of the range, and the groups of the low
// key bits a rotation flips hold few terms.
// pred holds on [x, i] and fails on [base, x) (pred(i) holds); returns x.
template <class P>
PLH_DEV uint32_t gallop_back(uint32_t base, uint32_t i, P pred) {
    uint32_t t_true = i, step = 1, L = base;
    while (t_true - base >= step) {
        const uint32_t t = t_true - step;
        if (!pred(t)) {
            L = t + 1;
            break;
        }
        t_true = t;
        step <<= 1;
    }
    uint32_t cnt = t_true - L;  // the first true lies in [L, t_true]
    while (cnt > 0) {
        const uint32_t half = cnt / 2, t = L + half;
        if (!pred(t)) {
            L = t + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    return L;
}

// pred holds on [i, x) and fails on [x, lim); returns x (lim when it holds to the end).
template <class P>
PLH_DEV uint32_t gallop_fwd(uint32_t i, uint32_t lim, P pred) {
    uint32_t t_end = i, step = 1, hi = lim;
    while (lim - t_end >= step) {
        const uint32_t t = t_end + step - 1;
        if (!pred(t)) {
            hi = t;
            break;
        }
        t_end = t + 1;
        step <<= 1;
    }
    uint32_t L = t_end, cnt = hi - L;  // the first false lies in [t_end, hi]
    while (cnt > 0) {
        const uint32_t half = cnt / 2, t = L + half;
        if (pred(t)) {
            L = t + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    return L;
}

// Rank of item i's new term k_i ^ M among the new terms: rank(i) (the
// anticommuting items before i) corrected at each set bit g of M: inside
// the group of items sharing i's bits above g ([base, lim) bounds the
// search), XOR with M swaps the bit-0 and bit-1 halves.  rank(t) = the
// anticommuting items before t.
template <int AS, bool JOB, class Rank>
PLH_DEV int64_t new_rank(const Sorted<AS, JOB>& S, const uint64_t* kw, uint32_t i, uint32_t base, uint32_t lim,
                         const uint64_t* px, const uint64_t* pz, Rank rank) {
    constexpr int XO = Sorted<AS, JOB>::XO;
    int64_t r = (int64_t)rank(i);
#pragma unroll
    for (int h = 0; h < 2 * AS; ++h) {
        uint64_t m = h < AS ? px[h] : pz[h - AS];
        const int wbase = 64 * (XO + h);
        while (m) {
            const int bit = 63 - __clzll(m);  // bit index within the word
            m &= ~(1ull << bit);
            const int g = wbase + (63 - bit);
            const int gw = g >> 6, gs = 63 - (g & 63);
            auto same_group = [&](uint32_t t) { return prefix_cmp(S, t, kw, g) == 0; };
            auto bit_set = [&](uint32_t t) { return ((S.word(t, gw) >> gs) & 1ull) != 0; };
            if (bit_at(kw, g)) {
                // group start, then its bit-1 half's start (both at or before i)
                const uint32_t lo = gallop_back(base, i, same_group);
                const uint32_t mid = gallop_back(lo, i, bit_set);
                r -= (int64_t)(rank(mid) - rank(lo));
            } else {
                // group end, then its bit-1 half's start (both after i)
                const uint32_t hi = gallop_fwd(i + 1, lim, same_group);
                const uint32_t mid = gallop_fwd(i + 1, hi, [&](uint32_t t) { return !bit_set(t); });
                r += (int64_t)(rank(hi) - rank(mid));
            }
        }
    }
    return r;
}

// ---- 2. new terms at their ranks (A: exclusive scan of a, A[n] = count)
template <int AS, bool JOB>
__global__ __launch_bounds__(kBlock) void k_rs_new(Sorted<AS, JOB> S, uint32_t n, RotParams<AS> rp, JobRotTables jt,
                                                   const uint32_t* __restrict__ a, const uint64_t* __restrict__ A,
                                                   const uint32_t* __restrict__ joff, NewTerms N) {
    constexpr int KW = Sorted<AS, JOB>::KW;
    constexpr int XO = Sorted<AS, JOB>::XO;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || !a[i]) return;
    uint64_t kw[KW + 1];
    S.key(i, kw);
    const uint64_t *px, *pz;
    uint32_t psf;
    double c, sn;
    const uint32_t job = S.job(i);
    rot_of<AS, JOB>(rp, jt, job, px, pz, psf, c, sn);
    // the item's job range (the groups below never leave it)
    const uint32_t base = JOB ? joff[job] : 0u, lim = JOB ? joff[job + 1] : n;
    const int64_t r = new_rank<AS, JOB>(S, kw, i, base, lim, px, pz, [&](uint32_t t) { return A[t]; });
    if (r < 0 || (uint64_t)r >= A[n]) return;  // never with sorted distinct input; keeps writes in bounds
    // the new term Q * ps (PauliSum::multiply(i, ps), :185-209), sign folded
    uint64_t c1 = 0, c2 = 0;
    const uint32_t f0 = S.f[i] & 1u;
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        phase_word(kw[XO + k], kw[XO + AS + k], px[k], pz[k], c1, c2);
        N.x[(int64_t)k * N.cap + r] = kw[XO + k] ^ px[k];
        N.z[(int64_t)k * N.cap + r] = kw[XO + AS + k] ^ pz[k];
    }
    const uint32_t raw = phase_of(f0, psf, c1, c2);
    const double nv = (-sn) * S.value(i);
    N.v[r] = (raw & 2u) ? -nv : nv;
    N.f[r] = (uint8_t)(raw & 1u);
    if (JOB) N.j[r] = S.job(i);
}

template <int AS, bool JOB>
PLH_DEV void new_key(const NewTerms& N, uint32_t r, uint64_t* kw) {
    constexpr int XO = JOB ? 1 : 0;
    if (JOB) kw[0] = N.j[r];
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        kw[XO + k] = N.x[(int64_t)k * N.cap + r];
        kw[XO + AS + k] = N.z[(int64_t)k * N.cap + r];
    }
    kw[2 * AS + XO] = (uint64_t)(N.f[r] & 1u) << 56;
}

// ---- 3a. merge: lower bounds in the other list, sums, keep flags.  Threads
// [0, n) take the old terms, [n, 2n) the new ones (A[n] of them).
template <int AS, bool JOB>
__global__ __launch_bounds__(kBlock) void k_rs_merge(Sorted<AS, JOB> S, uint32_t n, RotParams<AS> rp, JobRotTables jt,
                                                     const uint32_t* __restrict__ a, const uint64_t* __restrict__ A,
                                                     const uint32_t* __restrict__ joff, NewTerms N, Filters fl,
                                                     uint32_t* __restrict__ keep, uint32_t* __restrict__ lb_out,
                                                     double* __restrict__ val) {
    constexpr int KW = Sorted<AS, JOB>::KW;
    constexpr int XO = Sorted<AS, JOB>::XO;
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 2ull * n) return;
    const uint32_t nn = (uint32_t)A[n];
    uint64_t kw[KW + 1], ow[KW + 1];
    if (gid < n) {
        const uint32_t i = (uint32_t)gid;
        S.key(i, kw);
        double vo = S.value(i);
        if (a[i]) {
            const uint64_t *px, *pz;
            uint32_t psf;
            double c, sn;
            rot_of<AS, JOB>(rp, jt, S.job(i), px, pz, psf, c, sn);
            vo = vo * c;
        }
        // first new term >= k_i, inside the job's new terms
        const uint32_t job = S.job(i);
        uint32_t lo = JOB ? (uint32_t)A[joff[job]] : 0u;
        uint32_t cnt = (JOB ? (uint32_t)A[joff[job + 1]] : nn) - lo;
        while (cnt > 0) {
            const uint32_t half = cnt / 2, t = lo + half;
            new_key<AS, JOB>(N, t, ow);
            if (key_less<KW>(ow, kw)) {
                lo = t + 1;
                cnt -= half + 1;
            } else {
                cnt = half;
            }
        }
        double acc = 0.0;
        acc += vo;
        if (lo < (JOB ? (uint32_t)A[joff[job + 1]] : nn)) {
            new_key<AS, JOB>(N, lo, ow);
            if (key_eq<KW>(ow, kw)) acc += N.v[lo];
        }
        keep[i] = keep_of<AS, XO>(fl, kw, acc) ? 1u : 0u;
        lb_out[i] = lo;
        val[i] = acc;
        return;
    }
    const uint32_t r = (uint32_t)(gid - n);
    if (r >= nn) {
        keep[gid] = 0u;
        return;
    }
    new_key<AS, JOB>(N, r, kw);
    // first old term >= n_r, inside the job's old terms
    const uint32_t job = JOB ? N.j[r] : 0u;
    uint32_t lo = JOB ? joff[job] : 0u;
    uint32_t cnt = (JOB ? joff[job + 1] : n) - lo;
    while (cnt > 0) {
        const uint32_t half = cnt / 2, t = lo + half;
        S.key(t, ow);
        if (key_less<KW>(ow, kw)) {
            lo = t + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    bool dup = false;
    if (lo < (JOB ? joff[job + 1] : n)) {
        S.key(lo, ow);
        dup = key_eq<KW>(ow, kw);
    }
    keep[gid] = (!dup && keep_of<AS, XO>(fl, kw, N.v[r])) ? 1u : 0u;
    lb_out[gid] = lo;
}

// ---- 3b. kept terms to their output positions (C: exclusive scan of keep
// over both lists).
template <int AS, bool JOB>
__global__ __launch_bounds__(kBlock) void k_rs_write(Sorted<AS, JOB> S, uint32_t n, const uint64_t* __restrict__ A,
                                                     NewTerms N, const uint32_t* __restrict__ keep,
                                                     const uint32_t* __restrict__ lb, const double* __restrict__ val,
                                                     const uint64_t* __restrict__ C, int out_as, uint64_t* __restrict__ ox,
                                                     uint64_t* __restrict__ oz, uint8_t* __restrict__ of,
                                                     double* __restrict__ ov, uint32_t* __restrict__ oj, int64_t os) {
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 2ull * n || !keep[gid]) return;
    const uint64_t Cn = C[n];
    if (gid < n) {
        const uint32_t i = (uint32_t)gid;
        const uint64_t o = C[i] + (C[(uint64_t)n + lb[i]] - Cn);
        if (o >= C[2ull * n]) return;  // never for consistent lower bounds; keeps writes in bounds
        for (int k = 0; k < out_as; ++k) {
            ox[(int64_t)k * os + o] = S.x[(int64_t)k * S.stride + i];
            oz[(int64_t)k * os + o] = S.z[(int64_t)k * S.stride + i];
        }
        of[o] = S.f[i] & 1u;
        ov[o] = val[i];
        if (JOB) oj[o] = S.j[i];
        return;
    }
    const uint32_t r = (uint32_t)(gid - n);
    const uint64_t o = (C[gid] - Cn) + C[lb[gid]];
    if (o >= C[2ull * n]) return;
    for (int k = 0; k < out_as; ++k) {
        ox[(int64_t)k * os + o] = N.x[(int64_t)k * N.cap + r];
        oz[(int64_t)k * os + o] = N.z[(int64_t)k * N.cap + r];
    }
    of[o] = N.f[r];
    ov[o] = N.v[r];
    if (JOB) oj[o] = N.j[r];
}

size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

// ----------------------------------------------------------------------
// Batched sums (the server's sorted steps): the same rotation with the
// per-term state held as bit flags.  Old terms never search: the few new
// terms find their lower bounds among the old ones, and an old term learns
// the new terms below it (and its equal-key partner) from the sorted lower
// bounds that fall in its tile.  Flags are bits with per-tile counts, so a
// rank is a tile prefix, a word prefix and one popcount instead of a
// full-length scan.  Per old term the passes read its job and the words
// its job's P touches, then its value and flags, then the whole term
// to write it: about 20 + 13 + 45 bytes read and 45 written at two words
// per plane, against about 330 for the two-list merge above.
//
// Precondition (the server's sorted steps): every alive job's terms have
// passed a truncation under the job's weight limit since its keys last
// changed, so an old term, whose key a rotation keeps, is dropped only by
// its coefficient (zero, cutoff) or by its job being drained.

constexpr int kRjTile = 1024;              // terms per tile: 256 threads x 4
constexpr int kRjWords = kRjTile / kWave;  // flag words per tile
static_assert(kBlock * 4 == kRjTile, "a thread holds four terms of its tile");

// Flags over terms t < cap (cap = tiles * kRjTile; flags past the live
// count are 0): bits[w] = terms 64w .. 64w + 63, wpre[w] = set flags of
// w's tile before word w, tpre[t] = exclusive scan of the tile counts
// (tpre[tiles] = total).
struct BitRank {
    const uint64_t* bits;
    const uint32_t* wpre;
    const uint64_t* tpre;
    uint64_t cap;
    uint64_t tiles;
    PLH_DEV bool test(uint64_t i) const { return (bits[i >> 6] >> (i & 63)) & 1ull; }
    PLH_DEV uint64_t rank(uint64_t i) const {  // set flags before i
        if (i >= cap) return tpre[tiles];
        const uint64_t w = i >> 6;
        const uint64_t below = bits[w] & ((1ull << (i & 63)) - 1ull);
        return tpre[i / kRjTile] + wpre[w] + (uint64_t)popc(below);
    }
    PLH_DEV uint64_t total() const { return tpre[tiles]; }
};

// Where a tile's flags go: bits and in-tile word prefixes, the tile's
// count (null: not counted).
struct BitOut {
    uint64_t* bits;
    uint32_t* wpre;
    uint32_t* tcnt;
};

// A tile of 1024 terms is taken by a workgroup of 1024 / IPT threads, each
// holding IPT terms: 4 per thread (256 threads) for large sums, 1 per
// thread (1024 threads) for small ones, whose passes have too few waves to
// hide their loads' latency.  Term k (0..IPT-1) of this thread in tile t:
// t * 1024 + wave * 64 IPT + k * 64 + lane, so word t * 16 + wave * IPT + k
// holds it.
template <int IPT>
PLH_DEV uint64_t rj_item(uint64_t tile, int k) {
    return tile * kRjTile + (uint64_t)(threadIdx.x / kWave) * (IPT * kWave) + (uint64_t)k * kWave + lane_id();
}

// Every thread of the block: writes the tile's flag words, their in-tile
// prefixes and the tile count.  sh: kRjWords entries of LDS.
template <int IPT>
PLH_DEV void rj_put(const bool (&fl)[IPT], uint64_t tile, const BitOut& o, uint32_t* sh) {
    const int wv = (int)(threadIdx.x / kWave);
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint64_t b = ballot(fl[k]);
        if (lane_id() == 0) {
            o.bits[tile * kRjWords + (uint64_t)(wv * IPT + k)] = b;
            sh[wv * IPT + k] = (uint32_t)popc(b);
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)kRjWords) {
        uint32_t s = 0;
        for (unsigned w = 0; w < threadIdx.x; ++w) s += sh[w];
        if (o.wpre) o.wpre[tile * kRjWords + threadIdx.x] = s;
        if (o.tcnt && threadIdx.x == (unsigned)kRjWords - 1) o.tcnt[tile] = s + sh[kRjWords - 1];
    }
}

// First index in [lo, hi) with v[t] >= key (gt: > key).
PLH_DEV uint32_t u32_bound(const uint32_t* v, uint32_t lo, uint32_t hi, uint32_t key, bool gt) {
    uint32_t cnt = hi - lo;
    while (cnt > 0) {
        const uint32_t half = cnt / 2, t = lo + half;
        if (gt ? v[t] <= key : v[t] < key) {
            lo = t + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    return lo;
}

// Common key prefixes of neighbours in a sorted batch: l0[t] = the leading
// key bits (job word, x words, z words; at most 64 * (2 AS + 1)) that term t
// shares with term t - 1, 0 for t = 0 and for t >= n.  The terms sharing
// term i's first g bits (g >= 64: inside its job) are the range from the
// last t <= i with l0[t] < g to the first t > i with l0[t] < g, and inside
// it the terms with bit g set start at the t with l0[t] == g, if any.  Both
// are nearest-smaller queries on l0, answered through minima per 64 terms
// (l1), per tile (l2) and per 16 tiles (l3): a few contiguous loads near i
// instead of key compares at galloping distances.  A sum without a job plane
// reads as one job (job word 0), so its l0 is 64 + the common x/z prefix.
PLH_DEV int64_t i64min(int64_t a, int64_t b) { return a < b ? a : b; }

struct LcpIndex {
    const uint16_t* l0;
    const uint16_t* l1;
    const uint16_t* l2;
    const uint64_t* l3;  // tagged with the pass (bits 16 and up); the minimum in the low 16 bits
    uint64_t n;
    uint64_t tiles;
    PLH_DEV uint32_t l3v(uint64_t u) const { return (uint32_t)(l3[u] & 0xFFFFull); }
};

template <class T>
PLH_DEV T wave_min(T v) {
    return wave_last(wave_scan_op(v, (T)~T(0), [](T a, T b) { return a < b ? a : b; }));
}

// One sum's window, passed by value to the set-up kernel, which lays it out
// as a one-job window table.
template <int AS>
struct WindowArg {
    uint8_t wn;
    uint8_t kind[kJobWindow];
    uint8_t psf[kJobWindow];
    uint64_t p[kJobWindow * 2 * AS];
    double c[kJobWindow];
    double s[kJobWindow];
};

// The windows of a pass: per job from the step tables (jw), or a single
// sum's one window by value (a).
template <int AS, bool SINGLE>
struct WinView {
    JobWindows jw;
    WindowArg<AS> a;
    PLH_DEV uint32_t wn(uint32_t job) const { return SINGLE ? a.wn : jw.wn[job]; }
    PLH_DEV uint32_t kind(uint32_t job, uint32_t e) const {
        return SINGLE ? a.kind[e] : jw.kind[(size_t)job * kJobWindow + e];
    }
    PLH_DEV uint64_t p(uint32_t job, uint32_t e, int k) const {
        return SINGLE ? a.p[e * 2 * AS + k] : jw.p[((size_t)job * kJobWindow + e) * 2 * AS + k];
    }
    PLH_DEV uint32_t psf(uint32_t job, uint32_t e) const {
        return SINGLE ? a.psf[e] : jw.psf[(size_t)job * kJobWindow + e];
    }
    PLH_DEV double c(uint32_t job, uint32_t e) const { return SINGLE ? a.c[e] : jw.c[(size_t)job * kJobWindow + e]; }
    PLH_DEV double s(uint32_t job, uint32_t e) const { return SINGLE ? a.s[e] : jw.s[(size_t)job * kJobWindow + e]; }
};

// Qubit table entry (job, q): the window entries a set bit at qubit q
// anticommutes with, low 16 bits for the term's z bit (entries whose
// generator has X or Y there), high 16 bits for its x bit (Z or Y).  Entries
// of kind 0 (no change) are left out.
template <int AS, bool SINGLE>
PLH_DEV uint32_t qubit_entry(const WinView<AS, SINGLE>& win, uint32_t job, uint32_t q, bool alive) {
    const int wd = (int)(q >> 6), b = (int)(q & 63);
    uint32_t mz = 0, mx = 0;
    const int wn = alive ? (int)win.wn(job) : 0;
    for (int e = 0; e < wn; ++e) {
        if (win.kind(job, (uint32_t)e) != 4) continue;
        if ((win.p(job, (uint32_t)e, wd) >> b) & 1ull) mz |= 1u << e;       // X or Y at q: the term's z bit counts
        if ((win.p(job, (uint32_t)e, AS + wd) >> b) & 1ull) mx |= 1u << e;  // Z or Y at q: the term's x bit counts
    }
    return mz | (mx << 16);
}

// The pass tag of the per-job minima (jmin: tag << 8 | entry) and the
// 16-tile minima (l3: tag << 16 | value): it falls with every pass, so this
// pass's atomic minima win over whatever earlier passes left, and no reset
// is needed; the buffer is filled with ones once, when it is (re)allocated.
constexpr uint64_t kPassTagTop = (1ull << 47) - 1;

// ---- 0. set-up (batched sums): the qubit tables, one thread per (job,
// qubit); the first threads also find the job ranges (off, as k_rs_bounds).
// A term's anticommutation with entry e is the parity of bit e over its set
// bits, so one pass over a term's (few) set bits tests the whole window.
template <int AS>
__global__ void k_rj_setup(JobWindows jw, uint32_t njobs, const uint8_t* __restrict__ alive,
                           uint32_t* __restrict__ tbl, const uint32_t* __restrict__ J, uint32_t n,
                           uint32_t* __restrict__ off) {
    constexpr uint32_t Q = 64u * AS;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g <= njobs) {
        const uint32_t k = (uint32_t)g;
        uint32_t lo = 0, cnt = n;
        while (cnt > 0) {
            const uint32_t half = cnt / 2, t = lo + half;
            if (J[t] < k) {
                lo = t + 1;
                cnt -= half + 1;
            } else {
                cnt = half;
            }
        }
        off[k] = lo;
    }
    if (g >= (uint64_t)njobs * Q) return;
    const uint32_t job = (uint32_t)(g / Q);
    const WinView<AS, false> win{jw, {}};
    tbl[g] = qubit_entry<AS, false>(win, job, (uint32_t)(g % Q), !alive || alive[job]);
}

// ---- 1. per term the first rotation of its job's window it anticommutes
// with (fk, one byte), and per job the least of them (jmin).  A block stages
// the qubit tables of the jobs its tile spans (at most kStageJobs) in LDS;
// each term walks its set bits.
constexpr int kStageJobs = 8;

template <int AS>
PLH_DEV uint32_t first_hit(const uint32_t* t, const uint64_t* xw, const uint64_t* zw) {
    uint32_t par = 0;
#pragma unroll
    for (int w = 0; w < AS; ++w) {
        uint64_t m = xw[w] | zw[w];
        while (m) {
            const int b = __builtin_ctzll(m);
            m &= m - 1;
            const uint32_t e = t[64 * w + b];
            if ((zw[w] >> b) & 1ull) par ^= e & 0xFFFFu;
            if ((xw[w] >> b) & 1ull) par ^= e >> 16;
        }
    }
    return par ? (uint32_t)__builtin_ctz(par) : 0xFFu;
}

// A single sum (SINGLE) builds its one qubit table in LDS from the window.
template <int AS, bool SINGLE, int IPT>
__global__ __launch_bounds__(kRjTile / IPT) void k_rj_first(Sorted<AS, true> S, uint32_t n, const uint32_t* __restrict__ tbl,
                                                     WinView<AS, SINGLE> win, uint8_t* __restrict__ fk,
                                                     uint64_t* __restrict__ jmin, uint16_t* __restrict__ l0,
                                                     uint16_t* __restrict__ l1, uint16_t* __restrict__ l2,
                                                     uint64_t* __restrict__ l3, uint64_t tag) {
    constexpr uint32_t Q = 64u * AS;
    __shared__ uint32_t s_tbl[kStageJobs * Q];
    __shared__ uint32_t s_min[kStageJobs];
    __shared__ uint32_t s_lw[kRjWords];
    __shared__ uint32_t s_j0, s_nj;
    const uint64_t tile = blockIdx.x;
    const uint64_t i0 = tile * kRjTile;
    const uint64_t i1 = (i0 + kRjTile < (uint64_t)n ? i0 + kRjTile : (uint64_t)n) - 1;
    if (threadIdx.x == 0) {
        s_j0 = S.job((uint32_t)i0);
        s_nj = S.job((uint32_t)i1) - s_j0 + 1;
    }
    __syncthreads();
    const uint32_t j0 = s_j0, nj = s_nj;
    const bool staged = nj <= (uint32_t)kStageJobs;
    if constexpr (SINGLE) {
        for (uint32_t t = threadIdx.x; t < Q; t += kRjTile / IPT) s_tbl[t] = qubit_entry<AS, true>(win, 0, t, true);
    } else if (staged) {
        for (uint32_t t = threadIdx.x; t < nj * Q; t += kRjTile / IPT) s_tbl[t] = tbl[(size_t)j0 * Q + t];
    }
    if (threadIdx.x < (unsigned)kStageJobs) s_min[threadIdx.x] = 0xFFu;
    __syncthreads();
    uint32_t job[IPT];
    uint64_t xw[IPT][AS], zw[IPT][AS];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint64_t i = rj_item<IPT>(tile, k);
        job[k] = i < n ? S.job((uint32_t)i) : 0xFFFFFFFFu;
#pragma unroll
        for (int w = 0; w < AS; ++w) {
            xw[k][w] = i < n ? S.x[(int64_t)w * S.stride + i] : 0ull;
            zw[k][w] = i < n ? S.z[(int64_t)w * S.stride + i] : 0ull;
        }
    }
    // common prefix with the previous term (the LcpIndex arrays): the first
    // differing key word decides; l0 is 0 at t = 0 and past n
    {
        const int wv = (int)(threadIdx.x / kWave);
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const uint64_t i = rj_item<IPT>(tile, k);
            uint32_t v = 0;
            if (i < n && i > 0) {
                uint64_t d[2 * AS + 1];
                d[0] = (uint64_t)(job[k] ^ S.job((uint32_t)(i - 1)));
#pragma unroll
                for (int w = 0; w < AS; ++w) {
                    d[1 + w] = xw[k][w] ^ S.x[(int64_t)w * S.stride + i - 1];
                    d[1 + AS + w] = zw[k][w] ^ S.z[(int64_t)w * S.stride + i - 1];
                }
                v = 64u * (2 * AS + 1);
#pragma unroll
                for (int w = 2 * AS; w >= 0; --w)
                    if (d[w]) v = 64u * (uint32_t)w + (uint32_t)__clzll(d[w]);
            }
            l0[i] = (uint16_t)v;
            const uint32_t m = wave_min(v);
            if (lane_id() == 0) {
                l1[i >> 6] = (uint16_t)m;
                s_lw[wv * IPT + k] = m;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t m = s_lw[0];
            for (int w = 1; w < kRjWords; ++w) m = s_lw[w] < m ? s_lw[w] : m;
            l2[tile] = (uint16_t)m;
            atomicMin((unsigned long long*)&l3[tile >> 4], (unsigned long long)((tag << 16) | m));
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint64_t i = rj_item<IPT>(tile, k);
        uint32_t f = 0xFFu;
        if (i < n) {
            const uint32_t jb = job[k];
            f = staged ? first_hit<AS>(s_tbl + (jb - j0) * Q, xw[k], zw[k])
                       : first_hit<AS>(tbl + (size_t)jb * Q, xw[k], zw[k]);
            fk[i] = (uint8_t)f;
        }
        // per-job minimum, reduced in LDS first (one global atomic per job
        // and block: the per-job words would otherwise take an atomic from
        // every wave, all in a few L2 channels)
        const uint32_t jf = (uint32_t)__builtin_amdgcn_readfirstlane((int)job[k]);
        if (ballot(i < n && job[k] != jf) == 0ull) {
            const uint32_t m = wave_min(f);
            if (lane_id() == 0 && m < 0xFFu && jf != 0xFFFFFFFFu) {
                if (staged)
                    atomicMin(&s_min[jf - j0], m);
                else
                    atomicMin((unsigned long long*)&jmin[jf], (unsigned long long)((tag << 8) | m));
            }
        } else if (f < 0xFFu) {
            if (staged)
                atomicMin(&s_min[job[k] - j0], f);
            else
                atomicMin((unsigned long long*)&jmin[job[k]], (unsigned long long)((tag << 8) | f));
        }
    }
    __syncthreads();
    if (staged && threadIdx.x < nj && s_min[threadIdx.x] < 0xFFu)
        atomicMin((unsigned long long*)&jmin[j0 + threadIdx.x], (unsigned long long)((tag << 8) | s_min[threadIdx.x]));
}

// ---- 1b. per job the rotation it applies (the first of its window that
// changes a term: jmin) and the rotations it consumes (adv: up to and
// including that one; the whole window when no rotation of it changes a
// term); the first workgroup writes them as this step's rotation table.
// Every term is anticommuting when its first hit is its job's applied entry
// (every term's first hit is at least that one).
template <int AS, bool SINGLE>
PLH_DEV uint32_t rj_applied(const WinView<AS, SINGLE>& win, const uint64_t* jmin, uint64_t tag, uint32_t job) {
    const uint32_t w = win.wn(job);
    const uint64_t v = jmin[job];
    const uint32_t m = (v >> 8) == tag ? (uint32_t)(v & 0xFFu) : 0xFFu;  // no hit this pass: an older tag
    return (w != 0 && m < w) ? m : 0xFFu;
}

template <int AS, bool SINGLE, int IPT>
__global__ __launch_bounds__(kRjTile / IPT) void k_rj_anti(const uint32_t* __restrict__ J, uint32_t n, uint32_t njobs,
                                                    WinView<AS, SINGLE> win, const uint64_t* __restrict__ jmin,
                                                    uint64_t tag, const uint8_t* __restrict__ fk, JobRotTables jt,
                                                    uint8_t* __restrict__ kstar, uint8_t* __restrict__ adv, BitOut ab) {
    __shared__ uint32_t sh[kRjWords];
    const uint64_t tile = blockIdx.x;
    if (tile == 0) {
        uint8_t* kind = const_cast<uint8_t*>(jt.kind);
        uint64_t* p = const_cast<uint64_t*>(jt.p);
        uint8_t* psf = const_cast<uint8_t*>(jt.psf);
        double* c = const_cast<double*>(jt.c);
        double* sn = const_cast<double*>(jt.s);
        for (uint32_t t = threadIdx.x; t < njobs; t += kRjTile / IPT) {
            const uint32_t w = win.wn(t), m = rj_applied<AS, SINGLE>(win, jmin, tag, t);
            kind[t] = 0;
            kstar[t] = (uint8_t)m;
            adv[t] = (uint8_t)(m == 0xFFu ? w : m + 1);
            if (m == 0xFFu) continue;
            kind[t] = (uint8_t)win.kind(t, m);
#pragma unroll
            for (int k = 0; k < 2 * AS; ++k) p[(size_t)t * 2 * AS + k] = win.p(t, m, k);
            psf[t] = (uint8_t)win.psf(t, m);
            c[t] = win.c(t, m);
            sn[t] = win.s(t, m);
        }
    }
    bool fl[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint64_t i = rj_item<IPT>(tile, k);
        fl[k] = false;
        if (i < n) {
            const uint32_t ks = rj_applied<AS, SINGLE>(win, jmin, tag, J ? J[i] : 0u);
            fl[k] = ks != 0xFFu && fk[i] == ks;
        }
    }
    rj_put<IPT>(fl, tile, ab, sh);
}

// One workgroup of NT threads: exclusive scan of n tile counts (out[n] =
// total), in chunks of NT * 8.  sh: NT / 64 entries of LDS.
template <int NT>
PLH_DEV void rj_block_scan(const uint32_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n, uint64_t* sh) {
    constexpr int kIpt = 8;
    uint64_t carry = 0;
    for (uint64_t c0 = 0; c0 < n; c0 += (uint64_t)NT * kIpt) {
        const uint64_t base = c0 + (uint64_t)threadIdx.x * kIpt;
        uint32_t vals[kIpt];
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < kIpt; ++k) {
            vals[k] = base + k < n ? in[base + k] : 0u;
            acc += vals[k];
        }
        uint64_t tot;
        uint64_t run = carry + block_exclusive_scan<NT / kWave>(acc, tot, sh);
#pragma unroll
        for (int k = 0; k < kIpt; ++k) {
            if (base + k < n) out[base + k] = run;
            run += vals[k];
        }
        carry += tot;
        __syncthreads();  // sh is reused by the next chunk
    }
    if (threadIdx.x == 0) out[n] = carry;
}

// Sums of at most kLocalTiles tiles (2M terms) skip two launches: the new-
// term kernel scans the anticommuting tile counts itself, and the keep
// kernel finds its tile's new terms by a wave-wide search of the lower
// bounds.
constexpr uint64_t kLocalTiles = 2048;

// A wave's lower bound of v in the sorted a[0, m): the first r with
// a[r] >= v (m when none).  Every lane calls it; each round samples 64 evenly
// spaced entries and keeps the gap that holds the answer.
PLH_DEV uint32_t wave_lower_bound(const uint32_t* __restrict__ a, uint32_t m, uint64_t v) {
    uint32_t lo = 0, hi = m;  // the answer lies in [lo, hi]
    const uint32_t k = lane_id();
    while (hi - lo > (uint32_t)kWave) {
        const uint64_t len = hi - lo;
        const uint32_t pk = lo + (uint32_t)(len * k / kWave);
        const uint32_t c = (uint32_t)popc(ballot((uint64_t)a[pk] < v));  // samples below v: a prefix of k
        if (c == 0) return lo;
        const uint32_t plast = lo + (uint32_t)(len * (c - 1) / kWave);
        const uint32_t pnext = c < (uint32_t)kWave ? lo + (uint32_t)(len * c / kWave) : hi;
        lo = plast + 1;
        hi = pnext;
    }
    return lo + (uint32_t)popc(ballot(lo + k < hi && (uint64_t)a[lo + k] < v));
}

// Wave-wide searches on the common-prefix index (every lane calls them with
// the same arguments): one coalesced load and a ballot per level.
PLH_DEV uint64_t hi_bit(uint64_t b) { return 63u - (uint64_t)__clzll(b); }
PLH_DEV uint64_t lo_bit(uint64_t b) { return (uint64_t)__builtin_ctzll(b); }

// the last t in word w with l0[t] < thr (one exists)
PLH_DEV uint64_t wv_word_last(const LcpIndex& L, uint64_t w, uint32_t thr) {
    const uint64_t b = ballot(L.l0[64 * w + lane_id()] < thr);
    return 64 * w + (b ? hi_bit(b) : 0);
}
PLH_DEV uint64_t wv_tile_last(const LcpIndex& L, uint64_t T, uint32_t thr) {
    const uint32_t k = lane_id();
    const uint64_t b = ballot(k < 16 && L.l1[16 * T + k] < thr);
    return wv_word_last(L, 16 * T + (b ? hi_bit(b) : 0), thr);
}
PLH_DEV uint64_t wv_block_last(const LcpIndex& L, uint64_t u, uint32_t thr) {
    const uint32_t k = lane_id();
    const uint64_t T = 16 * u + k;
    const uint64_t b = ballot(k < 16 && T < L.tiles && L.l2[T] < thr);
    return wv_tile_last(L, 16 * u + (b ? hi_bit(b) : 0), thr);
}
// the last t < p with l0[t] < thr (p >= 1; l0[0] = 0 ends every search)
PLH_DEV uint64_t wv_prev_lt(const LcpIndex& L, uint64_t p, uint32_t thr) {
    const uint32_t k = lane_id();
    const uint64_t q = p - 1, w = q >> 6, T = w >> 4;
    uint64_t b = ballot(64 * w + k <= q && L.l0[64 * w + k] < thr);
    if (b) return 64 * w + hi_bit(b);
    b = ballot(k < 16 && (w & ~15ull) + k < w && L.l1[(w & ~15ull) + k] < thr);
    if (b) return wv_word_last(L, (w & ~15ull) + hi_bit(b), thr);
    b = ballot(k < 16 && (T & ~15ull) + k < T && L.l2[(T & ~15ull) + k] < thr);
    if (b) return wv_tile_last(L, (T & ~15ull) + hi_bit(b), thr);
    uint64_t U = T >> 4;  // blocks [0, U) remain
    while (U > 0) {
        const uint64_t u0 = U >= 64 ? U - 64 : 0;
        b = ballot(u0 + k < U && L.l3v(u0 + k) < thr);
        if (b) return wv_block_last(L, u0 + hi_bit(b), thr);
        U = u0;
    }
    return 0;
}
PLH_DEV uint64_t wv_word_first(const LcpIndex& L, uint64_t w, uint32_t thr) {
    const uint64_t b = ballot(L.l0[64 * w + lane_id()] < thr);
    return 64 * w + (b ? lo_bit(b) : 63);
}
PLH_DEV uint64_t wv_tile_first(const LcpIndex& L, uint64_t T, uint32_t thr) {
    const uint32_t k = lane_id();
    const uint64_t b = ballot(k < 16 && L.l1[16 * T + k] < thr);
    return wv_word_first(L, 16 * T + (b ? lo_bit(b) : 15), thr);
}
PLH_DEV uint64_t wv_block_first(const LcpIndex& L, uint64_t u, uint32_t thr) {
    const uint32_t k = lane_id();
    const uint64_t T = 16 * u + k;
    const uint64_t b = ballot(k < 16 && T < L.tiles && L.l2[T] < thr);
    return wv_tile_first(L, b ? 16 * u + lo_bit(b) : L.tiles - 1, thr);
}
// the first t > p with l0[t] < thr, n when there is none (l0 is 0 from n on)
PLH_DEV uint64_t wv_next_lt(const LcpIndex& L, uint64_t p, uint32_t thr) {
    const uint64_t s = p + 1;
    if (s >= L.n) return L.n;
    const uint32_t k = lane_id();
    const uint64_t w = s >> 6, T = w >> 4;
    auto clamp = [&](uint64_t t) { return t < L.n ? t : L.n; };
    uint64_t b = ballot(64 * w + k >= s && L.l0[64 * w + k] < thr);
    if (b) return clamp(64 * w + lo_bit(b));
    b = ballot(k < 16 && (w & ~15ull) + k > w && L.l1[(w & ~15ull) + k] < thr);
    if (b) return clamp(wv_word_first(L, (w & ~15ull) + lo_bit(b), thr));
    b = ballot(k < 16 && (T & ~15ull) + k > T && (T & ~15ull) + k < L.tiles && L.l2[(T & ~15ull) + k] < thr);
    if (b) return clamp(wv_tile_first(L, (T & ~15ull) + lo_bit(b), thr));
    const uint64_t nu = (L.tiles + 15) >> 4;
    for (uint64_t u0 = (T >> 4) + 1; u0 < nu; u0 += 64) {
        b = ballot(u0 + k < nu && L.l3v(u0 + k) < thr);
        if (b) return clamp(wv_block_first(L, u0 + lo_bit(b), thr));
    }
    return L.n;
}

// ---- 2. new terms at their ranks among the new terms (A: anticommuting
// flags).  The rank is new_rank's (the anticommuting count before the term,
// corrected at each set bit g of M by the halves of its group); the group
// bounds come from the common-prefix index, a wave at a time: each wave
// holds 64 consecutive terms, whose group starts (l0 < g) and bit-1 half
// starts (l0 == g) are two ballots, and a group that crosses the wave's edge
// takes one wave-wide search past that edge.  A group never crosses a job
// (l0 < 64 at a job's first term), so every lane tests its own job's g.
// Per new term r, hb[2r], hb[2r + 1] bound its lower bound among the old
// terms: the other half of its first set bit's group (the new key keeps the
// bits above that bit and flips it).
// With a_cnt set (at most kLocalTiles tiles), every workgroup first scans
// the anticommuting tile counts into LDS for its ranks, and workgroup 0
// stores the scan as A's tile prefixes for the later kernels.
template <int AS>
__global__ __launch_bounds__(kBlock) void k_rj_new(Sorted<AS, true> S, uint32_t n, JobRotTables jt, BitRank A,
                                                   LcpIndex L, NewTerms N, uint32_t* __restrict__ hb,
                                                   const uint32_t* __restrict__ a_cnt) {
    constexpr int KW = Sorted<AS, true>::KW;
    constexpr int XO = 1;
    __shared__ uint64_t s_pre[kLocalTiles + 1];
    __shared__ uint64_t s_red[kWavesPerBlock];
    if (a_cnt) {
        rj_block_scan<kBlock>(a_cnt, s_pre, A.tiles, s_red);
        __syncthreads();
        if (blockIdx.x == 0)
            for (uint64_t t = threadIdx.x; t <= A.tiles; t += kBlock) const_cast<uint64_t*>(A.tpre)[t] = s_pre[t];
        A.tpre = s_pre;
    }
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t i0 = i - lane_id();
    const uint32_t lane = lane_id();
    const bool in = i < n;
    const bool act = in && A.test(i);
    if (ballot(act) == 0ull) return;  // wave-uniform
    const uint32_t lz = in ? L.l0[i] : 0u;
    uint64_t kw[KW + 1] = {};
    if (act) S.key((uint32_t)i, kw);
    // this lane's job generator (every term of a rotated job, for the group
    // structure; none for other jobs)
    const uint32_t job = in ? S.job((uint32_t)i) : 0u;
    const bool rot = in && jt.kind[job] == 4;
    uint64_t m[2 * AS];
#pragma unroll
    for (int h = 0; h < 2 * AS; ++h) m[h] = rot ? jt.p[(size_t)job * 2 * AS + h] : 0ull;
    uint32_t w = 0;
#pragma unroll
    for (int h = 0; h < 2 * AS; ++h) w += (uint32_t)popc(m[h]);
    const uint32_t wmax = wave_last(wave_scan_op(w, 0u, [](uint32_t x, uint32_t y) { return x > y ? x : y; }));
    int64_t r = act ? (int64_t)A.rank(i) : 0;
    uint64_t h0 = 0, h1 = 0;
    const uint64_t below = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);  // lanes <= this one
    for (uint32_t k = 0; k < wmax; ++k) {
        const bool has = k < w;
        // the k-th set bit of M, most significant first (key bit g)
        uint32_t g = 0, kk = k;
        bool found = !has;
#pragma unroll
        for (int h = 0; h < 2 * AS; ++h) {
            const uint32_t c = (uint32_t)popc(m[h]);
            if (!found && kk < c) {
                uint64_t v = m[h];
                for (uint32_t t = 0; t < kk; ++t) v &= ~(1ull << (63 - __clzll(v)));
                g = (uint32_t)(64 * (XO + h) + __clzll(v));
                found = true;
            } else if (!found) {
                kk -= c;
            }
        }
        const uint32_t tlt = has ? g : 64u, tle = has ? g + 1 : 64u;  // no set bit: job starts only
        const uint64_t s_lt = ballot(lz < tlt), s_le = ballot(lz < tle);
        const bool one = has && act && bit_at(kw, (int)g);
        // searches past the wave's edges, for the lanes whose group crosses
        // them (lane 0's and lane 63's jobs)
        uint64_t p_le = 0, p_lt = 0, n_le = n, n_lt = n;
        if (ballot(act && has && (s_lt & below) == 0ull)) {
            const uint32_t t_le = __builtin_amdgcn_readfirstlane(tle), t_lt = __builtin_amdgcn_readfirstlane(tlt);
            p_le = wv_prev_lt(L, i0, t_le);
            p_lt = L.l0[p_le] < t_lt ? p_le : wv_prev_lt(L, p_le, t_lt);
        }
        if (ballot(act && has && (s_lt & ~below) == 0ull)) {
            const uint32_t t_le = (uint32_t)__builtin_amdgcn_readlane((int)tle, 63);
            const uint32_t t_lt = (uint32_t)__builtin_amdgcn_readlane((int)tlt, 63);
            n_le = wv_next_lt(L, i0 + 63, t_le);
            n_lt = (n_le >= n || L.l0[n_le] < t_lt) ? n_le : wv_next_lt(L, n_le, t_lt);
        }
        if (!(act && has)) continue;
        uint64_t lo, mid, hi;
        if (one) {
            // the bit-1 half starts at mid (the last t <= i with l0 <= g);
            // the group at lo (the last t < mid with l0 < g, or mid)
            const uint64_t bm = s_le & below;
            mid = bm ? i0 + hi_bit(bm) : p_le;
            if (bm) {
                const uint64_t ml = hi_bit(bm);
                const bool split = ((s_lt >> ml) & 1ull) == 0ull;
                const uint64_t bl = ml ? (s_lt & ((1ull << ml) - 1ull)) : 0ull;
                lo = split ? (bl ? i0 + hi_bit(bl) : p_lt) : mid;
            } else {
                lo = p_lt;
            }
            r -= (int64_t)(A.rank(mid) - A.rank(lo));
            if (k == 0) {
                h0 = lo;
                h1 = mid;
            }
        } else {
            // the bit-1 half starts at mid (the first t > i with l0 <= g)
            // when l0 there is g; the group ends at hi
            const uint64_t bm = s_le & ~below;
            mid = bm ? i0 + lo_bit(bm) : n_le;
            if (bm) {
                const uint64_t ml = lo_bit(bm);
                const bool split = ((s_lt >> ml) & 1ull) == 0ull && mid < n;
                const uint64_t bh = ml == 63 ? 0ull : (s_lt & ~((2ull << ml) - 1ull));
                hi = split ? (bh ? i0 + lo_bit(bh) : n_lt) : mid;
            } else {
                hi = n_lt;
            }
            r += (int64_t)(A.rank(hi) - A.rank(mid));
            if (k == 0) {
                h0 = mid;
                h1 = hi;
            }
        }
    }
    if (!act || w == 0 || r < 0 || (uint64_t)r >= A.total()) return;  // never with sorted distinct input
    const uint64_t* px = jt.p + (size_t)job * 2 * AS;
    const uint64_t* pz = px + AS;
    uint64_t c1 = 0, c2 = 0;
    const uint32_t f0 = S.f[i] & 1u;
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        phase_word(kw[XO + k], kw[XO + AS + k], px[k], pz[k], c1, c2);
        N.x[(int64_t)k * N.cap + r] = kw[XO + k] ^ px[k];
        N.z[(int64_t)k * N.cap + r] = kw[XO + AS + k] ^ pz[k];
    }
    const uint32_t raw = phase_of(f0, jt.psf[job], c1, c2);
    const double nv = (-jt.s[job]) * S.value((uint32_t)i);
    N.v[r] = (raw & 2u) ? -nv : nv;
    N.f[r] = (uint8_t)(raw & 1u);
    N.j[r] = job;
    hb[2 * r] = (uint32_t)h0;
    hb[2 * r + 1] = (uint32_t)h1;
}

// ---- 3. each new term's lower bound among the old terms, inside the range
// hb[2r] .. hb[2r + 1] (the other half of its first set bit's group, whose
// key words before that bit's word equal the new key's), whether it equals
// that old term (D), and whether it is kept as a term of its own (KN: not
// equal to an old term, and passing the filters)
template <int AS, int IPT>
__global__ __launch_bounds__(kRjTile / IPT) void k_rj_lb(Sorted<AS, true> S, uint32_t n, BitRank A, const uint32_t* __restrict__ hb,
                                                  JobRotTables jt, NewTerms N, Filters fl, uint32_t* __restrict__ lb,
                                                  BitOut dout, BitOut knout) {
    constexpr int KW = Sorted<AS, true>::KW;
    constexpr int XO = 1;
    __shared__ uint32_t sh[kRjWords];
    const uint64_t tile = blockIdx.x;
    const uint64_t m = A.total();
    bool dup[IPT], kn[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint64_t r = rj_item<IPT>(tile, k);
        dup[k] = false;
        kn[k] = false;
        if (r < m) {
            uint64_t kw[KW + 1];
            new_key<AS, true>(N, (uint32_t)r, kw);
            const uint32_t job = (uint32_t)kw[0];
            int w0 = KW;  // the word of the generator's first set bit
            for (int h = 2 * AS - 1; h >= 0; --h)
                if (jt.p[(size_t)job * 2 * AS + h]) w0 = XO + h;
            // clamped into the batch, so that no read leaves it
            const uint32_t hi = hb[2 * r + 1] < n ? hb[2 * r + 1] : n;
            uint32_t lo = hb[2 * r] < hi ? hb[2 * r] : hi;
            uint32_t cnt = hi - lo;
            while (cnt > 0) {
                const uint32_t half = cnt / 2, t = lo + half;
                // the probe's words from w0 on, loaded together (one latency
                // per probe), then compared
                uint64_t ow[KW + 1];
#pragma unroll
                for (int w = 0; w <= KW; ++w) ow[w] = w >= w0 ? S.word(t, w) : kw[w];
                bool less = false, decided = false;
#pragma unroll
                for (int w = 0; w <= KW; ++w) {
                    if (!decided && ow[w] != kw[w]) {
                        less = ow[w] < kw[w];
                        decided = true;
                    }
                }
                if (less) {
                    lo = t + 1;
                    cnt -= half + 1;
                } else {
                    cnt = half;
                }
            }
            if (lo < hi) {
                uint64_t ow[KW + 1];
#pragma unroll
                for (int w = 0; w <= KW; ++w) ow[w] = w >= w0 ? S.word(lo, w) : kw[w];
                dup[k] = key_eq<KW>(ow, kw);
            }
            kn[k] = !dup[k] && keep_of<AS, 1>(fl, kw, N.v[r]);
            lb[r] = lo;
        }
    }
    rj_put<IPT>(dup, tile, dout, sh);
    __syncthreads();  // sh is reused
    rj_put<IPT>(kn, tile, knout, sh);
}

// An old term's coefficient after the rotation: cos * v when it
// anticommutes, plus the equal new term's (-sin) * v when there is one
// (acc = 0; acc += old; acc += new, the reference's order).  u = the first
// new term whose lower bound lies above i.
template <int AS>
PLH_DEV double rj_old_value(const Sorted<AS, true>& S, uint32_t i, uint32_t job, const JobRotTables& jt,
                            const BitRank& A, const uint32_t* lb, const BitRank& D, const NewTerms& N, uint32_t r0,
                            uint32_t r1, uint32_t* u_out) {
    double vo = S.value(i);
    if (A.test(i)) vo = vo * jt.c[job];
    const uint32_t u = u32_bound(lb, r0, r1, i, true);
    double acc = 0.0;
    acc += vo;
    if (u > r0 && lb[u - 1] == i && D.test(u - 1)) acc += N.v[u - 1];
    *u_out = u;
    return acc;
}

PLH_DEV bool rj_keep_old(const Filters& fl, uint32_t job, double acc) {
    if ((fl.mode & kDropZero) && acc == 0.0) return false;
    if (fl.mode & kCutoff) {
        const double eps = fl.job_eps ? fl.job_eps[job] : fl.eps;
        if (eps > 0.0 && !(fabs(acc) >= eps)) return false;
    }
    return true;
}

// The first new term of each old tile: tr0[t] = the first r whose lower
// bound lies in tile t or later (t <= tiles; tr0[tiles] = m).  Thread r
// fills the tiles between its predecessor's lower bound and its own.
__global__ void k_rj_tiles(const uint32_t* __restrict__ lb, BitRank A, uint64_t tiles, uint32_t n,
                           uint32_t* __restrict__ tr0) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = A.total();
    if (r > m || r > n) return;
    // lower bounds are at most n (clamped so that the fill stays in tr0)
    auto tile_of = [&](uint32_t b) { return (int64_t)((b < n ? b : n) / kRjTile); };
    const int64_t prev = r > 0 ? tile_of(lb[r - 1]) : -1;
    const int64_t cur = r < m ? tile_of(lb[r]) : (int64_t)tiles;
    for (int64_t t = prev + 1; t <= cur; ++t) tr0[t] = (uint32_t)r;
}

// ---- 4. old terms' keep flags (KO)
// With search set (at most kLocalTiles tiles), the first wave finds the
// tile's new terms itself and stores tr0[tile] (and tr0[tiles]) for the write
// kernel; otherwise k_rj_tiles has filled tr0.
template <int AS, int IPT>
__global__ __launch_bounds__(kRjTile / IPT) void k_rj_keep(Sorted<AS, true> S, uint32_t n, JobRotTables jt, BitRank A,
                                                    const uint32_t* __restrict__ lb, uint32_t* __restrict__ tr0,
                                                    BitRank D, NewTerms N, Filters fl, BitOut koout, bool search) {
    __shared__ uint32_t sh[kRjWords];
    __shared__ uint32_t s_rr[2];
    const uint64_t tile = blockIdx.x;
    if (search) {
        if (threadIdx.x < (unsigned)kWave) {
            const uint64_t tiles = A.tiles;
            const uint32_t m = (uint32_t)A.total();
            const uint32_t r0 = wave_lower_bound(lb, m, tile * kRjTile);
            const uint32_t r1 = wave_lower_bound(lb, m, (tile + 1) * kRjTile);
            if (threadIdx.x == 0) {
                s_rr[0] = r0;
                s_rr[1] = r1;
                tr0[tile] = r0;
                if (tile + 1 == tiles) tr0[tiles] = r1;
            }
        }
        __syncthreads();
    }
    const uint32_t rr[2] = {search ? s_rr[0] : tr0[tile], search ? s_rr[1] : tr0[tile + 1]};
    bool ko[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint64_t i = rj_item<IPT>(tile, k);
        ko[k] = false;
        if (i < n) {
            const uint32_t job = S.job((uint32_t)i);
            if (!fl.job_alive || fl.job_alive[job]) {
                uint32_t u;
                const double acc = rj_old_value<AS>(S, (uint32_t)i, job, jt, A, lb, D, N, rr[0], rr[1], &u);
                ko[k] = rj_keep_old(fl, job, acc);
            }
        }
    }
    rj_put<IPT>(ko, tile, koout, sh);
}

// ---- 4b. the kept new (KN) and kept old (KO) terms' tile scans, then the
// output count KO + KN and each job's consumed rotations, posted to the
// host: every thread fences its stores to host memory before the block's
// barrier, and thread 0 then posts the count (the host polls its sequence).
// A single sum (no adv_host) posts its count, consumed rotations and applied
// window entry (0xFF: none) together.
struct RjPost {
    const uint32_t* kn_cnt;
    uint64_t* kn_pre;
    const uint32_t* ko_cnt;
    uint64_t* ko_pre;
    uint64_t tiles;
    uint64_t extra;
    uint64_t* out;
    const uint8_t* adv;
    const uint8_t* kstar;
    uint32_t njobs;
    uint8_t* adv_host;
    uint64_t* mbox;
    uint64_t seq;
};

// One workgroup, after the scans: the count to *out and the host.
PLH_DEV void rj_post(const RjPost& rp, uint64_t total) {
    if (rp.adv_host)
        for (uint32_t t = threadIdx.x; t < rp.njobs; t += blockDim.x) rp.adv_host[t] = rp.adv[t];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x != 0) return;
    *rp.out = total;
    if (rp.adv_host) {
        post_write(rp.mbox, rp.seq, 1, &total);
    } else {
        const uint64_t v[3] = {total, rp.adv[0], rp.kstar[0]};
        post_write(rp.mbox, rp.seq, 3, v);
    }
}

// One 1024-thread workgroup (the scans cover up to 2^21 tiles).
__global__ __launch_bounds__(kRjTile) void k_rj_counts(RjPost rp) {
    __shared__ uint64_t sh[kRjTile / kWave];
    rj_block_scan<kRjTile>(rp.kn_cnt, rp.kn_pre, rp.tiles, sh);
    rj_block_scan<kRjTile>(rp.ko_cnt, rp.ko_pre, rp.tiles, sh);
    __syncthreads();
    rj_post(rp, rp.ko_pre[rp.tiles] + rp.kn_pre[rp.tiles] + rp.extra);
}

// ---- 5. kept terms to their places: old term i after the kept old terms
// before it and the kept new terms below it; new term r after the kept new
// terms before it and the kept old terms below its lower bound; both after
// the placed terms (toff, when set) of the jobs before their own.  Blocks
// [0, tiles) take old tiles, [tiles, 2 tiles) new ones.  With local set (at
// most kLocalTiles tiles; the output allocated for the most terms it can
// take) every workgroup scans the KO and KN tile counts into dynamic LDS
// itself, and workgroup 0 stores them and posts the count, so the host's
// wait for it overlaps the writes.
template <int AS, int IPT>
__global__ __launch_bounds__(kRjTile / IPT) void k_rj_write(Sorted<AS, true> S, uint32_t n, uint64_t tiles, JobRotTables jt,
                                                     BitRank A, const uint32_t* __restrict__ lb,
                                                     const uint32_t* __restrict__ tr0, BitRank D, BitRank KO,
                                                     BitRank KN, NewTerms N, const uint64_t* __restrict__ d_total,
                                                     const uint32_t* __restrict__ toff,
                                                     uint64_t* __restrict__ ox, uint64_t* __restrict__ oz,
                                                     uint8_t* __restrict__ of, double* __restrict__ ov,
                                                     uint32_t* __restrict__ oj, int64_t os, RjPost rp, bool local) {
    extern __shared__ uint64_t dyn[];
    uint64_t total;
    if (local) {
        uint64_t* s_ko = dyn;
        uint64_t* s_kn = dyn + (tiles + 1);
        uint64_t* s_red = dyn + 2 * (tiles + 1);
        rj_block_scan<kRjTile / IPT>(rp.ko_cnt, s_ko, tiles, s_red);
        rj_block_scan<kRjTile / IPT>(rp.kn_cnt, s_kn, tiles, s_red);
        __syncthreads();
        KO.tpre = s_ko;
        KN.tpre = s_kn;
        total = s_ko[tiles] + s_kn[tiles] + rp.extra;
        if (blockIdx.x == 0) {
            for (uint64_t t = threadIdx.x; t <= tiles; t += kRjTile / IPT) {
                rp.ko_pre[t] = s_ko[t];
                rp.kn_pre[t] = s_kn[t];
            }
            rj_post(rp, total);
        }
    } else {
        total = *d_total;
    }
    const uint32_t m = (uint32_t)A.total();
    if (blockIdx.x < tiles) {
        const uint64_t tile = blockIdx.x;
        const uint32_t rr[2] = {tr0[tile], tr0[tile + 1]};
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const uint64_t i = rj_item<IPT>(tile, k);
            if (i >= n || !KO.test(i)) continue;
            const uint32_t job = S.job((uint32_t)i);
            uint32_t u;
            const double acc = rj_old_value<AS>(S, (uint32_t)i, job, jt, A, lb, D, N, rr[0], rr[1], &u);
            const uint64_t o = KO.rank(i) + KN.rank(u) + (toff ? toff[job] : 0u);
            if (o >= total) continue;  // never for consistent ranks; keeps writes in bounds
#pragma unroll
            for (int w = 0; w < AS; ++w) {
                ox[(int64_t)w * os + o] = S.x[(int64_t)w * S.stride + i];
                oz[(int64_t)w * os + o] = S.z[(int64_t)w * S.stride + i];
            }
            of[o] = S.f[i] & 1u;
            ov[o] = acc;
            if (oj) oj[o] = job;
        }
        return;
    }
    const uint64_t tile = blockIdx.x - tiles;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint64_t r = rj_item<IPT>(tile, k);
        if (r >= m || !KN.test(r)) continue;
        const uint64_t o = KN.rank(r) + KO.rank(lb[r]) + (toff ? toff[N.j[r]] : 0u);
        if (o >= total) continue;
#pragma unroll
        for (int w = 0; w < AS; ++w) {
            ox[(int64_t)w * os + o] = N.x[(int64_t)w * N.cap + r];
            oz[(int64_t)w * os + o] = N.z[(int64_t)w * N.cap + r];
        }
        of[o] = N.f[r];
        ov[o] = N.v[r];
        if (oj) oj[o] = N.j[r];
    }
}

// ---- 5b. a job-ordered sum of other jobs' terms (placed) into the output:
// term t of job j after the kept terms of the rotated jobs before j (their
// kept old terms before joff[j] and kept new terms before the first new
// term of job j) and after the placed terms before it.
template <int AS>
__global__ __launch_bounds__(kBlock) void k_rj_place(SumView P, const uint32_t* __restrict__ PJ, uint32_t pn,
                                                     const uint32_t* __restrict__ joff, BitRank A, BitRank KO,
                                                     BitRank KN, const uint64_t* __restrict__ d_total,
                                                     uint64_t* __restrict__ ox, uint64_t* __restrict__ oz,
                                                     uint8_t* __restrict__ of, double* __restrict__ ov,
                                                     uint32_t* __restrict__ oj, int64_t os) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= pn) return;
    const uint32_t job = PJ[t];
    const uint64_t b = joff[job];
    const uint64_t o = KO.rank(b) + KN.rank(A.rank(b)) + t;
    if (o >= *d_total) return;  // never for consistent ranks; keeps writes in bounds
#pragma unroll
    for (int w = 0; w < AS; ++w) {
        ox[(int64_t)w * os + o] = P.x[(int64_t)w * P.stride + t];
        oz[(int64_t)w * os + o] = P.z[(int64_t)w * P.stride + t];
    }
    of[o] = P.f[t];
    ov[o] = P.v[t];
    oj[o] = job;
}

// Flag arrays of `tiles` tiles: bits, in-tile prefixes, tile counts and
// their scan, carved from one workspace block.
struct FlagBuf {
    uint64_t* bits;
    uint32_t* wpre;
    uint32_t* tcnt;
    uint64_t* tpre;
    static size_t bytes(uint64_t tiles) {
        return up256(tiles * kRjWords * 8) + up256(tiles * kRjWords * 4) + up256(tiles * 4) + up256((tiles + 1) * 8);
    }
    void carve(char*& p, uint64_t tiles) {
        bits = (uint64_t*)p;
        p += up256(tiles * kRjWords * 8);
        wpre = (uint32_t*)p;
        p += up256(tiles * kRjWords * 4);
        tcnt = (uint32_t*)p;
        p += up256(tiles * 4);
        tpre = (uint64_t*)p;
        p += up256((tiles + 1) * 8);
    }
    BitOut out(bool counted = true) const { return BitOut{bits, counted ? wpre : nullptr, counted ? tcnt : nullptr}; }
    BitRank rank(uint64_t tiles) const { return BitRank{bits, wpre, tpre, tiles * kRjTile, tiles}; }
};

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t n, uint64_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// This pass's tag and the pass-tagged per-job minima (jmin, at least njobs)
// and 16-tile minima (l3, at least nl3), kept in their own workspace slot:
// filled with ones whenever the slot or its layout changes, after which
// every pass's tag is below every earlier value.
int rot_pass_state(plh_ctx* ctx, uint64_t njobs, uint64_t nl3, uint64_t** jmin, uint64_t** l3, uint64_t* tag) {
    const uint64_t jcap = std::max<uint64_t>(ctx->rot_jcap, std::max<uint64_t>(256, njobs));
    const uint64_t l3cap = std::max<uint64_t>(ctx->rot_l3cap, std::max<uint64_t>(1ull << 17, nl3));
    const size_t bytes = up256(jcap * 8) + up256(l3cap * 8);
    void* p;
    bool moved = false;
    PLH_TRY(ctx->ws.get(kWsRotE, bytes, &p, &moved));
    if (moved || jcap != ctx->rot_jcap || l3cap != ctx->rot_l3cap) {
        k_fill_u64<<<blocks_of(bytes / 8), kBlock, 0, ctx->stream>>>((uint64_t*)p, bytes / 8, ~0ull);
        PLH_HIP(hipGetLastError());
        ctx->rot_jcap = jcap;
        ctx->rot_l3cap = l3cap;
    }
    if (ctx->rot_epoch + 2 >= kPassTagTop) {
        set_error("sort-free rotation: pass tags exhausted");
        return PLH_ERR_STATE;
    }
    *tag = kPassTagTop - (++ctx->rot_epoch);
    *jmin = (uint64_t*)p;
    *l3 = (uint64_t*)((char*)p + up256(jcap * 8));
    return PLH_OK;
}

// A batched sum (adv_host: each job's consumed rotations), or a single sum
// as one job (no job plane, adv_host null, *consumed set).
template <int AS>
int rotate_sorted_jobs_as(plh_ctx* ctx, plh_sum* s, const JobWindows& jw, const Filters& fl, uint32_t njobs,
                          uint8_t* adv_host, const plh_sum* place, int* consumed = nullptr,
                          const WindowArg<AS>* single = nullptr) {
    hipStream_t st = ctx->stream;
    const uint64_t n = (uint64_t)s->n;
    const uint64_t tiles = (n + kRjTile - 1) / kRjTile;
    Sorted<AS, true> S{s->x, s->z, s->f, s->v, s->j, s->cap, s->as};
    // flags: anticommuting (A), new term equals an old one (D), new term kept
    // (KN), old term kept (KO); per job: its range, least window hit, applied
    // rotation and consumed count; per term its first window hit; the new
    // terms (at most n) and their lower bounds
    const size_t nj = njobs;
    const size_t job_bytes = up256((nj + 1) * 4) + up256(nj) + up256(nj * 2 * AS * 8) + up256(nj) +
                             2 * up256(nj * 8) + up256(nj) + up256(nj) + up256(16);
    // the common-prefix index: l0 per term (whole tiles), l1 per 64 terms,
    // l2 per tile, l3 per 16 tiles (with the per-job minima, in the
    // pass-tagged buffer)
    const uint64_t nl3 = (tiles + 15) / 16;
    const size_t lcp_bytes = up256(tiles * kRjTile * 2) + up256(tiles * kRjWords * 2) + up256(tiles * 2 + 16);
    uint64_t *jmin, *l3, tag;
    PLH_TRY(rot_pass_state(ctx, nj, nl3, &jmin, &l3, &tag));
    void* p;
    PLH_TRY(ctx->ws.get(kWsRotA, 4 * FlagBuf::bytes(tiles) + job_bytes + up256(n) + up256((tiles + 1) * 4) +
                                     up256(nj * 64 * AS * 4) + up256((nj + 1) * 4) + lcp_bytes + up256(n * 8),
                        &p));
    char* c = (char*)p;
    FlagBuf fa, fd, fkn, fko;
    fa.carve(c, tiles);
    fd.carve(c, tiles);
    fkn.carve(c, tiles);
    fko.carve(c, tiles);
    uint32_t* joff = (uint32_t*)c;
    c += up256((nj + 1) * 4);
    uint8_t* e_kind = (uint8_t*)c;
    c += up256(nj);
    uint64_t* e_p = (uint64_t*)c;
    c += up256(nj * 2 * AS * 8);
    uint8_t* e_psf = (uint8_t*)c;
    c += up256(nj);
    double* e_c = (double*)c;
    c += up256(nj * 8);
    double* e_s = (double*)c;
    c += up256(nj * 8);
    uint8_t* kstar = (uint8_t*)c;
    c += up256(nj);
    uint8_t* adv = (uint8_t*)c;
    c += up256(nj);
    uint64_t* d_total = (uint64_t*)c;
    c += up256(16);
    uint8_t* fk = (uint8_t*)c;
    c += up256(n);
    uint32_t* tr0 = (uint32_t*)c;
    c += up256((tiles + 1) * 4);
    uint32_t* qtbl = (uint32_t*)c;
    c += up256(nj * 64 * AS * 4);
    uint32_t* toff = (uint32_t*)c;  // job offsets of the placed terms
    c += up256((nj + 1) * 4);
    uint16_t* l0 = (uint16_t*)c;
    c += up256(tiles * kRjTile * 2);
    uint16_t* l1 = (uint16_t*)c;
    c += up256(tiles * kRjWords * 2);
    uint16_t* l2 = (uint16_t*)c;
    c += up256(tiles * 2 + 16);
    uint32_t* hb = (uint32_t*)c;  // per new term: the range of its lower bound
    const LcpIndex L{l0, l1, l2, l3, n, tiles};
    const uint64_t pn = place ? (uint64_t)place->n : 0;
    const JobRotTables jt{e_kind, e_p, e_psf, e_c, e_s};
    const size_t nb_planes = up256(n * 8 * AS);
    PLH_TRY(ctx->ws.get(kWsRotN, 2 * nb_planes + up256(n * 8) + up256(n) + up256(n * 4) + up256(n * 4), &p));
    NewTerms N;
    N.x = (uint64_t*)p;
    N.z = (uint64_t*)((char*)p + nb_planes);
    N.v = (double*)((char*)p + 2 * nb_planes);
    N.f = (uint8_t*)((char*)N.v + up256(n * 8));
    N.j = (uint32_t*)((char*)N.f + up256(n));
    N.cap = (int64_t)n;
    uint32_t* lb = (uint32_t*)((char*)N.j + up256(n * 4));
    const unsigned gt = (unsigned)tiles, g1 = blocks_of(n);
    if (pn) {
        k_rs_bounds<<<blocks_of((uint64_t)njobs + 1), kBlock, 0, st>>>(place->j, (uint32_t)pn, njobs, toff);
        PLH_HIP(hipGetLastError());
    }
    // small sums (at most kLocalTiles tiles) take one term per thread in
    // 1024-thread workgroups, large ones four per thread in 256
    const bool local = tiles <= kLocalTiles;
    auto pass = [&](auto ipt) -> int {
        constexpr int IPT = decltype(ipt)::value;
        {
            ProfScope ps(ctx, "rot_anti");
            if (single) {
                // one sum: its qubit table is built by each workgroup of the test
                const WinView<AS, true> win{JobWindows{}, *single};
                k_rj_first<AS, true, IPT><<<gt, kRjTile / IPT, 0, st>>>(S, (uint32_t)n, nullptr, win, fk, jmin, l0, l1, l2, l3, tag);
                PLH_HIP(hipGetLastError());
                k_rj_anti<AS, true, IPT><<<gt, kRjTile / IPT, 0, st>>>(s->j, (uint32_t)n, njobs, win, jmin, tag, fk, jt, kstar, adv,
                                                           fa.out());
            } else {
                k_rj_setup<AS><<<blocks_of(std::max<uint64_t>((uint64_t)nj * 64 * AS, nj + 1)), kBlock, 0, st>>>(
                    jw, njobs, fl.job_alive, qtbl, s->j, (uint32_t)n, joff);
                PLH_HIP(hipGetLastError());
                const WinView<AS, false> win{jw, {}};
                k_rj_first<AS, false, IPT><<<gt, kRjTile / IPT, 0, st>>>(S, (uint32_t)n, qtbl, win, fk, jmin, l0, l1, l2, l3, tag);
                PLH_HIP(hipGetLastError());
                k_rj_anti<AS, false, IPT><<<gt, kRjTile / IPT, 0, st>>>(s->j, (uint32_t)n, njobs, win, jmin, tag, fk, jt, kstar, adv,
                                                            fa.out());
            }
            PLH_HIP(hipGetLastError());
        }
        // small sums: the tile scan inside k_rj_new, the tile ranges inside k_rj_keep
        if (!local) PLH_TRY(scan_u32_to_u64(ctx, fa.tcnt, fa.tpre, tiles, fa.tpre + tiles));
        const BitRank A = fa.rank(tiles);
        {
            ProfScope ps(ctx, "rot_new");
            k_rj_new<AS><<<g1, kBlock, 0, st>>>(S, (uint32_t)n, jt, A, L, N, hb, local ? fa.tcnt : nullptr);
            PLH_HIP(hipGetLastError());
            k_rj_lb<AS, IPT><<<gt, kRjTile / IPT, 0, st>>>(S, (uint32_t)n, A, hb, jt, N, fl, lb, fd.out(false), fkn.out());
            PLH_HIP(hipGetLastError());
            if (!local) {
                k_rj_tiles<<<blocks_of(n + 1), kBlock, 0, st>>>(lb, A, tiles, (uint32_t)n, tr0);
                PLH_HIP(hipGetLastError());
            }
        }
        const BitRank D = fd.rank(tiles), KN = fkn.rank(tiles);
        {
            ProfScope ps(ctx, "rot_merge");
            k_rj_keep<AS, IPT><<<gt, kRjTile / IPT, 0, st>>>(S, (uint32_t)n, jt, A, lb, tr0, D, N, fl, fko.out(), local);
            PLH_HIP(hipGetLastError());
        }
        const BitRank KO = fko.rank(tiles);
        Post post;
        PLH_TRY(post_prepare(ctx, &post));
        const RjPost rp{fkn.tcnt, fkn.tpre, fko.tcnt, fko.tpre, tiles, pn, d_total, adv, kstar, njobs, adv_host,
                        post.mbox, post.seq};
        plh_sum tmp;
        tmp.ctx = ctx;
        tmp.as = s->as;
        tmp.nq = s->nq;
        tmp.has_job = s->has_job;
        uint64_t rb[3] = {0, 0, 0xFFu};
        auto launch_write = [&](bool lw) -> int {
            const size_t dyn = lw ? (2 * (tiles + 1) + kRjTile / IPT / kWave) * sizeof(uint64_t) : 0;
            ProfScope ps(ctx, "rot_write", (double)n * (16.0 * s->as + 13.0));
            k_rj_write<AS, IPT><<<2 * gt, kRjTile / IPT, dyn, st>>>(S, (uint32_t)n, tiles, jt, A, lb, tr0, D, KO, KN, N, d_total,
                                                        pn ? toff : nullptr, tmp.x, tmp.z, tmp.f, tmp.v, tmp.j, tmp.cap,
                                                        rp, lw);
            if (pn)
                k_rj_place<AS><<<blocks_of(pn), kBlock, 0, st>>>(
                    SumView{place->x, place->z, place->f, place->v, place->cap}, place->j, (uint32_t)pn, joff, A, KO, KN,
                    d_total, tmp.x, tmp.z, tmp.f, tmp.v, tmp.j, tmp.cap);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                sum_free_storage(&tmp);
                PLH_HIP(e);
            }
            return PLH_OK;
        };
        if (local) {
            // the output sized for every old term, one new term each and the
            // placed terms; the write posts the count
            PLH_TRY(sum_reserve(&tmp, (int64_t)(2 * n + pn), 0));
            PLH_TRY(launch_write(true));
            const int rc = post_wait(ctx, post, adv_host ? 1 : 3, rb);
            if (rc != PLH_OK) {
                sum_free_storage(&tmp);
                return rc;
            }
            if (!adv_host) {
                *consumed = (int)rb[1];
                if (rb[2] == 0xFFu) {  // no rotation of the window changes a term: the sum stays as it is
                    sum_free_storage(&tmp);
                    return PLH_OK;
                }
            }
        } else {
            k_rj_counts<<<1, kRjTile, 0, st>>>(rp);
            PLH_HIP(hipGetLastError());
            PLH_TRY(post_wait(ctx, post, adv_host ? 1 : 3, rb));
            if (!adv_host) {
                *consumed = (int)rb[1];
                if (rb[2] == 0xFFu) return PLH_OK;  // no rotation of the window changes a term: the sum stays as it is
            }
            PLH_TRY(sum_reserve(&tmp, (int64_t)std::max<uint64_t>(rb[0], 1), 0));
            if (rb[0]) PLH_TRY(launch_write(false));
        }
        const uint64_t total = rb[0];
        const int64_t nq = s->nq;
        sum_take_storage(s, &tmp);
        s->n = (int64_t)total;
        s->nq = nq;
        s->canonical = true;
        s->sorted = true;
        return PLH_OK;
    };
    return local ? pass(std::integral_constant<int, 1>{}) : pass(std::integral_constant<int, 4>{});
}

template <int AS, bool JOB>
int rotate_sorted_as(plh_ctx* ctx, plh_sum* s, const RotParams<AS>& rp, const JobRotTables& jt, const Filters& fl,
                     uint32_t njobs) {
    hipStream_t st = ctx->stream;
    const uint64_t n = (uint64_t)s->n;
    Sorted<AS, JOB> S{s->x, s->z, s->f, s->v, s->j, s->cap, s->as};
    // scratch: a (u32 n) and A (u64 n + 1); the new terms (n each plane);
    // keep (u32 2n), lower bounds (u32 2n), old sums (f64 n), C (u64 2n + 1)
    void* p;
    PLH_TRY(ctx->ws.get(kWsRotA, up256(n * 4) + (n + 1) * 8, &p));
    uint32_t* a = (uint32_t*)p;
    uint64_t* A = (uint64_t*)((char*)p + up256(n * 4));
    const size_t nb_planes = up256(n * 8 * AS);
    PLH_TRY(ctx->ws.get(kWsRotN, 2 * nb_planes + up256(n * 8) + up256(n) + up256(n * 4), &p));
    NewTerms N;
    N.x = (uint64_t*)p;
    N.z = (uint64_t*)((char*)p + nb_planes);
    N.v = (double*)((char*)p + 2 * nb_planes);
    N.f = (uint8_t*)((char*)N.v + up256(n * 8));
    N.j = (uint32_t*)((char*)N.f + up256(n));
    N.cap = (int64_t)n;
    PLH_TRY(ctx->ws.get(kWsRotM, up256(n * 8) + up256(n * 8) + up256(n * 8) + up256((2 * n + 1) * 8) +
                                     ((size_t)njobs + 1) * 4, &p));
    uint32_t* keep = (uint32_t*)p;
    uint32_t* lb = (uint32_t*)((char*)p + up256(n * 8));
    double* val = (double*)((char*)p + 2 * up256(n * 8));
    uint64_t* C = (uint64_t*)((char*)p + 3 * up256(n * 8));
    uint32_t* joff = (uint32_t*)((char*)C + up256((2 * n + 1) * 8));
    const unsigned g1 = blocks_of(n), g2 = blocks_of(2 * n);
    if (JOB) {
        k_rs_bounds<<<blocks_of((uint64_t)njobs + 1), kBlock, 0, st>>>(s->j, (uint32_t)n, njobs, joff);
        PLH_HIP(hipGetLastError());
    }
    {
        ProfScope ps(ctx, "rot_anti");
        k_rs_anti<AS, JOB><<<g1, kBlock, 0, st>>>(S, (uint32_t)n, rp, jt, a);
    }
    PLH_HIP(hipGetLastError());
    PLH_TRY(scan_u32_to_u64(ctx, a, A, n, A + n));
    {
        ProfScope ps(ctx, "rot_new");
        k_rs_new<AS, JOB><<<g1, kBlock, 0, st>>>(S, (uint32_t)n, rp, jt, a, A, joff, N);
    }
    {
        ProfScope ps(ctx, "rot_merge");
        k_rs_merge<AS, JOB><<<g2, kBlock, 0, st>>>(S, (uint32_t)n, rp, jt, a, A, joff, N, fl, keep, lb, val);
    }
    PLH_HIP(hipGetLastError());
    uint64_t total = 0;
    PLH_TRY(scan_u32_to_u64_read(ctx, keep, C, 2 * n, C + 2 * n, &total));
    plh_sum tmp;
    tmp.ctx = ctx;
    tmp.as = s->as;
    tmp.nq = s->nq;
    tmp.has_job = s->has_job;
    PLH_TRY(sum_reserve(&tmp, (int64_t)std::max<uint64_t>(total, 1), 0));
    if (total) {
        ProfScope ps(ctx, "rot_write", (double)total * (16.0 * s->as + 9.0));
        k_rs_write<AS, JOB><<<g2, kBlock, 0, st>>>(S, (uint32_t)n, A, N, keep, lb, val, C, s->as, tmp.x, tmp.z,
                                                   tmp.f, tmp.v, tmp.j, tmp.cap);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            sum_free_storage(&tmp);
            PLH_HIP(e);
        }
    }
    const int64_t nq = s->nq;
    sum_take_storage(s, &tmp);
    s->n = (int64_t)total;
    s->nq = nq;
    s->canonical = true;
    s->sorted = true;
    return PLH_OK;
}

// ----------------------------------------------------------------------
// A window of rotations on a single sorted sum (propagate's rotation list,
// a program's run of rotations): the first one that a term anticommutes
// with is applied, and the rotations before it are skipped.  They are exact
// no-ops when the sum has passed the same truncation since its keys last
// changed: a general rotation that no term anticommutes with changes no
// term, a rotation of kind 0 changes nothing, and the truncation after each
// finds nothing to drop.  The sum goes through the batched rotation as one
// job (its job word reads as 0).
constexpr int kSumWindow = 16;

// A window on a single sorted sum through the batched rotation as one job:
// the window test, the applied rotation and its output count in one pass
// with one read-back (the host learns the consumed count with the count).
template <int AS>
int rotate_window_single(plh_ctx* ctx, plh_sum* s, int w, const uint64_t* const* px, const uint64_t* const* pz,
                         const uint8_t* kind, const uint32_t* psf, const double* c, const double* sn,
                         const Filters& fl, int* consumed) {
    WindowArg<AS> a{};
    a.wn = (uint8_t)w;
    for (int e = 0; e < w; ++e) {
        a.kind[e] = kind[e];
        a.psf[e] = (uint8_t)psf[e];
        a.c[e] = c[e];
        a.s[e] = sn[e];
        for (int k = 0; k < s->as; ++k) {
            a.p[e * 2 * AS + k] = px[e][k];
            a.p[e * 2 * AS + AS + k] = pz[e][k];
        }
    }
    return rotate_sorted_jobs_as<AS>(ctx, s, JobWindows{}, fl, 1, nullptr, nullptr, consumed, &a);
}

Filters filters_of(const CollateSpec& spec) {
    return Filters{spec.mode, spec.max_weight, spec.eps, spec.job_eps, spec.job_maxw, spec.job_alive};
}

}  // namespace

int rotate_sorted(plh_ctx* ctx, plh_sum* s, const uint64_t* px, const uint64_t* pz, uint32_t psf, double c,
                  double sn, const CollateSpec& spec) {
    if (!s->sorted || s->has_job || (spec.mode & (kFold | kDropZero)) != (kFold | kDropZero)) {
        set_error("rotate_sorted: needs a sorted canonical single sum and a canonicalizing spec");
        return PLH_ERR_ARG;
    }
    if (s->n > (int64_t)0x7FFFFFFFll) {
        set_error("rotate_sorted: at most 2^31 - 1 terms");
        return PLH_ERR_SIZE;
    }
    if (s->n == 0) return PLH_OK;
    const Filters fl = filters_of(spec);
    switch (s->as <= 1 ? 1 : s->as == 2 ? 2 : s->as <= 4 ? 4 : s->as <= 8 ? 8 : 16) {
#define PLH_RS(ASV)                                                 \
    case ASV: {                                                     \
        RotParams<ASV> rp{};                                        \
        for (int k = 0; k < s->as; ++k) {                           \
            rp.px[k] = px[k];                                       \
            rp.pz[k] = pz[k];                                       \
        }                                                           \
        rp.psf = psf;                                               \
        rp.c = c;                                                   \
        rp.s = sn;                                                  \
        return rotate_sorted_as<ASV, false>(ctx, s, rp, JobRotTables{}, fl, 0); \
    }
        PLH_RS(1)
        PLH_RS(2)
        PLH_RS(4)
        PLH_RS(8)
        PLH_RS(16)
#undef PLH_RS
        default: return PLH_ERR_AS;
    }
}

int rotate_sorted_window(plh_ctx* ctx, plh_sum* s, int w, const uint64_t* const* px, const uint64_t* const* pz,
                         const uint8_t* kind, const uint32_t* psf, const double* c, const double* sn,
                         const CollateSpec& spec, int* consumed) {
    if (w < 1 || w > kSumWindow || !s->sorted || s->has_job || s->as > 2) {
        set_error("rotate_sorted_window: 1..16 rotations on a sorted single sum of at most two words per plane");
        return PLH_ERR_ARG;
    }
    if (s->n > (int64_t)0x7FFFFFFFll) {
        set_error("rotate_sorted_window: at most 2^31 - 1 terms");
        return PLH_ERR_SIZE;
    }
    if (s->n == 0) {
        *consumed = w;
        return PLH_OK;
    }
    const Filters fl = filters_of(spec);
    if (s->as == 1) return rotate_window_single<1>(ctx, s, w, px, pz, kind, psf, c, sn, fl, consumed);
    return rotate_window_single<2>(ctx, s, w, px, pz, kind, psf, c, sn, fl, consumed);
}

int rotate_sorted_jobs(plh_ctx* ctx, plh_sum* s, const JobWindows& jw, int njobs, const CollateSpec& spec,
                       uint8_t* adv_host, const plh_sum* place) {
    if (!s->sorted || !s->has_job || (spec.mode & (kFold | kDropZero)) != (kFold | kDropZero)) {
        set_error("rotate_sorted_jobs: needs a sorted canonical batched sum and a canonicalizing spec");
        return PLH_ERR_ARG;
    }
    if (s->n > (int64_t)0x7FFFFFFFll) {
        set_error("rotate_sorted_jobs: at most 2^31 - 1 terms");
        return PLH_ERR_SIZE;
    }
    if (njobs < 1 || !spec.job_alive || !adv_host) {
        set_error("rotate_sorted_jobs: needs at least one job slot, the job-alive table and the report buffer");
        return PLH_ERR_ARG;
    }
    if (s->n == 0) return PLH_ERR_ARG;  // the caller reports consumed rotations itself for an empty batch
    const Filters fl = filters_of(spec);
    if (place && place->n > 0 && (!place->has_job || place->as != s->as || place->n > (int64_t)0x7FFFFFFFll)) {
        set_error("rotate_sorted_jobs: the placed terms must be a batched sum of the same array size");
        return PLH_ERR_ARG;
    }
    if (s->as == 1) return rotate_sorted_jobs_as<1>(ctx, s, jw, fl, (uint32_t)njobs, adv_host, place);
    if (s->as == 2) return rotate_sorted_jobs_as<2>(ctx, s, jw, fl, (uint32_t)njobs, adv_host, place);
    set_error("rotate_sorted_jobs: batched sums support up to 128 qubits");
    return PLH_ERR_AS;
}

}  // namespace plh
