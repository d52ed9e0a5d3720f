// plh_collate.hip -- collation (sort_and_combine / canonicalize) on gfx950.
//
// What it computes (reference: PauliSum::sort_and_combine,
// include/PauliSumOperations.h:368-461, and canonicalize,
// include/PauliSumUtils.h:314-328): the items are stably sorted by the key
// (x words, z words, flags); each run of equal keys becomes one term whose
// coefficient is acc = 0; acc += v over the run in original-index order;
// canonicalize first folds the sign bit into v and afterwards drops exact
// zeros.  Propagation fuses the weight and cutoff filters
// (include/PauliSumUtils.h:377-393) into the same output pass.
//
// How.  Items are 32-bit ids (see plh_sources.hpp); keys are recomputed
// from the sources, which for products live in L2, so the uncombined
// product is never written.  A stable MSD radix sort on 8-bit digits of the
// live key bits runs in three tiers:
//   1. global passes (tile histogram -> per-bucket column scan -> stable
//      wave-level multi-split scatter) over every bucket larger than kMcap;
//      the first pass enumerates ids implicitly (no id array is read);
//   2. one workgroup per bucket <= kMcap finishes it: stable in-workgroup
//      MSD passes through L2-resident scratch, then LDS leaf sorts of
//      <= LCAP items (counting sort on the next bits + insertion sort on the
//      full (key, id)), counting the terms that survive the filters;
//   3. an exclusive scan of the per-bucket counts gives each bucket's
//      output offset and one workgroup per bucket writes its terms.
// Stability (ids ascend inside every bucket) makes runs of equal keys come
// out in original-index order, so the sums are the reference's sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "plh_block.hpp"
#include "plh_engine.hpp"
#include "plh_sources.hpp"

namespace plh {

namespace {

constexpr int kDigitBits = 8;
constexpr int kDigits = 1 << kDigitBits;
constexpr int kIpt = 8;                          // items per thread per round
constexpr uint32_t kRound = kBlock * kIpt;       // 2048 items per round
constexpr uint32_t kTileExplicit = 8 * kRound;   // 16384
constexpr uint32_t kTileImplicit = 32 * kRound;  // 65536
constexpr uint32_t kMcap = 16384;                // workgroup-finished bucket cap
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int kMaxKeyWords = 33;                 // 2*16 + 1
constexpr int kStack = 256;

template <int AS>
struct LeafDims {
    static constexpr int kCap = AS == 1 ? 2048 : AS == 2 ? 1024 : AS <= 4 ? 512 : AS <= 8 ? 256 : 128;
    static constexpr int kBins = kCap;  // counting-sort bins in a leaf
    static constexpr int kDl = AS == 1 ? 11 : AS == 2 ? 10 : AS <= 4 ? 9 : AS <= 8 ? 8 : 7;
    static constexpr int kPer = (kCap + kBlock - 1) / kBlock;  // slots per thread
};

// Kernel-argument block shared by the sort and emit kernels.
struct Cfg {
    const uint64_t* live;  // live mask, one u64 per key word (MSB-first bits)
    int kb;                // key bits = 64 * (2*AS + 1)
    uint32_t mode;
    uint32_t fmask;        // flags bits in the key: 1 (fold) or 3
    int64_t max_weight;
    double eps;
    int out_as;
    uint64_t* ox;
    uint64_t* oz;
    uint8_t* of;
    double* ov;
    int64_t ostride;
};

PLH_DEV int next_live(const uint64_t* live, int kb, int d) {
    while (d < kb) {
        const int w = d >> 6, o = d & 63;
        const uint64_t m = live[w] & (~0ull >> o);
        if (m) return (w << 6) + __clzll(m);
        d = (w + 1) << 6;
    }
    return kb;
}

inline int next_live_host(const uint64_t* live, int kb, int d) {
    while (d < kb) {
        const int w = d >> 6, o = d & 63;
        const uint64_t m = live[w] & (~0ull >> o);
        if (m) return (w << 6) + __builtin_clzll(m);
        d = (w + 1) << 6;
    }
    return kb;
}

PLH_DEV double fold_value(const Cfg& c, double base, uint32_t raw) {
    return ((c.mode & kFold) && (raw & 2u)) ? -base : base;
}

// ------------------------------------------------------- multi-split
// Stable rank of each item among same-digit items of the workgroup, in the
// item order (round, wave, sub-round, lane); running[] carries per-digit
// counts over the rounds of a tile.
struct MsplitShared {
    uint32_t wave_cnt[kWavesPerBlock][kDigits];
    uint32_t running[kDigits];
};

PLH_DEV void msplit_round(MsplitShared& sh, const uint32_t (&dig)[kIpt], uint32_t (&rank)[kIpt]) {
    const int w = threadIdx.x / kWave;
    uint32_t* flat = &sh.wave_cnt[0][0];
    for (int i = threadIdx.x; i < kWavesPerBlock * kDigits; i += kBlock) flat[i] = 0;
    __syncthreads();
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int s = 0; s < kIpt; ++s) {
        const uint32_t d = dig[s];
        const bool valid = d != kInvalid;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const bool bit = valid && ((d >> b) & 1u);
            const uint64_t bb = ballot(bit);
            m &= bit ? bb : ~bb;
        }
        uint32_t r = 0;
        if (valid) {
            const uint32_t lower = (uint32_t)popc(m & lt);
            const uint32_t base = sh.wave_cnt[w][d];
            r = base + lower;
            if (lower == 0) sh.wave_cnt[w][d] = base + (uint32_t)popc(m);
        }
        rank[s] = r;
    }
    __syncthreads();
    for (int dd = threadIdx.x; dd < kDigits; dd += kBlock) {
        uint32_t run = sh.running[dd];
#pragma unroll
        for (int ww = 0; ww < kWavesPerBlock; ++ww) {
            const uint32_t c = sh.wave_cnt[ww][dd];
            sh.wave_cnt[ww][dd] = run;
            run += c;
        }
        sh.running[dd] = run;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kIpt; ++s)
        if (dig[s] != kInvalid) rank[s] += sh.wave_cnt[w][dig[s]];
    __syncthreads();
}

// Position of item (round base r0, wave w, sub s, lane l) inside a round.
PLH_DEV uint32_t round_pos(int s) {
    return (uint32_t)((threadIdx.x / kWave) * (kWave * kIpt) + s * kWave + lane_id());
}

// Which big bucket a tile belongs to (tile_base is an exclusive scan).
PLH_DEV uint32_t find_bucket(const uint64_t* tile_base, uint32_t nbig, uint64_t t) {
    uint32_t lo = 0, hi = nbig;  // largest b with tile_base[b] <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tile_base[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// --------------------------------------------------------- global pass
template <class Src, bool IMPLICIT>
__global__ __launch_bounds__(kBlock) void k_msd_hist(Src src, const uint32_t* __restrict__ ids,
                                                     const uint64_t* __restrict__ bstart,
                                                     const uint64_t* __restrict__ bsize,
                                                     const uint64_t* __restrict__ tile_base,
                                                     uint32_t nbig, int pos, uint32_t tile_items,
                                                     uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kDigits];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t t = blockIdx.x;
    const uint32_t bi = find_bucket(tile_base, nbig, t);
    const uint64_t lt = t - tile_base[bi];
    const uint64_t beg = bstart[bi] + lt * tile_items;
    const uint64_t end = min(bstart[bi] + bsize[bi], beg + tile_items);
    for (uint64_t i = beg + threadIdx.x; i < end; i += kBlock) {
        uint32_t id;
        if (IMPLICIT) {
            id = src.dense_to_id((uint32_t)i);
            if (!src.valid(id)) continue;
        } else {
            id = ids[i];
        }
        atomicAdd(&h[key_window<kDigitBits>(src, id, pos)], 1u);
    }
    __syncthreads();
    hist[t * kDigits + threadIdx.x] = h[threadIdx.x];
}

// One workgroup per big bucket: per-digit prefix over its tiles (rewritten
// in place), sub-bucket starts and sizes.
__global__ __launch_bounds__(kBlock) void k_colscan(const uint64_t* __restrict__ tile_base,
                                                    uint32_t nbig, uint64_t ntiles,
                                                    const uint64_t* __restrict__ bstart,
                                                    uint32_t* __restrict__ hist,
                                                    uint64_t* __restrict__ sub_start,
                                                    uint64_t* __restrict__ sub_size) {
    __shared__ uint64_t sh[kWavesPerBlock];
    const uint32_t bi = blockIdx.x;
    const int d = threadIdx.x;
    const uint64_t t0 = tile_base[bi];
    const uint64_t t1 = (bi + 1 < nbig) ? tile_base[bi + 1] : ntiles;
    uint64_t run = 0;
    for (uint64_t t = t0; t < t1; ++t) {
        const uint32_t c = hist[t * kDigits + d];
        hist[t * kDigits + d] = (uint32_t)run;
        run += c;
    }
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan(run, tot, sh);
    sub_start[(uint64_t)bi * kDigits + d] = bstart[bi] + ex;
    sub_size[(uint64_t)bi * kDigits + d] = run;
}

template <class Src, bool IMPLICIT>
__global__ __launch_bounds__(kBlock) void k_msd_scatter(
    Src src, const uint32_t* __restrict__ ids_in, uint32_t* __restrict__ ids_out,
    const uint64_t* __restrict__ bstart, const uint64_t* __restrict__ bsize,
    const uint64_t* __restrict__ tile_base, uint32_t nbig, int pos, uint32_t tile_items,
    const uint32_t* __restrict__ hist, const uint64_t* __restrict__ sub_start) {
    __shared__ MsplitShared sh;
    __shared__ uint64_t base[kDigits];
    const uint64_t t = blockIdx.x;
    const uint32_t bi = find_bucket(tile_base, nbig, t);
    const uint64_t lt = t - tile_base[bi];
    const uint64_t beg = bstart[bi] + lt * tile_items;
    const uint64_t end = min(bstart[bi] + bsize[bi], beg + tile_items);
    base[threadIdx.x] = sub_start[(uint64_t)bi * kDigits + threadIdx.x] + hist[t * kDigits + threadIdx.x];
    sh.running[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t r0 = beg; r0 < end; r0 += kRound) {
        uint32_t dig[kIpt], rank[kIpt], idv[kIpt];
#pragma unroll
        for (int s = 0; s < kIpt; ++s) {
            const uint64_t i = r0 + round_pos(s);
            dig[s] = kInvalid;
            idv[s] = 0;
            if (i < end) {
                uint32_t id;
                bool ok = true;
                if (IMPLICIT) {
                    id = src.dense_to_id((uint32_t)i);
                    ok = src.valid(id);
                } else {
                    id = ids_in[i];
                }
                if (ok) {
                    idv[s] = id;
                    dig[s] = key_window<kDigitBits>(src, id, pos);
                }
            }
        }
        msplit_round(sh, dig, rank);
#pragma unroll
        for (int s = 0; s < kIpt; ++s)
            if (dig[s] != kInvalid) ids_out[base[dig[s]] + rank[s]] = idv[s];
    }
}

// ------------------------------------------------------------ bucket lists
// meta: bits 0-15 depth, bit 16 buffer, bit 17 uniform (no live bits left).
PLH_DEV uint32_t meta_depth(uint32_t m) { return m & 0xFFFFu; }
PLH_DEV uint32_t meta_buf(uint32_t m) { return (m >> 16) & 1u; }
PLH_DEV uint32_t meta_uniform(uint32_t m) { return (m >> 17) & 1u; }

__global__ void k_classify(const uint64_t* __restrict__ Lsize, const uint32_t* __restrict__ Lmeta,
                           uint64_t L, uint32_t* __restrict__ isbig) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L) return;
    isbig[e] = (Lsize[e] > kMcap && !meta_uniform(Lmeta[e])) ? 1u : 0u;
}

__global__ void k_gather_big(const uint64_t* __restrict__ Lstart, const uint64_t* __restrict__ Lsize,
                             const uint32_t* __restrict__ Lmeta, const uint32_t* __restrict__ isbig,
                             const uint64_t* __restrict__ bigpos, uint64_t L,
                             uint64_t* __restrict__ bstart, uint64_t* __restrict__ bsize,
                             uint32_t* __restrict__ bmeta, uint32_t* __restrict__ ntiles,
                             uint32_t tile_items) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L || !isbig[e]) return;
    const uint64_t b = bigpos[e];
    bstart[b] = Lstart[e];
    bsize[b] = Lsize[e];
    bmeta[b] = Lmeta[e];
    ntiles[b] = (uint32_t)((Lsize[e] + tile_items - 1) / tile_items);
}

__global__ void k_expand_count(const uint32_t* __restrict__ isbig, const uint64_t* __restrict__ bigpos,
                               const uint64_t* __restrict__ sub_size, uint64_t L,
                               uint32_t* __restrict__ cnt) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L) return;
    if (!isbig[e]) { cnt[e] = 1; return; }
    const uint64_t* ss = sub_size + bigpos[e] * kDigits;
    uint32_t c = 0;
    for (int d = 0; d < kDigits; ++d) c += ss[d] != 0;
    cnt[e] = c;
}

__global__ void k_expand_write(const uint64_t* __restrict__ Lstart, const uint64_t* __restrict__ Lsize,
                               const uint32_t* __restrict__ Lmeta, const uint32_t* __restrict__ isbig,
                               const uint64_t* __restrict__ bigpos, const uint64_t* __restrict__ sub_start,
                               const uint64_t* __restrict__ sub_size, const uint64_t* __restrict__ newpos,
                               uint64_t L, int child_depth, uint32_t child_uniform,
                               uint64_t* __restrict__ Nstart, uint64_t* __restrict__ Nsize,
                               uint32_t* __restrict__ Nmeta) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L) return;
    uint64_t o = newpos[e];
    if (!isbig[e]) {
        Nstart[o] = Lstart[e];
        Nsize[o] = Lsize[e];
        Nmeta[o] = Lmeta[e];
        return;
    }
    const uint32_t buf = meta_buf(Lmeta[e]) ^ 1u;
    const uint64_t* ss = sub_size + bigpos[e] * kDigits;
    const uint64_t* st = sub_start + bigpos[e] * kDigits;
    const uint32_t meta = (uint32_t)child_depth | (buf << 16) | (child_uniform << 17);
    for (int d = 0; d < kDigits; ++d) {
        if (!ss[d]) continue;
        Nstart[o] = st[d];
        Nsize[o] = ss[d];
        Nmeta[o] = meta;
        ++o;
    }
}

// -------------------------------------------------------------- leaves
template <int AS>
struct LeafShared {
    using D = LeafDims<AS>;
    uint64_t key[2 * AS][D::kCap];
    uint32_t id[D::kCap];
    uint32_t bin[D::kBins + 1];
    uint16_t perm[D::kCap];
    uint8_t fl[D::kCap];   // raw flags
    uint8_t head[D::kCap];
    uint32_t bigbin[D::kCap / 32 + 2][2];
    uint32_t nbigbin;
    uint32_t red[kWavesPerBlock];
};

template <int AS>
PLH_DEV uint64_t lds_word(const LeafShared<AS>& L, uint32_t fmask, int i, int w) {
    if (w < 2 * AS) return L.key[w][i];
    if (w == 2 * AS) return (uint64_t)(L.fl[i] & fmask) << 56;
    return 0ull;
}

template <int DB, int AS>
PLH_DEV uint32_t lds_window(const LeafShared<AS>& L, uint32_t fmask, int i, int pos) {
    const int w = pos >> 6, o = pos & 63;
    uint64_t hi = lds_word(L, fmask, i, w) << o;
    if (o + DB > 64) hi |= lds_word(L, fmask, i, w + 1) >> (64 - o);
    return (uint32_t)(hi >> (64 - DB));
}

// Compare slots a, b by (key, masked flags, id): -1, 0 (same key), 1 is not
// needed -- less() includes the id so the order is total.
template <int AS>
PLH_DEV bool slot_less(const LeafShared<AS>& L, uint32_t fmask, int a, int b) {
#pragma unroll
    for (int w = 0; w < 2 * AS; ++w) {
        const uint64_t p = L.key[w][a], q = L.key[w][b];
        if (p != q) return p < q;
    }
    const uint32_t fa = L.fl[a] & fmask, fb = L.fl[b] & fmask;
    if (fa != fb) return fa < fb;
    return L.id[a] < L.id[b];
}

template <int AS>
PLH_DEV bool slot_key_eq(const LeafShared<AS>& L, uint32_t fmask, int a, int b) {
#pragma unroll
    for (int w = 0; w < 2 * AS; ++w)
        if (L.key[w][a] != L.key[w][b]) return false;
    return (L.fl[a] & fmask) == (L.fl[b] & fmask);
}

template <int AS>
PLH_DEV bool keep_term(const Cfg& c, const uint64_t* xz, double acc) {
    if ((c.mode & kDropZero) && acc == 0.0) return false;
    if (c.mode & kMaxWeight) {
        int64_t wgt = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) wgt += popc(xz[k] | xz[AS + k]);
        if (wgt > c.max_weight) return false;
    }
    if ((c.mode & kCutoff) && !(fabs(acc) >= c.eps)) return false;
    return true;
}

template <int AS>
PLH_DEV bool keep_slot(const Cfg& c, const LeafShared<AS>& L, int slot, double acc) {
    uint64_t xz[2 * AS];
#pragma unroll
    for (int w = 0; w < 2 * AS; ++w) xz[w] = L.key[w][slot];
    return keep_term<AS>(c, xz, acc);
}

// Sort n <= kCap ids by (key, id), write them to dst (may equal src) and
// return how many output terms the sorted run produces.
template <int AS, class Src>
PLH_DEV uint32_t leaf_sort(const Src& src, const Cfg& c, const uint32_t* src_ids,
                           uint32_t* dst_ids, uint32_t n, int depth, LeafShared<AS>& L) {
    using D = LeafDims<AS>;
    const int tid = threadIdx.x;
    for (uint32_t i = tid; i < n; i += kBlock) {
        const uint32_t id = src_ids[i];
        uint64_t xz[2 * AS];
        uint32_t fl;
        src.key(id, xz, fl);
        L.id[i] = id;
#pragma unroll
        for (int w = 0; w < 2 * AS; ++w) L.key[w][i] = xz[w];
        L.fl[i] = (uint8_t)fl;
    }
    for (int b = tid; b < D::kBins; b += kBlock) L.bin[b] = 0;
    if (tid == 0) L.nbigbin = 0;
    __syncthreads();
    const int pos = next_live(c.live, c.kb, depth);
    constexpr int kPer = D::kPer;
    uint32_t dig[kPer], slot[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t i = tid + k * kBlock;
        dig[k] = 0;
        slot[k] = 0;
        if (i < n) {
            dig[k] = pos < c.kb ? lds_window<D::kDl, AS>(L, c.fmask, (int)i, pos) : 0u;
            slot[k] = atomicAdd(&L.bin[dig[k]], 1u);
        }
    }
    __syncthreads();
    {  // exclusive scan of the bins (blocked, kBins/kBlock per thread)
        constexpr int kBper = D::kBins / kBlock > 0 ? D::kBins / kBlock : 1;
        uint32_t v[kBper];
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < kBper; ++k) {
            const int b = tid * kBper + k;
            v[k] = b < D::kBins ? L.bin[b] : 0u;
            acc += v[k];
        }
        uint32_t tot;
        uint32_t run = block_exclusive_scan(acc, tot, L.red);
#pragma unroll
        for (int k = 0; k < kBper; ++k) {
            const int b = tid * kBper + k;
            if (b < D::kBins) L.bin[b] = run;
            run += v[k];
        }
        if (tid == 0) L.bin[D::kBins] = n;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t i = tid + k * kBlock;
        if (i < n) L.perm[L.bin[dig[k]] + slot[k]] = (uint16_t)i;
    }
    __syncthreads();
    // Small bins: insertion sort by one thread; large bins: rank sort below.
    for (int b = tid; b < D::kBins; b += kBlock) {
        const uint32_t s0 = L.bin[b], s1 = L.bin[b + 1];
        const uint32_t sz = s1 - s0;
        if (sz >= 2 && sz <= 32) {
            for (uint32_t p = s0 + 1; p < s1; ++p) {
                const uint16_t cur = L.perm[p];
                uint32_t q = p;
                while (q > s0 && slot_less<AS>(L, c.fmask, cur, L.perm[q - 1])) {
                    L.perm[q] = L.perm[q - 1];
                    --q;
                }
                L.perm[q] = cur;
            }
        } else if (sz > 32) {
            const uint32_t k = atomicAdd(&L.nbigbin, 1u);
            L.bigbin[k][0] = s0;
            L.bigbin[k][1] = sz;
        }
    }
    __syncthreads();
    const uint32_t nbb = L.nbigbin;
    uint16_t* tmp = reinterpret_cast<uint16_t*>(L.bin);  // bins no longer needed
    for (uint32_t bb = 0; bb < nbb; ++bb) {
        const uint32_t s0 = L.bigbin[bb][0], sz = L.bigbin[bb][1];
        for (uint32_t p = tid; p < sz; p += kBlock) {
            const uint16_t me = L.perm[s0 + p];
            uint32_t r = 0;
            for (uint32_t q = 0; q < sz; ++q) r += slot_less<AS>(L, c.fmask, L.perm[s0 + q], me) ? 1u : 0u;
            tmp[r] = me;
        }
        __syncthreads();
        for (uint32_t p = tid; p < sz; p += kBlock) L.perm[s0 + p] = tmp[p];
        __syncthreads();
    }
    for (uint32_t p = tid; p < n; p += kBlock)
        L.head[p] = (p == 0) || !slot_key_eq<AS>(L, c.fmask, L.perm[p], L.perm[p - 1]);
    __syncthreads();
    uint32_t kept = 0;
    for (uint32_t p = tid; p < n; p += kBlock) {
        if (!L.head[p]) continue;
        double acc = 0.0;
        uint32_t q = p;
        do {
            const int sl = L.perm[q];
            acc += fold_value(c, src.base_value(L.id[sl]), L.fl[sl]);
            ++q;
        } while (q < n && !L.head[q]);
        kept += keep_slot<AS>(c, L, L.perm[p], acc) ? 1u : 0u;
    }
    const uint32_t total_kept = block_sum(kept, L.red);
    for (uint32_t p = tid; p < n; p += kBlock) dst_ids[p] = L.id[L.perm[p]];
    __syncthreads();
    return total_kept;
}

// Run of identical keys already in id order: sequential sum, 0 or 1 term.
template <int AS, class Src>
PLH_DEV uint32_t uniform_count(const Src& src, const Cfg& c, const uint32_t* ids, uint64_t n,
                               uint32_t* red) {
    __shared__ uint32_t res;
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t id = ids[i];
            acc += fold_value(c, src.base_value(id), src.raw_flags(id));
        }
        uint64_t xz[2 * AS];
        uint32_t fl;
        src.key(ids[0], xz, fl);
        res = keep_term<AS>(c, xz, acc) ? 1u : 0u;
    }
    __syncthreads();
    const uint32_t r = res;
    __syncthreads();
    (void)red;
    return r;
}

struct StackEnt {
    uint16_t start;   // relative to the bucket start (bucket <= kMcap = 2^14)
    uint16_t size;    // 1..kMcap (kMcap itself stored as 0)
    uint16_t depth;
    uint16_t other;   // ids live in the other buffer
};

template <int AS>
struct SortShared {
    union U {
        LeafShared<AS> leaf;
        struct {
            MsplitShared ms;
            uint32_t hist[kDigits];
            uint32_t red[kWavesPerBlock];
        } msd;
    } u;
    StackEnt stk[kStack];
    uint16_t sub_start[kDigits];
    uint16_t sub_size[kDigits];  // 0..kMcap (kMcap stored as 0 never occurs: >= 2 digits or uniform)
    int sp;
    uint32_t kept;
};

// Stable in-workgroup MSD pass: ids of [0, n) from cur to nxt, grouped by
// the 8-bit window at pos; sub-bucket offsets/sizes into S.sub_*.
template <int AS, class Src>
PLH_DEV void wg_msd_pass(const Src& src, const uint32_t* cur, uint32_t* nxt, uint32_t n, int pos,
                         SortShared<AS>& S) {
    auto& M = S.u.msd;
    M.hist[threadIdx.x] = 0;
    M.ms.running[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock)
        atomicAdd(&M.hist[key_window<kDigitBits>(src, cur[i], pos)], 1u);
    __syncthreads();
    uint32_t tot;
    const uint32_t cnt = M.hist[threadIdx.x];
    const uint32_t ex = block_exclusive_scan(cnt, tot, M.red);
    S.sub_start[threadIdx.x] = (uint16_t)ex;
    S.sub_size[threadIdx.x] = (uint16_t)cnt;
    __syncthreads();
    for (uint32_t r0 = 0; r0 < n; r0 += kRound) {
        uint32_t dig[kIpt], rank[kIpt], idv[kIpt];
#pragma unroll
        for (int s = 0; s < kIpt; ++s) {
            const uint32_t i = r0 + round_pos(s);
            dig[s] = kInvalid;
            idv[s] = 0;
            if (i < n) {
                idv[s] = cur[i];
                dig[s] = key_window<kDigitBits>(src, idv[s], pos);
            }
        }
        msplit_round(M.ms, dig, rank);
#pragma unroll
        for (int s = 0; s < kIpt; ++s)
            if (dig[s] != kInvalid) nxt[S.sub_start[dig[s]] + rank[s]] = idv[s];
    }
    __syncthreads();
}

template <int AS, class Src>
__global__ __launch_bounds__(kBlock) void k_bucket_sort(Src src, Cfg c,
                                                        const uint64_t* __restrict__ Lstart,
                                                        const uint64_t* __restrict__ Lsize,
                                                        const uint32_t* __restrict__ Lmeta,
                                                        uint32_t* ids0, uint32_t* ids1,
                                                        uint32_t* __restrict__ kept_out) {
    using D = LeafDims<AS>;
    __shared__ SortShared<AS> S;
    const uint64_t e = blockIdx.x;
    const uint64_t start = Lstart[e];
    const uint64_t size = Lsize[e];
    const uint32_t meta = Lmeta[e];
    uint32_t* home = (meta_buf(meta) ? ids1 : ids0) + start;
    uint32_t* other = (meta_buf(meta) ? ids0 : ids1) + start;
    uint32_t kept = 0;
    if (meta_uniform(meta)) {
        kept = uniform_count<AS>(src, c, home, size, S.u.msd.red);
    } else if (size <= (uint64_t)D::kCap) {
        kept = leaf_sort<AS>(src, c, home, home, (uint32_t)size, (int)meta_depth(meta), S.u.leaf);
    } else {
        if (threadIdx.x == 0) {
            S.sp = 1;
            S.stk[0] = StackEnt{0, (uint16_t)(size & 0xFFFF), (uint16_t)meta_depth(meta), 0};
            S.kept = 0;
        }
        __syncthreads();
        while (S.sp > 0) {
            const StackEnt ent = S.stk[S.sp - 1];
            __syncthreads();
            if (threadIdx.x == 0) S.sp -= 1;
            const uint32_t esz = ent.size ? ent.size : kMcap;
            const uint32_t* cur = (ent.other ? other : home) + ent.start;
            uint32_t* nxt = (ent.other ? home : other) + ent.start;
            const int pos = next_live(c.live, c.kb, ent.depth);
            if (pos >= c.kb) {
                const uint32_t k = uniform_count<AS>(src, c, cur, esz, S.u.msd.red);
                if (ent.other)
                    for (uint32_t i = threadIdx.x; i < esz; i += kBlock) home[ent.start + i] = cur[i];
                if (threadIdx.x == 0) S.kept += k;
                __syncthreads();
                continue;
            }
            wg_msd_pass<AS>(src, cur, nxt, esz, pos, S);
            // nxt holds the sub-buckets; leaves go back to home.
            uint32_t g0 = 0, gn = 0;
            for (int d = 0; d < kDigits; ++d) {
                const uint32_t sz = S.sub_size[d];
                if (sz == 0) continue;
                const uint32_t so = S.sub_start[d];
                const bool big = sz > (uint32_t)D::kCap;
                if (gn > 0 && (big || gn + sz > (uint32_t)D::kCap)) {
                    const uint32_t k = leaf_sort<AS>(src, c, nxt + g0, home + ent.start + g0, gn, pos, S.u.leaf);
                    if (threadIdx.x == 0) S.kept += k;
                    gn = 0;
                }
                if (big) {
                    if (threadIdx.x == 0) {
                        S.stk[S.sp] = StackEnt{(uint16_t)(ent.start + so), (uint16_t)(sz & 0xFFFF),
                                               (uint16_t)(pos + kDigitBits), (uint16_t)(ent.other ^ 1)};
                        S.sp += 1;
                    }
                    continue;
                }
                if (gn == 0) g0 = so;
                gn += sz;
            }
            if (gn > 0) {
                const uint32_t k = leaf_sort<AS>(src, c, nxt + g0, home + ent.start + g0, gn, pos, S.u.leaf);
                if (threadIdx.x == 0) S.kept += k;
            }
            __syncthreads();
        }
        kept = S.kept;
    }
    if (threadIdx.x == 0) kept_out[e] = kept;
}

// ----------------------------------------------------------------- emit
template <int AS, class Src>
__global__ __launch_bounds__(kBlock) void k_emit(Src src, Cfg c, const uint64_t* __restrict__ Lstart,
                                                 const uint64_t* __restrict__ Lsize,
                                                 const uint32_t* __restrict__ Lmeta,
                                                 const uint32_t* __restrict__ ids0,
                                                 const uint32_t* __restrict__ ids1,
                                                 const uint64_t* __restrict__ out_off) {
    using D = LeafDims<AS>;
    constexpr int kCh = D::kCap;
    constexpr int kPer = D::kPer;
    __shared__ LeafShared<AS> L;  // key/id/fl/head used; perm unused
    // Open (not yet closed) run carried between chunks, double-buffered so
    // the chunk that reads the old state never races the one writing it.
    __shared__ uint64_t okey[2][2 * AS];
    __shared__ uint32_t ofl[2];
    __shared__ double osum[2];
    __shared__ int has_open, new_open;
    __shared__ uint64_t red64[kWavesPerBlock];
    const uint64_t e = blockIdx.x;
    const uint64_t size = Lsize[e];
    const uint32_t meta = Lmeta[e];
    const uint32_t* ids = (meta_buf(meta) ? ids1 : ids0) + Lstart[e];
    uint64_t out_pos = out_off[e];
    const int tid = threadIdx.x;
    if (tid == 0) has_open = 0;
    __syncthreads();
    int ob = 0;  // buffer holding the open run of the previous chunk
    for (uint64_t c0 = 0; c0 < size; c0 += kCh, ob ^= 1) {
        const uint32_t m = (uint32_t)min((uint64_t)kCh, size - c0);
        const bool last_chunk = c0 + m == size;
        for (uint32_t i = tid; i < m; i += kBlock) {
            const uint32_t id = ids[c0 + i];
            uint64_t xz[2 * AS];
            uint32_t fl;
            src.key(id, xz, fl);
            L.id[i] = id;
#pragma unroll
            for (int w = 0; w < 2 * AS; ++w) L.key[w][i] = xz[w];
            L.fl[i] = (uint8_t)fl;
        }
        if (tid == 0) new_open = 0;
        __syncthreads();
        for (uint32_t i = tid; i < m; i += kBlock) {
            bool h;
            if (i == 0) {
                h = !has_open;
                if (!h) {
                    bool eq = (L.fl[0] & c.fmask) == ofl[ob];
#pragma unroll
                    for (int w = 0; w < 2 * AS; ++w) eq = eq && (L.key[w][0] == okey[ob][w]);
                    h = !eq;
                }
            } else {
                h = !slot_key_eq<AS>(L, c.fmask, (int)i, (int)i - 1);
            }
            L.head[i] = h ? 1 : 0;
        }
        __syncthreads();
        // Blocked ownership: thread t owns segment starts in [t*kPer, t*kPer+kPer).
        double sums[kPer];
        uint16_t starts[kPer];
        int nk = 0;
        const uint32_t p0 = (uint32_t)tid * kPer;
        for (int k = 0; k < kPer; ++k) {
            const uint32_t p = p0 + k;
            if (p >= m) break;
            const bool cont = (p == 0) && !L.head[0];
            if (!L.head[p] && !cont) continue;
            double acc = cont ? osum[ob] : 0.0;
            uint32_t q = p;
            do {
                acc += fold_value(c, src.base_value(L.id[q]), L.fl[q]);
                ++q;
            } while (q < m && !L.head[q]);
            if (q == m && !last_chunk) {
                // Still open: carried into the next chunk.
#pragma unroll
                for (int w = 0; w < 2 * AS; ++w) okey[ob ^ 1][w] = L.key[w][p];
                ofl[ob ^ 1] = L.fl[p] & c.fmask;
                osum[ob ^ 1] = acc;
                new_open = 1;
                continue;
            }
            if (keep_slot<AS>(c, L, (int)p, acc)) {
                sums[nk] = acc;
                starts[nk] = (uint16_t)p;
                ++nk;
            }
        }
        uint64_t tot;
        const uint64_t ex = block_exclusive_scan((uint64_t)nk, tot, red64);
        for (int k = 0; k < nk; ++k) {
            const uint64_t o = out_pos + ex + k;
            const int p = starts[k];
            for (int w = 0; w < c.out_as; ++w) {
                c.ox[(int64_t)w * c.ostride + o] = L.key[w][p];
                c.oz[(int64_t)w * c.ostride + o] = L.key[AS + w][p];
            }
            c.of[o] = (uint8_t)(L.fl[p] & c.fmask);
            c.ov[o] = sums[k];
        }
        out_pos += tot;
        __syncthreads();
        if (tid == 0) has_open = new_open;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ driver
struct Lists {
    uint64_t* start;
    uint64_t* size;
    uint32_t* meta;
};

int get_lists(plh_ctx* ctx, int slot, uint64_t cap, Lists* out) {
    void* p = nullptr;
    const size_t bytes = (size_t)cap * (2 * sizeof(uint64_t) + sizeof(uint32_t)) + 64;
    PLH_TRY(ctx->ws.get(slot, bytes, &p));
    char* b = (char*)p;
    out->start = (uint64_t*)b;
    out->size = (uint64_t*)(b + (size_t)cap * 8);
    out->meta = (uint32_t*)(b + (size_t)cap * 16);
    return PLH_OK;
}

inline unsigned grid1(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <int AS, class Src>
int collate_run(plh_ctx* ctx, const Src& src, uint64_t n_dense, const uint64_t* h_live,
                uint32_t mode, int64_t max_weight, double eps, plh_sum* out, int out_as) {
    hipStream_t st = ctx->stream;
    constexpr int kWordsK = 2 * AS + 1;
    const int kb = 64 * kWordsK;
    if (n_dense == 0) {
        out->n = 0;
        return PLH_OK;
    }
    if (n_dense > 0xFFFFFFFFull) {
        set_error("collation supports at most 2^32 - 1 items");
        return PLH_ERR_SIZE;
    }
    void* p;
    PLH_TRY(ctx->ws.get(kWsIds0, n_dense * 4, &p));
    uint32_t* ids0 = (uint32_t*)p;
    PLH_TRY(ctx->ws.get(kWsIds1, n_dense * 4, &p));
    uint32_t* ids1 = (uint32_t*)p;
    PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
    uint64_t* misc = (uint64_t*)p;  // [0..32] live, [40] total, [48..52] root
    uint64_t* d_live = misc;
    uint64_t* d_total = misc + 40;
    uint64_t* root_start = misc + 48;
    uint64_t* root_size = misc + 49;
    uint64_t* root_tiles = misc + 50;  // tile_base of the single root bucket
    uint64_t hroot[3] = {0, n_dense, 0};
    PLH_HIP(hipMemcpyAsync(d_live, h_live, sizeof(uint64_t) * kWordsK, hipMemcpyHostToDevice, st));
    PLH_HIP(hipMemcpyAsync(root_start, hroot, sizeof(hroot), hipMemcpyHostToDevice, st));

    Cfg cfg{};
    cfg.live = d_live;
    cfg.kb = kb;
    cfg.mode = mode;
    cfg.fmask = (mode & kFold) ? 1u : 3u;
    cfg.max_weight = max_weight;
    cfg.eps = eps;
    cfg.out_as = out_as;

    // ---- pass 0: implicit ids, one root bucket.
    int depth = 0;
    int pos = next_live_host(h_live, kb, depth);
    if (pos >= kb) pos = 0;  // all keys zero: one bucket, still sorted by id below
    const uint64_t t0 = (n_dense + kTileImplicit - 1) / kTileImplicit;
    PLH_TRY(ctx->ws.get(kWsHist, t0 * kDigits * 4, &p));
    uint32_t* hist = (uint32_t*)p;
    PLH_TRY(ctx->ws.get(kWsSub, 2 * kDigits * 8, &p));
    uint64_t* sub_start = (uint64_t*)p;
    uint64_t* sub_size = sub_start + kDigits;
    {
        ProfScope ps(ctx, "msd_hist_first");
        k_msd_hist<Src, true><<<dim3((unsigned)t0), kBlock, 0, st>>>(src, nullptr, root_start, root_size,
                                                                     root_tiles, 1u, pos, kTileImplicit, hist);
    }
    {
        ProfScope ps(ctx, "colscan");
        k_colscan<<<1, kBlock, 0, st>>>(root_tiles, 1u, t0, root_start, hist, sub_start, sub_size);
    }
    {
        ProfScope ps(ctx, "msd_scatter_first", 4.0 * (double)n_dense);
        k_msd_scatter<Src, true><<<dim3((unsigned)t0), kBlock, 0, st>>>(
            src, nullptr, ids0, root_start, root_size, root_tiles, 1u, pos, kTileImplicit, hist, sub_start);
    }
    PLH_HIP(hipGetLastError());
    depth = pos + kDigitBits;

    // Initial list: the non-empty root sub-buckets (buffer 0).
    std::vector<uint64_t> hsz(kDigits), hst(kDigits);
    PLH_HIP(hipMemcpyAsync(hsz.data(), sub_size, kDigits * 8, hipMemcpyDeviceToHost, st));
    PLH_HIP(hipMemcpyAsync(hst.data(), sub_start, kDigits * 8, hipMemcpyDeviceToHost, st));
    PLH_HIP(hipStreamSynchronize(st));
    const uint32_t uni0 = next_live_host(h_live, kb, depth) >= kb ? 1u : 0u;
    std::vector<uint64_t> ls, lz;
    std::vector<uint32_t> lm;
    for (int d = 0; d < kDigits; ++d) {
        if (!hsz[d]) continue;
        ls.push_back(hst[d]);
        lz.push_back(hsz[d]);
        lm.push_back((uint32_t)depth | (0u << 16) | (uni0 << 17));
    }
    uint64_t n_valid = 0;
    for (uint64_t v : lz) n_valid += v;
    uint64_t L = ls.size();
    if (L == 0) {
        out->n = 0;
        return PLH_OK;
    }
    Lists A, B;
    PLH_TRY(get_lists(ctx, kWsListA, L, &A));
    PLH_HIP(hipMemcpyAsync(A.start, ls.data(), L * 8, hipMemcpyHostToDevice, st));
    PLH_HIP(hipMemcpyAsync(A.size, lz.data(), L * 8, hipMemcpyHostToDevice, st));
    PLH_HIP(hipMemcpyAsync(A.meta, lm.data(), L * 4, hipMemcpyHostToDevice, st));
    int cur_slot = kWsListA;

    // ---- global passes over buckets larger than kMcap.
    for (int guard = 0; guard < 64; ++guard) {
        PLH_TRY(ctx->ws.get(kWsFlags, L * 4 + 64, &p));
        uint32_t* isbig = (uint32_t*)p;
        PLH_TRY(ctx->ws.get(kWsOffs, L * 8 + 64, &p));
        uint64_t* bigpos = (uint64_t*)p;
        k_classify<<<grid1(L), kBlock, 0, st>>>(A.size, A.meta, L, isbig);
        PLH_TRY(scan_u32_to_u64(ctx, isbig, bigpos, L, d_total));
        uint64_t nbig = 0;
        PLH_TRY(read_u64(ctx, d_total, &nbig));
        if (nbig == 0) break;
        pos = next_live_host(h_live, kb, depth);
        PLH_TRY(ctx->ws.get(kWsBig, nbig * (8 + 8 + 4 + 4 + 8) + 64, &p));
        uint64_t* bstart = (uint64_t*)p;
        uint64_t* bsize = bstart + nbig;
        uint64_t* tbase = bsize + nbig;
        uint32_t* bmeta = (uint32_t*)(tbase + nbig);
        uint32_t* ntiles = bmeta + nbig;
        k_gather_big<<<grid1(L), kBlock, 0, st>>>(A.start, A.size, A.meta, isbig, bigpos, L, bstart,
                                                  bsize, bmeta, ntiles, kTileExplicit);
        PLH_TRY(scan_u32_to_u64(ctx, ntiles, tbase, nbig, d_total));
        uint64_t T = 0;
        PLH_TRY(read_u64(ctx, d_total, &T));
        uint64_t pass_items = 0;  // ids moved by this pass (for the timing records)
        if (ctx->prof) {
            PLH_TRY(ctx->ws.get(kWsScan1, nbig * 8 + 64, &p));
            PLH_TRY(scan_u64(ctx, bsize, (uint64_t*)p, nbig, d_total));
            PLH_TRY(read_u64(ctx, d_total, &pass_items));
        }
        PLH_TRY(ctx->ws.get(kWsHist, T * kDigits * 4, &p));
        hist = (uint32_t*)p;
        PLH_TRY(ctx->ws.get(kWsSub, nbig * kDigits * 16, &p));
        sub_start = (uint64_t*)p;
        sub_size = sub_start + nbig * kDigits;
        // All big buckets of a pass sit in the same buffer (they were all
        // produced by the previous pass).
        uint32_t hmeta = 0;
        PLH_HIP(hipMemcpyAsync(&hmeta, bmeta, 4, hipMemcpyDeviceToHost, st));
        PLH_HIP(hipStreamSynchronize(st));
        const uint32_t* in_ids = ((hmeta >> 16) & 1u) ? ids1 : ids0;
        uint32_t* out_ids = ((hmeta >> 16) & 1u) ? ids0 : ids1;
        {
            ProfScope ps(ctx, "msd_hist", 4.0 * (double)pass_items);
            k_msd_hist<Src, false><<<dim3((unsigned)T), kBlock, 0, st>>>(src, in_ids, bstart, bsize, tbase,
                                                                         (uint32_t)nbig, pos, kTileExplicit, hist);
        }
        {
            ProfScope ps(ctx, "colscan");
            k_colscan<<<dim3((unsigned)nbig), kBlock, 0, st>>>(tbase, (uint32_t)nbig, T, bstart, hist,
                                                               sub_start, sub_size);
        }
        {
            ProfScope ps(ctx, "msd_scatter", 8.0 * (double)pass_items);
            k_msd_scatter<Src, false><<<dim3((unsigned)T), kBlock, 0, st>>>(
                src, in_ids, out_ids, bstart, bsize, tbase, (uint32_t)nbig, pos, kTileExplicit, hist, sub_start);
        }
        PLH_HIP(hipGetLastError());
        depth = pos + kDigitBits;
        const uint32_t uni = next_live_host(h_live, kb, depth) >= kb ? 1u : 0u;
        PLH_TRY(ctx->ws.get(kWsScan0, L * 4 + 64, &p));
        uint32_t* cnt = (uint32_t*)p;
        PLH_TRY(ctx->ws.get(kWsScan1, L * 8 + 64, &p));
        uint64_t* newpos = (uint64_t*)p;
        k_expand_count<<<grid1(L), kBlock, 0, st>>>(isbig, bigpos, sub_size, L, cnt);
        PLH_TRY(scan_u32_to_u64(ctx, cnt, newpos, L, d_total));
        uint64_t L2 = 0;
        PLH_TRY(read_u64(ctx, d_total, &L2));
        const int nslot = cur_slot == kWsListA ? kWsListB : kWsListA;
        PLH_TRY(get_lists(ctx, nslot, L2, &B));
        k_expand_write<<<grid1(L), kBlock, 0, st>>>(A.start, A.size, A.meta, isbig, bigpos, sub_start,
                                                    sub_size, newpos, L, depth, uni, B.start, B.size,
                                                    B.meta);
        PLH_HIP(hipGetLastError());
        std::swap(A, B);
        cur_slot = nslot;
        L = L2;
    }

    // ---- finish every bucket in one workgroup, count surviving terms.
    PLH_TRY(ctx->ws.get(kWsKept, L * 4 + 64, &p));
    uint32_t* kept = (uint32_t*)p;
    PLH_TRY(ctx->ws.get(kWsOffs, L * 8 + 64, &p));
    uint64_t* offs = (uint64_t*)p;
    {
        ProfScope ps(ctx, "bucket_sort", 8.0 * (double)n_valid);
        k_bucket_sort<AS, Src><<<dim3((unsigned)L), kBlock, 0, st>>>(src, cfg, A.start, A.size, A.meta,
                                                                     ids0, ids1, kept);
    }
    PLH_HIP(hipGetLastError());
    PLH_TRY(scan_u32_to_u64(ctx, kept, offs, L, d_total));
    uint64_t total = 0;
    PLH_TRY(read_u64(ctx, d_total, &total));
    PLH_TRY(sum_reserve(out, (int64_t)total, 0));
    out->n = (int64_t)total;
    if (total == 0) return PLH_OK;
    cfg.ox = out->x;
    cfg.oz = out->z;
    cfg.of = out->f;
    cfg.ov = out->v;
    cfg.ostride = out->cap;
    {
        ProfScope ps(ctx, "emit", (double)total * (16.0 * out_as + 9.0));
        k_emit<AS, Src><<<dim3((unsigned)L), kBlock, 0, st>>>(src, cfg, A.start, A.size, A.meta, ids0, ids1,
                                                              offs);
    }
    PLH_HIP(hipGetLastError());
    return PLH_OK;
}

int padded_as(int as) {
    if (as <= 1) return 1;
    if (as == 2) return 2;
    if (as <= 4) return 4;
    if (as <= 8) return 8;
    if (as <= 16) return 16;
    return -1;
}

template <int AS>
int collate_terms_as(plh_ctx* ctx, const plh_sum* in, const CollateSpec& spec, plh_sum* out) {
    TermSource<AS> src;
    src.s = SumView{in->x, in->z, in->f, in->v, in->cap};
    src.as = in->as;
    src.mode = spec.mode;
    src.n = (uint32_t)in->n;
    uint64_t orw[32];
    PLH_TRY(op_or_words(ctx, in, orw));
    uint64_t live[kMaxKeyWords] = {0};
    for (int k = 0; k < in->as; ++k) {
        live[k] = orw[k];
        live[AS + k] = orw[in->as + k];
    }
    live[2 * AS] = (spec.mode & kFold) ? (1ull << 56) : (3ull << 56);
    return collate_run<AS>(ctx, src, (uint64_t)in->n, live, spec.mode, spec.max_weight, spec.eps, out,
                           in->as);
}

template <int AS, int KIND>
int collate_pairs_as(plh_ctx* ctx, const plh_sum* Ls, const plh_sum* Rs, int by_i,
                     const CollateSpec& spec, plh_sum* out) {
    // KIND 0: hi = R index, lo = L index.  KIND 1: hi = L index, lo = R index.
    const uint64_t n_lo = KIND == 0 ? (uint64_t)Ls->n : (uint64_t)Rs->n;
    const uint64_t n_hi = KIND == 0 ? (uint64_t)Rs->n : (uint64_t)Ls->n;
    if (n_lo == 0 || n_hi == 0) {
        out->n = 0;
        return PLH_OK;
    }
    uint32_t lo_bits = 0;
    while ((1ull << lo_bits) < n_lo) ++lo_bits;
    if ((n_hi << lo_bits) > (1ull << 32) || n_lo * n_hi > 0xFFFFFFFFull + 1ull) {
        set_error("collated product: the pair ids (rows << ceil(log2(cols))) exceed 2^32");
        return PLH_ERR_SIZE;
    }
    if (n_lo * n_hi > 0xFFFFFFFFull) {
        set_error("collated product: more than 2^32 - 1 pairs");
        return PLH_ERR_SIZE;
    }
    PairSource<AS, KIND> src;
    src.L = SumView{Ls->x, Ls->z, Ls->f, Ls->v, Ls->cap};
    src.R = SumView{Rs->x, Rs->z, Rs->f, Rs->v, Rs->cap};
    src.as = Ls->as;
    src.mode = spec.mode;
    src.lo_bits = lo_bits;
    src.n_lo = (uint32_t)n_lo;
    src.n_hi = (uint32_t)n_hi;
    src.div_lo = make_fastdiv((uint32_t)n_lo);
    src.by_i = (uint32_t)by_i;
    uint64_t ol[32], orr[32];
    PLH_TRY(op_or_words(ctx, Ls, ol));
    PLH_TRY(op_or_words(ctx, Rs, orr));
    uint64_t live[kMaxKeyWords] = {0};
    for (int k = 0; k < Ls->as; ++k) {
        live[k] = ol[k] | orr[k];
        live[AS + k] = ol[Ls->as + k] | orr[Ls->as + k];
    }
    live[2 * AS] = (spec.mode & kFold) ? (1ull << 56) : (3ull << 56);
    return collate_run<AS>(ctx, src, n_lo * n_hi, live, spec.mode, spec.max_weight, spec.eps, out,
                           Ls->as);
}

}  // namespace

int collate_terms(plh_ctx* ctx, plh_sum* in, const CollateSpec& spec, plh_sum* out) {
    if (in->n > (int64_t)0xFFFFFFFFll) {
        set_error("collation supports at most 2^32 - 1 terms");
        return PLH_ERR_SIZE;
    }
    // The output is written to a scratch sum when it aliases the input.
    plh_sum tmp;
    plh_sum* dst = out;
    const bool alias = (out == in);
    if (alias) {
        tmp.ctx = ctx;
        tmp.as = in->as;
        tmp.nq = in->nq;
        dst = &tmp;
    }
    dst->as = in->as;
    dst->nq = in->nq;
    int rc;
    switch (padded_as(in->as)) {
        case 1: rc = collate_terms_as<1>(ctx, in, spec, dst); break;
        case 2: rc = collate_terms_as<2>(ctx, in, spec, dst); break;
        case 4: rc = collate_terms_as<4>(ctx, in, spec, dst); break;
        case 8: rc = collate_terms_as<8>(ctx, in, spec, dst); break;
        case 16: rc = collate_terms_as<16>(ctx, in, spec, dst); break;
        default: set_error("unsupported array_size"); return PLH_ERR_AS;
    }
    if (rc != PLH_OK) {
        if (alias) sum_free_storage(&tmp);
        return rc;
    }
    dst->as = in->as;
    dst->nq = in->nq;
    if (alias) {
        PLH_HIP(hipStreamSynchronize(ctx->stream));
        sum_free_storage(in);
        in->x = tmp.x;
        in->z = tmp.z;
        in->f = tmp.f;
        in->v = tmp.v;
        in->cap = tmp.cap;
        in->n = tmp.n;
    }
    out->canonical = (spec.mode & kFold) && (spec.mode & kDropZero);
    return PLH_OK;
}

int collate_pairs(plh_ctx* ctx, const plh_sum* L, const plh_sum* R, int kind, int by_i,
                  const CollateSpec& spec, plh_sum* out) {
    if (L->as != R->as) {
        set_error("operands have different array sizes");
        return PLH_ERR_ARG;
    }
    if (out == L || out == R) {
        set_error("output must not alias an operand");
        return PLH_ERR_ARG;
    }
    out->as = L->as;
    int rc;
    const int pa = padded_as(L->as);
#define PLH_PAIR(ASV)                                                                     \
    case ASV:                                                                             \
        rc = kind == 0 ? collate_pairs_as<ASV, 0>(ctx, L, R, by_i, spec, out)             \
                       : collate_pairs_as<ASV, 1>(ctx, L, R, by_i, spec, out);            \
        break;
    switch (pa) {
        PLH_PAIR(1)
        PLH_PAIR(2)
        PLH_PAIR(4)
        PLH_PAIR(8)
        PLH_PAIR(16)
        default: set_error("unsupported array_size"); return PLH_ERR_AS;
    }
#undef PLH_PAIR
    if (rc != PLH_OK) return rc;
    out->as = L->as;
    out->canonical = (spec.mode & kFold) && (spec.mode & kDropZero);
    return PLH_OK;
}

}  // namespace plh
