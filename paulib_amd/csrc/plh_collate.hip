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
// live key bits runs in two tiers:
//   1. global passes (tile histogram -> per-bucket column scan -> stable
//      wave-level multi-split scatter) until every bucket fits a leaf; the
//      first pass enumerates ids implicitly (no id array is read).  After a
//      pass, runs of consecutive small sub-buckets are merged into one leaf
//      (they share the parent's prefix and are in key order);
//   2. one workgroup per leaf (<= LCAP ids) sorts it in LDS -- counting sort
//      on the bits where its keys first differ, insertion / rank sort per
//      bin on the full (key, id) -- sums each run in id order, applies the
//      filters and writes the output terms at the leaf's own start (the
//      final position whenever no run collapses and nothing is dropped).
//      Otherwise an exclusive scan of the per-leaf counts and one copy
//      kernel compact the output.
// Stability (ids ascend inside every bucket) and the (key, id) leaf order
// make runs of equal keys come out in original-index order, so the sums are
// the reference's sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
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
constexpr uint32_t kTileMaxExplicit = 16 * kRound;  // 32768 ids per tile at most
constexpr uint32_t kTileMaxImplicit = 128 * kRound; // 262144 pairs per tile at most
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int kMaxKeyWords = 33;                    // 2*16 + 1
constexpr int kLeafThreads = 512;
constexpr int kLeafWaves = kLeafThreads / kWave;
static_assert(kBlock == kDigits, "pass kernels map one thread to one digit");

// Leaf geometry for keys of KW words (plus flags): ids per leaf sized so a
// leaf's LDS (key words, value, sort key, two permutations, flags per item,
// plus the bins) fits its budget.  Single-word keys take ~52 KB, so three
// leaves share a CU's 160 KB (the leaf is latency-bound: on the 65536 x 65536
// product on 64 qubits three resident leaves of 1280 ids beat two of 1792 by
// 17%, and four of 896 lose to three); wider keys take ~78 KB, two per CU.
template <int AS, int KW>
struct LeafDims {
    static constexpr int kLdsBudget = (AS == 1 ? 52 : 78) * 1024;
    static constexpr int kDl = AS <= 2 ? 10 : AS <= 4 ? 9 : AS <= 8 ? 8 : 7;
    static constexpr int kBins = 1 << kDl;  // counting-sort bins
    static constexpr int kPerItem = 8 * KW + 8 + 8 + 2 + 2 + 1;
    static constexpr int kCap = ((kLdsBudget - 4 * (kBins + 1) - 1024) / kPerItem) / 256 * 256;
    static constexpr int kPer = (kCap + kLeafThreads - 1) / kLeafThreads;
    static constexpr int kSlotBits = kCap <= 256 ? 8 : kCap <= 512 ? 9 : kCap <= 1024 ? 10 : 11;
    static_assert(kCap >= 256 && kCap <= 2048, "leaf capacity");
};

// Tile size for a pass over `items` ids: about kTileTarget tiles (four per
// CU: the chip fills while the column scans stay short), whole rounds,
// capped (the cap is rounded down to whole rounds).
constexpr uint64_t kTileTarget = 1024;
PLH_HD uint32_t pick_tile(uint64_t items, uint32_t cap) {
    uint64_t t = items / kTileTarget;
    t = (t + kRound - 1) / kRound * kRound;
    const uint64_t c = cap < kRound ? kRound : (uint64_t)cap / kRound * kRound;
    if (t < kRound) t = kRound;
    if (t > c) t = c;
    return (uint32_t)t;
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
// b % 8), and each XCD has its own L2.  This maps workgroup b < nb to a tile
// so that XCD x takes one contiguous range of the nb tiles (a bijection on
// [0, nb)): neighbouring tiles, whose outputs meet in shared cache lines, then
// mostly write through the same L2.
PLH_DEV uint64_t xcd_contiguous(uint64_t b, uint64_t nb) {
    const uint64_t x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

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
    uint32_t* oj;              // output job plane (batched sums) or null
    const double* job_eps;     // per-job cutoff (batched sums) or null
    const int64_t* job_maxw;   // per-job max weight (< 0: none) or null
    const uint8_t* job_alive;  // per-job keep mask or null
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

PLH_DEV double fold_value(const Cfg& c, double base, uint32_t raw) {
    return ((c.mode & kFold) && (raw & 2u)) ? -base : base;
}

// ------------------------------------------------------- multi-split
// Stable rank of each item among same-digit items of the workgroup, in the
// item order (round, wave, sub-round, lane); running[] carries per-digit
// counts over the rounds of a tile.
// An item id with the 32 key bits that follow the root digit (the carry
// window [W0, W0 + 32), MSB-first).  The ids move through the passes with
// their window, so a pass whose digit lies inside it reads the digit here
// instead of gathering the key from the source.
struct IdKey {
    uint32_t id;
    uint32_t bits;
};
constexpr int kCarryBits = 32;

struct MsplitShared {
    uint32_t wave_cnt[kWavesPerBlock][kDigits];
    uint32_t running[kDigits];
    uint32_t rbefore[kDigits];       // running[] before this round
    uint32_t roff[kDigits];          // this round's items of digit d start at roff[d] in stage
    uint32_t scan[kWavesPerBlock];
    IdKey stage_id[kRound];          // this round's items in (digit, rank) order
    uint8_t stage_dig[kRound];
};
static_assert(kDigits <= 256, "stage_dig holds a digit in one byte");

// Returns the number of valid items of the round (uniform).
PLH_DEV uint32_t msplit_round(MsplitShared& sh, const uint32_t (&dig)[kIpt], uint32_t (&rank)[kIpt]) {
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
    // one digit per thread (kBlock == kDigits): wave offsets, the running
    // count, and this round's count of the digit
    const int dd = threadIdx.x;
    uint32_t run = sh.running[dd];
    sh.rbefore[dd] = run;
    const uint32_t r_before = run;
#pragma unroll
    for (int ww = 0; ww < kWavesPerBlock; ++ww) {
        const uint32_t c = sh.wave_cnt[ww][dd];
        sh.wave_cnt[ww][dd] = run;
        run += c;
    }
    sh.running[dd] = run;
    uint32_t total;
    const uint32_t off = block_exclusive_scan(run - r_before, total, sh.scan);  // has barriers
    sh.roff[dd] = off;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kIpt; ++s)
        if (dig[s] != kInvalid) rank[s] += sh.wave_cnt[w][dig[s]];
    __syncthreads();
    return total;
}

// The round's ids staged in LDS in (digit, rank) order, then written out by
// consecutive threads: each digit's run of the round goes to consecutive
// addresses from base[d] + rank, so a wave's stores fall into a few runs
// instead of 64 scattered words.  Targets are those of the unstaged write.
PLH_DEV void msplit_write(MsplitShared& sh, uint32_t total, const uint32_t (&dig)[kIpt],
                          const uint32_t (&rank)[kIpt], const IdKey (&idv)[kIpt], const uint64_t* base,
                          IdKey* __restrict__ ids_out) {
#pragma unroll
    for (int s = 0; s < kIpt; ++s) {
        const uint32_t d = dig[s];
        if (d == kInvalid) continue;
        const uint32_t slot = sh.roff[d] + (rank[s] - sh.rbefore[d]);
        sh.stage_id[slot] = idv[s];
        sh.stage_dig[slot] = (uint8_t)d;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < total; j += kBlock) {
        const uint32_t d = sh.stage_dig[j];
        ids_out[base[d] + sh.rbefore[d] + (j - sh.roff[d])] = sh.stage_id[j];
    }
    // the next round's first barrier orders these reads before its writes
}

// Position of item (round base r0, wave w, sub s, lane l) inside a round.
PLH_DEV uint32_t round_pos(int s) {
    return (uint32_t)((threadIdx.x / kWave) * (kWave * kIpt) + s * kWave + lane_id());
}

// --------------------------------------------------------- global passes
// Root digit position: the first live bit (0 when none: all keys equal).
PLH_DEV int root_pos(const uint64_t* live, int kb) {
    const int pos = next_live(live, kb, 0);
    return pos >= kb ? 0 : pos;
}

// Start of the carry window: the first key bit after the root digit.
PLH_DEV int carry_start(const uint64_t* live, int kb) { return root_pos(live, kb) + kDigitBits; }
// A pass reads its digit from the carry window when the digit lies inside it.
PLH_DEV bool in_carry(int pos, int w0) { return pos >= w0 && pos + kDigitBits <= w0 + kCarryBits; }
PLH_DEV uint32_t carry_digit(uint32_t bits, int pos, int w0) {
    return (uint32_t)(bits << (pos - w0)) >> (32 - kDigitBits);
}
// The window's bits from pos on, MSB-aligned; bits past the window read 0.
PLH_DEV uint64_t carry_rest(uint32_t bits, int pos, int w0) { return ((uint64_t)bits << 32) << (pos - w0); }

constexpr int kHistU = 16;  // items per thread per round of the histogram kernels

// First pass: the root bucket, ids enumerated implicitly (dense index i is
// item i of the source), one digit position for the whole root.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_root_hist(Src src, uint64_t n_dense, const uint64_t* __restrict__ live,
                                                      int kb, uint32_t tile_items, uint32_t* __restrict__ hist) {
    __shared__ uint32_t hw[kWavesPerBlock][kDigits];  // one histogram per wave
    const int pos = root_pos(live, kb);
    for (int i = threadIdx.x; i < kWavesPerBlock * kDigits; i += kBlock) (&hw[0][0])[i] = 0;
    uint32_t* h = hw[threadIdx.x / kWave];
    __syncthreads();
    const uint64_t t = blockIdx.x;
    const uint64_t beg = t * tile_items;
    const uint64_t end = min(n_dense, beg + tile_items);
    for (uint64_t i0 = beg + threadIdx.x; i0 < end; i0 += (uint64_t)kBlock * kHistU) {
        uint32_t dg[kHistU];
#pragma unroll
        for (int u = 0; u < kHistU; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kBlock;
            dg[u] = kInvalid;
            if (i < end) {
                const uint32_t id = src.dense_to_id((uint32_t)i);
                if (src.valid(id)) dg[u] = key_window<kDigitBits>(src, id, pos);
            }
        }
#pragma unroll
        for (int u = 0; u < kHistU; ++u)
            if (dg[u] != kInvalid) atomicAdd(&h[dg[u]], 1u);
    }
    __syncthreads();
    uint32_t sum = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) sum += hw[w][threadIdx.x];
    hist[t * kDigits + threadIdx.x] = sum;
}

// Column scans of per-tile digit histograms (in place: each count becomes
// the exclusive prefix of its digit over the earlier tiles).  kColGroups
// groups of kDigits threads take consecutive slices of the tiles: a sum
// pass per slice, then a prefix-writing pass from the slice's offset, with
// kColChunk loads in flight per thread.  Returns the digit total in group
// 0's thread d (0 in the others).
constexpr int kColGroups = 4;
constexpr int kColThreads = kDigits * kColGroups;
constexpr int kColChunk = 32;

template <int G>
PLH_DEV uint64_t column_scan(uint64_t t0, uint64_t t1, uint32_t* __restrict__ hist, uint64_t (&part)[G][kDigits]) {
    const int d = (int)threadIdx.x % kDigits, g = (int)threadIdx.x / kDigits;
    const uint64_t n = t1 - t0, per = (n + G - 1) / G;
    const uint64_t a = t0 + min(n, (uint64_t)g * per), b = t0 + min(n, (uint64_t)(g + 1) * per);
    uint64_t sum = 0;
    for (uint64_t t = a; t < b; t += kColChunk) {
        uint32_t c[kColChunk];
#pragma unroll
        for (int k = 0; k < kColChunk; ++k) c[k] = (t + k < b) ? hist[(t + k) * kDigits + d] : 0u;
#pragma unroll
        for (int k = 0; k < kColChunk; ++k) sum += c[k];
    }
    part[g][d] = sum;
    __syncthreads();
    uint64_t run = 0;
    for (int q = 0; q < g; ++q) run += part[q][d];
    for (uint64_t t = a; t < b; t += kColChunk) {
        uint32_t c[kColChunk];
#pragma unroll
        for (int k = 0; k < kColChunk; ++k) c[k] = (t + k < b) ? hist[(t + k) * kDigits + d] : 0u;
#pragma unroll
        for (int k = 0; k < kColChunk; ++k) {
            if (t + k < b) hist[(t + k) * kDigits + d] = (uint32_t)run;
            run += c[k];
        }
    }
    uint64_t tot_d = 0;
    if (g == 0)
        for (int q = 0; q < G; ++q) tot_d += part[q][d];
    return tot_d;
}

// The root's column scan and its digit offsets (one workgroup).
__global__ __launch_bounds__(kColThreads) void k_root_colscan(uint64_t ntiles, uint32_t* __restrict__ hist,
                                                              uint64_t* __restrict__ sub_start,
                                                              uint64_t* __restrict__ sub_size) {
    __shared__ uint64_t part[kColGroups][kDigits];
    __shared__ uint64_t sh[kColThreads / kWave];
    const uint64_t tot_d = column_scan<kColGroups>(0, ntiles, hist, part);
    // digit totals in the first group's threads, a scan over the whole
    // workgroup (the other threads add 0 after them)
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan<kColThreads / kWave>(tot_d, tot, sh);
    if (threadIdx.x < kDigits) {
        sub_start[threadIdx.x] = ex;
        sub_size[threadIdx.x] = tot_d;
    }
}

template <class Src>
__global__ __launch_bounds__(kBlock) void k_root_scatter(Src src, IdKey* __restrict__ ids_out, uint64_t n_dense,
                                                         const uint64_t* __restrict__ live, int kb,
                                                         uint32_t tile_items, const uint32_t* __restrict__ hist,
                                                         const uint64_t* __restrict__ sub_start) {
    __shared__ MsplitShared sh;
    const int pos = root_pos(live, kb);
    __shared__ uint64_t base[kDigits];
    const uint64_t t = blockIdx.x;
    const uint64_t beg = t * tile_items;
    const uint64_t end = min(n_dense, beg + tile_items);
    base[threadIdx.x] = sub_start[threadIdx.x] + hist[t * kDigits + threadIdx.x];
    sh.running[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t r0 = beg; r0 < end; r0 += kRound) {
        uint32_t dig[kIpt], rank[kIpt];
        IdKey idv[kIpt];
#pragma unroll
        for (int s = 0; s < kIpt; ++s) {
            const uint64_t i = r0 + round_pos(s);
            dig[s] = kInvalid;
            idv[s] = IdKey{0u, 0u};
            if (i < end) {
                const uint32_t id = src.dense_to_id((uint32_t)i);
                if (src.valid(id)) {
                    // the root digit and the carry window after it, in one read
                    const uint64_t win = key_bits<kDigitBits + kCarryBits>(src, id, pos);
                    idv[s] = IdKey{id, (uint32_t)win};
                    dig[s] = (uint32_t)(win >> kCarryBits);
                }
            }
        }
        const uint32_t total = msplit_round(sh, dig, rank);
        msplit_write(sh, total, dig, rank, idv, base, ids_out);
    }
}

// ---------------------------------------------- root pass of a product by groups
// Each x/z word of a product key is the XOR of its operands' words, so when
// the root digit lies in the x/z words, the digit of pair (rhs i, lhs j) is
// dl(j) ^ dr(i), and so are the 32 carried bits when they lie there too.
// Root bucket d then holds, for each rhs row i in order, the lhs terms of
// digit d ^ dr(i) in index order: the ids in ascending order, exactly what
// the scatter writes.  The lhs are grouped by digit once (stably); a block of
// rhs rows counts its pairs per bucket from the group sizes alone, and its
// ids are written group by group, with no key gather and no ranking.

// The 40-bit window [pos, pos + 40) of one operand record's x/z words (words
// from 2*AS on read 0; callers use it only where the window lies in them).
template <int AS>
PLH_DEV uint64_t rec_window(const uint64_t* rec, int pos) {
    constexpr int W = kDigitBits + kCarryBits;
    const int w = pos >> 6, o = pos & 63;
    uint64_t hi = (w < 2 * AS ? rec[w] : 0ull) << o;
    if (o + W > 64) hi |= (w + 1 < 2 * AS ? rec[w + 1] : 0ull) >> (64 - o);
    return hi >> (64 - W);
}

template <class Src>
struct GroupRoot : std::false_type {};
template <int AS>
struct GroupRoot<PairSource<AS, 0>> : std::true_type {};

constexpr int kGrpThreads = 1024;  // the grouping workgroup
constexpr uint32_t kGrpRows = 32;  // rhs rows per block of the generator

// Stable grouping of the n lhs terms by root digit, in one workgroup:
// gcnt[0..255] the group sizes, gcnt[256..511] their starts; position p of
// the grouped order holds lrec[p] = carried bits << 32 | j and lgrp[p] = the
// digit.  A chunk of 1024 terms is ranked by 8 ballots per wave and per-wave
// digit counts.
template <int AS>
__global__ __launch_bounds__(kGrpThreads) void k_group_lhs(PackedView<AS> L, uint32_t n,
                                                          const uint64_t* __restrict__ live, int kb,
                                                          uint64_t* __restrict__ lrec, uint8_t* __restrict__ lgrp,
                                                          uint32_t* __restrict__ gcnt) {
    constexpr int kW = kGrpThreads / kWave;
    __shared__ uint32_t h[kDigits];
    __shared__ uint32_t run[kDigits];
    __shared__ uint32_t wc[kW][kDigits];
    const int pos = root_pos(live, kb);
    if (threadIdx.x < (unsigned)kDigits) h[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += kGrpThreads)
        atomicAdd(&h[(uint32_t)(rec_window<AS>(L.rec(j), pos) >> kCarryBits)], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0;
        for (int d = 0; d < kDigits; ++d) {
            run[d] = r;
            r += h[d];
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)kDigits) {
        gcnt[threadIdx.x] = h[threadIdx.x];
        gcnt[kDigits + threadIdx.x] = run[threadIdx.x];
    }
    const int w = (int)threadIdx.x / kWave;
    const uint64_t lt = lanemask_lt();
    for (uint32_t c = 0; c < n; c += kGrpThreads) {
        for (int i = threadIdx.x; i < kW * kDigits; i += kGrpThreads) (&wc[0][0])[i] = 0;
        __syncthreads();
        const uint32_t j = c + threadIdx.x;
        const bool valid = j < n;
        const uint64_t win = valid ? rec_window<AS>(L.rec(j), pos) : 0ull;
        const uint32_t d = (uint32_t)(win >> kCarryBits);
        uint64_t m = ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const bool bit = valid && ((d >> b) & 1u);
            const uint64_t bb = ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t lower = (uint32_t)popc(m & lt);
        if (valid && lower == 0) wc[w][d] = (uint32_t)popc(m);
        __syncthreads();
        if (threadIdx.x < (unsigned)kDigits) {
            const int dd = (int)threadIdx.x;
            uint32_t r = run[dd];
#pragma unroll
            for (int ww = 0; ww < kW; ++ww) {
                const uint32_t t = wc[ww][dd];
                wc[ww][dd] = r;
                r += t;
            }
            run[dd] = r;
        }
        __syncthreads();
        if (valid) {
            const uint32_t p = wc[w][d] + lower;
            lrec[p] = (uint64_t)(uint32_t)win << 32 | j;
            lgrp[p] = (uint8_t)d;
        }
        __syncthreads();  // before the next chunk clears wc
    }
}

// Pairs per root bucket of each block of kGrpRows rhs rows: hist[b][d] =
// sum over the block's rows i of the size of lhs group d ^ dr(i).
template <int AS>
__global__ __launch_bounds__(kBlock) void k_group_hist(PackedView<AS> R, uint32_t m, const uint64_t* __restrict__ live,
                                                       int kb, const uint32_t* __restrict__ gcnt,
                                                       uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[kDigits];
    __shared__ uint32_t dr[kGrpRows];
    const int pos = root_pos(live, kb);
    const uint32_t i0 = blockIdx.x * kGrpRows;
    const uint32_t nr = min(kGrpRows, m - i0);
    cnt[threadIdx.x] = gcnt[threadIdx.x];
    if (threadIdx.x < nr) dr[threadIdx.x] = (uint32_t)(rec_window<AS>(R.rec(i0 + threadIdx.x), pos) >> kCarryBits);
    __syncthreads();
    uint32_t s = 0;
    for (uint32_t r = 0; r < nr; ++r) s += cnt[threadIdx.x ^ dr[r]];
    hist[(uint64_t)blockIdx.x * kDigits + threadIdx.x] = s;
}

// The root ids of one block of rhs rows (blockIdx.x) for one chunk of the
// grouped lhs positions (blockIdx.y).  Row r's pairs with the lhs of group g
// start at bucket g ^ dr(r)'s next free slot; adj[r][g] is that slot minus
// the group's start, so position p goes to adj[r][g(p)] + p (mod 2^32: every
// target is below n_dense <= 2^32).  A wave's 64 positions fall in one group
// or a few, so each row's stores are runs of consecutive ids.
template <int AS>
__global__ __launch_bounds__(kBlock) void k_group_gen(PairSource<AS, 0> src, uint32_t m, uint32_t n, uint32_t chunk,
                                                      const uint64_t* __restrict__ live, int kb,
                                                      const uint64_t* __restrict__ lrec,
                                                      const uint8_t* __restrict__ lgrp,
                                                      const uint32_t* __restrict__ gcnt,
                                                      const uint32_t* __restrict__ hist,
                                                      const uint64_t* __restrict__ sub_start,
                                                      IdKey* __restrict__ ids_out) {
    __shared__ uint32_t adj[kGrpRows][kDigits];
    __shared__ uint32_t cnt[kDigits];
    __shared__ uint32_t gst[kDigits];
    __shared__ uint32_t dr[kGrpRows];
    __shared__ uint32_t rw[kGrpRows];
    const int pos = root_pos(live, kb);
    const bool xw = pos + kDigitBits + kCarryBits <= 2 * AS * 64;  // carried bits lie in the x/z words
    // row blocks of one chunk write next to each other in every bucket: an
    // XCD takes a contiguous run of them (linear order x + y * gridDim.x)
    const uint64_t lin = xcd_contiguous(blockIdx.x + (uint64_t)blockIdx.y * gridDim.x,
                                        (uint64_t)gridDim.x * gridDim.y);
    const uint32_t bx = (uint32_t)(lin % gridDim.x), by = (uint32_t)(lin / gridDim.x);
    const uint32_t i0 = bx * kGrpRows;
    const uint32_t nr = min(kGrpRows, m - i0);
    cnt[threadIdx.x] = gcnt[threadIdx.x];
    gst[threadIdx.x] = gcnt[kDigits + threadIdx.x];
    if (threadIdx.x < nr) {
        const uint64_t win = rec_window<AS>(src.R.rec(i0 + threadIdx.x), pos);
        dr[threadIdx.x] = (uint32_t)(win >> kCarryBits);
        rw[threadIdx.x] = (uint32_t)win;
    }
    __syncthreads();
    {
        const uint32_t d = threadIdx.x;
        uint64_t next = sub_start[d] + hist[(uint64_t)bx * kDigits + d];
        for (uint32_t r = 0; r < nr; ++r) {
            const uint32_t g = d ^ dr[r];
            adj[r][g] = (uint32_t)(next - gst[g]);
            next += cnt[g];
        }
    }
    __syncthreads();
    const uint32_t p0 = by * chunk;
    const uint32_t p1 = min(n, p0 + chunk);
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += kBlock) {
        const uint64_t lr = lrec[p];
        const uint32_t g = lgrp[p];
        const uint32_t j = (uint32_t)lr, lw = (uint32_t)(lr >> 32);
        for (uint32_t r = 0; r < nr; ++r) {
            const uint32_t id = ((i0 + r) << src.lo_bits) | j;
            const uint32_t bits = xw ? (lw ^ rw[r]) : (uint32_t)key_bits<kDigitBits + kCarryBits>(src, id, pos);
            ids_out[(size_t)(uint32_t)(adj[r][g] + p)] = IdKey{id, bits};
        }
    }
}

// ------------------------------------------------------------ bucket lists
// meta: bits 0-15 depth (key bits before it are equal across the bucket),
// bit 16 id buffer, bit 17 uniform (no live bit at or after depth).
PLH_HD uint32_t meta_depth(uint32_t m) { return m & 0xFFFFu; }
PLH_HD uint32_t meta_buf(uint32_t m) { return (m >> 16) & 1u; }
PLH_HD uint32_t meta_uniform(uint32_t m) { return (m >> 17) & 1u; }
PLH_HD uint32_t make_meta(uint32_t depth, uint32_t buf, uint32_t uni) {
    return depth | (buf << 16) | (uni << 17);
}

// The big buckets of one pass (device arrays of nbig entries each).
struct PassBuckets {
    uint64_t* start;
    uint64_t* size;
    uint64_t* tbase;  // exclusive scan of ntiles
    uint64_t* kor;    // OR / AND over the bucket of the key bits from pos to its word's end
    uint64_t* kand;
    uint32_t* meta;
    uint32_t* pos;     // digit position of this pass (first live bit >= depth)
    uint32_t* ntiles;
    uint32_t* skip;    // 0: split on the digit; else offset of the first differing bit (>= 8)
};

// The bucket list of a pass: entries at or beyond *dL are stale.  isbig
// marks buckets to split; bigsz feeds the count of items moved this pass.
__global__ void k_classify(const uint64_t* __restrict__ Lsize, const uint32_t* __restrict__ Lmeta,
                           const uint64_t* __restrict__ dL, uint64_t Lb, uint32_t cap,
                           uint32_t* __restrict__ isbig, uint64_t* __restrict__ bigsz) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Lb) return;
    const bool big = e < *dL && Lsize[e] > cap && !meta_uniform(Lmeta[e]);
    isbig[e] = big ? 1u : 0u;
    bigsz[e] = big ? Lsize[e] : 0ull;
}

__global__ void k_gather_big(const uint64_t* __restrict__ Lstart, const uint64_t* __restrict__ Lsize,
                             const uint32_t* __restrict__ Lmeta, const uint32_t* __restrict__ isbig,
                             const uint64_t* __restrict__ bigpos, uint64_t Lb, PassBuckets B) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Lb || !isbig[e]) return;
    const uint64_t b = bigpos[e];
    B.start[b] = Lstart[e];
    B.size[b] = Lsize[e];
    B.meta[b] = Lmeta[e];
}

__global__ void k_bucket_setup(PassBuckets B, uint64_t nbig, uint32_t tile, const uint64_t* __restrict__ live,
                               int kb) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbig) return;
    B.pos[b] = (uint32_t)next_live(live, kb, (int)meta_depth(B.meta[b]));
    B.ntiles[b] = (uint32_t)((B.size[b] + tile - 1) / tile);
    B.kor[b] = 0ull;
    B.kand[b] = ~0ull;
}

// Which big bucket a tile belongs to (tbase is an exclusive scan).
PLH_DEV uint32_t find_bucket(const uint64_t* tbase, uint32_t nbig, uint64_t t) {
    uint32_t lo = 0, hi = nbig;  // largest b with tbase[b] <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tbase[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// The digit at pos, and in `rest` the key bits from pos to the end of its
// word, MSB-aligned and zero-filled (64 - pos % 64 bits are examined).
template <class Src>
PLH_DEV uint32_t digit_and_rest(const Src& src, uint32_t id, int pos, uint64_t& rest) {
    const int w = pos >> 6, o = pos & 63;
    const uint64_t hi = src.word(id, w) << o;
    rest = hi;
    uint64_t d = hi;
    if (o + kDigitBits > 64) d |= src.word(id, w + 1) >> (64 - o);
    return (uint32_t)(d >> (64 - kDigitBits));
}

// Digit histogram of a tile at its bucket's position, plus the OR and AND
// of the key bits from there to the end of that word (so a bucket whose
// digit is constant can jump straight to its first differing bit).  Grid
// is an upper bound on the tile count; tiles at or beyond *dT exit at once.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_pass_hist(Src src, const IdKey* __restrict__ ids0,
                                                      const IdKey* __restrict__ ids1, PassBuckets B,
                                                      uint32_t nbig, const uint64_t* __restrict__ dT,
                                                      uint32_t tile_items, uint32_t* __restrict__ hist,
                                                      const uint64_t* __restrict__ live, int kb) {
    __shared__ uint32_t hw[kWavesPerBlock][kDigits];  // one histogram per wave: fewer colliding atomics
    __shared__ unsigned long long sor, sand;
    const uint64_t t = blockIdx.x;
    if (t >= *dT) return;
    for (int i = threadIdx.x; i < kWavesPerBlock * kDigits; i += kBlock) (&hw[0][0])[i] = 0;
    uint32_t* h = hw[threadIdx.x / kWave];
    if (threadIdx.x == 0) {
        sor = 0ull;
        sand = ~0ull;
    }
    __syncthreads();
    const uint32_t bi = find_bucket(B.tbase, nbig, t);
    const uint64_t beg = B.start[bi] + (t - B.tbase[bi]) * tile_items;
    const uint64_t end = min(B.start[bi] + B.size[bi], beg + tile_items);
    const int pos = (int)B.pos[bi];
    const int w0 = carry_start(live, kb);
    const bool carried = in_carry(pos, w0);  // uniform over the tile
    const IdKey* ids = meta_buf(B.meta[bi]) ? ids1 : ids0;
    uint64_t o = 0ull, a = ~0ull;
    // kHistU items per thread per round: their id loads, then their key
    // gathers (when the digit is not carried), are all in flight before the
    // first LDS atomic
    for (uint64_t i0 = beg + threadIdx.x; i0 < end; i0 += (uint64_t)kBlock * kHistU) {
        IdKey idv[kHistU];
        uint32_t dg[kHistU];
#pragma unroll
        for (int u = 0; u < kHistU; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kBlock;
            idv[u] = i < end ? ids[i] : IdKey{0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < kHistU; ++u) {
            dg[u] = kInvalid;
            if (i0 + (uint64_t)u * kBlock < end) {
                uint64_t rest;
                if (carried) {
                    dg[u] = carry_digit(idv[u].bits, pos, w0);
                    rest = carry_rest(idv[u].bits, pos, w0);
                } else {
                    dg[u] = digit_and_rest(src, idv[u].id, pos, rest);
                }
                o |= rest;
                a &= rest;
            }
        }
#pragma unroll
        for (int u = 0; u < kHistU; ++u)
            if (dg[u] != kInvalid) atomicAdd(&h[dg[u]], 1u);
    }
    o = wave_or(o);
    a = wave_and(a);
    if (lane_id() == 0) {
        atomicOr(&sor, (unsigned long long)o);
        atomicAnd(&sand, (unsigned long long)a);
    }
    __syncthreads();
    {
        uint32_t sum = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) sum += hw[w][threadIdx.x];
        hist[t * kDigits + threadIdx.x] = sum;
    }
    if (threadIdx.x == 0) {
        atomicOr((unsigned long long*)&B.kor[bi], sor);
        atomicAnd((unsigned long long*)&B.kand[bi], sand);
    }
}

// One workgroup per big bucket: column scan, and whether the digit splits.
// The children of a split bucket, in digit order: sub-buckets larger than
// `cap` stay single (depth pos + 8); runs of consecutive small ones are
// merged greedily while their total fits `cap` (a run of several
// sub-buckets keeps the parent's depth, pos).  Empty sub-buckets are
// skipped.  Read sequentially: a group starts at a non-empty sub-bucket h
// and ends before the first non-empty sub-bucket that is large or would
// take the total over `cap`; the next group starts there.
struct Children {
    uint64_t* start;
    uint64_t* size;
    uint32_t* meta;
    uint32_t* count;
};

// Orders one wave's LDS accesses around it: a wave issues its LDS
// instructions in order, so lanes see each other's earlier writes once the
// compiler keeps the order (no workgroup barrier needed).
PLH_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of one wave's grouping (group_children).
struct GroupShared {
    uint64_t P[kDigits + 1];   // exclusive prefix of the sizes
    uint32_t NZ[kDigits + 1];  // exclusive count of non-empty sub-buckets
    uint32_t J[kDigits + 1];   // successor under pointer doubling
    uint32_t M[kDigits + 1];   // reached from the first group's head
};

// Index of the first set flag among bits k+1..3 of a lane's 4-bit mask,
// else of the first flagged element in a later lane, else kDigits.
PLH_DEV uint32_t next_flagged(uint32_t mine, int k, uint32_t later) {
    const uint32_t rest = mine & (0xEu << k) & 0xFu;
    return rest ? 4u * lane_id() + (uint32_t)__builtin_ctz(rest) : later;
}
PLH_DEV uint32_t first_flagged_after_lane(uint32_t mine) {
    const uint32_t l = lane_id();
    const uint64_t any = ballot(mine != 0);
    const uint64_t above = l == kWave - 1 ? 0ull : any & (~0ull << (l + 1));
    const int l2 = above ? __ffsll((long long)above) - 1 : (int)l;
    const uint32_t first = mine ? (uint32_t)__builtin_ctz(mine) : 0u;
    const uint32_t f2 = (uint32_t)__shfl((int)first, l2, kWave);
    return above ? 4u * (uint32_t)l2 + f2 : (uint32_t)kDigits;
}

// The grouping of 256 sub-buckets held by one wave (lane l has digits 4l ..
// 4l+3), computed in parallel: each sub-bucket d gets the end nxt[d] of a
// group that would start there (a binary search on the size prefix, capped
// by the next large sub-bucket; a large one ends at the next non-empty
// one), and the groups actually formed are the heads reached from the
// first non-empty sub-bucket by following nxt (8 rounds of pointer
// doubling cover chains of up to 256).  emit(o, start, size, parts) is
// called by the lane owning group o's head (o counts groups in digit
// order).  Returns the number of groups (uniform).  Runs on one wave (the
// other waves of its workgroup need not take part) with wave-local syncs.
template <class F>
PLH_DEV uint32_t group_children(const uint64_t (&st)[4], const uint64_t (&ss)[4], uint32_t cap, GroupShared& g,
                                F&& emit) {
    const uint32_t l = lane_id();
    uint32_t bigm = 0, nzm = 0;
    uint64_t lp[4];
    uint32_t lnz[4];
    uint64_t run = 0;
    uint32_t nrun = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        lp[k] = run;
        lnz[k] = nrun;
        run += ss[k];
        nrun += ss[k] != 0;
        bigm |= (ss[k] > cap ? 1u : 0u) << k;
        nzm |= (ss[k] != 0 ? 1u : 0u) << k;
    }
    const uint64_t pex = wave_inclusive_scan(run) - run;
    const uint32_t nex = wave_inclusive_scan(nrun) - nrun;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        g.P[4 * l + k] = pex + lp[k];
        g.NZ[4 * l + k] = nex + lnz[k];
    }
    if (l == kWave - 1) {
        g.P[kDigits] = pex + run;
        g.NZ[kDigits] = nex + nrun;
        g.J[kDigits] = kDigits;
        g.M[kDigits] = 0;
    }
    const uint32_t big_later = first_flagged_after_lane(bigm);
    const uint32_t nz_later = first_flagged_after_lane(nzm);
    const uint64_t nzl = ballot(nzm != 0);
    if (!nzl) return 0;  // every sub-bucket empty (uniform)
    const int lc = __ffsll((long long)nzl) - 1;
    const uint32_t c0 = 4u * (uint32_t)lc + (uint32_t)__builtin_ctz((uint32_t)__shfl((int)nzm, lc, kWave));
    wave_sync();
    uint32_t nxt[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t d = 4 * l + k;
        if ((bigm >> k) & 1u) {
            nxt[k] = next_flagged(nzm, k, nz_later);
        } else {
            // first m in [d+1, 256] with P[m] > P[d] + cap (257: none)
            const uint64_t t = g.P[d] + cap;
            uint32_t lo = d + 1, n = kDigits + 1 - lo;
            while (n > 0) {
                const uint32_t half = n / 2, m = lo + half;
                if (g.P[m] <= t) {
                    lo = m + 1;
                    n -= half + 1;
                } else {
                    n = half;
                }
            }
            nxt[k] = min(next_flagged(bigm, k, big_later), lo - 1);
        }
        g.J[d] = nxt[k];
        g.M[d] = d == c0 ? 1u : 0u;
    }
    wave_sync();
    for (int r = 0; r < 8; ++r) {
        uint32_t j[4], jj[4], m[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            j[k] = g.J[4 * l + k];
            jj[k] = g.J[j[k]];
            m[k] = g.M[4 * l + k];
        }
        wave_sync();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (m[k]) g.M[j[k]] = 1u;
            g.J[4 * l + k] = jj[k];
        }
        wave_sync();
    }
    uint32_t head = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) head |= (g.M[4 * l + k] ? 1u : 0u) << k;
    const uint32_t hc = (uint32_t)__builtin_popcount(head);
    const uint32_t hinc = wave_inclusive_scan(hc);
    uint32_t o = hinc - hc;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!((head >> k) & 1u)) continue;
        const uint32_t d = 4 * l + k;
        emit(o, st[k], g.P[nxt[k]] - g.P[d], g.NZ[nxt[k]] - g.NZ[d]);
        ++o;
    }
    return wave_last(hinc);
}

// Per big bucket (one workgroup each): the column scan of its tiles'
// histograms, whether its digit is constant (skip), and its children for
// the next list (group_children on wave 0).  A bucket whose digit was
// constant has one child: itself, deeper, in the same id buffer.
// G groups of kDigits threads per bucket: 4 when buckets hold many tiles
// (the scan's serial part shrinks), 1 when they hold few.
template <int G>
__global__ __launch_bounds__(kDigits * G) void k_pass_colscan(PassBuckets B, uint32_t nbig,
                                                             const uint64_t* __restrict__ dT,
                                                             uint32_t* __restrict__ hist,
                                                             uint64_t* __restrict__ sub_start,
                                                             uint64_t* __restrict__ sub_size,
                                                             const uint64_t* __restrict__ live, int kb,
                                                             uint32_t cap, Children C) {
    __shared__ uint64_t part[G][kDigits];
    __shared__ uint64_t sh[kDigits * G / kWave];
    __shared__ uint64_t gst[kDigits], gss[kDigits];
    __shared__ uint32_t skip_sh;
    __shared__ GroupShared g;
    const uint32_t bi = blockIdx.x;
    const uint64_t t0 = B.tbase[bi];
    const uint64_t t1 = (bi + 1 < nbig) ? B.tbase[bi + 1] : *dT;
    const uint64_t base = (uint64_t)bi * kDigits;
    const uint64_t tot_d = column_scan<G>(t0, t1, hist, part);
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan<kDigits * G / kWave>(tot_d, tot, sh);
    if (threadIdx.x < kDigits) {
        const uint64_t s0 = B.start[bi] + ex;
        sub_start[base + threadIdx.x] = s0;
        sub_size[base + threadIdx.x] = tot_d;
        gst[threadIdx.x] = s0;
        gss[threadIdx.x] = tot_d;
    }
    const uint32_t pos = B.pos[bi];
    if (threadIdx.x == 0) {
        // first differing bit among the examined ones, or the end of what was
        // examined: the carry window for a carried digit, else the word
        const int w0 = carry_start(live, kb);
        const uint64_t diff = B.kor[bi] ^ B.kand[bi];
        const uint32_t end =
            in_carry((int)pos, w0) ? (uint32_t)(w0 + kCarryBits - (int)pos) : 64u - (pos & 63u);
        const uint32_t off = diff ? (uint32_t)__clzll(diff) : end;
        const uint32_t skip = off >= (uint32_t)kDigitBits ? off : 0u;
        B.skip[bi] = skip;
        skip_sh = skip;
    }
    __syncthreads();
    if (threadIdx.x >= kWave) return;
    const int lane = (int)threadIdx.x;
    const uint32_t skip = skip_sh;
    if (skip) {
        if (lane == 0) {
            const int depth = min((int)(pos + skip), kb);
            C.start[base] = B.start[bi];
            C.size[base] = B.size[bi];
            C.meta[base] = make_meta((uint32_t)depth, meta_buf(B.meta[bi]), next_live(live, kb, depth) >= kb ? 1u : 0u);
            C.count[bi] = 1;
        }
        return;
    }
    uint64_t st[4], ss[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        st[k] = gst[4 * lane + k];
        ss[k] = gss[4 * lane + k];
    }
    const uint32_t buf = meta_buf(B.meta[bi]) ^ 1u;
    const int child_depth = min((int)pos + kDigitBits, kb);
    const uint32_t child_uniform = next_live(live, kb, child_depth) >= kb ? 1u : 0u;
    const uint32_t n = group_children(st, ss, cap, g, [&](uint32_t o, uint64_t s0, uint64_t sz, uint32_t parts) {
        C.start[base + o] = s0;
        C.size[base + o] = sz;
        C.meta[base + o] =
            parts == 1 ? make_meta((uint32_t)child_depth, buf, child_uniform) : make_meta(pos, buf, 0u);
    });
    if (lane == 0) C.count[bi] = n;
}

// Stable scatter of a tile into its bucket's sub-buckets (other id buffer);
// buckets whose digit is constant are left in place.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_pass_scatter(Src src, IdKey* __restrict__ ids0,
                                                         IdKey* __restrict__ ids1, PassBuckets B, uint32_t nbig,
                                                         const uint64_t* __restrict__ dT, uint32_t tile_items,
                                                         const uint32_t* __restrict__ hist,
                                                         const uint64_t* __restrict__ sub_start,
                                                         const uint64_t* __restrict__ live, int kb) {
    __shared__ MsplitShared sh;
    __shared__ uint64_t base[kDigits];
    const uint64_t nt = *dT;
    if (blockIdx.x >= nt) return;
    const uint64_t t = xcd_contiguous(blockIdx.x, nt);
    const uint32_t bi = find_bucket(B.tbase, nbig, t);
    if (B.skip[bi]) return;
    const uint64_t beg = B.start[bi] + (t - B.tbase[bi]) * tile_items;
    const uint64_t end = min(B.start[bi] + B.size[bi], beg + tile_items);
    const int pos = (int)B.pos[bi];
    const int w0 = carry_start(live, kb);
    const bool carried = in_carry(pos, w0);
    const bool from1 = meta_buf(B.meta[bi]) != 0;
    const IdKey* ids_in = from1 ? ids1 : ids0;
    IdKey* ids_out = from1 ? ids0 : ids1;
    base[threadIdx.x] = sub_start[(uint64_t)bi * kDigits + threadIdx.x] + hist[t * kDigits + threadIdx.x];
    sh.running[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t r0 = beg; r0 < end; r0 += kRound) {
        uint32_t dig[kIpt], rank[kIpt];
        IdKey idv[kIpt];
        // every id load of the round, then every key gather, before the split
#pragma unroll
        for (int s = 0; s < kIpt; ++s) {
            const uint64_t i = r0 + round_pos(s);
            idv[s] = i < end ? ids_in[i] : IdKey{0u, 0u};
        }
#pragma unroll
        for (int s = 0; s < kIpt; ++s) {
            dig[s] = kInvalid;
            if (r0 + round_pos(s) < end)
                dig[s] = carried ? carry_digit(idv[s].bits, pos, w0) : key_window<kDigitBits>(src, idv[s].id, pos);
        }
        const uint32_t total = msplit_round(sh, dig, rank);
        msplit_write(sh, total, dig, rank, idv, base, ids_out);
    }
}

PLH_DEV uint32_t expand_count_one(uint64_t e, const uint32_t* isbig, const uint64_t* bigpos, const Children& C) {
    return isbig[e] ? C.count[bigpos[e]] : 1u;
}

__global__ void k_expand_count(const uint32_t* __restrict__ isbig, const uint64_t* __restrict__ bigpos, Children C,
                               uint64_t L, uint32_t* __restrict__ cnt) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < L) cnt[e] = expand_count_one(e, isbig, bigpos, C);
}



// ---- short lists: one workgroup does a list step that otherwise takes a
// kernel, a scan (one to three launches) and another kernel.
constexpr int kPlanThreads = 1024;
constexpr int kPlanWaves = kPlanThreads / kWave;
constexpr int kPlanPer = 16;
constexpr uint64_t kPlanMax = (uint64_t)kPlanThreads * kPlanPer;  // list entries

// classify + scan + gather of the big buckets of a list of L entries, then
// their setup (digit position, tile count, tile base) for tiles of
// pick_tile(items, tile_cap) ids, the size the host derives from the same
// item count; thread 0 posts the pass's read-back {L, big buckets, their
// items, valid items}.  Entry e is thread e / kPlanPer's.  One workgroup of
// kPlanThreads; L <= kPlanMax.
PLH_DEV void plan_list(const uint64_t* Lstart, const uint64_t* Lsize, const uint32_t* Lmeta, uint64_t L, uint32_t cap,
                       uint32_t* isbig, uint64_t* bigpos, const PassBuckets& B, uint64_t* d_nbig,
                       uint64_t* d_items, uint32_t tile_cap, const uint64_t* live, int kb, uint64_t* d_T,
                       uint64_t n_valid, uint64_t* mbox, uint64_t seq, uint64_t* sh) {
    const uint64_t e0 = (uint64_t)threadIdx.x * kPlanPer;
    static_assert(kPlanPer <= 32, "big flags in one word");
    uint32_t big = 0;
    uint64_t cnt = 0, items = 0;
#pragma unroll
    for (int k = 0; k < kPlanPer; ++k) {
        const uint64_t e = e0 + k;
        const uint64_t sz = e < L ? Lsize[e] : 0ull;
        const bool b = e < L && sz > cap && !meta_uniform(Lmeta[e]);
        big |= (b ? 1u : 0u) << k;
        cnt += b;
        items += b ? sz : 0ull;
    }
    uint64_t nbig, tot_items;
    const uint64_t off0 = block_exclusive_scan<kPlanWaves>(cnt, nbig, sh);
    (void)block_exclusive_scan<kPlanWaves>(items, tot_items, sh);
    const uint32_t tile = pick_tile(tot_items, tile_cap);
    uint64_t off = off0, nt_sum = 0;
    for (int k = 0; k < kPlanPer; ++k) {
        const uint64_t e = e0 + k;
        if (e >= L) break;
        const uint32_t bk = (big >> k) & 1u;
        isbig[e] = bk;
        bigpos[e] = off;
        if (bk) {
            const uint32_t meta = Lmeta[e];
            const uint64_t sz = Lsize[e];
            const uint32_t nt = (uint32_t)((sz + tile - 1) / tile);
            B.start[off] = Lstart[e];
            B.size[off] = sz;
            B.meta[off] = meta;
            B.pos[off] = (uint32_t)next_live(live, kb, (int)meta_depth(meta));
            B.ntiles[off] = nt;
            B.kor[off] = 0ull;
            B.kand[off] = ~0ull;
            nt_sum += nt;
            ++off;
        }
    }
    uint64_t T;
    uint64_t tb = block_exclusive_scan<kPlanWaves>(nt_sum, T, sh);
    for (uint64_t b = off0; b < off; ++b) {
        B.tbase[b] = tb;
        tb += B.ntiles[b];  // written by this thread above
    }
    if (threadIdx.x == 0) {
        *d_nbig = nbig;
        *d_items = tot_items;
        *d_T = T;
        const uint64_t info[4] = {L, nbig, tot_items, n_valid};
        post_write(mbox, seq, 4, info);
    }
}

__global__ __launch_bounds__(kPlanThreads) void k_plan_small(const uint64_t* __restrict__ Lstart,
                                                             const uint64_t* __restrict__ Lsize,
                                                             const uint32_t* __restrict__ Lmeta,
                                                             const uint64_t* __restrict__ dL, uint32_t cap,
                                                             uint32_t* __restrict__ isbig, uint64_t* __restrict__ bigpos,
                                                             PassBuckets B, uint64_t* __restrict__ d_nbig,
                                                             uint64_t* __restrict__ d_items, uint32_t tile_cap,
                                                             const uint64_t* __restrict__ live, int kb,
                                                             uint64_t* __restrict__ d_T,
                                                             const uint64_t* __restrict__ d_nvalid, uint64_t* mbox,
                                                             uint64_t seq) {
    __shared__ uint64_t sh[kPlanWaves];
    plan_list(Lstart, Lsize, Lmeta, *dL, cap, isbig, bigpos, B, d_nbig, d_items, tile_cap, live, kb, d_T, *d_nvalid,
              mbox, seq, sh);
}

// bucket setup + scan of the tile counts.
__global__ __launch_bounds__(kPlanThreads) void k_setup_small(PassBuckets B, uint64_t nbig, uint32_t tile,
                                                              const uint64_t* __restrict__ live, int kb,
                                                              uint64_t* __restrict__ d_T) {
    __shared__ uint64_t sh[kPlanWaves];
    const uint64_t b0 = (uint64_t)threadIdx.x * kPlanPer;
    uint32_t nt[kPlanPer];
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < kPlanPer; ++k) {
        const uint64_t b = b0 + k;
        nt[k] = 0;
        if (b < nbig) {
            B.pos[b] = (uint32_t)next_live(live, kb, (int)meta_depth(B.meta[b]));
            nt[k] = (uint32_t)((B.size[b] + tile - 1) / tile);
            B.ntiles[b] = nt[k];
            B.kor[b] = 0ull;
            B.kand[b] = ~0ull;
        }
        acc += nt[k];
    }
    uint64_t T;
    uint64_t run = block_exclusive_scan<kPlanWaves>(acc, T, sh);
#pragma unroll
    for (int k = 0; k < kPlanPer; ++k) {
        const uint64_t b = b0 + k;
        if (b < nbig) B.tbase[b] = run;
        run += nt[k];
    }
    if (threadIdx.x == 0) *d_T = T;
}

// child count + scan + child write; the new length goes to *d_L and is
// returned.  Entry e belongs to thread e % kPlanThreads in round e /
// kPlanThreads, so a short list spreads over the workgroup.
// Copies the children of entries r0 .. r0+nent-1 (their output offsets
// relative to out0 in roff, in LDS) to out0 .. out0+tot-1, one output slot
// per thread: slot q belongs to the entry whose offset is the last one <= q
// (every entry has at least one child, so the offsets increase strictly).
PLH_DEV void expand_copy(const uint32_t* roff, uint32_t nent, uint64_t r0, uint64_t out0, uint32_t tot,
                         const uint64_t* Lstart, const uint64_t* Lsize, const uint32_t* Lmeta, const uint32_t* isbig,
                         const uint64_t* bigpos, const Children& C, uint64_t* Nstart, uint64_t* Nsize,
                         uint32_t* Nmeta) {
    for (uint32_t q = threadIdx.x; q < tot; q += blockDim.x) {
        uint32_t lo = 0, n = nent;  // last t with roff[t] <= q
        while (n > 1) {
            const uint32_t half = n / 2;
            if (roff[lo + half] <= q) {
                lo += half;
                n -= half;
            } else {
                n = half;
            }
        }
        const uint64_t ee = r0 + lo, out = out0 + q;
        const uint32_t k = q - roff[lo];
        if (!isbig[ee]) {
            Nstart[out] = Lstart[ee];
            Nsize[out] = Lsize[ee];
            Nmeta[out] = Lmeta[ee];
        } else {
            const uint64_t src = bigpos[ee] * kDigits + k;
            Nstart[out] = C.start[src];
            Nsize[out] = C.size[src];
            Nmeta[out] = C.meta[src];
        }
    }
}

// Long lists: one workgroup per kPlanThreads consecutive entries, their
// children copied by all its threads (newpos: exclusive scan of the counts).
__global__ __launch_bounds__(kPlanThreads) void k_expand_write(
    const uint64_t* __restrict__ Lstart, const uint64_t* __restrict__ Lsize, const uint32_t* __restrict__ Lmeta,
    const uint32_t* __restrict__ isbig, const uint64_t* __restrict__ bigpos, Children C,
    const uint64_t* __restrict__ newpos, uint64_t L, uint64_t* __restrict__ Nstart, uint64_t* __restrict__ Nsize,
    uint32_t* __restrict__ Nmeta) {
    __shared__ uint32_t roff[kPlanThreads];
    const uint64_t r0 = (uint64_t)blockIdx.x * kPlanThreads;
    const uint32_t nent = (uint32_t)min<uint64_t>(L - r0, (uint64_t)kPlanThreads);
    const uint64_t base = newpos[r0];
    if (threadIdx.x < nent) roff[threadIdx.x] = (uint32_t)(newpos[r0 + threadIdx.x] - base);  // < 1024 * 256
    __syncthreads();
    const uint64_t last = r0 + nent - 1;
    const uint32_t tot = roff[nent - 1] + expand_count_one(last, isbig, bigpos, C);
    expand_copy(roff, nent, r0, base, tot, Lstart, Lsize, Lmeta, isbig, bigpos, C, Nstart, Nsize, Nmeta);
}

// The round's output slots are written by all threads: slot q of the round
// belongs to the entry whose offset is the last one <= q (every entry has at
// least one child, so the offsets in `roff` increase strictly), so a big
// bucket's up to 256 children are copied in parallel, not by its thread.
PLH_DEV uint64_t expand_list(const uint64_t* Lstart, const uint64_t* Lsize, const uint32_t* Lmeta,
                             const uint32_t* isbig, const uint64_t* bigpos, const Children& C, uint64_t L,
                             uint64_t* Nstart, uint64_t* Nsize, uint32_t* Nmeta, uint64_t* d_L, uint64_t* sh,
                             uint32_t* roff) {
    uint64_t carry = 0;
    for (uint64_t r0 = 0; r0 < L; r0 += kPlanThreads) {
        const uint64_t e = r0 + threadIdx.x;
        const uint64_t c = e < L ? expand_count_one(e, isbig, bigpos, C) : 0ull;
        uint64_t tot;
        const uint64_t o = block_exclusive_scan<kPlanWaves>(c, tot, sh);
        roff[threadIdx.x] = (uint32_t)o;  // < kPlanThreads * kDigits
        __syncthreads();
        const uint32_t nent = (uint32_t)min<uint64_t>(L - r0, (uint64_t)kPlanThreads);
        expand_copy(roff, nent, r0, carry, (uint32_t)tot, Lstart, Lsize, Lmeta, isbig, bigpos, C, Nstart, Nsize,
                    Nmeta);
        carry += tot;
        __syncthreads();  // roff is rewritten by the next round
    }
    if (threadIdx.x == 0) *d_L = carry;
    return carry;
}

__global__ __launch_bounds__(kPlanThreads) void k_expand_small(
    const uint64_t* __restrict__ Lstart, const uint64_t* __restrict__ Lsize, const uint32_t* __restrict__ Lmeta,
    const uint32_t* __restrict__ isbig, const uint64_t* __restrict__ bigpos, Children C, uint64_t L,
    uint64_t* __restrict__ Nstart, uint64_t* __restrict__ Nsize, uint32_t* __restrict__ Nmeta,
    uint64_t* __restrict__ d_L) {
    __shared__ uint64_t sh[kPlanWaves];
    __shared__ uint32_t roff[kPlanThreads];
    (void)expand_list(Lstart, Lsize, Lmeta, isbig, bigpos, C, L, Nstart, Nsize, Nmeta, d_L, sh, roff);
}

// The expansion into the next list and that list's plan in one workgroup
// (the new list's entries, written by all threads, are read back after a
// workgroup barrier).  The plan's big-bucket flags go to a second set of
// list arrays (isbig2, bigpos2): the expansion still reads the first.
// Needs L <= kPlanMax and a next list of at most kPlanMax entries.
__global__ __launch_bounds__(kPlanThreads) void k_expand_plan_small(
    const uint64_t* __restrict__ Lstart, const uint64_t* __restrict__ Lsize, const uint32_t* __restrict__ Lmeta,
    const uint32_t* __restrict__ isbig, const uint64_t* __restrict__ bigpos, Children C, uint64_t L,
    uint64_t* __restrict__ Nstart, uint64_t* __restrict__ Nsize, uint32_t* __restrict__ Nmeta,
    uint64_t* __restrict__ d_L, uint32_t cap, uint32_t* __restrict__ isbig2, uint64_t* __restrict__ bigpos2,
    PassBuckets B2, uint64_t* __restrict__ d_nbig, uint64_t* __restrict__ d_items, uint32_t tile_cap,
    const uint64_t* __restrict__ live, int kb, uint64_t* __restrict__ d_T, const uint64_t* __restrict__ d_nvalid,
    uint64_t* mbox, uint64_t seq) {
    __shared__ uint64_t sh[kPlanWaves];
    __shared__ uint32_t roff[kPlanThreads];
    const uint64_t L2 =
        expand_list(Lstart, Lsize, Lmeta, isbig, bigpos, C, L, Nstart, Nsize, Nmeta, d_L, sh, roff);
    __syncthreads();  // the new list's entries, written by other threads
    plan_list(Nstart, Nsize, Nmeta, L2, cap, isbig2, bigpos2, B2, d_nbig, d_items, tile_cap, live, kb, d_T,
              *d_nvalid, mbox, seq, sh);
}

// --------------------------------------------------------------- leaves
template <int AS, int KW>
struct LeafLds {
    using D = LeafDims<AS, KW>;
    uint64_t key[KW][D::kCap];         // key words (job word first when present, then x, z)
    double val[D::kCap];               // folded coefficients; run sums at head slots
    uint64_t bkey[D::kCap];            // sort keys, in bin order
    uint16_t perm[D::kCap];            // bin order -> slot (identity when all keys are equal)
    uint16_t perm2[D::kCap];           // (key, id) order -> slot; the spare array holds run heads, then rslot
    uint8_t fl[D::kCap];               // raw flags
    uint32_t bin[D::kBins + 1];        // counting-sort bin starts
    unsigned long long diff[KW + 1];   // OR of each key word over the leaf
    unsigned long long dand[KW + 1];   // AND of each key word over the leaf
    uint64_t vmask[KW + 1];            // varying bits of each key word: diff & ~dand
    uint64_t cmask[KW + 1][6];         // compress masks of diff
    int nvary[KW + 1];                 // varying bits per word
    int crun[KW + 1];                  // 1 + shift when a word's varying bits are one run, else 0
    int pos;
    uint32_t nvtot;                    // varying bits in the sort key (<= 64)
    uint32_t pcs;                      // set sort-key bits, summed over the leaf
    uint32_t dup;                      // nonzero when two items share a sort key
    uint32_t red[kLeafWaves];
};

template <int AS, int KW>
PLH_DEV uint64_t lds_word(const LeafLds<AS, KW>& L, uint32_t fmask, int i, int w) {
    if (w < KW) return L.key[w][i];
    if (w == KW) return (uint64_t)(L.fl[i] & fmask) << 56;
    return 0ull;
}

template <int DB, int AS, int KW>
PLH_DEV uint32_t lds_window(const LeafLds<AS, KW>& L, uint32_t fmask, int i, int pos) {
    const int w = pos >> 6, o = pos & 63;
    uint64_t hi = lds_word(L, fmask, i, w) << o;
    if (o + DB > 64) hi |= lds_word(L, fmask, i, w + 1) >> (64 - o);
    return (uint32_t)(hi >> (64 - DB));
}

template <int AS, int KW>
PLH_DEV uint64_t lds_window64(const LeafLds<AS, KW>& L, uint32_t fmask, int i, int pos) {
    const int w = pos >> 6, o = pos & 63;
    uint64_t v = lds_word(L, fmask, i, w) << o;
    if (o) v |= lds_word(L, fmask, i, w + 1) >> (64 - o);
    return v;
}

// Bit compress (Hacker's Delight 7-4): the bits of x under mask m, moved to
// the low end in their order.  The six shift masks depend on m only, so a
// leaf computes them once per key word.
PLH_HD void compress_masks(uint64_t m, uint64_t* mv) {
    uint64_t mk = ~m << 1;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        uint64_t mp = mk ^ (mk << 1);
        mp ^= mp << 2;
        mp ^= mp << 4;
        mp ^= mp << 8;
        mp ^= mp << 16;
        mp ^= mp << 32;
        mv[i] = mp & m;
        m = (m ^ mv[i]) | (mv[i] >> (1 << i));
        mk &= ~mp;
    }
}

PLH_HD uint64_t compress_bits(uint64_t x, uint64_t m, const uint64_t* mv) {
    x &= m;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const uint64_t t = x & mv[i];
        x = (x ^ t) | (t >> (1 << i));
    }
    return x;
}

// Sort key of item i: its key bits at the leaf's varying positions (bits
// where some item differs), most significant first, the first 64 of them,
// left-aligned.  Bits that do not vary inside a leaf do not order it, so
// this is an order-preserving key that is complete when at most 64 bits vary.
template <int AS, int KW>
PLH_DEV uint64_t varying_key(const LeafLds<AS, KW>& L, const uint64_t (&kw)[KW + 1]) {
    uint64_t K = 0;
    int used = 0;
#pragma unroll
    for (int w = 0; w < KW + 1; ++w) {
        const int nv = L.nvary[w];
        if (nv == 0 || used >= 64) continue;
        const uint64_t m = L.vmask[w];
        const uint64_t x = kw[w];
        // a contiguous run of varying bits needs only a mask and a shift
        const uint64_t c = L.crun[w] ? (x & m) >> (L.crun[w] - 1) : compress_bits(x, m, L.cmask[w]);
        const int room = 64 - used;
        if (nv <= room) {
            K = nv == 64 ? c : (K << nv) | c;
            used += nv;
        } else {
            K = (K << room) | (c >> (nv - room));
            used = 64;
        }
    }
    return used == 0 ? 0ull : used < 64 ? K << (64 - used) : K;
}

// Sparse form of a sort key K: the positions of its first 9 set bits, 7 bits
// each, most significant first, as 64 - position (absent bits: 0).  An
// earlier set bit makes K larger, so these codes order as K does on all they
// cover; they spread sparse keys (few set bits) over the counting-sort bins,
// where K's own leading bits are mostly 0.
PLH_DEV uint64_t sparse_code(uint64_t K) {
    uint64_t S = 0;
#pragma unroll
    for (int f = 0; f < 9; ++f) {
        if (!K) break;
        const int b = __clzll(K);
        S |= (uint64_t)(64 - b) << (57 - 7 * f);
        K &= ~(1ull << (63 - b));
    }
    return S;
}

// (key, masked flags, id) order.  A leaf's ids ascend with its slots (the
// passes are stable), so on a full key tie the slot decides.
template <int AS, int KW>
PLH_DEV bool slot_less(const LeafLds<AS, KW>& L, uint32_t fmask, int a, int b) {
#pragma unroll
    for (int w = 0; w < KW; ++w) {
        const uint64_t p = L.key[w][a], q = L.key[w][b];
        if (p != q) return p < q;
    }
    const uint32_t fa = L.fl[a] & fmask, fb = L.fl[b] & fmask;
    if (fa != fb) return fa < fb;
    return a < b;
}

template <int AS, int KW>
PLH_DEV bool slot_key_eq(const LeafLds<AS, KW>& L, uint32_t fmask, int a, int b) {
#pragma unroll
    for (int w = 0; w < KW; ++w)
        if (L.key[w][a] != L.key[w][b]) return false;
    return (L.fl[a] & fmask) == (L.fl[b] & fmask);
}

// Filters of one output term; kw holds the key words (job word first when
// XO = 1).  Per-job limits and the job-alive mask apply to batched sums.
template <int AS, int XO>
PLH_DEV bool keep_term(const Cfg& c, const uint64_t* kw, double acc) {
    if ((c.mode & kDropZero) && acc == 0.0) return false;
    const uint32_t job = XO ? (uint32_t)kw[0] : 0u;
    if (XO && c.job_alive && !c.job_alive[job]) return false;
    if (c.mode & kMaxWeight) {
        int64_t wgt = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) wgt += popc(kw[XO + k] | kw[XO + AS + k]);
        const int64_t mw = (XO && c.job_maxw) ? c.job_maxw[job] : c.max_weight;
        if (mw >= 0 && wgt > mw) return false;
    }
    if (c.mode & kCutoff) {
        const double eps = (XO && c.job_eps) ? c.job_eps[job] : c.eps;
        if (eps > 0.0 && !(fabs(acc) >= eps)) return false;  // cutoff() runs only for eps > 0
    }
    return true;
}

// Output stores are non-temporal: the output is not read again by the
// collation, and streaming it past the caches keeps the operand records
// the leaves gather from resident.
template <int AS, int XO>
PLH_DEV void write_term(const Cfg& c, uint64_t o, const uint64_t* kw, uint32_t fl, double v) {
    for (int w = 0; w < c.out_as; ++w) {
        __builtin_nontemporal_store(kw[XO + w], &c.ox[(int64_t)w * c.ostride + o]);
        __builtin_nontemporal_store(kw[XO + AS + w], &c.oz[(int64_t)w * c.ostride + o]);
    }
    if (XO && c.oj) __builtin_nontemporal_store((uint32_t)kw[0], &c.oj[o]);
    __builtin_nontemporal_store((uint8_t)fl, &c.of[o]);
    __builtin_nontemporal_store(v, &c.ov[o]);
}

// One leaf per workgroup: sort, sum runs, filter, write at out_base.
// out_base = the leaf's start in id space (first try) or its scanned
// offset; a leaf larger than LCAP is a run of identical keys (uniform).
template <int AS, class Src>
__global__ __launch_bounds__(kLeafThreads) void k_leaf(Src src, Cfg c, const uint64_t* __restrict__ Lstart,
                                                       const uint64_t* __restrict__ Lsize,
                                                       const uint32_t* __restrict__ Lmeta,
                                                       const IdKey* __restrict__ ids0,
                                                       const IdKey* __restrict__ ids1,
                                                       uint32_t* __restrict__ kept_out) {
    constexpr int KW = Src::kKW;   // key words before the flags word
    constexpr int XO = Src::kXOff;  // 1 when the job word leads the key
    using D = LeafDims<AS, KW>;
    constexpr int kPer = D::kPer;
    __shared__ LeafLds<AS, KW> L;
    const int tid = threadIdx.x;
    // one leaf per workgroup (grid = the leaf count); neighbouring leaves'
    // outputs are adjacent, so an XCD takes a contiguous run of leaves
    const uint64_t e = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint64_t start = Lstart[e];
    const uint64_t size = Lsize[e];
    const uint32_t meta = Lmeta[e];
    const IdKey* ids = (meta_buf(meta) ? ids1 : ids0) + start;
    const uint64_t out_base = start;

    if (size > (uint64_t)D::kCap) {
        // identical keys, ids ascending (stable passes): one run, summed in
        // id order by one thread from values all threads stage in LDS
        double acc = 0.0;
        for (uint64_t c0 = 0; c0 < size; c0 += D::kCap) {
            const uint32_t m = (uint32_t)min<uint64_t>(size - c0, (uint64_t)D::kCap);
            for (uint32_t i = tid; i < m; i += kLeafThreads) {
                const uint32_t id = ids[c0 + i].id;
                L.val[i] = fold_value(c, src.base_value(id), src.raw_flags(id));
            }
            __syncthreads();
            if (tid == 0)
                for (uint32_t i = 0; i < m; ++i) acc += L.val[i];
            __syncthreads();
        }
        if (tid == 0) {
            uint64_t xz[KW];
            uint32_t fl;
            src.key(ids[0].id, xz, fl);
            const bool keep = keep_term<AS, XO>(c, xz, acc);
            if (keep) write_term<AS, XO>(c, out_base, xz, fl & c.fmask, acc);
            kept_out[e] = keep ? 1u : 0u;
        }
        return;
    }
    const uint32_t n = (uint32_t)size;
    if (tid < KW + 1) {
        L.diff[tid] = 0;
        L.dand[tid] = ~0ull;
    }
    // ---- load: keys, raw flags, folded values (loads of kPer items batched).
    // Each thread keeps its items' key words and flags in registers for the
    // varying-bit masks and the sort keys.
    uint64_t xz[kPer][KW];
    uint32_t fl[kPer];
    {
        uint32_t idv[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t p = tid + k * kLeafThreads;
            idv[k] = p < n ? ids[p].id : 0u;
        }
        // every gather into registers first, then the LDS stores: the
        // stores cannot then hold back the next item's loads
        double bv[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t p = tid + k * kLeafThreads;
            if (p < n) src.key_value(idv[k], xz[k], fl[k], bv[k]);
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t p = tid + k * kLeafThreads;
            if (p < n) {
#pragma unroll
                for (int w = 0; w < KW; ++w) L.key[w][p] = xz[k][w];
                L.fl[p] = (uint8_t)fl[k];
                L.val[p] = fold_value(c, bv[k], fl[k]);
            }
        }
    }
    // ---- bits where the keys differ: set in some item and clear in another
    {
        uint64_t vo[KW + 1], va[KW + 1];
#pragma unroll
        for (int w = 0; w < KW + 1; ++w) {
            vo[w] = 0;
            va[w] = ~0ull;
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t p = tid + k * kLeafThreads;
            if (p < n) {
#pragma unroll
                for (int w = 0; w < KW; ++w) {
                    vo[w] |= xz[k][w];
                    va[w] &= xz[k][w];
                }
                const uint64_t fw = (uint64_t)(fl[k] & 0xFFu & c.fmask) << 56;
                vo[KW] |= fw;
                va[KW] &= fw;
            }
        }
        __syncthreads();  // the accumulators are set; the LDS stores are done
#pragma unroll
        for (int w = 0; w < KW + 1; ++w) {
            const uint64_t o = wave_or(vo[w]);
            const uint64_t a = wave_and(va[w]);
            if (lane_id() == 0) {
                if (o) atomicOr(&L.diff[w], (unsigned long long)o);
                if (~a) atomicAnd(&L.dand[w], (unsigned long long)a);
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        int pos = -1;
        for (int w = 0; w < KW + 1 && pos < 0; ++w) {
            const uint64_t m = L.diff[w] & ~L.dand[w];
            if (m) pos = 64 * w + __clzll(m);
        }
        L.pos = pos;
        L.pcs = 0;
        L.dup = 0;
        int tot = 0;
        for (int w = 0; w < KW + 1; ++w) tot += popc(L.diff[w] & ~L.dand[w]);
        L.nvtot = (uint32_t)min(tot, 64);
    }
    if (tid < KW + 1) {
        const uint64_t m = L.diff[tid] & ~L.dand[tid];
        L.vmask[tid] = m;
        L.nvary[tid] = popc(m);
        compress_masks(m, L.cmask[tid]);
        const int tz = m ? __ffsll((long long)m) - 1 : 0;
        L.crun[tid] = (m && (((m >> tz) + 1) & (m >> tz)) == 0) ? tz + 1 : 0;
    }
    for (int b = tid; b < D::kBins; b += kLeafThreads) L.bin[b] = 0;
    __syncthreads();
    const int pos = L.pos;
    // sp: sorted position -> slot (perm when all keys are equal, else
    // perm2); spare: the other array, free once the order is known
    uint16_t* sp = L.perm;
    uint16_t* spare = L.perm2;
    if (pos < 0) {
        // all keys equal; ids ascend -> identity order
        for (uint32_t p = tid; p < n; p += kLeafThreads) L.perm[p] = (uint16_t)p;
        __syncthreads();
    } else {
        // Sort key: the leaf's varying bits (see varying_key), or for sparse
        // keys (under 1 bit in 8 set) their set-bit positions (sparse_code).
        uint64_t K[kPer];
        {
            uint32_t pc = 0;
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const uint32_t p = tid + k * kLeafThreads;
                K[k] = 0ull;
                if (p < n) {
                    uint64_t kw[KW + 1];
#pragma unroll
                    for (int w = 0; w < KW; ++w) kw[w] = xz[k][w];
                    kw[KW] = (uint64_t)(fl[k] & 0xFFu & c.fmask) << 56;
                    K[k] = varying_key<AS, KW>(L, kw);
                }
                pc += (uint32_t)popc(K[k]);
            }
            pc = wave_sum(pc);
            if (lane_id() == 0) atomicAdd(&L.pcs, pc);
        }
        __syncthreads();
        const bool sparse = 8u * L.pcs < L.nvtot * n;
        // With few varying bits the key is complete in the low-padded sort
        // key, and the slot (= id order) fills the padding: every sort key
        // is then distinct and already carries the (key, id) order.
        const bool complete = !sparse && L.nvtot + D::kSlotBits <= 64;
        // complete sort keys carry the slot, so they never tie even where
        // the keys do: run heads must then be found from the keys
        if (complete && tid == 0) L.dup = 1u;
        // counting sort on the kDl leading bits of the sort key ...
        uint32_t dig[kPer], slot[kPer];
        uint64_t S[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t p = tid + k * kLeafThreads;
            dig[k] = 0;
            slot[k] = 0;
            S[k] = 0;
            if (p < n) {
                S[k] = sparse ? sparse_code(K[k]) : complete ? (K[k] | p) : K[k];
                dig[k] = (uint32_t)(S[k] >> (64 - D::kDl));
                slot[k] = atomicAdd(&L.bin[dig[k]], 1u);
            }
        }
        __syncthreads();
        {
            constexpr int kBper = (D::kBins + kLeafThreads - 1) / kLeafThreads;
            uint32_t v[kBper];
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < kBper; ++k) {
                const int b = tid * kBper + k;
                v[k] = b < D::kBins ? L.bin[b] : 0u;
                acc += v[k];
            }
            uint32_t tot;
            uint32_t run = block_exclusive_scan<kLeafWaves>(acc, tot, L.red);
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
            const uint32_t p = tid + k * kLeafThreads;
            if (p < n) {
                const uint32_t q = L.bin[dig[k]] + slot[k];
                L.perm[q] = (uint16_t)p;
                L.bkey[q] = S[k];
            }
        }
        __syncthreads();
        // ... then each item's rank inside its bin on the full (key, id).
        // Work is assigned by bin-ordered position, so the lanes of a wave
        // mostly share a bin.  The sort keys, read in bin order four at a
        // time, decide every comparison but ties, which compare the whole
        // key (and then the slot).
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t q = tid + k * kLeafThreads;
            if (q < n) {
                const int me = L.perm[q];
                const uint64_t ms = L.bkey[q];
                const uint32_t b = (uint32_t)(ms >> (64 - D::kDl));
                const uint32_t s0 = L.bin[b], s1 = L.bin[b + 1];
                uint32_t r = s0;
                if (s1 - s0 > 1) {
                    uint32_t lt = 0, eq = 0;
                    uint32_t t = s0;
                    for (; t + 4 <= s1; t += 4) {
                        const uint64_t a0 = L.bkey[t], a1 = L.bkey[t + 1], a2 = L.bkey[t + 2], a3 = L.bkey[t + 3];
                        lt += (uint32_t)(a0 < ms) + (uint32_t)(a1 < ms) + (uint32_t)(a2 < ms) + (uint32_t)(a3 < ms);
                        eq += (uint32_t)(a0 == ms) + (uint32_t)(a1 == ms) + (uint32_t)(a2 == ms) + (uint32_t)(a3 == ms);
                    }
                    for (; t < s1; ++t) {
                        const uint64_t a = L.bkey[t];
                        lt += (uint32_t)(a < ms);
                        eq += (uint32_t)(a == ms);
                    }
                    r += lt;
                    if (eq > 1) {  // other items share this sort key (eq counts this one too)
                        L.dup = 1u;
                        for (t = s0; t < s1; ++t)
                            if (t != q && L.bkey[t] == ms && slot_less<AS, KW>(L, c.fmask, L.perm[t], me)) ++r;
                    }
                }
                L.perm2[r] = (uint16_t)me;
            }
        }
        __syncthreads();
        sp = L.perm2;
        spare = L.perm;
    }
    // ---- values in sorted order and run heads, by all threads (the sort
    // keys and the second permutation are free until the output step)
    // When no two items share a sort key (a sort key is a function of the
    // key and the masked flags), every item is a run of its own: the run
    // heads and the sorted values are not needed.
    const bool distinct = pos >= 0 && L.dup == 0u;
    double* vs = reinterpret_cast<double*>(L.bkey);
    uint16_t* head = spare;
    if (!distinct) {
        for (uint32_t q = tid; q < n; q += kLeafThreads) {
            const int sq = sp[q];
            vs[q] = L.val[sq];
            head[q] = (q == 0 || !slot_key_eq<AS, KW>(L, c.fmask, sq, sp[q - 1])) ? 1 : 0;
        }
        __syncthreads();
    }
    // ---- runs: thread t owns sorted positions [t*kPer, t*kPer + kPer);
    // each run is summed acc = 0; acc += v in (key, id) order
    // Slots are indexed by k (compile-time after unrolling), never by a
    // running count, so sums[] and kept[] stay in registers.
    double sums[kPer];
    bool kept[kPer];
    uint32_t nk = 0;
    const uint32_t q0 = (uint32_t)tid * kPer;
    if (distinct) {
        // Items are owned strided (q = tid + k * kLeafThreads), so that when
        // every item is kept, item q goes to out_base + q straight from
        // registers and a wave's stores are consecutive.
        double dsum[kPer];
        bool dkeep[kPer];
        uint64_t dw[kPer][KW];
        uint32_t dfl[kPer];
        uint32_t dn = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t q = tid + k * kLeafThreads;
            dkeep[k] = false;
            dsum[k] = 0.0;
            dfl[k] = 0;
            if (q >= n) continue;
            const int s = sp[q];
            double acc = 0.0;
            acc += L.val[s];
#pragma unroll
            for (int w = 0; w < KW; ++w) dw[k][w] = L.key[w][s];
            dfl[k] = L.fl[s] & c.fmask;
            if (keep_term<AS, XO>(c, dw[k], acc)) {
                dsum[k] = acc;
                dkeep[k] = true;
                ++dn;
            }
        }
        uint32_t dtotal;
        (void)block_exclusive_scan<kLeafWaves>(dn, dtotal, L.red);
        if (dtotal == n) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const uint32_t q = tid + k * kLeafThreads;
                if (q >= n) continue;
                const uint64_t o = out_base + q;
                for (int w = 0; w < c.out_as; ++w) {
                    __builtin_nontemporal_store(dw[k][XO + w], &c.ox[(int64_t)w * c.ostride + o]);
                    __builtin_nontemporal_store(dw[k][XO + AS + w], &c.oz[(int64_t)w * c.ostride + o]);
                }
                if (XO && c.oj) __builtin_nontemporal_store((uint32_t)dw[k][0], &c.oj[o]);
                __builtin_nontemporal_store((uint8_t)dfl[k], &c.of[o]);
                __builtin_nontemporal_store(dsum[k], &c.ov[o]);
            }
            if (tid == 0) kept_out[e] = n;
            return;
        }
        // Some items are dropped: keep marks by sorted position (in the free
        // sort-key array) and the sums by slot, then the kept items of the
        // contiguous positions below, as for runs.
        uint8_t* kmark = reinterpret_cast<uint8_t*>(L.bkey);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t q = tid + k * kLeafThreads;
            if (q >= n) continue;
            kmark[q] = dkeep[k] ? 1 : 0;
            if (dkeep[k]) L.val[sp[q]] = dsum[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t q = q0 + k;
            kept[k] = q < n && kmark[q] != 0;
            sums[k] = kept[k] ? L.val[sp[q]] : 0.0;
            nk += kept[k] ? 1u : 0u;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t q = q0 + k;
            kept[k] = false;
            sums[k] = 0.0;
            if (q >= n || !head[q]) continue;
            double acc = 0.0;
            acc += vs[q];
            for (uint32_t r = q + 1; r < n;) {
                if (r + 4 <= n) {
                    const uint16_t h0 = head[r], h1 = head[r + 1], h2 = head[r + 2], h3 = head[r + 3];
                    const double v0 = vs[r], v1 = vs[r + 1], v2 = vs[r + 2], v3 = vs[r + 3];
                    if (h0) break;
                    acc += v0;
                    if (h1) break;
                    acc += v1;
                    if (h2) break;
                    acc += v2;
                    if (h3) break;
                    acc += v3;
                    r += 4;
                } else {
                    if (head[r]) break;
                    acc += vs[r];
                    ++r;
                }
            }
            uint64_t xz[KW];
            const int s = sp[q];
#pragma unroll
            for (int w = 0; w < KW; ++w) xz[w] = L.key[w][s];
            if (keep_term<AS, XO>(c, xz, acc)) {
                sums[k] = acc;
                kept[k] = true;
                ++nk;
            }
        }
    }
    uint32_t total;
    const uint32_t r0 = block_exclusive_scan<kLeafWaves>(nk, total, L.red);
    // the scan's barriers order the run sums (and their reads of head[])
    // before these writes; sp is unchanged since the sort
    uint16_t* rslot = spare;
    {
        uint32_t r = r0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (!kept[k]) continue;
            const int s = sp[q0 + k];
            rslot[r++] = (uint16_t)s;
            L.val[s] = sums[k];
        }
    }
    __syncthreads();
    for (uint32_t r = tid; r < total; r += kLeafThreads) {
        const int s = rslot[r];
        const uint64_t o = out_base + r;
        for (int w = 0; w < c.out_as; ++w) {
            __builtin_nontemporal_store(L.key[XO + w][s], &c.ox[(int64_t)w * c.ostride + o]);
            __builtin_nontemporal_store(L.key[XO + AS + w][s], &c.oz[(int64_t)w * c.ostride + o]);
        }
        if (XO && c.oj) __builtin_nontemporal_store((uint32_t)L.key[0][s], &c.oj[o]);
        __builtin_nontemporal_store((uint8_t)(L.fl[s] & c.fmask), &c.of[o]);
        __builtin_nontemporal_store(L.val[s], &c.ov[o]);
    }
    if (tid == 0) kept_out[e] = total;
}

// Compaction after the first write: leaf e's kept terms move from its start
// to its scanned offset (out of place).
__global__ __launch_bounds__(kBlock) void k_compact(const uint64_t* __restrict__ Lstart,
                                                    const uint32_t* __restrict__ kept,
                                                    const uint64_t* __restrict__ offs, int as,
                                                    const uint64_t* __restrict__ sx, const uint64_t* __restrict__ sz,
                                                    const uint8_t* __restrict__ sf, const double* __restrict__ sv,
                                                    int64_t sstride, uint64_t* __restrict__ dx,
                                                    uint64_t* __restrict__ dz, uint8_t* __restrict__ df,
                                                    double* __restrict__ dv, int64_t dstride,
                                                    const uint32_t* __restrict__ sj, uint32_t* __restrict__ dj) {
    const uint64_t e = blockIdx.x;
    const uint64_t s0 = Lstart[e], d0 = offs[e];
    const uint32_t k = kept[e];
    for (uint32_t i = threadIdx.x; i < k; i += kBlock) {
        for (int w = 0; w < as; ++w) {
            dx[(int64_t)w * dstride + d0 + i] = sx[(int64_t)w * sstride + s0 + i];
            dz[(int64_t)w * dstride + d0 + i] = sz[(int64_t)w * sstride + s0 + i];
        }
        df[d0 + i] = sf[s0 + i];
        dv[d0 + i] = sv[s0 + i];
        if (dj) dj[d0 + i] = sj[s0 + i];
    }
}

// ------------------------------------------------------- live key bits
// live[w] = OR over the items of key word w (a superset is fine: bits that
// never vary cost at most a histogram pass).  The flags word's live bits
// are preset by the mode; the rest is reduced on the device, nothing is
// read back.
struct LiveWords {
    uint64_t w[kMaxKeyWords + 1];
};

__global__ void k_live_init(uint64_t* __restrict__ live, int nwords, LiveWords preset) {
    const int w = threadIdx.x;
    if (w < nwords) live[w] = preset.w[w];
}

// One grid row per plane of a sum: x words go to key word xo + k, z words to
// xo + AS + k, the job plane (when j is set) to key word 0.
__global__ __launch_bounds__(kBlock) void k_live_or(SumView s, const uint32_t* __restrict__ j, int as, int key_as,
                                                    int xo, int64_t n, unsigned long long* __restrict__ live) {
    const int r = blockIdx.y;
    uint64_t acc = 0;
    int kw;
    if (r < 2 * as) {
        const bool isx = r < as;
        const int k = isx ? r : r - as;
        const uint64_t* p = (isx ? s.x : s.z) + (int64_t)k * s.stride;
        for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock)
            acc |= p[t];
        kw = xo + (isx ? 0 : key_as) + k;
    } else {
        for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock)
            acc |= j[t];
        kw = 0;
    }
    acc = wave_or(acc);
    if (lane_id() == 0 && acc) atomicOr(&live[kw], (unsigned long long)acc);
}

// The root's children as the first bucket list (buffer 0), grouped like a
// split bucket's, and that list's plan, in one workgroup: wave 0 groups the
// root's sub-buckets (wave-local) and writes the list length and the number
// of valid items, then all threads plan the list.
__global__ __launch_bounds__(kPlanThreads) void k_root_list_plan(
    const uint64_t* __restrict__ sub_start, const uint64_t* __restrict__ sub_size, uint32_t cap,
    const uint64_t* __restrict__ live, int kb, uint64_t* __restrict__ Nstart, uint64_t* __restrict__ Nsize,
    uint32_t* __restrict__ Nmeta, uint64_t* __restrict__ d_L, uint64_t* __restrict__ d_nvalid,
    uint32_t* __restrict__ isbig, uint64_t* __restrict__ bigpos, PassBuckets B, uint64_t* __restrict__ d_nbig,
    uint64_t* __restrict__ d_items, uint32_t tile_cap, uint64_t* __restrict__ d_T, uint64_t* mbox, uint64_t seq) {
    __shared__ GroupShared g;
    __shared__ uint64_t sh[kPlanWaves];
    __shared__ uint64_t s_len, s_valid;
    if (threadIdx.x < kWave) {
        const int lane = (int)threadIdx.x;
        uint64_t st[4], ss[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            st[k] = sub_start[4 * lane + k];
            ss[k] = sub_size[4 * lane + k];
        }
        const int pos = root_pos(live, kb);
        const int depth = min(pos + kDigitBits, kb);
        const uint32_t uni = next_live(live, kb, depth) >= kb ? 1u : 0u;
        const uint32_t n = group_children(st, ss, cap, g, [&](uint32_t o, uint64_t s0, uint64_t sz, uint32_t parts) {
            Nstart[o] = s0;
            Nsize[o] = sz;
            Nmeta[o] = parts == 1 ? make_meta((uint32_t)depth, 0u, uni) : make_meta((uint32_t)pos, 0u, 0u);
        });
        if (lane == 0) {
            const uint64_t nv = n ? g.P[kDigits] : 0ull;
            *d_L = n;
            *d_nvalid = nv;
            s_len = n;
            s_valid = nv;
        }
    }
    __syncthreads();  // the list, written by wave 0
    plan_list(Nstart, Nsize, Nmeta, s_len, cap, isbig, bigpos, B, d_nbig, d_items, tile_cap, live, kb, d_T, s_valid,
              mbox, seq, sh);
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

// The root pass by lhs groups applies to products of at least 2^22 pairs
// with 2^11..2^22 lhs terms and 2^9 rhs rows or more (below that the
// grouping and its read of the live mask cost more than the scatter saves),
// and only when the root digit lies in the x/z words: some x/z bit before
// the words' last 8 is live.
template <int AS>
int root_by_groups_ok(plh_ctx* ctx, const PairSource<AS, 0>& src, uint64_t n_dense, const uint64_t* d_live,
                      bool* ok) {
    *ok = false;
    if (n_dense < (1ull << 22) || src.n_lo < (1u << 11) || src.n_lo > (1u << 22) || src.n_hi < (1u << 9))
        return PLH_OK;
    uint64_t w[2 * AS];
    PLH_TRY(read_u64n(ctx, d_live, 2 * AS, w));
    for (int k = 0; k < 2 * AS; ++k)
        if (w[k]) {
            *ok = 64 * k + __builtin_clzll(w[k]) + kDigitBits <= 2 * AS * 64;
            break;
        }
    return PLH_OK;
}

template <int AS>
int root_by_groups(plh_ctx* ctx, const PairSource<AS, 0>& src, const uint64_t* d_live, int kb, IdKey* ids0,
                   uint64_t* sub_start, uint64_t* sub_size) {
    hipStream_t st = ctx->stream;
    const uint32_t n = src.n_lo, m = src.n_hi;
    const size_t lbytes = ((size_t)n * 8 + 255) & ~(size_t)255;
    const size_t gbytes = ((size_t)n + 255) & ~(size_t)255;
    void* p = nullptr;
    PLH_TRY(ctx->ws.get(kWsGroup, lbytes + gbytes + 2 * kDigits * 4, &p));
    uint64_t* lrec = (uint64_t*)p;
    uint8_t* lgrp = (uint8_t*)p + lbytes;
    uint32_t* gcnt = (uint32_t*)((char*)p + lbytes + gbytes);
    const uint64_t nb = (m + kGrpRows - 1) / kGrpRows;
    PLH_TRY(ctx->ws.get(kWsHist, nb * kDigits * 4, &p));
    uint32_t* hist = (uint32_t*)p;
    // chunks of the lhs positions: about 4096 workgroups, at least 1024
    // positions each
    uint64_t nc = std::max<uint64_t>(1, (4096 + nb - 1) / nb);
    nc = std::min<uint64_t>(nc, std::max<uint64_t>(1, n / 1024));
    const uint32_t chunk = (uint32_t)(((n + nc - 1) / nc + kBlock - 1) / kBlock * kBlock);
    nc = (n + chunk - 1) / chunk;
    {
        ProfScope ps(ctx, "msd_hist_first");
        k_group_lhs<AS><<<1, kGrpThreads, 0, st>>>(src.L, n, d_live, kb, lrec, lgrp, gcnt);
        k_group_hist<AS><<<dim3((unsigned)nb), kBlock, 0, st>>>(src.R, m, d_live, kb, gcnt, hist);
    }
    {
        ProfScope ps(ctx, "colscan");
        k_root_colscan<<<1, kColThreads, 0, st>>>(nb, hist, sub_start, sub_size);
    }
    {
        ProfScope ps(ctx, "msd_scatter_first", 8.0 * (double)n * (double)m);
        k_group_gen<AS><<<dim3((unsigned)nb, (unsigned)nc), kBlock, 0, st>>>(src, m, n, chunk, d_live, kb, lrec, lgrp,
                                                                            gcnt, hist, sub_start, ids0);
    }
    PLH_HIP(hipGetLastError());
    return PLH_OK;
}

template <int AS, class Src>
int collate_run(plh_ctx* ctx, const Src& src, uint64_t n_dense, const CollateSpec& spec, plh_sum* out,
                int out_as) {
    // The caller has filled the live mask: misc[0 .. kKW] of kWsMisc.
    const uint32_t mode = spec.mode;
    const int64_t max_weight = spec.max_weight;
    const double eps = spec.eps;
    using D = LeafDims<AS, Src::kKW>;
    hipStream_t st = ctx->stream;
    constexpr int kWordsK = Src::kKW + 1;
    const int kb = 64 * kWordsK;
    const uint32_t cap = (uint32_t)D::kCap;
    out->n = 0;
    if (n_dense == 0) return PLH_OK;
    if (n_dense > (1ull << 32)) {
        set_error("collation supports at most 2^32 items");
        return PLH_ERR_SIZE;
    }
    void* p;
    PLH_TRY(ctx->ws.get(kWsIds0, n_dense * sizeof(IdKey), &p));
    IdKey* ids0 = (IdKey*)p;
    PLH_TRY(ctx->ws.get(kWsIds1, n_dense * sizeof(IdKey), &p));
    IdKey* ids1 = (IdKey*)p;
    PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
    uint64_t* misc = (uint64_t*)p;  // [0..33] live, [40] total, [56..61] pass counts, root position
    uint64_t* d_live = misc;
    uint64_t* d_total = misc + 40;
    // pass bookkeeping read back once per pass: list length, big buckets,
    // their items, valid items; then tiles and the root position
    uint64_t* d_L = misc + 56;
    uint64_t* d_nbig = misc + 57;
    uint64_t* d_items = misc + 58;
    uint64_t* d_nvalid = misc + 59;
    uint64_t* d_T = misc + 60;

    Cfg cfg{};
    cfg.live = d_live;
    cfg.kb = kb;
    cfg.mode = mode;
    cfg.fmask = (mode & kFold) ? 1u : 3u;
    cfg.max_weight = max_weight;
    cfg.eps = eps;
    cfg.out_as = out_as;
    cfg.job_eps = spec.job_eps;
    cfg.job_maxw = spec.job_maxw;
    cfg.job_alive = spec.job_alive;

    // ---- pass 0: implicit ids, one root bucket (its digit position is the
    // first live bit, found by each kernel from the live mask).
    PLH_TRY(ctx->ws.get(kWsSub, 2 * kDigits * 8, &p));
    uint64_t* sub_start = (uint64_t*)p;
    uint64_t* sub_size = sub_start + kDigits;
    uint32_t* hist = nullptr;  // per-tile digit histograms (the later passes reuse the slot)
    bool grouped = false;
    if constexpr (GroupRoot<Src>::value) {
        PLH_TRY(root_by_groups_ok<AS>(ctx, src, n_dense, d_live, &grouped));
        if (grouped) PLH_TRY(root_by_groups<AS>(ctx, src, d_live, kb, ids0, sub_start, sub_size));
    }
    if (!grouped) {
        const uint32_t tile0 =
            pick_tile(n_dense, ctx->tune.tile_root_max ? ctx->tune.tile_root_max : kTileMaxImplicit);
        const uint64_t t0 = (n_dense + tile0 - 1) / tile0;
        PLH_TRY(ctx->ws.get(kWsHist, t0 * kDigits * 4, &p));
        hist = (uint32_t*)p;
        {
            ProfScope ps(ctx, "msd_hist_first");
            k_root_hist<Src><<<dim3((unsigned)t0), kBlock, 0, st>>>(src, n_dense, d_live, kb, tile0, hist);
        }
        {
            ProfScope ps(ctx, "colscan");
            k_root_colscan<<<1, kColThreads, 0, st>>>(t0, hist, sub_start, sub_size);
        }
        {
            ProfScope ps(ctx, "msd_scatter_first", 8.0 * (double)n_dense);
            k_root_scatter<Src><<<dim3((unsigned)t0), kBlock, 0, st>>>(src, ids0, n_dense, d_live, kb, tile0, hist,
                                                                        sub_start);
        }
    }
    PLH_HIP(hipGetLastError());
    // list arrays (isbig u32, cnt u32, bigpos u64, bigsz u64, newpos u64) for a
    // list bound of lb entries, in one of two alternating slots
    struct ListArrays {
        uint32_t* isbig;
        uint32_t* cnt;
        uint64_t* bigpos;
        uint64_t* bigsz;
        uint64_t* newpos;
    };
    auto list_arrays = [&](int slot, uint64_t lb, ListArrays* la) -> int {
        void* q = nullptr;
        PLH_TRY(ctx->ws.get(slot, lb * 32 + 256, &q));
        la->isbig = (uint32_t*)q;
        la->cnt = la->isbig + lb;
        la->bigpos = (uint64_t*)((char*)q + ((lb * 8 + 15) & ~(uint64_t)15));
        la->bigsz = la->bigpos + lb;
        la->newpos = la->bigsz + lb;
        return PLH_OK;
    };
    // the big-bucket arrays, sized by the bound lb before the counts are known
    auto pass_buckets = [&](uint64_t lb, PassBuckets* pb) -> int {
        void* q = nullptr;
        PLH_TRY(ctx->ws.get(kWsPassB, lb * 56 + 256, &q));
        pb->start = (uint64_t*)q;
        pb->size = pb->start + lb;
        pb->tbase = pb->size + lb;
        pb->kor = pb->tbase + lb;
        pb->kand = pb->kor + lb;
        pb->meta = (uint32_t*)(pb->kand + lb);
        pb->pos = pb->meta + lb;
        pb->ntiles = pb->pos + lb;
        pb->skip = pb->ntiles + lb;
        return PLH_OK;
    };
    const uint32_t tile_cap = ctx->tune.tile_pass_max ? ctx->tune.tile_pass_max : kTileMaxExplicit;
    int pl_slot = kWsPassL;
    Post post;
    // ---- the root's children as the first list, and that list's plan
    Lists A, B;
    PLH_TRY(get_lists(ctx, kWsListA, kDigits, &A));
    {
        ListArrays LA0;
        PLH_TRY(list_arrays(pl_slot, kDigits, &LA0));
        PassBuckets PB0;
        PLH_TRY(pass_buckets(kDigits, &PB0));
        PLH_TRY(post_prepare(ctx, &post));
        k_root_list_plan<<<1, kPlanThreads, 0, st>>>(sub_start, sub_size, cap, d_live, kb, A.start, A.size, A.meta,
                                                     d_L, d_nvalid, LA0.isbig, LA0.bigpos, PB0, d_nbig, d_items,
                                                     tile_cap, d_T, post.mbox, post.seq);
    }
    PLH_HIP(hipGetLastError());
    int cur_slot = kWsListA;
    uint64_t Lb = kDigits;  // bound on the list length
    uint64_t L = 0, n_valid = 0;

    // ---- global passes until every non-uniform bucket fits a leaf.  Grids
    // of the list kernels use the bound Lb; the exact counts come back in
    // one read per pass.
    bool planned = true;  // the pass's plan ran inside the previous kernel (root list or expansion)
    for (int guard = 0; guard < 2 * kb; ++guard) {
        ListArrays LA;
        PLH_TRY(list_arrays(pl_slot, Lb, &LA));
        uint32_t* isbig = LA.isbig;
        uint32_t* cnt = LA.cnt;
        uint64_t* bigpos = LA.bigpos;
        uint64_t* bigsz = LA.bigsz;
        uint64_t* newpos = LA.newpos;
        uint64_t info[4];
        PassBuckets PB;
        PLH_TRY(pass_buckets(Lb, &PB));
        // the plan kernel also sets the buckets up (and so does the fused expansion)
        const bool plan_small = planned || Lb <= kPlanMax;
        if (planned) {
            // plan and read-back already issued
        } else if (plan_small) {
            PLH_TRY(post_prepare(ctx, &post));
            k_plan_small<<<1, kPlanThreads, 0, st>>>(A.start, A.size, A.meta, d_L, cap, isbig, bigpos, PB, d_nbig,
                                                    d_items, tile_cap, d_live, kb, d_T, d_nvalid, post.mbox,
                                                    post.seq);
        } else {
            k_classify<<<grid1(Lb), kBlock, 0, st>>>(A.size, A.meta, d_L, Lb, cap, isbig, bigsz);
            PLH_TRY(scan_u32_to_u64(ctx, isbig, bigpos, Lb, d_nbig));
            PLH_TRY(scan_u64(ctx, bigsz, newpos, Lb, d_items));
            k_gather_big<<<grid1(Lb), kBlock, 0, st>>>(A.start, A.size, A.meta, isbig, bigpos, Lb, PB);
        }
        PLH_HIP(hipGetLastError());
        if (plan_small)
            PLH_TRY(post_wait(ctx, post, 4, info));
        else
            PLH_TRY(read_u64n(ctx, d_L, 4, info));
        planned = false;
        L = info[0];
        n_valid = info[3];
        const uint64_t nbig = info[1], pass_items = info[2];
        if (nbig == 0) break;
        const uint32_t tile = pick_tile(pass_items, tile_cap);
        if (plan_small) {
            // set up by k_plan_small
        } else if (nbig <= kPlanMax) {
            k_setup_small<<<1, kPlanThreads, 0, st>>>(PB, nbig, tile, d_live, kb, d_T);
        } else {
            k_bucket_setup<<<grid1(nbig), kBlock, 0, st>>>(PB, nbig, tile, d_live, kb);
            PLH_TRY(scan_u32_to_u64(ctx, PB.ntiles, PB.tbase, nbig, d_T));
        }
        const uint64_t Tb = pass_items / tile + nbig;  // >= sum of ceil(size / tile)
        PLH_TRY(ctx->ws.get(kWsHist, Tb * kDigits * 4, &p));
        hist = (uint32_t*)p;
        PLH_TRY(ctx->ws.get(kWsSub, nbig * kDigits * 16, &p));
        sub_start = (uint64_t*)p;
        sub_size = sub_start + nbig * kDigits;
        {
            ProfScope ps(ctx, "msd_hist", 8.0 * (double)pass_items);
            k_pass_hist<Src><<<dim3((unsigned)Tb), kBlock, 0, st>>>(src, ids0, ids1, PB, (uint32_t)nbig, d_T, tile,
                                                                     hist, d_live, kb);
        }
        PLH_TRY(ctx->ws.get(kWsChild, nbig * kDigits * 20 + nbig * 4 + 256, &p));
        Children CH;
        CH.start = (uint64_t*)p;
        CH.size = CH.start + nbig * kDigits;
        CH.meta = (uint32_t*)(CH.size + nbig * kDigits);
        CH.count = CH.meta + nbig * kDigits;
        {
            ProfScope ps(ctx, "colscan");
            if (pass_items / tile >= 64 * nbig)  // many tiles per bucket
                k_pass_colscan<kColGroups><<<dim3((unsigned)nbig), kColThreads, 0, st>>>(
                    PB, (uint32_t)nbig, d_T, hist, sub_start, sub_size, d_live, kb, cap, CH);
            else
                k_pass_colscan<1><<<dim3((unsigned)nbig), kDigits, 0, st>>>(PB, (uint32_t)nbig, d_T, hist, sub_start,
                                                                            sub_size, d_live, kb, cap, CH);
        }
        {
            ProfScope ps(ctx, "msd_scatter", 16.0 * (double)pass_items);
            k_pass_scatter<Src><<<dim3((unsigned)Tb), kBlock, 0, st>>>(src, ids0, ids1, PB, (uint32_t)nbig, d_T,
                                                                        tile, hist, sub_start, d_live, kb);
        }
        PLH_HIP(hipGetLastError());
        // each split bucket has at most kDigits children; every entry holds >= 1 item
        const uint64_t Lb2 = std::min<uint64_t>(L + (kDigits - 1) * nbig, n_valid);
        const int nslot = cur_slot == kWsListA ? kWsListB : kWsListA;
        const int npl = pl_slot == kWsPassL ? kWsPassL2 : kWsPassL;
        PLH_TRY(get_lists(ctx, nslot, Lb2, &B));
        if (L <= kPlanMax && Lb2 <= kPlanMax) {
            // expansion + the next pass's plan (its read-back posted)
            ListArrays LA2;
            PLH_TRY(list_arrays(npl, Lb2, &LA2));
            PassBuckets PB2;
            PLH_TRY(pass_buckets(Lb2, &PB2));
            PLH_TRY(post_prepare(ctx, &post));
            k_expand_plan_small<<<1, kPlanThreads, 0, st>>>(A.start, A.size, A.meta, isbig, bigpos, CH, L, B.start,
                                                           B.size, B.meta, d_L, cap, LA2.isbig, LA2.bigpos, PB2,
                                                           d_nbig, d_items, tile_cap, d_live, kb, d_T, d_nvalid,
                                                           post.mbox, post.seq);
            planned = true;
        } else if (L <= kPlanMax) {
            k_expand_small<<<1, kPlanThreads, 0, st>>>(A.start, A.size, A.meta, isbig, bigpos, CH, L, B.start, B.size,
                                                      B.meta, d_L);
        } else {
            k_expand_count<<<grid1(L), kBlock, 0, st>>>(isbig, bigpos, CH, L, cnt);
            PLH_TRY(scan_u32_to_u64(ctx, cnt, newpos, L, d_L));
            k_expand_write<<<dim3((unsigned)((L + kPlanThreads - 1) / kPlanThreads)), kPlanThreads, 0, st>>>(
                A.start, A.size, A.meta, isbig, bigpos, CH, newpos, L, B.start, B.size, B.meta);
        }
        PLH_HIP(hipGetLastError());
        std::swap(A, B);
        cur_slot = nslot;
        pl_slot = npl;
        Lb = Lb2;
    }

    if (L == 0) return PLH_OK;  // no valid item (out->n is 0)
    // ---- leaves: sort, sum, filter, write at each leaf's start.
    PLH_TRY(sum_reserve(out, (int64_t)n_valid, 0));
    cfg.ox = out->x;
    cfg.oz = out->z;
    cfg.of = out->f;
    cfg.ov = out->v;
    cfg.oj = out->j;
    cfg.ostride = out->cap;
    PLH_TRY(ctx->ws.get(kWsKept, L * 4 + 64, &p));
    uint32_t* kept = (uint32_t*)p;
    {
        ProfScope ps(ctx, "leaf", (double)n_valid * (16.0 * out_as + 9.0));
        k_leaf<AS, Src><<<dim3((unsigned)L), kLeafThreads, 0, st>>>(src, cfg, A.start, A.size, A.meta, ids0,
                                                                    ids1, kept);
    }
    PLH_HIP(hipGetLastError());
    PLH_TRY(ctx->ws.get(kWsOffs, L * 8 + 64, &p));
    uint64_t* offs = (uint64_t*)p;
    uint64_t total = 0;
    PLH_TRY(scan_u32_to_u64_read(ctx, kept, offs, L, d_total, &total));
    if (total != n_valid) {
        // Runs collapsed or terms were dropped: compact out of place.
        plh_sum tmp;
        tmp.ctx = ctx;
        tmp.as = out->as;
        tmp.nq = out->nq;
        tmp.has_job = out->has_job;
        PLH_TRY(sum_reserve(&tmp, (int64_t)total, 0));
        if (total) {
            ProfScope ps(ctx, "compact", (double)total * 2.0 * (16.0 * out_as + 9.0));
            k_compact<<<dim3((unsigned)L), kBlock, 0, st>>>(A.start, kept, offs, out_as, out->x, out->z, out->f,
                                                            out->v, out->cap, tmp.x, tmp.z, tmp.f, tmp.v, tmp.cap,
                                                            out->j, tmp.j);
            PLH_HIP(hipGetLastError());
        }
        sum_take_storage(out, &tmp);
    }
    out->n = (int64_t)total;
    return PLH_OK;
}

// Pack a sum into records of 2*AS + 2 words (x[AS], z[AS], value,
// flags | job << 32); words beyond the sum's own array size are zero.  With
// live set, the same pass ORs the sum's words into the collation's live
// mask (x words at key word xo, z words at xo + AS, the job plane at 0).
template <int AS>
__global__ __launch_bounds__(kBlock) void k_pack(SumView s, const uint32_t* __restrict__ j, int as, int64_t n,
                                                 uint64_t* __restrict__ rec, unsigned long long* __restrict__ live,
                                                 int xo) {
    uint64_t ox[AS], oz[AS];
    uint32_t oj = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k) ox[k] = oz[k] = 0ull;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) {
        uint64_t* r = rec + (size_t)t * (2 * AS + 2);
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            const uint64_t x = k < as ? s.x[(int64_t)k * s.stride + t] : 0ull;
            const uint64_t z = k < as ? s.z[(int64_t)k * s.stride + t] : 0ull;
            r[k] = x;
            r[AS + k] = z;
            ox[k] |= x;
            oz[k] |= z;
        }
        const uint32_t jt = j ? j[t] : 0u;
        oj |= jt;
        r[2 * AS] = (uint64_t)__double_as_longlong(s.v[t]);
        r[2 * AS + 1] = (uint64_t)s.f[t] | ((uint64_t)jt << 32);
    }
    if (!live) return;
    // OR over the wave, then over the workgroup in LDS, then one global
    // atomic per word per workgroup
    __shared__ unsigned long long acc[2 * AS + 1];
    if (threadIdx.x < 2 * AS + 1) acc[threadIdx.x] = 0ull;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        const uint64_t a = wave_or(ox[k]), b = wave_or(oz[k]);
        if (lane_id() == 0) {
            if (a) atomicOr(&acc[k], (unsigned long long)a);
            if (b) atomicOr(&acc[AS + k], (unsigned long long)b);
        }
    }
    oj = wave_or(oj);
    if (lane_id() == 0 && oj) atomicOr(&acc[2 * AS], (unsigned long long)oj);
    __syncthreads();
    if (threadIdx.x < 2 * AS) {
        const int w = (int)threadIdx.x;
        const uint64_t v = acc[w];
        if (v) atomicOr(&live[w < AS ? xo + w : xo + AS + (w - AS)], (unsigned long long)v);
    } else if (threadIdx.x == 2 * AS && j && acc[2 * AS]) {
        atomicOr(&live[0], acc[2 * AS]);
    }
}

template <int AS>
int pack_sum(plh_ctx* ctx, const plh_sum* in, int slot, const uint64_t** out, uint64_t* live = nullptr,
             int xo = 0) {
    void* p;
    PLH_TRY(ctx->ws.get(slot, (size_t)(in->n ? in->n : 1) * (2 * AS + 2) * 8, &p));
    if (in->n) {
        const unsigned g = (unsigned)std::min<uint64_t>(grid1((uint64_t)in->n), 512);
        k_pack<AS><<<g, kBlock, 0, ctx->stream>>>(SumView{in->x, in->z, in->f, in->v, in->cap},
                                                  in->has_job ? in->j : nullptr, in->as, in->n, (uint64_t*)p,
                                                  (unsigned long long*)live, xo);
    }
    PLH_HIP(hipGetLastError());
    *out = (const uint64_t*)p;
    return PLH_OK;
}

// Records of a collated sum beyond this many bytes are not made (the
// leaf then reads the planes).
constexpr uint64_t kMaxPackBytes = 16ull << 30;

int padded_as(int as) {
    if (as <= 1) return 1;
    if (as == 2) return 2;
    if (as <= 4) return 4;
    if (as <= 8) return 8;
    if (as <= 16) return 16;
    return -1;
}

// Live mask of a collation: every key word preset from `preset` (or zero),
// the flags word's bits set by the mode.
int live_init(plh_ctx* ctx, int key_words, uint32_t mode, uint64_t** live, const LiveWords* preset = nullptr) {
    void* p;
    PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
    *live = (uint64_t*)p;
    LiveWords w{};
    if (preset) w = *preset;
    w.w[key_words] = (mode & kFold) ? (1ull << 56) : (3ull << 56);
    k_live_init<<<1, 64, 0, ctx->stream>>>(*live, key_words + 1, w);
    PLH_HIP(hipGetLastError());
    return PLH_OK;
}

// OR a sum's planes into the live mask (x words at key word xo, z words at
// xo + key_as, the job plane at key word 0 when the sum has one).
int live_or(plh_ctx* ctx, const plh_sum* s, int key_as, int xo, uint64_t* live) {
    if (s->n == 0) return PLH_OK;
    const int rows = 2 * s->as + (s->has_job ? 1 : 0);
    const unsigned gx = (unsigned)std::min<uint64_t>(grid1((uint64_t)s->n), 512);
    k_live_or<<<dim3(gx, (unsigned)rows), kBlock, 0, ctx->stream>>>(SumView{s->x, s->z, s->f, s->v, s->cap},
                                                                      s->has_job ? s->j : nullptr, s->as, key_as, xo,
                                                                      s->n, (unsigned long long*)live);
    PLH_HIP(hipGetLastError());
    return PLH_OK;
}

template <int AS, bool JOB>
int collate_terms_as(plh_ctx* ctx, const plh_sum* in, const CollateSpec& spec, plh_sum* out) {
    TermSource<AS, JOB> src;
    src.s = SumView{in->x, in->z, in->f, in->v, in->cap};
    src.j = in->j;
    src.rec = nullptr;
    src.as = in->as;
    src.mode = spec.mode;
    src.n = (uint32_t)in->n;
    uint64_t* live;
    PLH_TRY(live_init(ctx, TermSource<AS, JOB>::kKW, spec.mode, &live));
    if ((uint64_t)in->n * TermSource<AS, JOB>::kRec * 8 <= kMaxPackBytes)
        PLH_TRY(pack_sum<AS>(ctx, in, kWsPackL, &src.rec, live, TermSource<AS, JOB>::kXOff));
    else
        PLH_TRY(live_or(ctx, in, AS, TermSource<AS, JOB>::kXOff, live));
    return collate_run<AS>(ctx, src, (uint64_t)in->n, spec, out, in->as);
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
    if ((n_hi << lo_bits) > (1ull << 32)) {
        set_error("collated product: the pair ids (rows << ceil(log2(cols))) exceed 2^32");
        return PLH_ERR_SIZE;
    }
    PairSource<AS, KIND> src;
    // a product's words are XORs of the operands' words: OR(L) | OR(R) covers them
    uint64_t* live;
    PLH_TRY(live_init(ctx, 2 * AS, spec.mode, &live));
    PLH_TRY(pack_sum<AS>(ctx, Ls, kWsPackL, &src.L.r, live, 0));
    PLH_TRY(pack_sum<AS>(ctx, Rs, kWsPackR, &src.R.r, live, 0));
    src.Lp = SumView{Ls->x, Ls->z, Ls->f, Ls->v, Ls->cap};
    src.Rp = SumView{Rs->x, Rs->z, Rs->f, Rs->v, Rs->cap};
    src.as = Ls->as;
    src.mode = spec.mode;
    src.lo_bits = lo_bits;
    src.n_lo = (uint32_t)n_lo;
    src.n_hi = (uint32_t)n_hi;
    src.div_lo = make_fastdiv((uint32_t)n_lo);
    src.by_i = (uint32_t)by_i;
    return collate_run<AS>(ctx, src, n_lo * n_hi, spec, out, Ls->as);
}

// Collation of a sum after a rotation's general path (RotSource): 2n ids,
// the new terms' bits are the old ones XOR P, so OR(sum) | P covers them.
template <int AS, bool JOB>
int collate_rotation_as(plh_ctx* ctx, const plh_sum* in, const RotParams<AS>& rp, const JobRotTables& jt,
                        const uint64_t* p_or, const CollateSpec& spec, plh_sum* out) {
    using S = RotSource<AS, JOB>;
    S src;
    src.s = SumView{in->x, in->z, in->f, in->v, in->cap};
    src.j = in->j;
    src.rec = nullptr;
    src.as = in->as;
    src.mode = spec.mode;
    src.n = (uint32_t)in->n;
    src.rp = rp;
    src.jt = jt;
    LiveWords add{};
    for (int k = 0; k < 2 * AS; ++k) add.w[S::kXOff + k] = p_or[k];
    uint64_t* live;
    PLH_TRY(live_init(ctx, S::kKW, spec.mode, &live, &add));
    if ((uint64_t)in->n * S::kRec * 8 <= kMaxPackBytes)
        PLH_TRY(pack_sum<AS>(ctx, in, kWsPackL, &src.rec, live, S::kXOff));
    else
        PLH_TRY(live_or(ctx, in, AS, S::kXOff, live));
    return collate_run<AS>(ctx, src, 2 * (uint64_t)in->n, spec, out, in->as);
}

// Runs f(dst) with dst = out, or a scratch sum when out aliases in, whose
// storage then replaces in's.
template <class F>
int with_output(plh_ctx* ctx, plh_sum* in, plh_sum* out, F&& f) {
    plh_sum tmp;
    plh_sum* dst = out;
    const bool alias = (out == in);
    if (alias) {
        tmp.ctx = ctx;
        tmp.as = in->as;
        tmp.nq = in->nq;
        tmp.has_job = in->has_job;
        dst = &tmp;
    }
    dst->as = in->as;
    dst->nq = in->nq;
    dst->has_job = in->has_job;
    const int rc = f(dst);
    if (rc != PLH_OK) {
        if (alias) sum_free_storage(&tmp);
        return rc;
    }
    dst->as = in->as;
    dst->nq = in->nq;
    if (alias) {
        const int64_t n = tmp.n;
        sum_take_storage(in, &tmp);
        in->n = n;
    }
    return PLH_OK;
}

}  // namespace

// ------------------------------------------------------- product batches
// One concatenated record array for the lhs (or rhs) sums of a batch: term
// t of the concatenation belongs to the sum whose prefix range holds it.
struct PackJob {
    const uint64_t* x;
    const uint64_t* z;
    const uint8_t* f;
    const double* v;
    int64_t stride;
    uint32_t first;  // first term in the concatenation
    uint32_t n;
};

template <int AS>
__global__ __launch_bounds__(kBlock) void k_pack_multi(const PackJob* __restrict__ jobs, uint32_t njobs, int as,
                                                       uint64_t total, uint64_t* __restrict__ rec,
                                                       unsigned long long* __restrict__ live) {
    uint64_t ox[AS], oz[AS];
#pragma unroll
    for (int k = 0; k < AS; ++k) ox[k] = oz[k] = 0ull;
    for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (uint64_t)gridDim.x * kBlock) {
        uint32_t lo = 0, hi = njobs;  // last job with first <= t
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].first <= t) lo = mid; else hi = mid;
        }
        const PackJob J = jobs[lo];
        const uint64_t i = t - J.first;
        uint64_t* r = rec + t * (2 * AS + 2);
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            const uint64_t x = k < as ? J.x[(int64_t)k * J.stride + i] : 0ull;
            const uint64_t z = k < as ? J.z[(int64_t)k * J.stride + i] : 0ull;
            r[k] = x;
            r[AS + k] = z;
            ox[k] |= x;
            oz[k] |= z;
        }
        r[2 * AS] = (uint64_t)__double_as_longlong(J.v[i]);
        r[2 * AS + 1] = (uint64_t)J.f[i];
    }
    __shared__ unsigned long long acc[2 * AS];
    if (threadIdx.x < 2 * AS) acc[threadIdx.x] = 0ull;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        const uint64_t a = wave_or(ox[k]), b = wave_or(oz[k]);
        if (lane_id() == 0) {
            if (a) atomicOr(&acc[k], (unsigned long long)a);
            if (b) atomicOr(&acc[AS + k], (unsigned long long)b);
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * AS && acc[threadIdx.x]) atomicOr(&live[1 + threadIdx.x], acc[threadIdx.x]);
}

// Job ranges of a job-ordered batched sum: offsets[k] = first term of job k.
__global__ void k_job_offsets(const uint32_t* __restrict__ J, int64_t n, uint32_t njobs, int64_t* __restrict__ off) {
    const uint32_t k = (uint32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k > njobs) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (J[mid] < k) lo = mid + 1; else hi = mid;
    }
    off[k] = lo;
}

// Destination of one job when a batched sum is split into per-job sums.
struct SplitDst {
    uint64_t* x;
    uint64_t* z;
    uint8_t* f;
    double* v;
    int64_t stride;
    int64_t off;  // the job's first term in the batch
};

// Term t of a job-ordered batch goes to its job's sum at t - off.
__global__ __launch_bounds__(kBlock) void k_split_jobs(SumView b, const uint32_t* __restrict__ J, int as, int64_t n,
                                                       const SplitDst* __restrict__ dst) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) {
        const SplitDst d = dst[J[t]];
        const int64_t i = t - d.off;
        for (int w = 0; w < as; ++w) {
            d.x[(int64_t)w * d.stride + i] = b.x[(int64_t)w * b.stride + t];
            d.z[(int64_t)w * d.stride + i] = b.z[(int64_t)w * b.stride + t];
        }
        d.f[i] = b.f[t];
        d.v[i] = b.v[t];
    }
}

template <int AS>
int collate_products_as(plh_ctx* ctx, int njobs, const plh_sum* const* lhs, const plh_sum* const* rhs,
                        const CollateSpec& spec, plh_sum* out, uint64_t total_pairs) {
    using S = MultiPairSource<AS>;
    hipStream_t st = ctx->stream;
    // host tables: jobs, pack descriptors, id chunks
    std::vector<ProductJob> jobs((size_t)njobs);
    std::vector<PackJob> pl((size_t)njobs), pr((size_t)njobs);
    uint64_t off = 0, lt = 0, rt = 0;
    for (int k = 0; k < njobs; ++k) {
        const plh_sum* a = lhs[k];
        const plh_sum* b = rhs[k];
        jobs[(size_t)k] = ProductJob{off, (uint32_t)a->n, (uint32_t)lt, (uint32_t)rt,
                                     make_fastdiv(a->n > 0 ? (uint32_t)a->n : 1u)};
        pl[(size_t)k] = PackJob{a->x, a->z, a->f, a->v, a->cap, (uint32_t)lt, (uint32_t)a->n};
        pr[(size_t)k] = PackJob{b->x, b->z, b->f, b->v, b->cap, (uint32_t)rt, (uint32_t)b->n};
        off += (uint64_t)a->n * (uint64_t)b->n;
        lt += (uint64_t)a->n;
        rt += (uint64_t)b->n;
    }
    const uint64_t nchunks = (total_pairs >> S::kChunkBits) + 1;
    std::vector<uint32_t> chunk((size_t)nchunks);
    {
        uint32_t k = 0;
        for (uint64_t c = 0; c < nchunks; ++c) {
            const uint64_t id = c << S::kChunkBits;
            while (k + 1 < (uint32_t)njobs && jobs[(size_t)k + 1].off <= id) ++k;
            chunk[(size_t)c] = k;
        }
    }
    const size_t b_jobs = (size_t)njobs * sizeof(ProductJob), b_pack = (size_t)njobs * sizeof(PackJob);
    const size_t o_pl = (b_jobs + 255) & ~(size_t)255, o_pr = o_pl + ((b_pack + 255) & ~(size_t)255),
                 o_ch = o_pr + ((b_pack + 255) & ~(size_t)255);
    void* p;
    PLH_TRY(ctx->ws.get(kWsJobs, o_ch + nchunks * 4 + 256, &p));
    char* dj = (char*)p;
    PLH_HIP(hipMemcpyAsync(dj, jobs.data(), b_jobs, hipMemcpyHostToDevice, st));
    PLH_HIP(hipMemcpyAsync(dj + o_pl, pl.data(), b_pack, hipMemcpyHostToDevice, st));
    PLH_HIP(hipMemcpyAsync(dj + o_pr, pr.data(), b_pack, hipMemcpyHostToDevice, st));
    PLH_HIP(hipMemcpyAsync(dj + o_ch, chunk.data(), nchunks * 4, hipMemcpyHostToDevice, st));
    // live mask: the job word's bits up to the last job, the operands' words, the mode's flags
    LiveWords preset{};
    uint64_t jm = 0;
    while (jm < (uint64_t)(njobs - 1)) jm = (jm << 1) | 1ull;
    preset.w[0] = jm;
    uint64_t* live;
    PLH_TRY(live_init(ctx, S::kKW, spec.mode, &live, &preset));
    PLH_TRY(ctx->ws.get(kWsPackL, (size_t)(lt ? lt : 1) * S::kRec * 8, &p));
    uint64_t* lrec = (uint64_t*)p;
    PLH_TRY(ctx->ws.get(kWsPackR, (size_t)(rt ? rt : 1) * S::kRec * 8, &p));
    uint64_t* rrec = (uint64_t*)p;
    const int as = lhs[0]->as;
    if (lt)
        k_pack_multi<AS><<<(unsigned)std::min<uint64_t>(grid1(lt), 1024), kBlock, 0, st>>>(
            (const PackJob*)(dj + o_pl), (uint32_t)njobs, as, lt, lrec, (unsigned long long*)live);
    if (rt)
        k_pack_multi<AS><<<(unsigned)std::min<uint64_t>(grid1(rt), 1024), kBlock, 0, st>>>(
            (const PackJob*)(dj + o_pr), (uint32_t)njobs, as, rt, rrec, (unsigned long long*)live);
    PLH_HIP(hipGetLastError());
    S src;
    src.L = lrec;
    src.R = rrec;
    src.jobs = (const ProductJob*)dj;
    src.chunk_job = (const uint32_t*)(dj + o_ch);
    src.njobs = (uint32_t)njobs;
    src.mode = spec.mode;
    // the host tables above are read by the copies before collate_run's first read-back returns
    return collate_run<AS>(ctx, src, total_pairs, spec, out, as);
}

int collate_rotation(plh_ctx* ctx, plh_sum* in, const uint64_t* px, const uint64_t* pz, uint32_t psf, double c,
                     double sn, const CollateSpec& spec, plh_sum* out) {
    if (in->n > (int64_t)0x7FFFFFFFll) {
        set_error("a fused rotation supports at most 2^31 - 1 terms");
        return PLH_ERR_SIZE;
    }
    if (in->has_job) {
        set_error("collate_rotation: batched sums rotate per job");
        return PLH_ERR_ARG;
    }
    const int pa = padded_as(in->as);
    uint64_t p_or[2 * 16] = {0};
    int rc = with_output(ctx, in, out, [&](plh_sum* dst) -> int {
        switch (pa) {
#define PLH_ROT(ASV)                                                                              \
    case ASV: {                                                                                   \
        RotParams<ASV> rp{};                                                                      \
        for (int k = 0; k < in->as; ++k) {                                                        \
            rp.px[k] = px[k];                                                                     \
            rp.pz[k] = pz[k];                                                                     \
            p_or[k] = px[k];                                                                      \
            p_or[ASV + k] = pz[k];                                                                \
        }                                                                                         \
        rp.psf = psf;                                                                             \
        rp.c = c;                                                                                 \
        rp.s = sn;                                                                                \
        return collate_rotation_as<ASV, false>(ctx, in, rp, JobRotTables{}, p_or, spec, dst);   \
    }
            PLH_ROT(1)
            PLH_ROT(2)
            PLH_ROT(4)
            PLH_ROT(8)
            PLH_ROT(16)
#undef PLH_ROT
            default: set_error("unsupported array_size"); return PLH_ERR_AS;
        }
    });
    if (rc != PLH_OK) return rc;
    out->canonical = (spec.mode & kFold) && (spec.mode & kDropZero);
    out->sorted = out->canonical;
    return PLH_OK;
}

int collate_rotation_jobs(plh_ctx* ctx, plh_sum* in, const JobRotTables& jt, const uint64_t* p_or,
                          const CollateSpec& spec, plh_sum* out) {
    if (in->n > (int64_t)0x7FFFFFFFll) {
        set_error("a fused rotation supports at most 2^31 - 1 terms");
        return PLH_ERR_SIZE;
    }
    if (!in->has_job) {
        set_error("collate_rotation_jobs: the sum has no job plane");
        return PLH_ERR_ARG;
    }
    const int pa = padded_as(in->as);
    int rc = with_output(ctx, in, out, [&](plh_sum* dst) -> int {
        if (pa == 1) return collate_rotation_as<1, true>(ctx, in, RotParams<1>{}, jt, p_or, spec, dst);
        if (pa == 2) return collate_rotation_as<2, true>(ctx, in, RotParams<2>{}, jt, p_or, spec, dst);
        set_error("batched sums support up to 128 qubits");
        return PLH_ERR_AS;
    });
    if (rc != PLH_OK) return rc;
    out->canonical = (spec.mode & kFold) && (spec.mode & kDropZero);
    out->sorted = out->canonical;
    return PLH_OK;
}

int collate_products(plh_ctx* ctx, int njobs, const plh_sum* const* lhs, const plh_sum* const* rhs,
                     const CollateSpec& spec, plh_sum* out, int64_t* offsets) {
    if (njobs <= 0) {
        set_error("collate_products: no jobs");
        return PLH_ERR_ARG;
    }
    const int as = lhs[0]->as;
    uint64_t total = 0;
    for (int k = 0; k < njobs; ++k) {
        if (lhs[k]->as != as || rhs[k]->as != as) {
            set_error("collate_products: every operand of a batch must have the same array size");
            return PLH_ERR_ARG;
        }
        if (lhs[k] == out || rhs[k] == out) {
            set_error("collate_products: the output must not alias an operand");
            return PLH_ERR_ARG;
        }
        total += (uint64_t)lhs[k]->n * (uint64_t)rhs[k]->n;
    }
    if (total > (1ull << 32)) {
        set_error("collate_products: more than 2^32 pairs in one batch (split it)");
        return PLH_ERR_SIZE;
    }
    sum_free_storage(out);
    out->as = as;
    out->has_job = true;
    out->n = 0;
    int rc;
    switch (padded_as(as)) {
        case 1: rc = collate_products_as<1>(ctx, njobs, lhs, rhs, spec, out, total); break;
        case 2: rc = collate_products_as<2>(ctx, njobs, lhs, rhs, spec, out, total); break;
        default: set_error("product batches support up to 128 qubits"); return PLH_ERR_AS;
    }
    if (rc != PLH_OK) return rc;
    if (out->n > 0 && !out->j) {
        set_error("collate_products: output lost its job plane");
        return PLH_ERR_HIP;
    }
    out->canonical = (spec.mode & kFold) && (spec.mode & kDropZero);
    out->sorted = out->canonical;
    void* p;
    PLH_TRY(ctx->ws.get(kWsTmpSum, ((size_t)njobs + 1) * 8 + 64, &p));
    int64_t* doff = (int64_t*)p;
    if (out->n > 0) {
        k_job_offsets<<<(unsigned)((njobs + 1 + 255) / 256), 256, 0, ctx->stream>>>(out->j, out->n, (uint32_t)njobs,
                                                                                    doff);
        PLH_HIP(hipGetLastError());
        PLH_HIP(hipMemcpyAsync(offsets, doff, ((size_t)njobs + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
        PLH_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        for (int k = 0; k <= njobs; ++k) offsets[k] = 0;
    }
    return PLH_OK;
}

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
        tmp.has_job = in->has_job;
        dst = &tmp;
    }
    dst->as = in->as;
    dst->nq = in->nq;
    dst->has_job = in->has_job;
    int rc;
    if (in->has_job) {
        switch (padded_as(in->as)) {
            case 1: rc = collate_terms_as<1, true>(ctx, in, spec, dst); break;
            case 2: rc = collate_terms_as<2, true>(ctx, in, spec, dst); break;
            default: set_error("batched sums support up to 128 qubits"); return PLH_ERR_AS;
        }
    } else {
        switch (padded_as(in->as)) {
            case 1: rc = collate_terms_as<1, false>(ctx, in, spec, dst); break;
            case 2: rc = collate_terms_as<2, false>(ctx, in, spec, dst); break;
            case 4: rc = collate_terms_as<4, false>(ctx, in, spec, dst); break;
            case 8: rc = collate_terms_as<8, false>(ctx, in, spec, dst); break;
            case 16: rc = collate_terms_as<16, false>(ctx, in, spec, dst); break;
            default: set_error("unsupported array_size"); return PLH_ERR_AS;
        }
    }
    if (rc != PLH_OK) {
        if (alias) sum_free_storage(&tmp);
        return rc;
    }
    dst->as = in->as;
    dst->nq = in->nq;
    if (alias) {
        const int64_t n = tmp.n;
        sum_take_storage(in, &tmp);
        in->n = n;
    }
    out->canonical = (spec.mode & kFold) && (spec.mode & kDropZero);
    out->sorted = out->canonical;
    return PLH_OK;
}

int split_jobs(plh_ctx* ctx, const plh_sum* batch, int njobs, const int64_t* offsets, plh_sum* const* outs) {
    if (!batch->has_job || (batch->n > 0 && !batch->j)) {
        set_error("split_jobs: the batch has no job plane");
        return PLH_ERR_ARG;
    }
    if (offsets[0] != 0 || offsets[njobs] != batch->n) {
        set_error("split_jobs: offsets do not cover the batch");
        return PLH_ERR_ARG;
    }
    std::vector<SplitDst> d((size_t)njobs);
    for (int k = 0; k < njobs; ++k) {
        const int64_t cnt = offsets[k + 1] - offsets[k];
        if (cnt < 0) {
            set_error("split_jobs: offsets decrease");
            return PLH_ERR_ARG;
        }
        plh_sum* o = outs[k];
        o->as = batch->as;
        o->has_job = false;
        PLH_TRY(sum_reserve(o, cnt, 0));
        o->n = cnt;
        o->canonical = batch->canonical;
        o->sorted = batch->sorted;
        d[(size_t)k] = SplitDst{o->x, o->z, o->f, o->v, o->cap, offsets[k]};
    }
    if (batch->n == 0) return PLH_OK;
    void* p;
    PLH_TRY(ctx->ws.get(kWsJobs, (size_t)njobs * sizeof(SplitDst) + 256, &p));
    PLH_HIP(hipMemcpyAsync(p, d.data(), (size_t)njobs * sizeof(SplitDst), hipMemcpyHostToDevice, ctx->stream));
    const uint64_t n = (uint64_t)batch->n;
    k_split_jobs<<<(unsigned)std::min<uint64_t>(grid1(n), 4096), kBlock, 0, ctx->stream>>>(
        SumView{batch->x, batch->z, batch->f, batch->v, batch->cap}, batch->j, batch->as, batch->n,
        (const SplitDst*)p);
    PLH_HIP(hipGetLastError());
    // the host table is read by the copy before the caller's synchronisation returns
    PLH_HIP(hipStreamSynchronize(ctx->stream));
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
    out->sorted = out->canonical;
    return PLH_OK;
}

}  // namespace plh
