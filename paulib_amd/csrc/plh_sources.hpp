// plh_sources.hpp -- key sources for the collation engine.
//
// Collation sorts 32-bit item ids; a source turns an id into the item's
// sort key, coefficient and output term.  The key is the reference's
// sort_and_combine key (include/PauliSumOperations.h:376-386): x words from
// word 0, then z words, then the flags byte, all unsigned and most
// significant first.  It is addressed as key words W[0..2AS]:
//   W[w] = x[w] (w < AS), z[w-AS] (AS <= w < 2AS), flags << 56 (w = 2AS),
// and a global bit position g counts from the most significant bit of W[0]
// (g = 64*w + 63 - bit).
//
//   TermSource  a materialised sum in HBM; id = term index.
//   PairSource  the all-pairs product (or commutator) of two sums in HBM,
//               computed on the fly: id = (hi << lo_bits) | lo, whose
//               numeric order is the reference's output order.
#pragma once
#include "plh_common.hpp"

namespace plh {

// Collation modes (flags into the engine).
enum : uint32_t {
    kFold = 1u,       // canonicalize: fold sign bit into the coefficient
    kDropZero = 2u,   // canonicalize: drop exact-zero sums
    kMaxWeight = 4u,  // truncate_weight filter on output terms
    kCutoff = 8u,     // cutoff filter (|c| >= eps) on output terms
};

struct SumView {  // plane-major SoA: x[k*stride + t]
    const uint64_t* x;
    const uint64_t* z;
    const uint8_t* f;
    const double* v;
    int64_t stride;
};

// JOB = true: a batched sum whose terms carry a job id (plane j[t]); the job
// id is the leading key word, so one collation keeps every job's terms
// contiguous and sorted exactly as that job alone would be.
template <int AS, bool JOB = false>
struct TermSource {
    static constexpr int kAS = AS;
    static constexpr int kXOff = JOB ? 1 : 0;    // index of x[0] in the key
    static constexpr int kKW = 2 * AS + kXOff;  // key words before the flags word
    static constexpr int kRec = 2 * AS + 2;     // packed record words
    SumView s;
    const uint32_t* j;    // job plane (JOB only)
    const uint64_t* rec;  // the terms packed as records (x[AS], z[AS], value,
                          // flags | job << 32), or null: whole keys come from
                          // one record instead of one load per plane
    int as;               // real words per plane (<= AS)
    uint32_t mode;        // kFold ...
    uint32_t n;           // items

    PLH_DEV bool valid(uint32_t) const { return true; }
    PLH_DEV uint32_t dense_to_id(uint32_t t) const { return t; }
    PLH_DEV uint32_t flag(uint32_t id) const {
        const uint32_t fl = s.f[id];
        return (mode & kFold) ? (fl & 1u) : fl;
    }
    PLH_DEV uint64_t word(uint32_t id, int w) const {
        if (JOB) {
            if (w == 0) return j[id];
            w -= 1;
        }
        if (w < AS) return w < as ? s.x[(int64_t)w * s.stride + id] : 0ull;
        if (w < 2 * AS) return (w - AS) < as ? s.z[(int64_t)(w - AS) * s.stride + id] : 0ull;
        if (w == 2 * AS) return (uint64_t)flag(id) << 56;
        return 0ull;
    }
    // Key words and the RAW flags byte (the engine masks the sign bit).
    PLH_DEV void key(uint32_t id, uint64_t* kw, uint32_t& fl) const {
        if (rec) {
            const uint64_t* r = rec + (size_t)id * kRec;
            const uint64_t tail = r[kRec - 1];
            if (JOB) kw[0] = tail >> 32;
#pragma unroll
            for (int k = 0; k < AS; ++k) {
                kw[kXOff + k] = r[k];
                kw[kXOff + AS + k] = r[AS + k];
            }
            fl = (uint32_t)tail & 0xFFu;
            return;
        }
        if (JOB) kw[0] = j[id];
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            kw[kXOff + k] = k < as ? s.x[(int64_t)k * s.stride + id] : 0ull;
            kw[kXOff + AS + k] = k < as ? s.z[(int64_t)k * s.stride + id] : 0ull;
        }
        fl = s.f[id];
    }
    PLH_DEV uint32_t raw_flags(uint32_t id) const {
        return rec ? (uint32_t)rec[(size_t)id * kRec + kRec - 1] & 0xFFu : s.f[id];
    }
    // Coefficient before the sign fold.
    PLH_DEV double base_value(uint32_t id) const {
        return rec ? __longlong_as_double((long long)rec[(size_t)id * kRec + 2 * AS]) : s.v[id];
    }
    PLH_DEV void key_value(uint32_t id, uint64_t* kw, uint32_t& fl, double& v) const {
        key(id, kw, fl);
        v = base_value(id);
    }
};

// One rotation exp(-i theta/2 P) on the general path of evolve
// (include/PauliSumOperations.h:148-180): ps = P with the i bit flipped,
// c = cos(theta), s = sin(theta) (full angle, from the host).
template <int AS>
struct RotParams {
    uint64_t px[AS], pz[AS];
    uint32_t psf;  // flags of ps
    double c, s;
};

// Per-job rotations of a batched sum (device tables indexed by job slot):
// kind 4 = general path this step; P as x[AS] then z[AS] per slot.
struct JobRotTables {
    const uint8_t* kind;
    const uint64_t* p;
    const uint8_t* psf;
    const double* c;
    const double* s;
};

// Per-job rotation windows of a sorted server step (device tables by job
// slot): wn[slot] rotations, entry slot * kJobWindow + k with its rotation
// kind (0 no change, 4 general), ps flags, cos, sin and P (x[AS] then z[AS]).
constexpr int kJobWindow = 16;
struct JobWindows {
    const uint8_t* wn;
    const uint8_t* kind;
    const uint64_t* p;
    const uint8_t* psf;
    const double* c;
    const double* s;
};

// A sum after evolve's general path, without materialising the new terms.
// id < n: term id, its value times c when it anticommutes with P; id >= n:
// the new term (term id - n) * ps with value (-s) * v, valid only when that
// term anticommutes with P.  The numeric id order is the reference's order
// after the new terms are appended (old terms, then new ones in old-index
// order), so sort_and_combine over these ids sums exactly as it does.
// JOB: a batched sum; each term rotates by its job's P when that job's
// rotation is on the general path this step, else it passes through.
template <int AS, bool JOB = false>
struct RotSource {
    static constexpr int kAS = AS;
    static constexpr int kXOff = JOB ? 1 : 0;
    static constexpr int kKW = 2 * AS + kXOff;
    static constexpr int kRec = 2 * AS + 2;
    SumView s;
    const uint32_t* j;    // job plane (JOB only)
    const uint64_t* rec;  // the old terms as records (x[AS], z[AS], value, flags | job << 32)
    int as;
    uint32_t mode;
    uint32_t n;           // old terms; ids run over [0, 2n)
    RotParams<AS> rp;     // single sum
    JobRotTables jt;      // batched sum

    PLH_DEV uint32_t dense_to_id(uint32_t t) const { return t; }
    // The rotation that applies to old term t: false when none (batched sums).
    PLH_DEV bool rot_of(uint32_t t, const uint64_t*& px, const uint64_t*& pz, uint32_t& psf, double& c,
                        double& sn) const {
        if (!JOB) {
            px = rp.px;
            pz = rp.pz;
            psf = rp.psf;
            c = rp.c;
            sn = rp.s;
            return true;
        }
        const uint32_t job = rec ? (uint32_t)(rec[(size_t)t * kRec + kRec - 1] >> 32) : j[t];
        if (jt.kind[job] != 4) return false;
        px = jt.p + (size_t)job * 2 * AS;
        pz = px + AS;
        psf = jt.psf[job];
        c = jt.c[job];
        sn = jt.s[job];
        return true;
    }
    PLH_DEV void load(uint32_t t, uint64_t* x, uint64_t* z, uint32_t& fl, double& v, uint32_t& job) const {
        if (rec) {
            const uint64_t* r = rec + (size_t)t * kRec;
#pragma unroll
            for (int k = 0; k < AS; ++k) {
                x[k] = r[k];
                z[k] = r[AS + k];
            }
            v = __longlong_as_double((long long)r[2 * AS]);
            fl = (uint32_t)r[kRec - 1] & 0xFFu;
            job = (uint32_t)(r[kRec - 1] >> 32);
            return;
        }
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            x[k] = k < as ? s.x[(int64_t)k * s.stride + t] : 0ull;
            z[k] = k < as ? s.z[(int64_t)k * s.stride + t] : 0ull;
        }
        v = s.v[t];
        fl = s.f[t];
        job = JOB ? j[t] : 0u;
    }
    PLH_DEV static bool anti(const uint64_t* x, const uint64_t* z, const uint64_t* px, const uint64_t* pz) {
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) acc ^= (x[k] & pz[k]) ^ (z[k] & px[k]);
        return popc(acc) & 1;
    }
    // The item's key words (x, z) and raw flags, and its coefficient before
    // the sign fold; false for a new-term id whose old term commutes.
    PLH_DEV bool item(uint32_t id, uint64_t* x, uint64_t* z, uint32_t& fl, double& v, uint32_t& job) const {
        const bool nw = id >= n;
        const uint32_t t = nw ? id - n : id;
        load(t, x, z, fl, v, job);
        const uint64_t *px, *pz;
        uint32_t psf;
        double c, sn;
        if (!rot_of(t, px, pz, psf, c, sn)) return !nw;
        const bool a = anti(x, z, px, pz);
        if (!nw) {
            if (a) v = v * c;
            return true;
        }
        if (!a) return false;
        uint64_t c1 = 0, c2 = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            phase_word(x[k], z[k], px[k], pz[k], c1, c2);
            x[k] ^= px[k];
            z[k] ^= pz[k];
        }
        fl = phase_of(fl, psf, c1, c2);
        v = (-sn) * v;
        return true;
    }
    PLH_DEV bool valid(uint32_t id) const {
        if (id < n) return true;
        uint64_t x[AS], z[AS];
        uint32_t fl, job;
        double v;
        return item(id, x, z, fl, v, job);
    }
    PLH_DEV uint64_t word(uint32_t id, int w) const {
        if (JOB && w == 0) {
            const uint32_t t = id >= n ? id - n : id;
            return j[t];
        }
        const int kw = w - kXOff;
        if (kw > 2 * AS) return 0ull;
        const bool nw = id >= n;
        const uint32_t t = nw ? id - n : id;
        if (kw < 2 * AS) {
            const bool isx = kw < AS;
            const int k = isx ? kw : kw - AS;
            uint64_t v = k < as ? (isx ? s.x : s.z)[(int64_t)k * s.stride + t] : 0ull;
            if (nw) {
                const uint64_t *px, *pz;
                uint32_t psf;
                double c, sn;
                if (rot_of(t, px, pz, psf, c, sn)) v ^= isx ? px[k] : pz[k];
            }
            return v;
        }
        uint64_t x[AS], z[AS];
        uint32_t fl, job;
        double v;
        item(id, x, z, fl, v, job);
        return (uint64_t)((mode & kFold) ? (fl & 1u) : fl) << 56;
    }
    PLH_DEV void key(uint32_t id, uint64_t* kw, uint32_t& fl) const {
        uint64_t x[AS], z[AS];
        uint32_t job;
        double v;
        item(id, x, z, fl, v, job);
        if (JOB) kw[0] = job;
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            kw[kXOff + k] = x[k];
            kw[kXOff + AS + k] = z[k];
        }
    }
    PLH_DEV uint32_t raw_flags(uint32_t id) const {
        uint64_t x[AS], z[AS];
        uint32_t fl, job;
        double v;
        item(id, x, z, fl, v, job);
        return fl;
    }
    PLH_DEV double base_value(uint32_t id) const {
        uint64_t x[AS], z[AS];
        uint32_t fl, job;
        double v;
        item(id, x, z, fl, v, job);
        return v;
    }
    PLH_DEV void key_value(uint32_t id, uint64_t* kw, uint32_t& fl, double& v) const {
        uint64_t x[AS], z[AS];
        uint32_t job;
        item(id, x, z, fl, v, job);
        if (JOB) kw[0] = job;
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            kw[kXOff + k] = x[k];
            kw[kXOff + AS + k] = z[k];
        }
    }
};

// A term packed as one record of 2*AS + 2 words: x[AS], z[AS], value,
// flags -- a pair's key and coefficient come from two records (two cache
// lines) instead of eight plane loads.
template <int AS>
struct PackedView {
    static constexpr int kRec = 2 * AS + 2;
    const uint64_t* r;
    PLH_DEV const uint64_t* rec(uint32_t i) const { return r + (size_t)i * kRec; }
};

// KIND 0: product.  hi = rhs index i, lo = lhs index j, term = lhs[j]*rhs[i],
//         value lhs.v[j]*rhs.v[i]  (PauliSum::multiply,
//         include/PauliSumOperations.h:255-306).
// KIND 1: commutator.  hi = a index i, lo = b index j, kept only when the
//         pair anticommutes, term = a[i]*b[j] (+1 phase for i[A,B]), value
//         (2*a.v[i])*b.v[j]  (commutator_impl, :315-346).
template <int AS, int KIND>
struct PairSource {
    static constexpr int kAS = AS;
    static constexpr int kXOff = 0;
    static constexpr int kKW = 2 * AS;
    PackedView<AS> L;  // left factor of the product (records: key + value)
    PackedView<AS> R;  // right factor
    SumView Lp, Rp;    // the same operands as planes (one key word per load)
    int as;
    uint32_t mode;
    uint32_t lo_bits;
    uint32_t n_lo, n_hi;
    FastDiv div_lo;  // dense t -> (t / n_lo, t % n_lo)
    uint32_t by_i;   // commutator: multiply by i

    PLH_DEV void split(uint32_t id, uint32_t& li, uint32_t& ri) const {
        const uint32_t hi = id >> lo_bits;
        const uint32_t lo = id & ((1u << lo_bits) - 1u);
        if (KIND == 0) { li = lo; ri = hi; } else { li = hi; ri = lo; }
    }
    PLH_DEV uint32_t dense_to_id(uint32_t t) const {
        const uint32_t hi = fdiv(t, div_lo);
        const uint32_t lo = t - hi * n_lo;
        return (hi << lo_bits) | lo;
    }
    PLH_DEV bool valid(uint32_t id) const {
        if (KIND == 0) return true;
        uint32_t li, ri;
        split(id, li, ri);
        const uint64_t* a = L.rec(li);
        const uint64_t* b = R.rec(ri);
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) acc ^= (a[k] & b[AS + k]) ^ (a[AS + k] & b[k]);
        return (popc(acc) & 1) != 0;
    }
    PLH_DEV uint32_t raw_flags_of(const uint64_t* a, const uint64_t* b) const {
        uint64_t c1 = 0, c2 = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) phase_word(a[k], a[AS + k], b[k], b[AS + k], c1, c2);
        uint32_t fl = phase_of((uint32_t)a[2 * AS + 1], (uint32_t)b[2 * AS + 1], c1, c2);
        if (KIND == 1) fl = (fl + by_i) & 3u;
        return fl;
    }
    PLH_DEV uint64_t word(uint32_t id, int w) const {
        uint32_t li, ri;
        split(id, li, ri);
        if (w < AS) return w < as ? (Lp.x[(int64_t)w * Lp.stride + li] ^ Rp.x[(int64_t)w * Rp.stride + ri]) : 0ull;
        if (w < 2 * AS) {
            const int k = w - AS;
            return k < as ? (Lp.z[(int64_t)k * Lp.stride + li] ^ Rp.z[(int64_t)k * Rp.stride + ri]) : 0ull;
        }
        if (w == 2 * AS) {
            const uint32_t fl = raw_flags_of(L.rec(li), R.rec(ri));
            return (uint64_t)((mode & kFold) ? (fl & 1u) : fl) << 56;
        }
        return 0ull;
    }
    PLH_DEV void key(uint32_t id, uint64_t* xz, uint32_t& fl) const {
        uint32_t li, ri;
        split(id, li, ri);
        const uint64_t* a = L.rec(li);
        const uint64_t* b = R.rec(ri);
#pragma unroll
        for (int w = 0; w < 2 * AS; ++w) xz[w] = a[w] ^ b[w];
        fl = raw_flags_of(a, b);  // raw flags; the engine masks the sign bit
    }
    PLH_DEV uint32_t raw_flags(uint32_t id) const {
        uint32_t li, ri;
        split(id, li, ri);
        return raw_flags_of(L.rec(li), R.rec(ri));
    }
    // Coefficient before the sign fold: lhs.v[j]*rhs.v[i] (product) or
    // (2*a.v[i])*b.v[j] (commutator), one IEEE multiply each.
    PLH_DEV double base_value(uint32_t id) const {
        uint32_t li, ri;
        split(id, li, ri);
        const double va = __longlong_as_double((long long)L.rec(li)[2 * AS]);
        const double vb = __longlong_as_double((long long)R.rec(ri)[2 * AS]);
        return (KIND == 0) ? va * vb : (2.0 * va) * vb;
    }
    PLH_DEV void key_value(uint32_t id, uint64_t* kw, uint32_t& fl, double& v) const {
        uint32_t li, ri;
        split(id, li, ri);
        const uint64_t* a = L.rec(li);
        const uint64_t* b = R.rec(ri);
#pragma unroll
        for (int k = 0; k < 2 * AS; ++k) kw[k] = a[k] ^ b[k];
        fl = raw_flags_of(a, b);
        const double va = __longlong_as_double((long long)a[2 * AS]);
        const double vb = __longlong_as_double((long long)b[2 * AS]);
        v = (KIND == 0) ? va * vb : (2.0 * va) * vb;
    }
};

// Many independent products canonicalize(lhs_k * rhs_k) in one collation.
// Ids are dense over the concatenated pair ranges: job k owns
// [off_k, off_k + n_k * m_k), and inside it id - off_k = i * n_k + j is the
// reference's order out[i*n + j] = lhs[j] * rhs[i].  The job index leads the
// key, so each job's terms are collated exactly as they would be alone.
// Operands are packed records (x[AS], z[AS], value, flags) concatenated over
// the jobs; chunk_job[c] is the job that owns id c << kChunkBits.
struct ProductJob {
    uint64_t off;    // first id of the job
    uint32_t n;      // lhs terms (the fast index j)
    uint32_t lbase;  // first lhs record of the job
    uint32_t rbase;  // first rhs record
    FastDiv div;     // divides by n
};

template <int AS>
struct MultiPairSource {
    static constexpr int kAS = AS;
    static constexpr int kXOff = 1;
    static constexpr int kKW = 2 * AS + 1;
    static constexpr int kRec = 2 * AS + 2;
    static constexpr int kChunkBits = 12;
    const uint64_t* L;   // lhs records of all jobs
    const uint64_t* R;   // rhs records of all jobs
    const ProductJob* jobs;
    const uint32_t* chunk_job;
    uint32_t njobs;
    uint32_t mode;

    PLH_DEV bool valid(uint32_t) const { return true; }
    PLH_DEV uint32_t dense_to_id(uint32_t t) const { return t; }
    // job of id: from the chunk table, then forward while the next job starts at or before id
    PLH_DEV uint32_t job_of(uint32_t id) const {
        uint32_t k = chunk_job[id >> kChunkBits];
        while (k + 1 < njobs && jobs[k + 1].off <= id) ++k;
        return k;
    }
    PLH_DEV void split(uint32_t id, uint32_t& k, const uint64_t*& a, const uint64_t*& b) const {
        k = job_of(id);
        const ProductJob J = jobs[k];
        const uint32_t local = (uint32_t)(id - J.off);
        const uint32_t i = fdiv(local, J.div);
        const uint32_t jj = local - i * J.n;
        a = L + (size_t)(J.lbase + jj) * kRec;
        b = R + (size_t)(J.rbase + i) * kRec;
    }
    PLH_DEV static uint32_t flags_of(const uint64_t* a, const uint64_t* b) {
        uint64_t c1 = 0, c2 = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) phase_word(a[k], a[AS + k], b[k], b[AS + k], c1, c2);
        return phase_of((uint32_t)a[2 * AS + 1], (uint32_t)b[2 * AS + 1], c1, c2);
    }
    PLH_DEV uint64_t word(uint32_t id, int w) const {
        uint32_t k;
        const uint64_t *a, *b;
        split(id, k, a, b);
        if (w == 0) return k;
        if (w <= 2 * AS) return a[w - 1] ^ b[w - 1];
        if (w == 2 * AS + 1) {
            const uint32_t fl = flags_of(a, b);
            return (uint64_t)((mode & kFold) ? (fl & 1u) : fl) << 56;
        }
        return 0ull;
    }
    PLH_DEV void key(uint32_t id, uint64_t* kw, uint32_t& fl) const {
        uint32_t k;
        const uint64_t *a, *b;
        split(id, k, a, b);
        kw[0] = k;
#pragma unroll
        for (int q = 0; q < 2 * AS; ++q) kw[1 + q] = a[q] ^ b[q];
        fl = flags_of(a, b);
    }
    PLH_DEV uint32_t raw_flags(uint32_t id) const {
        uint32_t k;
        const uint64_t *a, *b;
        split(id, k, a, b);
        return flags_of(a, b);
    }
    PLH_DEV double base_value(uint32_t id) const {
        uint32_t k;
        const uint64_t *a, *b;
        split(id, k, a, b);
        return __longlong_as_double((long long)a[2 * AS]) * __longlong_as_double((long long)b[2 * AS]);
    }
    PLH_DEV void key_value(uint32_t id, uint64_t* kw, uint32_t& fl, double& v) const {
        uint32_t k;
        const uint64_t *a, *b;
        split(id, k, a, b);
        kw[0] = k;
#pragma unroll
        for (int q = 0; q < 2 * AS; ++q) kw[1 + q] = a[q] ^ b[q];
        fl = flags_of(a, b);
        v = __longlong_as_double((long long)a[2 * AS]) * __longlong_as_double((long long)b[2 * AS]);
    }
};

// Window of DB <= 64 bits starting at global bit position pos (MSB-first);
// bits past the key's last word read 0.
template <int DB, class Src>
PLH_DEV uint64_t key_bits(const Src& src, uint32_t id, int pos) {
    static_assert(DB > 0 && DB <= 64, "window wider than a word");
    const int w = pos >> 6;
    const int o = pos & 63;
    uint64_t hi = src.word(id, w) << o;
    if (o + DB > 64) hi |= src.word(id, w + 1) >> (64 - o);
    return hi >> (64 - DB);
}
template <int DB, class Src>
PLH_DEV uint32_t key_window(const Src& src, uint32_t id, int pos) {
    static_assert(DB <= 32, "use key_bits for windows wider than 32 bits");
    return (uint32_t)key_bits<DB>(src, id, pos);
}

}  // namespace plh
