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

template <int AS>
struct TermSource {
    static constexpr int kWords = 2 * AS + 1;
    SumView s;
    int as;           // real words per plane (<= AS)
    uint32_t mode;    // kFold ...
    uint32_t n;       // items

    PLH_DEV bool valid(uint32_t) const { return true; }
    PLH_DEV uint32_t dense_to_id(uint32_t t) const { return t; }
    PLH_DEV uint32_t flag(uint32_t id) const {
        const uint32_t fl = s.f[id];
        return (mode & kFold) ? (fl & 1u) : fl;
    }
    PLH_DEV uint64_t word(uint32_t id, int w) const {
        if (w < AS) return w < as ? s.x[(int64_t)w * s.stride + id] : 0ull;
        if (w < 2 * AS) return (w - AS) < as ? s.z[(int64_t)(w - AS) * s.stride + id] : 0ull;
        if (w == 2 * AS) return (uint64_t)flag(id) << 56;
        return 0ull;
    }
    // Key words and the RAW flags byte (the engine masks the sign bit).
    PLH_DEV void key(uint32_t id, uint64_t* xz, uint32_t& fl) const {
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            xz[k] = k < as ? s.x[(int64_t)k * s.stride + id] : 0ull;
            xz[AS + k] = k < as ? s.z[(int64_t)k * s.stride + id] : 0ull;
        }
        fl = s.f[id];
    }
    PLH_DEV uint32_t raw_flags(uint32_t id) const { return s.f[id]; }
    // Coefficient before the sign fold.
    PLH_DEV double base_value(uint32_t id) const { return s.v[id]; }
};

// KIND 0: product.  hi = rhs index i, lo = lhs index j, term = lhs[j]*rhs[i],
//         value lhs.v[j]*rhs.v[i]  (PauliSum::multiply,
//         include/PauliSumOperations.h:255-306).
// KIND 1: commutator.  hi = a index i, lo = b index j, kept only when the
//         pair anticommutes, term = a[i]*b[j] (+1 phase for i[A,B]), value
//         (2*a.v[i])*b.v[j]  (commutator_impl, :315-346).
template <int AS, int KIND>
struct PairSource {
    static constexpr int kWords = 2 * AS + 1;
    SumView L;       // left factor of the product
    SumView R;       // right factor
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
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            if (k < as) {
                acc ^= (L.x[(int64_t)k * L.stride + li] & R.z[(int64_t)k * R.stride + ri]) ^
                       (L.z[(int64_t)k * L.stride + li] & R.x[(int64_t)k * R.stride + ri]);
            }
        }
        return (popc(acc) & 1) != 0;
    }
    PLH_DEV uint32_t raw_flags(uint32_t li, uint32_t ri) const {
        uint64_t c1 = 0, c2 = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            if (k < as) {
                phase_word(L.x[(int64_t)k * L.stride + li], L.z[(int64_t)k * L.stride + li],
                           R.x[(int64_t)k * R.stride + ri], R.z[(int64_t)k * R.stride + ri], c1, c2);
            }
        }
        uint32_t fl = phase_of(L.f[li], R.f[ri], c1, c2);
        if (KIND == 1) fl = (fl + by_i) & 3u;
        return fl;
    }
    PLH_DEV uint32_t flag(uint32_t id) const {
        uint32_t li, ri;
        split(id, li, ri);
        const uint32_t fl = raw_flags(li, ri);
        return (mode & kFold) ? (fl & 1u) : fl;
    }
    PLH_DEV uint64_t word(uint32_t id, int w) const {
        uint32_t li, ri;
        split(id, li, ri);
        if (w < AS)
            return w < as ? (L.x[(int64_t)w * L.stride + li] ^ R.x[(int64_t)w * R.stride + ri]) : 0ull;
        if (w < 2 * AS) {
            const int k = w - AS;
            return k < as ? (L.z[(int64_t)k * L.stride + li] ^ R.z[(int64_t)k * R.stride + ri]) : 0ull;
        }
        if (w == 2 * AS) {
            const uint32_t fl = raw_flags(li, ri);
            return (uint64_t)((mode & kFold) ? (fl & 1u) : fl) << 56;
        }
        return 0ull;
    }
    PLH_DEV void key(uint32_t id, uint64_t* xz, uint32_t& fl) const {
        uint32_t li, ri;
        split(id, li, ri);
        uint64_t c1 = 0, c2 = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            uint64_t ax = 0, az = 0, bx = 0, bz = 0;
            if (k < as) {
                ax = L.x[(int64_t)k * L.stride + li];
                az = L.z[(int64_t)k * L.stride + li];
                bx = R.x[(int64_t)k * R.stride + ri];
                bz = R.z[(int64_t)k * R.stride + ri];
            }
            xz[k] = ax ^ bx;
            xz[AS + k] = az ^ bz;
            phase_word(ax, az, bx, bz, c1, c2);
        }
        uint32_t f = phase_of(L.f[li], R.f[ri], c1, c2);
        if (KIND == 1) f = (f + by_i) & 3u;
        fl = f;  // raw flags; the engine masks the sign bit
    }
    PLH_DEV uint32_t raw_flags(uint32_t id) const {
        uint32_t li, ri;
        split(id, li, ri);
        return raw_flags(li, ri);
    }
    // Coefficient before the sign fold: lhs.v[j]*rhs.v[i] (product) or
    // (2*a.v[i])*b.v[j] (commutator), one IEEE multiply each.
    PLH_DEV double base_value(uint32_t id) const {
        uint32_t li, ri;
        split(id, li, ri);
        return (KIND == 0) ? L.v[li] * R.v[ri] : (2.0 * L.v[li]) * R.v[ri];
    }
};

// Window of DB bits starting at global bit position pos (MSB-first).
template <int DB, class Src>
PLH_DEV uint32_t key_window(const Src& src, uint32_t id, int pos) {
    const int w = pos >> 6;
    const int o = pos & 63;
    uint64_t hi = src.word(id, w) << o;
    if (o + DB > 64) hi |= src.word(id, w + 1) >> (64 - o);
    return (uint32_t)(hi >> (64 - DB));
}

}  // namespace plh
