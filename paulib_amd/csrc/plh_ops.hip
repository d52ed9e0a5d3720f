// plh_ops.hip -- element-wise and pair kernels of the PauliSum engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "plh_block.hpp"
#include "plh_engine.hpp"
#include "plh_sources.hpp"

namespace plh {

namespace {

constexpr int kOpsIpt = 16;  // items per thread in the ordered-compaction kernels
constexpr uint64_t kOpsTile = (uint64_t)kBlock * kOpsIpt;

int padded(int as) {
    if (as <= 1) return 1;
    if (as == 2) return 2;
    if (as <= 4) return 4;
    if (as <= 8) return 8;
    if (as <= 16) return 16;
    return -1;
}

// ------------------------------------------------------------ raw product
// out[i*n + j] = lhs[j] * rhs[i] (include/PauliSumOperations.h:255-306).
// One row i per block-row; the rhs term is wave-uniform, lhs reads and all
// writes are coalesced; planes use non-temporal stores (streamed output).
template <int AS>
__global__ __launch_bounds__(kBlock) void k_mul_raw(
    int as, const uint64_t* __restrict__ lx, const uint64_t* __restrict__ lz,
    const uint8_t* __restrict__ lf, const double* __restrict__ lv, int64_t ls,
    const uint64_t* __restrict__ rx, const uint64_t* __restrict__ rz,
    const uint8_t* __restrict__ rf, const double* __restrict__ rv, int64_t rs,
    uint64_t* __restrict__ ox, uint64_t* __restrict__ oz, uint8_t* __restrict__ of,
    double* __restrict__ ov, int64_t os, int64_t n, uint32_t jblocks, int reps) {
    const uint64_t i = blockIdx.x / jblocks;
    const uint64_t jb = blockIdx.x - i * jblocks;
    uint64_t bx[AS], bz[AS];
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        bx[k] = k < as ? rx[(int64_t)k * rs + i] : 0ull;
        bz[k] = k < as ? rz[(int64_t)k * rs + i] : 0ull;
    }
    const uint32_t bf = rf[i];
    const double bv = rv[i];
    const int64_t row = (int64_t)i * n;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r >= reps) break;
        const int64_t j = (int64_t)jb * (reps * kBlock) + r * kBlock + threadIdx.x;
        if (j >= n) break;
        uint64_t c1 = 0, c2 = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            if (k < as) {
                const uint64_t ax = lx[(int64_t)k * ls + j], az = lz[(int64_t)k * ls + j];
                __builtin_nontemporal_store(ax ^ bx[k], &ox[(int64_t)k * os + row + j]);
                __builtin_nontemporal_store(az ^ bz[k], &oz[(int64_t)k * os + row + j]);
                phase_word(ax, az, bx[k], bz[k], c1, c2);
            }
        }
        of[row + j] = (uint8_t)phase_of(lf[j], bf, c1, c2);
        __builtin_nontemporal_store(lv[j] * bv, &ov[row + j]);
    }
}

// ------------------------------------------------------- commutator (raw)
// Anticommuting pairs of a (i) and b (j) in (i, j) order, compacted with a
// counting pass and an ordered write pass (no atomics, deterministic).
template <int AS>
PLH_DEV bool anticommutes(int as, const uint64_t* ax, const uint64_t* az, int64_t as_,
                          uint64_t ia, const uint64_t* bx, const uint64_t* bz, int64_t bs,
                          uint64_t jb) {
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k)
        if (k < as)
            acc ^= (ax[(int64_t)k * as_ + ia] & bz[(int64_t)k * bs + jb]) ^
                   (az[(int64_t)k * as_ + ia] & bx[(int64_t)k * bs + jb]);
    return popc(acc) & 1;
}

template <int AS>
__global__ __launch_bounds__(kBlock) void k_comm_count(int as, SumView A, SumView B, uint64_t m,
                                                       uint64_t total, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t red[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    uint32_t c = 0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= total) break;
        const uint64_t i = t / m, j = t - i * m;
        c += anticommutes<AS>(as, A.x, A.z, A.stride, i, B.x, B.z, B.stride, j) ? 1u : 0u;
    }
    const uint32_t s = block_sum(c, red);
    if (threadIdx.x == 0) cnt[blockIdx.x] = s;
}

template <int AS>
__global__ __launch_bounds__(kBlock) void k_comm_write(int as, SumView A, SumView B, uint64_t m,
                                                       uint64_t total, uint32_t by_i,
                                                       const uint64_t* __restrict__ boff,
                                                       uint64_t* __restrict__ ox, uint64_t* __restrict__ oz,
                                                       uint8_t* __restrict__ of, double* __restrict__ ov,
                                                       int64_t os, uint64_t limit) {
    __shared__ uint32_t red[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    uint32_t mask = 0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= total) break;
        const uint64_t i = t / m, j = t - i * m;
        if (anticommutes<AS>(as, A.x, A.z, A.stride, i, B.x, B.z, B.stride, j)) mask |= 1u << k;
    }
    uint32_t tot;
    uint64_t o = boff[blockIdx.x] + block_exclusive_scan((uint32_t)__popc(mask), tot, red);
    for (int k = 0; k < kOpsIpt; ++k) {
        if (!(mask >> k & 1u)) continue;
        if (o >= limit) break;  // output slots from `limit` on are not the caller's
        const uint64_t t = base + k;
        const uint64_t i = t / m, j = t - i * m;
        uint64_t c1 = 0, c2 = 0;
        for (int w = 0; w < as; ++w) {
            const uint64_t ax = A.x[(int64_t)w * A.stride + i], az = A.z[(int64_t)w * A.stride + i];
            const uint64_t bx = B.x[(int64_t)w * B.stride + j], bz = B.z[(int64_t)w * B.stride + j];
            ox[(int64_t)w * os + o] = ax ^ bx;
            oz[(int64_t)w * os + o] = az ^ bz;
            phase_word(ax, az, bx, bz, c1, c2);
        }
        of[o] = (uint8_t)((phase_of(A.f[i], B.f[j], c1, c2) + by_i) & 3u);
        ov[o] = (2.0 * A.v[i]) * B.v[j];
        ++o;
    }
}

// ----------------------------------------------------------- Clifford
// H, S, CNOT exactly as include/PauliSumClifford.h:43-81, 212-239.
__global__ void k_clifford(int gate, uint64_t* __restrict__ x, uint64_t* __restrict__ z,
                           uint8_t* __restrict__ f, int64_t stride, int64_t n, int wc, uint64_t mc,
                           int wt, uint64_t mt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (gate == 0) {  // H
        uint64_t xw = x[(int64_t)wc * stride + i], zw = z[(int64_t)wc * stride + i];
        const uint32_t xc = (xw & mc) != 0, zc = (zw & mc) != 0;
        f[i] ^= (uint8_t)((xc & zc) << 1);
        xw ^= zw & mc;
        zw ^= xw & mc;
        xw ^= zw & mc;
        x[(int64_t)wc * stride + i] = xw;
        z[(int64_t)wc * stride + i] = zw;
    } else if (gate == 1) {  // S
        const uint64_t xw = x[(int64_t)wc * stride + i], zw = z[(int64_t)wc * stride + i];
        const uint32_t xc = (xw & mc) != 0, zc = (zw & mc) != 0;
        f[i] ^= (uint8_t)((xc & zc) << 1);
        z[(int64_t)wc * stride + i] = zw ^ (xw & mc);
    } else {  // CNOT(control c, target t)
        const uint64_t xc = (x[(int64_t)wc * stride + i] & mc) != 0;
        const uint64_t zc = (z[(int64_t)wc * stride + i] & mc) != 0;
        const uint64_t xt = (x[(int64_t)wt * stride + i] & mt) != 0;
        const uint64_t zt = (z[(int64_t)wt * stride + i] & mt) != 0;
        f[i] ^= (uint8_t)((xc & zt & (1ull ^ xt ^ zc)) << 1);
        x[(int64_t)wt * stride + i] ^= (0ull - xc) & mt;
        z[(int64_t)wc * stride + i] ^= (0ull - zt) & mc;
    }
}

// A run of Clifford ops in one pass: each thread holds its term (AS words
// per plane and the flags) in registers and applies the ops in order, H, S
// and CNOT as k_clifford, table gates as k_clifford_table.  Words are picked
// by compare so the term stays in registers.
template <int AS>
__global__ void k_clifford_run(uint64_t* __restrict__ x, uint64_t* __restrict__ z, uint8_t* __restrict__ f,
                               int64_t stride, int64_t n, int as, CliffordRun run) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t xw[AS], zw[AS];
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        xw[k] = k < as ? x[(int64_t)k * stride + i] : 0ull;
        zw[k] = k < as ? z[(int64_t)k * stride + i] : 0ull;
    }
    uint32_t fl = f[i];
    for (int o = 0; o < run.nops; ++o) {
        const uint32_t op = run.op[o];
        const uint32_t kind = op >> 28, q0 = (op >> 14) & 0x3FFFu, q1 = op & 0x3FFFu;
        const int wc = (int)(q0 >> 6), wt = (int)(q1 >> 6);
        const uint64_t mc = 1ull << (q0 & 63), mt = 1ull << (q1 & 63);
        uint64_t xcw = 0, zcw = 0, xtw = 0, ztw = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            if (k == wc) {
                xcw = xw[k];
                zcw = zw[k];
            }
            if (k == wt) {
                xtw = xw[k];
                ztw = zw[k];
            }
        }
        const uint64_t xc = (xcw & mc) != 0, zc = (zcw & mc) != 0;
        if (kind == 0) {  // H: X <-> Z, Y -> -Y
            fl ^= (uint32_t)((xc & zc) << 1);
            const uint64_t d = (xcw ^ zcw) & mc;
#pragma unroll
            for (int k = 0; k < AS; ++k)
                if (k == wc) {
                    xw[k] ^= d;
                    zw[k] ^= d;
                }
        } else if (kind == 1) {  // S
            fl ^= (uint32_t)((xc & zc) << 1);
#pragma unroll
            for (int k = 0; k < AS; ++k)
                if (k == wc) zw[k] = zcw ^ (xcw & mc);
        } else if (kind == 2) {  // CNOT(q0 -> q1)
            const uint64_t xt = (xtw & mt) != 0, zt = (ztw & mt) != 0;
            fl ^= (uint32_t)((xc & zt & (1ull ^ xt ^ zc)) << 1);
#pragma unroll
            for (int k = 0; k < AS; ++k) {
                if (k == wt) xw[k] ^= (0ull - xc) & mt;
                if (k == wc) zw[k] ^= (0ull - zt) & mc;
            }
        } else {  // a table gate on q0 (and q1 when kind 4)
            const bool two = kind == 4;
            uint32_t b = (uint32_t)xc | ((uint32_t)zc << 1);
            if (two) b |= ((uint32_t)((xtw & mt) != 0) << 2) | ((uint32_t)((ztw & mt) != 0) << 3);
            fl ^= ((run.sgn[o] >> b) & 1u) << 1;
            const uint32_t e = (uint32_t)(run.tbl[o] >> (4 * b)) & 15u;
#pragma unroll
            for (int k = 0; k < AS; ++k) {
                if (k == wc) {
                    xw[k] = (xw[k] & ~mc) | ((e & 1u) ? mc : 0ull);
                    zw[k] = (zw[k] & ~mc) | ((e & 2u) ? mc : 0ull);
                }
                if (two && k == wt) {
                    xw[k] = (xw[k] & ~mt) | ((e & 4u) ? mt : 0ull);
                    zw[k] = (zw[k] & ~mt) | ((e & 8u) ? mt : 0ull);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        if (k < as) {
            x[(int64_t)k * stride + i] = xw[k];
            z[(int64_t)k * stride + i] = zw[k];
        }
    }
    f[i] = (uint8_t)fl;
}

// ------------------------------------------------ table-driven Clifford
// Every gate of include/PauliSumClifford.h acts on a term only through the
// bits of its one or two qubits, b = xc | zc<<1 | xt<<2 | zt<<3.  A 16-entry
// table gives the new local bits (one nibble per entry, `tbl`) and whether
// the sign flips (bit b of `sgn`).  A wave reads 64 consecutive words of each
// touched plane; the control and target words are loaded once when they are
// the same word.  Gates that only flip signs (X, Y, Z) write only flags.
__global__ void k_clifford_table(uint64_t* __restrict__ x, uint64_t* __restrict__ z,
                                 uint8_t* __restrict__ f, int64_t stride, int64_t n, int wc,
                                 uint64_t mc, int wt, uint64_t mt, uint64_t tbl, uint32_t sgn,
                                 int two, int writes_bits) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t oc = (int64_t)wc * stride + i, ot = (int64_t)wt * stride + i;
    uint64_t xcw = x[oc], zcw = z[oc];
    uint32_t b = (uint32_t)((xcw & mc) != 0) | ((uint32_t)((zcw & mc) != 0) << 1);
    const bool same = two && wt == wc;
    uint64_t xtw = xcw, ztw = zcw;
    if (two) {
        if (!same) {
            xtw = x[ot];
            ztw = z[ot];
        }
        b |= ((uint32_t)((xtw & mt) != 0) << 2) | ((uint32_t)((ztw & mt) != 0) << 3);
    }
    if ((sgn >> b) & 1u) f[i] ^= (uint8_t)2;
    if (!writes_bits) return;
    const uint32_t e = (uint32_t)(tbl >> (4 * b)) & 15u;
    xcw = (xcw & ~mc) | ((e & 1u) ? mc : 0ull);
    zcw = (zcw & ~mc) | ((e & 2u) ? mc : 0ull);
    if (two) {
        if (same) {
            xcw = (xcw & ~mt) | ((e & 4u) ? mt : 0ull);
            zcw = (zcw & ~mt) | ((e & 8u) ? mt : 0ull);
        } else {
            x[ot] = (xtw & ~mt) | ((e & 4u) ? mt : 0ull);
            z[ot] = (ztw & ~mt) | ((e & 8u) ? mt : 0ull);
        }
    }
    x[oc] = xcw;
    z[oc] = zcw;
}

// elementwise_multiply: term i <- term_i * rhs_i with the pair product's
// phase (include/PauliSumOperations.h:220-250).  Each thread reads both
// terms' words before writing its own, so rhs may be the sum itself.
template <int AS>
__global__ void k_elementwise(int as, uint64_t* x, uint64_t* z, uint8_t* f, double* v, int64_t stride,
                              const uint64_t* bx, const uint64_t* bz, const uint8_t* bf, const double* bv,
                              int64_t bstride, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t c1 = 0, c2 = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        if (k < as) {
            const uint64_t ax = x[(int64_t)k * stride + i], az = z[(int64_t)k * stride + i];
            const uint64_t rx = bx[(int64_t)k * bstride + i], rz = bz[(int64_t)k * bstride + i];
            phase_word(ax, az, rx, rz, c1, c2);
            x[(int64_t)k * stride + i] = ax ^ rx;
            z[(int64_t)k * stride + i] = az ^ rz;
        }
    }
    f[i] = (uint8_t)phase_of(f[i], bf[i], c1, c2);
    v[i] = v[i] * bv[i];
}

// -------------------------------------------------------------- rotation
struct Rot {
    uint64_t px[16];
    uint64_t pz[16];
    uint32_t psf;  // flags of ps = P with the i bit flipped
    double c, s;
};

template <int AS>
PLH_DEV bool term_anticommutes(int as, const uint64_t* x, const uint64_t* z, int64_t stride,
                               uint64_t t, const Rot& r) {
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k)
        if (k < as) acc ^= (x[(int64_t)k * stride + t] & r.pz[k]) ^ (z[(int64_t)k * stride + t] & r.px[k]);
    return popc(acc) & 1;
}

// In-place term * ps (PauliSum::multiply(i, ps), :185-209).
template <int AS>
PLH_DEV void term_times_ps(int as, uint64_t* x, uint64_t* z, uint8_t* f, int64_t stride, uint64_t t,
                           const Rot& r) {
    uint64_t c1 = 0, c2 = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        if (k < as) {
            const uint64_t ax = x[(int64_t)k * stride + t], az = z[(int64_t)k * stride + t];
            x[(int64_t)k * stride + t] = ax ^ r.px[k];
            z[(int64_t)k * stride + t] = az ^ r.pz[k];
            phase_word(ax, az, r.px[k], r.pz[k], c1, c2);
        }
    }
    f[t] = (uint8_t)phase_of(f[t], r.psf, c1, c2);
}

// Fast paths (:121-147): kind 1 flips the sign, 2 flips and multiplies,
// 3 multiplies; only terms that anticommute with P.
template <int AS>
__global__ void k_rot_inplace(int as, int kind, uint64_t* __restrict__ x, uint64_t* __restrict__ z,
                              uint8_t* __restrict__ f, int64_t stride, int64_t n, Rot r) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (!term_anticommutes<AS>(as, x, z, stride, t, r)) return;
    if (kind == 1 || kind == 2) f[t] ^= 2u;
    if (kind >= 2) term_times_ps<AS>(as, x, z, f, stride, t, r);
}

template <int AS>
__global__ __launch_bounds__(kBlock) void k_rot_count(int as, const uint64_t* __restrict__ x,
                                                      const uint64_t* __restrict__ z, int64_t stride,
                                                      int64_t n, Rot r, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t red[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    uint32_t c = 0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= (uint64_t)n) break;
        c += term_anticommutes<AS>(as, x, z, stride, t, r) ? 1u : 0u;
    }
    const uint32_t s = block_sum(c, red);
    if (threadIdx.x == 0) cnt[blockIdx.x] = s;
}

// General path (:148-180): anticommuting term j is scaled by cos in place
// and term_j * ps with value (-sin) * v_j is appended at n + rank(j).
template <int AS>
__global__ __launch_bounds__(kBlock) void k_rot_general(int as, uint64_t* __restrict__ x,
                                                        uint64_t* __restrict__ z, uint8_t* __restrict__ f,
                                                        double* __restrict__ v, int64_t stride, int64_t n,
                                                        Rot r, const uint64_t* __restrict__ boff) {
    __shared__ uint32_t red[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    uint32_t mask = 0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= (uint64_t)n) break;
        if (term_anticommutes<AS>(as, x, z, stride, t, r)) mask |= 1u << k;
    }
    uint32_t tot;
    uint64_t o = (uint64_t)n + boff[blockIdx.x] + block_exclusive_scan((uint32_t)__popc(mask), tot, red);
    const double neg_s = -r.s;
    for (int k = 0; k < kOpsIpt; ++k) {
        if (!(mask >> k & 1u)) continue;
        const uint64_t t = base + k;
        const double orig = v[t];
        v[t] = orig * r.c;
        uint64_t c1 = 0, c2 = 0;
#pragma unroll
        for (int w = 0; w < AS; ++w) {
            if (w < as) {
                const uint64_t ax = x[(int64_t)w * stride + t], az = z[(int64_t)w * stride + t];
                x[(int64_t)w * stride + o] = ax ^ r.px[w];
                z[(int64_t)w * stride + o] = az ^ r.pz[w];
                phase_word(ax, az, r.px[w], r.pz[w], c1, c2);
            }
        }
        f[o] = (uint8_t)phase_of(f[t], r.psf, c1, c2);
        v[o] = neg_s * orig;
        ++o;
    }
}

// ---------------------------------------------------------------- filters
template <int AS>
PLH_DEV bool filter_keep(int as, const uint64_t* x, const uint64_t* z, const double* v, int64_t stride,
                         uint64_t t, uint32_t mode, int64_t max_weight, double eps) {
    if (mode & kMaxWeight) {
        int64_t w = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k)
            if (k < as) w += popc(x[(int64_t)k * stride + t] | z[(int64_t)k * stride + t]);
        if (w > max_weight) return false;
    }
    if ((mode & kCutoff) && !(fabs(v[t]) >= eps)) return false;
    return true;
}

template <int AS>
__global__ __launch_bounds__(kBlock) void k_filter_count(int as, SumView s, int64_t n, uint32_t mode,
                                                         int64_t max_weight, double eps,
                                                         uint32_t* __restrict__ cnt) {
    __shared__ uint32_t red[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    uint32_t c = 0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= (uint64_t)n) break;
        c += filter_keep<AS>(as, s.x, s.z, s.v, s.stride, t, mode, max_weight, eps) ? 1u : 0u;
    }
    const uint32_t tot = block_sum(c, red);
    if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

template <int AS>
__global__ __launch_bounds__(kBlock) void k_filter_write(int as, SumView s, int64_t n, uint32_t mode,
                                                         int64_t max_weight, double eps,
                                                         const uint64_t* __restrict__ boff,
                                                         uint64_t* __restrict__ ox, uint64_t* __restrict__ oz,
                                                         uint8_t* __restrict__ of, double* __restrict__ ov,
                                                         int64_t os, const uint32_t* __restrict__ sj,
                                                         uint32_t* __restrict__ oj) {
    __shared__ uint32_t red[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    uint32_t mask = 0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= (uint64_t)n) break;
        if (filter_keep<AS>(as, s.x, s.z, s.v, s.stride, t, mode, max_weight, eps)) mask |= 1u << k;
    }
    uint32_t tot;
    uint64_t o = boff[blockIdx.x] + block_exclusive_scan((uint32_t)__popc(mask), tot, red);
    for (int k = 0; k < kOpsIpt; ++k) {
        if (!(mask >> k & 1u)) continue;
        const uint64_t t = base + k;
        for (int w = 0; w < as; ++w) {
            ox[(int64_t)w * os + o] = s.x[(int64_t)w * s.stride + t];
            oz[(int64_t)w * os + o] = s.z[(int64_t)w * s.stride + t];
        }
        of[o] = s.f[t];
        ov[o] = s.v[t];
        if (oj) oj[o] = sj[t];
        ++o;
    }
}

// ------------------------------------------------------- expectation / OR
// Deterministic two-level reduction: each block sums its contiguous range
// with a fixed tree; block partials are summed in block order.
__global__ __launch_bounds__(kBlock) void k_zse_partial(int as, SumView s, int64_t n,
                                                        double* __restrict__ part) {
    __shared__ double red[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    double acc = 0.0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= (uint64_t)n) break;
        bool has_xy = false;
        for (int w = 0; w < as; ++w) has_xy |= s.x[(int64_t)w * s.stride + t] != 0;
        const uint32_t fl = s.f[t];
        if (has_xy || (fl & 1u)) continue;
        acc += (fl & 2u) ? -s.v[t] : s.v[t];
    }
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_zse_final(const double* __restrict__ part, uint64_t nb,
                                                      double* __restrict__ out) {
    __shared__ double red[kWavesPerBlock];
    double acc = 0.0;
    // Each thread owns a contiguous slice so the order is fixed.
    const uint64_t per = (nb + kBlock - 1) / kBlock;
    for (uint64_t i = threadIdx.x * per; i < min(nb, (threadIdx.x + 1) * per); ++i) acc += part[i];
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) *out = tot;
}

inline unsigned nblk(uint64_t n, uint64_t per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ====================================================================== API
int op_mul_raw(int as, const uint64_t* lx, const uint64_t* lz, const uint8_t* lf, const double* lv,
               int64_t lstride, const uint64_t* rx, const uint64_t* rz, const uint8_t* rf,
               const double* rv, int64_t rstride, uint64_t* ox, uint64_t* oz, uint8_t* of, double* ov,
               int64_t ostride, int64_t n, int64_t m, hipStream_t st) {
    if (n == 0 || m == 0) return PLH_OK;
    // 4 x 256 lhs terms per workgroup; 1 x 256 when that would leave the
    // chip underfilled (fewer than 8192 workgroups, e.g. 1000 x 1000)
    const int reps = ((uint64_t)((n + 4 * kBlock - 1) / (4 * kBlock)) * (uint64_t)m >= 8192) ? 4 : 1;
    const uint32_t jblocks = (uint32_t)((n + reps * kBlock - 1) / (reps * kBlock));
    const uint64_t grid = (uint64_t)jblocks * (uint64_t)m;
    if (grid > 0x7FFFFFFFull) {
        set_error("raw product grid too large");
        return PLH_ERR_SIZE;
    }
#define PLH_MUL(ASV)                                                                           \
    case ASV:                                                                                  \
        k_mul_raw<ASV><<<dim3((unsigned)grid), kBlock, 0, st>>>(as, lx, lz, lf, lv, lstride, rx, rz, \
                                                               rf, rv, rstride, ox, oz, of, ov,     \
                                                               ostride, n, jblocks, reps);         \
        break;
    switch (padded(as)) {
        PLH_MUL(1) PLH_MUL(2) PLH_MUL(4) PLH_MUL(8) PLH_MUL(16)
        default: set_error("unsupported array_size"); return PLH_ERR_AS;
    }
#undef PLH_MUL
    PLH_HIP(hipGetLastError());
    return PLH_OK;
}

int op_commutator_raw(plh_ctx* ctx, const plh_sum* a, const plh_sum* b, int by_i, plh_sum* out) {
    hipStream_t st = ctx->stream;
    const uint64_t n = (uint64_t)a->n, m = (uint64_t)b->n;
    const uint64_t total = n * m;
    out->as = a->as;
    out->nq = a->nq > b->nq ? a->nq : b->nq;
    out->n = 0;
    if (total == 0) return PLH_OK;
    const uint64_t nb = (total + kOpsTile - 1) / kOpsTile;
    void* p;
    PLH_TRY(ctx->ws.get(kWsScan0, nb * 4 + 64, &p));
    uint32_t* cnt = (uint32_t*)p;
    PLH_TRY(ctx->ws.get(kWsScan1, nb * 8 + 64, &p));
    uint64_t* off = (uint64_t*)p;
    PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
    uint64_t* d_total = (uint64_t*)p + 40;
    SumView A{a->x, a->z, a->f, a->v, a->cap}, B{b->x, b->z, b->f, b->v, b->cap};
#define PLH_CC(ASV) \
    case ASV: k_comm_count<ASV><<<dim3((unsigned)nb), kBlock, 0, st>>>(a->as, A, B, m, total, cnt); break;
    switch (padded(a->as)) {
        PLH_CC(1) PLH_CC(2) PLH_CC(4) PLH_CC(8) PLH_CC(16)
        default: set_error("unsupported array_size"); return PLH_ERR_AS;
    }
#undef PLH_CC
    PLH_TRY(scan_u32_to_u64(ctx, cnt, off, nb, d_total));
    uint64_t k = 0;
    PLH_TRY(read_u64(ctx, d_total, &k));
    PLH_TRY(sum_reserve(out, (int64_t)k, 0));
    out->n = (int64_t)k;
    if (k == 0) return PLH_OK;
#define PLH_CW(ASV)                                                                                   \
    case ASV:                                                                                         \
        k_comm_write<ASV><<<dim3((unsigned)nb), kBlock, 0, st>>>(a->as, A, B, m, total, (uint32_t)by_i, \
                                                                off, out->x, out->z, out->f, out->v,   \
                                                                out->cap, k);                         \
        break;
    switch (padded(a->as)) {
        PLH_CW(1) PLH_CW(2) PLH_CW(4) PLH_CW(8) PLH_CW(16)
        default: return PLH_ERR_AS;
    }
#undef PLH_CW
    PLH_HIP(hipGetLastError());
    out->canonical = false;
    out->sorted = false;
    return PLH_OK;
}

int op_clifford(plh_ctx* ctx, plh_sum* s, int gate, int64_t q0, int64_t q1) {
    if (q0 < 0 || q0 >= 64 * (int64_t)s->as || (gate == 2 && (q1 < 0 || q1 >= 64 * (int64_t)s->as))) {
        set_error("qubit index outside the sum's words");
        return PLH_ERR_ARG;
    }
    if (s->n == 0) return PLH_OK;
    const int wc = (int)(q0 / 64), wt = gate == 2 ? (int)(q1 / 64) : 0;
    const uint64_t mc = 1ull << (q0 % 64), mt = gate == 2 ? 1ull << (q1 % 64) : 0ull;
    k_clifford<<<nblk((uint64_t)s->n, kBlock), kBlock, 0, ctx->stream>>>(gate, s->x, s->z, s->f, s->cap,
                                                                          s->n, wc, mc, wt, mt);
    PLH_HIP(hipGetLastError());
    s->sorted = false;  // keys stay distinct (canonical holds) but their order does not
    return PLH_OK;
}

// Clifford tables.  A local Pauli on (c, t) is i^k * P(b), b = xc | zc<<1 |
// xt<<2 | zt<<3, with bits (1,1) on a qubit meaning Y (the flags convention
// of include/PauliStringCore.h:34-35).  A gate is given by its images of the
// generators Xc, Zc (and Xt, Zt); conjugation is multiplicative, so the image
// of i^(xc&zc + xt&zt) Xc^xc Zc^zc Xt^xt Zt^zt is the ordered product of the
// images, and its phase is 0 or 2 (the sign flip).
namespace {
struct Loc {
    uint32_t b, k;
};

// i^k of sigma_a sigma_b on one qubit, codes x | z<<1 (1 X, 2 Z, 3 Y):
// XY = iZ, YZ = iX, ZX = iY, and the reversed products take -i.
uint32_t site_phase(uint32_t a, uint32_t b) {
    if (a == 0 || b == 0 || a == b) return 0;
    const bool cyclic = (a == 1 && b == 3) || (a == 3 && b == 2) || (a == 2 && b == 1);
    return cyclic ? 1u : 3u;
}

Loc loc_mul(Loc a, Loc b) {
    uint32_t k = a.k + b.k;
    for (int q = 0; q < 2; ++q) k += site_phase((a.b >> (2 * q)) & 3u, (b.b >> (2 * q)) & 3u);
    return Loc{a.b ^ b.b, k & 3u};
}

// "-ZY": optional sign, then the Pauli on c and (for two-qubit images) on t.
Loc loc_parse(const char* s) {
    Loc r{0, 0};
    if (*s == '-') {
        r.k = 2;
        ++s;
    }
    for (int q = 0; s[q]; ++q) {
        const char ch = s[q];
        const uint32_t code = ch == 'X' ? 1u : ch == 'Z' ? 2u : ch == 'Y' ? 3u : 0u;
        r.b |= code << (2 * q);
    }
    return r;
}

struct CliffTable {
    uint64_t tbl;
    uint32_t sgn;
};

CliffTable table_from_images(const char* const img[4]) {
    Loc g[4];
    for (int j = 0; j < 4; ++j) g[j] = loc_parse(img[j]);
    CliffTable t{0, 0};
    for (uint32_t b = 0; b < 16; ++b) {
        Loc acc{0, ((b & 1u) & (b >> 1 & 1u)) + ((b >> 2 & 1u) & (b >> 3 & 1u))};
        for (int j = 0; j < 4; ++j)
            if (b >> j & 1u) acc = loc_mul(acc, g[j]);
        t.tbl |= (uint64_t)acc.b << (4 * b);
        t.sgn |= (acc.k >> 1 & 1u) << b;
    }
    return t;
}

// Applies a, then b.
CliffTable table_then(CliffTable a, CliffTable b) {
    CliffTable t{0, 0};
    for (uint32_t i = 0; i < 16; ++i) {
        const uint32_t m = (uint32_t)(a.tbl >> (4 * i)) & 15u;
        t.tbl |= (uint64_t)((b.tbl >> (4 * m)) & 15u) << (4 * i);
        t.sgn |= (((a.sgn >> i) ^ (b.sgn >> m)) & 1u) << i;
    }
    return t;
}

// Images of Xc, Zc, Xt, Zt (qubit c first) for the gate ids of
// plh_apply_clifford, read off the per-term updates of
// include/PauliSumClifford.h:44-490.
const char* const kImages[20][4] = {
    {"X", "Z", "IX", "IZ"},     // I
    {"X", "-Z", "IX", "IZ"},    // X
    {"-X", "-Z", "IX", "IZ"},   // Y
    {"-X", "Z", "IX", "IZ"},    // Z
    {"Z", "X", "IX", "IZ"},     // H
    {"Y", "Z", "IX", "IZ"},     // S
    {"-X", "Y", "IX", "IZ"},    // H_YZ
    {"Y", "-Z", "IX", "IZ"},    // H_XY
    {"X", "-Y", "IX", "IZ"},    // sqrt_X
    {"-Z", "X", "IX", "IZ"},    // sqrt_Y
    {"XX", "ZI", "IX", "ZZ"},   // CNOT
    {"XZ", "ZI", "ZX", "IZ"},   // CZ
    {"XY", "ZI", "ZX", "ZZ"},   // CY
    {"XI", "ZX", "IX", "XZ"},   // XC_X
    {"XI", "ZZ", "XX", "IZ"},   // XC_Z
    {"XI", "ZY", "XX", "XZ"},   // XC_Y
    {"XX", "ZX", "IX", "YZ"},   // YC_X
    {"XZ", "ZZ", "YX", "IZ"},   // YC_Z
    {"XY", "ZY", "YX", "YZ"},   // YC_Y
    {"IX", "IZ", "XI", "ZI"},   // SWAP
};

// gate 20, iSWAP: SWAP, then CZ, then S on c, then S on t
// (include/PauliSumClifford.h:493-498), composed into one table.
CliffTable clifford_table(int gate) {
    if (gate < 20) return table_from_images(kImages[gate]);
    static const char* const s_t[4] = {"X", "Z", "IY", "IZ"};
    CliffTable t = table_then(table_from_images(kImages[19]), table_from_images(kImages[11]));
    t = table_then(t, table_from_images(kImages[5]));
    return table_then(t, table_from_images(s_t));
}
}  // namespace

int clifford_table_of(int gate, uint64_t* tbl, uint32_t* sgn) {
    if (gate < 0 || gate >= kNumGates) return PLH_ERR_ARG;
    const CliffTable t = clifford_table(gate);
    *tbl = t.tbl;
    *sgn = t.sgn;
    return PLH_OK;
}

int op_clifford_table(plh_ctx* ctx, plh_sum* s, int gate, int64_t q0, int64_t q1) {
    if (gate < 0 || gate >= kNumGates) {
        set_error("unknown Clifford gate id");
        return PLH_ERR_ARG;
    }
    const bool two = gate >= kFirstTwoQubitGate;
    const int64_t nqb = 64 * (int64_t)s->as;
    if (q0 < 0 || q0 >= nqb || (two && (q1 < 0 || q1 >= nqb))) {
        set_error("qubit index outside the sum's words");
        return PLH_ERR_ARG;
    }
    if (two && q0 == q1) {
        set_error("a two-qubit gate needs two different qubits");
        return PLH_ERR_ARG;
    }
    if (gate == 0 || s->n == 0) return PLH_OK;  // apply_I is a no-op
    const CliffTable t = clifford_table(gate);
    const bool sign_only = gate >= 1 && gate <= 3;
    const int wc = (int)(q0 / 64), wt = two ? (int)(q1 / 64) : wc;
    const uint64_t mc = 1ull << (q0 % 64), mt = two ? 1ull << (q1 % 64) : 0ull;
    // algorithmic bytes: the touched words read (and written unless the gate
    // only flips signs), the flags read and written
    const double words = (two && wt != wc) ? 4.0 : 2.0;
    ProfScope ps(ctx, "clifford_table", (double)s->n * (words * 8.0 * (sign_only ? 1.0 : 2.0) + 2.0));
    k_clifford_table<<<nblk((uint64_t)s->n, kBlock), kBlock, 0, ctx->stream>>>(
        s->x, s->z, s->f, s->cap, s->n, wc, mc, wt, mt, t.tbl, t.sgn, two ? 1 : 0, sign_only ? 0 : 1);
    PLH_HIP(hipGetLastError());
    s->canonical = false;
    s->sorted = false;
    return PLH_OK;
}

int op_clifford_run(plh_ctx* ctx, plh_sum* s, const CliffordRun& run, bool any_table) {
    if (run.nops < 1 || run.nops > kCliffordRunMax) return PLH_ERR_ARG;
    if (s->n > 0) {
        const unsigned g = nblk((uint64_t)s->n, kBlock);
        switch (s->as <= 1 ? 1 : s->as == 2 ? 2 : s->as <= 4 ? 4 : s->as <= 8 ? 8 : 16) {
            case 1: k_clifford_run<1><<<g, kBlock, 0, ctx->stream>>>(s->x, s->z, s->f, s->cap, s->n, s->as, run); break;
            case 2: k_clifford_run<2><<<g, kBlock, 0, ctx->stream>>>(s->x, s->z, s->f, s->cap, s->n, s->as, run); break;
            case 4: k_clifford_run<4><<<g, kBlock, 0, ctx->stream>>>(s->x, s->z, s->f, s->cap, s->n, s->as, run); break;
            case 8: k_clifford_run<8><<<g, kBlock, 0, ctx->stream>>>(s->x, s->z, s->f, s->cap, s->n, s->as, run); break;
            default:
                k_clifford_run<16><<<g, kBlock, 0, ctx->stream>>>(s->x, s->z, s->f, s->cap, s->n, s->as, run);
                break;
        }
        PLH_HIP(hipGetLastError());
    }
    if (any_table) s->canonical = false;
    s->sorted = false;  // keys stay distinct but their order does not
    return PLH_OK;
}

int op_elementwise(plh_ctx* ctx, plh_sum* s, const plh_sum* r) {
    if (r->as != s->as || r->n != s->n || r->has_job || s->has_job) {
        set_error("elementwise_multiply needs two sums of the same length and word count");
        return PLH_ERR_ARG;
    }
    if (s->n == 0) return PLH_OK;
    const int pa = padded(s->as);
    if (pa < 0) return PLH_ERR_AS;
    hipStream_t st = ctx->stream;
    // both terms read (words, flags, value), the left one written
    ProfScope ps(ctx, "elementwise", (double)s->n * 3.0 * (16.0 * s->as + 9.0));
#define PLH_EW(ASV)                                                                                       \
    case ASV:                                                                                             \
        k_elementwise<ASV><<<nblk((uint64_t)s->n, kBlock), kBlock, 0, st>>>(                              \
            s->as, s->x, s->z, s->f, s->v, s->cap, r->x, r->z, r->f, r->v, r->cap, s->n);                 \
        break;
    switch (pa) { PLH_EW(1) PLH_EW(2) PLH_EW(4) PLH_EW(8) PLH_EW(16) }
#undef PLH_EW
    PLH_HIP(hipGetLastError());
    s->canonical = false;
    s->sorted = false;
    return PLH_OK;
}

static bool close_one(double v) { return std::fabs(v - 1.0) <= 1e-14; }  // isCloseOne, :19-21

int rotation_kind(double cos_t, double sin_t) {
    if (close_one(cos_t)) return 0;
    if (close_one(-cos_t)) return 1;
    if (close_one(sin_t)) return 2;
    if (close_one(-sin_t)) return 3;
    return 4;
}

int op_rotation(plh_ctx* ctx, plh_sum* s, const uint64_t* px, const uint64_t* pz, uint8_t pflags,
                double cos_t, double sin_t, int* appended_general) {
    hipStream_t st = ctx->stream;
    *appended_general = 0;
    Rot r{};
    for (int k = 0; k < s->as; ++k) {
        r.px[k] = px[k];
        r.pz[k] = pz[k];
    }
    r.psf = (uint32_t)(pflags ^ 1u);
    r.c = cos_t;
    r.s = sin_t;
    const int kind = rotation_kind(cos_t, sin_t);
    if (kind == 0) return PLH_OK;
    if (s->n == 0) return PLH_OK;
    const int pa = padded(s->as);
    if (pa < 0) return PLH_ERR_AS;
    if (kind != 4) {
#define PLH_RI(ASV)                                                                                 \
    case ASV:                                                                                       \
        k_rot_inplace<ASV><<<nblk((uint64_t)s->n, kBlock), kBlock, 0, st>>>(s->as, kind, s->x, s->z, \
                                                                            s->f, s->cap, s->n, r); \
        break;
        switch (pa) { PLH_RI(1) PLH_RI(2) PLH_RI(4) PLH_RI(8) PLH_RI(16) }
#undef PLH_RI
        PLH_HIP(hipGetLastError());
        s->canonical = false;
        s->sorted = false;
        return PLH_OK;
    }
    const uint64_t nb = nblk((uint64_t)s->n, kOpsTile);
    void* p;
    PLH_TRY(ctx->ws.get(kWsScan0, nb * 4 + 64, &p));
    uint32_t* cnt = (uint32_t*)p;
    PLH_TRY(ctx->ws.get(kWsScan1, nb * 8 + 64, &p));
    uint64_t* off = (uint64_t*)p;
    PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
    uint64_t* d_total = (uint64_t*)p + 40;
#define PLH_RC(ASV)                                                                                      \
    case ASV:                                                                                            \
        k_rot_count<ASV><<<dim3((unsigned)nb), kBlock, 0, st>>>(s->as, s->x, s->z, s->cap, s->n, r, cnt); \
        break;
    switch (pa) { PLH_RC(1) PLH_RC(2) PLH_RC(4) PLH_RC(8) PLH_RC(16) }
#undef PLH_RC
    PLH_TRY(scan_u32_to_u64(ctx, cnt, off, nb, d_total));
    uint64_t k = 0;
    PLH_TRY(read_u64(ctx, d_total, &k));
    if (k == 0) return PLH_OK;
    PLH_TRY(sum_reserve(s, s->n + (int64_t)k, s->n));
#define PLH_RG(ASV)                                                                                    \
    case ASV:                                                                                          \
        k_rot_general<ASV><<<dim3((unsigned)nb), kBlock, 0, st>>>(s->as, s->x, s->z, s->f, s->v, s->cap, \
                                                                 s->n, r, off);                        \
        break;
    switch (pa) { PLH_RG(1) PLH_RG(2) PLH_RG(4) PLH_RG(8) PLH_RG(16) }
#undef PLH_RG
    PLH_HIP(hipGetLastError());
    s->n += (int64_t)k;
    s->canonical = false;
    s->sorted = false;
    *appended_general = 1;
    return PLH_OK;
}

int op_filter(plh_ctx* ctx, plh_sum* s, uint32_t mode, int64_t max_weight, double eps) {
    hipStream_t st = ctx->stream;
    if (s->n == 0 || !(mode & (kMaxWeight | kCutoff))) return PLH_OK;
    const uint64_t nb = nblk((uint64_t)s->n, kOpsTile);
    void* p;
    PLH_TRY(ctx->ws.get(kWsScan0, nb * 4 + 64, &p));
    uint32_t* cnt = (uint32_t*)p;
    PLH_TRY(ctx->ws.get(kWsScan1, nb * 8 + 64, &p));
    uint64_t* off = (uint64_t*)p;
    PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
    uint64_t* d_total = (uint64_t*)p + 40;
    SumView v{s->x, s->z, s->f, s->v, s->cap};
    const int pa = padded(s->as);
#define PLH_FC(ASV)                                                                                     \
    case ASV:                                                                                           \
        k_filter_count<ASV><<<dim3((unsigned)nb), kBlock, 0, st>>>(s->as, v, s->n, mode, max_weight, eps, \
                                                                  cnt);                                 \
        break;
    switch (pa) {
        PLH_FC(1) PLH_FC(2) PLH_FC(4) PLH_FC(8) PLH_FC(16)
        default: return PLH_ERR_AS;
    }
#undef PLH_FC
    PLH_TRY(scan_u32_to_u64(ctx, cnt, off, nb, d_total));
    uint64_t k = 0;
    PLH_TRY(read_u64(ctx, d_total, &k));
    if (k == (uint64_t)s->n) return PLH_OK;
    plh_sum tmp;
    tmp.ctx = ctx;
    tmp.as = s->as;
    tmp.nq = s->nq;
    tmp.has_job = s->has_job;
    PLH_TRY(sum_reserve(&tmp, (int64_t)k, 0));
    if (k) {
#define PLH_FW(ASV)                                                                                     \
    case ASV:                                                                                           \
        k_filter_write<ASV><<<dim3((unsigned)nb), kBlock, 0, st>>>(s->as, v, s->n, mode, max_weight, eps, \
                                                                  off, tmp.x, tmp.z, tmp.f, tmp.v,      \
                                                                  tmp.cap, s->j, tmp.j);                \
        break;
        switch (pa) { PLH_FW(1) PLH_FW(2) PLH_FW(4) PLH_FW(8) PLH_FW(16) }
#undef PLH_FW
        PLH_HIP(hipGetLastError());
    }
    const bool canon = s->canonical, srt = s->sorted;
    sum_take_storage(s, &tmp);
    s->n = (int64_t)k;
    s->canonical = canon;  // a filter of a canonical sum stays canonical
    s->sorted = srt;
    return PLH_OK;
}

// ------------------------------------------------------------------ trim
// PauliSum::trim(max_terms) (include/PauliSumUtils.h:330-343) keeps the
// max_terms terms of largest |c| and compacts them in term order.  The
// magnitude's IEEE bits order like the magnitude, so the k-th largest is
// found by an 8-round radix select on those 63 bits (device-side, no
// read-back); a term is kept when its key is above that threshold, or equal
// to it and among the first k_rem such terms by index.
namespace {

PLH_DEV uint64_t mag_key(double v) { return (uint64_t)__double_as_longlong(v) & 0x7FFFFFFFFFFFFFFFull; }

// sel: [0] key prefix found so far, [1] its mask, [2] terms still to take among keys with that prefix
__global__ void k_trim_init(uint64_t* __restrict__ sel, uint32_t* __restrict__ ghist, uint64_t k) {
    if (threadIdx.x < 3) sel[threadIdx.x] = threadIdx.x == 2 ? k : 0ull;
    ghist[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kBlock) void k_trim_hist(const double* __restrict__ v, int64_t n,
                                                      const uint64_t* __restrict__ sel, int shift,
                                                      uint32_t* __restrict__ ghist, const uint8_t* __restrict__ prot) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t prefix = sel[0], mask = sel[1];
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) {
        if (prot && prot[t]) continue;  // protected terms are kept outside the selection
        const uint64_t k = mag_key(v[t]);
        if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&ghist[threadIdx.x], h[threadIdx.x]);
}

// The digit holding the k_rem-th largest key with the prefix (from the top).
__global__ void k_trim_pick(uint64_t* __restrict__ sel, uint32_t* __restrict__ ghist, int shift) {
    if (threadIdx.x == 0) {
        const uint64_t k = sel[2];
        uint64_t run = 0;
        int d = 255;
        for (; d > 0; --d) {
            if (run + ghist[d] >= k) break;
            run += ghist[d];
        }
        sel[2] = k - run;
        sel[0] |= (uint64_t)d << shift;
        sel[1] |= 255ull << shift;
    }
    __syncthreads();
    ghist[threadIdx.x] = 0;
}

// Per block of kOpsTile terms: keys above the threshold, and keys equal to it.
__global__ __launch_bounds__(kBlock) void k_trim_count(const double* __restrict__ v, int64_t n,
                                                       const uint64_t* __restrict__ sel, uint32_t* __restrict__ cgt,
                                                       uint32_t* __restrict__ ctie, const uint8_t* __restrict__ prot) {
    __shared__ uint32_t red[kWavesPerBlock];
    const uint64_t T = sel[0];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    uint32_t g = 0, e = 0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= (uint64_t)n) break;
        const uint64_t key = mag_key(v[t]);
        const bool pr = prot && prot[t];
        g += pr || key > T;
        e += !pr && key == T;
    }
    const uint32_t tg = block_sum(g, red);
    const uint32_t te = block_sum(e, red);
    if (threadIdx.x == 0) {
        cgt[blockIdx.x] = tg;
        ctie[blockIdx.x] = te;
    }
}

// Kept terms per block: every key above the threshold, and the block's share
// of the first sel[2] ties (toff = ties in earlier blocks).
__global__ void k_trim_keepcnt(const uint32_t* __restrict__ cgt, const uint32_t* __restrict__ ctie,
                               const uint64_t* __restrict__ toff, const uint64_t* __restrict__ sel, uint64_t nb,
                               uint32_t* __restrict__ ckeep) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const uint64_t kr = sel[2];
    const uint64_t take = toff[b] >= kr ? 0 : min((uint64_t)ctie[b], kr - toff[b]);
    ckeep[b] = cgt[b] + (uint32_t)take;
}

__global__ __launch_bounds__(kBlock) void k_trim_write(int as, SumView s, int64_t n, const uint64_t* __restrict__ sel,
                                                       const uint64_t* __restrict__ toff,
                                                       const uint64_t* __restrict__ koff, uint64_t* __restrict__ ox,
                                                       uint64_t* __restrict__ oz, uint8_t* __restrict__ of,
                                                       double* __restrict__ ov, int64_t os, uint64_t kmax,
                                                       const uint32_t* __restrict__ sj, uint32_t* __restrict__ oj,
                                                       const uint8_t* __restrict__ prot) {
    __shared__ uint32_t red[kWavesPerBlock];
    const uint64_t T = sel[0], kr = sel[2];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    uint32_t gt = 0, tie = 0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= (uint64_t)n) break;
        const uint64_t key = mag_key(s.v[t]);
        const bool pr = prot && prot[t];
        if (pr || key > T) gt |= 1u << k;
        if (!pr && key == T) tie |= 1u << k;
    }
    uint32_t tt;
    uint64_t r = toff[blockIdx.x] + block_exclusive_scan((uint32_t)__popc(tie), tt, red);  // rank of my first tie
    uint32_t keep = gt;
    for (int k = 0; k < kOpsIpt; ++k) {
        if (!(tie >> k & 1u)) continue;
        if (r < kr) keep |= 1u << k;
        ++r;
    }
    uint32_t tk;
    uint64_t o = koff[blockIdx.x] + block_exclusive_scan((uint32_t)__popc(keep), tk, red);
    for (int k = 0; k < kOpsIpt; ++k) {
        if (!(keep >> k & 1u)) continue;
        if (o >= kmax) break;  // the select keeps exactly kmax terms; never write past them
        const uint64_t t = base + k;
        for (int w = 0; w < as; ++w) {
            ox[(int64_t)w * os + o] = s.x[(int64_t)w * s.stride + t];
            oz[(int64_t)w * os + o] = s.z[(int64_t)w * s.stride + t];
        }
        of[o] = s.f[t];
        ov[o] = s.v[t];
        if (oj) oj[o] = sj[t];
        ++o;
    }
}

// prot[t] = 1 when term t's (x, z) pattern is one of the nk sorted distinct
// patterns in keys (x words then z words, as words each); per-block counts.
__global__ __launch_bounds__(kBlock) void k_trim_protect(int as, SumView s, int64_t n, const uint64_t* __restrict__ keys,
                                                         uint32_t nk, uint8_t* __restrict__ prot,
                                                         uint32_t* __restrict__ cnt) {
    __shared__ uint32_t red[kWavesPerBlock];
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t hit = 0;
    if (t < n) {
        uint32_t lo = 0, c = nk;  // first key >= term t
        while (c > 0) {
            const uint32_t half = c / 2, m = lo + half;
            const uint64_t* kk = keys + (size_t)m * 2 * as;
            int cmp = 0;
            for (int w = 0; w < 2 * as && cmp == 0; ++w) {
                const uint64_t a = kk[w];
                const uint64_t b = w < as ? s.x[(int64_t)w * s.stride + t] : s.z[(int64_t)(w - as) * s.stride + t];
                cmp = a < b ? -1 : a > b ? 1 : 0;
            }
            if (cmp < 0) {
                lo = m + 1;
                c -= half + 1;
            } else {
                c = half;
            }
        }
        if (lo < nk) {
            const uint64_t* kk = keys + (size_t)lo * 2 * as;
            bool eq = true;
            for (int w = 0; w < 2 * as; ++w)
                eq = eq && kk[w] == (w < as ? s.x[(int64_t)w * s.stride + t] : s.z[(int64_t)(w - as) * s.stride + t]);
            hit = eq ? 1u : 0u;
        }
        prot[t] = (uint8_t)hit;
    }
    const uint32_t tot = block_sum(hit, red);
    if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// v[t] *= c (PauliSum::scale, include/PauliSumUtils.h:395-398).
__global__ void k_scale(double* __restrict__ v, int64_t n, double c) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) v[t] *= c;
}

// Fixed-order partial sums of one selection per block: mode 0 lookup (terms
// with the query's (x, z), real relative phase, sign folded;
// include/PauliSumUtils.h:253-268), 1 trace (identity terms, real phase,
// sign folded; :231-243), 2 frobenius_norm2 (v^2; :403-410).
struct Query {
    uint64_t x[16], z[16];
    uint32_t f;
};
__global__ __launch_bounds__(kBlock) void k_select_partial(int mode, int as, SumView s, int64_t n, Query q,
                                                           double* __restrict__ part) {
    __shared__ double red[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kOpsTile + (uint64_t)threadIdx.x * kOpsIpt;
    double acc = 0.0;
    for (int k = 0; k < kOpsIpt; ++k) {
        const uint64_t t = base + k;
        if (t >= (uint64_t)n) break;
        const double v = s.v[t];
        if (mode == 2) {
            acc += v * v;
            continue;
        }
        bool match = true;
        for (int w = 0; w < as; ++w) {
            const uint64_t xw = s.x[(int64_t)w * s.stride + t], zw = s.z[(int64_t)w * s.stride + t];
            match = match && (mode == 0 ? (xw == q.x[w] && zw == q.z[w]) : (xw == 0 && zw == 0));
        }
        if (!match) continue;
        const uint32_t rel = mode == 0 ? ((uint32_t)s.f[t] - q.f) & 3u : (uint32_t)s.f[t] & 3u;
        if (rel & 1u) continue;
        acc += (rel & 2u) ? -v : v;
    }
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

}  // namespace

int op_trim(plh_ctx* ctx, plh_sum* s, int64_t max_terms) { return op_trim_keep(ctx, s, max_terms, nullptr); }

int op_scale(plh_ctx* ctx, plh_sum* s, double c) {
    if (s->n == 0) return PLH_OK;
    k_scale<<<nblk((uint64_t)s->n, kBlock), kBlock, 0, ctx->stream>>>(s->v, s->n, c);
    PLH_HIP(hipGetLastError());
    // keys and order unchanged; a zero factor leaves zeros (as the reference)
    if (c == 0.0) {
        s->canonical = false;
        s->sorted = false;
    }
    return PLH_OK;
}

int op_select_sum(plh_ctx* ctx, const plh_sum* s, int mode, const uint64_t* qx, const uint64_t* qz, uint32_t qf,
                  double* out) {
    if (s->n == 0) {
        *out = 0.0;
        return PLH_OK;
    }
    Query q{};
    for (int w = 0; w < s->as && mode == 0; ++w) {
        q.x[w] = qx[w];
        q.z[w] = qz[w];
    }
    q.f = qf & 3u;
    hipStream_t st = ctx->stream;
    const uint64_t nb = nblk((uint64_t)s->n, kOpsTile);
    void* p;
    PLH_TRY(ctx->ws.get(kWsScan1, nb * 8 + 64, &p));
    double* part = (double*)p;
    PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
    double* d_out = (double*)p + 41;
    k_select_partial<<<dim3((unsigned)nb), kBlock, 0, st>>>(mode, s->as, SumView{s->x, s->z, s->f, s->v, s->cap},
                                                            s->n, q, part);
    k_zse_final<<<1, kBlock, 0, st>>>(part, nb, d_out);
    PLH_HIP(hipGetLastError());
    uint64_t bits = 0;
    PLH_TRY(read_u64(ctx, (const uint64_t*)d_out, &bits));
    std::memcpy(out, &bits, sizeof(double));
    return PLH_OK;
}

int op_trim_keep(plh_ctx* ctx, plh_sum* s, int64_t max_terms, const plh_sum* keep) {
    if (max_terms < 0) return PLH_ERR_ARG;
    if (s->n <= max_terms) return PLH_OK;
    hipStream_t st = ctx->stream;
    const uint64_t n = (uint64_t)s->n;
    // trim(max_terms, keep) (include/PauliSumUtils.h:348-375): terms whose
    // (x, z) pattern is in keep are never dropped; the others compete for
    // the max_terms - protected places left
    uint8_t* prot = nullptr;
    uint64_t nprot = 0;
    if (keep && keep->n > 0) {
        const int as = s->as;
        std::vector<uint64_t> kx((size_t)keep->as * keep->n), kz((size_t)keep->as * keep->n);
        std::vector<uint8_t> kf((size_t)keep->n);
        std::vector<double> kv((size_t)keep->n);
        for (int w = 0; w < keep->as; ++w) {
            PLH_HIP(hipMemcpyAsync(kx.data() + (size_t)w * keep->n, keep->x + (int64_t)w * keep->cap,
                                   (size_t)keep->n * 8, hipMemcpyDeviceToHost, st));
            PLH_HIP(hipMemcpyAsync(kz.data() + (size_t)w * keep->n, keep->z + (int64_t)w * keep->cap,
                                   (size_t)keep->n * 8, hipMemcpyDeviceToHost, st));
        }
        PLH_HIP(hipStreamSynchronize(st));
        std::vector<std::vector<uint64_t>> pats((size_t)keep->n, std::vector<uint64_t>((size_t)2 * as, 0));
        for (int64_t t = 0; t < keep->n; ++t)
            for (int w = 0; w < std::min(as, keep->as); ++w) {
                pats[(size_t)t][(size_t)w] = kx[(size_t)w * keep->n + t];
                pats[(size_t)t][(size_t)(as + w)] = kz[(size_t)w * keep->n + t];
            }
        std::sort(pats.begin(), pats.end());
        pats.erase(std::unique(pats.begin(), pats.end()), pats.end());
        std::vector<uint64_t> flat;
        flat.reserve(pats.size() * 2 * as);
        for (auto& v : pats) flat.insert(flat.end(), v.begin(), v.end());
        const uint64_t nb = nblk(n, kBlock);
        void* p;
        PLH_TRY(ctx->ws.get(kWsKeep, flat.size() * 8 + n + 256 + nb * 4 + (nb + 1) * 8 + 256, &p));
        uint64_t* dkeys = (uint64_t*)p;
        prot = (uint8_t*)(dkeys + flat.size());
        uint32_t* cnt = (uint32_t*)((char*)prot + ((n + 255) & ~(uint64_t)255));
        uint64_t* coff = (uint64_t*)(cnt + nb + (nb & 1));
        PLH_HIP(hipMemcpyAsync(dkeys, flat.data(), flat.size() * 8, hipMemcpyHostToDevice, st));
        k_trim_protect<<<(unsigned)nb, kBlock, 0, st>>>(as, SumView{s->x, s->z, s->f, s->v, s->cap}, s->n, dkeys,
                                                       (uint32_t)pats.size(), prot, cnt);
        PLH_HIP(hipGetLastError());
        PLH_TRY(scan_u32_to_u64_read(ctx, cnt, coff, nb, coff + nb, &nprot));
    }
    const uint64_t budget = (uint64_t)max_terms > nprot ? (uint64_t)max_terms - nprot : 0;
    const uint64_t k = nprot + budget;  // kept terms (fewer than n: n > max_terms >= budget + ...)
    plh_sum tmp;
    tmp.ctx = ctx;
    tmp.as = s->as;
    tmp.nq = s->nq;
    tmp.has_job = s->has_job;
    PLH_TRY(sum_reserve(&tmp, (int64_t)k, 0));
    if (k > 0) {
        const uint64_t nb = nblk(n, kOpsTile);
        void* p;
        PLH_TRY(ctx->ws.get(kWsScan0, nb * 12 + 256, &p));
        uint32_t* cgt = (uint32_t*)p;
        uint32_t* ctie = cgt + nb;
        uint32_t* ckeep = ctie + nb;
        PLH_TRY(ctx->ws.get(kWsScan1, nb * 16 + 256, &p));
        uint64_t* toff = (uint64_t*)p;
        uint64_t* koff = toff + nb;
        PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
        uint64_t* sel = (uint64_t*)p + 64;            // 3 words
        uint32_t* ghist = (uint32_t*)((uint64_t*)p + 128);  // 256 counters
        uint64_t* d_total = (uint64_t*)p + 40;
        const unsigned grid = (unsigned)std::min<uint64_t>(nblk(n, kBlock), 2048);
        k_trim_init<<<1, 256, 0, st>>>(sel, ghist, budget);
        for (int shift = 56; shift >= 0; shift -= 8) {
            k_trim_hist<<<grid, kBlock, 0, st>>>(s->v, s->n, sel, shift, ghist, prot);
            k_trim_pick<<<1, 256, 0, st>>>(sel, ghist, shift);
        }
        k_trim_count<<<dim3((unsigned)nb), kBlock, 0, st>>>(s->v, s->n, sel, cgt, ctie, prot);
        PLH_HIP(hipGetLastError());
        PLH_TRY(scan_u32_to_u64(ctx, ctie, toff, nb, d_total));
        k_trim_keepcnt<<<(unsigned)nblk(nb, 256), 256, 0, st>>>(cgt, ctie, toff, sel, nb, ckeep);
        PLH_TRY(scan_u32_to_u64(ctx, ckeep, koff, nb, d_total));
        k_trim_write<<<dim3((unsigned)nb), kBlock, 0, st>>>(s->as, SumView{s->x, s->z, s->f, s->v, s->cap}, s->n,
                                                            sel, toff, koff, tmp.x, tmp.z, tmp.f, tmp.v, tmp.cap, k,
                                                            s->j, tmp.j, prot);
        PLH_HIP(hipGetLastError());
    }
    const bool canon = s->canonical, srt = s->sorted;
    sum_take_storage(s, &tmp);
    s->n = (int64_t)k;
    s->canonical = canon;  // a subset in term order of a canonical sum stays canonical
    s->sorted = srt;
    return PLH_OK;
}

int op_zero_state_expectation(plh_ctx* ctx, const plh_sum* s, double* out) {
    hipStream_t st = ctx->stream;
    if (s->n == 0) {
        *out = 0.0;
        return PLH_OK;
    }
    const uint64_t nb = nblk((uint64_t)s->n, kOpsTile);
    void* p;
    PLH_TRY(ctx->ws.get(kWsScan1, nb * 8 + 64, &p));
    double* part = (double*)p;
    PLH_TRY(ctx->ws.get(kWsMisc, 4096, &p));
    double* d_out = (double*)p + 41;
    SumView v{s->x, s->z, s->f, s->v, s->cap};
    k_zse_partial<<<dim3((unsigned)nb), kBlock, 0, st>>>(s->as, v, s->n, part);
    k_zse_final<<<1, kBlock, 0, st>>>(part, nb, d_out);
    PLH_HIP(hipGetLastError());
    uint64_t bits = 0;
    PLH_TRY(read_u64(ctx, (const uint64_t*)d_out, &bits));
    std::memcpy(out, &bits, sizeof(double));
    return PLH_OK;
}


}  // namespace plh

// ---------------------------------------------- commutator without allocation
// plh_launch_commutator_soa_noalloc: the raw commutator of plh_commutator
// (same kernels, same deterministic (i over a, j over b) order) into
// caller-allocated planes, for callers that cannot allocate device memory
// (graph capture, serving loops, harnesses that forbid it).  Its scratch --
// the per-tile anticommuting counts, their exclusive scan and the scan's
// block sums -- is the file-scope array `scratch` below, one per device,
// sized for kNoAllocMaxPairs pairs.  Host code reaches the array only
// through the per-device address from hipGetSymbolAddress (the array's name
// is a host-side stand-in, not a device address).  Every call on a device
// shares its array: calls that overlap on one device race.
namespace plh {
namespace {

constexpr uint64_t kNoAllocMaxPairs = 1ull << 32;
constexpr uint64_t kNoAllocTiles = kNoAllocMaxPairs / kOpsTile;  // 2^20 tiles of 4096 pairs
constexpr size_t kNoAllocCntBytes = kNoAllocTiles * 4;
constexpr size_t kNoAllocOffBytes = kNoAllocTiles * 8;
constexpr size_t kNoAllocScratchBytes = kNoAllocCntBytes + kNoAllocOffBytes + (64u << 10);  // ~12 MiB

}  // namespace
}  // namespace plh

static __device__ __align__(256) unsigned char scratch[plh::kNoAllocScratchBytes];

namespace plh {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int device_scratch(void** out) {
    int dev = 0;
    PLH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("plh_launch_commutator_soa_noalloc: device index beyond " + std::to_string(kMaxDevices));
        return PLH_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        PLH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("plh_launch_commutator_soa_noalloc: no device address for the scratch array");
            return PLH_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PLH_OK;
}

__global__ void k_store_u64(unsigned long long* p, unsigned long long v) { *p = v; }

}  // namespace
}  // namespace plh

extern "C" int plh_launch_commutator_soa_noalloc(int array_size, int multiply_by_i, const uint64_t* a_x,
                                                 const uint64_t* a_z, const uint8_t* a_flags, const double* a_val,
                                                 const uint64_t* b_x, const uint64_t* b_z, const uint8_t* b_flags,
                                                 const double* b_val, uint64_t* out_x, uint64_t* out_z,
                                                 uint8_t* out_flags, double* out_val, unsigned long long* d_count,
                                                 int64_t n, int64_t m, int64_t capacity, void* stream) {
    using namespace plh;
    if (array_size < 1 || array_size > 16) {
        set_error("plh_launch_commutator_soa_noalloc: unsupported array_size " + std::to_string(array_size) +
                  " (supported: 1..16)");
        return PLH_ERR_AS;
    }
    if (n < 0 || m < 0 || capacity < 0 || !d_count) {
        set_error("plh_launch_commutator_soa_noalloc: negative size or null count");
        return PLH_ERR_ARG;
    }
    if (n > 0 && m > 0 && (uint64_t)n * (uint64_t)m > kNoAllocMaxPairs) {
        set_error("plh_launch_commutator_soa_noalloc: n*m = " + std::to_string((uint64_t)n * (uint64_t)m) +
                  " pairs exceeds the scratch's limit of 2^32 pairs");
        return PLH_ERR_SIZE;
    }
    if ((n > 0 && (!a_x || !a_z || !a_flags || !a_val)) || (m > 0 && (!b_x || !b_z || !b_flags || !b_val)) ||
        (capacity > 0 && (!out_x || !out_z || !out_flags || !out_val))) {
        set_error("plh_launch_commutator_soa_noalloc: null plane");
        return PLH_ERR_ARG;
    }
    void* sp = nullptr;
    PLH_TRY(device_scratch(&sp));
    hipStream_t st = (hipStream_t)stream;
    const uint64_t total = (uint64_t)n * (uint64_t)m;
    if (total == 0) {
        k_store_u64<<<1, 1, 0, st>>>(d_count, 0ull);
        PLH_HIP(hipGetLastError());
        return PLH_OK;
    }
    const uint64_t nb = (total + kOpsTile - 1) / kOpsTile;
    uint32_t* cnt = (uint32_t*)sp;
    uint64_t* off = (uint64_t*)((char*)sp + kNoAllocCntBytes);
    uint64_t* bsum = (uint64_t*)((char*)sp + kNoAllocCntBytes + kNoAllocOffBytes);
    static_assert(kNoAllocTiles <= 1ull << 20 && (64u << 10) >= ((1ull << 20) / 2048 + 1) * 8,
                  "block sums of the scan fit the scratch's tail");
    const SumView A{a_x, a_z, a_flags, a_val, n}, B{b_x, b_z, b_flags, b_val, m};
    const uint32_t by_i = multiply_by_i ? 1u : 0u;
#define PLH_NA(ASV)                                                                                              \
    case ASV:                                                                                                    \
        k_comm_count<ASV><<<dim3((unsigned)nb), kBlock, 0, st>>>(array_size, A, B, (uint64_t)m, total, cnt);     \
        PLH_HIP(hipGetLastError());                                                                              \
        PLH_TRY(scan_u32_to_u64_scratch(st, cnt, off, nb, (uint64_t*)d_count, bsum));                            \
        k_comm_write<ASV><<<dim3((unsigned)nb), kBlock, 0, st>>>(array_size, A, B, (uint64_t)m, total, by_i, off, \
                                                                 out_x, out_z, out_flags, out_val, capacity,     \
                                                                 (uint64_t)capacity);                           \
        break;
    switch (padded(array_size)) {
        PLH_NA(1) PLH_NA(2) PLH_NA(4) PLH_NA(8) PLH_NA(16)
        default: return PLH_ERR_AS;
    }
#undef PLH_NA
    PLH_HIP(hipGetLastError());
    return PLH_OK;
}
