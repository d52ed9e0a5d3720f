// plh_serve.hip -- batched propagation: many jobs in one device-resident sum.
//
// A batched sum is a plh_sum with a job plane: term t belongs to job slot
// j[t].  The slot is the leading key word of every collation, so the batch
// keeps each job's terms contiguous and ordered exactly as that job alone
// would be, under its own cutoff and weight limit.  A full step advances
// every job in flight by its Clifford gates up to and including its next
// rotation (Heisenberg order) and collates the batch; a sorted step keeps
// the order and advances each job whose next op is an order-keeping
// rotation through a window of up to 16 of them without a sort (the first
// one that changes a term is applied, the ones before it are no-ops).
// Either takes a fixed number of launches, so the per-gate launch and
// synchronisation cost of the plain path is paid once per step for all jobs
// instead of once per job.
//
// Per job the semantics are those of a Clifford + rotation program run by
// propagate() (include/PauliPropagation.h:53-63) with the Clifford methods
// (include/PauliSumClifford.h:43-81, 212-239): Clifford gates conjugate in
// place; a rotation is evolve (include/PauliSumOperations.h:106-183) followed
// by apply_truncation (include/PauliPropagation.h:35-47), fused into one
// collation.  The fusion is exact while the job's (x, z, imag) keys are
// distinct, which submission checks and every step preserves.  Steps with
// Clifford gates only (the program's leading gates) still fold and order
// the job's terms but apply no truncation, as propagate() would not.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <string>
#include <unordered_map>
#include <vector>

#include "plh_block.hpp"
#include "plh_engine.hpp"
#include "plh_sched.hpp"
#include "plh_sources.hpp"

namespace plh {

namespace {

constexpr int kIpt = 16;
// Per-step job tables (device, one upload per step), one entry per slot.
struct StepTables {
    const uint32_t* cl_off;   // [slots + 1] range of the slot's Clifford ops
    const uint32_t* cl_op;    // kind << 28 | q0 << 14 | q1 (kind 0 H, 1 S, 2 CNOT); kind 3: apply_table_op;
                              // kind 5: gate layer (low 28 bits: its index in layers)
    const uint64_t* layers;   // kLayerWords words per gate layer (GateLayer)
    const uint8_t* rot_kind;  // 0 none, 1 sign flip, 2 flip + multiply, 3 multiply, 4 general
    const uint8_t* rot_psf;   // flags of ps = P with the i bit flipped
    const double* rot_c;
    const double* rot_s;
    const uint64_t* rot_p;    // per slot: x[AS] then z[AS]
    double* eps;              // collation filters of this step
    int64_t* maxw;
    uint8_t* alive;
    const uint8_t* isc;      // mixed steps: 1 for a C job (collated apart)
    const uint8_t* alive_r;  // mixed steps: alive and an R job
};

// Tables of the gate set (16 nibbles of new local bits, 16 sign bits per
// gate, as k_clifford_table takes them), written once per device when a
// server is created.
__constant__ uint64_t c_gate_tbl[kNumGates];
__constant__ uint32_t c_gate_sgn[kNumGates];

// A table gate on a term held in registers: op = 3 << 28 | gate << 20 |
// q0 << 10 | q1 (qubits below 1024).
template <int AS>
PLH_DEV void apply_table_op(uint32_t op, uint64_t* x, uint64_t* z, uint32_t& f) {
    const uint32_t g = (op >> 20) & 0xFFu, q0 = (op >> 10) & 0x3FFu, q1 = op & 0x3FFu;
    const bool two = g >= (uint32_t)kFirstTwoQubitGate;
    const int wc = (int)(q0 >> 6), wt = (int)(q1 >> 6);
    const uint64_t mc = 1ull << (q0 & 63), mt = two ? 1ull << (q1 & 63) : 0ull;
    uint32_t b = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        if (k == wc) b |= (uint32_t)((x[k] & mc) != 0) | ((uint32_t)((z[k] & mc) != 0) << 1);
        if (two && k == wt) b |= ((uint32_t)((x[k] & mt) != 0) << 2) | ((uint32_t)((z[k] & mt) != 0) << 3);
    }
    const uint32_t e = (uint32_t)(c_gate_tbl[g] >> (4 * b)) & 15u;
    f ^= ((c_gate_sgn[g] >> b) & 1u) << 1;
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        if (k == wc) {
            x[k] = (x[k] & ~mc) | ((e & 1u) ? mc : 0ull);
            z[k] = (z[k] & ~mc) | ((e & 2u) ? mc : 0ull);
        }
        if (two && k == wt) {
            x[k] = (x[k] & ~mt) | ((e & 4u) ? mt : 0ull);
            z[k] = (z[k] & ~mt) | ((e & 8u) ? mt : 0ull);
        }
    }
}

// A gate layer: H, S and CNOT ops on pairwise distinct qubits, as masks
// (qubits below 128): H on h, S on s, CNOT(c -> c + d) for each control
// bit c of cm, one shift d (|d| < 64) per layer.  Its ops commute (each
// reads and writes only its own qubits, and sign flips add), so the layer
// is all of them at once, from the term's bits before the layer.
constexpr int kLayerWords = 7;  // h[2], s[2], cm[2], d

// v's bits moved d (1..63) qubits up (out bit q + d = v bit q) or down.
template <int AS>
PLH_DEV void bits_up(const uint64_t* v, int d, uint64_t* out) {
#pragma unroll
    for (int k = AS - 1; k >= 0; --k) out[k] = (v[k] << d) | (k > 0 ? v[k - 1] >> (64 - d) : 0ull);
}
template <int AS>
PLH_DEV void bits_down(const uint64_t* v, int d, uint64_t* out) {
#pragma unroll
    for (int k = 0; k < AS; ++k) out[k] = (v[k] >> d) | (k + 1 < AS ? v[k + 1] << (64 - d) : 0ull);
}

template <int AS>
PLH_DEV void apply_gate_layer(const uint64_t* L, uint64_t* x, uint64_t* z, uint32_t& f) {
    const int64_t d = (int64_t)L[6];
    const int ad = (int)(d < 0 ? -d : d);
    uint64_t xt[AS], zt[AS], xu[AS], xc[AS];
    // the target qubits' bits at their controls' positions
    if (d > 0) {
        bits_down<AS>(x, ad, xt);
        bits_down<AS>(z, ad, zt);
    } else if (d < 0) {
        bits_up<AS>(x, ad, xt);
        bits_up<AS>(z, ad, zt);
    }
    uint32_t par = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        const uint64_t h = L[k], sm = L[2 + k], cm = d != 0 ? L[4 + k] : 0ull;
        par += (uint32_t)popc(x[k] & z[k] & (h | sm));                         // H and S: Y gets -1
        xc[k] = x[k] & cm;
        if (d != 0) par += (uint32_t)popc(xc[k] & zt[k] & ~(xt[k] ^ z[k]) & cm);  // CNOT: xc zt (1 ^ xt ^ zc)
    }
    // the controls' x bits at their targets
    if (d > 0) bits_up<AS>(xc, ad, xu);
    else if (d < 0) bits_down<AS>(xc, ad, xu);
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        const uint64_t h = L[k], sm = L[2 + k], cm = d != 0 ? L[4 + k] : 0ull;
        const uint64_t dh = (x[k] ^ z[k]) & h;
        const uint64_t nx = x[k] ^ dh ^ (d != 0 ? xu[k] : 0ull);
        const uint64_t nz = z[k] ^ dh ^ (x[k] & sm) ^ (zt[k] & cm);
        x[k] = nx;
        z[k] = nz;
    }
    f ^= (par & 1u) << 1;
}

// One Clifford conjugation of a term held in registers: the update of the
// plain k_clifford kernel (H, S, CNOT of include/PauliSumClifford.h), or a
// table gate.  Words are selected by compare so the term stays in registers.
template <int AS>
PLH_DEV void apply_clifford_op(uint32_t op, uint64_t* x, uint64_t* z, uint32_t& f, const uint64_t* layers) {
    const uint32_t kind = op >> 28, q0 = (op >> 14) & 0x3FFFu, q1 = op & 0x3FFFu;
    if (kind == 5) {
        apply_gate_layer<AS>(layers + (size_t)(op & 0x0FFFFFFFu) * kLayerWords, x, z, f);
        return;
    }
    if (kind == 3) {
        apply_table_op<AS>(op, x, z, f);
        return;
    }
    const int wc = (int)(q0 >> 6);
    const uint64_t mc = 1ull << (q0 & 63);
    uint64_t xw = 0, zw = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k)
        if (k == wc) { xw = x[k]; zw = z[k]; }
    const uint32_t xc = (xw & mc) != 0, zc = (zw & mc) != 0;
    if (kind == 0) {  // H: X <-> Z, Y -> -Y
        f ^= (xc & zc) << 1;
        const uint64_t d = (xw ^ zw) & mc;
#pragma unroll
        for (int k = 0; k < AS; ++k)
            if (k == wc) { x[k] = xw ^ d; z[k] = zw ^ d; }
    } else if (kind == 1) {  // S
        f ^= (xc & zc) << 1;
#pragma unroll
        for (int k = 0; k < AS; ++k)
            if (k == wc) z[k] = zw ^ (xw & mc);
    } else {  // CNOT(q0 -> q1)
        const int wt = (int)(q1 >> 6);
        const uint64_t mt = 1ull << (q1 & 63);
        uint64_t xtw = 0, ztw = 0;
#pragma unroll
        for (int k = 0; k < AS; ++k)
            if (k == wt) { xtw = x[k]; ztw = z[k]; }
        const uint64_t xt = (xtw & mt) != 0, zt = (ztw & mt) != 0;
        f ^= (uint32_t)(((uint64_t)xc & zt & (1ull ^ xt ^ (uint64_t)zc)) << 1);
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            if (k == wt) x[k] ^= (0ull - (uint64_t)xc) & mt;
            if (k == wc) z[k] ^= (0ull - zt) & mc;
        }
    }
}

template <int AS>
PLH_DEV bool anticommutes_p(const uint64_t* x, const uint64_t* z, const uint64_t* p) {
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k) acc ^= (x[k] & p[AS + k]) ^ (z[k] & p[k]);
    return popc(acc) & 1;
}

// term * ps in registers (PauliSum::multiply(i, ps), include/PauliSumOperations.h:185-209).
template <int AS>
PLH_DEV uint32_t times_p(uint64_t* x, uint64_t* z, uint32_t f, const uint64_t* p, uint32_t psf) {
    uint64_t c1 = 0, c2 = 0;
#pragma unroll
    for (int k = 0; k < AS; ++k) {
        phase_word(x[k], z[k], p[k], p[AS + k], c1, c2);
        x[k] ^= p[k];
        z[k] ^= p[AS + k];
    }
    return phase_of(f, psf, c1, c2);
}

// Clifford ops, then the fast-path rotations (:121-147) in place; the
// general path (kind 4) happens inside the step's collation.
template <int AS>
__global__ __launch_bounds__(kBlock) void k_batch_gates(uint64_t* __restrict__ X, uint64_t* __restrict__ Z,
                                                        uint8_t* __restrict__ F, const uint32_t* __restrict__ J,
                                                        int64_t stride, int64_t n, StepTables tb, int ipt) {
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * (uint64_t)ipt;
    for (int k = 0; k < ipt; ++k) {
        const uint64_t t = base + (uint64_t)k * kBlock + threadIdx.x;
        if (t >= (uint64_t)n) break;
        const uint32_t job = J[t];
        const uint32_t o0 = tb.cl_off[job], o1 = tb.cl_off[job + 1];
        const uint32_t kind = tb.rot_kind[job];
        if (o0 == o1 && (kind == 0 || kind == 4)) continue;
        uint64_t x[AS], z[AS];
#pragma unroll
        for (int w = 0; w < AS; ++w) {
            x[w] = X[(int64_t)w * stride + t];
            z[w] = Z[(int64_t)w * stride + t];
        }
        uint32_t f = F[t];
        for (uint32_t o = o0; o < o1; ++o) apply_clifford_op<AS>(tb.cl_op[o], x, z, f, tb.layers);
        if (kind >= 1 && kind <= 3) {
            const uint64_t* p = tb.rot_p + (size_t)job * 2 * AS;
            if (anticommutes_p<AS>(x, z, p)) {
                if (kind == 1 || kind == 2) f ^= 2u;
                if (kind >= 2) f = times_p<AS>(x, z, f, p, tb.rot_psf[job]);
            }
        }
#pragma unroll
        for (int w = 0; w < AS; ++w) {
            X[(int64_t)w * stride + t] = x[w];
            Z[(int64_t)w * stride + t] = z[w];
        }
        F[t] = (uint8_t)f;
    }
}

// <0|O_job|0> (include/PauliSumUtils.h:162-177) for finished jobs of a
// job-ordered batch, in fixed order: kZseParts workgroups per job each sum a
// fixed slice of the job's contiguous range (a slice depends only on the
// job's length, so a job's sum is the same alone or in any mix), then one
// thread per job adds the slices in order.  The first slice's workgroup also
// returns the job's range.
constexpr int kZseParts = 64;

__global__ __launch_bounds__(kBlock) void k_batch_zse_part(int as, SumView s, const uint32_t* __restrict__ J,
                                                           int64_t n, const uint32_t* __restrict__ jobs,
                                                           double* __restrict__ part, int64_t* __restrict__ range) {
    __shared__ double red[kWavesPerBlock];
    __shared__ int64_t s_ab[2];
    const uint32_t f = blockIdx.x / kZseParts, q = blockIdx.x % kZseParts;
    if (threadIdx.x == 0) {
        const uint32_t job = jobs[f];
        int64_t lo = 0, hi = n;  // first index with J >= job
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (J[mid] < job) lo = mid + 1; else hi = mid;
        }
        s_ab[0] = lo;
        hi = n;
        while (lo < hi) {  // first index with J > job
            const int64_t mid = (lo + hi) >> 1;
            if (J[mid] <= job) lo = mid + 1; else hi = mid;
        }
        s_ab[1] = lo;
    }
    __syncthreads();
    const int64_t a = s_ab[0], b = s_ab[1], len = b - a;
    const int64_t i0 = a + len * q / kZseParts, i1 = a + len * (q + 1) / kZseParts;
    double acc = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) {
        bool has_xy = false;
        for (int w = 0; w < as; ++w) has_xy |= s.x[(int64_t)w * s.stride + i] != 0;
        const uint32_t fl = s.f[i];
        if (has_xy || (fl & 1u)) continue;
        acc += (fl & 2u) ? -s.v[i] : s.v[i];
    }
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tot;
        if (q == 0) {
            range[2 * f] = a;
            range[2 * f + 1] = b;
        }
    }
}

__global__ void k_batch_zse_final(const double* __restrict__ part, uint32_t F, double* __restrict__ out) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    double acc = 0.0;
    for (int q = 0; q < kZseParts; ++q) acc += part[(size_t)f * kZseParts + q];
    out[f] = acc;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned blocks_for(uint64_t n, uint64_t per) { return (unsigned)((n + per - 1) / per); }

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

using sched::Job;

struct Result {
    int64_t handle = -1;
    double expectation = 0.0;
    int64_t terms = 0;
    double latency_s = 0.0;
    std::vector<uint64_t> x, z;
    std::vector<uint8_t> f;
    std::vector<double> v;
};

}  // namespace
}  // namespace plh

struct plh_server {
    plh_ctx* ctx = nullptr;
    int as = 2;
    bool keep_terms = false;
    plh_sum* batch = nullptr;     // job-plane sum holding every job in flight (sorted after each step)
    plh_sum* incoming = nullptr;  // observables admitted for the coming step
    plh_sum* side = nullptr;      // a mixed step's C jobs, collated apart
    plh::sched::Scheduler* sched = nullptr;
    std::deque<plh::Result> done;
    std::unordered_map<int64_t, plh::Result> kept;
    int64_t next_handle = 0;
    void* d_step = nullptr;
    size_t step_cap = 0;
    std::vector<uint8_t> h_step;
    // per slot, the rotations a sorted step consumed (host-pinned, written
    // by the device; adv_dev is its device address)
    uint8_t* adv_host = nullptr;
    uint8_t* adv_dev = nullptr;
    std::chrono::steady_clock::time_point t0;
    int64_t steps = 0;
    // set when a step's device work failed part-way: the batch may hold
    // half-applied gates, so the server refuses further steps
    bool failed = false;
    std::string failure;
};

namespace plh {
namespace {

double now_s(const plh_server* s) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

bool keys_distinct(const Job& j) {
    const int a = j.job_as;
    std::vector<std::vector<uint64_t>> keys((size_t)j.n0);
    for (int64_t t = 0; t < j.n0; ++t) {
        auto& k = keys[(size_t)t];
        for (int w = 0; w < a; ++w) k.push_back(j.ox[(size_t)(w * j.n0 + t)]);
        for (int w = 0; w < a; ++w) k.push_back(j.oz[(size_t)(w * j.n0 + t)]);
        k.push_back(j.of[(size_t)t] & 1u);
    }
    std::sort(keys.begin(), keys.end());
    return std::adjacent_find(keys.begin(), keys.end()) == keys.end();
}

// ---- per-job trim (apply_truncation's last step, include/PauliPropagation.h:
// 43-46; PauliSum::trim, include/PauliSumUtils.h:330-343): a job whose step
// truncated and that holds more than its max_terms keeps its max_terms terms
// of largest |c|, in term order; among equal magnitudes the lower index
// first (as the engine's single-sum trim).  One workgroup per job selects
// the max_terms-th largest |c| by its IEEE bits (non-negative doubles order
// as their bits), eight 8-bit digits, then finds the tie cut in term order.
constexpr int kTrimThreads = 1024;

__global__ void k_job_offsets(const uint32_t* __restrict__ J, uint32_t n, uint32_t njobs,
                              uint32_t* __restrict__ off) {
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

PLH_DEV uint64_t mag_bits(double v) { return (uint64_t)__double_as_longlong(fabs(v)); }

__global__ __launch_bounds__(kTrimThreads) void k_job_trim_select(const double* __restrict__ v,
                                                                  const uint32_t* __restrict__ joff,
                                                                  const int64_t* __restrict__ maxt,
                                                                  uint8_t* __restrict__ trim,
                                                                  uint64_t* __restrict__ thr,
                                                                  uint32_t* __restrict__ cut) {
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_prefix, s_mask, s_need;
    __shared__ uint32_t red[kTrimThreads / kWave];
    __shared__ uint32_t s_cut;
    const uint32_t job = blockIdx.x, tid = threadIdx.x;
    const uint32_t a = joff[job], b = joff[job + 1];
    const int64_t m = maxt[job];
    if (m < 0 || (int64_t)(b - a) <= m) {
        if (tid == 0) trim[job] = 0;
        return;
    }
    if (tid == 0) {
        s_prefix = 0;
        s_mask = 0;
        s_need = (uint64_t)m;  // the m-th largest magnitude (m >= 0; m == 0 keeps nothing)
    }
    __syncthreads();
    if (m > 0) {
        for (int shift = 56; shift >= 0; shift -= 8) {
            for (uint32_t d = tid; d < 256; d += kTrimThreads) hist[d] = 0;
            __syncthreads();
            const uint64_t prefix = s_prefix, mask = s_mask;
            for (uint32_t t = a + tid; t < b; t += kTrimThreads) {
                const uint64_t key = mag_bits(v[t]);
                if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
            }
            __syncthreads();
            if (tid == 0) {
                uint64_t need = s_need, cum = 0;
                int d = 255;
                for (; d > 0; --d) {
                    if (cum + hist[d] >= need) break;
                    cum += hist[d];
                }
                s_need = need - cum;
                s_prefix = prefix | ((uint64_t)d << shift);
                s_mask = mask | (255ull << shift);
            }
            __syncthreads();
        }
    }
    // ties: the s_need lowest-index terms of magnitude T are kept
    const uint64_t T = s_prefix;
    const uint64_t need = m > 0 ? s_need : 0;
    if (tid == 0) s_cut = a;
    __syncthreads();
    if (need > 0) {
        uint32_t run = 0;
        for (uint32_t c0 = a; c0 < b; c0 += kTrimThreads) {
            const uint32_t t = c0 + tid;
            const uint32_t eq = (t < b && mag_bits(v[t]) == T) ? 1u : 0u;
            uint32_t tot;
            const uint32_t ex = block_exclusive_scan<kTrimThreads / kWave>(eq, tot, red);
            if (eq && run + ex + 1 == need) s_cut = t + 1;
            run += tot;
            __syncthreads();
            if (run >= need) break;
        }
    }
    __syncthreads();
    if (tid == 0) {
        trim[job] = 1;
        thr[job] = m > 0 ? T : ~0ull;  // m == 0: nothing passes
        cut[job] = s_cut;
    }
}

__global__ __launch_bounds__(kBlock) void k_job_trim_flags(const double* __restrict__ v,
                                                           const uint32_t* __restrict__ J, uint32_t n,
                                                           const uint8_t* __restrict__ trim,
                                                           const uint64_t* __restrict__ thr,
                                                           const uint32_t* __restrict__ cut,
                                                           uint32_t* __restrict__ keep) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t job = J[t];
    bool k = true;
    if (trim[job]) {
        const uint64_t key = mag_bits(v[t]), T = thr[job];
        k = T != ~0ull && (key > T || (key == T && t < cut[job]));
    }
    keep[t] = k ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void k_job_trim_write(SumView s, const uint32_t* __restrict__ J, int as,
                                                           uint32_t n, const uint32_t* __restrict__ keep,
                                                           const uint64_t* __restrict__ pos, uint64_t* __restrict__ ox,
                                                           uint64_t* __restrict__ oz, uint8_t* __restrict__ of,
                                                           double* __restrict__ ov, uint32_t* __restrict__ oj,
                                                           int64_t os) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n || !keep[t]) return;
    const uint64_t o = pos[t];
    for (int w = 0; w < as; ++w) {
        ox[(int64_t)w * os + o] = s.x[(int64_t)w * s.stride + t];
        oz[(int64_t)w * os + o] = s.z[(int64_t)w * s.stride + t];
    }
    of[o] = s.f[t];
    ov[o] = s.v[t];
    oj[o] = J[t];
}

// Trims the jobs of the batch whose entry of maxt is set and that hold more
// terms than it (the batch stays job-ordered and sorted).
int trim_jobs(plh_ctx* ctx, plh_sum* B, const int64_t* d_maxt, uint32_t slots) {
    hipStream_t st = ctx->stream;
    const uint64_t n = (uint64_t)B->n;
    if (n == 0) return PLH_OK;
    if (n > 0xFFFFFFFFull) {
        set_error("plh_server_step: batch too large to trim");
        return PLH_ERR_SIZE;
    }
    void* p;
    const size_t sj = ((size_t)slots + 1 + 63) & ~(size_t)63;
    PLH_TRY(ctx->ws.get(kWsKeep, sj * (4 + 1 + 8 + 4) + 256 + ((n * 4 + 255) & ~(size_t)255) + (n + 1) * 8, &p));
    char* c = (char*)p;
    uint32_t* joff = (uint32_t*)c;
    c += sj * 4;
    uint8_t* trim = (uint8_t*)c;
    c += sj;
    uint64_t* thr = (uint64_t*)(((uintptr_t)c + 7) & ~(uintptr_t)7);
    c = (char*)(thr + sj);
    uint32_t* cut = (uint32_t*)c;
    c += sj * 4 + 256;
    uint32_t* keep = (uint32_t*)(((uintptr_t)c + 255) & ~(uintptr_t)255);
    uint64_t* pos = (uint64_t*)((char*)keep + ((n * 4 + 255) & ~(size_t)255));
    k_job_offsets<<<blocks_for((uint64_t)slots + 1, 256), 256, 0, st>>>(B->j, (uint32_t)n, slots, joff);
    k_job_trim_select<<<dim3(slots), kTrimThreads, 0, st>>>(B->v, joff, d_maxt, trim, thr, cut);
    k_job_trim_flags<<<blocks_for(n, kBlock), kBlock, 0, st>>>(B->v, B->j, (uint32_t)n, trim, thr, cut, keep);
    PLH_HIP(hipGetLastError());
    uint64_t total = 0;
    PLH_TRY(scan_u32_to_u64_read(ctx, keep, pos, n, pos + n, &total));
    if (total == n) return PLH_OK;
    plh_sum tmp;
    tmp.ctx = ctx;
    tmp.as = B->as;
    tmp.nq = B->nq;
    tmp.has_job = true;
    PLH_TRY(sum_reserve(&tmp, (int64_t)std::max<uint64_t>(total, 1), 0));
    if (total) {
        k_job_trim_write<<<blocks_for(n, kBlock), kBlock, 0, st>>>(SumView{B->x, B->z, B->f, B->v, B->cap}, B->j,
                                                                    B->as, (uint32_t)n, keep, pos, tmp.x, tmp.z,
                                                                    tmp.f, tmp.v, tmp.j, tmp.cap);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            sum_free_storage(&tmp);
            PLH_HIP(e);
        }
    }
    const int64_t nq = B->nq;
    const bool sorted = B->sorted, canonical = B->canonical;
    sum_take_storage(B, &tmp);
    B->n = (int64_t)total;
    B->nq = nq;
    B->sorted = sorted;
    B->canonical = canonical;
    return PLH_OK;
}

// ---- a mixed step's C jobs: their terms gathered out of the batch (job
// order kept).  cpre = exclusive scan over slots of the C jobs' term counts
// (one workgroup), its total posted to the host.
__global__ __launch_bounds__(kBlock) void k_c_sizes(const uint32_t* __restrict__ joff,
                                                    const uint8_t* __restrict__ isc, uint32_t slots,
                                                    uint64_t* __restrict__ cpre, uint64_t* mbox, uint64_t seq) {
    __shared__ uint64_t sh[kWavesPerBlock];
    uint64_t carry = 0;
    for (uint32_t c0 = 0; c0 < slots; c0 += kBlock) {
        const uint32_t t = c0 + threadIdx.x;
        const uint64_t v = (t < slots && isc[t]) ? (uint64_t)(joff[t + 1] - joff[t]) : 0ull;
        uint64_t tot;
        const uint64_t ex = block_exclusive_scan(v, tot, sh);
        if (t < slots) cpre[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        cpre[slots] = carry;
        post_write(mbox, seq, 1, &carry);
    }
}

// One thread per gathered term o: its job is the slot t with cpre[t] <= o <
// cpre[t + 1] (a search of the slots' prefix; a job that is not gathered
// has an empty range), its source joff[t] + (o - cpre[t]).
__global__ __launch_bounds__(kBlock) void k_c_gather(SumView b, int as, uint32_t slots,
                                                     const uint32_t* __restrict__ joff,
                                                     const uint64_t* __restrict__ cpre, uint64_t total,
                                                     uint64_t* __restrict__ ox, uint64_t* __restrict__ oz,
                                                     uint8_t* __restrict__ of, double* __restrict__ ov,
                                                     uint32_t* __restrict__ oj, int64_t os) {
    const uint64_t o = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= total) return;
    uint32_t lo = 0, cnt = slots;  // the last t with cpre[t] <= o
    while (cnt > 0) {
        const uint32_t half = cnt / 2, t = lo + half;
        if (cpre[t + 1] <= o) {
            lo = t + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    const uint32_t job = lo;
    if (job >= slots) return;  // never: o < cpre[slots]
    const uint64_t i = (uint64_t)joff[job] + (o - cpre[job]);
    for (int w = 0; w < as; ++w) {
        ox[(int64_t)w * os + o] = b.x[(int64_t)w * b.stride + i];
        oz[(int64_t)w * os + o] = b.z[(int64_t)w * b.stride + i];
    }
    of[o] = b.f[i];
    ov[o] = b.v[i];
    oj[o] = job;
}

// T = the terms of the batch's C jobs (isc, device), in job order, with room
// for `extra` more terms.
int gather_jobs(plh_ctx* ctx, const plh_sum* B, const uint8_t* isc, uint32_t slots, int64_t extra, plh_sum* T) {
    hipStream_t st = ctx->stream;
    const uint64_t n = (uint64_t)B->n;
    void* p;
    const size_t sj = ((size_t)slots + 1 + 63) & ~(size_t)63;
    PLH_TRY(ctx->ws.get(kWsTmpSum, sj * 4 + sj * 8 + 256, &p));
    uint32_t* joff = (uint32_t*)p;
    uint64_t* cpre = (uint64_t*)((char*)p + sj * 4);
    k_job_offsets<<<blocks_for((uint64_t)slots + 1, 256), 256, 0, st>>>(B->j, (uint32_t)n, slots, joff);
    Post post;
    PLH_TRY(post_prepare(ctx, &post));
    k_c_sizes<<<1, kBlock, 0, st>>>(joff, isc, slots, cpre, post.mbox, post.seq);
    PLH_HIP(hipGetLastError());
    uint64_t total = 0;
    PLH_TRY(post_wait(ctx, post, 1, &total));
    T->n = 0;
    PLH_TRY(sum_reserve(T, std::max<int64_t>((int64_t)total + extra, 1), 0));
    if (total) {
        k_c_gather<<<blocks_for(total, kBlock), kBlock, 0, st>>>(SumView{B->x, B->z, B->f, B->v, B->cap}, B->as, slots,
                                                                  joff, cpre, total, T->x, T->z, T->f, T->v, T->j,
                                                                  T->cap);
        PLH_HIP(hipGetLastError());
    }
    T->n = (int64_t)total;
    T->canonical = false;
    T->sorted = false;
    return PLH_OK;
}

// Append a job's observable to the incoming sum under its slot.  On failure
// it holds what it held before (its storage may have grown).
int admit(plh_server* s, const Job& j, int slot) {
    const int64_t n = j.n0;
    if (n == 0) return PLH_OK;
    plh_sum tmp;
    tmp.ctx = s->ctx;
    tmp.as = s->as;
    tmp.nq = 64 * s->as;
    tmp.has_job = true;
    PLH_TRY(sum_reserve(&tmp, n, 0));
    hipStream_t st = s->ctx->stream;
    hipError_t e = hipSuccess;
    for (int k = 0; k < s->as; ++k) {
        uint64_t* dx = tmp.x + (int64_t)k * tmp.cap;
        uint64_t* dz = tmp.z + (int64_t)k * tmp.cap;
        if (k < j.job_as) {
            if (e == hipSuccess) e = hipMemcpyAsync(dx, j.ox.data() + (size_t)k * n, (size_t)n * 8, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipMemcpyAsync(dz, j.oz.data() + (size_t)k * n, (size_t)n * 8, hipMemcpyHostToDevice, st);
        } else {
            if (e == hipSuccess) e = hipMemsetAsync(dx, 0, (size_t)n * 8, st);
            if (e == hipSuccess) e = hipMemsetAsync(dz, 0, (size_t)n * 8, st);
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(tmp.f, j.of.data(), (size_t)n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(tmp.v, j.ov.data(), (size_t)n * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        k_fill_u32<<<blocks_for((uint64_t)n, 256), 256, 0, st>>>(tmp.j, n, (uint32_t)slot);
        e = hipGetLastError();
    }
    int rc = PLH_OK;
    if (e != hipSuccess) {
        set_error(std::string("HIP: ") + hipGetErrorString(e) + " while admitting a job");
        rc = PLH_ERR_HIP;
    } else {
        tmp.n = n;
        rc = op_append(s->ctx, s->incoming, &tmp);
    }
    if (hipStreamSynchronize(st) != hipSuccess && rc == PLH_OK) rc = PLH_ERR_HIP;  // staging is host memory of j
    sum_free_storage(&tmp);
    return rc;
}

// The device part of a step: table upload, gates, one collation for every
// job, and the finished jobs' expectations (and terms); their results go
// to *res.  Nothing of the scheduler changes here; a sorted step settles the
// plan (the jobs' next ops) from the rotations the device consumed.
// The step's Clifford ops per slot, with each maximal run of consecutive H,
// S and CNOT ops on pairwise distinct qubits (CNOTs of one shift q1 - q0)
// made into one gate layer; table gates, and layers of one op, stay single
// ops.  off2/ops2 replace the plan's off/ops for the device.
void compile_layers(const sched::StepPlan& pl, std::vector<uint32_t>* off2, std::vector<uint32_t>* ops2,
                    std::vector<uint64_t>* layers) {
    const size_t S = pl.off.empty() ? 0 : pl.off.size() - 1;
    off2->assign(S + 1, 0);
    ops2->clear();
    layers->clear();
    for (size_t i = 0; i < S; ++i) {
        (*off2)[i] = (uint32_t)ops2->size();
        uint64_t L[kLayerWords] = {};
        uint64_t used[2] = {0, 0};
        int nin = 0;
        uint32_t first = 0;
        auto flush = [&]() {
            if (nin == 1) {
                ops2->push_back(first);
            } else if (nin > 1) {
                ops2->push_back((5u << 28) | (uint32_t)(layers->size() / kLayerWords));
                layers->insert(layers->end(), L, L + kLayerWords);
            }
            for (int k = 0; k < kLayerWords; ++k) L[k] = 0;
            used[0] = used[1] = 0;
            nin = 0;
        };
        for (uint32_t o = pl.off[i]; o < pl.off[i + 1]; ++o) {
            const uint32_t op = pl.ops[o];
            const uint32_t kind = op >> 28, q0 = (op >> 14) & 0x3FFFu, q1 = op & 0x3FFFu;
            const bool two = kind == 2;
            if (kind > 2 || q0 >= 128 || (two && (q1 >= 128 || q1 == q0))) {
                flush();
                ops2->push_back(op);
                continue;
            }
            const int64_t d = two ? (int64_t)q1 - (int64_t)q0 : 0;
            const bool free0 = !((used[q0 >> 6] >> (q0 & 63)) & 1ull);
            const bool free1 = !two || !((used[q1 >> 6] >> (q1 & 63)) & 1ull);
            const bool shift_ok = !two || d == (int64_t)L[6] || (int64_t)L[6] == 0;
            if (!(free0 && free1 && shift_ok && (!two || (d > -64 && d < 64)))) flush();
            if (two && (d <= -64 || d >= 64)) {  // not a layer CNOT
                ops2->push_back(op);
                continue;
            }
            used[q0 >> 6] |= 1ull << (q0 & 63);
            if (kind == 0) L[q0 >> 6] |= 1ull << (q0 & 63);
            if (kind == 1) L[2 + (q0 >> 6)] |= 1ull << (q0 & 63);
            if (two) {
                used[q1 >> 6] |= 1ull << (q1 & 63);
                L[4 + (q0 >> 6)] |= 1ull << (q0 & 63);
                L[6] = (uint64_t)d;
            }
            if (nin++ == 0) first = op;
        }
        flush();
    }
    (*off2)[S] = (uint32_t)ops2->size();
}

int step_device(plh_server* s, sched::StepPlan& pl, sched::StepKind kind, std::vector<Result>* res) {
    plh_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const int AS = s->as;
    plh_sum* B = s->batch;
    struct Part {
        const void* src;
        size_t bytes;
    };
    constexpr int kParts = 20;
    std::vector<uint8_t> alive_r(pl.alive.size(), 0);
    for (size_t i = 0; i < alive_r.size(); ++i) alive_r[i] = pl.alive[i] && !(i < pl.isc.size() && pl.isc[i]);
    std::vector<uint32_t> off2, ops2;
    std::vector<uint64_t> layers;
    compile_layers(pl, &off2, &ops2, &layers);
    const Part parts[kParts] = {{off2.data(), off2.size() * 4}, {ops2.data(), ops2.size() * 4},
                                {pl.rk.data(), pl.rk.size()},       {pl.rpsf.data(), pl.rpsf.size()},
                                {pl.rc.data(), pl.rc.size() * 8},   {pl.rs.data(), pl.rs.size() * 8},
                                {pl.rp.data(), pl.rp.size() * 8},   {pl.eps.data(), pl.eps.size() * 8},
                                {pl.maxw.data(), pl.maxw.size() * 8}, {pl.alive.data(), pl.alive.size()},
                                {pl.wn.data(), pl.wn.size()},       {pl.wk.data(), pl.wk.size()},
                                {pl.wp.data(), pl.wp.size() * 8},   {pl.wpsf.data(), pl.wpsf.size()},
                                {pl.wc.data(), pl.wc.size() * 8},   {pl.ws.data(), pl.ws.size() * 8},
                                {pl.maxt.data(), pl.maxt.size() * 8}, {pl.isc.data(), pl.isc.size()},
                                {alive_r.data(), alive_r.size()},   {layers.data(), layers.size() * 8}};
    static_assert(sched::kWindow == kJobWindow, "scheduler and engine agree on the window depth");
    size_t total = 0;
    for (const Part& p : parts) total += align256(std::max<size_t>(p.bytes, 1));
    PLH_HIP(hipStreamSynchronize(st));  // the previous step's upload has been consumed
    if (total > s->step_cap) {
        PLH_HIP(hipFree(s->d_step));
        s->d_step = nullptr;
        s->step_cap = 0;
        PLH_HIP(hipMalloc(&s->d_step, total * 2));
        s->step_cap = total * 2;
    }
    s->h_step.assign(total, 0);
    void* dptr[kParts];
    size_t o = 0;
    for (int i = 0; i < kParts; ++i) {
        if (parts[i].bytes) std::memcpy(s->h_step.data() + o, parts[i].src, parts[i].bytes);
        dptr[i] = (char*)s->d_step + o;
        o += align256(std::max<size_t>(parts[i].bytes, 1));
    }
    StepTables tb;
    tb.cl_off = (const uint32_t*)dptr[0];
    tb.cl_op = (const uint32_t*)dptr[1];
    tb.rot_kind = (const uint8_t*)dptr[2];
    tb.rot_psf = (const uint8_t*)dptr[3];
    tb.rot_c = (const double*)dptr[4];
    tb.rot_s = (const double*)dptr[5];
    tb.rot_p = (const uint64_t*)dptr[6];
    tb.eps = (double*)dptr[7];
    tb.maxw = (int64_t*)dptr[8];
    tb.alive = (uint8_t*)dptr[9];
    tb.isc = (const uint8_t*)dptr[17];
    tb.alive_r = (const uint8_t*)dptr[18];
    tb.layers = (const uint64_t*)dptr[19];
    PLH_HIP(hipMemcpyAsync(s->d_step, s->h_step.data(), total, hipMemcpyHostToDevice, st));

    CollateSpec sc;
    sc.mode = kFold | kDropZero | kMaxWeight | kCutoff;
    sc.max_weight = -1;
    sc.eps = 0.0;
    sc.job_eps = tb.eps;
    sc.job_maxw = tb.maxw;
    sc.job_alive = tb.alive;
    // gates (Cliffords and fast-path rotations in place) and one collation
    // (general rotations' new terms, fold, combine, filters, drained jobs
    // out) over a job-plane sum
    auto gates_and_collate = [&](plh_sum* X) -> int {
        // up to kIpt terms per thread, fewer while that leaves under 8192
        // workgroups (a small sum of C jobs would not fill the device)
        int ipt = kIpt;
        while (ipt > 1 && blocks_for((uint64_t)X->n, (uint64_t)kBlock * ipt) < 8192) ipt /= 2;
        const uint64_t nb = blocks_for((uint64_t)X->n, (uint64_t)kBlock * ipt);
        {
            ProfScope ps(ctx, "batch_gates", (double)X->n * (2.0 * (16.0 * AS + 1.0) + 4.0));
            if (AS == 1)
                k_batch_gates<1><<<dim3((unsigned)nb), kBlock, 0, st>>>(X->x, X->z, X->f, X->j, X->cap, X->n, tb, ipt);
            else
                k_batch_gates<2><<<dim3((unsigned)nb), kBlock, 0, st>>>(X->x, X->z, X->f, X->j, X->cap, X->n, tb, ipt);
            PLH_HIP(hipGetLastError());
        }
        if (pl.any_general) {
            JobRotTables jt{tb.rot_kind, tb.rot_p, tb.rot_psf, tb.rot_c, tb.rot_s};
            return collate_rotation_jobs(ctx, X, jt, pl.p_or.data(), sc, X);
        }
        return collate_terms(ctx, X, sc, X);
    };
    if (kind == sched::kFullStep) {
        // ---- every job: the admitted observables join the batch, then gates
        // and one collation
        if (s->incoming->n > 0) {
            PLH_TRY(op_append(ctx, B, s->incoming));
            s->incoming->n = 0;
        }
        if (B->n > 0) PLH_TRY(gates_and_collate(B));
    } else {
        // ---- C jobs (a Clifford op next, or admitted): their terms apart,
        // gates and a collation
        plh_sum* T = s->side;
        T->n = 0;
        if (pl.any_c) {
            PLH_TRY(gather_jobs(ctx, B, tb.isc, (uint32_t)s->sched->slots(), s->incoming->n, T));
            if (s->incoming->n > 0) {
                PLH_TRY(op_append(ctx, T, s->incoming));
                s->incoming->n = 0;
            }
            if (T->n > 0) PLH_TRY(gates_and_collate(T));
        }
        // ---- R jobs: each one's rotation window without a sort; the C jobs'
        // collated terms go between them (the C and drained jobs' terms in the
        // batch are dropped)
        CollateSpec sr = sc;
        sr.job_alive = tb.alive_r;
        const JobWindows jw{(const uint8_t*)dptr[10], (const uint8_t*)dptr[11], (const uint64_t*)dptr[12],
                            (const uint8_t*)dptr[13], (const double*)dptr[14], (const double*)dptr[15]};
        PLH_TRY(rotate_sorted_jobs(ctx, B, jw, s->sched->slots(), sr, s->adv_dev, T->n > 0 ? T : nullptr));
        if (!s->sched->settle(&pl, s->adv_host)) {
            set_error("plh_server_step: the device reported an out-of-range rotation count");
            return PLH_ERR_HIP;
        }
    }

    // ---- apply_truncation's trim for the jobs that have max_terms
    if (pl.any_trim && B->n > 0) PLH_TRY(trim_jobs(ctx, B, (const int64_t*)dptr[16], (uint32_t)s->sched->slots()));

    // ---- finished jobs: expectation on the job-ordered batch
    if (pl.finishing.empty()) return PLH_OK;
    const size_t F = pl.finishing.size();
    std::vector<double> he(F, 0.0);
    std::vector<int64_t> hr(2 * F, 0);
    if (B->n > 0) {
        void* p;
        PLH_TRY(ctx->ws.get(kWsTmpSum, align256(F * 4) + F * 24 + align256(F * kZseParts * 8) + 256, &p));
        uint32_t* d_jobs = (uint32_t*)p;
        double* d_e = (double*)((char*)p + align256(F * 4));
        int64_t* d_rng = (int64_t*)(d_e + F);
        double* d_part = (double*)((char*)p + align256(F * 4) + align256(F * 24));
        std::vector<uint32_t> hj(F);
        for (size_t i = 0; i < F; ++i) hj[i] = (uint32_t)pl.finishing[i];
        PLH_HIP(hipMemcpyAsync(d_jobs, hj.data(), F * 4, hipMemcpyHostToDevice, st));
        k_batch_zse_part<<<dim3((unsigned)(F * kZseParts)), kBlock, 0, st>>>(
            B->as, SumView{B->x, B->z, B->f, B->v, B->cap}, B->j, B->n, d_jobs, d_part, d_rng);
        PLH_HIP(hipGetLastError());
        k_batch_zse_final<<<dim3((unsigned)((F + kBlock - 1) / kBlock)), kBlock, 0, st>>>(d_part, (uint32_t)F, d_e);
        PLH_HIP(hipGetLastError());
        PLH_HIP(hipMemcpyAsync(he.data(), d_e, F * 8, hipMemcpyDeviceToHost, st));
        PLH_HIP(hipMemcpyAsync(hr.data(), d_rng, F * 16, hipMemcpyDeviceToHost, st));
        PLH_HIP(hipStreamSynchronize(st));
    }
    const double t = now_s(s);
    res->resize(F);
    for (size_t i = 0; i < F; ++i) {
        const int slot = pl.finishing[i];
        const Job& j = s->sched->job(slot);
        Result& r = (*res)[i];
        r.handle = j.handle;
        r.expectation = he[i];
        r.terms = hr[2 * i + 1] - hr[2 * i];
        r.latency_s = t - j.t_submit;
        if (!s->keep_terms) continue;
        const int64_t a = hr[2 * i], m = r.terms;
        if (a < 0 || m < 0 || a + m > B->n) {
            set_error("plh_server_step: a finished job's range lies outside the batch");
            return PLH_ERR_HIP;
        }
        r.x.resize((size_t)j.job_as * m);
        r.z.resize((size_t)j.job_as * m);
        r.f.resize((size_t)m);
        r.v.resize((size_t)m);
        if (m > 0) {
            for (int k = 0; k < j.job_as; ++k) {
                PLH_HIP(hipMemcpyAsync(r.x.data() + (size_t)k * m, B->x + (int64_t)k * B->cap + a, (size_t)m * 8,
                                       hipMemcpyDeviceToHost, st));
                PLH_HIP(hipMemcpyAsync(r.z.data() + (size_t)k * m, B->z + (int64_t)k * B->cap + a, (size_t)m * 8,
                                       hipMemcpyDeviceToHost, st));
            }
            PLH_HIP(hipMemcpyAsync(r.f.data(), B->f + a, (size_t)m, hipMemcpyDeviceToHost, st));
            PLH_HIP(hipMemcpyAsync(r.v.data(), B->v + a, (size_t)m * 8, hipMemcpyDeviceToHost, st));
            PLH_HIP(hipStreamSynchronize(st));
        }
    }
    return PLH_OK;
}

}  // namespace
}  // namespace plh

using namespace plh;

extern "C" {

int plh_server_create(plh_ctx* ctx, int array_size, int max_jobs, int64_t hbm_budget_bytes, int keep_terms,
                      plh_server** out) {
    static_assert(sched::kGateBase == kProgramGateBase && sched::kGates == kNumGates &&
                      sched::kFirstTwoQubit == kFirstTwoQubitGate,
                  "scheduler and engine agree on the program op kinds");
    if (!ctx || !out || max_jobs <= 0 || max_jobs > (1 << 16)) {
        set_error("plh_server_create: bad context, output or max_jobs (1..65536)");
        return PLH_ERR_ARG;
    }
    if (array_size != 1 && array_size != 2) {
        set_error("batched propagation supports 1 or 2 words per plane (up to 128 qubits)");
        return PLH_ERR_AS;
    }
    PLH_HIP(hipSetDevice(ctx->device));
    {
        uint64_t tbl[kNumGates];
        uint32_t sgn[kNumGates];
        for (int g = 0; g < kNumGates; ++g) PLH_TRY(clifford_table_of(g, &tbl[g], &sgn[g]));
        PLH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gate_tbl), tbl, sizeof(tbl), 0, hipMemcpyHostToDevice));
        PLH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gate_sgn), sgn, sizeof(sgn), 0, hipMemcpyHostToDevice));
    }
    size_t free_b = 0, total_b = 0;
    PLH_HIP(hipMemGetInfo(&free_b, &total_b));
    plh_server* s = new plh_server();
    s->ctx = ctx;
    s->as = array_size;
    s->keep_terms = keep_terms != 0;
    s->sched = new sched::Scheduler(array_size, max_jobs,
                                    hbm_budget_bytes > 0 ? hbm_budget_bytes : (int64_t)(free_b / 2));
    int rc = PLH_OK;
    if (hipHostMalloc((void**)&s->adv_host, (size_t)max_jobs, hipHostMallocCoherent | hipHostMallocMapped) !=
            hipSuccess ||
        hipHostGetDevicePointer((void**)&s->adv_dev, s->adv_host, 0) != hipSuccess) {
        set_error("plh_server_create: pinned report buffer");
        rc = PLH_ERR_HIP;
    }
    for (plh_sum** ps : {&s->batch, &s->incoming, &s->side}) {
        if (rc == PLH_OK) rc = plh_sum_create(ctx, array_size, 64 * array_size, 0, ps);
        if (rc == PLH_OK) {
            (*ps)->has_job = true;
            sum_free_storage(*ps);
            rc = sum_reserve(*ps, 4096, 0);
        }
    }
    if (rc != PLH_OK) {
        for (plh_sum* q : {s->batch, s->incoming, s->side})
            if (q) plh_sum_destroy(q);
        if (s->adv_host) (void)hipHostFree(s->adv_host);
        delete s->sched;
        delete s;
        return rc;
    }
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return PLH_OK;
}

int plh_server_destroy(plh_server* s) {
    if (!s) return PLH_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    plh_sum_destroy(s->batch);
    plh_sum_destroy(s->incoming);
    plh_sum_destroy(s->side);
    (void)hipFree(s->d_step);
    if (s->adv_host) (void)hipHostFree(s->adv_host);
    delete s->sched;
    delete s;
    return PLH_OK;
}

int plh_server_submit(plh_server* s, int64_t n, const uint64_t* x, const uint64_t* z, const uint8_t* f,
                      const double* v, int64_t num_qubits, int64_t nops, const int32_t* kind, const int64_t* q0,
                      const int64_t* q1, const uint64_t* px, const uint64_t* pz, const double* theta,
                      int64_t max_weight, double cutoff, int64_t max_terms, int64_t* handle) {
    if (!s || n < 0 || nops < 0 || !handle || (n > 0 && (!x || !z || !f || !v)) ||
        (nops > 0 && (!kind || !q0 || !q1 || !px || !pz || !theta))) {
        set_error("plh_server_submit: null pointer or negative size");
        return PLH_ERR_ARG;
    }
    if (s->failed) {
        set_error("plh_server_submit: the server failed in an earlier step (" + s->failure + ")");
        return PLH_ERR_STATE;
    }
    if (num_qubits < 1 || num_qubits > 64 * s->as) {
        set_error("plh_server_submit: job qubit count exceeds the server's array size");
        return PLH_ERR_ARG;
    }
    if (n > ((int64_t)1 << 31)) {
        set_error("plh_server_submit: observable too large for a batched job");
        return PLH_ERR_SIZE;
    }
    for (int64_t l = 0; l < nops; ++l) {
        const bool table = kind[l] >= kProgramGateBase && kind[l] < kProgramGateBase + kNumGates;
        const bool two = kind[l] == 2 || (table && kind[l] - kProgramGateBase >= kFirstTwoQubitGate);
        const bool clifford = (kind[l] >= 0 && kind[l] <= 2) || table;
        if (((kind[l] < 0 || kind[l] > 3) && !table) || (clifford && (q0[l] < 0 || q0[l] >= num_qubits)) ||
            (two && (q1[l] < 0 || q1[l] >= num_qubits || q1[l] == q0[l]))) {
            set_error("plh_server_submit: invalid op in job program");
            return PLH_ERR_ARG;
        }
    }
    Job j;
    j.job_as = (int)((num_qubits + 63) / 64);
    j.n0 = n;
    const int a = j.job_as;
    if (n > 0) {
        j.ox.assign(x, x + (size_t)a * n);
        j.oz.assign(z, z + (size_t)a * n);
        j.of.assign(f, f + n);
        j.ov.assign(v, v + n);
    }
    if (nops > 0) {
        j.kind.assign(kind, kind + nops);
        j.q0.assign(q0, q0 + nops);
        j.q1.assign(q1, q1 + nops);
        j.px.assign(px, px + (size_t)a * nops);
        j.pz.assign(pz, pz + (size_t)a * nops);
        j.theta.assign(theta, theta + nops);
    }
    if (!keys_distinct(j)) {
        set_error("plh_server_submit: observable has repeated (x, z, imag) keys; canonicalize it first");
        return PLH_ERR_ARG;
    }
    j.next = nops - 1;
    j.max_weight = max_weight;
    j.cutoff = cutoff;
    j.max_terms = max_terms;
    j.handle = s->next_handle++;
    j.t_submit = now_s(s);
    *handle = j.handle;
    s->sched->queue().push_back(std::move(j));
    return PLH_OK;
}

int plh_server_step(plh_server* s, int* active_out, int* queued_out) {
    if (!s) return PLH_ERR_ARG;
    if (s->failed) {
        set_error("plh_server_step: the server failed in an earlier step (" + s->failure + ")");
        return PLH_ERR_STATE;
    }
    sched::Scheduler& sc = *s->sched;
    plh_sum* B = s->batch;

    // ---- admission: a free slot, and the batch with the new job fits the
    // HBM budget.  The job's observable goes to the incoming sum, which the
    // step consumes (a full step appends it to the batch, a mixed step
    // collates it with the C jobs).  A job that cannot be staged stays
    // queued; the server stays usable.
    {
        int64_t in_flight = B->n + s->incoming->n;
        for (int slot; (slot = sc.admission_slot(in_flight)) >= 0;) {
            const int64_t n0 = sc.queue().front().n0;
            PLH_TRY(admit(s, sc.queue().front(), slot));
            sc.admitted(slot);
            in_flight += n0;
        }
    }
    // ---- the step's kind (Scheduler::choose_step)
    const sched::StepKind kind = sc.choose_step(B->sorted, B->n);

    // ---- the step: planned, run on the device, then committed
    sched::StepPlan pl;
    sc.plan(&pl, kind);
    std::vector<Result> res;
    const int rc = step_device(s, pl, kind, &res);
    if (rc != PLH_OK) {
        s->failed = true;
        s->failure = last_error();
        return rc;
    }
    for (Result& r : res) {
        if (s->keep_terms) {
            Result meta;
            meta.handle = r.handle;
            meta.expectation = r.expectation;
            meta.terms = r.terms;
            meta.latency_s = r.latency_s;
            s->done.push_back(std::move(meta));
            s->kept[r.handle] = std::move(r);
        } else {
            s->done.push_back(std::move(r));
        }
    }
    if (sc.commit(pl)) B->n = 0;  // nothing running: drained terms go without a collation
    ++s->steps;
    if (active_out) *active_out = sc.running();
    if (queued_out) *queued_out = (int)sc.queue().size();
    return PLH_OK;
}

int plh_server_poll(plh_server* s, int max_results, int64_t* handles, double* expectation, int64_t* terms,
                    double* latency_s, int* count) {
    if (!s || !count || max_results < 0) return PLH_ERR_ARG;
    int n = 0;
    while (n < max_results && !s->done.empty()) {
        const Result& r = s->done.front();
        if (handles) handles[n] = r.handle;
        if (expectation) expectation[n] = r.expectation;
        if (terms) terms[n] = r.terms;
        if (latency_s) latency_s[n] = r.latency_s;
        ++n;
        s->done.pop_front();
    }
    *count = n;
    return PLH_OK;
}

int plh_server_take_terms(plh_server* s, int64_t handle, uint64_t* x, uint64_t* z, uint8_t* f, double* v) {
    if (!s) return PLH_ERR_ARG;
    auto it = s->kept.find(handle);
    if (it == s->kept.end()) {
        set_error("plh_server_take_terms: no kept terms for this handle (server made without keep_terms, "
                  "job not finished, or terms already taken)");
        return PLH_ERR_ARG;
    }
    const Result& r = it->second;
    if (x && !r.x.empty()) std::memcpy(x, r.x.data(), r.x.size() * 8);
    if (z && !r.z.empty()) std::memcpy(z, r.z.data(), r.z.size() * 8);
    if (f && !r.f.empty()) std::memcpy(f, r.f.data(), r.f.size());
    if (v && !r.v.empty()) std::memcpy(v, r.v.data(), r.v.size() * 8);
    s->kept.erase(it);
    return PLH_OK;
}

int64_t plh_server_batch_terms(const plh_server* s) { return s && s->batch ? s->batch->n : -1; }

int64_t plh_server_steps(const plh_server* s) { return s ? s->steps : -1; }

}  // extern "C"
