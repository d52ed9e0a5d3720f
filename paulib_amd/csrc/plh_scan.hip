// plh_scan.hip -- device-wide exclusive scans and small host helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>

#include "plh_block.hpp"
#include "plh_engine.hpp"

namespace plh {

namespace {

constexpr int kScanIpt = 8;
constexpr uint64_t kScanTile = (uint64_t)kBlock * kScanIpt;  // elements per block

template <class TIn>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const TIn* __restrict__ in, uint64_t n,
                                                        uint64_t* __restrict__ bsum) {
    __shared__ uint64_t sh[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanIpt; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i < n) acc += (uint64_t)in[i];
    }
    const uint64_t tot = block_sum(acc, sh);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of nb block sums in place; total out.
__global__ __launch_bounds__(kBlock) void k_scan_bsum(uint64_t* __restrict__ bsum, uint64_t nb,
                                                      uint64_t* __restrict__ total) {
    __shared__ uint64_t sh[kWavesPerBlock];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t c0 = 0; c0 < nb; c0 += kBlock) {
        const uint64_t i = c0 + threadIdx.x;
        const uint64_t v = i < nb ? bsum[i] : 0;
        uint64_t tot;
        const uint64_t ex = block_exclusive_scan(v, tot, sh);
        const uint64_t c = carry;
        if (i < nb) bsum[i] = c + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

template <class TIn>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const TIn* __restrict__ in,
                                                       uint64_t* __restrict__ out, uint64_t n,
                                                       const uint64_t* __restrict__ bsum) {
    __shared__ uint64_t sh[kWavesPerBlock];
    // Blocked: thread t owns elements [t*ipt, t*ipt+ipt) of this block.
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanIpt;
    uint64_t vals[kScanIpt];
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanIpt; ++k) {
        const uint64_t i = base + k;
        vals[k] = i < n ? (uint64_t)in[i] : 0;
        acc += vals[k];
    }
    uint64_t tot;
    uint64_t run = block_exclusive_scan(acc, tot, sh) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanIpt; ++k) {
        const uint64_t i = base + k;
        if (i < n) out[i] = run;
        run += vals[k];
    }
}

// One workgroup for short inputs: blocked chunks of kScanTile with a carry.
template <class TIn>
__global__ __launch_bounds__(kBlock) void k_scan_small(const TIn* __restrict__ in, uint64_t* __restrict__ out,
                                                       uint64_t n, uint64_t* __restrict__ total, uint64_t* mbox,
                                                       uint64_t seq) {
    __shared__ uint64_t sh[kWavesPerBlock];
    uint64_t carry = 0;
    for (uint64_t c0 = 0; c0 < n; c0 += kScanTile) {
        const uint64_t base = c0 + (uint64_t)threadIdx.x * kScanIpt;
        uint64_t vals[kScanIpt];
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < kScanIpt; ++k) {
            const uint64_t i = base + k;
            vals[k] = i < n ? (uint64_t)in[i] : 0;
            acc += vals[k];
        }
        uint64_t tot;
        uint64_t run = carry + block_exclusive_scan(acc, tot, sh);
#pragma unroll
        for (int k = 0; k < kScanIpt; ++k) {
            const uint64_t i = base + k;
            if (i < n) out[i] = run;
            run += vals[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *total = carry;
        if (mbox) post_write(mbox, seq, 1, &carry);
    }
}

constexpr uint64_t kScanSmall = 16 * kScanTile;  // one workgroup up to this many elements

// With `post`, a one-workgroup scan also posts the total to the mailbox
// (*posted set); a longer one leaves that to the caller.
template <class TIn>
int scan_impl(plh_ctx* ctx, const TIn* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total,
              const Post* post = nullptr, bool* posted = nullptr) {
    hipStream_t st = ctx->stream;
    if (n <= kScanSmall) {
        k_scan_small<TIn><<<1, kBlock, 0, st>>>(d_in, d_out, n, d_total, post ? post->mbox : nullptr,
                                                post ? post->seq : 0ull);
        PLH_HIP(hipGetLastError());
        if (posted) *posted = post != nullptr;
        return PLH_OK;
    }
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    void* p = nullptr;
    PLH_TRY(ctx->ws.get(kWsScan2, (size_t)(nb ? nb : 1) * sizeof(uint64_t), &p));
    uint64_t* bsum = (uint64_t*)p;
    if (nb) {
        k_scan_reduce<TIn><<<dim3((unsigned)nb), dim3(kBlock), 0, st>>>(d_in, n, bsum);
    }
    k_scan_bsum<<<1, kBlock, 0, st>>>(bsum, nb, d_total);
    if (nb) {
        k_scan_apply<TIn><<<dim3((unsigned)nb), dim3(kBlock), 0, st>>>(d_in, d_out, n, bsum);
    }
    PLH_HIP(hipGetLastError());
    return PLH_OK;
}

// Copies n <= 64 device words into the pinned mailbox (words 1..n), then,
// after a system-scope fence, the sequence number into word 0 (one wave).
__global__ __launch_bounds__(kWave) void k_post(const uint64_t* __restrict__ src, int n, uint64_t* mbox,
                                                uint64_t seq) {
    const int i = (int)threadIdx.x;
    if (i < n) __hip_atomic_store(&mbox[1 + i], src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    if (i == 0) __hip_atomic_store(&mbox[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace

int scan_u32_to_u64_scratch(hipStream_t st, const uint32_t* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total,
                            uint64_t* bsum) {
    if (n <= kScanSmall) {
        k_scan_small<uint32_t><<<1, kBlock, 0, st>>>(d_in, d_out, n, d_total, nullptr, 0ull);
        PLH_HIP(hipGetLastError());
        return PLH_OK;
    }
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    k_scan_reduce<uint32_t><<<dim3((unsigned)nb), dim3(kBlock), 0, st>>>(d_in, n, bsum);
    k_scan_bsum<<<1, kBlock, 0, st>>>(bsum, nb, d_total);
    k_scan_apply<uint32_t><<<dim3((unsigned)nb), dim3(kBlock), 0, st>>>(d_in, d_out, n, bsum);
    PLH_HIP(hipGetLastError());
    return PLH_OK;
}

uint64_t scan_block_sums(uint64_t n) { return n <= kScanSmall ? 0 : (n + kScanTile - 1) / kScanTile; }

int scan_u32_to_u64(plh_ctx* ctx, const uint32_t* d_in, uint64_t* d_out, uint64_t n,
                    uint64_t* d_total) {
    return scan_impl<uint32_t>(ctx, d_in, d_out, n, d_total);
}

int scan_u64(plh_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total) {
    return scan_impl<uint64_t>(ctx, d_in, d_out, n, d_total);
}

int read_u64(plh_ctx* ctx, const uint64_t* d_src, uint64_t* out) { return read_u64n(ctx, d_src, 1, out); }

int post_prepare(plh_ctx* ctx, Post* p) {
    if (!ctx->stage.p) PLH_TRY(ctx->stage.init());
    p->mbox = ctx->stage.dp;
    p->seq = ++ctx->stage.seq;
    return PLH_OK;
}

// Polls the mailbox for p.seq; after 2 ms it synchronises the stream once (a
// long-running predecessor, or an error, which the synchronisation
// reports), after which the values must be there.
int post_wait(plh_ctx* ctx, const Post& p, int n, uint64_t* out) {
    const auto t0 = std::chrono::steady_clock::now();
    bool synced = false;
    while (__atomic_load_n(&ctx->stage.p[0], __ATOMIC_ACQUIRE) != p.seq) {
        if (synced) {
            set_error("read-back: the posting kernel completed without writing the mailbox");
            return PLH_ERR_HIP;
        }
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
            PLH_HIP(hipStreamSynchronize(ctx->stream));
            synced = true;
        }
    }
    for (int i = 0; i < n; ++i) out[i] = __atomic_load_n(&ctx->stage.p[1 + i], __ATOMIC_RELAXED);
    return PLH_OK;
}

int scan_u32_to_u64_read(plh_ctx* ctx, const uint32_t* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total,
                         uint64_t* host_total) {
    Post post;
    PLH_TRY(post_prepare(ctx, &post));
    bool posted = false;
    PLH_TRY(scan_impl<uint32_t>(ctx, d_in, d_out, n, d_total, &post, &posted));
    if (posted) return post_wait(ctx, post, 1, host_total);
    return read_u64(ctx, d_total, host_total);
}

// A read-back posted by a one-wave kernel into the pinned mailbox and polled
// by the host: the host sees the values as soon as they land, instead of
// waking from a stream synchronisation after a copy.
int read_u64n(plh_ctx* ctx, const uint64_t* d_src, int n, uint64_t* out) {
    if (n < 1 || n > 64) return PLH_ERR_ARG;
    Post post;
    PLH_TRY(post_prepare(ctx, &post));
    k_post<<<1, kWave, 0, ctx->stream>>>(d_src, n, post.mbox, post.seq);
    PLH_HIP(hipGetLastError());
    return post_wait(ctx, post, n, out);
}

static hipEvent_t take_event(plh_ctx* c) {
    if (!c->prof_pool.empty()) {
        hipEvent_t e = c->prof_pool.back();
        c->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

ProfScope::ProfScope(plh_ctx* c, const char* name, double alg_bytes) : c_(c) {
    if (!c_ || !c_->prof) return;
    hipEvent_t a = take_event(c_), b = take_event(c_);
    if (!a || !b) return;
    if (hipEventRecord(a, c_->stream) != hipSuccess) return;
    idx_ = (int)c_->prof_recs.size();
    c_->prof_recs.push_back(ProfRec{name, a, b, alg_bytes});
}

ProfScope::~ProfScope() {
    if (idx_ >= 0) (void)hipEventRecord(c_->prof_recs[(size_t)idx_].b, c_->stream);
}

int HostStage::init() {
    PLH_HIP(hipHostMalloc((void**)&p, 72 * sizeof(uint64_t), hipHostMallocCoherent | hipHostMallocMapped));
    for (int i = 0; i < 72; ++i) p[i] = 0;
    seq = 0;
    PLH_HIP(hipHostGetDevicePointer((void**)&dp, p, 0));
    return PLH_OK;
}
HostStage::~HostStage() {
    if (p) (void)hipHostFree(p);
}

}  // namespace plh
