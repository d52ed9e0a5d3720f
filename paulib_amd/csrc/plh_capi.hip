// plh_capi.hip -- the extern "C" boundary (include/paulib_hip.h).
//
// Each entry point validates its arguments, runs stream-ordered work on the
// context's stream and returns a status code; HIP errors become
// PLH_ERR_HIP/PLH_ERR_OOM with text in plh_last_error() (the reference's
// wrapper throws std::runtime_error("CUDA: ...") instead,
// include/gpu/GpuPauliSum.h:25-30; the C++ mirror header turns these codes
// back into that exception).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <thread>
#include <cstring>
#include <string>

#include "plh_angles.hpp"
#include "plh_engine.hpp"
#include "plh_sources.hpp"

namespace plh {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

}  // namespace plh

using namespace plh;

namespace {

bool valid_as(int as) { return as >= 1 && as <= 16; }

int check_sum(const plh_sum* s) {
    if (!s || !s->ctx) {
        set_error("null sum");
        return PLH_ERR_ARG;
    }
    return PLH_OK;
}

int check_pair(const plh_sum* a, const plh_sum* b) {
    PLH_TRY(check_sum(a));
    PLH_TRY(check_sum(b));
    if (a->ctx != b->ctx) {
        set_error("operands belong to different contexts");
        return PLH_ERR_ARG;
    }
    if (a->as != b->as) {
        set_error("operands have different array sizes");
        return PLH_ERR_ARG;
    }
    return PLH_OK;
}

CollateSpec canon_spec() {
    CollateSpec c;
    c.mode = kFold | kDropZero;
    return c;
}

CollateSpec trunc_spec(int64_t max_weight, double cutoff) {
    CollateSpec c = canon_spec();
    if (max_weight >= 0) {
        c.mode |= kMaxWeight;
        c.max_weight = max_weight;
    }
    if (cutoff > 0.0) {
        c.mode |= kCutoff;
        c.eps = cutoff;
    }
    return c;
}

// One rotation of evolve + apply_truncation (the body of propagate,
// include/PauliPropagation.h:56-61).  When the sum is canonical and the
// general path appended new terms, every (x, z, imag) key holds at most one
// old and one new term, so evolve's sort_and_combine followed by
// canonicalize equals a single fold-combine pass (IEEE addition of two
// terms is commutative and sign-symmetric); the filters are per-term
// predicates of the canonical terms, so they fuse into the same pass.
int rotation_step_core(plh_sum* s, const uint64_t* px, const uint64_t* pz, uint8_t pf, double theta,
                       bool truncate, int64_t max_weight, double cutoff) {
    double c, sn;
    angle_cos_sin(theta, &c, &sn);
    const bool was_canonical = s->canonical;
    if (rotation_kind(c, sn) == 4) {
        // general path: evolve's new terms are produced inside the collation
        // (append + sort_and_combine, and with distinct input keys also the
        // truncation's canonicalize: every key then has at most two addends)
        const uint32_t psf = (uint32_t)(pf ^ 1u);
        // a sorted canonical input rotates without a sort (rotate_sorted); one
        // with distinct keys in another order (after Clifford gates) takes
        // the single fused collation
        if (truncate && s->sorted) return rotate_sorted(s->ctx, s, px, pz, psf, c, sn, trunc_spec(max_weight, cutoff));
        if (truncate && was_canonical)
            return collate_rotation(s->ctx, s, px, pz, psf, c, sn, trunc_spec(max_weight, cutoff), s);
        CollateSpec sc;
        PLH_TRY(collate_rotation(s->ctx, s, px, pz, psf, c, sn, sc, s));
        s->canonical = false;
        s->sorted = false;
        if (!truncate) return PLH_OK;
        return collate_terms(s->ctx, s, trunc_spec(max_weight, cutoff), s);
    }
    int appended = 0;
    PLH_TRY(op_rotation(s->ctx, s, px, pz, pf, c, sn, &appended));
    if (!truncate) {
        if (appended) {
            CollateSpec sc;
            PLH_TRY(collate_terms(s->ctx, s, sc, s));
            s->canonical = false;
            s->sorted = false;
        }
        return PLH_OK;
    }
    if (appended && !was_canonical) {
        CollateSpec sc;
        PLH_TRY(collate_terms(s->ctx, s, sc, s));
    }
    return collate_terms(s->ctx, s, trunc_spec(max_weight, cutoff), s);
}

// apply_truncation's last step, trim (when max_terms >= 0), follows the
// collation (include/PauliPropagation.h:35-47).
int rotation_step(plh_sum* s, const uint64_t* px, const uint64_t* pz, uint8_t pf, double theta, bool truncate,
                  int64_t max_weight, double cutoff, int64_t max_terms) {
    PLH_TRY(rotation_step_core(s, px, pz, pf, theta, truncate, max_weight, cutoff));
    if (truncate && max_terms >= 0) PLH_TRY(op_trim(s->ctx, s, max_terms));
    return PLH_OK;
}

// Rotations l, l - 1, ... (Heisenberg order: the last op first) through the
// window of rotate_sorted_window when the sum qualifies: sorted, at most two
// words per plane, and `clean` (it passed this truncation since its keys last
// changed).  rot_at(l, &px, &pz, &pf, &theta) describes op l, false when it
// is not a rotation.  The window holds the rotations from l down whose path
// is general or no change (up to 16).  *used = the rotations consumed, 0 when
// the window does not apply (the caller runs op l as before).
template <class RotAt>
int rotation_window_step(plh_sum* s, bool clean, int64_t l, RotAt rot_at, int64_t max_weight, double cutoff,
                         int64_t max_terms, int64_t* used) {
    *used = 0;
    if (!clean || !s->sorted || s->as > 2 || s->n > (int64_t)0x7FFFFFFFll) return PLH_OK;
    constexpr int K = 16;
    const uint64_t* px[K];
    const uint64_t* pz[K];
    uint8_t kind[K];
    uint32_t psf[K];
    double c[K], sn[K];
    int w = 0;
    for (int64_t t = l; t >= 0 && w < K; --t, ++w) {
        uint8_t pf = 0;
        double theta = 0.0;
        if (!rot_at(t, &px[w], &pz[w], &pf, &theta)) break;
        angle_cos_sin(theta, &c[w], &sn[w]);
        const int k = rotation_kind(c[w], sn[w]);
        if (k != 0 && k != 4) break;
        kind[w] = (uint8_t)k;
        psf[w] = (uint32_t)(pf ^ 1u);
    }
    if (w == 0) return PLH_OK;
    int consumed = 0;
    PLH_TRY(rotate_sorted_window(s->ctx, s, w, px, pz, kind, psf, c, sn, trunc_spec(max_weight, cutoff), &consumed));
    if (max_terms >= 0) PLH_TRY(op_trim(s->ctx, s, max_terms));  // a no-op unless a rotation added terms
    *used = consumed;
    return PLH_OK;
}

// Host copy of len bytes spread over up to 8 threads (the destination's
// first-touch page faults are spread with it).
void parallel_memcpy(void* dst, const void* src, size_t len) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t T = std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, len >> 20));
    if (T <= 1) {
        std::memcpy(dst, src, len);
        return;
    }
    const size_t part = (len + T - 1) / T;
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (size_t t = 1; t < T; ++t) {
        const size_t o = t * part;
        if (o >= len) break;
        th.emplace_back([=] { std::memcpy((char*)dst + o, (const char*)src + o, std::min(part, len - o)); });
    }
    std::memcpy(dst, src, std::min(part, len));
    for (auto& t : th) t.join();
}

// Device-to-host copy into pageable memory, ordered after the context's
// earlier work.  Large copies go through the pinned XferStage chunks: the DMA
// of chunk k+1 is issued before the host copies chunk k out, and a chunk's
// buffer is reused only after its host copy has finished.  Small ones are
// one direct async copy (the caller synchronises the stream).
int copy_to_host(plh_ctx* ctx, void* dst, const void* src, size_t bytes) {
    hipStream_t st = ctx->stream;
    if (bytes < XferStage::kMinStaged) {
        PLH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        return PLH_OK;
    }
    PLH_TRY(ctx->xfer.init());
    const size_t C = XferStage::kChunk;
    const size_t nch = (bytes + C - 1) / C;
    auto issue = [&](size_t k) -> int {
        const size_t off = k * C, len = std::min(C, bytes - off);
        PLH_HIP(hipMemcpyAsync(ctx->xfer.buf[k & 1], (const char*)src + off, len, hipMemcpyDeviceToHost, st));
        PLH_HIP(hipEventRecord(ctx->xfer.ev[k & 1], st));
        return PLH_OK;
    };
    PLH_TRY(issue(0));
    for (size_t k = 0; k < nch; ++k) {
        if (k + 1 < nch) PLH_TRY(issue(k + 1));  // its buffer's last chunk (k - 1) is already copied out
        PLH_HIP(hipEventSynchronize(ctx->xfer.ev[k & 1]));
        const size_t off = k * C;
        parallel_memcpy((char*)dst + off, ctx->xfer.buf[k & 1], std::min(C, bytes - off));
    }
    return PLH_OK;
}

// Host-to-device copy from pageable memory, ordered on the context's stream.
// Large copies go through the pinned chunks: the host copy of chunk k+1 into
// one chunk runs while chunk k's DMA runs from the other, and a chunk is
// refilled only after its previous DMA has finished (its event).  The caller
// synchronises the stream before the source may change.
int copy_to_device(plh_ctx* ctx, void* dst, const void* src, size_t bytes) {
    hipStream_t st = ctx->stream;
    if (bytes < XferStage::kMinStaged) {
        PLH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        return PLH_OK;
    }
    PLH_TRY(ctx->xfer.init());
    const size_t C = XferStage::kChunk;
    for (size_t k = 0, off = 0; off < bytes; ++k, off += C) {
        const size_t len = std::min(C, bytes - off);
        PLH_HIP(hipEventSynchronize(ctx->xfer.ev[k & 1]));  // the chunk's previous DMA is done
        parallel_memcpy(ctx->xfer.buf[k & 1], (const char*)src + off, len);
        PLH_HIP(hipMemcpyAsync((char*)dst + off, ctx->xfer.buf[k & 1], len, hipMemcpyHostToDevice, st));
        PLH_HIP(hipEventRecord(ctx->xfer.ev[k & 1], st));
    }
    return PLH_OK;
}

}  // namespace

namespace plh {
int XferStage::init() {
    for (int b = 0; b < 2; ++b) {
        if (!buf[b]) PLH_HIP(hipHostMalloc(&buf[b], kChunk, hipHostMallocDefault));
        if (!ev[b]) PLH_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
    }
    return PLH_OK;
}

XferStage::~XferStage() {
    for (int b = 0; b < 2; ++b) {
        if (buf[b]) (void)hipHostFree(buf[b]);
        if (ev[b]) (void)hipEventDestroy(ev[b]);
    }
}
}  // namespace plh

extern "C" {

const char* plh_last_error(void) { return last_error(); }
int plh_version(void) { return 1; }
const char* plh_build_info(void) {
    return "libpaulib_hip: gfx950 (wave64), -O3 -ffp-contract=off, no packed-fp32";
}

int plh_ctx_create(int device, void* stream, plh_ctx** out) {
    if (!out) return PLH_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        set_error("no HIP device visible");
        return PLH_ERR_NOGPU;
    }
    if (device < 0 || device >= ndev) {
        set_error("device index out of range");
        return PLH_ERR_ARG;
    }
    PLH_HIP(hipSetDevice(device));
    plh_ctx* c = new plh_ctx();
    c->device = device;
    c->heap.set_device(device);
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            set_error(std::string("HIP: ") + hipGetErrorString(e));
            return PLH_ERR_HIP;
        }
        c->own_stream = true;
    }
    *out = c;
    return PLH_OK;
}

int plh_ctx_destroy(plh_ctx* ctx) {
    if (!ctx) return PLH_OK;
    (void)hipStreamSynchronize(ctx->stream);
    for (plh_sum* s : ctx->sums) {  // sums outliving their context become empty and unusable
        sum_free_storage(s);
        s->ctx = nullptr;
        s->n = 0;
    }
    ctx->sums.clear();
    for (auto& r : ctx->prof_recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ctx->prof_pool) (void)hipEventDestroy(e);
    ctx->ws.release();
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PLH_OK;
}

int plh_ctx_synchronize(plh_ctx* ctx) {
    if (!ctx) return PLH_ERR_ARG;
    PLH_HIP(hipStreamSynchronize(ctx->stream));
    return PLH_OK;
}

void* plh_ctx_stream(plh_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int64_t plh_ctx_scratch_bytes(plh_ctx* ctx) {
    return ctx ? (int64_t)(ctx->ws.bytes_held() + ctx->heap.bytes_held() - ctx->heap.bytes_live()) : 0;
}

int plh_ctx_trim(plh_ctx* ctx) {
    if (!ctx) return PLH_ERR_ARG;
    PLH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->ws.release();
    ctx->heap.trim();
    return PLH_OK;
}

int plh_ctx_profile(plh_ctx* ctx, int enable) {
    if (!ctx) return PLH_ERR_ARG;
    PLH_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& r : ctx->prof_recs) {
        ctx->prof_pool.push_back(r.a);
        ctx->prof_pool.push_back(r.b);
    }
    ctx->prof_recs.clear();
    ctx->prof = enable != 0;
    alloc_clock() = AllocClock{};
    return PLH_OK;
}

int plh_ctx_profile_read(plh_ctx* ctx, int max_entries, char* names, double* total_ms, int64_t* launches,
                         double* alg_bytes, int* count) {
    if (!ctx || !count) return PLH_ERR_ARG;
    PLH_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<std::string> keys;
    std::vector<double> ms;
    std::vector<int64_t> cnt;
    std::vector<double> byt;
    for (auto& r : ctx->prof_recs) {
        float t = 0.f;
        PLH_HIP(hipEventElapsedTime(&t, r.a, r.b));
        size_t k = 0;
        while (k < keys.size() && keys[k] != r.name) ++k;
        if (k == keys.size()) {
            keys.push_back(r.name);
            ms.push_back(0.0);
            cnt.push_back(0);
            byt.push_back(0.0);
        }
        ms[k] += t;
        cnt[k] += 1;
        byt[k] += r.bytes;
        ctx->prof_pool.push_back(r.a);
        ctx->prof_pool.push_back(r.b);
    }
    ctx->prof_recs.clear();
    AllocClock& ac = alloc_clock();
    if (ac.mallocs) {
        keys.push_back("host:heap_map");
        ms.push_back(ac.malloc_ms);
        cnt.push_back(ac.mallocs);
        byt.push_back(ac.malloc_bytes);
    }
    if (ac.frees) {
        keys.push_back("host:heap_unmap");
        ms.push_back(ac.free_ms);
        cnt.push_back(ac.frees);
        byt.push_back(0.0);
    }
    ac = AllocClock{};
    const int n = (int)keys.size() < max_entries ? (int)keys.size() : max_entries;
    for (int i = 0; i < n; ++i) {
        if (names) {
            std::memset(names + 32 * i, 0, 32);
            std::strncpy(names + 32 * i, keys[(size_t)i].c_str(), 31);
        }
        if (total_ms) total_ms[i] = ms[(size_t)i];
        if (launches) launches[i] = cnt[(size_t)i];
        if (alg_bytes) alg_bytes[i] = byt[(size_t)i];
    }
    *count = n;
    return PLH_OK;
}

int plh_sum_create(plh_ctx* ctx, int array_size, int64_t num_qubits, int64_t capacity, plh_sum** out) {
    if (!ctx || !out) return PLH_ERR_ARG;
    *out = nullptr;
    if (!valid_as(array_size)) {
        set_error("unsupported array_size " + std::to_string(array_size) + " (supported: 1..16)");
        return PLH_ERR_AS;
    }
    if (num_qubits < 0 || num_qubits > 64 * (int64_t)array_size) {
        set_error("qubit count exceeds capacity");
        return PLH_ERR_ARG;
    }
    plh_sum* s = new plh_sum();
    s->ctx = ctx;
    s->as = array_size;
    s->nq = num_qubits;
    int rc = sum_reserve(s, capacity > 0 ? capacity : 16, 0);
    if (rc != PLH_OK) {
        delete s;
        return rc;
    }
    ctx->sums.push_back(s);
    *out = s;
    return PLH_OK;
}

int plh_sum_destroy(plh_sum* s) {
    if (!s) return PLH_OK;
    if (s->ctx) {
        auto& v = s->ctx->sums;
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] == s) {
                v[i] = v.back();
                v.pop_back();
                break;
            }
    }
    sum_free_storage(s);  // back to the context's cache, stream-ordered
    delete s;
    return PLH_OK;
}

int plh_sum_upload(plh_sum* s, int64_t n, const uint64_t* x, const uint64_t* z, const uint8_t* flags,
                   const double* val) {
    PLH_TRY(check_sum(s));
    if (n < 0 || (n > 0 && (!x || !z || !flags || !val))) return PLH_ERR_ARG;
    PLH_TRY(sum_reserve(s, n, 0));
    plh_ctx* ctx = s->ctx;
    for (int k = 0; k < s->as && n > 0; ++k) {
        PLH_TRY(copy_to_device(ctx, s->x + (int64_t)k * s->cap, x + (int64_t)k * n, (size_t)n * 8));
        PLH_TRY(copy_to_device(ctx, s->z + (int64_t)k * s->cap, z + (int64_t)k * n, (size_t)n * 8));
    }
    if (n > 0) {
        PLH_TRY(copy_to_device(ctx, s->f, flags, (size_t)n));
        PLH_TRY(copy_to_device(ctx, s->v, val, (size_t)n * 8));
    }
    PLH_HIP(hipStreamSynchronize(ctx->stream));
    s->n = n;
    s->canonical = false;
    s->sorted = false;
    return PLH_OK;
}

int plh_sum_download(const plh_sum* s, uint64_t* x, uint64_t* z, uint8_t* flags, double* val) {
    PLH_TRY(check_sum(s));
    const int64_t n = s->n;
    if (n > 0 && (!x || !z || !flags || !val)) return PLH_ERR_ARG;
    if (n == 0) return PLH_OK;
    return plh_sum_download_range(s, 0, n, x, z, flags, val);
}

int plh_sum_download_range(const plh_sum* s, int64_t offset, int64_t count, uint64_t* x, uint64_t* z,
                           uint8_t* flags, double* val) {
    PLH_TRY(check_sum(s));
    if (offset < 0 || count < 0 || offset > s->n || count > s->n - offset) {
        set_error("plh_sum_download_range: range outside the sum");
        return PLH_ERR_ARG;
    }
    if (count == 0) return PLH_OK;
    plh_ctx* ctx = s->ctx;
    for (int k = 0; k < s->as; ++k) {
        if (x)
            PLH_TRY(copy_to_host(ctx, x + (int64_t)k * count, s->x + (int64_t)k * s->cap + offset,
                                 (size_t)count * 8));
        if (z)
            PLH_TRY(copy_to_host(ctx, z + (int64_t)k * count, s->z + (int64_t)k * s->cap + offset,
                                 (size_t)count * 8));
    }
    if (flags) PLH_TRY(copy_to_host(ctx, flags, s->f + offset, (size_t)count));
    if (val) PLH_TRY(copy_to_host(ctx, val, s->v + offset, (size_t)count * 8));
    PLH_HIP(hipStreamSynchronize(ctx->stream));
    return PLH_OK;
}

int64_t plh_sum_length(const plh_sum* s) { return s ? s->n : -1; }
int64_t plh_sum_num_qubits(const plh_sum* s) { return s ? s->nq : -1; }
int plh_sum_set_num_qubits(plh_sum* s, int64_t q) {
    PLH_TRY(check_sum(s));
    if (q < 0 || q > 64 * (int64_t)s->as) return PLH_ERR_ARG;
    s->nq = q;
    return PLH_OK;
}
int plh_sum_array_size(const plh_sum* s) { return s ? s->as : -1; }

int plh_sum_device_planes(plh_sum* s, uint64_t** x, uint64_t** z, uint8_t** flags, double** val,
                          int64_t* stride) {
    PLH_TRY(check_sum(s));
    if (x) *x = s->x;
    if (z) *z = s->z;
    if (flags) *flags = s->f;
    if (val) *val = s->v;
    if (stride) *stride = s->cap;
    return PLH_OK;
}

int plh_sum_copy(const plh_sum* src, plh_sum* dst) {
    PLH_TRY(check_pair(src, dst));
    if (src == dst) return PLH_OK;
    dst->n = 0;
    PLH_TRY(op_append(dst->ctx, dst, src));
    dst->nq = src->nq;
    dst->canonical = src->canonical;
    dst->sorted = src->sorted;
    return PLH_OK;
}

int plh_launch_mul_soa(int array_size, const uint64_t* lhs_x, const uint64_t* lhs_z,
                       const uint8_t* lhs_flags, const double* lhs_val, const uint64_t* rhs_x,
                       const uint64_t* rhs_z, const uint8_t* rhs_flags, const double* rhs_val,
                       uint64_t* out_x, uint64_t* out_z, uint8_t* out_flags, double* out_val, int n,
                       int m, void* stream) {
    if (!valid_as(array_size)) {
        set_error("plh_launch_mul_soa: unsupported array_size=" + std::to_string(array_size) +
                  " (supported: 1-16, up to 1024 qubits)");
        return PLH_ERR_AS;
    }
    if (n < 0 || m < 0) return PLH_ERR_ARG;
    const int64_t total = (int64_t)n * m;
    return op_mul_raw(array_size, lhs_x, lhs_z, lhs_flags, lhs_val, n, rhs_x, rhs_z, rhs_flags, rhs_val, m,
                      out_x, out_z, out_flags, out_val, total, n, m, (hipStream_t)stream);
}

int plh_multiply(plh_sum* lhs, const plh_sum* rhs) {
    PLH_TRY(check_pair(lhs, rhs));
    const int64_t n = lhs->n, m = rhs->n;
    if (n > 0x7FFFFFFF || m > 0x7FFFFFFF) {
        set_error("plh_multiply: input exceeds 2^31-1 terms");
        return PLH_ERR_SIZE;
    }
    plh_sum out;
    out.ctx = lhs->ctx;
    out.as = lhs->as;
    out.nq = lhs->nq;
    PLH_TRY(sum_reserve(&out, n * m, 0));
    int rc;
    {
        // algorithmic bytes: the n*m output terms written, both operands read
        const double term_bytes = 16.0 * lhs->as + 9.0;
        ProfScope ps(lhs->ctx, "mul_raw", term_bytes * ((double)n * (double)m + (double)(n + m)));
        rc = op_mul_raw(lhs->as, lhs->x, lhs->z, lhs->f, lhs->v, lhs->cap, rhs->x, rhs->z, rhs->f, rhs->v,
                        rhs->cap, out.x, out.z, out.f, out.v, out.cap, n, m, lhs->ctx->stream);
    }
    if (rc != PLH_OK) {
        sum_free_storage(&out);
        return rc;
    }
    sum_take_storage(lhs, &out);
    lhs->n = n * m;
    lhs->canonical = false;
    lhs->sorted = false;
    return PLH_OK;
}

int plh_multiply_canonical(const plh_sum* lhs, const plh_sum* rhs, plh_sum* out) {
    PLH_TRY(check_pair(lhs, rhs));
    PLH_TRY(check_sum(out));
    out->nq = lhs->nq;
    PLH_TRY(collate_pairs(lhs->ctx, lhs, rhs, 0, 0, canon_spec(), out));
    out->nq = lhs->nq;
    return plh_ctx_synchronize(lhs->ctx);
}

// Products of at least this many pairs run alone in
// plh_batch_multiply_canonical_each: from there a product fills the device
// by itself and batching only adds the job word to its keys.
constexpr int64_t kSoloPairsDefault = (int64_t)1 << 22;
// The batched collation's 32-bit ids bound the pairs of one collation.
constexpr int64_t kBatchPairsMax = (int64_t)1 << 31;

int plh_batch_multiply_canonical(int njobs, const plh_sum* const* lhs, const plh_sum* const* rhs, plh_sum* out,
                                 int64_t* offsets) {
    if (njobs <= 0 || !lhs || !rhs || !offsets) {
        set_error("plh_batch_multiply_canonical: bad job list");
        return PLH_ERR_ARG;
    }
    PLH_TRY(check_sum(out));
    int64_t nq = 0;
    for (int k = 0; k < njobs; ++k) {
        PLH_TRY(check_pair(lhs[k], rhs[k]));
        if (lhs[k]->ctx != out->ctx) {
            set_error("plh_batch_multiply_canonical: operands and output must share one context");
            return PLH_ERR_ARG;
        }
        if (lhs[k]->n > 0x7FFFFFFF || rhs[k]->n > 0x7FFFFFFF) {
            set_error("plh_batch_multiply_canonical: input exceeds 2^31-1 terms");
            return PLH_ERR_SIZE;
        }
        nq = std::max<int64_t>(nq, lhs[k]->nq);
    }
    PLH_TRY(collate_products(out->ctx, njobs, lhs, rhs, canon_spec(), out, offsets));
    out->nq = nq;
    return plh_ctx_synchronize(out->ctx);
}

int plh_batch_multiply_canonical_each(int njobs, const plh_sum* const* lhs, const plh_sum* const* rhs,
                                      plh_sum* const* outs, int64_t solo_pairs) {
    if (njobs <= 0 || !lhs || !rhs || !outs || solo_pairs < 0) {
        set_error("plh_batch_multiply_canonical_each: bad job list or threshold");
        return PLH_ERR_ARG;
    }
    plh_ctx* ctx = nullptr;
    for (int k = 0; k < njobs; ++k) {
        PLH_TRY(check_pair(lhs[k], rhs[k]));
        PLH_TRY(check_sum(outs[k]));
        if (!ctx) ctx = outs[k]->ctx;
        if (lhs[k]->ctx != ctx || outs[k]->ctx != ctx) {
            set_error("plh_batch_multiply_canonical_each: operands and outputs must share one context");
            return PLH_ERR_ARG;
        }
        if (lhs[k]->n > 0x7FFFFFFF || rhs[k]->n > 0x7FFFFFFF) {
            set_error("plh_batch_multiply_canonical_each: input exceeds 2^31-1 terms");
            return PLH_ERR_SIZE;
        }
    }
    {  // every output distinct, and none an operand of any job
        std::vector<const plh_sum*> o(outs, outs + njobs), ops;
        for (int k = 0; k < njobs; ++k) {
            ops.push_back(lhs[k]);
            ops.push_back(rhs[k]);
        }
        std::sort(o.begin(), o.end());
        std::sort(ops.begin(), ops.end());
        bool bad = std::adjacent_find(o.begin(), o.end()) != o.end();
        for (const plh_sum* x : o) bad |= std::binary_search(ops.begin(), ops.end(), x);
        if (bad) {
            set_error("plh_batch_multiply_canonical_each: outputs must be distinct and must not alias an operand");
            return PLH_ERR_ARG;
        }
    }
    if (solo_pairs == 0) solo_pairs = ctx->tune.solo_pairs > 0 ? ctx->tune.solo_pairs : kSoloPairsDefault;
    const CollateSpec spec = canon_spec();
    // large products (and those over 128 qubits) alone: one collation each,
    // written straight into their output
    auto solo = [&](int k) { return lhs[k]->n * rhs[k]->n >= solo_pairs || lhs[k]->as > 2; };
    for (int k = 0; k < njobs; ++k) {
        if (!solo(k)) continue;
        PLH_TRY(collate_pairs(ctx, lhs[k], rhs[k], 0, 0, spec, outs[k]));
        outs[k]->nq = std::max(lhs[k]->nq, rhs[k]->nq);
    }
    // the rest batched: per array size, in job order, up to kBatchPairsMax pairs per collation
    plh_sum tmp;
    tmp.ctx = ctx;
    tmp.has_job = true;
    int rc = PLH_OK;
    for (int as = 1; as <= 2 && rc == PLH_OK; ++as) {
        std::vector<int> group;
        int64_t gp = 0;
        auto flush = [&]() -> int {
            if (group.empty()) return PLH_OK;
            const int g = (int)group.size();
            std::vector<const plh_sum*> L((size_t)g), R((size_t)g);
            std::vector<plh_sum*> O((size_t)g);
            for (int i = 0; i < g; ++i) {
                L[(size_t)i] = lhs[group[(size_t)i]];
                R[(size_t)i] = rhs[group[(size_t)i]];
                O[(size_t)i] = outs[group[(size_t)i]];
            }
            std::vector<int64_t> off((size_t)g + 1, 0);
            tmp.as = as;
            PLH_TRY(collate_products(ctx, g, L.data(), R.data(), spec, &tmp, off.data()));
            PLH_TRY(split_jobs(ctx, &tmp, g, off.data(), O.data()));
            for (int i = 0; i < g; ++i) O[(size_t)i]->nq = std::max(L[(size_t)i]->nq, R[(size_t)i]->nq);
            group.clear();
            gp = 0;
            return PLH_OK;
        };
        for (int k = 0; k < njobs && rc == PLH_OK; ++k) {
            if (lhs[k]->as != as || solo(k)) continue;
            const int64_t pairs = lhs[k]->n * rhs[k]->n;
            if (gp + pairs > kBatchPairsMax) rc = flush();
            group.push_back(k);
            gp += pairs;
        }
        if (rc == PLH_OK) rc = flush();
    }
    (void)hipStreamSynchronize(ctx->stream);
    sum_free_storage(&tmp);
    PLH_TRY(rc);
    return plh_ctx_synchronize(ctx);
}

int plh_commutator(const plh_sum* a, const plh_sum* b, int multiply_by_i, plh_sum* out) {
    PLH_TRY(check_pair(a, b));
    PLH_TRY(check_sum(out));
    if (out == a || out == b) return PLH_ERR_ARG;
    if (a->n > 0x7FFFFFFF || b->n > 0x7FFFFFFF) {
        set_error("plh_commutator: input exceeds 2^31-1 terms");
        return PLH_ERR_SIZE;
    }
    PLH_TRY(op_commutator_raw(a->ctx, a, b, multiply_by_i ? 1 : 0, out));
    return plh_ctx_synchronize(a->ctx);
}

int plh_commutator_canonical(const plh_sum* a, const plh_sum* b, int multiply_by_i, plh_sum* out) {
    PLH_TRY(check_pair(a, b));
    PLH_TRY(check_sum(out));
    PLH_TRY(collate_pairs(a->ctx, a, b, 1, multiply_by_i ? 1 : 0, canon_spec(), out));
    out->nq = a->nq > b->nq ? a->nq : b->nq;
    return plh_ctx_synchronize(a->ctx);
}

int plh_sort_and_combine(plh_sum* s) {
    PLH_TRY(check_sum(s));
    CollateSpec sc;
    PLH_TRY(collate_terms(s->ctx, s, sc, s));
    s->canonical = false;
    s->sorted = false;
    return plh_ctx_synchronize(s->ctx);
}

int plh_canonicalize(plh_sum* s) {
    PLH_TRY(check_sum(s));
    PLH_TRY(collate_terms(s->ctx, s, canon_spec(), s));
    return plh_ctx_synchronize(s->ctx);
}

int plh_add(plh_sum* lhs, const plh_sum* rhs) {
    PLH_TRY(check_pair(lhs, rhs));
    if (lhs == rhs) {
        plh_sum tmp;
        tmp.ctx = lhs->ctx;
        tmp.as = lhs->as;
        tmp.nq = lhs->nq;
        int rc = op_append(lhs->ctx, &tmp, rhs);
        if (rc == PLH_OK) rc = op_append(lhs->ctx, lhs, &tmp);
        (void)hipStreamSynchronize(lhs->ctx->stream);
        sum_free_storage(&tmp);
        PLH_TRY(rc);
    } else {
        PLH_TRY(op_append(lhs->ctx, lhs, rhs));
    }
    return plh_sort_and_combine(lhs);
}

int plh_sub(plh_sum* lhs, const plh_sum* rhs) {
    PLH_TRY(check_pair(lhs, rhs));
    // append_scaled(lhs, -1, rhs), then sort_and_combine (operator-=,
    // include/PauliSumCore.h:221-250): rhs is copied, its copy negated
    plh_sum tmp;
    tmp.ctx = lhs->ctx;
    tmp.as = lhs->as;
    tmp.nq = lhs->nq;
    int rc = op_append(lhs->ctx, &tmp, rhs);
    if (rc == PLH_OK) rc = op_scale(lhs->ctx, &tmp, -1.0);
    if (rc == PLH_OK) rc = op_append(lhs->ctx, lhs, &tmp);
    (void)hipStreamSynchronize(lhs->ctx->stream);
    sum_free_storage(&tmp);
    PLH_TRY(rc);
    return plh_sort_and_combine(lhs);
}

int plh_scale(plh_sum* s, double c) {
    PLH_TRY(check_sum(s));
    PLH_TRY(op_scale(s->ctx, s, c));
    return plh_ctx_synchronize(s->ctx);
}

int plh_lookup(const plh_sum* s, int array_size, const uint64_t* x, const uint64_t* z, uint8_t flags,
               double* out) {
    PLH_TRY(check_sum(s));
    if (!out || array_size != s->as || (array_size > 0 && (!x || !z))) {
        set_error("plh_lookup: null output or a pattern of another array size");
        return PLH_ERR_ARG;
    }
    return op_select_sum(s->ctx, s, 0, x, z, flags, out);
}

int plh_trace(const plh_sum* s, double* out) {
    PLH_TRY(check_sum(s));
    if (!out) return PLH_ERR_ARG;
    double sum = 0.0;
    PLH_TRY(op_select_sum(s->ctx, s, 1, nullptr, nullptr, 0, &sum));
    *out = std::ldexp(sum, (int)s->nq);
    return PLH_OK;
}

int plh_frobenius_norm2(const plh_sum* s, double* out) {
    PLH_TRY(check_sum(s));
    if (!out) return PLH_ERR_ARG;
    return op_select_sum(s->ctx, s, 2, nullptr, nullptr, 0, out);
}

int plh_zero_state_expectation(const plh_sum* s, double* out) {
    PLH_TRY(check_sum(s));
    if (!out) return PLH_ERR_ARG;
    return op_zero_state_expectation(s->ctx, s, out);
}

int plh_truncate_weight(plh_sum* s, int64_t max_weight) {
    PLH_TRY(check_sum(s));
    PLH_TRY(op_filter(s->ctx, s, kMaxWeight, max_weight, 0.0));
    return plh_ctx_synchronize(s->ctx);
}

int plh_cutoff(plh_sum* s, double eps) {
    PLH_TRY(check_sum(s));
    PLH_TRY(op_filter(s->ctx, s, kCutoff, 0, eps));
    return plh_ctx_synchronize(s->ctx);
}

int plh_apply_H(plh_sum* s, int64_t q) {
    PLH_TRY(check_sum(s));
    PLH_TRY(op_clifford(s->ctx, s, 0, q, 0));
    return plh_ctx_synchronize(s->ctx);
}

int plh_apply_S(plh_sum* s, int64_t q) {
    PLH_TRY(check_sum(s));
    PLH_TRY(op_clifford(s->ctx, s, 1, q, 0));
    return plh_ctx_synchronize(s->ctx);
}

int plh_apply_CNOT(plh_sum* s, int64_t control, int64_t target) {
    PLH_TRY(check_sum(s));
    PLH_TRY(op_clifford(s->ctx, s, 2, control, target));
    return plh_ctx_synchronize(s->ctx);
}

int plh_apply_clifford(plh_sum* s, int gate, int64_t q0, int64_t q1) {
    PLH_TRY(check_sum(s));
    PLH_TRY(op_clifford_table(s->ctx, s, gate, q0, q1));
    return plh_ctx_synchronize(s->ctx);
}

#define PLH_GATE1(NAME, ID) \
    int plh_apply_##NAME(plh_sum* s, int64_t q) { return plh_apply_clifford(s, ID, q, 0); }
#define PLH_GATE2(NAME, ID) \
    int plh_apply_##NAME(plh_sum* s, int64_t c, int64_t t) { return plh_apply_clifford(s, ID, c, t); }
PLH_GATE1(I, 0)
PLH_GATE1(X, 1)
PLH_GATE1(Y, 2)
PLH_GATE1(Z, 3)
PLH_GATE1(H_YZ, 6)
PLH_GATE1(H_XY, 7)
PLH_GATE1(sqrt_X, 8)
PLH_GATE1(sqrt_Y, 9)
PLH_GATE2(CZ, 11)
PLH_GATE2(CY, 12)
PLH_GATE2(XC_X, 13)
PLH_GATE2(XC_Z, 14)
PLH_GATE2(XC_Y, 15)
PLH_GATE2(YC_X, 16)
PLH_GATE2(YC_Z, 17)
PLH_GATE2(YC_Y, 18)
PLH_GATE2(SWAP, 19)
PLH_GATE2(iSWAP, 20)
#undef PLH_GATE1
#undef PLH_GATE2

int plh_clifford_table(int gate, uint64_t* table, uint32_t* sign) {
    if (!table || !sign) return PLH_ERR_ARG;
    const int rc = clifford_table_of(gate, table, sign);
    if (rc) set_error("unknown Clifford gate id");
    return rc;
}

// apply_RX / apply_RY / apply_RZ (include/PauliSumOperations.h:28-104):
// every term Q that anticommutes with P_q keeps cos(angle/2)*v and gains
// (iP_q) Q with sin(angle/2)*v, appended in term order, then
// sort_and_combine.  For such Q, (iP) Q = Q (-iP), so this is the fused
// rotation collation with ps = -iP (flags 3) and s = -sin(angle/2): its new
// coefficient (-s) * v is sin(angle/2) * v exactly.  There are no fast
// paths: at angle 0 the zero-valued new terms are kept, as in the reference.
static int rotation_axis(plh_sum* s, int axis, int64_t q, double angle) {
    PLH_TRY(check_sum(s));
    if (q < 0 || q >= 64 * (int64_t)s->as) {
        set_error("qubit index outside the sum's words");
        return PLH_ERR_ARG;
    }
    if (s->has_job) {
        set_error("apply_R*: batched sums are not supported");
        return PLH_ERR_ARG;
    }
    if (s->n == 0) return plh_ctx_synchronize(s->ctx);
    uint64_t px[16] = {0}, pz[16] = {0};
    const uint64_t m = 1ull << (q % 64);
    if (axis != 3) px[q / 64] = m;  // X or Y
    if (axis != 1) pz[q / 64] = m;  // Z or Y
    double ch, sh;
    angle_cos_sin(angle / 2, &ch, &sh);
    CollateSpec sc;
    PLH_TRY(collate_rotation(s->ctx, s, px, pz, 3u, ch, -sh, sc, s));
    s->canonical = false;
    s->sorted = false;
    return plh_ctx_synchronize(s->ctx);
}

int plh_apply_RX(plh_sum* s, int64_t q, double angle) { return rotation_axis(s, 1, q, angle); }
int plh_apply_RY(plh_sum* s, int64_t q, double angle) { return rotation_axis(s, 2, q, angle); }
int plh_apply_RZ(plh_sum* s, int64_t q, double angle) { return rotation_axis(s, 3, q, angle); }

int plh_elementwise_multiply(plh_sum* s, const plh_sum* rhs) {
    PLH_TRY(check_sum(s));
    PLH_TRY(check_sum(rhs));
    if (rhs->ctx != s->ctx) {
        set_error("elementwise_multiply: both sums must belong to one context");
        return PLH_ERR_ARG;
    }
    PLH_TRY(op_elementwise(s->ctx, s, rhs));
    return plh_ctx_synchronize(s->ctx);
}

int plh_evolve(plh_sum* s, int64_t nrot, const uint64_t* px, const uint64_t* pz, const uint8_t* pflags,
               const double* theta) {
    PLH_TRY(check_sum(s));
    if (nrot < 0 || (nrot > 0 && (!px || !pz || !theta))) return PLH_ERR_ARG;
    for (int64_t r = 0; r < nrot; ++r) {
        PLH_TRY(rotation_step(s, px + r * s->as, pz + r * s->as, pflags ? pflags[r] : 0, theta[r], false, -1,
                              0.0, -1));
    }
    return plh_ctx_synchronize(s->ctx);
}

int plh_apply_truncation(plh_sum* s, int64_t max_weight, double cutoff, int64_t max_terms) {
    return plh_apply_truncation_keep(s, max_weight, cutoff, max_terms, nullptr);
}

int plh_apply_truncation_keep(plh_sum* s, int64_t max_weight, double cutoff, int64_t max_terms,
                              const plh_sum* keep) {
    PLH_TRY(check_sum(s));
    if (keep) PLH_TRY(check_pair(s, keep));
    PLH_TRY(collate_terms(s->ctx, s, trunc_spec(max_weight, cutoff), s));
    if (max_terms >= 0) PLH_TRY(op_trim_keep(s->ctx, s, max_terms, keep));
    return plh_ctx_synchronize(s->ctx);
}

int plh_trim(plh_sum* s, int64_t max_terms) {
    PLH_TRY(check_sum(s));
    if (max_terms < 0) {
        set_error("plh_trim: max_terms must be >= 0");
        return PLH_ERR_ARG;
    }
    PLH_TRY(op_trim(s->ctx, s, max_terms));
    return plh_ctx_synchronize(s->ctx);
}

int plh_trim_keep(plh_sum* s, int64_t max_terms, const plh_sum* keep) {
    PLH_TRY(check_sum(s));
    if (keep) PLH_TRY(check_pair(s, keep));
    if (max_terms < 0) {
        set_error("plh_trim_keep: max_terms must be >= 0");
        return PLH_ERR_ARG;
    }
    PLH_TRY(op_trim_keep(s->ctx, s, max_terms, keep));
    return plh_ctx_synchronize(s->ctx);
}

int plh_propagate(plh_sum* s, int64_t nrot, const uint64_t* px, const uint64_t* pz, const uint8_t* pflags,
                  const double* theta, int64_t max_weight, double cutoff, int64_t max_terms) {
    PLH_TRY(check_sum(s));
    if (nrot < 0 || (nrot > 0 && (!px || !pz || !theta))) return PLH_ERR_ARG;
    auto rot_at = [&](int64_t t, const uint64_t** rx, const uint64_t** rz, uint8_t* pf, double* th) {
        *rx = px + t * s->as;
        *rz = pz + t * s->as;
        *pf = pflags ? pflags[t] : 0;
        *th = theta[t];
        return true;
    };
    bool clean = false;  // truncated since the keys last changed
    for (int64_t l = nrot - 1; l >= 0;) {
        int64_t used = 0;
        PLH_TRY(rotation_window_step(s, clean, l, rot_at, max_weight, cutoff, max_terms, &used));
        if (used == 0) {
            PLH_TRY(rotation_step(s, px + l * s->as, pz + l * s->as, pflags ? pflags[l] : 0, theta[l], true,
                                  max_weight, cutoff, max_terms));
            used = 1;
        }
        clean = true;
        l -= used;
    }
    return plh_ctx_synchronize(s->ctx);
}

int plh_run_program(plh_sum* s, int64_t nops, const int32_t* kind, const int64_t* q0, const int64_t* q1,
                    const uint64_t* px, const uint64_t* pz, const double* theta, int64_t max_weight,
                    double cutoff, int64_t max_terms) {
    PLH_TRY(check_sum(s));
    if (nops < 0 || (nops > 0 && (!kind || !q0 || !q1 || !px || !pz || !theta))) return PLH_ERR_ARG;
    auto rot_at = [&](int64_t t, const uint64_t** rx, const uint64_t** rz, uint8_t* pf, double* th) {
        if (kind[t] != 3) return false;
        *rx = px + t * s->as;
        *rz = pz + t * s->as;
        *pf = 0;
        *th = theta[t];
        return true;
    };
    bool clean = false;  // truncated since the keys last changed (Clifford ops change them)
    for (int64_t l = nops; l-- > 0;) {
        if (kind[l] == 3) {
            int64_t used = 0;
            PLH_TRY(rotation_window_step(s, clean, l, rot_at, max_weight, cutoff, max_terms, &used));
            clean = true;
            if (used > 0) {
                l -= used - 1;
                continue;
            }
        } else {
            clean = false;
            // a run of consecutive gate ops (in the order they act) in one pass
            if (s->as <= 16) {
                CliffordRun run;
                run.nops = 0;
                bool any_table = false;
                int64_t t = l;
                const int64_t nqb = 64 * (int64_t)s->as;
                for (; t >= 0 && run.nops < kCliffordRunMax && kind[t] != 3; --t) {
                    const int32_t kd = kind[t];
                    const bool table = kd >= kProgramGateBase && kd < kProgramGateBase + kNumGates;
                    if (!(kd == 0 || kd == 1 || kd == 2 || table)) break;  // reported by the single-op path below
                    const int gate = table ? kd - kProgramGateBase : -1;
                    const bool two = kd == 2 || (table && gate >= kFirstTwoQubitGate);
                    if (q0[t] < 0 || q0[t] >= nqb || (two && (q1[t] < 0 || q1[t] >= nqb)) || (two && q0[t] == q1[t]))
                        break;  // reported by the single-op path below
                    if (table && gate == 0) continue;  // apply_I
                    uint32_t k4 = kd <= 2 ? (uint32_t)kd : (two ? 4u : 3u);
                    const int o = run.nops++;
                    run.op[o] = (k4 << 28) | ((uint32_t)q0[t] << 14) | (two ? (uint32_t)q1[t] : 0u);
                    run.tbl[o] = 0;
                    run.sgn[o] = 0;
                    if (table) {
                        PLH_TRY(clifford_table_of(gate, &run.tbl[o], &run.sgn[o]));
                        any_table = true;
                    }
                }
                if (t < l) {  // ops l .. t + 1 taken (apply_I ones only skipped)
                    if (run.nops > 0) PLH_TRY(op_clifford_run(s->ctx, s, run, any_table));
                    l = t + 1;
                    continue;
                }
            }
        }
        switch (kind[l]) {
            case 0: PLH_TRY(op_clifford(s->ctx, s, 0, q0[l], 0)); break;
            case 1: PLH_TRY(op_clifford(s->ctx, s, 1, q0[l], 0)); break;
            case 2: PLH_TRY(op_clifford(s->ctx, s, 2, q0[l], q1[l])); break;
            case 3:
                PLH_TRY(rotation_step(s, px + l * s->as, pz + l * s->as, 0, theta[l], true, max_weight, cutoff,
                                      max_terms));
                break;
            default:
                if (kind[l] < kProgramGateBase || kind[l] >= kProgramGateBase + kNumGates) {
                    set_error("unknown op kind");
                    return PLH_ERR_ARG;
                }
                PLH_TRY(op_clifford_table(s->ctx, s, kind[l] - kProgramGateBase, q0[l], q1[l]));
                break;
        }
    }
    return plh_ctx_synchronize(s->ctx);
}

}  // extern "C"
