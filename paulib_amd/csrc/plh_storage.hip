// plh_storage.hip -- storage of device-resident sums: one heap block per
// sum holding its planes, re-laid out when the array size or the job plane
// changes, grown with or without keeping terms, handed between sums, and
// appended to.  Host code only (stream-ordered copies on the context's
// stream), so the CPU replay of the server's bookkeeping
// (tests/cpp/sched_sim.cpp) compiles it against a simulated device.
#include <hip/hip_runtime.h>

#include "plh_engine.hpp"

namespace plh {

void sum_free_storage(plh_sum* s) {
    if (s->block && s->ctx) s->ctx->heap.free(s->block);
    s->block = nullptr;
    s->x = s->z = nullptr;
    s->f = nullptr;
    s->v = nullptr;
    s->j = nullptr;
    s->cap = 0;
}

void sum_take_storage(plh_sum* dst, plh_sum* src) {
    sum_free_storage(dst);
    dst->block = src->block;
    dst->block_as = src->block_as;
    dst->block_job = src->block_job;
    dst->x = src->x;
    dst->z = src->z;
    dst->f = src->f;
    dst->v = src->v;
    dst->j = src->j;
    dst->cap = src->cap;
    src->block = nullptr;
    src->x = src->z = nullptr;
    src->f = nullptr;
    src->v = nullptr;
    src->j = nullptr;
    src->cap = 0;
}

// Layout of a sum's block: x planes, z planes, values, job ids (batched
// sums), flags.
static size_t sum_bytes(int as, int64_t cap, bool job) {
    return (size_t)cap * (16 * (size_t)as + 9 + (job ? 4 : 0));
}

int sum_reserve(plh_sum* s, int64_t cap, int64_t keep) {
    // the block's planes were laid out for block_as words and a job plane
    // or not: a sum whose array size or job plane changed needs a new block
    const bool same_layout = s->block_as == s->as && s->block_job == s->has_job;
    if (cap <= s->cap && s->block && same_layout) return PLH_OK;
    if (keep > 0 && s->block && !same_layout) {
        set_error("sum_reserve: cannot keep terms across an array-size or job-plane change");
        return PLH_ERR_ARG;
    }
    int64_t nc = cap < 16 ? 16 : cap;
    if (s->cap > 0 && nc < s->cap + s->cap / 2 && keep > 0) nc = s->cap + s->cap / 2;  // amortise appends
    // nothing to keep: give the old block back first, so the two need not
    // coexist (a sum near the HBM size can still be re-reserved); the sum is
    // empty from here on, so a failed allocation leaves it empty, not with
    // terms over released planes
    if (keep <= 0) {
        sum_free_storage(s);
        s->n = 0;
        s->canonical = false;
        s->sorted = false;
    }
    if (!s->ctx) {
        set_error("the sum's context has been destroyed");
        return PLH_ERR_ARG;
    }
    void* blk = nullptr;
    PLH_TRY(s->ctx->heap.alloc(sum_bytes(s->as, nc, s->has_job), &blk));
    char* b = (char*)blk;
    uint64_t* nx = (uint64_t*)b;
    uint64_t* nz = nx + (size_t)s->as * nc;
    double* nv = (double*)(nz + (size_t)s->as * nc);
    uint32_t* nj = s->has_job ? (uint32_t*)(nv + nc) : nullptr;
    uint8_t* nf = s->has_job ? (uint8_t*)(nj + nc) : (uint8_t*)(nv + nc);
    hipStream_t st = s->ctx ? s->ctx->stream : nullptr;
    if (keep > 0 && s->block) {
        for (int k = 0; k < s->as; ++k) {
            PLH_HIP(hipMemcpyAsync(nx + (int64_t)k * nc, s->x + (int64_t)k * s->cap, (size_t)keep * 8,
                                   hipMemcpyDeviceToDevice, st));
            PLH_HIP(hipMemcpyAsync(nz + (int64_t)k * nc, s->z + (int64_t)k * s->cap, (size_t)keep * 8,
                                   hipMemcpyDeviceToDevice, st));
        }
        PLH_HIP(hipMemcpyAsync(nf, s->f, (size_t)keep, hipMemcpyDeviceToDevice, st));
        PLH_HIP(hipMemcpyAsync(nv, s->v, (size_t)keep * 8, hipMemcpyDeviceToDevice, st));
        if (s->has_job && s->j) PLH_HIP(hipMemcpyAsync(nj, s->j, (size_t)keep * 4, hipMemcpyDeviceToDevice, st));
    }
    sum_free_storage(s);  // stream-ordered reuse: no synchronisation needed
    s->block = blk;
    s->block_as = s->as;
    s->block_job = s->has_job;
    s->j = nj;
    s->x = nx;
    s->z = nz;
    s->f = nf;
    s->v = nv;
    s->cap = nc;
    return PLH_OK;
}


int op_append(plh_ctx* ctx, plh_sum* dst, const plh_sum* src) {
    hipStream_t st = ctx->stream;
    if (src->n == 0) return PLH_OK;
    const int64_t n0 = dst->n;
    PLH_TRY(sum_reserve(dst, n0 + src->n, n0));
    for (int k = 0; k < dst->as; ++k) {
        PLH_HIP(hipMemcpyAsync(dst->x + (int64_t)k * dst->cap + n0, src->x + (int64_t)k * src->cap,
                               (size_t)src->n * 8, hipMemcpyDeviceToDevice, st));
        PLH_HIP(hipMemcpyAsync(dst->z + (int64_t)k * dst->cap + n0, src->z + (int64_t)k * src->cap,
                               (size_t)src->n * 8, hipMemcpyDeviceToDevice, st));
    }
    PLH_HIP(hipMemcpyAsync(dst->f + n0, src->f, (size_t)src->n, hipMemcpyDeviceToDevice, st));
    PLH_HIP(hipMemcpyAsync(dst->v + n0, src->v, (size_t)src->n * 8, hipMemcpyDeviceToDevice, st));
    if (dst->has_job && src->has_job)
        PLH_HIP(hipMemcpyAsync(dst->j + n0, src->j, (size_t)src->n * 4, hipMemcpyDeviceToDevice, st));
    dst->n = n0 + src->n;
    dst->canonical = false;
    dst->sorted = false;
    return PLH_OK;
}

}  // namespace plh
