// plh_tune.hip -- per-device autotuning of the collation engine's tile caps
// and of the product-batch threshold, with a text cache.
//
// The sweep collates seeded products on 64 qubits (sites non-I with
// probability 0.25, the large-product distribution) and keeps the fastest
// cap (minimum of three timed runs after one warm-up, HIP events on the
// context's stream): the root pass's cap on a 32768 x 32768 product, then
// the later passes' cap on a 65536 x 65536 product (their tile count grows
// with the size, so that cap is picked at full size).  A cap replaces the
// built-in one only when it is at least 2% faster.  The batch
// threshold is where a product's own run time stops dominating the fixed
// cost of a call: threshold = t_call * P / (f - 1), with t_call the time of a
// tiny product call, P the solo pairs/s and f the per-pair cost of a batched
// product over a solo one (from 4 products of 4096 x 4096 pairs), rounded
// down to a power of two in [2^16, 2^26].
//
// Cache lines: "<device signature> <tile_root_max> <tile_pass_max> <solo_pairs>",
// one per device signature (architecture, CU count, memory size).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/paulib_hip.h"
#include "plh_engine.hpp"

namespace plh {
namespace {

constexpr uint32_t kRootCaps[] = {65536, 131072, 262144, 524288};
constexpr uint32_t kPassCaps[] = {16384, 32768, 65536};
constexpr uint32_t kRootDefault = 262144;  // the engine's built-in caps (plh_collate.hip)
constexpr uint32_t kPassDefault = 32768;

std::string device_signature(int device) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) {
        (void)hipGetLastError();
        return "unknown";
    }
    std::ostringstream o;
    o << p.gcnArchName << "/cu" << p.multiProcessorCount << "/mem" << (p.totalGlobalMem >> 30) << "G";
    std::string s = o.str();
    for (char& c : s)
        if (c == ' ' || c == ':') c = '_';
    return s;
}

bool load_cache(const std::string& path, const std::string& sig, Tuning& t) {
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string s;
        uint64_t a = 0, b = 0;
        int64_t c = 0;
        if ((is >> s >> a >> b >> c) && s == sig && a && b && c > 0) {
            t.tile_root_max = (uint32_t)a;
            t.tile_pass_max = (uint32_t)b;
            t.solo_pairs = c;
            return true;
        }
    }
    return false;
}

void store_cache(const std::string& path, const std::string& sig, const Tuning& t) {
    std::vector<std::string> keep;
    {
        std::ifstream f(path);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream is(line);
            std::string s;
            if ((is >> s) && s != sig) keep.push_back(line);
        }
    }
    std::ofstream f(path, std::ios::trunc);
    for (auto& l : keep) f << l << "\n";
    f << sig << " " << t.tile_root_max << " " << t.tile_pass_max << " " << t.solo_pairs << "\n";
}

// A seeded sum of n terms on 64 qubits: each site non-I with probability
// 1/4 (X, Y, Z alike), coefficients uniform on [-1, 1).
int make_sum(plh_ctx* ctx, int64_t n, uint64_t seed, plh_sum** out) {
    uint64_t r = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        r ^= r << 13;
        r ^= r >> 7;
        r ^= r << 17;
        return r;
    };
    std::vector<uint64_t> x((size_t)n), z((size_t)n);
    std::vector<uint8_t> f((size_t)n, 0);
    std::vector<double> v((size_t)n);
    for (int64_t t = 0; t < n; ++t) {
        uint64_t xw = 0, zw = 0;
        for (int q = 0; q < 64; ++q) {
            const uint64_t u = next() % 12;  // 0..8: I, 9: X, 10: Y, 11: Z
            if (u == 9 || u == 10) xw |= 1ull << q;
            if (u == 10 || u == 11) zw |= 1ull << q;
        }
        x[(size_t)t] = xw;
        z[(size_t)t] = zw;
        v[(size_t)t] = (double)(next() >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    PLH_TRY(plh_sum_create(ctx, 1, 64, n, out));
    const int rc = plh_sum_upload(*out, n, x.data(), z.data(), f.data(), v.data());
    if (rc != PLH_OK) plh_sum_destroy(*out);
    return rc;
}

double wall_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Best of three timed collated products (after one warm-up), in ms of device time.
int time_product(plh_ctx* ctx, plh_sum* a, plh_sum* b, plh_sum* out, double* best) {
    hipEvent_t e0, e1;
    PLH_HIP(hipEventCreate(&e0));
    PLH_HIP(hipEventCreate(&e1));
    int rc = PLH_OK;
    *best = 1e30;
    for (int rep = 0; rep < 4 && rc == PLH_OK; ++rep) {
        (void)hipEventRecord(e0, ctx->stream);
        rc = plh_multiply_canonical(a, b, out);
        (void)hipEventRecord(e1, ctx->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0) *best = std::min(*best, (double)ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int autotune(plh_ctx* ctx, Tuning& t) {
    plh_sum *a = nullptr, *b = nullptr, *out = nullptr, *ba = nullptr, *bb = nullptr;
    int rc = make_sum(ctx, 32768, 1, &a);
    if (rc == PLH_OK) rc = make_sum(ctx, 32768, 2, &b);
    if (rc == PLH_OK) rc = plh_sum_create(ctx, 1, 64, 0, &out);
    // root-pass tile cap on the 32768 x 32768 product; a cap other than the
    // built-in one must win by 2% (run-to-run noise is about that)
    Tuning pick = t;
    double best = 1e30, base = 1e30;
    for (uint32_t cap : kRootCaps) {
        if (rc != PLH_OK) break;
        ctx->tune.tile_root_max = cap;
        double ms = 0;
        rc = time_product(ctx, a, b, out, &ms);
        if (rc != PLH_OK) break;
        if (cap == kRootDefault) base = ms;
        if (ms < best) {
            best = ms;
            pick.tile_root_max = cap;
        }
    }
    if (best > 0.98 * base) {
        pick.tile_root_max = kRootDefault;
        best = base;
    }
    ctx->tune = pick;
    const double pairs_per_s = 32768.0 * 32768.0 / (best * 1e-3);
    // later passes' tile cap on a 65536 x 65536 product: their tile count,
    // not the root's, changes with the size, so it is picked at full size
    if (rc == PLH_OK) rc = make_sum(ctx, 65536, 5, &ba);
    if (rc == PLH_OK) rc = make_sum(ctx, 65536, 6, &bb);
    best = 1e30;
    base = 1e30;
    for (uint32_t cap : kPassCaps) {
        if (rc != PLH_OK) break;
        ctx->tune.tile_pass_max = cap;
        double ms = 0;
        rc = time_product(ctx, ba, bb, out, &ms);
        if (rc != PLH_OK) break;
        if (cap == kPassDefault) base = ms;
        if (ms < best) {
            best = ms;
            pick.tile_pass_max = cap;
        }
    }
    if (best > 0.98 * base) pick.tile_pass_max = kPassDefault;
    for (plh_sum* p : {ba, bb})
        if (p) plh_sum_destroy(p);
    ctx->tune = pick;
    // batch threshold
    double t_call = 1e30, f = 1.35;
    plh_sum *sa = nullptr, *sb = nullptr;
    if (rc == PLH_OK) rc = make_sum(ctx, 256, 3, &sa);
    if (rc == PLH_OK) rc = make_sum(ctx, 256, 4, &sb);
    for (int rep = 0; rep < 6 && rc == PLH_OK; ++rep) {
        const double w0 = wall_ms();
        rc = plh_multiply_canonical(sa, sb, out);
        if (rep > 0) t_call = std::min(t_call, wall_ms() - w0);
    }
    plh_sum* mid[8] = {nullptr};
    for (int k = 0; k < 8 && rc == PLH_OK; ++k) rc = make_sum(ctx, 4096, 10 + (uint64_t)k, &mid[k]);
    if (rc == PLH_OK) {
        double solo = 1e30, batch = 1e30;
        std::vector<int64_t> off(5);
        plh_sum* bout = nullptr;
        rc = plh_sum_create(ctx, 1, 64, 0, &bout);
        const plh_sum* L[4] = {mid[0], mid[1], mid[2], mid[3]};
        const plh_sum* R[4] = {mid[4], mid[5], mid[6], mid[7]};
        for (int rep = 0; rep < 4 && rc == PLH_OK; ++rep) {
            double w0 = wall_ms();
            for (int k = 0; k < 4 && rc == PLH_OK; ++k) rc = plh_multiply_canonical(L[k], R[k], out);
            const double ts = wall_ms() - w0;
            w0 = wall_ms();
            if (rc == PLH_OK) rc = plh_batch_multiply_canonical(4, L, R, bout, off.data());
            const double tb = wall_ms() - w0;
            if (rep > 0) {
                solo = std::min(solo, ts - 4 * t_call);  // the products' own time
                batch = std::min(batch, tb - t_call);
            }
        }
        if (bout) plh_sum_destroy(bout);
        if (rc == PLH_OK && solo > 0 && batch > 0) f = batch / solo;
    }
    for (plh_sum* p : mid)
        if (p) plh_sum_destroy(p);
    for (plh_sum* p : {a, b, out, sa, sb})
        if (p) plh_sum_destroy(p);
    if (rc != PLH_OK) return rc;
    const double cross = t_call * 1e-3 * pairs_per_s / std::max(f - 1.0, 0.05);
    int64_t solo_pairs = (int64_t)1 << 16;
    while (solo_pairs * 2 <= (int64_t)cross && solo_pairs < ((int64_t)1 << 26)) solo_pairs *= 2;
    ctx->tune.solo_pairs = solo_pairs;
    t = ctx->tune;
    return PLH_OK;
}

}  // namespace
}  // namespace plh

using namespace plh;

extern "C" {

int plh_ctx_autotune(plh_ctx* ctx, const char* cache_path, int force) {
    if (!ctx) return PLH_ERR_ARG;
    const std::string sig = device_signature(ctx->device);
    Tuning t;
    if (cache_path && !force && load_cache(cache_path, sig, t)) {
        ctx->tune = t;
        return PLH_OK;
    }
    const Tuning saved = ctx->tune;
    const int rc = autotune(ctx, t);
    if (rc != PLH_OK) {
        ctx->tune = saved;
        return rc;
    }
    if (cache_path) store_cache(cache_path, sig, t);
    (void)plh_ctx_trim(ctx);  // the sweep's large scratch is not the caller's working set
    return PLH_OK;
}

int plh_ctx_tuning(const plh_ctx* ctx, int64_t* tile_root_max, int64_t* tile_pass_max, int64_t* solo_pairs) {
    if (!ctx) return PLH_ERR_ARG;
    if (tile_root_max) *tile_root_max = ctx->tune.tile_root_max;
    if (tile_pass_max) *tile_pass_max = ctx->tune.tile_pass_max;
    if (solo_pairs) *solo_pairs = ctx->tune.solo_pairs;
    return PLH_OK;
}

}  // extern "C"
