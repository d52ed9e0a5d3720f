// plh_block.hpp -- wave/workgroup primitives (wave64).
#pragma once
#include "plh_common.hpp"

namespace plh {

// Cross-lane moves through DPP (GFX9 encoding) rather than ds_bpermute:
// row_shr:k shifts within each row of 16 lanes, row_bcast:15 hands lane
// 15 of a row to the next row, row_bcast:31 hands lane 31 to rows 2 and 3.
// Lanes without a source, and rows outside the row mask, keep `old`.
namespace dpp {
constexpr int kRowShr1 = 0x111;
constexpr int kRowShr2 = 0x112;
constexpr int kRowShr4 = 0x114;
constexpr int kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;
}  // namespace dpp

template <int CTRL, int ROWS, class T>
PLH_DEV T dpp_move(T v, T old) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lane values");
    if constexpr (sizeof(T) == 4) {
        const int r = __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v), CTRL,
                                                  ROWS, 0xF, false);
        return __builtin_bit_cast(T, r);
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v), o = __builtin_bit_cast(uint64_t, old);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)o, (int)(uint32_t)u, CTRL, ROWS,
                                                                  0xF, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(o >> 32), (int)(uint32_t)(u >> 32),
                                                                  CTRL, ROWS, 0xF, false);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}

// Lane 63's value, in every lane.
template <class T>
PLH_DEV T wave_last(T v) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), kWave - 1));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, kWave - 1);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), kWave - 1);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}

// Inclusive scan of op (identity id) over the 64 lanes: a scan inside each
// row by shifts of 1, 2, 4 and 8, then rows 1 and 3 take the totals of rows
// 0 and 2, then rows 2 and 3 take lane 31's (rows 0-1).  Every lane must be
// active.  The combining order is fixed, so floating-point results are
// deterministic.
template <class T, class Op>
PLH_DEV T wave_scan_op(T v, T id, Op op) {
    v = op(v, dpp_move<dpp::kRowShr1, 0xF>(v, id));
    v = op(v, dpp_move<dpp::kRowShr2, 0xF>(v, id));
    v = op(v, dpp_move<dpp::kRowShr4, 0xF>(v, id));
    v = op(v, dpp_move<dpp::kRowShr8, 0xF>(v, id));
    v = op(v, dpp_move<dpp::kRowBcast15, 0xA>(v, id));
    v = op(v, dpp_move<dpp::kRowBcast31, 0xC>(v, id));
    return v;
}

template <class T>
PLH_DEV T wave_inclusive_scan(T v) {
    return wave_scan_op(v, T(0), [](T a, T b) { return a + b; });
}

template <class T>
PLH_DEV T wave_sum(T v) {
    return wave_last(wave_inclusive_scan(v));
}

template <class T>
PLH_DEV T wave_or(T v) {
    return wave_last(wave_scan_op(v, T(0), [](T a, T b) { return a | b; }));
}

template <class T>
PLH_DEV T wave_and(T v) {
    return wave_last(wave_scan_op(v, (T)~T(0), [](T a, T b) { return a & b; }));
}

// Exclusive scan across the NW waves of a workgroup; `sh` holds NW entries.
// Contains barriers: call from uniform control flow.
template <int NW = kWavesPerBlock, class T>
PLH_DEV T block_exclusive_scan(T v, T& total, T* sh) {
    const uint32_t lane = lane_id();
    const int w = threadIdx.x / kWave;
    const T inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) sh[w] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const T s = sh[i];
        if (i < w) pre += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return pre + inc - v;
}

template <int NW = kWavesPerBlock, class T>
PLH_DEV T block_sum(T v, T* sh) {
    T total;
    (void)block_exclusive_scan<NW>(v, total, sh);
    return total;
}

// One thread posts n values to a host-pinned coherent mailbox: words
// 1..n, then, after a system-scope fence, the sequence number in word 0
// (the host polls word 0).
PLH_DEV void post_write(uint64_t* mbox, uint64_t seq, int n, const uint64_t* v) {
    for (int i = 0; i < n; ++i) __hip_atomic_store(&mbox[1 + i], v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    __hip_atomic_store(&mbox[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace plh
