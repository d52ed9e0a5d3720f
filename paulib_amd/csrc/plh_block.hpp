// plh_block.hpp -- wave/workgroup primitives (wave64).
#pragma once
#include "plh_common.hpp"

namespace plh {

template <class T>
PLH_DEV T wave_inclusive_scan(T v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const T t = __shfl_up(v, o, kWave);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

template <class T>
PLH_DEV T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
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
