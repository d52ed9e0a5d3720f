// plh_common.hpp -- shared device helpers for the PauliSum engine (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define PLH_DEV __device__ __forceinline__
#define PLH_HD __host__ __device__ __forceinline__

namespace plh {

constexpr int kWave = 64;
constexpr int kBlock = 256;            // threads per workgroup everywhere
constexpr int kWavesPerBlock = kBlock / kWave;

PLH_DEV uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
PLH_DEV uint64_t lanemask_lt() {
    const uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}
PLH_DEV uint64_t ballot(bool p) { return __ballot(p); }
PLH_DEV int popc(uint64_t w) { return __popcll(w); }

// Division of a 32-bit numerator by a runtime-invariant divisor d >= 1
// (Granlund-Montgomery round-up method): q = (umulhi(n, m) + n) >> s.
struct FastDiv {
    uint32_t d, m, s;
};
inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{d, 0, 0};
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.s = l;
    f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    return f;
}
PLH_DEV uint32_t fdiv(uint32_t n, const FastDiv& f) {
    const uint32_t hi = __umulhi(n, f.m);
    return (uint32_t)(((uint64_t)hi + n) >> f.s);
}

// Symplectic pair product of one word: updates the bit-sliced mod-4 phase
// counters (c1 = low bit per site, c2 = high bit), as in
// PauliString::multiply, include/PauliStringCore.h:149-174.
PLH_DEV void phase_word(uint64_t ax, uint64_t az, uint64_t bx, uint64_t bz, uint64_t& c1,
                        uint64_t& c2) {
    const uint64_t xbz = ax & bz;
    const uint64_t anti = (bx & az) ^ xbz;
    c2 ^= (c1 ^ ax ^ az ^ bx ^ bz ^ xbz) & anti;
    c1 ^= anti;
}
PLH_DEV uint32_t phase_of(uint32_t fa, uint32_t fb, uint64_t c1, uint64_t c2) {
    return (fa + fb + (uint32_t)popc(c1) + 2u * ((uint32_t)popc(c2) & 1u)) & 3u;
}

}  // namespace plh
