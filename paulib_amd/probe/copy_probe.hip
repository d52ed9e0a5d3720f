// copy_probe.hip -- a plain device-to-device copy of a known size.
//
// Run under rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE): the counter read
// for k_copy divided by the bytes it moves (MiB * 1024 KiB each way) is the
// scale that turns the same counters of the engine's kernels into HBM bytes.
// Also prints the copy's own bandwidth, the attainable HBM rate on this card.
//
// usage: copy_probe [MiB=1024] [reps=5]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                          \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));         \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_fill(uint4* __restrict__ a, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = make_uint4((unsigned)i, (unsigned)(i >> 32), 0x5a5a5a5au, 1u);
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    if (mib == 0 || mib > 65536 || reps <= 0 || reps > 1000) {
        std::fprintf(stderr, "usage: copy_probe [MiB in 1..65536] [reps in 1..1000]\n");
        return 2;
    }
    const size_t bytes = mib << 20, n = bytes / sizeof(uint4);
    const unsigned grid = (unsigned)((n + 255) / 256);
    uint4 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    k_fill<<<grid, 256>>>(a, n);
    CK(hipGetLastError());
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    k_copy<<<grid, 256>>>(a, b, n);  // warm-up
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) k_copy<<<grid, 256>>>(a, b, n);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    uint4 probe{};
    CK(hipMemcpy(&probe, b + (n - 1), sizeof(uint4), hipMemcpyDeviceToHost));
    const bool ok = probe.x == (unsigned)(n - 1) && probe.z == 0x5a5a5a5au;
    std::printf("copy %zu MiB x %d: %.3f ms/copy, %.1f GB/s (read + write), KiB read/written per copy = %zu, check %s\n",
                mib, reps, ms / reps, 2.0 * (double)bytes / (ms / reps * 1e-3) / 1e9, bytes >> 10,
                ok ? "ok" : "FAILED");
    CK(hipFree(a));
    CK(hipFree(b));
    return ok ? 0 : 1;
}
