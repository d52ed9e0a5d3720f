// A simulated device for host-only tests of the engine's memory
// bookkeeping: the virtual-memory calls of the context heap
// (paulib_amd/csrc/plh_heap.hip) over a fake address reservation, with a cap
// on physical chunks (sim::cap_chunks).  Nothing is ever dereferenced at
// these addresses.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string>

#include "../../paulib_amd/csrc/plh_engine.hpp"

// ------------------------------------------------------------ fake device
namespace sim {
constexpr uintptr_t kBase = 0x100000000000ull;
constexpr size_t kTotal = 32 * plh::DeviceHeap::kChunk;  // "device memory": 32 chunks
size_t cap_chunks = 32;
size_t created = 0;
std::set<uintptr_t> mapped;                // chunk addresses
std::map<uintptr_t, uintptr_t> map_handle;  // chunk address -> handle
std::set<uintptr_t> handles;
uintptr_t next_handle = 1;
size_t reserved = 0;
inline int fail(const char* what) {
    std::printf("FAIL: %s\n", what);
    std::exit(1);
}
}  // namespace sim

extern "C" {
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
    *v = 1;
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* f, size_t* t) {
    *f = sim::kTotal;
    *t = sim::kTotal;
    return hipSuccess;
}
hipError_t hipMemAddressReserve(void** p, size_t size, size_t, void*, unsigned long long) {
    if (sim::reserved) sim::fail("second reservation");
    sim::reserved = size;
    *p = (void*)sim::kBase;
    return hipSuccess;
}
hipError_t hipMemAddressFree(void*, size_t) {
    sim::reserved = 0;
    return hipSuccess;
}
hipError_t hipMemCreate(hipMemGenericAllocationHandle_t* h, size_t size, const hipMemAllocationProp*,
                        unsigned long long) {
    if (size != plh::DeviceHeap::kChunk) sim::fail("chunk size");
    if (sim::created >= sim::cap_chunks) return hipErrorOutOfMemory;
    ++sim::created;
    const uintptr_t id = sim::next_handle++;
    sim::handles.insert(id);
    *h = (hipMemGenericAllocationHandle_t)id;
    return hipSuccess;
}
hipError_t hipMemRelease(hipMemGenericAllocationHandle_t h) {
    if (!sim::handles.erase((uintptr_t)h)) sim::fail("release of an unknown handle");
    --sim::created;
    return hipSuccess;
}
hipError_t hipMemMap(void* p, size_t size, size_t, hipMemGenericAllocationHandle_t h, unsigned long long) {
    const uintptr_t a = (uintptr_t)p;
    if (size != plh::DeviceHeap::kChunk || (a - sim::kBase) % size) sim::fail("map shape");
    if (a < sim::kBase || a + size > sim::kBase + sim::reserved) sim::fail("map outside the reservation");
    if (sim::mapped.count(a)) sim::fail("map over a mapped chunk");
    if (!sim::handles.count((uintptr_t)h)) sim::fail("map of an unknown handle");
    for (auto& kv : sim::map_handle)
        if (kv.second == (uintptr_t)h) sim::fail("handle mapped twice");
    sim::mapped.insert(a);
    sim::map_handle[a] = (uintptr_t)h;
    return hipSuccess;
}
hipError_t hipMemSetAccess(void* p, size_t, const hipMemAccessDesc*, size_t) {
    if (!sim::mapped.count((uintptr_t)p)) sim::fail("access on an unmapped chunk");
    return hipSuccess;
}
hipError_t hipMemUnmap(void* p, size_t) {
    if (!sim::mapped.erase((uintptr_t)p)) sim::fail("unmap of an unmapped chunk");
    sim::map_handle.erase((uintptr_t)p);
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "simulated"; }
}

namespace plh {
static std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const char* last_error() { return g_err.c_str(); }
}  // namespace plh

