// plh_heap.hip -- device memory of a context: a virtual address range
// backed chunk by chunk and carved into ranges, and the workspace slots
// drawn from it.  Host code only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>

#include "plh_engine.hpp"

namespace plh {

namespace {

constexpr size_t kAlign = 256;
constexpr size_t kSlotDoubleMax = (size_t)256 << 20;  // workspace slots grow x2 up to here, then by 1/8

size_t round_up(size_t v, size_t q) { return (v + q - 1) / q * q; }

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

AllocClock& alloc_clock() {
    static thread_local AllocClock c;
    return c;
}

// The reservation spans four times the device memory, so that a request
// always finds contiguous addresses however the live ranges lie.
int DeviceHeap::reserve() {
    int vmm = 0;
    PLH_HIP(hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, device_));
    if (!vmm) {
        set_error("device heap: the device does not support virtual memory management");
        return PLH_ERR_HIP;
    }
    size_t free_b = 0, total_b = 0;
    PLH_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t size = round_up(4 * total_b, kChunk);
    void* p = nullptr;
    PLH_HIP(hipMemAddressReserve(&p, size, kChunk, nullptr, 0));
    va_ = (char*)p;
    va_size_ = size;
    slot_.assign(size / kChunk, hipMemGenericAllocationHandle_t{});
    slot_live_.assign(size / kChunk, 0);
    free_[(uintptr_t)va_] = size;
    return PLH_OK;
}

size_t DeviceHeap::bytes_held() const { return (mapped_ + pool_.size()) * kChunk; }

bool DeviceHeap::holds(const void* p, size_t n) const {
    const uintptr_t a = (uintptr_t)p;
    for (const auto& kv : used_)
        if (a >= kv.first && a + n <= kv.first + kv.second) return true;
    return false;
}

void DeviceHeap::add_live(uintptr_t lo, uintptr_t hi, bool add) {
    for (size_t i = slot_of(lo); i <= slot_of(hi - 1); ++i) {
        const uintptr_t c0 = (uintptr_t)va_ + i * kChunk;
        const size_t part = std::min<uintptr_t>(hi, c0 + kChunk) - std::max<uintptr_t>(lo, c0);
        if (add)
            slot_live_[i] += part;
        else
            slot_live_[i] -= part;
    }
}

// Unmaps every mapped chunk without live bytes, except those under
// [keep_lo, keep_hi) when keep_hi > keep_lo, into the pool, after the
// device has finished the work queued on it.
void DeviceHeap::reclaim(uintptr_t keep_lo, uintptr_t keep_hi) {
    (void)hipDeviceSynchronize();
    const bool keep = keep_hi > keep_lo;
    const size_t k0 = keep ? slot_of(keep_lo) : 1, k1 = keep ? slot_of(keep_hi - 1) : 0;
    for (size_t i = 0; i < slot_.size(); ++i) {
        if (!slot_[i] || slot_live_[i] != 0 || (i >= k0 && i <= k1)) continue;
        const double t0 = now_ms();
        if (hipMemUnmap(va_ + i * kChunk, kChunk) != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        alloc_clock().free_ms += now_ms() - t0;
        alloc_clock().frees += 1;
        pool_.push_back(slot_[i]);
        slot_[i] = hipMemGenericAllocationHandle_t{};
        --mapped_;
    }
}

int DeviceHeap::map_slot(size_t i, uintptr_t lo, uintptr_t hi) {
    char* at = va_ + i * kChunk;
    hipMemAccessDesc acc = {};
    acc.location.type = hipMemLocationTypeDevice;
    acc.location.id = device_;
    acc.flags = hipMemAccessFlagsProtReadWrite;
    for (int attempt = 0; attempt < 2; ++attempt) {
        const double t0 = now_ms();
        hipMemGenericAllocationHandle_t h{};
        bool fresh = false;
        if (!pool_.empty()) {
            h = pool_.back();
            pool_.pop_back();
        } else {
            hipMemAllocationProp prop = {};
            prop.type = hipMemAllocationTypePinned;
            prop.location.type = hipMemLocationTypeDevice;
            prop.location.id = device_;
            if (hipMemCreate(&h, kChunk, &prop, 0) != hipSuccess) {
                (void)hipGetLastError();
                reclaim(lo, hi);
                continue;
            }
            fresh = true;
        }
        if (hipMemMap(at, kChunk, 0, h, 0) != hipSuccess) {
            (void)hipGetLastError();
            if (fresh)
                (void)hipMemRelease(h);
            else
                pool_.push_back(h);
            reclaim(lo, hi);
            continue;
        }
        if (hipMemSetAccess(at, kChunk, &acc, 1) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipMemUnmap(at, kChunk);
            if (fresh)
                (void)hipMemRelease(h);
            else
                pool_.push_back(h);
            reclaim(lo, hi);
            continue;
        }
        slot_[i] = h;
        ++mapped_;
        alloc_clock().malloc_ms += now_ms() - t0;
        alloc_clock().mallocs += 1;
        if (fresh) alloc_clock().malloc_bytes += (double)kChunk;
        return PLH_OK;
    }
    set_error("device allocation of " + std::to_string(hi - lo) + " bytes failed: out of device memory (context maps " +
              std::to_string(mapped_ * kChunk) + " bytes, " + std::to_string(live_) + " live)");
    return PLH_ERR_OOM;
}

int DeviceHeap::back(uintptr_t lo, uintptr_t hi) {
    for (size_t i = slot_of(lo); i <= slot_of(hi - 1); ++i)
        if (!slot_[i]) PLH_TRY(map_slot(i, lo, hi));
    return PLH_OK;
}

int DeviceHeap::alloc(size_t bytes, void** out) {
    if (!va_) PLH_TRY(reserve());
    const size_t a = round_up(bytes ? bytes : 1, kAlign);
    // the free range whose first a bytes need the fewest unmapped chunks,
    // then the smallest such range, then the lowest address
    auto best = free_.end();
    size_t best_cost = ~(size_t)0;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < a) continue;
        size_t cost = 0;
        for (size_t i = slot_of(it->first); i <= slot_of(it->first + a - 1); ++i) cost += slot_[i] ? 0 : 1;
        if (cost < best_cost || (cost == best_cost && it->second < best->second)) {
            best = it;
            best_cost = cost;
        }
    }
    if (best == free_.end()) {
        set_error("device heap: address reservation exhausted for " + std::to_string(a) + " bytes");
        return PLH_ERR_OOM;
    }
    const uintptr_t start = best->first;
    const size_t len = best->second;
    PLH_TRY(back(start, start + a));
    free_.erase(start);
    if (len > a) free_[start + a] = len - a;
    used_[start] = a;
    add_live(start, start + a, true);
    live_ += a;
    *out = (void*)start;
    return PLH_OK;
}

void DeviceHeap::free(void* p) {
    if (!p) return;
    auto u = used_.find((uintptr_t)p);
    if (u == used_.end()) return;
    const uintptr_t start = u->first;
    size_t size = u->second;
    used_.erase(u);
    add_live(start, start + size, false);
    live_ -= size;
    uintptr_t s0 = start;
    auto next = free_.find(start + size);
    if (next != free_.end()) {
        size += next->second;
        free_.erase(next);
    }
    auto prev = free_.lower_bound(start);
    if (prev != free_.begin()) {
        --prev;
        if (prev->first + prev->second == start) {
            prev->second += size;
            return;
        }
    }
    free_[s0] = size;
}

void DeviceHeap::trim() {
    if (!va_) return;
    reclaim(0, 0);
    for (auto h : pool_) (void)hipMemRelease(h);
    pool_.clear();
}

void DeviceHeap::release() {
    if (!va_) return;
    (void)hipDeviceSynchronize();
    for (size_t i = 0; i < slot_.size(); ++i) {
        if (!slot_[i]) continue;
        (void)hipMemUnmap(va_ + i * kChunk, kChunk);
        (void)hipMemRelease(slot_[i]);
    }
    for (auto h : pool_) (void)hipMemRelease(h);
    (void)hipMemAddressFree(va_, va_size_);
    va_ = nullptr;
    slot_.clear();
    slot_live_.clear();
    pool_.clear();
    free_.clear();
    used_.clear();
    live_ = mapped_ = 0;
}

// ------------------------------------------------------------ workspace

int Workspace::get(int slot, size_t bytes, void** out, bool* moved) {
    if (moved) *moved = false;
    if (slot < 0) return PLH_ERR_ARG;
    if ((size_t)slot >= bufs_.size()) bufs_.resize((size_t)slot + 1);
    Buf& b = bufs_[(size_t)slot];
    if (bytes == 0) bytes = 1;
    if (b.n < bytes) {
        // Grow geometrically so a sum that grows step by step (propagation)
        // moves each slot only a few times: x2 (at least 1 MiB) up to 256 MiB,
        // then by an eighth.  The new range is taken before the old one is
        // released, so a failed request leaves the slot as it was; only when
        // the two cannot coexist is the old one released first (its contents
        // are not kept either way).  Release is stream-ordered: later work on
        // the stream may reuse the range.
        const size_t want = bytes <= kSlotDoubleMax ? std::max<size_t>(2 * bytes, (size_t)1 << 20) : bytes + bytes / 8;
        void* np = nullptr;
        const int rc = heap_->alloc(want, &np);
        if (rc != PLH_OK) {
            if (!b.p) return rc;
            heap_->free(b.p);
            b.p = nullptr;
            b.n = 0;
            PLH_TRY(heap_->alloc(want, &np));
        } else {
            heap_->free(b.p);
        }
        b.p = np;
        b.n = want;
        if (moved) *moved = true;
    }
    *out = b.p;
    return PLH_OK;
}

void Workspace::release() {
    for (auto& b : bufs_) {
        heap_->free(b.p);
        b.p = nullptr;
        b.n = 0;
    }
}

size_t Workspace::bytes_held() const {
    size_t t = 0;
    for (auto& b : bufs_) t += b.n;
    return t;
}

}  // namespace plh
