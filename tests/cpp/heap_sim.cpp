// Host-only check of the context heap's bookkeeping (paulib_amd/csrc/
// plh_heap.hip) against a simulated device: the virtual-memory calls the
// heap makes are defined here, over a fake address reservation, with a cap
// on physical chunks.  A seeded sequence of allocations and releases of
// mixed sizes runs against a shadow model; after every step:
//   - live ranges are disjoint, 256-byte aligned and inside the reservation;
//   - every chunk under a live range is mapped;
//   - the live byte count and each chunk's live count match the shadow;
//   - no more chunks exist than the cap (created = mapped + pooled).
// An allocation the cap cannot meet must fail with PLH_ERR_OOM and leave
// the heap usable.  Exit status 0 on success; the first failure is printed.
#include <cstdio>
#include <map>
#include <random>
#include <set>

#include "fake_vmm.hpp"

#include "../../paulib_amd/csrc/plh_heap.hip"

// ------------------------------------------------------------ the check
int main() {
    using plh::DeviceHeap;
    const size_t C = DeviceHeap::kChunk;
    std::mt19937_64 rng(20241203);
    for (int round = 0; round < 3; ++round) {
        DeviceHeap heap;
        heap.set_device(0);
        std::map<uintptr_t, size_t> live;  // shadow: start -> requested bytes (rounded)
        size_t ooms = 0;
        for (int step = 0; step < 4000; ++step) {
            const bool drop = !live.empty() && (rng() % 100 < 45 || live.size() > 40);
            if (drop) {
                auto it = live.begin();
                std::advance(it, (long)(rng() % live.size()));
                heap.free((void*)it->first);
                live.erase(it);
            } else {
                size_t n;
                const unsigned r = rng() % 100;
                if (r < 60)
                    n = 1 + rng() % (1 << 20);
                else if (r < 90)
                    n = 1 + rng() % (64ull << 20);
                else
                    n = 1 + rng() % (3 * C);
                void* p = nullptr;
                const int rc = heap.alloc(n, &p);
                if (rc == PLH_ERR_OOM) {
                    ++ooms;
                    continue;
                }
                if (rc != PLH_OK) sim::fail("alloc error other than OOM");
                const size_t a = (n + 255) / 256 * 256;
                const uintptr_t s = (uintptr_t)p;
                if (s % 256) sim::fail("misaligned range");
                if (s < sim::kBase || s + a > sim::kBase + sim::reserved) sim::fail("range outside the reservation");
                auto nx = live.lower_bound(s);
                if (nx != live.end() && nx->first < s + a) sim::fail("overlaps the next live range");
                if (nx != live.begin()) {
                    auto pv = std::prev(nx);
                    if (pv->first + pv->second > s) sim::fail("overlaps the previous live range");
                }
                for (uintptr_t c = (s - sim::kBase) / C * C + sim::kBase; c < s + a; c += C)
                    if (!sim::mapped.count(c)) sim::fail("live range over an unmapped chunk");
                live[s] = a;
            }
            size_t tot = 0;
            for (auto& kv : live) tot += kv.second;
            if (tot != heap.bytes_live()) sim::fail("live byte count");
            if (sim::created > sim::cap_chunks) sim::fail("more chunks than the cap");
            if (heap.bytes_held() != sim::created * C) sim::fail("held bytes != created chunks");
            if (step % 500 == 499) {
                heap.trim();
                // after trim only chunks under live ranges stay
                std::set<uintptr_t> need;
                for (auto& kv : live)
                    for (uintptr_t c = (kv.first - sim::kBase) / C * C + sim::kBase; c < kv.first + kv.second; c += C)
                        need.insert(c);
                if (need != sim::mapped) sim::fail("trim left unused chunks mapped or unmapped a used one");
                if (sim::created != need.size()) sim::fail("trim kept pooled chunks");
            }
        }
        // a request beyond the cap fails cleanly and the heap still works
        void* p = nullptr;
        if (heap.alloc((sim::cap_chunks + 1) * C, &p) != PLH_ERR_OOM) sim::fail("oversized request did not fail");
        for (auto& kv : live) heap.free((void*)kv.first);
        live.clear();
        if (heap.bytes_live() != 0) sim::fail("live bytes after releasing all");
        if (heap.alloc(sim::cap_chunks * C, &p) != PLH_OK) sim::fail("full-size request after release");
        heap.free(p);
        std::printf("round %d: %zu OOMs on the way, %zu chunks held at the end\n", round, ooms, sim::created);
        sim::cap_chunks = round == 0 ? 12 : 64;
    }
    if (sim::created != 0 || !sim::mapped.empty() || sim::reserved) sim::fail("heap destructor left memory behind");
    std::printf("ok\n");
    return 0;
}
