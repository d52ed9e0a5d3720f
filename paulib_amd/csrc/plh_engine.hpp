// plh_engine.hpp -- host-side engine state shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/paulib_hip.h"

namespace plh {

// Thread-local last-error text behind plh_last_error().
void set_error(const std::string& msg);
const char* last_error();

#define PLH_HIP(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            ::plh::set_error(std::string("HIP: ") + hipGetErrorString(_e) + " at " + \
                             __FILE__ + ":" + std::to_string(__LINE__));           \
            return _e == hipErrorOutOfMemory ? PLH_ERR_OOM : PLH_ERR_HIP;          \
        }                                                                          \
    } while (0)

#define PLH_TRY(expr)            \
    do {                         \
        int _rc = (expr);        \
        if (_rc != PLH_OK) return _rc; \
    } while (0)

// Host time spent mapping and unmapping device memory (and the bytes of
// physical memory created), reported with the kernel profile.
struct AllocClock {
    double malloc_ms = 0.0, free_ms = 0.0, malloc_bytes = 0.0;
    int64_t mallocs = 0, frees = 0;
};
AllocClock& alloc_clock();

// Device memory of a context: one reserved virtual address range, backed
// chunk by chunk (2 GiB) with physical memory from hipMemCreate, and
// carved into byte ranges.  A chunk's physical memory is cleared once, when
// it is first created: new physical memory on this device costs about
// 1 s per 60 GB, so a chunk whose ranges are all released stays mapped for
// later requests, and when a request needs backing that is not there,
// unused chunks are unmapped and mapped again where the request lies (a
// remap keeps the memory and costs well under a millisecond).  Ranges are
// placed where they need the fewest new chunks (then best fit), split on
// allocation and merged with free neighbours on release, so the physical
// memory held is the live bytes plus partly used chunks, whatever the
// order of requests.  Every kernel of a context runs on its one stream, so
// a range released by one call is handed to the next without a
// synchronisation; only unmapping waits for the device.
class DeviceHeap {
   public:
    ~DeviceHeap() { release(); }
    void set_device(int device) { device_ = device; }
    int alloc(size_t bytes, void** out);
    void free(void* p);
    // Unmaps the chunks without live ranges and gives their physical
    // memory back to the device (synchronises).
    void trim();
    size_t bytes_held() const;
    size_t bytes_live() const { return live_; }
    // Whether [p, p + n) lies inside one live range (a check for host-side
    // bookkeeping tests: device work may only touch live ranges).
    bool holds(const void* p, size_t n) const;
    static constexpr size_t kChunk = (size_t)2 << 30;

   private:
    int reserve();
    int back(uintptr_t lo, uintptr_t hi);
    int map_slot(size_t i, uintptr_t lo, uintptr_t hi);
    void reclaim(uintptr_t keep_lo, uintptr_t keep_hi);
    void release();
    size_t slot_of(uintptr_t a) const { return (size_t)((a - (uintptr_t)va_) / kChunk); }
    void add_live(uintptr_t lo, uintptr_t hi, bool add);
    int device_ = 0;
    char* va_ = nullptr;
    size_t va_size_ = 0;
    std::vector<hipMemGenericAllocationHandle_t> slot_;  // handle mapped at chunk i, or null
    std::vector<size_t> slot_live_;                      // live bytes inside chunk i
    std::vector<hipMemGenericAllocationHandle_t> pool_;  // created, not mapped
    std::map<uintptr_t, size_t> free_;                   // free address ranges: start -> size
    std::unordered_map<uintptr_t, size_t> used_;         // live ranges: start -> size
    size_t live_ = 0, mapped_ = 0;
};

// Grow-only device scratch owned by a context, drawn from its heap.  Slots
// are independent and stay until plh_ctx_trim, so a repeated call of one
// shape allocates nothing; get() keeps the old contents only when no
// reallocation happens.
class Workspace {
   public:
    explicit Workspace(DeviceHeap* heap) : heap_(heap) {}
    ~Workspace() { release(); }
    // moved (optional): set when the slot was (re)allocated by this call, so
    // its contents are not what the last caller left there
    int get(int slot, size_t bytes, void** out, bool* moved = nullptr);
    void release();
    size_t bytes_held() const;

   private:
    struct Buf {
        void* p = nullptr;
        size_t n = 0;
    };
    DeviceHeap* heap_;
    std::vector<Buf> bufs_;
};

// Small pinned host staging buffer for count read-backs.
// Host-pinned, GPU-coherent mailbox for small read-backs: word 0 is a
// sequence number, words 1..64 the values (written by a kernel on the
// context's stream, polled by the host).
struct HostStage {
    uint64_t* p = nullptr;   // host address
    uint64_t* dp = nullptr;  // device address of the same memory
    uint64_t seq = 0;
    int init();
    ~HostStage();
};

// Pinned staging for large device-to-host copies into pageable memory (a
// caller's vectors): two host-pinned chunks, so the DMA of chunk k+1 runs
// while the host copies chunk k out, spread over a few threads.
struct XferStage {
    static constexpr size_t kChunk = 64ull << 20;
    static constexpr size_t kMinStaged = 1ull << 20;  // smaller copies go direct
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int init();
    ~XferStage();
};

}  // namespace plh

// Device-resident PauliSum<N,double>: plane-major SoA, x[k*cap + t].
struct plh_sum {
    plh_ctx* ctx = nullptr;
    void* block = nullptr;  // one allocation holding x, z, v, f
    int as = 1;
    int64_t nq = 64;
    int64_t n = 0;
    int64_t cap = 0;
    uint64_t* x = nullptr;
    uint64_t* z = nullptr;
    uint8_t* f = nullptr;
    double* v = nullptr;
    uint32_t* j = nullptr;   // job plane of a batched sum (has_job), else null
    bool has_job = false;
    int block_as = 0;        // layout of the block: words per plane and job plane
    bool block_job = false;
    bool canonical = false;  // (x, z, imag) keys distinct (the fused rotation collation relies on it)
    bool sorted = false;     // sorted by (x, z, imag) with distinct keys and no zeros, as a folding
                             // collation leaves it (the sort-free rotation relies on it)
};

namespace plh {
// Kernel timing with HIP events on the context's stream (enabled by
// plh_ctx_profile); records are aggregated per kernel name on read.
struct ProfRec {
    const char* name;
    hipEvent_t a, b;
    double bytes;  // algorithmic HBM bytes of the bracketed launch(es)
};
}  // namespace plh

namespace plh {
// Engine parameters picked per device by plh_ctx_autotune (0 = built-in
// default): the tile caps of the root pass and of the later radix passes
// (items per workgroup), and the batch threshold of
// plh_batch_multiply_canonical_each.
struct Tuning {
    uint32_t tile_root_max = 0;
    uint32_t tile_pass_max = 0;
    int64_t solo_pairs = 0;
};
}  // namespace plh

struct plh_ctx {
    int device = 0;
    plh::Tuning tune;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    plh::DeviceHeap heap;  // declared before ws: ws returns its ranges on destruction
    plh::Workspace ws{&heap};
    plh::HostStage stage;
    plh::XferStage xfer;
    bool prof = false;
    std::vector<plh_sum*> sums;  // live sums; orphaned (ctx = null) when the context goes
    std::vector<plh::ProfRec> prof_recs;
    std::vector<hipEvent_t> prof_pool;
    // sort-free rotation passes: a pass number tags the per-job and 16-tile
    // minima, which then need no reset (kWsRotE; capacities of its layout)
    uint64_t rot_epoch = 0;
    uint64_t rot_jcap = 0, rot_l3cap = 0;
};

namespace plh {
// RAII: events around the launches issued inside the scope.
class ProfScope {
   public:
    ProfScope(plh_ctx* c, const char* name, double alg_bytes = 0.0);
    ~ProfScope();

   private:
    plh_ctx* c_;
    int idx_ = -1;
};
}  // namespace plh

namespace plh {

// Workspace slots.
enum : int {
    kWsIds0 = 0,
    kWsIds1,
    kWsHist,
    kWsListA,
    kWsListB,
    kWsBig,
    kWsSub,
    kWsScan0,
    kWsScan1,
    kWsScan2,
    kWsKept,
    kWsOffs,
    kWsMisc,
    kWsFlags,
    kWsTmpSum,
    kWsPackL,
    kWsPackR,
    kWsPassL,   // per-pass arrays over the bucket list (alternating with kWsPassL2)
    kWsPassL2,
    kWsPassB,   // per-pass arrays over the big buckets
    kWsChild,   // children of the big buckets of a pass
    kWsJobs,    // product batch: job table and id-chunk table
    kWsRotA,    // sort-free rotation: anticommuting flags and their scan
    kWsRotN,    // sort-free rotation: the new terms, sorted
    kWsRotM,    // sort-free rotation: merge flags, lower bounds, sums, output scan
    kWsKeep,    // trim with keep: protected patterns and flags
    kWsRotE,    // sort-free rotation: per-job minima and 16-tile minima, epoch-tagged, kept between passes
    kWsGroup,   // collated product: the lhs grouped by root digit and the group sizes
    kWsCount
};

// (Re)allocate a sum to hold at least cap terms; keeps the first keep
// terms when growing.
int sum_reserve(plh_sum* s, int64_t cap, int64_t keep);
void sum_free_storage(plh_sum* s);
// dst takes src's storage (dst's old storage is released; src is left empty).
void sum_take_storage(plh_sum* dst, plh_sum* src);

// Device-wide exclusive scan of n u32 values into u64 offsets; the total
// is written to *d_total (device).  Stream-ordered.
int scan_u32_to_u64(plh_ctx* ctx, const uint32_t* d_in, uint64_t* d_out, uint64_t n,
                    uint64_t* d_total);
// The same on a given stream with a caller-provided buffer of
// scan_block_sums(n) u64 block sums (nothing drawn from a workspace).
int scan_u32_to_u64_scratch(hipStream_t st, const uint32_t* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total,
                            uint64_t* bsum);
uint64_t scan_block_sums(uint64_t n);
// Same for u64 inputs.
int scan_u64(plh_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total);

// Read one u64 from device memory (synchronises the stream).
int read_u64(plh_ctx* ctx, const uint64_t* d_src, uint64_t* out);
// Read n <= 64 consecutive u64 from device memory (one synchronisation).
int read_u64n(plh_ctx* ctx, const uint64_t* d_src, int n, uint64_t* out);
// A read-back posted by a kernel of the caller's: post_prepare gives the
// mailbox address and sequence number to hand to the kernel (post_write),
// post_wait polls for it after the launch.
struct Post {
    uint64_t* mbox = nullptr;
    uint64_t seq = 0;
};
int post_prepare(plh_ctx* ctx, Post* p);
int post_wait(plh_ctx* ctx, const Post& p, int n, uint64_t* out);
// scan_u32_to_u64 whose total also comes back to the host (*host_total).
int scan_u32_to_u64_read(plh_ctx* ctx, const uint32_t* d_in, uint64_t* d_out, uint64_t n, uint64_t* d_total,
                         uint64_t* host_total);

// Collation (plh_collate.hip).
struct CollateSpec {
    uint32_t mode = 0;      // kFold | kDropZero | kMaxWeight | kCutoff
    int64_t max_weight = 0;
    double eps = 0.0;
    // batched sums: per-job limits and keep mask (device arrays) or null
    const double* job_eps = nullptr;
    const int64_t* job_maxw = nullptr;
    const uint8_t* job_alive = nullptr;
};
// out = collate(in); out may alias in (the result replaces it).
int collate_terms(plh_ctx* ctx, plh_sum* in, const CollateSpec& spec, plh_sum* out);
// out = collate(evolve's general path applied to in): the rotation by P
// (psf = flags of P with the i bit flipped, c/sn its cos and sin) is folded
// into the collation, nothing is materialised; out may alias in.
int collate_rotation(plh_ctx* ctx, plh_sum* in, const uint64_t* px, const uint64_t* pz, uint32_t psf, double c,
                     double sn, const CollateSpec& spec, plh_sum* out);
// The same for a batched sum: each job rotates by its own P (device tables
// by job slot; jobs whose kind is not 4 pass through); p_or = OR of every
// job's P words (x then z, 2 * padded array size).
struct JobRotTables;
int collate_rotation_jobs(plh_ctx* ctx, plh_sum* in, const JobRotTables& jt, const uint64_t* p_or,
                          const CollateSpec& spec, plh_sum* out);
// In place: canonicalize(evolve(s)) for one general-path rotation followed
// by spec's filters, for a canonical s (sorted by (x, z, imag), distinct
// keys), without a sort (plh_rotate.hip); psf = flags of ps.
int rotate_sorted(plh_ctx* ctx, plh_sum* s, const uint64_t* px, const uint64_t* pz, uint32_t psf, double c,
                  double sn, const CollateSpec& spec);
// Rotations e = 0 .. w - 1 (w <= 16; kind 0 no change or 4 general, psf =
// flags of ps) on a sorted canonical sum of at most two words per plane that
// has passed spec's truncation since its keys last changed: the first
// rotation a term anticommutes with is applied as rotate_sorted, the ones
// before it are no-ops; *consumed = its index + 1, or w when none changes a
// term (the sum is then untouched).
int rotate_sorted_window(plh_ctx* ctx, plh_sum* s, int w, const uint64_t* const* px, const uint64_t* const* pz,
                         const uint8_t* kind, const uint32_t* psf, const double* c, const double* sn,
                         const CollateSpec& spec, int* consumed);
// The same for a canonical batched sum of job slots 0 .. njobs - 1, each job
// over its own window of rotations (JobWindows): a job applies the first
// rotation of its window that changes a term and consumes the rotations up
// to it (all of them when none does); jobs with an empty window pass
// through, under their own filters.  adv_host (device address of host-pinned
// memory, njobs bytes) receives the rotations each job consumed; it is
// complete when the call returns.  Every job's terms must have passed its
// truncation since its keys last changed.  place (or null): a sorted batched
// sum of other jobs (not alive under spec) whose terms go into the output
// between the rotated jobs', in job order.
struct JobWindows;
int rotate_sorted_jobs(plh_ctx* ctx, plh_sum* s, const JobWindows& jw, int njobs, const CollateSpec& spec,
                       uint8_t* adv_host, const plh_sum* place = nullptr);
// out = canonicalize(lhs[k] * rhs[k]) for k < njobs as one batched sum
// (job plane = k, jobs in order); offsets[k] .. offsets[k + 1] is job k's
// range (njobs + 1 entries, host).  Sum of n_k * m_k <= 2^32.
int collate_products(plh_ctx* ctx, int njobs, const plh_sum* const* lhs, const plh_sum* const* rhs,
                     const CollateSpec& spec, plh_sum* out, int64_t* offsets);
// outs[k] = job k's range offsets[k] .. offsets[k + 1] (host) of a
// job-ordered batched sum, for k < njobs, copied in one launch.
int split_jobs(plh_ctx* ctx, const plh_sum* batch, int njobs, const int64_t* offsets, plh_sum* const* outs);
// out = collate(L x R) for kind 0 (product, out[i*n+j] = L[j]*R[i]) or
// kind 1 (commutator of L with R, by_i adds the i factor).
int collate_pairs(plh_ctx* ctx, const plh_sum* L, const plh_sum* R, int kind, int by_i,
                  const CollateSpec& spec, plh_sum* out);

// Ops (plh_ops.hip).
int op_mul_raw(int as, const uint64_t* lx, const uint64_t* lz, const uint8_t* lf,
               const double* lv, int64_t lstride, const uint64_t* rx, const uint64_t* rz,
               const uint8_t* rf, const double* rv, int64_t rstride, uint64_t* ox, uint64_t* oz,
               uint8_t* of, double* ov, int64_t ostride, int64_t n, int64_t m, hipStream_t stream);
int op_commutator_raw(plh_ctx* ctx, const plh_sum* a, const plh_sum* b, int by_i, plh_sum* out);
int op_clifford(plh_ctx* ctx, plh_sum* s, int gate, int64_t q0, int64_t q1);
// The whole gate set of include/PauliSumClifford.h by id (0 I .. 9 sqrt_Y
// single-qubit, 10 CNOT .. 20 iSWAP two-qubit), one table-driven launch.
constexpr int kNumGates = 21;
constexpr int kFirstTwoQubitGate = 10;
// Program op kinds (plh_run_program, plh_server_submit): 0 H, 1 S, 2 CNOT,
// 3 rotation, and kProgramGateBase + id for any gate of the table set.
constexpr int kProgramGateBase = 16;
int op_clifford_table(plh_ctx* ctx, plh_sum* s, int gate, int64_t q0, int64_t q1);
int clifford_table_of(int gate, uint64_t* tbl, uint32_t* sgn);
// A run of up to kCliffordRunMax Clifford ops in one pass over the terms
// (each thread applies the whole run to its term in registers).  op[k]:
// kind << 28 | q0 << 14 | q1, kind 0 H, 1 S, 2 CNOT(q0 -> q1), 3 a one-qubit
// table gate, 4 a two-qubit table gate (tbl[k], sgn[k] as clifford_table_of
// gives them); qubits below 16384.
constexpr int kCliffordRunMax = 64;
struct CliffordRun {
    int nops;
    uint32_t op[kCliffordRunMax];
    uint64_t tbl[kCliffordRunMax];
    uint32_t sgn[kCliffordRunMax];
};
int op_clifford_run(plh_ctx* ctx, plh_sum* s, const CliffordRun& run, bool any_table);
int op_elementwise(plh_ctx* ctx, plh_sum* s, const plh_sum* rhs);
int op_rotation(plh_ctx* ctx, plh_sum* s, const uint64_t* px, const uint64_t* pz, uint8_t pflags,
                double cos_t, double sin_t, int* appended_general);
// evolve's path for (cos, sin) (include/PauliSumOperations.h:121-180): 0 no
// change, 1 sign flip, 2 sign flip and * ps, 3 * ps, 4 general (new terms).
int rotation_kind(double cos_t, double sin_t);
int op_filter(plh_ctx* ctx, plh_sum* s, uint32_t mode, int64_t max_weight, double eps);
// PauliSum::trim(max_terms): keep the max_terms largest |c| in term order
// (equal magnitudes: lower index first).
int op_trim(plh_ctx* ctx, plh_sum* s, int64_t max_terms);
// PauliSum::trim(max_terms, keep) (include/PauliSumUtils.h:348-375): terms
// whose (x, z) pattern appears in keep are kept; the others compete for the
// places left (keep null: op_trim).
int op_trim_keep(plh_ctx* ctx, plh_sum* s, int64_t max_terms, const plh_sum* keep);
// PauliSum::scale (include/PauliSumUtils.h:395-398): v *= c.
int op_scale(plh_ctx* ctx, plh_sum* s, double c);
// Fixed-order sums of a selection: mode 0 lookup of the pattern (qx, qz) with
// phase flags qf, 1 trace's identity sum (before the 2^n factor), 2
// frobenius_norm2.
int op_select_sum(plh_ctx* ctx, const plh_sum* s, int mode, const uint64_t* qx, const uint64_t* qz, uint32_t qf,
                  double* out);
int op_zero_state_expectation(plh_ctx* ctx, const plh_sum* s, double* out);
int op_append(plh_ctx* ctx, plh_sum* dst, const plh_sum* src);

}  // namespace plh
