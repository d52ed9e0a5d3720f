// CPU replay of the batched propagation server's host-side bookkeeping
// (paulib_amd/csrc/plh_sched.hpp, the scheduler the server runs) with the
// real sum storage (plh_storage.hip) and context heap (plh_heap.hip) over a
// simulated device (fake_vmm.hpp).  The device work of a step is replaced by
// a model of its memory traffic: the gates read and write the batch, the
// collation takes its workspace slots and an output block of the valid
// items, compacts into a second block when terms collapse or drop, and
// replaces the batch; a finished job's expectation and terms are read from
// its range.  Every simulated copy, fill, read and write must lie inside a
// live heap range when it is issued (stream-ordered use after release
// would be a fault on the device).
//
// Input (stdin): the jobs of a server test and, per job, the term count
// after each of its rotations as computed by the oracle's run_program on the
// program's suffix (count_at[l] = terms after ops l .. nops-1).  The replay
// checks, after every step:
//   - each job's planned ops are its program's ops in reverse order, each
//     exactly once, and each step ends at a rotation (or at op 0);
//   - the batch holds exactly the terms of the jobs in it (the shadow);
//   - no slot is reused while the terms of its previous job are still in
//     the batch;
//   - every job finishes once, with the oracle's final term count.
// Runs with injected failures (device memory capped so that allocations
// fail, and device errors at random phases of a step) check that an
// admission failure leaves the job queued and the batch as it was, that a
// failure in a step's device work makes the server refuse further steps,
// and that the batch and workspace stay consistent either way.  Built with
// -fsanitize=address,undefined by tests/test_sched_sim.py.
#include <cinttypes>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <random>
#include <vector>

#include "fake_vmm.hpp"

namespace sim {
plh::DeviceHeap* heap = nullptr;
bool is_device(const void* p) {
    const uintptr_t a = (uintptr_t)p;
    return a >= kBase && a < kBase + reserved;
}
void check_live(const void* p, size_t n, const char* what) {
    if (n == 0 || !is_device(p)) return;
    if (!heap || !heap->holds(p, n)) {
        std::printf("FAIL: %s touches device memory outside a live range (%p, %zu bytes)\n", what, p, n);
        std::exit(1);
    }
}
}  // namespace sim

extern "C" {
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t) {
    sim::check_live(dst, n, "copy destination");
    sim::check_live(src, n, "copy source");
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int, size_t n, hipStream_t) {
    sim::check_live(dst, n, "fill");
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
}

namespace plh {
HostStage::~HostStage() {}
XferStage::~XferStage() {}
}  // namespace plh

#include "../../paulib_amd/csrc/plh_heap.hip"
#include "../../paulib_amd/csrc/plh_sched.hpp"
#include "../../paulib_amd/csrc/plh_storage.hip"

using plh::sched::Job;
using plh::sched::Scheduler;
using plh::sched::StepPlan;

namespace {

int fail(const char* fmt, long long a = 0, long long b = 0) {
    std::printf("FAIL: ");
    std::printf(fmt, a, b);
    std::printf("\n");
    std::exit(1);
}

struct JobSpec {
    Job job;
    std::vector<int64_t> count_at;  // terms after ops l .. nops-1 (-1: not a step end); [nops] = n0
};

// The replayed server: the real scheduler and storage, a model of the
// device work, and a shadow of which job's terms the batch holds.
struct Replay {
    plh_ctx ctx;
    plh_sum batch;
    Scheduler sc;
    bool keep_terms = true;
    bool failed = false;
    int64_t scale = 1;  // bytes model: every count times scale
    std::mt19937_64 rng;
    double p_fail = 0.0;  // probability of a device error at each phase
    std::map<int, int64_t> shadow;       // slot -> terms of its job in the batch
    std::map<int, int64_t> pending;      // slot -> terms of its admitted job in the incoming sum
    plh_sum incoming, side;              // plh_serve.hip: admitted observables; a mixed step's C jobs
    std::map<int, int64_t> slot_handle;  // slot -> handle of the job whose terms those are
    std::map<int64_t, const JobSpec*> spec;
    std::map<int64_t, int64_t> finished;  // handle -> final terms
    std::map<int64_t, int64_t> applied;   // handle -> next op expected (Heisenberg order)
    int admission_errors = 0;
    int sorted_steps = 0;

    Replay(int as, int slots, int64_t budget, uint64_t seed) : sc(as, slots, budget), rng(seed) {
        ctx.heap.set_device(0);
        batch.ctx = &ctx;
        for (plh_sum* q : {&batch, &incoming, &side}) {
            q->ctx = &ctx;
            q->as = as;
            q->nq = 64 * as;
            q->has_job = true;
        }
    }
    ~Replay() {
        plh::sum_free_storage(&batch);
        plh::sum_free_storage(&incoming);
        plh::sum_free_storage(&side);
    }

    bool injected() { return p_fail > 0.0 && std::uniform_real_distribution<double>(0, 1)(rng) < p_fail; }
    int device_error() {
        plh::set_error("simulated device error");
        return PLH_ERR_HIP;
    }
    int64_t shadow_total() const {
        int64_t t = 0;
        for (auto& kv : shadow) t += kv.second;
        return t;
    }
    // The batch's terms lie in live storage (whatever happened before).
    void check_consistent(const char* when) {
        if (batch.n > batch.cap) fail("batch count %lld over its capacity %lld", batch.n, batch.cap);
        if (batch.n > 0) {
            if (!batch.block) fail("batch holds %lld terms and no storage", batch.n);
            sim::check_live(batch.x, (size_t)batch.cap * 8 * batch.as, when);
            sim::check_live(batch.f, (size_t)batch.n, when);
            sim::check_live(batch.j, (size_t)batch.n * 4, when);
        }
    }
    // ... and exactly the terms of the jobs the shadow says it holds.
    void check_batch(const char* when) {
        if (batch.n != shadow_total() * scale) fail("batch holds %lld terms, the shadow %lld", batch.n, shadow_total() * scale);
        check_consistent(when);
    }

    // plh_serve.hip admit(): a staging sum, uploads and the slot fill into
    // it, then op_append onto the incoming sum.
    int admit(const Job& j, int slot) {
        const int64_t n = j.n0 * scale;
        if (n == 0) return PLH_OK;
        plh_sum tmp;
        tmp.ctx = &ctx;
        tmp.as = sc.as();
        tmp.nq = 64 * sc.as();
        tmp.has_job = true;
        PLH_TRY(plh::sum_reserve(&tmp, n, 0));
        for (int k = 0; k < sc.as(); ++k) {
            (void)hipMemsetAsync(tmp.x + (int64_t)k * tmp.cap, 0, (size_t)n * 8, nullptr);
            (void)hipMemsetAsync(tmp.z + (int64_t)k * tmp.cap, 0, (size_t)n * 8, nullptr);
        }
        sim::check_live(tmp.j, (size_t)n * 4, "slot fill");
        int rc = PLH_OK;
        if (injected()) {
            rc = device_error();
        } else {
            tmp.n = n;
            rc = plh::op_append(&ctx, &incoming, &tmp);
        }
        plh::sum_free_storage(&tmp);
        return rc;
    }

    // plh_serve.hip step_device(): a full step appends the incoming sum to
    // the batch, then gates and one collation; a mixed step gathers the C
    // jobs' terms with the incoming ones into the side sum and collates it,
    // then rotates the R jobs without a sort, the side sum placed between
    // them.  The finished jobs' ranges are read last.
    // One collation's memory (with_output + collate_run's storage): X is
    // replaced by `out` terms.
    int collate_model(plh_sum& X, int64_t n_valid, int64_t out) {
        const int64_t n = X.n;
        if (injected()) return device_error();
        sim::check_live(X.x, (size_t)n * 8, "gates");
        sim::check_live(X.f, (size_t)n, "gates");
        sim::check_live(X.j, (size_t)n * 4, "gates");
        void* p = nullptr;
        PLH_TRY(ctx.ws.get(plh::kWsMisc, 4096, &p));
        PLH_TRY(ctx.ws.get(plh::kWsPackL, (size_t)n * (2 * sc.as() + 2) * 8, &p));
        PLH_TRY(ctx.ws.get(plh::kWsIds0, (size_t)n_valid * 8, &p));
        PLH_TRY(ctx.ws.get(plh::kWsIds1, (size_t)n_valid * 8, &p));
        if (injected()) return device_error();
        plh_sum tmp;
        tmp.ctx = &ctx;
        tmp.as = X.as;
        tmp.nq = X.nq;
        tmp.has_job = true;
        auto run = [&]() -> int {
            PLH_TRY(plh::sum_reserve(&tmp, std::max<int64_t>(n_valid, 1), 0));
            sim::check_live(X.x, (size_t)n * 8, "leaf reads");
            if (n_valid) sim::check_live(tmp.x, (size_t)n_valid * 8, "leaf writes");
            if (n_valid) sim::check_live(tmp.j, (size_t)n_valid * 4, "leaf writes");
            if (injected()) return device_error();
            if (out != n_valid) {  // collapsed runs or dropped terms: compaction
                plh_sum c2;
                c2.ctx = &ctx;
                c2.as = tmp.as;
                c2.nq = tmp.nq;
                c2.has_job = true;
                const int rc = plh::sum_reserve(&c2, std::max<int64_t>(out, 1), 0);
                if (rc != PLH_OK) {
                    plh::sum_free_storage(&c2);
                    return rc;
                }
                if (out) sim::check_live(c2.v, (size_t)out * 8, "compaction writes");
                plh::sum_take_storage(&tmp, &c2);
            }
            tmp.n = out;
            return PLH_OK;
        };
        const int rc = run();
        if (rc != PLH_OK) {
            plh::sum_free_storage(&tmp);
            return rc;
        }
        plh::sum_take_storage(&X, &tmp);
        X.n = out;
        X.sorted = true;
        return PLH_OK;
    }

    int device(StepPlan& pl, plh::sched::StepKind kind, std::map<int, int64_t>* after) {
        if (kind == plh::sched::kSortedStep) {
            // the device's report: each R job consumes 1 .. wn of its window
            // (here drawn at random: any count is a valid report)
            std::vector<uint8_t> adv((size_t)sc.slots(), 0);
            for (int i = 0; i < sc.slots(); ++i)
                if (pl.wn[(size_t)i] > 0)
                    adv[(size_t)i] = (uint8_t)(1 + std::uniform_int_distribution<int>(0, pl.wn[(size_t)i] - 1)(rng));
            if (!sc.settle(&pl, adv.data())) fail("settle refused a valid report");
            verify_settled(pl, adv);
        }
        // the terms each running slot holds once this step has run
        after->clear();
        int64_t out = 0, out_c = 0, in_c = 0, valid_c = 0;
        for (int slot = 0; slot < sc.slots(); ++slot) {
            if (sc.state(slot) != plh::sched::kRunning) continue;  // draining: dropped
            const Job& j = sc.job(slot);
            const JobSpec* sp = spec.at(j.handle);
            const int64_t c = sp->count_at[(size_t)(pl.next[(size_t)slot] + 1)];
            if (c < 0) fail("the step of a job ends at op %lld, which is neither a rotation nor op 0", pl.next[(size_t)slot] + 1);
            (*after)[slot] = c;
            out += c * scale;
            const bool is_c = kind == plh::sched::kFullStep || pl.isc[(size_t)slot];
            if (is_c) {
                const int64_t held = (shadow.count(slot) ? shadow.at(slot) : 0) + (pending.count(slot) ? pending.at(slot) : 0);
                in_c += held * scale;
                out_c += c * scale;
                valid_c += (pl.rk[(size_t)slot] == 4 ? 2 : 1) * held * scale;  // candidate new terms
            }
        }
        if (kind == plh::sched::kFullStep) {
            PLH_TRY(plh::op_append(&ctx, &batch, &incoming));
            incoming.n = 0;
            pending.clear();
            if (batch.n > 0) {
                const int64_t nv = batch.n + (valid_c - in_c);
                PLH_TRY(collate_model(batch, nv, out));
            }
        } else {
            // C jobs apart: gather, the incoming terms, collation
            side.n = 0;
            if (pl.any_c) {
                void* p = nullptr;
                PLH_TRY(ctx.ws.get(plh::kWsTmpSum, (size_t)sc.slots() * 16 + 1024, &p));
                if (injected()) return device_error();
                const int64_t gathered = in_c - incoming.n;
                PLH_TRY(plh::sum_reserve(&side, std::max<int64_t>(gathered + incoming.n, 1), 0));
                if (gathered) sim::check_live(side.x, (size_t)gathered * 8, "gather writes");
                side.n = gathered;
                PLH_TRY(plh::op_append(&ctx, &side, &incoming));
                incoming.n = 0;
                pending.clear();
                if (side.n > 0) PLH_TRY(collate_model(side, valid_c, out_c));
            }
            // R jobs: plh_rotate.hip rotate_sorted_jobs, the side sum placed
            const int64_t n = batch.n;
            void* p = nullptr;
            PLH_TRY(ctx.ws.get(plh::kWsRotA, (size_t)n + (size_t)n / 2 + 4096 + (size_t)sc.slots() * 600, &p));
            PLH_TRY(ctx.ws.get(plh::kWsRotN, (size_t)n * (16 * sc.as() + 17) + 1536, &p));
            if (injected()) return device_error();
            sim::check_live(batch.x, (size_t)n * 8, "sorted rotation reads");
            sim::check_live(batch.j, (size_t)n * 4, "sorted rotation reads");
            plh_sum tmp;
            tmp.ctx = &ctx;
            tmp.as = batch.as;
            tmp.nq = batch.nq;
            tmp.has_job = true;
            const int rc = plh::sum_reserve(&tmp, std::max<int64_t>(out, 1), 0);
            if (rc != PLH_OK) {
                plh::sum_free_storage(&tmp);
                return rc;
            }
            if (out) sim::check_live(tmp.v, (size_t)out * 8, "sorted rotation writes");
            if (side.n) sim::check_live(side.v, (size_t)side.n * 8, "placed terms");
            if (injected()) {
                plh::sum_free_storage(&tmp);
                return device_error();
            }
            plh::sum_take_storage(&batch, &tmp);
            batch.n = out;
            batch.sorted = true;
        }
        if (pl.finishing.empty()) return PLH_OK;
        if (out > 0) {
            void* p = nullptr;
            PLH_TRY(ctx.ws.get(plh::kWsTmpSum, pl.finishing.size() * 32 + 512, &p));
            if (injected()) return device_error();
            for (int slot : pl.finishing) {
                const int64_t m = after->at(slot) * scale;
                if (keep_terms && m > 0) {
                    int64_t a = 0;  // the job-ordered batch: slots ascending
                    for (auto& kv : *after)
                        if (kv.first < slot) a += kv.second * scale;
                    sim::check_live(batch.x + a, (size_t)m * 8, "download of a finished job");
                    sim::check_live(batch.v + a, (size_t)m * 8, "download of a finished job");
                }
            }
        }
        return PLH_OK;
    }

    // One plh_server_step: returns its status.
    int step() {
        if (failed) return PLH_ERR_STATE;
        int64_t in_flight = (batch.n + incoming.n) / scale;  // the scheduler counts unscaled terms
        for (int slot; (slot = sc.admission_slot(in_flight)) >= 0;) {
            if (shadow.count(slot) || pending.count(slot))
                fail("slot %lld reused while its previous job's terms are in the batch", slot);
            const Job& q = sc.queue().front();
            const int64_t h = q.handle, n0 = q.n0, nb = batch.n, ni = incoming.n;
            const int rc = admit(q, slot);
            if (rc != PLH_OK) {
                // the job stays queued, the batch and the incoming sum keep their terms
                if (sc.queue().empty() || sc.queue().front().handle != h) fail("a failed admission lost the job");
                if (batch.n != nb || incoming.n != ni) fail("a failed admission changed the batch or incoming count");
                check_batch("after a failed admission");
                ++admission_errors;
                return rc;
            }
            sc.admitted(slot);
            pending[slot] = n0;
            slot_handle[slot] = h;
            in_flight += n0;
        }
        const plh::sched::StepKind kind = sc.choose_step(batch.sorted, batch.n);
        const bool full = kind == plh::sched::kFullStep;
        if (!full && !batch.sorted) fail("a mixed step on an unsorted batch");
        StepPlan pl;
        sc.plan(&pl, kind);
        verify_plan(pl, kind);
        std::map<int, int64_t> after;
        const int64_t nb = batch.n;
        const int rc = device(pl, kind, &after);
        if (rc != PLH_OK) {
            // before the collation replaced the batch it is unchanged;
            // after, it is the collation's result; consistent either way
            failed = true;
            if (batch.n != nb && batch.n != 0 && !batch.block) fail("a failed step left terms without storage");
            check_consistent("after a failed step");
            return rc;
        }
        // the collation dropped the draining jobs and applied this step
        shadow = after;
        for (int slot : pl.finishing) {
            const int64_t h = sc.job(slot).handle;
            if (finished.count(h)) fail("job %lld finished twice", h);
            finished[h] = after.at(slot);
        }
        for (size_t i = 0; i < pl.next.size(); ++i)
            if (sc.state((int)i) == plh::sched::kRunning) applied[sc.job((int)i).handle] = pl.next[i];
        if (sc.commit(pl)) {
            batch.n = 0;
            shadow.clear();
        }
        sorted_steps += !full;
        if (incoming.n != 0 || !pending.empty()) fail("a step left admitted terms unconsumed");
        check_batch("after a step");
        return PLH_OK;
    }

    // After a sorted step's report: each advancing job moved by its count,
    // waiting jobs stayed, and exactly the jobs that passed op 0 finish.
    void verify_settled(const StepPlan& pl, const std::vector<uint8_t>& adv) {
        for (int i = 0; i < sc.slots(); ++i) {
            if (sc.state(i) != plh::sched::kRunning) continue;
            const Job& j = sc.job(i);
            if (pl.wn[(size_t)i] > 0 && pl.next[(size_t)i] != j.next - (int64_t)adv[(size_t)i])
                fail("job %lld settled at the wrong op", j.handle);
            const bool fin = std::find(pl.finishing.begin(), pl.finishing.end(), i) != pl.finishing.end();
            if (fin != (pl.next[(size_t)i] < 0)) fail("finishing flag of job %lld", j.handle);
        }
        std::vector<uint8_t> bad = adv;
        for (size_t i = 0; i < bad.size(); ++i)
            if (sc.state((int)i) == plh::sched::kRunning) bad[i] = (uint8_t)(pl.wn[i] + 1);
        StepPlan copy = pl;
        if (std::any_of(pl.wn.begin(), pl.wn.end(), [](uint8_t w) { return w > 0; }) && sc.settle(&copy, bad.data()))
            fail("settle accepted a count past a window");
    }

    // The plan's ops for each running job are its program's next ops, in
    // reverse order, up to and including one rotation.
    void verify_plan(const StepPlan& pl, plh::sched::StepKind kind) {
        for (int i = 0; i < sc.slots(); ++i) {
            if (sc.state(i) != plh::sched::kRunning) {
                if (pl.off[(size_t)i] != pl.off[(size_t)i + 1] || pl.alive[(size_t)i]) fail("a slot that does not run has ops");
                continue;
            }
            const Job& j = sc.job(i);
            auto it = applied.find(j.handle);
            const int64_t expect_next = it == applied.end() ? (int64_t)j.kind.size() - 1 : it->second;
            if (j.next != expect_next) fail("job %lld resumes at op %lld", j.handle, j.next);
            const bool r_job = kind == plh::sched::kSortedStep && sc.sorted_ready(i) && !sc.fresh(i);
            if (kind == plh::sched::kSortedStep && pl.isc[(size_t)i] != (r_job ? 0 : (pl.any_c ? 1 : 0)))
                fail("job %lld is %lld in a mixed step's C set", j.handle, pl.isc[(size_t)i]);
            if (kind == plh::sched::kSortedStep && !r_job && pl.wn[(size_t)i] != 0) fail("a C job has a window (job %lld)", j.handle);
            if (kind == plh::sched::kSortedStep && !r_job && !pl.any_c) {  // waits: untouched
                if (sc.fresh(i)) fail("an admitted job waits (job %lld)", j.handle);
                if (pl.off[(size_t)i] != pl.off[(size_t)i + 1] || pl.rk[(size_t)i] != 0 || pl.maxw[(size_t)i] != -1 ||
                    pl.eps[(size_t)i] != 0.0 || pl.next[(size_t)i] != j.next || !pl.alive[(size_t)i])
                    fail("a waiting job is not left alone (job %lld)", j.handle);
                continue;
            }
            if (r_job) {
                if (pl.off[(size_t)i] != pl.off[(size_t)i + 1]) fail("Clifford ops for an R job (job %lld)", j.handle);
                if (pl.rk[(size_t)i] != 0 || pl.next[(size_t)i] != j.next || !pl.alive[(size_t)i] ||
                    std::find(pl.finishing.begin(), pl.finishing.end(), i) != pl.finishing.end())
                    fail("a mixed step's plan decides for R job %lld before the device", j.handle);
                // the window: the job's next rotations in order, each keeping the
                // order, up to the window depth or the first op that does not
                const int wn = pl.wn[(size_t)i];
                if (wn < 1 || wn > plh::sched::kWindow) fail("window of job %lld holds %lld rotations", j.handle, wn);
                const size_t A = (size_t)sc.as(), ja = (size_t)j.job_as;
                int64_t l = j.next;
                for (int e = 0; e < wn; ++e, --l) {
                    if (l < 0 || j.kind[(size_t)l] != 3) fail("window entry %lld of job %lld is not a rotation", e, j.handle);
                    double c, sn;
                    plh::angle_cos_sin(j.theta[(size_t)l], &c, &sn);
                    const uint8_t want = plh::sched::rotation_kind_of(c, sn);
                    const size_t ent = (size_t)i * plh::sched::kWindow + (size_t)e;
                    if (pl.wk[ent] != want || (want != 0 && want != 4)) fail("window entry %lld of job %lld: kind", e, j.handle);
                    if (pl.wc[ent] != c || pl.ws[ent] != sn) fail("window entry %lld of job %lld: angle", e, j.handle);
                    for (size_t k = 0; k < A; ++k)
                        if (pl.wp[ent * 2 * A + k] != (k < ja ? j.px[(size_t)l * ja + k] : 0ull) ||
                            pl.wp[ent * 2 * A + A + k] != (k < ja ? j.pz[(size_t)l * ja + k] : 0ull))
                            fail("window entry %lld of job %lld: generator", e, j.handle);
                }
                if (wn < plh::sched::kWindow && l >= 0 && j.kind[(size_t)l] == 3) {
                    double c, sn;
                    plh::angle_cos_sin(j.theta[(size_t)l], &c, &sn);
                    const uint8_t k = plh::sched::rotation_kind_of(c, sn);
                    if (k == 0 || k == 4) fail("window of job %lld stops before an order-keeping rotation", j.handle);
                }
                if (pl.maxw[(size_t)i] != j.max_weight || pl.eps[(size_t)i] != j.cutoff) fail("truncation of job %lld", j.handle);
                continue;
            }
            int64_t l = j.next;
            for (uint32_t o = pl.off[(size_t)i]; o < pl.off[(size_t)i + 1]; ++o, --l) {
                if (l < 0) fail("job %lld planned past its first op", j.handle);
                const uint32_t op = pl.ops[o];
                const int k = j.kind[(size_t)l];
                if (k == 3) fail("a rotation among the Clifford ops of job %lld (op %lld)", j.handle, l);
                const uint32_t kind = op >> 28;
                if (k >= plh::sched::kGateBase) {
                    if (kind != 3 || ((op >> 20) & 0xFFu) != (uint32_t)(k - plh::sched::kGateBase) ||
                        ((op >> 10) & 0x3FFu) != (uint32_t)j.q0[(size_t)l])
                        fail("table gate op %lld of job %lld encoded wrongly", l, j.handle);
                } else if (kind != (uint32_t)k || ((op >> 14) & 0x3FFFu) != (uint32_t)j.q0[(size_t)l] ||
                           (k == 2 && (op & 0x3FFFu) != (uint32_t)j.q1[(size_t)l])) {
                    fail("Clifford op %lld of job %lld encoded wrongly", l, j.handle);
                }
            }
            // then either one rotation or the program's start
            if (l >= 0) {
                if (j.kind[(size_t)l] != 3) fail("job %lld's step stops at a non-rotation op %lld", j.handle, l);
                if (pl.next[(size_t)i] != l - 1) fail("job %lld's next op after the step is %lld", j.handle, pl.next[(size_t)i]);
                double c, sn;
                plh::angle_cos_sin(j.theta[(size_t)l], &c, &sn);
                const uint8_t want = plh::sched::rotation_kind_of(c, sn);
                if (pl.rk[(size_t)i] != want) fail("rotation kind of job %lld", j.handle);
                if (pl.maxw[(size_t)i] != j.max_weight || pl.eps[(size_t)i] != j.cutoff) fail("truncation of job %lld", j.handle);
            } else {
                if (pl.next[(size_t)i] != -1 || pl.rk[(size_t)i] != 0 || pl.maxw[(size_t)i] != -1)
                    fail("job %lld's last Clifford step", j.handle);
            }
            const bool fin = std::find(pl.finishing.begin(), pl.finishing.end(), i) != pl.finishing.end();
            if (fin != (pl.next[(size_t)i] < 0)) fail("finishing flag of job %lld", j.handle);
        }
    }
};

std::vector<JobSpec> read_jobs(FILE* in, int* slots, int* as) {
    int nj = 0;
    if (std::fscanf(in, "%d %d %d", &nj, slots, as) != 3) fail("input header");
    std::vector<JobSpec> jobs((size_t)nj);
    for (int k = 0; k < nj; ++k) {
        JobSpec& s = jobs[(size_t)k];
        Job& j = s.job;
        long long nq, n0, nops, mw;
        double cut;
        if (std::fscanf(in, "%lld %lld %lld %lld %lf", &nq, &n0, &nops, &mw, &cut) != 5) fail("job header %lld", k);
        j.job_as = (int)((nq + 63) / 64);
        j.n0 = n0;
        j.max_weight = mw;
        j.cutoff = cut;
        j.kind.resize((size_t)nops);
        j.q0.resize((size_t)nops);
        j.q1.resize((size_t)nops);
        j.theta.resize((size_t)nops);
        j.px.assign((size_t)nops * j.job_as, 0);
        j.pz.assign((size_t)nops * j.job_as, 0);
        s.count_at.assign((size_t)nops + 1, -1);
        for (long long l = 0; l < nops; ++l) {
            long long q0, q1, cnt;
            int kind;
            if (std::fscanf(in, "%d %lld %lld %lf %lld", &kind, &q0, &q1, &j.theta[(size_t)l], &cnt) != 5) fail("op %lld", l);
            j.kind[(size_t)l] = kind;
            j.q0[(size_t)l] = q0;
            j.q1[(size_t)l] = q1;
            s.count_at[(size_t)l] = cnt;
            for (int w = 0; w < j.job_as; ++w) {
                unsigned long long x, z;
                if (std::fscanf(in, "%llx %llx", &x, &z) != 2) fail("rotation words");
                j.px[(size_t)l * j.job_as + w] = x;
                j.pz[(size_t)l * j.job_as + w] = z;
            }
        }
        s.count_at[(size_t)nops] = n0;
        j.next = nops - 1;
    }
    return jobs;
}

int g_sorted_steps = 0;

// Submits every job, then steps until all are done or the server fails; a
// failed admission is retried at the next step, as a caller would.
void run(const std::vector<JobSpec>& jobs, int slots, int as, int64_t scale, size_t cap_chunks, double p_fail,
         uint64_t seed, int* n_failed_runs, int* n_adm_errors, int* n_complete = nullptr) {
    sim::cap_chunks = cap_chunks;
    const int64_t budget = (int64_t)(cap_chunks * plh::DeviceHeap::kChunk / 2);
    Replay r(as, slots, budget, seed);
    sim::heap = &r.ctx.heap;
    r.scale = scale;
    r.p_fail = p_fail;
    for (size_t k = 0; k < jobs.size(); ++k) {
        Job j = jobs[k].job;
        j.handle = (int64_t)k;
        r.spec[j.handle] = &jobs[k];
        r.sc.queue().push_back(std::move(j));
    }
    bool stuck = false;
    for (int guard = 0;; ++guard) {
        if (guard > 1000000) fail("no progress after a million steps");
        const int rc = r.step();
        if (rc != PLH_OK && r.failed) {
            if (r.step() != PLH_ERR_STATE) fail("a failed server accepted another step");
            ++*n_failed_runs;
            break;
        }
        if (rc != PLH_OK && r.sc.running() == 0 && r.batch.n == 0 && p_fail == 0.0) {
            stuck = true;  // the device cannot hold the queue's first job even alone
            break;
        }
        if (r.sc.running() == 0 && r.sc.queue().empty()) break;
    }
    *n_adm_errors += r.admission_errors;
    g_sorted_steps += r.sorted_steps;
    if (!r.failed && !stuck) {
        if (r.finished.size() != jobs.size()) fail("%lld of %lld jobs finished", (long long)r.finished.size(), (long long)jobs.size());
        for (auto& kv : r.finished) {
            const int64_t want = jobs[(size_t)kv.first].count_at[0];
            if (kv.second != want) fail("job %lld ended with %lld terms", kv.first, kv.second);
        }
        if (r.batch.n != 0) fail("terms left in the batch after the last job");
        if (n_complete) ++*n_complete;
    }
    plh::sum_free_storage(&r.batch);
    plh::sum_free_storage(&r.incoming);
    plh::sum_free_storage(&r.side);
    r.ctx.ws.release();
    if (r.ctx.heap.bytes_live() != 0) fail("%lld bytes still live after the server went", (long long)r.ctx.heap.bytes_live());
    sim::heap = nullptr;
}

}  // namespace

int main() {
    int slots = 0, as = 0;
    const std::vector<JobSpec> jobs = read_jobs(stdin, &slots, &as);
    int failed_runs = 0, adm_errors = 0;
    // plain replay at the test's sizes, then with a single slot and with
    // many more slots than jobs
    for (int s : {slots, 1, 64}) run(jobs, s, as, 1, 32, 0.0, 1, &failed_runs, &adm_errors);
    if (failed_runs || adm_errors) fail("failures without injection");
    if (g_sorted_steps == 0) fail("no sorted step in the clean replays");
    std::printf("clean replays ok (%zu jobs, %d sorted steps)\n", jobs.size(), g_sorted_steps);
    // sizes scaled up against a small device, so that admissions and
    // collations run out of memory, and device errors at random phases
    const int64_t max_terms = [&] {
        int64_t m = 1;
        for (auto& j : jobs)
            for (int64_t c : j.count_at) m = std::max(m, c);
        return m;
    }();
    // the largest job then needs about a chunk of device memory in flight
    const int64_t scale = std::max<int64_t>(1, (int64_t)((double)plh::DeviceHeap::kChunk / (200.0 * (double)max_terms)));
    int complete = 0;
    for (uint64_t seed = 1; seed <= 200; ++seed) {
        const size_t cap = 2 + seed % 7;
        const double p = seed % 3 == 0 ? 0.0 : 0.002 * (double)(seed % 5);
        run(jobs, slots, as, scale, cap, p, seed, &failed_runs, &adm_errors, &complete);
    }
    if (complete == 0 || failed_runs == 0) fail("the injected runs did not mix completions (%lld) and failures (%lld)", complete, failed_runs);
    std::printf("injected (200 runs, terms x%lld, 2..8 chunks): %d completed, %d ended by a failed step, "
                "%d admission failures retried; ok\n", (long long)scale, complete, failed_runs, adm_errors);
    if (sim::created != 0 || !sim::mapped.empty() || sim::reserved) fail("heap left memory behind");
    return 0;
}
