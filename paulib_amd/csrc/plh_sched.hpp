// plh_sched.hpp -- host-side bookkeeping of the batched propagation server:
// the job queue, the slots and their states, admission by HBM budget, and
// the per-step tables (Clifford ops up to each job's next rotation, that
// rotation's fast-path kind, the step's truncation).  Pure C++ with no HIP
// calls, so the same code runs in the server (plh_serve.hip) and in the
// CPU replay of its bookkeeping (tests/cpp/sched_sim.cpp).
//
// A step is planned without changing any job (plan), the device work runs,
// and only then is the plan committed (commit): a step whose device work
// fails leaves every job where it was, and the server refuses further
// steps, since the batch may hold half-applied gates.
//
// Per job the program order is that of propagate()
// (include/PauliPropagation.h:53-63): ops from the last to the first, a
// rotation (evolve, include/PauliSumOperations.h:106-183) followed by
// apply_truncation (include/PauliPropagation.h:35-47); Clifford ops alone
// are not followed by a truncation.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <vector>

#include "plh_angles.hpp"

namespace plh {
namespace sched {

// Program op kinds: 0 H, 1 S, 2 CNOT, 3 rotation, kGateBase + id for a gate
// of the table set.
constexpr int kGateBase = 16;
constexpr int kGates = 21;
constexpr int kFirstTwoQubit = 10;

struct Job {
    int64_t handle = -1;
    int job_as = 1;
    int64_t n0 = 0;
    std::vector<uint64_t> ox, oz;  // observable planes, x[k*n0 + t]
    std::vector<uint8_t> of;
    std::vector<double> ov;
    std::vector<int32_t> kind;
    std::vector<int64_t> q0, q1;
    std::vector<uint64_t> px, pz;  // per op: job_as words
    std::vector<double> theta;
    int64_t next = -1;  // next op to apply (Heisenberg order: last to first)
    int64_t max_weight = -1;
    double cutoff = 0.0;
    int64_t max_terms = -1;  // trim after each truncation (< 0: none)
    double t_submit = 0.0;
};

enum SlotState : int { kFree = 0, kRunning = 1, kDraining = 2 };

// A full step advances every running job by its Clifford ops up to and
// including its next rotation and collates the batch.  A mixed step keeps
// the batch sorted and advances every running job too: a job whose next op
// is a rotation that keeps the order (its truncation alone, or the general
// path) and whose terms are in the sorted batch advances by a sort-free
// rotation window (an R job); the others (a Clifford op or a key-moving fast
// path next, or admitted this step) take the full step's ops on their own
// terms, which are collated apart and placed between the R jobs' (C jobs).
enum StepKind : int { kFullStep = 0, kSortedStep = 1 };

// A sorted step hands each advancing job a window of up to kWindow
// consecutive order-keeping rotations.  The device applies the first one
// that changes the job's terms (a general rotation with an anticommuting
// term) and reports how many rotations the job consumed: the ones before it
// are exact no-ops, since the job's terms already passed the truncation
// that follows each of them, and a rotation that no term anticommutes with
// changes no term.
constexpr int kWindow = 16;

// One step's device tables (one entry per slot) and its effect on the jobs.
struct StepPlan {
    std::vector<uint32_t> off, ops;  // ops[off[i] .. off[i+1]) are slot i's Clifford ops
    std::vector<uint8_t> rk, rpsf, alive;
    std::vector<double> rc, rs, eps;
    std::vector<int64_t> maxw;
    std::vector<int64_t> maxt;    // per slot: max_terms when its step truncates and it trims, else -1
    bool any_trim = false;
    // mixed steps: per slot 1 for a C job (collated apart), and whether any
    std::vector<uint8_t> isc;
    bool any_c = false;
    std::vector<uint64_t> rp;     // per slot: x[AS] then z[AS] of its rotation
    std::vector<uint64_t> p_or;   // OR of the general-path rotations' words (x then z)
    std::vector<int64_t> next;    // per slot: the job's next op after this step
    std::vector<int> finishing;   // slots whose job ends with this step
    bool any_general = false;
    // sorted steps: per slot its window (wn rotations, entries slot * kWindow + k:
    // rotation kind, ps flags, cos, sin, P as x[AS] then z[AS]); next and
    // finishing are set by Scheduler::settle once the device has reported
    // how far each job advanced
    std::vector<uint8_t> wn, wk, wpsf;
    std::vector<double> wc, ws;
    std::vector<uint64_t> wp;
};

inline bool close_one(double v) { return std::fabs(v - 1.0) <= 1e-14; }  // isCloseOne, :19-21

// evolve's path for (cos, sin) (include/PauliSumOperations.h:121-147): 0 no
// change, 1 sign flip, 2 sign flip and * ps, 3 * ps, 4 general.
inline uint8_t rotation_kind_of(double c, double s) {
    if (close_one(c)) return 0;
    if (close_one(-c)) return 1;
    if (close_one(s)) return 2;
    if (close_one(-s)) return 3;
    return 4;
}

// Device bytes per term in flight: the batch (x, z, v, j, f), the collation
// output, its two id arrays (an id and 32 carried key bits each) and packed
// key records.
inline double bytes_per_term(int as) { return 2.0 * (16.0 * as + 13.0) + 16.0 + (2.0 * as + 2.0) * 8.0; }

class Scheduler {
   public:
    Scheduler(int as, int slots, int64_t budget)
        : as_(as), slots_(slots), budget_(budget), active_((size_t)slots), state_((size_t)slots, kFree) {}

    int as() const { return as_; }
    int slots() const { return slots_; }
    int64_t budget() const { return budget_; }
    std::deque<Job>& queue() { return queue_; }
    const std::deque<Job>& queue() const { return queue_; }
    Job& job(int slot) { return active_[(size_t)slot]; }
    int state(int slot) const { return state_[(size_t)slot]; }
    int running() const {
        int r = 0;
        for (int s : state_) r += s == kRunning;
        return r;
    }

    // The slot the queue's first job would be admitted to, or -1: a free
    // slot, and (unless nothing runs) room for the batch plus 64 times the
    // new job's terms within the budget.
    int admission_slot(int64_t terms_in_flight) const {
        if (queue_.empty()) return -1;
        int slot = -1;
        for (int i = 0; i < slots_ && slot < 0; ++i)
            if (state_[(size_t)i] == kFree) slot = i;
        if (slot < 0) return -1;
        const Job& q = queue_.front();
        const double need =
            ((double)terms_in_flight + 64.0 * (double)std::max<int64_t>(q.n0, 1)) * bytes_per_term(as_);
        if (running() > 0 && need > (double)budget_) return -1;
        return slot;
    }
    // The queue's first job, whose terms the batch now holds, runs in slot.
    void admitted(int slot) {
        active_[(size_t)slot] = std::move(queue_.front());
        queue_.pop_front();
        state_[(size_t)slot] = kRunning;
        fresh_[(size_t)slot] = 1;
    }
    // Whether the slot's job was admitted for the coming step (its terms are
    // not in the sorted batch yet).
    bool fresh(int slot) const { return fresh_[(size_t)slot] != 0; }

    // Whether a running job can advance in a sorted step.
    bool sorted_ready(int slot) const {
        const Job& j = active_[(size_t)slot];
        if (state_[(size_t)slot] != kRunning || j.next < 0 || j.kind[(size_t)j.next] != 3) return false;
        double c, sn;
        angle_cos_sin(j.theta[(size_t)j.next], &c, &sn);
        const uint8_t k = rotation_kind_of(c, sn);
        return k == 0 || k == 4;
    }
    int sorted_ready_count() const {
        int r = 0;
        for (int i = 0; i < slots_; ++i) r += sorted_ready(i);
        return r;
    }

    // The next step's kind: a mixed step while the batch is sorted and some
    // running job can advance without a sort, a full step otherwise.
    StepKind choose_step(bool batch_sorted, int64_t batch_terms) const {
        const bool full = !batch_sorted || batch_terms == 0 || sorted_ready_count() == 0;
        return full ? kFullStep : kSortedStep;
    }

    // Whether this mixed step takes its C jobs (their Clifford layers and the
    // admitted jobs): when at least an eighth of the running jobs are C, or
    // one of them has waited kMaxWait steps; else they wait (their terms stay
    // in the batch untouched) and only R jobs advance, which saves a small
    // collation's launches and read-backs per step.
    static constexpr int kMaxWait = 8;
    bool take_c() const {
        int c = 0, run = 0, waited = 0;
        for (int i = 0; i < slots_; ++i) {
            if (state_[(size_t)i] != kRunning) continue;
            ++run;
            if (fresh(i)) return true;  // admitted jobs' terms are outside the batch
            if (!sorted_ready(i)) {
                ++c;
                waited = std::max(waited, (int)wait_[(size_t)i]);
            }
        }
        return c > 0 && (8 * c >= run || waited >= kMaxWait);
    }

    // This step's tables; no job changes.  In a mixed step the R jobs get
    // their windows and the C jobs the full step's tables (or, when the step
    // does not take them, nothing: they wait).
    void plan(StepPlan* p, StepKind kind = kFullStep) const {
        const bool with_c = kind == kFullStep || take_c();
        const size_t S = (size_t)slots_, A = (size_t)as_;
        p->off.assign(S + 1, 0);
        p->ops.clear();
        p->rk.assign(S, 0);
        p->rpsf.assign(S, 1);  // program generators carry flags 0, so ps has flags 1
        p->alive.assign(S, 0);
        p->rc.assign(S, 0.0);
        p->rs.assign(S, 0.0);
        p->eps.assign(S, 0.0);
        p->maxw.assign(S, -1);
        p->maxt.assign(S, -1);
        p->any_trim = false;
        p->isc.assign(S, 0);
        p->any_c = false;
        p->rp.assign(S * 2 * A, 0);
        p->p_or.assign(2 * A, 0);
        p->next.assign(S, -1);
        p->finishing.clear();
        p->any_general = false;
        const size_t SW = kind == kSortedStep ? S * (size_t)kWindow : 0;
        p->wn.assign(kind == kSortedStep ? S : 0, 0);
        p->wk.assign(SW, 0);
        p->wpsf.assign(SW, 1);
        p->wc.assign(SW, 0.0);
        p->ws.assign(SW, 0.0);
        p->wp.assign(SW * 2 * A, 0);
        for (size_t i = 0; i < S; ++i) {
            p->off[i] = (uint32_t)p->ops.size();
            if (state_[i] != kRunning) continue;
            p->alive[i] = 1;
            const Job& j = active_[i];
            int64_t next = j.next;
            if (kind == kSortedStep && sorted_ready((int)i) && !fresh((int)i)) {
                p->next[i] = next;  // settled after the device
                // the window: rotations next, next - 1, ... that keep the order
                int w = 0;
                const size_t ja = (size_t)j.job_as;
                for (int64_t l = next; w < kWindow && l >= 0 && j.kind[(size_t)l] == 3; --l, ++w) {
                    double c, sn;
                    angle_cos_sin(j.theta[(size_t)l], &c, &sn);
                    const uint8_t rk = rotation_kind_of(c, sn);
                    if (rk != 0 && rk != 4) break;
                    const size_t e = i * (size_t)kWindow + (size_t)w;
                    p->wk[e] = rk;
                    p->wc[e] = c;
                    p->ws[e] = sn;
                    for (size_t k = 0; k < ja; ++k) {
                        p->wp[e * 2 * A + k] = j.px[(size_t)l * ja + k];
                        p->wp[e * 2 * A + A + k] = j.pz[(size_t)l * ja + k];
                    }
                }
                p->wn[i] = (uint8_t)w;
                p->maxw[i] = j.max_weight;
                p->eps[i] = j.cutoff;
                p->maxt[i] = j.max_terms;
                p->any_trim |= j.max_terms >= 0;
                continue;
            }
            if (kind == kSortedStep) {
                if (!with_c) {
                    p->next[i] = next;  // waits, terms untouched
                    continue;
                }
                p->isc[i] = 1;
                p->any_c = true;
            }
            while (next >= 0 && j.kind[(size_t)next] != 3) {
                const int64_t l = next--;
                const uint32_t k = (uint32_t)j.kind[(size_t)l];
                if (k >= (uint32_t)kGateBase) {  // table gate: qubits < 1024 (at most 16 words)
                    const uint32_t g = k - (uint32_t)kGateBase;
                    const uint32_t t = g >= (uint32_t)kFirstTwoQubit ? (uint32_t)j.q1[(size_t)l] : 0u;
                    p->ops.push_back((3u << 28) | (g << 20) | ((uint32_t)j.q0[(size_t)l] << 10) | t);
                } else {
                    p->ops.push_back((k << 28) | ((uint32_t)j.q0[(size_t)l] << 14) |
                                     (k == 2 ? (uint32_t)j.q1[(size_t)l] : 0u));
                }
            }
            if (next >= 0) {
                const int64_t l = next--;
                double c, sn;
                angle_cos_sin(j.theta[(size_t)l], &c, &sn);
                const uint8_t kind = rotation_kind_of(c, sn);
                p->rk[i] = kind;
                const size_t ja = (size_t)j.job_as;
                if (kind == 4) {
                    p->any_general = true;
                    for (size_t k = 0; k < ja; ++k) {
                        p->p_or[k] |= j.px[(size_t)l * ja + k];
                        p->p_or[A + k] |= j.pz[(size_t)l * ja + k];
                    }
                }
                p->rc[i] = c;
                p->rs[i] = sn;
                for (size_t k = 0; k < ja; ++k) {
                    p->rp[i * 2 * A + k] = j.px[(size_t)l * ja + k];
                    p->rp[i * 2 * A + A + k] = j.pz[(size_t)l * ja + k];
                }
                p->maxw[i] = j.max_weight;  // apply_truncation follows every rotation
                p->eps[i] = j.cutoff;
                p->maxt[i] = j.max_terms;
                p->any_trim |= j.max_terms >= 0;
            }
            p->next[i] = next;
            if (next < 0) p->finishing.push_back((int)i);
        }
        p->off[S] = (uint32_t)p->ops.size();
    }

    // A sorted step's device work reported adv[slot] rotations consumed per
    // slot (1 .. wn[slot] for an advancing job, 0 otherwise): the jobs' next
    // ops and the finishing jobs follow.  False when a report is out of range.
    bool settle(StepPlan* p, const uint8_t* adv) const {
        for (size_t i = 0; i < (size_t)slots_; ++i) {
            if (state_[i] != kRunning) continue;
            const int a = adv[i], w = i < p->wn.size() ? p->wn[i] : 0;
            if ((w == 0 && a != 0) || (w > 0 && (a < 1 || a > w))) return false;
            if (w == 0) continue;  // a C job: planned
            p->next[i] = active_[i].next - a;
            if (p->next[i] < 0) p->finishing.push_back((int)i);
        }
        return true;
    }

    // The step's device work is done: draining slots (their terms left in
    // this step's collation) become free, every running job moves to its
    // planned next op, and finishing jobs' slots start draining.  Returns
    // true when nothing runs any more (the batch can then be emptied).
    bool commit(const StepPlan& p) {
        for (size_t i = 0; i < (size_t)slots_; ++i) {
            const bool waited = state_[i] == kRunning && i < p.wn.size() && p.wn[i] == 0 &&
                                !(i < p.isc.size() && p.isc[i]) && p.next[i] == active_[i].next && !p.wn.empty();
            wait_[i] = waited ? (uint8_t)std::min(255, wait_[i] + 1) : 0;
        }
        for (auto& f : fresh_) f = 0;
        for (auto& s : state_)
            if (s == kDraining) s = kFree;
        for (size_t i = 0; i < (size_t)slots_; ++i)
            if (state_[i] == kRunning) active_[i].next = p.next[i];
        for (int slot : p.finishing) {
            state_[(size_t)slot] = kDraining;
            active_[(size_t)slot] = Job();
        }
        if (running() == 0) {
            for (auto& s : state_) s = kFree;
            return true;
        }
        return false;
    }

   private:
    int as_, slots_;
    int64_t budget_;
    std::deque<Job> queue_;
    std::vector<Job> active_;  // by slot
    std::vector<int> state_;
    std::vector<uint8_t> fresh_ = std::vector<uint8_t>((size_t)slots_, 0);
    std::vector<uint8_t> wait_ = std::vector<uint8_t>((size_t)slots_, 0);  // mixed steps waited at a C op
};

}  // namespace sched
}  // namespace plh
