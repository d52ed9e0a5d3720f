#!/usr/bin/env python3
"""Benchmark: collated PauliSum product throughput + propagation-job latency.

Step = canonicalize(A * B) for two 65,536-term sums on 64 qubits (sites
non-identity w.p. 0.25, U[-1,1] coefficients, seeded), computed by the fused
product->collation path with inputs and output resident in HBM.  value =
n*m / (seconds per step) (collated product terms/s).  Propagation latency:
sequential jobs of Z_0 + Z_31 through the depth-20 64-qubit brickwork
(truncation |c| < 1e-4, weight > 8) on the same device, p50/p99 reported.
Batched products: 512 uneven independent products (C3 distribution) in a
few collations, against the plain drop-in looped over the same jobs.
Serving: open-loop Poisson streams of propagation jobs (weight-1..3
observables on 32..128 qubits, brickwork depth 5..40, cutoffs 1e-5..1e-3)
through the continuous-batching server (plh_server), once far above
capacity (1000 jobs arriving at 500/s: the completion rate while the server
is saturated is its capacity) and once at 70% of that capacity: p50/p99
latency from arrival and jobs/s, beside the plain drop-in's jobs/s on 100
jobs drawn at random from the stream and run one after another (their
results compared with the served ones).

roofline: the kernel with the largest share of the timed region, timed
with HIP events on the engine's stream during the timed steps; achieved =
its algorithmic bytes per launch / its mean launch time (the leaf kernel's
bytes are 25 B per collated term at one word per plane).
cpu_baseline: the reference's own multiply + canonicalize (oracle/_ref, its
OpenMP build) on a 65,536 x m_s slice of the same product, on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_profile(ctx):
    from paulib_amd import _lib

    n = 64
    names = C.create_string_buffer(32 * n)
    ms = (C.c_double * n)()
    launches = (C.c_int64 * n)()
    byt = (C.c_double * n)()
    cnt = C.c_int()
    _lib.call("plh_ctx_profile_read", ctx.handle, n, names, ms, launches, byt, C.byref(cnt))
    out = {}
    for i in range(cnt.value):
        name = names.raw[32 * i:32 * i + 32].split(b"\0", 1)[0].decode()
        out[name] = dict(ms=ms[i], launches=int(launches[i]), bytes=byt[i])
    return out


def serve_stream(ctx, jobs, max_jobs: int, label: str):
    """An open-loop stream through the batched server: jobs are submitted at
    their arrival times while the server steps; per job its latency from
    arrival, its expectation and final term count."""
    from paulib_amd import TruncationOptions
    from paulib_amd.serve import Server

    n_jobs = len(jobs)
    arrival, finish, res = {}, {}, {}
    order = {}
    in_flight_max = 0
    t_done = []
    with Server(ctx, array_size=2, max_jobs=max_jobs) as srv:
        t0 = time.perf_counter()
        i = 0
        last_log = t0
        while len(finish) < n_jobs:
            now = time.perf_counter() - t0
            while i < n_jobs and jobs[i][0] <= now:
                _, O, prog, cut = jobs[i]
                h = srv.submit(O, prog, TruncationOptions(max_weight=8, coefficient_cutoff=cut))
                arrival[h] = t0 + jobs[i][0]
                order[h] = i
                i += 1
            running, queued = srv.step()
            in_flight_max = max(in_flight_max, running)
            t = time.perf_counter()
            for r in srv.poll():
                finish[r.handle] = t
                t_done.append(t - t0)
                res[order[r.handle]] = (r.expectation, r.terms)
            if running == 0 and queued == 0 and i < n_jobs:
                time.sleep(max(0.0, jobs[i][0] - (time.perf_counter() - t0)))
            if t - last_log > 5.0:
                log(f"serving {label}: {len(finish)}/{n_jobs} done, {running} running, {queued} queued, "
                    f"batch {srv.batch_terms} terms, step {srv.steps}")
                last_log = t
        steps = srv.steps
    lat = np.array([finish[h] - arrival[h] for h in finish]) * 1e3
    span = max(finish.values()) - (t0 + jobs[0][0])
    # completion rate between the 10% and 90% completions: the server's rate
    # while it is saturated, without the start-up and the drain of the last
    # (longest) jobs
    t_done.sort()
    a, b = int(0.1 * n_jobs), int(0.9 * n_jobs)
    steady = (b - a) / (t_done[b] - t_done[a]) if b > a and t_done[b] > t_done[a] else 0.0
    return {"jobs": n_jobs, "jobs_per_s": n_jobs / span, "steady_jobs_per_s": steady,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)), "steps": steps,
            "max_running": in_flight_max}, res


def serving(ctx, n_jobs: int, rate: float, n_plain: int, max_jobs: int, sustain_frac: float, sustain_jobs: int):
    """C4: an open-loop Poisson stream of propagation jobs through the
    batched server at `rate` (above capacity: all slots busy, which measures
    the server's throughput), the same kind of stream at sustain_frac of the
    measured capacity (latency at a load the server keeps up with), and the
    plain drop-in (run_program per job, one after another) on n_plain jobs
    drawn at random from the first stream, whose expectations and term
    counts are compared with the served ones."""
    from paulib_amd import PauliSum, TruncationOptions, run_program
    from paulib_amd import workloads as W

    jobs = W.serving_jobs(n_jobs, rate, seed=9)
    over, served = serve_stream(ctx, jobs, max_jobs, "overload")
    log(f"serving at {rate:g}/s: {over['jobs_per_s']:.2f} jobs/s over the stream, {over['steady_jobs_per_s']:.2f} "
        f"jobs/s saturated, p50 {over['p50_ms']:.0f} ms, p99 {over['p99_ms']:.0f} ms")
    rate_s = sustain_frac * over["steady_jobs_per_s"]
    sus = {}
    if sustain_jobs > 0:
        jobs_s = W.serving_jobs(sustain_jobs, rate_s, seed=10)
        sus, _ = serve_stream(ctx, jobs_s, max_jobs, "sustained")
        log(f"serving at {rate_s:.2f}/s: {sus['jobs_per_s']:.2f} jobs/s, p50 {sus['p50_ms']:.0f} ms, "
            f"p99 {sus['p99_ms']:.0f} ms")

    rng = np.random.Generator(np.random.PCG64(2024))
    sample = sorted(rng.choice(n_jobs, size=min(n_plain, n_jobs), replace=False).tolist())
    t_plain, max_diff, term_mismatch = [], 0.0, 0
    for k in sample:
        _, O, prog, cut = jobs[k]
        d = PauliSum.from_host(ctx, O)
        t0 = time.perf_counter()
        run_program(d, prog, TruncationOptions(max_weight=8, coefficient_cutoff=cut))
        e = d.zero_state_expectation()
        t_plain.append(time.perf_counter() - t0)
        e_s, n_s = served[k]
        max_diff = max(max_diff, abs(e - e_s))
        term_mismatch += int(n_s != len(d))
        d.close()
    plain_rate = len(t_plain) / sum(t_plain)
    rots_all = [p.num_rotations() for _, _, p, _ in jobs]
    rots_smp = [jobs[k][2].num_rotations() for k in sample]
    # the plain loop's rate over the whole stream, from its per-rotation time on the sample
    plain_rate_stream = plain_rate * float(np.mean(rots_smp)) / float(np.mean(rots_all))
    log(f"serving: plain loop {plain_rate:.2f} jobs/s on {len(sample)} sampled jobs "
        f"({plain_rate_stream:.2f} scaled to the stream's rotations); served vs plain: max |<O> diff| {max_diff:.2e}, "
        f"{term_mismatch} term-count mismatches")
    return {
        "workload": (f"{n_jobs} propagation jobs, Poisson arrivals at {rate:g}/s: weight-1..3 observable on "
                     "32..128 qubits, brickwork depth 5..40, |c| cutoff 1e-5..1e-3 (log-uniform), weight <= 8"),
        "rates": ("jobs_per_s = jobs / whole stream span; steady_jobs_per_s = completions between the 10% and 90% "
                  "marks / their span (saturated server); speedup_vs_plain = steady rate / the plain loop's rate "
                  "scaled to the stream's mean rotations"),
        "mode": "continuous batching (plh_server), open loop",
        "jobs": n_jobs, "rotations_mean": float(np.mean(rots_all)), "rotations_max": int(max(rots_all)),
        "jobs_per_s": over["jobs_per_s"], "steady_jobs_per_s": over["steady_jobs_per_s"],
        "p50_ms": over["p50_ms"], "p99_ms": over["p99_ms"],
        "steps": over["steps"], "max_running": over["max_running"],
        "sustained": {"rate_per_s": rate_s, "fraction_of_capacity": sustain_frac, **sus},
        "plain_jobs_per_s": plain_rate, "plain_sample_jobs": len(sample),
        "plain_sample_rotations_mean": float(np.mean(rots_smp)),
        "plain_jobs_per_s_stream": plain_rate_stream,
        "speedup_vs_plain": over["steady_jobs_per_s"] / plain_rate_stream,
        "speedup_vs_plain_whole_stream": over["jobs_per_s"] / plain_rate_stream,
        "served_vs_plain": {"max_abs_expectation_diff": max_diff, "term_count_mismatches": term_mismatch},
    }, served


def batched_products(ctx, count: int, budget_pairs: int, solo_pairs: int):
    """C3: `count` independent collated products of uneven sizes (terms
    log-uniform in [64, 65536], qubits from {16, 32, 64, 128}), three ways:
    - batched, one output per job (plh_batch_multiply_canonical_each): jobs
      in groups whose outputs fit at once (budget_pairs pairs); inside a
      group products of at least solo_pairs pairs run alone and the rest
      share one collation;
    - one collation per group (plh_batch_multiply_canonical, per array
      size): every product of the group in one job-keyed collation;
    - the plain drop-in (plh_multiply_canonical) looped over the jobs.
    Outputs are reused from group to group (results are discarded)."""
    from paulib_amd import PauliSum, batch_multiply_canonical, batch_multiply_canonical_each
    from paulib_amd import workloads as W

    spec = W.batch_products(count)
    t0 = time.perf_counter()
    L, R, pairs = [], [], []
    for nq, n, m, seed in spec:
        L.append(PauliSum.from_host(ctx, W.random_sum(n, nq, 0.75, "uniform", seed)))
        R.append(PauliSum.from_host(ctx, W.random_sum(m, nq, 0.75, "uniform", seed + 1)))
        pairs.append(n * m)
    log(f"batch: {count} jobs, {sum(pairs):.4g} pairs, inputs built in {time.perf_counter() - t0:.1f} s")

    def grouped(keys):
        groups = []
        for a in keys:
            idx = [k for k in range(count) if a is None or L[k].array_size == a]
            cur, tot = [], 0
            for k in idx:
                if cur and tot + pairs[k] > budget_pairs:
                    groups.append(cur)
                    cur, tot = [], 0
                cur.append(k)
                tot += pairs[k]
            if cur:
                groups.append(cur)
        return groups

    if solo_pairs <= 0:  # the engine's threshold: the context's tuned one, else its default
        solo_pairs = ctx.tuning()["solo_pairs"] or (1 << 22)
    each_groups = grouped([None])
    one_groups = grouped([1, 2])
    terms = 0

    # a group's outputs are released before the next group runs (their
    # memory stays mapped in the context's heap for the next outputs)
    def run_each():
        nonlocal terms
        terms = 0
        for g in each_groups:
            o = batch_multiply_canonical_each([L[k] for k in g], [R[k] for k in g], solo_pairs=solo_pairs)
            terms += sum(len(x) for x in o)
            for x in o:
                x.close()

    # One output sum at a time: the largest 128-qubit jobs alone need most of
    # the HBM, so the output of one array size goes back to the context's
    # heap before the other's jobs run.
    def run_one():
        out, cur = None, None
        for g in one_groups:
            a = L[g[0]].array_size
            if a != cur:
                if out is not None:
                    out.close()
                out, cur = PauliSum(ctx, a, 64 * a), a
            out, off = batch_multiply_canonical([L[k] for k in g], [R[k] for k in g], out)
        out.close()

    def run_plain(order):
        from paulib_amd import _lib

        out, cur = None, None
        for k in order:
            a = L[k].array_size
            if a != cur:
                if out is not None:
                    out.close()
                out, cur = PauliSum(ctx, a, 64 * a), a
            _lib.call("plh_multiply_canonical", L[k].handle, R[k].handle, out.handle)
        out.close()

    def timed(f, *a):
        ctx.synchronize()
        t0 = time.perf_counter()
        f(*a)
        ctx.synchronize()
        return time.perf_counter() - t0

    # the products below the solo threshold (the ones batching can change):
    # their share of the plain loop, and the batch against the loop on them
    small = sorted((k for k in range(count) if pairs[k] < solo_pairs and L[k].array_size <= 2),
                   key=lambda k: L[k].array_size)

    def run_small_batched():
        for a in (1, 2):
            ks = [k for k in small if L[k].array_size == a]
            for g0 in range(0, len(ks), 4096):
                o = batch_multiply_canonical_each([L[k] for k in ks[g0:g0 + 4096]], [R[k] for k in ks[g0:g0 + 4096]],
                                                  solo_pairs=solo_pairs)
                for x in o:
                    x.close()

    run_each()  # warm-up: grows the context's heap to the workload's working set
    t_each = timed(run_each)
    t_one = timed(run_one)
    t_plain = timed(run_plain, sorted(range(count), key=lambda k: L[k].array_size))
    run_small_batched()
    t_small_batch = timed(run_small_batched)
    t_small_plain = timed(run_plain, small)
    for s_ in L + R:
        s_.close()
    ctx.trim()
    total = sum(pairs)
    solo = sum(1 for k in range(count) if pairs[k] >= solo_pairs or L[k].array_size > 2)
    log(f"batch: {len(each_groups)} groups ({solo} products alone, {count - solo} batched), {t_each:.3f} s "
        f"({total / t_each:.4g} pairs/s); one collation per group {t_one:.3f} s; plain loop {t_plain:.3f} s "
        f"({total / t_plain:.4g} pairs/s)")
    log(f"batch: the {len(small)} products below {solo_pairs} pairs take {t_small_plain:.3f} s of the plain loop "
        f"({100 * t_small_plain / t_plain:.1f}%); batched {t_small_batch:.3f} s ({t_small_plain / t_small_batch:.2f}x)")
    return {
        "workload": (f"{count} independent collated products, terms per sum log-uniform in [64, 65536], "
                     "qubits from {16, 32, 64, 128}, sites non-I w.p. 0.25, U[-1,1] FP64"),
        "jobs": count, "pairs": total, "collated_terms": terms, "groups": len(each_groups),
        "solo_pairs": solo_pairs, "products_alone": solo,
        "batch_s": t_each, "pairs_per_s": total / t_each, "jobs_per_s": count / t_each,
        "one_collation_per_group_s": t_one, "one_collation_groups": len(one_groups),
        "plain_loop_s": t_plain, "plain_pairs_per_s": total / t_plain, "speedup_vs_plain": t_plain / t_each,
        "small_products": {"jobs": len(small), "pairs": int(sum(pairs[k] for k in small)),
                           "plain_loop_s": t_small_plain, "share_of_plain_loop": t_small_plain / t_plain,
                           "batch_s": t_small_batch, "speedup_vs_plain": t_small_plain / t_small_batch},
    }


def clifford_gates(ctx, log2_terms: int, reps: int):
    """The table-driven Clifford kernel on a resident 64-qubit sum: per gate
    kind, the kernel time from HIP events on the engine's stream and the
    algorithmic bytes (touched words read and written, flags read and
    written) against the HBM peak."""
    from paulib_amd import PauliSum, _lib
    from paulib_amd.host import HostSum

    n = 1 << log2_terms
    rng = np.random.default_rng(12)
    x = rng.integers(0, 1 << 63, size=(1, n), dtype=np.uint64) << np.uint64(1)
    z = rng.integers(0, 1 << 63, size=(1, n), dtype=np.uint64)
    h = HostSum(1, 64, x, z, np.zeros(n, np.uint8), rng.uniform(-1, 1, n))
    d = PauliSum.from_host(ctx, h)
    out = {"workload": f"{n} terms on 64 qubits, resident", "terms": n}
    for name, c, t in [("H", 5, 0), ("X", 7, 0), ("CZ", 3, 40), ("iSWAP", 12, 50)]:
        d.apply_gate(name, c, t)
        ctx.synchronize()
        _lib.call("plh_ctx_profile", ctx.handle, 1)
        for _ in range(reps):
            d.apply_gate(name, c, t)
        ctx.synchronize()
        prof = kernel_profile(ctx)["clifford_table"]
        _lib.call("plh_ctx_profile", ctx.handle, 0)
        ms = prof["ms"] / prof["launches"]
        gbs = prof["bytes"] / prof["launches"] / (ms / 1e3) / 1e9
        out[name] = {"ms": ms, "GB/s": gbs, "frac": gbs * 1e9 / HBM_PEAK}
        log(f"gate {name}: {ms:.3f} ms per launch, {gbs:.0f} GB/s")
    d.close()
    return out


def raw_products(ctx, reps: int):
    """The uncombined all-pairs product (PauliSum::multiply, k_mul_raw) at the
    shapes of the reference's own GPU benchmark (500 qubits, 8 words per
    plane; bench/GpuMulBench.cpp) and at 65536 x 65536 on 64 qubits: kernel
    time from HIP events on the engine's stream, pairs/s and the HBM roofline
    (output written, operands read: 16 * words + 9 bytes per term)."""
    from paulib_amd import PauliSum, _lib
    from paulib_amd import workloads as W

    shapes = [(1000, 100, 500), (1000, 500, 500), (1000, 1000, 500), (1000, 2000, 500), (1000, 5000, 500),
              (10000, 10000, 500), (65536, 65536, 64)]
    out = {}
    for n, m, nq in shapes:
        if nq == 64:
            A, B = W.product_65536x65536_q64(seed=7, n=n, m=m)
        else:
            A, B = W.random_words_sum(n, nq, 101), W.random_words_sum(m, nq, 102)
        a, b = PauliSum.from_host(ctx, A), PauliSum.from_host(ctx, B)
        r = max(1, reps if n * m <= 10 ** 8 else 2)
        lhs = a.copy()
        lhs.multiply(b)  # warm-up: output storage
        lhs.close()
        ctx.synchronize()
        _lib.call("plh_ctx_profile", ctx.handle, 1)
        for _ in range(r):
            lhs = a.copy()
            lhs.multiply(b)
            lhs.close()
        ctx.synchronize()
        prof = kernel_profile(ctx)["mul_raw"]
        _lib.call("plh_ctx_profile", ctx.handle, 0)
        ms = prof["ms"] / prof["launches"]
        gbs = prof["bytes"] / prof["launches"] / (ms / 1e3) / 1e9
        key = f"{n}x{m}_q{nq}"
        out[key] = {"ms": ms, "pairs_per_s": n * m / (ms / 1e3), "GB/s": gbs, "frac": gbs * 1e9 / HBM_PEAK,
                    "words": A.array_size}
        log(f"raw product {key}: {ms:.4f} ms, {n * m / (ms / 1e3):.3e} pairs/s, {gbs:.0f} GB/s")
        a.close(), b.close()
    ctx.trim()
    return out


def collate_leg(ctx, reps: int):
    """sort_and_combine of a materialised 1000 x m product on 100 qubits
    (2 words per plane), timed alone as the reference's
    bench/CollateBench.cpp times it; operands as random_pauli_sum (uniform
    words, random flags)."""
    from paulib_amd import PauliSum
    from paulib_amd import workloads as W

    A = W.random_words_sum(1000, 100, 42)
    a = PauliSum.from_host(ctx, A)
    out = {}
    for m in (100, 500, 1000, 2000):
        b = PauliSum.from_host(ctx, W.random_words_sum(m, 100, 123))
        ts = []
        for r in range(reps + 1):
            d = a.copy()
            d.multiply(b)
            ctx.synchronize()
            t0 = time.perf_counter()
            d.sort_and_combine()  # synchronises
            if r > 0:
                ts.append(time.perf_counter() - t0)
            d.close()
        out[f"1000x{m}_q100"] = {"ms_best": min(ts) * 1e3, "ms_median": float(np.median(ts)) * 1e3,
                                 "terms_per_s": 1000 * m / float(np.median(ts))}
        log(f"collate 1000x{m}: {min(ts) * 1e3:.3f} ms best")
        b.close()
    a.close()
    return out


def propagation_trim_leg(ctx):
    """propagate(Z_0, mixed-field Ising circuit n = 40, 10 layers) with
    max_terms 2^14, 2^16, 2^18 (the reference's bench/PropagationBench.cpp):
    ms per gate."""
    from paulib_amd import PauliSum, TruncationOptions, propagate
    from paulib_amd import workloads as W

    O, circ = W.mixed_field_ising(40, 10)
    out = {"workload": "mixed-field Ising, 40 qubits, 10 layers, 1190 rotations, observable Z_0"}
    for log2m in (14, 16, 18):
        for r in range(2):  # the first run warms the workspace
            d = PauliSum.from_host(ctx, O)
            t0 = time.perf_counter()
            propagate(d, circ, TruncationOptions(max_terms=1 << log2m))
            e = d.zero_state_expectation()
            t = time.perf_counter() - t0
            terms = len(d)
            d.close()
        out[f"trim_2^{log2m}"] = {"s": t, "ms_per_gate": 1e3 * t / len(circ), "final_terms": terms,
                                  "expectation": e}
        log(f"propagation trim 2^{log2m}: {t:.3f} s, {1e3 * t / len(circ):.3f} ms/gate, {terms} terms")
    return out


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A, B, target_s: float, gpu_c2=None, gpu_trim=None):
    """Reference multiply + canonicalize on a rhs slice sized for ~target_s,
    and the reference's propagation legs in full (the depth-20 brickwork
    configuration and the PropagationBench trims), each compared with the
    GPU's result of the same workload when given."""
    import oracle

    kind = "reference" if oracle.available("reference") else "oracle"
    from paulib_amd.host import HostSum

    def run(m):
        Bs = HostSum(B.array_size, B.num_qubits, B.x[:, :m], B.z[:, :m], B.flags[:m], B.value[:m])
        a = oracle.CpuSum.from_host(kind, A)
        b = oracle.CpuSum.from_host(kind, Bs)
        t0 = time.perf_counter()
        a.multiply(b)
        a.canonicalize()
        return time.perf_counter() - t0

    m = 8
    t = run(m)  # warm-up and calibration
    m = int(min(len(B), max(8, m * target_s / max(t, 1e-3))))
    t = run(m)
    threads = os.environ.get("OMP_NUM_THREADS") or str(os.cpu_count())
    if kind == "reference":
        import ctypes

        h = ctypes.CDLL(str(oracle.REF_LIB))
        threads = str(h.ref_num_threads())
    # the 1000 x 1000, 32-qubit case in full: best of 5 after 3 warm-ups
    from paulib_amd import workloads as W

    A0, B0 = W.product_1000x1000_q32(seed=1)
    c0 = {}
    for op in ("product", "commutator"):
        ts = []
        for r in range(8):
            a = oracle.CpuSum.from_host(kind, A0)
            b = oracle.CpuSum.from_host(kind, B0)
            t0 = time.perf_counter()
            if op == "product":
                a.multiply(b)
                a.canonicalize()
            else:
                a.commutator(b).canonicalize()
            if r >= 3:
                ts.append(time.perf_counter() - t0)
        c0[op + "_ms_best"] = min(ts) * 1e3
    # the uncombined product at one shape of the reference's GPU benchmark
    # (bench/GpuMulBench.cpp: 1000 x 1000 on 500 qubits), best of 5
    Ar, Br = W.random_words_sum(1000, 500, 101), W.random_words_sum(1000, 500, 102)
    ts = []
    for r in range(8):
        a = oracle.CpuSum.from_host(kind, Ar)
        b = oracle.CpuSum.from_host(kind, Br)
        t0 = time.perf_counter()
        a.multiply(b)
        if r >= 3:
            ts.append(time.perf_counter() - t0)
    c0["raw_product_1000x1000_q500_ms_best"] = min(ts) * 1e3
    # sort_and_combine of the 1000 x 1000 product on 100 qubits (CollateBench)
    Ac, Bc = W.random_words_sum(1000, 100, 42), W.random_words_sum(1000, 100, 123)
    ts = []
    for r in range(4):
        a = oracle.CpuSum.from_host(kind, Ac)
        a.multiply(oracle.CpuSum.from_host(kind, Bc))
        t0 = time.perf_counter()
        a.sort_and_combine()
        if r >= 1:
            ts.append(time.perf_counter() - t0)
    c0["collate_1000x1000_q100_ms_best"] = min(ts) * 1e3
    # propagate with max_terms 2^14..2^18 on the mixed-field Ising circuit (PropagationBench)
    O, circ = W.mixed_field_ising(40, 10)
    for log2m in (14, 16, 18):
        o = oracle.CpuSum.from_host(kind, O)
        t0 = time.perf_counter()
        o.propagate(circ, None, 0.0, 1 << log2m)
        tp = time.perf_counter() - t0
        c0[f"propagation_trim_2^{log2m}_ms_per_gate"] = 1e3 * tp / len(circ)
        g = (gpu_trim or {}).get(f"trim_2^{log2m}")
        if g:
            e = o.zero_state_expectation()
            c0[f"propagation_trim_2^{log2m}_agrees"] = {
                "terms_cpu": len(o), "terms_gpu": g["final_terms"], "expectation_cpu": e,
                "expectation_gpu": g["expectation"], "abs_diff": abs(e - g["expectation"])}
    # the depth-20 brickwork propagation (Z0 + Z31, 64 qubits) in full
    Oc, prog, tr = W.brickwork_q64_d20()
    o = oracle.CpuSum.from_host(kind, Oc)
    t0 = time.perf_counter()
    o.run_program(prog, tr["max_weight"], tr["cutoff"])
    e = o.zero_state_expectation()
    tc2 = time.perf_counter() - t0
    c2 = {"ms": tc2 * 1e3, "terms": len(o), "expectation": e}
    if gpu_c2:
        c2.update(terms_gpu=gpu_c2["terms"], expectation_gpu=gpu_c2["expectation"],
                  abs_diff=abs(e - gpu_c2["expectation"]), speedup_gpu=tc2 * 1e3 / gpu_c2["p50_ms"])
    log(f"cpu baseline: C2 {tc2:.1f} s, {len(o)} terms, <O> = {e:.10e}")
    return {
        "c0": c0,
        "c2": c2,
        "value": len(A) * m / t,
        "unit": "collated product terms/s",
        "cores": int(threads),
        "kind": "reference" if kind == "reference" else "port",
        "cpu": cpu_model(),
        "build": ("reference headers via oracle/ref_bridge.cpp: g++ -std=c++20 -O3 -march=x86-64-v3 -ffast-math "
                  "-fopenmp (the reference's flags, -march=native replaced so the library built elsewhere runs here)"
                  if kind == "reference" else "oracle/pauli_oracle.c: gcc -O2 -ffp-contract=off, one thread"),
        "sample": f"{len(A)} x {m} rhs slice of the 65536 x 65536 q64 product (multiply + canonicalize), {t:.2f} s",
    }


def _cpu_kind():
    import oracle

    return "reference" if oracle.available("reference") else "oracle"


def cpu_batch_sample(count: int, target_s: float, gpu):
    """C3 on the host cores: products of the same batch in a seeded random
    order (those of at most 2^24 pairs, so that a sample holds many of
    them), multiply + canonicalize each, until target_s has passed."""
    import oracle
    from paulib_amd import workloads as W

    kind = _cpu_kind()
    spec = W.batch_products(count)
    rng = np.random.Generator(np.random.PCG64(77))
    order = [int(k) for k in rng.permutation(len(spec)) if spec[k][1] * spec[k][2] <= 1 << 24]
    pairs, t, jobs = 0, 0.0, 0
    for k in order:
        nq, n, m, seed = spec[k]
        a = oracle.CpuSum.from_host(kind, W.random_sum(n, nq, 0.75, "uniform", seed))
        b = oracle.CpuSum.from_host(kind, W.random_sum(m, nq, 0.75, "uniform", seed + 1))
        t0 = time.perf_counter()
        a.multiply(b)
        a.canonicalize()
        t += time.perf_counter() - t0
        pairs += n * m
        jobs += 1
        del a, b
        if t >= target_s:
            break
    rate = pairs / t
    log(f"cpu baseline: C3 sample {jobs} products, {pairs:.4g} pairs in {t:.1f} s ({rate:.3g} pairs/s)")
    return {"jobs": jobs, "pairs": pairs, "s": t, "pairs_per_s": rate,
            "sample": (f"{jobs} of the {count} products in a seeded random order (those of <= 2^24 pairs), "
                       "multiply + canonicalize each"),
            "gpu_batch_pairs_per_s": gpu["pairs_per_s"], "speedup_gpu_batch": gpu["pairs_per_s"] / rate}


def cpu_serving_sample(n_jobs: int, rate_hz: float, target_s: float, gpu, served):
    """C4 on the host cores: jobs of the same stream (the plain loop's seeded
    sample order) through the reference's run_program and
    zero_state_expectation, until target_s has passed; the rate is scaled to
    the stream's mean rotations like the plain GPU loop's, and each job's
    expectation and term count are compared with the served ones."""
    import oracle
    from paulib_amd import workloads as W

    kind = _cpu_kind()
    jobs = W.serving_jobs(n_jobs, rate_hz, seed=9)
    rng = np.random.Generator(np.random.PCG64(2024))
    order = rng.permutation(n_jobs).tolist()
    t, done, rots, max_diff, mism = 0.0, 0, [], 0.0, 0
    for k in order:
        _, O, prog, cut = jobs[k]
        o = oracle.CpuSum.from_host(kind, O)
        t0 = time.perf_counter()
        o.run_program(prog, 8, cut)
        e = o.zero_state_expectation()
        t += time.perf_counter() - t0
        done += 1
        rots.append(prog.num_rotations())
        if k in served:
            e_s, n_s = served[k]
            max_diff = max(max_diff, abs(e - e_s))
            mism += int(n_s != len(o))
        del o
        if t >= target_s:
            break
    rate = done / t
    rots_all = [p.num_rotations() for _, _, p, _ in jobs]
    rate_stream = rate * float(np.mean(rots)) / float(np.mean(rots_all))
    log(f"cpu baseline: C4 sample {done} jobs in {t:.1f} s ({rate:.3f} jobs/s, {rate_stream:.3f} scaled to the "
        f"stream); served vs cpu: max |<O> diff| {max_diff:.2e}, {mism} term-count mismatches")
    return {"jobs": done, "s": t, "jobs_per_s": rate, "sample_rotations_mean": float(np.mean(rots)),
            "jobs_per_s_stream": rate_stream,
            "sample": (f"{done} jobs of the {n_jobs}-job stream in a seeded random order, run_program + "
                       "zero_state_expectation each (max_weight 8, the job's cutoff)"),
            "served_steady_jobs_per_s": gpu["steady_jobs_per_s"],
            "speedup_served": gpu["steady_jobs_per_s"] / rate_stream,
            "served_vs_cpu": {"max_abs_expectation_diff": max_diff, "term_count_mismatches": mism}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--m", type=int, default=65536)
    ap.add_argument("--prop-jobs", type=int, default=5)
    ap.add_argument("--serve-jobs", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=500.0)
    ap.add_argument("--serve-plain", type=int, default=100)
    ap.add_argument("--serve-sustain", type=float, default=0.7)
    ap.add_argument("--serve-sustain-jobs", type=int, default=400)
    ap.add_argument("--serve-slots", type=int, default=256)
    ap.add_argument("--batch-jobs", type=int, default=512)
    ap.add_argument("--batch-pairs", type=int, default=1 << 31)
    ap.add_argument("--solo-pairs", type=int, default=0)  # 0: the context's tuned threshold
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--gate-log2-terms", type=int, default=26)
    ap.add_argument("--commutator-steps", type=int, default=2)
    ap.add_argument("--raw-reps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-propagation", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune-cache", default=os.environ.get("PLH_TUNE_CACHE",
                                                          str(ROOT / "paulib_amd" / "_build" / "tune_cache.txt")))
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "traffic_product_65536x65536_q64.json"))
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    from paulib_amd import Context, PauliSum, TruncationOptions, _lib, run_program
    from paulib_amd import workloads as W

    log(f"inputs: {args.n} x {args.m} sums on 64 qubits")
    A, B = W.product_65536x65536_q64(seed=7, n=args.n, m=args.m)
    ctx = Context(0)
    tuning = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuning = ctx.autotune(args.tune_cache)
        log(f"autotune: {tuning} ({time.perf_counter() - t0:.1f} s, cache {args.tune_cache})")
    dA, dB = PauliSum.from_host(ctx, A), PauliSum.from_host(ctx, B)  # first upload: allocation
    dA.close(), dB.close()
    ctx.synchronize()
    t0 = time.perf_counter()
    dA, dB = PauliSum.from_host(ctx, A), PauliSum.from_host(ctx, B)
    ctx.synchronize()
    t_upload = time.perf_counter() - t0
    out = PauliSum(ctx, 1, 64)
    pairs = args.n * args.m

    def step():
        _lib.call("plh_multiply_canonical", dA.handle, dB.handle, out.handle)

    for i in range(args.warmup):
        t0 = time.perf_counter()
        step()
        ctx.synchronize()
        log(f"warmup {i}: {time.perf_counter() - t0:.3f} s, {len(out)} terms")
    _lib.call("plh_ctx_profile", ctx.handle, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    prof = kernel_profile(ctx)
    _lib.call("plh_ctx_profile", ctx.handle, 0)
    t_out = len(out)
    log(f"timed: {dt * 1e3:.2f} ms/step, {t_out} collated terms")
    for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"]):
        log(f"  {k:20s} {v['ms'] / args.steps:9.3f} ms/step  {v['launches'] / args.steps:6.1f} launches/step")

    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    per_launch_s = dom["ms"] / dom["launches"] / 1e3
    per_launch_bytes = dom["bytes"] / dom["launches"]
    achieved = per_launch_bytes / per_launch_s / 1e9 if per_launch_bytes else 0.0
    # HBM bytes per launch of the dominant kernel: a profile constant from the
    # FETCH_SIZE / WRITE_SIZE counter runs (rocprofv3 cannot run inside the
    # bench), only for the workload it was measured on
    traffic, traffic_src = None, None
    tp = Path(args.traffic)
    if tp.exists() and args.n == 65536 and args.m == 65536:
        try:
            tj = json.loads(tp.read_text())
            traffic = tj.get(dom_name)
            traffic_src = f"{tp.relative_to(ROOT) if tp.is_relative_to(ROOT) else tp}: {tj.get('_about', '')}"
        except Exception:
            traffic = None
    b_alg_job = (args.n + args.m) * 25 + t_out * 25
    result = {
        "metric": "PauliSum product terms/s (collated) + p50/p99 propagation-job latency",
        "value": pairs / dt,
        "unit": "collated product terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 words + f64 coefficients",
        "data": "synthetic, seeded (sites non-I w.p. 0.25, U[-1,1] FP64)",
        "config": {"workload": f"product_{args.n}x{args.m}_q64 canonicalize(A*B)", "pairs": pairs,
                   "collated_terms": t_out, "qubits": 64, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved * 1e9 / HBM_PEAK, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "launch_ms": per_launch_s * 1e3, "alg_bytes_per_launch": per_launch_bytes},
        "job_roofline": {"alg_bytes": b_alg_job, "frac": b_alg_job / dt / HBM_PEAK},
        "kernels_ms_per_step": {k: v["ms"] / args.steps for k, v in prof.items()},
        "tuning": tuning,
    }
    # PCIe-inclusive rate of the host-buffer drop-in (multiply_canonical_on_gpu):
    # measured upload of both operands, the device time, and the output's
    # download at the rate measured on a range of 2^27 terms (pageable host
    # memory, as the drop-in writes into the caller's vectors).
    cnt = min(1 << 27, t_out)
    out.range_to_host(0, min(1 << 20, t_out))
    t0 = time.perf_counter()
    rng_host = out.range_to_host(0, cnt)
    t_rng = time.perf_counter() - t0
    d2h = cnt * 25 / t_rng
    up = PauliSum.from_host(ctx, rng_host)  # allocates; the timed upload reuses the storage
    ctx.synchronize()
    t0 = time.perf_counter()
    up.upload(rng_host)
    ctx.synchronize()
    h2d = cnt * 25 / (time.perf_counter() - t0)
    up.close()
    del rng_host
    t_pcie = t_upload + dt + t_out * 25 / d2h
    result["pcie_inclusive"] = {
        "value": pairs / t_pcie, "unit": "collated product terms/s", "upload_s": t_upload, "device_s": dt,
        "download_s_est": t_out * 25 / d2h, "d2h_GBps": d2h / 1e9, "h2d_GBps": h2d / 1e9,
        "note": f"download time from a measured {cnt}-term range download; not the headline value"}
    log(f"pcie-inclusive: upload {t_upload:.3f} s, D2H {d2h / 1e9:.1f} GB/s, H2D {h2d / 1e9:.1f} GB/s, "
        f"{pairs / t_pcie:.3e} terms/s")
    out.close()
    if args.commutator_steps > 0:
        cout = PauliSum(ctx, 1, 64)
        _lib.call("plh_commutator_canonical", dA.handle, dB.handle, 0, cout.handle)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.commutator_steps):
            _lib.call("plh_commutator_canonical", dA.handle, dB.handle, 0, cout.handle)
        ctx.synchronize()
        dt_c = (time.perf_counter() - t0) / args.commutator_steps
        result["commutator"] = {"workload": f"canonicalize([A, B]) at {args.n} x {args.m}, 64 qubits",
                                "ms_per_step": dt_c * 1e3, "pairs_per_s": pairs / dt_c,
                                "collated_terms": len(cout), "steps": args.commutator_steps}
        log(f"commutator: {dt_c * 1e3:.1f} ms, {len(cout)} collated terms")
        cout.close()
    dA.close(), dB.close()
    ctx.trim()

    # the 1000 x 1000, 32-qubit case (A*B and [A, B], canonicalized): device
    # time with resident operands and the time with upload and download
    A0, B0 = W.product_1000x1000_q32(seed=1)
    c0 = {"workload": "product_1000x1000_q32 canonicalize(A*B) and canonicalize([A,B])"}
    for name, fn in [("product", "plh_multiply_canonical"), ("commutator", "plh_commutator_canonical")]:
        dev_t, inc_t = [], []
        for r in range(8):
            t0 = time.perf_counter()
            a, b = PauliSum.from_host(ctx, A0), PauliSum.from_host(ctx, B0)
            ctx.synchronize()
            t1 = time.perf_counter()
            o = PauliSum(ctx, 1, 32)
            if fn == "plh_multiply_canonical":
                _lib.call(fn, a.handle, b.handle, o.handle)
            else:
                _lib.call(fn, a.handle, b.handle, 0, o.handle)
            ctx.synchronize()
            t2 = time.perf_counter()
            h = o.to_host()
            t3 = time.perf_counter()
            if r >= 3:
                dev_t.append(t2 - t1)
                inc_t.append(t3 - t0)
            a.close(), b.close(), o.close()
        c0[name] = {"device_ms_best": min(dev_t) * 1e3, "device_ms_median": float(np.median(dev_t)) * 1e3,
                    "with_transfers_ms_median": float(np.median(inc_t)) * 1e3, "collated_terms": len(h),
                    "pairs_per_s": 1e6 / float(np.median(dev_t))}
        log(f"c0 {name}: {min(dev_t) * 1e3:.3f} ms device, {float(np.median(inc_t)) * 1e3:.3f} ms with transfers")
    result["c0"] = c0

    if not args.no_propagation and args.prop_jobs > 0:
        O, prog, tr = W.brickwork_q64_d20()
        lat = []
        for j in range(args.prop_jobs + 1):
            d = PauliSum.from_host(ctx, O)
            t0 = time.perf_counter()
            run_program(d, prog, TruncationOptions(max_weight=tr["max_weight"], coefficient_cutoff=tr["cutoff"]))
            e = d.zero_state_expectation()
            t = time.perf_counter() - t0
            terms_c2 = len(d)
            log(f"propagation job {j}: {t * 1e3:.1f} ms, {terms_c2} terms, <O> = {e:.10f}")
            if j > 0:  # the first job warms the workspace
                lat.append(t)
            d.close()
        lat = np.array(lat) * 1e3
        result["propagation"] = {"workload": "brickwork_q64_d20 (Z0+Z31, depth 20, w<=8, |c|>=1e-4)",
                                 "mode": "sequential jobs, plain drop-in", "jobs": len(lat),
                                 "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                                 "jobs_per_s": float(1e3 / lat.mean()), "expectation": e, "terms": terms_c2}
    if args.raw_reps > 0:
        result["raw_products"] = raw_products(ctx, args.raw_reps)
        result["collate"] = collate_leg(ctx, 5)
    if not args.no_propagation:
        result["propagation_trim"] = propagation_trim_leg(ctx)
    if args.gate_log2_terms > 0:
        result["clifford_gates"] = clifford_gates(ctx, args.gate_log2_terms, 20)
    if args.batch_jobs > 0:
        result["batched_products"] = batched_products(ctx, args.batch_jobs, args.batch_pairs, args.solo_pairs)
    if args.serve_jobs > 0:
        result["serving"], served = serving(ctx, args.serve_jobs, args.serve_rate, args.serve_plain,
                                            args.serve_slots, args.serve_sustain, args.serve_sustain_jobs)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(A, B, args.cpu_seconds, result.get("propagation"),
                                              result.get("propagation_trim"))
        if args.batch_jobs > 0:
            result["cpu_baseline"]["c3"] = cpu_batch_sample(args.batch_jobs, args.cpu_seconds,
                                                            result["batched_products"])
        if args.serve_jobs > 0:
            result["cpu_baseline"]["c4"] = cpu_serving_sample(args.serve_jobs, args.serve_rate,
                                                              1.5 * args.cpu_seconds, result["serving"], served)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
