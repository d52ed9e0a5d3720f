"""Continuous-batching propagation server (plh_server_* in include/paulib_hip.h).

Jobs are (observable, Clifford + rotation program, truncation) triples, as
for ``run_program``.  Admitted jobs share one device-resident sum tagged by
job slot; every ``step()`` advances each running job by one rotation (and
the Clifford gates before it) for all jobs at once, and jobs that finish
report ``zero_state_expectation`` of their propagated observable.  New jobs
are admitted at step boundaries while slots and HBM allow, so requests that
arrive while others run join the batch (continuous batching).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import Context, TruncationOptions, _f64, _i32, _i64, _u8, _u64
from .host import HostSum, words_for


@dataclass
class JobResult:
    handle: int
    expectation: float
    terms: int
    latency_s: float


class Server:
    """plh_server: up to ``max_jobs`` jobs in flight on one device."""

    def __init__(self, ctx: Context, array_size: int = 2, max_jobs: int = 256,
                 hbm_budget_bytes: int = 0, keep_terms: bool = False):
        self.ctx = ctx
        self.array_size = array_size
        self.keep_terms = keep_terms
        h = C.c_void_p()
        _lib.call("plh_server_create", ctx.handle, array_size, max_jobs, int(hbm_budget_bytes),
                  1 if keep_terms else 0, C.byref(h))
        self._h = h
        self._nq = {}

    def close(self):
        if self._h:
            _lib.lib().plh_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ jobs
    def submit(self, O: HostSum, prog, opts: TruncationOptions | None = None) -> int:
        """Queue a job; O must have distinct (x, z, imag) keys (canonical)."""
        opts = opts or TruncationOptions()
        nq = int(O.num_qubits)
        a = words_for(nq)
        if prog.num_qubits != nq or prog.array_size != a:
            raise ValueError("program and observable qubit counts differ")
        x = np.ascontiguousarray(O.x[:a], dtype=np.uint64)
        z = np.ascontiguousarray(O.z[:a], dtype=np.uint64)
        if O.array_size > a and (np.any(O.x[a:]) or np.any(O.z[a:])):
            raise ValueError("observable has bits beyond its qubit count")
        f = np.ascontiguousarray(O.flags, dtype=np.uint8)
        v = np.ascontiguousarray(O.value, dtype=np.float64)
        kind = np.ascontiguousarray(prog.kind, dtype=np.int32)
        q0 = np.ascontiguousarray(prog.q0, dtype=np.int64)
        q1 = np.ascontiguousarray(prog.q1, dtype=np.int64)
        px = np.ascontiguousarray(prog.px, dtype=np.uint64)
        pz = np.ascontiguousarray(prog.pz, dtype=np.uint64)
        th = np.ascontiguousarray(prog.theta, dtype=np.float64)
        mw, cut, mt = opts.args()
        h = C.c_int64()
        _lib.call("plh_server_submit", self._h, len(f), _u64(x), _u64(z), _u8(f), _f64(v), nq, len(kind),
                  _i32(kind), _i64(q0), _i64(q1), _u64(px), _u64(pz), _f64(th), mw, cut, mt, C.byref(h))
        self._nq[int(h.value)] = nq
        return int(h.value)

    def step(self) -> tuple[int, int]:
        """Advance every running job by one rotation; (running, queued)."""
        r, q = C.c_int(), C.c_int()
        _lib.call("plh_server_step", self._h, C.byref(r), C.byref(q))
        return r.value, q.value

    def poll(self, max_results: int = 4096) -> list[JobResult]:
        hs = np.zeros(max_results, np.int64)
        ex = np.zeros(max_results, np.float64)
        tm = np.zeros(max_results, np.int64)
        lt = np.zeros(max_results, np.float64)
        n = C.c_int()
        _lib.call("plh_server_poll", self._h, max_results, _i64(hs), _f64(ex), _i64(tm), _f64(lt), C.byref(n))
        return [JobResult(int(hs[i]), float(ex[i]), int(tm[i]), float(lt[i])) for i in range(n.value)]

    def take_terms(self, result: JobResult) -> HostSum:
        """Final sum of a finished job (keep_terms servers), canonical order."""
        nq = self._nq.pop(result.handle)
        a = words_for(nq)
        m = result.terms
        x = np.zeros((a, m), np.uint64)
        z = np.zeros((a, m), np.uint64)
        f = np.zeros(m, np.uint8)
        v = np.zeros(m, np.float64)
        _lib.call("plh_server_take_terms", self._h, result.handle, _u64(x), _u64(z), _u8(f), _f64(v))
        return HostSum(array_size=a, num_qubits=nq, x=x, z=z, flags=f, value=v)

    def drain(self) -> list[JobResult]:
        """Step until every submitted job has finished; results in completion order."""
        out = []
        while True:
            running, queued = self.step()
            out.extend(self.poll())
            if running == 0 and queued == 0:
                break
        out.extend(self.poll())
        return out

    @property
    def batch_terms(self) -> int:
        return int(_lib.lib().plh_server_batch_terms(self._h))

    @property
    def steps(self) -> int:
        return int(_lib.lib().plh_server_steps(self._h))
