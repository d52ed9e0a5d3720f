"""CPU replay of the batched server's bookkeeping (tests/cpp/sched_sim.cpp):
the server's scheduler (paulib_amd/csrc/plh_sched.hpp), sum storage
(plh_storage.hip) and heap (plh_heap.hip), built for the host with
AddressSanitizer and UndefinedBehaviorSanitizer, driven through the mixed
job list of tests/test_gpu_serve.py (14 jobs through 6 slots, so queued
jobs take reused slots) with per-step term counts from the oracle's
run_program, then with device memory capped and device errors injected."""
import subprocess
from pathlib import Path

import numpy as np

from helpers import cpu
from test_gpu_serve import _mixed_jobs

ROOT = Path(__file__).resolve().parent.parent


def _suffix(prog, l):
    from paulib_amd.workloads import Program

    a = prog.array_size
    return Program(a, prog.num_qubits, prog.kind[l:], prog.q0[l:], prog.q1[l:], prog.px[l * a:], prog.pz[l * a:],
                   prog.theta[l:])


def _job_text(O, prog, opts) -> str:
    """One job: header, then per op its kind, qubits, angle, the oracle's
    term count after ops l .. end (-1 where no step ends) and its Pauli words."""
    nops = len(prog)
    kind = np.asarray(prog.kind)
    ends = set(int(l) for l in np.nonzero(kind == 3)[0]) | {0}
    lines = [f"{O.num_qubits} {len(O)} {nops} {-1 if opts.max_weight is None else opts.max_weight} "
             f"{opts.coefficient_cutoff!r}"]
    a = prog.array_size
    for l in range(nops):
        cnt = -1
        if l in ends:
            o = cpu("oracle", O)
            o.run_program(_suffix(prog, l), opts.max_weight, opts.coefficient_cutoff)
            cnt = len(o)
        words = " ".join(f"{int(prog.px[l * a + w]):x} {int(prog.pz[l * a + w]):x}" for w in range(a))
        lines.append(f"{int(kind[l])} {int(prog.q0[l])} {int(prog.q1[l])} {float(prog.theta[l])!r} {cnt} {words}")
    return "\n".join(lines)


def test_server_bookkeeping_replay(tmp_path):
    exe = tmp_path / "sched_sim"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", str(ROOT / "tests" / "cpp" / "sched_sim.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    jobs = _mixed_jobs(14, seed=3)
    text = f"{len(jobs)} 6 2\n" + "\n".join(_job_text(O, prog, opts) for O, prog, opts in jobs) + "\n"
    r = subprocess.run([str(exe)], input=text, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert r.stdout.strip().endswith("ok"), r.stdout
