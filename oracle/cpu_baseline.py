"""ORACLE -- test infrastructure only: the timed CPU baseline of bench.py.

Runs the OpenMP form of the CPU restatement (oracle.cpp's *_omp path: the
reference's level-set sweeps of omp.h:1094-1119 and the fused PCG passes of
src/solve.cpp:122-269) on one workload, as its own process so that the
OpenMP runtime starts with the thread placement given in its environment
(bench.py sets OMP_NUM_THREADS and OMP_PROC_BIND=close / OMP_PLACES=cores).
Setup (host factorization at one thread, the CPU restatement's SpTRSV
preparation and operator build) and solve are timed apart, as
benchmarks/src/benchmark.cpp:585-588 reports them; the solve time is the
median of `--repeat` solves. Prints one JSON object.

    python -m oracle.cpu_baseline --workload sddm3d_128 --threads 16 --repeat 3
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def physical_cores() -> tuple[int, int]:
    """(physical cores, logical CPUs) among the CPUs this process may run on,
    from the kernel's topology files (one core = one (package, core id))."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            cores.add((base.joinpath("physical_package_id").read_text().strip(),
                       base.joinpath("core_id").read_text().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return len(cores), len(cpus)


def baseline_threads() -> tuple[int, str]:
    """Threads for the baseline: every physical core this process may use,
    capped by OMP_NUM_THREADS when the environment sets it (a share of a
    larger host); and a sentence saying how the number came about."""
    phys, logical = physical_cores()
    env = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(env) if env.isdigit() and int(env) > 0 else 0
    t = min(phys, cap) if cap else phys
    why = f"{phys} physical cores / {logical} logical CPUs in this process's affinity"
    if cap:
        why += f", OMP_NUM_THREADS={cap}"
    return max(1, t), why


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="sddm3d_128")
    ap.add_argument("--threads", type=int, default=0, help="0: every physical core (capped by OMP_NUM_THREADS)")
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-8)
    args = ap.parse_args()

    import numpy as np

    import apxchol_amd as ax
    import oracle
    from apxchol_amd import matgen as mg

    threads, why = baseline_threads()
    if args.threads > 0:
        threads = args.threads
    A, b = mg.workload(args.workload)
    t0 = time.perf_counter()
    F = ax.factorize(A, seed=0)
    t_fact = time.perf_counter() - t0
    t0 = time.perf_counter()
    o = oracle.Oracle(A, F.export(), F.sddm)
    t_prep = time.perf_counter() - t0
    times, iters, x = [], [], None
    for _ in range(max(1, args.repeat)):
        t0 = time.perf_counter()
        x, it, _ = o.solve_omp(b, threads, tol=args.tol)
        times.append(time.perf_counter() - t0)
        iters.append(it)
    t_solve = float(np.median(times))
    true_res = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    print(json.dumps({"workload": args.workload, "n": int(A.shape[0]), "threads": threads, "threads_from": why,
                      "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
                      "omp_places": os.environ.get("OMP_PLACES", ""),
                      "setup_factorize_s": t_fact, "factor_threads": F.threads, "setup_prep_s": t_prep, "solve_s": t_solve,
                      "solve_times_s": times, "iterations": iters[0], "true_residual": true_res}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
