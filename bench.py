#!/usr/bin/env python3
"""Benchmark of the MI355X apxchol solve path on one GPU.

Workload (default): 3-D 7-point grid SDDM 128^3 (n = 2,097,152), edge weights
exp(N(0,1)) (seed 1), diagonal excess U[0,1e-3) * degree (seed 2), one N(0,1)
right-hand side (seed 0), factor seed 0, tol 1e-8. A step is one full PCG
solve with the right-hand side already resident in HBM. The headline value is
the mixed path (fp32 factor values and fp32 sweeps, fp64 SpMV / CG); the plain
fp64-sweep path is timed beside it. Every solve is graded by the true residual
||b - A x|| / ||b|| recomputed in fp64 on the host, and the first solution is
compared with the CPU oracle's fp64 solve of the same system.

The roofline object covers the kernel with the largest total time, measured
live with HIP events around its launches during the timed solves.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0


def sweep_bytes(info: dict, backward: bool, vec_bytes: int, fp16: bool) -> float:
    """Algorithmic HBM bytes of one sweep: every stored off-diagonal once
    (index + value), the row pointers, the diagonal, the batch table, the
    right-hand side read, the output written once and gathered once."""
    m, off = info["m"], info["offdiag"]
    nb = info["bck_batches"] if backward else info["fwd_batches"]
    ent = 6 if fp16 else 8
    rhs = (vec_bytes if backward else 8)
    extra = (4 * m) if (fp16 and backward) else 0
    return off * ent + 4 * (m + 1) + 4 * m + 12 * nb + m * rhs + 2 * m * vec_bytes + extra


def spmv_bytes(info: dict) -> float:
    n, nnz = info["n"], info["op_nnz"]
    return nnz * (4 + (4 if info["op_fp32"] else 8)) + 4 * (n + 1) + 16 * n


def percentile(xs, q):
    s = sorted(xs)
    if not s:
        return float("nan")
    k = min(len(s) - 1, max(0, int(round(q / 100.0 * (len(s) - 1)))))
    return s[k]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sddm3d_128")
    ap.add_argument("--iters-per-check", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solve path is single-GPU (--gpus {args.gpus} is not supported)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import apxchol_amd as ax
    from apxchol_amd import matgen as mg

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    A, b = mg.workload(args.workload)
    t_gen = time.perf_counter() - t0
    n = A.shape[0]
    print(f"[bench] workload {args.workload}: n={n} nnz={A.nnz} generated in {t_gen:.1f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    F = ax.factorize(A, seed=0)
    t_fact = time.perf_counter() - t0
    print(f"[bench] host factorization {t_fact:.1f}s nnz(L)={F.nnz} rounds={F.nrounds}", file=sys.stderr, flush=True)

    bd = torch.tensor(b, device=dev)
    xd = torch.empty_like(bd)
    bnorm = float(np.linalg.norm(b))

    def run(precision: str):
        t0 = time.perf_counter()
        s = ax.Solver(A, factor=F, precision=precision, iters_per_check=args.iters_per_check)
        t_setup = time.perf_counter() - t0
        for _ in range(args.warmup):
            s.solve(bd, out=xd)
        torch.cuda.synchronize()
        s.set_timing(True)
        s.take_timing()
        lat, iters, res = [], [], []
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            t1 = time.perf_counter()
            r = s.solve(bd, out=xd)
            lat.append(time.perf_counter() - t1)
            iters.append(r.iterations)
            res.append(r.residual)
        torch.cuda.synchronize()
        total = time.perf_counter() - t_start
        timing = s.take_timing()
        s.set_timing(False)
        x = xd.cpu().numpy()
        true_res = float(np.linalg.norm(b - A @ x) / bnorm)
        info = dict(s.info)
        s.close()
        return dict(total=total, lat=lat, iters=iters, rec_res=res, x=x, true_res=true_res, timing=timing,
                    info=info, setup=t_setup)

    mixed = run("fp32")
    print(f"[bench] mixed: {args.steps / mixed['total']:.2f} solves/s iters={mixed['iters'][0]} "
          f"true_res={mixed['true_res']:.2e}", file=sys.stderr, flush=True)
    plain = None if args.no_fp64 else run("fp64")
    if plain:
        print(f"[bench] fp64:  {args.steps / plain['total']:.2f} solves/s iters={plain['iters'][0]} "
              f"true_res={plain['true_res']:.2e}", file=sys.stderr, flush=True)

    # Roofline of the dominant kernel of the mixed solve (HIP events, live).
    info = mixed["info"]
    fp16 = bool(info["fp16_storage"])
    tm = mixed["timing"]
    kb = {"fwd_sweep": sweep_bytes(info, False, 4, fp16), "bck_sweep": sweep_bytes(info, True, 4, fp16),
          "spmv": spmv_bytes(info)}
    dom = max(kb, key=lambda k: tm[k][0])
    ms_total, launches = tm[dom]
    avg_s = ms_total / 1e3 / max(launches, 1)
    achieved = kb[dom] / avg_s / 1e9
    per_kernel = {k: {"avg_us": tm[k][0] * 1e3 / max(tm[k][1], 1), "launches": tm[k][1],
                      "bytes": kb[k], "gbps": kb[k] / (tm[k][0] / 1e3 / max(tm[k][1], 1)) / 1e9 if tm[k][1] else 0.0}
                  for k in kb}
    per_kernel["vector"] = {"avg_us": tm["vector"][0] * 1e3 / max(tm["vector"][1], 1), "launches": tm["vector"][1]}

    cpu = None
    x_cpu_err = None
    if not args.no_cpu_baseline:
        import oracle
        fac = F.export()
        t0 = time.perf_counter()
        o = oracle.Oracle(A, fac, F.sddm)
        t_or_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        xo, ito, _ = o.solve(b, tol=1e-8)
        t_cpu = time.perf_counter() - t0
        x_cpu_err = float(np.linalg.norm(mixed["x"] - xo) / np.linalg.norm(xo))
        cpu = {"value": 1.0 / t_cpu, "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"one full fp64 PCG solve of {args.workload} (n={n}) by the one-thread CPU oracle "
                         f"({ito} iterations, {t_cpu:.1f}s; oracle setup {t_or_setup:.1f}s not counted)"}
        print(f"[bench] cpu oracle: {t_cpu:.1f}s/solve, {ito} iterations", file=sys.stderr, flush=True)

    value = args.steps / mixed["total"]
    out = {
        "metric": "solves/s + p50/p99 latency at rel-res 1e-8 (mixed vs FP64); SpTRSV HBM GB/s",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["total"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 sweeps + fp64 CG",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "nnz": int(A.nnz), "tol": 1e-8, "factor_seed": 0,
                   "factor_nnz": F.nnz, "levels": info["levels"], "rhs": "N(0,1) seed 0"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None, "algorithmic_bytes": kb[dom],
                     "avg_launch_us": avg_s * 1e6},
        "cpu_baseline": cpu,
        "detail": {
            "mixed": {"solves_per_s": value, "p50_ms": 1e3 * percentile(mixed["lat"], 50),
                      "p99_ms": 1e3 * percentile(mixed["lat"], 99), "iterations": mixed["iters"][0],
                      "true_residual": mixed["true_res"], "err_vs_cpu_fp64": x_cpu_err,
                      "setup_s": mixed["setup"]},
            "fp64": None if not plain else {
                "solves_per_s": args.steps / plain["total"], "p50_ms": 1e3 * percentile(plain["lat"], 50),
                "p99_ms": 1e3 * percentile(plain["lat"], 99), "iterations": plain["iters"][0],
                "true_residual": plain["true_res"], "setup_s": plain["setup"]},
            "mixed_over_fp64": None if not plain else value / (args.steps / plain["total"]),
            "kernels_mixed": per_kernel,
            "host_factorization_s": t_fact,
            "solver_info": info,
        },
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
