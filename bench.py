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

Beside the headline, blocked k-right-hand-side solves (k = 32, fp32 sweeps,
autotuned SpMV form and block width) are timed on the same system, against
its plain fp64 single-vector rate, and on the 64^3 anisotropic grid
(weights 1/10/100 x exp(N(0,0.5))); the R-MAT scale-22 Laplacian (edge
factor 16, weights U[1,10), per-component mean-projected N(0,1) right-hand
side) is timed in mixed and fp64 single solves and one blocked k = 32 solve;
the serving mix's factors of at most 300k rows, and 16 small 2-D grids, are
solved one after another and as one grouped solve (many small factors in one
launch per kernel kind); the serving mix runs 8 resident factors
(2-D grids 1e4..1e6, 3-D grids to 96^3, R-MAT scales 16..20) under Poisson
arrivals and reports p50 / p99 latency and solves/s.

The CPU baseline is the OpenMP form of the CPU restatement of the reference's
solve path (level-set sweeps with thin levels on one thread and fat levels
split over the team, fused PCG passes with thread-ordered partial sums),
timed first, in a child process, on every physical core this process may use
(capped by OMP_NUM_THREADS), bound and unbound; setup and the median of three
solves are reported apart. The one-thread oracle's solve is the checker.

The timed solves replay each PCG iteration as a captured HIP graph. After
them, a timing pass of --timing-solves solves launches eagerly with HIP events
around every kernel on the solve's stream; the roofline object covers the
single kernel (sweep_levels, sweep_tail or the SpMV) with the largest total
time there: algorithmic bytes per launch (from the solver's layout: per level
off-diagonals x (4 + value width) + rows x (8 + input + output width) + 12 per
batch; SpMV nnz x (4 + value width) + 4(n+1) + 16n) over its average launch
duration.

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


def spmv_bytes(info: dict) -> float:
    n, nnz = info["n"], info["op_nnz"]
    return nnz * (4 + (4 if info["op_fp32"] else 8)) + 4 * (n + 1) + 16 * n


def kernel_source_digest() -> str:
    """sha256 over the HIP kernel sources and the shared device layouts: a
    traffic profile applies to the kernels it was measured on only."""
    import hashlib
    h = hashlib.sha256()
    csrc = ROOT / "apxchol_amd" / "csrc"
    for f in sorted(list((csrc / "kernels").glob("*")) + [csrc / "device.hpp"]):
        if f.suffix in (".hip", ".hpp"):
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()


def percentile(xs, q):
    s = sorted(xs)
    if not s:
        return float("nan")
    k = min(len(s) - 1, max(0, int(round(q / 100.0 * (len(s) - 1)))))
    return s[k]


def run_serving(args, ax, mg, np) -> dict:
    """The serving mix: 8 resident factors, Poisson arrivals at a fixed rate,
    one N(0,1) right-hand side per request with tol drawn from {1e-6, 1e-8,
    1e-10}; mixed precision, continuous batching with the single lanes and
    narrow blocks on (and, for comparison, the main blocks alone). Every
    result is graded by its true residual."""
    import threading
    t0 = time.perf_counter()
    mats, solvers = [], []
    for name in mg.SERVING_FACTORS:
        L, _ = mg.workload(name)
        F = ax.factorize(L, seed=0)
        S = ax.Solver(L, factor=F, precision="fp32")
        mats.append((name, L, F.sddm))
        solvers.append(S)
        print(f"[bench] serving: factor {name} (n={L.shape[0]}) resident after {time.perf_counter() - t0:.1f}s",
              file=sys.stderr, flush=True)
    t_setup = time.perf_counter() - t0
    print(f"[bench] serving: {len(mats)} factors resident ({t_setup:.1f}s)", file=sys.stderr, flush=True)
    rng = np.random.default_rng(6)
    tols = [1e-6, 1e-8, 1e-10]
    plan = []
    for i in range(args.serving_requests):
        f = int(rng.integers(len(mats)))
        plan.append((f, float(tols[int(rng.integers(3))]), float(rng.exponential(1.0 / args.serving_qps))))
    rhs = [mg.request_rhs(mats[f][1], mats[f][2], 1000 + i) for i, (f, _, _) in enumerate(plan)]

    def offered(lanes: bool) -> dict:
        """The Poisson arrivals into a server over the resident solvers (a
        solver serves one server at a time, so each pass has its own); lanes:
        the single lane and the narrow block on (the defaults), else the main
        block alone."""
        label = "serving" if lanes else "serving, main block only"
        srv = ax.Server(slots=args.serving_slots, iters_per_check=2, single_lane=lanes,
                        narrow_slots=args.serving_narrow if lanes else 0)
        fids = [srv.add(S) for S in solvers]
        ids = [None] * len(plan)
        errors = []

        def produce():
            try:
                for i, (f, tol, gap) in enumerate(plan):
                    time.sleep(gap)
                    ids[i] = srv.submit(fids[f], rhs[i], tol)
            except Exception as e:   # a refused submission ends the run with its reason
                errors.append(e)

        th = threading.Thread(target=produce)
        t1 = time.perf_counter()
        th.start()
        th.join()
        if errors:
            srv.close()
            raise errors[0]
        print(f"[bench] {label}: {len(plan)} requests submitted", file=sys.stderr, flush=True)
        worst = 0.0
        lat = [[] for _ in mats]
        for i, (f, tol, _) in enumerate(plan):
            # A request the server has not finished within 120 s ends the
            # measurement with that reason (the serving section records it).
            try:
                x, r = srv.wait(ids[i], timeout=120)
            except TimeoutError:
                st = srv.stats()
                srv.close()
                raise RuntimeError(f"request {i} (factor {mats[f][0]}) not done after 120 s; "
                                   f"{st['completed']} of {st['submitted']} completed")
            L = mats[f][1]
            tr = float(np.linalg.norm(rhs[i] - L @ x) / np.linalg.norm(rhs[i]))
            worst = max(worst, tr / tol)
            lat[f].append(r.timings["pcg_ms"])
            if (i + 1) % 64 == 0:
                print(f"[bench] {label}: {i + 1} of {len(plan)} results graded ({time.perf_counter() - t1:.1f}s)",
                      file=sys.stderr, flush=True)
        wall = time.perf_counter() - t1
        st = srv.stats()
        srv.close()
        print(f"[bench] {label}: {st['solves_per_s']:.1f} solves/s p50 {st['p50_ms']:.1f} ms "
              f"p99 {st['p99_ms']:.1f} ms ({st['single_lane_completed']} on single lanes)", file=sys.stderr,
              flush=True)
        return {"solves_per_s": st["solves_per_s"], "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                "mean_ms": st["mean_ms"], "wall_s": wall, "worst_true_residual_over_tol": worst,
                "memory_used_bytes": st["memory_used"], "single_lane_completed": st["single_lane_completed"],
                "by_factor": {mats[f][0]: {"requests": len(v), "p50_ms": percentile(v, 50), "p99_ms": percentile(v, 99)}
                              for f, v in enumerate(lat) if v}}

    main = offered(True)
    out = {"factors": [m[0] for m in mats], "requests": len(plan), "qps_offered": args.serving_qps,
           "slots": args.serving_slots, "narrow_slots": args.serving_narrow, "single_lane": True, **main,
           "setup_s": t_setup}
    if args.serving_blocks_only:
        out["main_block_only"] = offered(False)
    # Capacity: the same requests, `serving_capacity_reps` times over,
    # submitted at once to a fresh server over the same resident solvers (the
    # first server released them), so the rate is what the device sustains
    # rather than the offered load (repeating the set keeps the last
    # stragglers from deciding the rate).
    cap = ax.Server(slots=args.serving_slots, iters_per_check=2, single_lane=True, narrow_slots=args.serving_narrow)
    cfids = [cap.add(S) for S in solvers]
    cplan = [(i, f, tol) for _ in range(max(1, args.serving_capacity_reps)) for i, (f, tol, _) in enumerate(plan)]
    t2 = time.perf_counter()
    cids = [cap.submit(cfids[f], rhs[i], tol) for i, f, tol in cplan]
    cworst = 0.0
    for (i, f, tol), rid in zip(cplan, cids):
        x, r = cap.wait(rid, timeout=300)
        L = mats[f][1]
        cworst = max(cworst, float(np.linalg.norm(rhs[i] - L @ x) / np.linalg.norm(rhs[i])) / tol)
    cst = cap.stats()
    cap.close()
    out["capacity"] = {"requests": len(cplan), "submitted_at_once": True, "solves_per_s": cst["solves_per_s"],
                       "p50_ms": cst["p50_ms"], "p99_ms": cst["p99_ms"], "wall_s": time.perf_counter() - t2,
                       "worst_true_residual_over_tol": cworst,
                       "single_lane_completed": cst["single_lane_completed"]}
    print(f"[bench] serving capacity (all {len(cplan)} requests at once): {cst['solves_per_s']:.1f} solves/s "
          f"p50 {cst['p50_ms']:.1f} ms p99 {cst['p99_ms']:.1f} ms", file=sys.stderr, flush=True)
    return out


def run_cpu_baseline(args) -> dict | None:
    """The OpenMP CPU path (oracle/cpu_baseline.py) in a child process on
    every physical core this process may use (capped by OMP_NUM_THREADS):
    once with the threads bound close to one core each (OMP_PROC_BIND=close,
    OMP_PLACES=cores) and once unbound, setup and the median of three solves
    timed apart in each; the faster placement is the baseline. A failure
    there is reported and leaves the baseline out."""
    import subprocess
    sys.path.insert(0, str(ROOT))
    from oracle.cpu_baseline import baseline_threads
    threads, why = baseline_threads()
    runs = {}
    for label, extra in (("bound", {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}),
                         ("unbound", {"OMP_PROC_BIND": "false", "OMP_PLACES": ""})):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), **extra)
        if not extra["OMP_PLACES"]:
            env.pop("OMP_PLACES")
        print(f"[bench] cpu baseline ({label}): OpenMP solve of {args.workload} on {threads} threads ({why}) starts",
              file=sys.stderr, flush=True)
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--workload", args.workload,
                            "--threads", str(threads), "--repeat", "3"], cwd=str(ROOT), env=env,
                           capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            print(f"[bench] cpu baseline ({label}) failed ({r.returncode}): {r.stderr[-400:]}", file=sys.stderr,
                  flush=True)
            continue
        c = json.loads(r.stdout.strip().splitlines()[-1])
        print(f"[bench] cpu baseline ({label}): {c['solve_s']:.3f}s/solve on {c['threads']} threads, "
              f"{c['iterations']} iterations (setup {c['setup_factorize_s']:.1f}s + {c['setup_prep_s']:.1f}s)",
              file=sys.stderr, flush=True)
        runs[label] = c
    if not runs:
        return None
    best = min(runs, key=lambda k: runs[k]["solve_s"])
    out = dict(runs[best])
    out["placement"] = best
    out["solve_s_by_placement"] = {k: v["solve_s"] for k, v in runs.items()}
    return out


def run_batched(args, ax, mg, np, torch, dev, name, A, F, plain_rate, seed0) -> dict:
    """One blocked PCG of k right-hand sides (fp32 sweeps, graph-replayed
    iterations) after the autotuner has picked the SpMV form and block width;
    N(0,1) right-hand sides of seeds seed0 .. seed0+k-1, projected onto the
    range where the matrix is a Laplacian. Without plain_rate the plain fp64
    single-vector solve of the same system (seed 0 right-hand side) is timed
    for the ratio. Every column is graded by its true residual."""
    if A is None:
        A, _ = mg.workload(name)
        F = ax.factorize(A, seed=0)
    n = A.shape[0]
    k = args.batch_k
    B = np.stack([mg.request_rhs(A, F.sddm, seed0 + c) for c in range(k)])
    Bd = torch.tensor(B, device=dev)
    s = ax.Solver(A, factor=F, precision="fp32", iters_per_check=args.iters_per_check)
    tune = s.autotune(max_block_width=64)
    s.solve_batch(Bd)
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        Xb, rb = s.solve_batch(Bd)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / reps
    Xh = Xb.cpu().numpy()
    worst = max(float(np.linalg.norm(B[c] - A @ Xh[c]) / np.linalg.norm(B[c])) for c in range(k))
    s.close()
    if plain_rate is None:
        p = ax.Solver(A, factor=F, precision="fp64", iters_per_check=args.iters_per_check)
        bd = torch.tensor(mg.request_rhs(A, F.sddm, 0), device=dev)
        xd = torch.empty_like(bd)
        p.solve(bd, out=xd)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            p.solve(bd, out=xd)
        torch.cuda.synchronize()
        plain_rate = 3 / (time.perf_counter() - t0)
        p.close()
    rate = k / tb
    print(f"[bench] blocked {name} k={k}: {rate:.1f} solves/s ({rate / plain_rate:.1f}x the fp64 single-vector rate)",
          file=sys.stderr, flush=True)
    return {"n": n, "k": k, "solves_per_s": rate, "ms_per_block": tb * 1e3,
            "iterations_max": max(r.iterations for r in rb), "worst_true_residual": worst,
            "fp64_single_solves_per_s": plain_rate, "over_fp64_single": rate / plain_rate,
            "autotune": {"spmv": "sell" if tune["spmv_sell"] else f"csr lanes {tune['spmv_lanes']}",
                         "block_width": tune["block_width"]}}


# The reference's own published result cells for unit 3-D grid Laplacians
# (results/cells/grids/grid3d_{side}__apxchol_v1__bg_tree_vec_pool___t16__{cpu,gpu}.json):
# iterations and solve time (wall minus setup) of its GPU path (CUDA dataflow
# sweeps, RTX 4090 Laptop GPU) and its CPU path (16 threads, Ryzen 9 7945HX).
REFERENCE_CELLS = {
    200: {"gpu": {"iterations": 30, "solve_s": 0.7292422}, "cpu": {"iterations": 31, "solve_s": 4.493158}},
    250: {"gpu": {"iterations": 32, "solve_s": 1.460878}, "cpu": {"iterations": 31, "solve_s": 10.11121}},
}


def run_reference_cell(args, ax, mg, np, torch, dev, side: int) -> dict:
    """A reference result cell's workload: the unit 3-D grid Laplacian of the
    given side with the reference benchmark's right-hand side (mt19937_64
    normal draws, b = L g, centred, normalised) and factor seed 42, solved to
    1e-8 in mixed precision and in fp64 (median of 3 timed solves after a
    warm-up), beside the cell's published iterations and solve time."""
    sys.path.insert(0, str(ROOT / "tests"))
    import refcells
    t0 = time.perf_counter()
    print(f"[bench] reference cell grid3d_{side}: generation and host factorization start", file=sys.stderr,
          flush=True)
    A = mg.grid_laplacian((side, side, side))
    b = refcells.make_rhs(A)
    F = ax.factorize(A, seed=refcells.FACTOR_SEED)
    print(f"[bench] reference cell grid3d_{side}: factor ready after {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    bd = torch.tensor(b, device=dev)
    out = {"n": int(A.shape[0]), "published": REFERENCE_CELLS.get(side)}
    for prec in ("fp32", "fp64"):
        s = ax.Solver(A, factor=F, precision=prec)
        x = torch.empty_like(bd)
        s.solve(bd, tol=1e-8, maxiter=500, out=x)
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            t1 = time.perf_counter()
            r = s.solve(bd, tol=1e-8, maxiter=500, out=x)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t1)
        s.close()
        xs = x.cpu().numpy()
        key = "mixed" if prec == "fp32" else "fp64"
        out[key] = {"solve_s": float(np.median(ts)), "iterations": r.iterations,
                    "graded_residual": refcells.graded_residual(A, b, xs)}
        print(f"[bench] reference cell grid3d_{side} {key}: {1e3 * out[key]['solve_s']:.1f} ms, "
              f"{r.iterations} iterations", file=sys.stderr, flush=True)
    pub = REFERENCE_CELLS.get(side)
    if pub:
        out["mixed_speedup_vs_published_gpu"] = pub["gpu"]["solve_s"] / out["mixed"]["solve_s"]
    return out


def run_second_workload(args, ax, mg, np, torch, dev) -> dict:
    """The R-MAT configuration (default scale 22, mean-projected per component
    N(0,1) right-hand side, tol 1e-8): single solves in mixed precision and
    in fp64 (graph replay, one warm-up then --second-steps timed solves), and
    one blocked k-right-hand-side solve against the fp64 single-solve rate."""
    name = args.second_workload
    t0 = time.perf_counter()
    print(f"[bench] {name}: generation and host factorization start", file=sys.stderr, flush=True)
    A, b = mg.workload(name)
    F = ax.factorize(A, seed=0)
    t_fact = time.perf_counter() - t0
    print(f"[bench] {name}: n={A.shape[0]} nnz={A.nnz} factor ready after {t_fact:.1f}s", file=sys.stderr, flush=True)
    bd = torch.tensor(b, device=dev)
    xd = torch.empty_like(bd)
    out = {"n": int(A.shape[0]), "nnz": int(A.nnz), "factor_nnz": F.nnz, "generation_and_factorization_s": t_fact}
    for prec, label in (("fp32", "mixed"), ("fp64", "fp64")):
        s = ax.Solver(A, factor=F, precision=prec)
        s.solve(bd, out=xd)
        torch.cuda.synchronize()
        lat, iters = [], []
        for _ in range(args.second_steps):
            t1 = time.perf_counter()
            r = s.solve(bd, out=xd)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t1)
            iters.append(r.iterations)
        x = xd.cpu().numpy()
        tr = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
        out[label] = {"solves_per_s": len(lat) / sum(lat), "p50_ms": 1e3 * percentile(lat, 50),
                      "p99_ms": 1e3 * percentile(lat, 99), "iterations": iters[0], "true_residual": tr}
        s.close()
        print(f"[bench] {name} {label}: {out[label]['solves_per_s']:.2f} solves/s iters={iters[0]} true_res={tr:.2e}",
              file=sys.stderr, flush=True)
    out["mixed_over_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    if args.batch_k > 0:
        out["batched_mixed"] = run_batched(args, ax, mg, np, torch, dev, name, A, F, out["fp64"]["solves_per_s"], 100)
    return out


def run_group(args, ax, mg, np, torch, dev, which: str) -> dict:
    """Many small factors in one launch, mixed precision, one N(0,1)
    right-hand side each (projected per component on a Laplacian), solved
    one after another with their own solves and as one grouped solve (every
    kernel kind of an iteration one launch over all members); both timed
    after a warm-up, graph replay. which = "serving": the serving mix's
    factors of at most 300k rows (different sizes and kinds); "grids16": 16
    2-D 100 x 100 grids with exp(N(0,1)) weights of seeds 1..16."""
    if which == "serving":
        # the mix's factors above 300k rows (grid2d_1000, grid3d_96, rmat_20) stay out
        gen = [(nm, lambda nm=nm: mg.workload(nm)[0]) for nm in mg.SERVING_FACTORS
               if nm not in ("grid2d_1000", "grid3d_96", "rmat_20")]
    else:
        gen = [(f"grid2d_100_w{k}", lambda k=k: mg.grid_laplacian((100, 100), "lognormal", 1.0, seed=k))
               for k in range(1, 17)]
    names, solvers, bs, mats = [], [], [], []
    for nm, make in gen:
        A = make()
        i = len(names)
        names.append(nm)
        F = ax.factorize(A, seed=0)
        solvers.append(ax.Solver(A, factor=F, precision="fp32"))
        b = mg.request_rhs(A, F.sddm, 500 + i)
        bs.append(torch.tensor(b, device=dev))
        mats.append((A, b))
    g = ax.Group(solvers)
    reps = args.group_reps

    def own_round():
        return [s.solve(b) for s, b in zip(solvers, bs)]

    own_round()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        own = own_round()
    torch.cuda.synchronize()
    t_own = (time.perf_counter() - t0) / reps
    g.solve(bs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        xs, rs = g.solve(bs)
    torch.cuda.synchronize()
    t_grp = (time.perf_counter() - t0) / reps
    worst = max(float(np.linalg.norm(b - A @ x.cpu().numpy()) / np.linalg.norm(b)) for (A, b), x in zip(mats, xs))
    same = all(torch.equal(x.view(torch.int64), o.x.view(torch.int64)) for x, o in zip(xs, own))
    launches = sum(7 + sum(s.info["level_launches"]) + 2 * sum(s.info["tail_runs"]) for s in solvers)
    out = {"factors": names, "n": [int(A.shape[0]) for A, _ in mats], "one_after_another_ms": 1e3 * t_own,
           "group_ms": 1e3 * t_grp, "solves_per_s_one_after_another": len(names) / t_own,
           "solves_per_s_group": len(names) / t_grp, "speedup": t_own / t_grp,
           "iterations": [r.iterations for r in rs], "worst_true_residual": worst, "same_bits_as_own_solves": same,
           "launches_per_iteration_group": g.launches_per_iteration,
           "launches_per_iteration_own_sum": launches}
    g.close()
    for s in solvers:
        s.close()
    print(f"[bench] group ({which}) of {len(names)} small factors: {1e3 * t_grp:.1f} ms vs {1e3 * t_own:.1f} ms "
          f"one after another ({t_own / t_grp:.2f}x), bits equal: {same}", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sddm3d_128")
    ap.add_argument("--iters-per-check", type=int, default=2, help="as the library default (apx_options_default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--timing-solves", type=int, default=2)
    ap.add_argument("--batch-k", type=int, default=32, help="right-hand sides of the blocked measurements (0: skip)")
    ap.add_argument("--batch-workload", default="aniso3d_64",
                    help="second blocked workload (default: the 64^3 anisotropic grid; empty: none)")
    ap.add_argument("--second-workload", default="rmat_22",
                    help="R-MAT configuration timed beside the headline (single mixed / fp64, blocked k; empty: none)")
    ap.add_argument("--second-steps", type=int, default=3)
    ap.add_argument("--group-reps", type=int, default=5,
                    help="grouped solves of the serving mix's small factors against their own solves (0: skip)")
    ap.add_argument("--serving-requests", type=int, default=256, help="requests of the serving measurement (0: skip)")
    ap.add_argument("--serving-qps", type=float, default=25.0)
    ap.add_argument("--serving-slots", type=int, default=32)
    ap.add_argument("--reference-cell", type=int, default=200,
                    help="side of the unit 3-D grid of a reference result cell to solve (0: skip)")
    ap.add_argument("--serving-narrow", type=int, default=4, help="columns of each factor's narrow block")
    ap.add_argument("--serving-capacity-reps", type=int, default=4,
                    help="times the request set is submitted at once for the capacity pass")
    ap.add_argument("--serving-blocks-only", type=int, default=1,
                    help="also run the arrivals on the main blocks alone (no single lane, no narrow block)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run 3 plain device copies of 256 MiB (calibration of the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solve path is single-GPU (--gpus {args.gpus} is not supported)",
              file=sys.stderr)
        return 2

    # The CPU baseline runs first, as a child process, before this process
    # touches the GPU and while nothing else competes for the host cores.
    cpu_run = None if args.no_cpu_baseline else run_cpu_baseline(args)

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
    print(f"[bench] host factorization of {args.workload} starts", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    F = ax.factorize(A, seed=0)
    t_fact = time.perf_counter() - t0
    print(f"[bench] host factorization {t_fact:.1f}s on {F.threads} threads nnz(L)={F.nnz} rounds={F.nrounds}", file=sys.stderr, flush=True)

    bd = torch.tensor(b, device=dev)
    xd = torch.empty_like(bd)
    bnorm = float(np.linalg.norm(b))

    def run(precision: str):
        t0 = time.perf_counter()
        s = ax.Solver(A, factor=F, precision=precision, iters_per_check=args.iters_per_check)
        t_setup = time.perf_counter() - t0
        for _ in range(args.warmup):
            s.solve(bd, out=xd)
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
        x = xd.cpu().numpy()
        true_res = float(np.linalg.norm(b - A @ x) / bnorm)
        # Kernel timing pass (HIP events around every kernel on the solve's
        # stream; the solves launch eagerly while it is on).
        s.set_timing(True)
        s.take_timing()
        torch.cuda.synchronize()
        for _ in range(args.timing_solves):
            s.solve(bd, out=xd)
        torch.cuda.synchronize()
        timing = s.take_timing()
        s.set_timing(False)
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

    # Roofline of the dominant single kernel of the mixed solve, from the
    # event-timed pass: algorithmic bytes per launch / average launch time.
    info = mixed["info"]
    tm = mixed["timing"]
    lb, ll = info["level_bytes"], info["level_launches"]
    tb, tr = info["tail_bytes"], info["tail_runs"]
    per_launch_bytes = {
        "sweep_levels": (lb[0] + lb[1]) / max(ll[0] + ll[1], 1),
        "sweep_tail": (tb[0] + tb[1]) / max(tr[0] + tr[1], 1),
        "spmv": spmv_bytes(info),
    }
    per_kernel = {}
    for k, bpl in per_launch_bytes.items():
        ms, cnt = tm[k]
        avg = ms / 1e3 / cnt if cnt else float("nan")
        per_kernel[k] = {"avg_us": avg * 1e6, "launches": cnt, "total_ms": ms, "bytes_per_launch": bpl,
                         "gbps": bpl / avg / 1e9 if cnt else 0.0}
    for k in ("sweep_couple", "vector", "fwd_sweep", "bck_sweep"):
        ms, cnt = tm[k]
        per_kernel[k] = {"avg_us": ms * 1e3 / max(cnt, 1), "launches": cnt, "total_ms": ms}
    dom = max(per_launch_bytes, key=lambda k: tm[k][0])
    achieved = per_kernel[dom]["gbps"]
    # HBM bytes per launch of that kernel from the PMC passes recorded with
    # the profiles (FETCH_SIZE / WRITE_SIZE, copy-calibrated), when present.
    # Only a profile measured on these kernel sources counts (its digest).
    traffic = None
    tf = ROOT / "profiles" / f"traffic_{args.workload}.json"
    if tf.exists():
        tj = json.loads(tf.read_text())
        if tj.get("kernel") == dom and tj.get("kernel_sources_sha256") == kernel_source_digest():
            traffic = tj["bytes_per_launch"]
    # Whole-sweep SpTRSV rate (level launches + tail runs, gaps included).
    sweep_gbps = {}
    for di, k in ((0, "fwd_sweep"), (1, "bck_sweep")):
        ms, cnt = tm[k]
        sweep_gbps[k] = (lb[di] + tb[di]) / (ms / 1e3 / max(cnt, 1)) / 1e9 if cnt else 0.0

    # Blocked k-RHS solves (mixed): the 128^3 system itself, against its plain
    # single-vector fp64 rate, and the 64^3 anisotropic block workload.
    batched = None
    if args.batch_k > 0:
        plain_rate = (args.steps / plain["total"]) if plain else None
        batched = {args.workload: run_batched(args, ax, mg, np, torch, dev, args.workload, A, F, plain_rate, 100)}
        if args.batch_workload and args.batch_workload != args.workload:
            batched[args.batch_workload] = run_batched(args, ax, mg, np, torch, dev, args.batch_workload, None, None,
                                                       None, 0)

    second = None
    if args.second_workload:
        second = run_second_workload(args, ax, mg, np, torch, dev)
    cells = None
    if args.reference_cell > 0:
        cells = {f"grid3d_{args.reference_cell}": run_reference_cell(args, ax, mg, np, torch, dev,
                                                                     args.reference_cell)}
    group = None if args.group_reps <= 0 else {w: run_group(args, ax, mg, np, torch, dev, w)
                                               for w in ("serving", "grids16")}

    # The serving mix is reported beside the headline: a failure there is
    # recorded in its section (with the server's reason) and does not void
    # the measured single-solve line.
    serving = None
    if args.serving_requests > 0:
        try:
            serving = run_serving(args, ax, mg, np)
        except Exception as e:
            print(f"[bench] serving failed: {e}", file=sys.stderr, flush=True)
            serving = {"error": str(e)}

    cpu = None
    x_cpu_err = None
    if not args.no_cpu_baseline:
        # The checker: the one-thread oracle's fp64 solve of the same system.
        import oracle
        o = oracle.Oracle(A, F.export(), F.sddm)
        t0 = time.perf_counter()
        xo, ito, _ = o.solve(b, tol=1e-8)
        t_one = time.perf_counter() - t0
        x_cpu_err = float(np.linalg.norm(mixed["x"] - xo) / np.linalg.norm(xo))
        print(f"[bench] cpu oracle (one thread, the checker): {t_one:.1f}s/solve, {ito} iterations",
              file=sys.stderr, flush=True)
        if cpu_run is not None:
            c = cpu_run
            cpu = {"value": 1.0 / c["solve_s"], "unit": "solves/s", "cores": c["threads"], "kind": "port",
                   "sample": f"the OpenMP CPU restatement (level-set sweeps, fused passes) solving {args.workload} "
                             f"(n={n}) to 1e-8 in fp64, median of {len(c['solve_times_s'])} solves "
                             f"({c['iterations']} iterations, true residual {c['true_residual']:.2e}) on "
                             f"{c['threads']} threads ({c['threads_from']}); median solve by thread placement "
                             + ", ".join(f"{k} {v:.3f}s" for k, v in c["solve_s_by_placement"].items())
                             + f" (bound: OMP_PROC_BIND=close, OMP_PLACES=cores), the faster one taken; "
                             f"setup timed apart: host factorization "
                             f"{c['setup_factorize_s']:.2f}s ({c.get('factor_threads', 1)} threads, the one-thread factor's bits), SpTRSV preparation and operator "
                             f"{c['setup_prep_s']:.2f}s; the same solve on one thread took {t_one:.2f}s",
                   "solve_s": c["solve_s"], "setup_s": c["setup_factorize_s"] + c["setup_prep_s"],
                   "one_thread_solve_s": t_one}

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
                   "factor_nnz": F.nnz, "levels": info["levels"], "rhs": "N(0,1) seed 0",
                   "iters_per_check": args.iters_per_check},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": per_kernel[dom]["bytes_per_launch"],
                     "avg_launch_us": per_kernel[dom]["avg_us"]},
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
            "batched_mixed": batched,
            "second_workload": None if second is None else {args.second_workload: second},
            "reference_cell": cells,
            "group_small_factors": group,
            "serving": serving,
            "kernels_mixed": per_kernel,
            "sptrsv_sweep_gbps_mixed": sweep_gbps,
            "host_factorization_s": t_fact,
            "host_factorization_threads": F.threads,
            "solver_info": info,
        },
    }
    print(json.dumps(out))
    if args.copy_probe:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] copy probe: 3 x 256 MiB device copies", file=sys.stderr, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
