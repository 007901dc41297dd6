"""The benchmark configurations at their full sizes against the oracle:

* R-MAT scale 22 (n = 4,194,304, ~131 M operator entries): the fp64 GPU
  preconditioner application matches the oracle's (cpu_solver::apply) to
  1e-12 relative on two consecutive applications, and the mixed-precision
  PCG (fp32 sweeps, fp64 CG) reaches the recomputed residual 1e-8 within 2
  iterations of the oracle's fp64 solve;
* the 64^3 anisotropic grid (axis weights 1/10/100 x exp(N(0,0.5))) with 32
  right-hand sides in one blocked fp32 PCG: every column reaches the
  recomputed residual 1e-8, and sampled columns converge within 2 iterations
  of the oracle's single solve of that column;
* the serving mix of 8 resident factors (2-D grids 1e4..1e6, 3-D grids 48^3
  and 96^3, R-MAT scales 16..20): every request meets its own tolerance, and
  one request per factor gives the same bits alone as in the mix.

The oracle's solves here run on the OpenMP form of the CPU restatement (its
sweeps are the one-thread bits; only the reductions' order differs), the
applications on the one-thread form."""
import threading
import time

import numpy as np
import pytest

import apxchol_amd as ax
import oracle
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def true_res(A, b, x):
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def rmat22():
    print("[case] R-MAT s22: generation and factorization start", flush=True)
    t0 = time.perf_counter()
    A, b = mg.workload("rmat_22")
    F = ax.factorize(A, seed=0)
    print(f"[case] R-MAT s22: factor ready after {time.perf_counter() - t0:.0f}s; oracle setup", flush=True)
    o = oracle.Oracle(A, F.export(), F.sddm)
    return A, b, F, o


def test_rmat22_fp64_apply_matches_oracle(gpu, rmat22):
    import torch
    A, b, F, o = rmat22
    print("[case] R-MAT s22 fp64 apply vs oracle", flush=True)
    s = ax.Solver(A, factor=F, precision="fp64")
    o.reset_apply_count()
    rng = np.random.default_rng(31)
    for _ in range(2):
        r = mg.project_range(A, rng.standard_normal(A.shape[0]))
        z = s.apply(torch.tensor(r, device=gpu)).cpu().numpy()
        zo = o.apply(r)
        err = np.linalg.norm(z - zo) / np.linalg.norm(zo)
        assert err <= 1e-12, err
    s.close()


def test_rmat22_mixed_pcg_within_two_iterations_of_oracle(gpu, rmat22):
    import torch
    A, b, F, o = rmat22
    print("[case] R-MAT s22 mixed PCG vs oracle", flush=True)
    s = ax.Solver(A, factor=F, precision="fp32")
    res = s.solve(torch.tensor(b, device=gpu), tol=1e-8)
    x = res.x.cpu().numpy()
    s.close()
    _, it_o, _ = o.solve_omp(b, ORACLE_THREADS, tol=1e-8)
    print(f"R-MAT s22: GPU mixed {res.iterations} iterations, oracle {it_o}, true residual {true_res(A, b, x):.2e}")
    assert res.status == 1
    assert abs(res.iterations - it_o) <= 2
    assert true_res(A, b, x) <= 1e-8


def test_aniso64_k32_blocked_columns_against_oracle(gpu):
    import torch
    print("[case] aniso3d_64 k=32 blocked fp32 PCG", flush=True)
    A, _ = mg.workload("aniso3d_64")
    F = ax.factorize(A, seed=0)
    k = 32
    B = np.stack([mg.request_rhs(A, F.sddm, c) for c in range(k)])
    s = ax.Solver(A, factor=F, precision="fp32")
    s.autotune(max_block_width=64)
    X, res = s.solve_batch(torch.tensor(B, device=gpu), tol=1e-8)
    X = X.cpu().numpy()
    s.close()
    for c in range(k):
        assert res[c].status == 1, (c, res[c].status)
        assert true_res(A, B[c], X[c]) <= 1e-8, (c, true_res(A, B[c], X[c]))
    o = oracle.Oracle(A, F.export(), F.sddm)
    for c in (0, 9, 17, 31):
        _, it_o, _ = o.solve_omp(B[c], ORACLE_THREADS, tol=1e-8)
        assert abs(res[c].iterations - it_o) <= 2, (c, res[c].iterations, it_o)


def test_serving_eight_factors(gpu):
    """The serving mix's 8 resident factors: 48 requests with Poisson
    arrivals and tolerances from {1e-6, 1e-8, 1e-10}; each result meets its
    tolerance, and one request per factor is replayed alone in a fresh
    server (same block width) with the same bits (the first request of a
    factor finds its single lane free, in the mix and alone)."""
    print("[case] serving with the 8 configured factors", flush=True)
    mats = {}
    for name in mg.SERVING_FACTORS:
        A, _ = mg.workload(name)
        mats[name] = (A, ax.factorize(A, seed=0))
    rng = np.random.default_rng(6)
    names = list(mats)
    reqs = []
    for i in range(48):
        name = names[i % len(names)] if i < len(names) else names[int(rng.integers(len(names)))]
        A, F = mats[name]
        tol = [1e-6, 1e-8, 1e-10][int(rng.integers(3))]
        reqs.append((name, mg.request_rhs(A, F.sddm, 1000 + i), tol, float(rng.exponential(1.0 / 100.0))))
    srv = ax.Server(slots=32, iters_per_check=2)
    fid = {name: srv.add(ax.Solver(A, factor=F, precision="fp32")) for name, (A, F) in mats.items()}
    ids = [None] * len(reqs)

    def produce():
        for i, (name, b, tol, gap) in enumerate(reqs):
            time.sleep(gap)
            ids[i] = srv.submit(fid[name], b, tol)

    th = threading.Thread(target=produce)
    th.start()
    th.join()
    mixed = [srv.wait(r, timeout=120) for r in ids]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs)
    for (name, b, tol, _), (x, r) in zip(reqs, mixed):
        A, _ = mats[name]
        assert r.status == 1, (name, r.status)
        assert np.linalg.norm(b - A @ x) <= 2 * tol * np.linalg.norm(b), (name, tol)
    alone = ax.Server(slots=32, iters_per_check=2)
    fid2 = {name: alone.add(ax.Solver(A, factor=F, precision="fp32")) for name, (A, F) in mats.items()}
    for i in range(len(names)):   # the first request of every factor
        name, b, tol, _ = reqs[i]
        x1, r1 = alone.wait(alone.submit(fid2[name], b, tol), timeout=120)
        xm, rm = mixed[i]
        assert r1.iterations == rm.iterations, (i, name)
        assert np.array_equal(x1.view(np.uint64), xm.view(np.uint64)), (i, name)
    alone.close()
