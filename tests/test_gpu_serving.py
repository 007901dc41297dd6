"""GPU tests of the serving loop (continuous batching over resident factors).

Bars: with one block per factor (no single lane, no narrow block), every
request's result is bit-identical whether it ran alone in its factor's block
or admitted mid-flight among others (the per-column arithmetic does not
depend on the other columns or the slot); with the defaults, a request alone
gets the bits of its own single-vector solve, and in the mix the bits of
exactly one of its paths (that solve, or its column in the narrow or the
main block); every result meets its own tolerance (true residual below
2 x tol) and agrees with a single-vector solve within one iteration; factors
beyond the memory budget are refused."""
import threading
import time

import numpy as np
import pytest

import apxchol_amd as ax
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu

FACTORS = {
    "grid2d_64": lambda: mg.grid_laplacian((64, 64)),
    "sddm3d_16": lambda: mg.add_excess(mg.grid_laplacian((16, 16, 16), "lognormal")),
    "rmat_12": lambda: mg.rmat_laplacian(12),
}
_cache = {}


def factor(name):
    if name not in _cache:
        A = FACTORS[name]()
        _cache[name] = (A, ax.factorize(A, seed=0))
    return _cache[name]


def make_requests(count, seed):
    rng = np.random.default_rng(seed)
    names = list(FACTORS)
    reqs = []
    for i in range(count):
        name = names[int(rng.integers(len(names)))]
        A, F = factor(name)
        b = rng.standard_normal(A.shape[0])
        if not F.sddm:
            b = mg.project_range(A, b)
        tol = [1e-6, 1e-8, 1e-10][int(rng.integers(3))]
        reqs.append((name, b, tol, float(rng.exponential(0.004))))
    return reqs


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("slots", [8, 32])
def test_serving_mix_equals_alone(gpu, slots):
    print(f"[case] serving slots={slots}", flush=True)
    reqs = make_requests(36, 7)
    # In the mix: Poisson-like arrivals from a producer thread while the
    # server iterates.
    mix = ax.Server(slots=slots, iters_per_check=2, single_lane=False, narrow_slots=0)
    fid = {name: mix.add(ax.Solver(factor(name)[0], factor=factor(name)[1], precision="fp32")) for name in FACTORS}
    ids = [None] * len(reqs)

    def produce():
        for i, (name, b, tol, gap) in enumerate(reqs):
            time.sleep(gap)
            ids[i] = mix.submit(fid[name], b, tol)

    th = threading.Thread(target=produce)
    th.start()
    th.join()
    mixed = [mix.wait(r, timeout=120) for r in ids]
    st = mix.stats()
    assert st["completed"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0 and st["solves_per_s"] > 0
    mix.close()
    # Alone: the same requests one at a time, each alone in its block.
    alone = ax.Server(slots=slots, iters_per_check=2, single_lane=False, narrow_slots=0)
    fid2 = {name: alone.add(ax.Solver(factor(name)[0], factor=factor(name)[1], precision="fp32")) for name in FACTORS}
    for i, (name, b, tol, _) in enumerate(reqs):
        x1, r1 = alone.wait(alone.submit(fid2[name], b, tol), timeout=120)
        xm, rm = mixed[i]
        assert r1.status == rm.status == 1, (i, r1.status, rm.status)
        assert r1.iterations == rm.iterations, i
        assert np.array_equal(bits(x1), bits(xm)), i
        A, _ = factor(name)
        assert np.linalg.norm(b - A @ xm) <= 2 * tol * np.linalg.norm(b), (i, name, tol)
    alone.close()


def block_alone(reqs, width):
    """Each request alone in a block of `width` columns (no single lane, no
    narrow block)."""
    srv = ax.Server(slots=width, iters_per_check=2, single_lane=False, narrow_slots=0)
    f = {name: srv.add(ax.Solver(factor(name)[0], factor=factor(name)[1], precision="fp32")) for name in FACTORS}
    out = [srv.wait(srv.submit(f[name], b, tol), timeout=120) for name, b, tol, _ in reqs]
    srv.close()
    return out


def test_requests_take_the_single_lane_or_a_block(gpu):
    """Defaults (single lane on, a 4-column narrow block before an 8-column
    main block): a request submitted alone runs on its factor's single lane
    and gets the bits of Solver.solve(b, tol); among Poisson arrivals each
    request gets the bits of exactly one of its paths -- that solve, its
    column alone in a 4-wide block, or alone in an 8-wide block -- and the
    single lane and the narrow block both take requests."""
    print("[case] serving single lane + narrow block", flush=True)
    reqs = make_requests(36, 7)
    own = {name: ax.Solver(factor(name)[0], factor=factor(name)[1], precision="fp32") for name in FACTORS}
    singles = [own[name].solve(b, tol=tol, maxiter=200) for name, b, tol, _ in reqs]
    for o in own.values():
        o.close()
    narrow, main = block_alone(reqs, 4), block_alone(reqs, 8)
    alone = ax.Server(slots=8, iters_per_check=2)
    fa = {name: alone.add(ax.Solver(factor(name)[0], factor=factor(name)[1], precision="fp32")) for name in FACTORS}
    for i, (name, b, tol, _) in enumerate(reqs[:9]):
        x1, r1 = alone.wait(alone.submit(fa[name], b, tol), timeout=120)
        assert r1.iterations == singles[i].iterations, i
        assert np.array_equal(bits(x1), bits(singles[i].x)), i
    alone.close()
    mix = ax.Server(slots=8, iters_per_check=2)
    fm = {name: mix.add(ax.Solver(factor(name)[0], factor=factor(name)[1], precision="fp32")) for name in FACTORS}
    ids = [None] * len(reqs)

    def produce():
        for i, (name, b, tol, gap) in enumerate(reqs):
            time.sleep(gap / 4)
            ids[i] = mix.submit(fm[name], b, tol)

    th = threading.Thread(target=produce)
    th.start()
    th.join()
    mixed = [mix.wait(r, timeout=120) for r in ids]
    st = mix.stats()
    mix.close()
    paths = {"single": 0, "narrow": 0, "main": 0}
    for i, (name, b, tol, _) in enumerate(reqs):
        xm, rm = mixed[i]
        A, _ = factor(name)
        assert rm.status == 1 and np.linalg.norm(b - A @ xm) <= 2 * tol * np.linalg.norm(b), (i, name, tol)
        refs = {"single": (singles[i].x, singles[i].iterations), "narrow": (narrow[i][0], narrow[i][1].iterations),
                "main": (main[i][0], main[i][1].iterations)}
        hit = [k for k, (x, it) in refs.items() if np.array_equal(bits(xm), bits(x)) and rm.iterations == it]
        assert hit, (i, name)
        paths[hit[0]] += 1
    print(f"[case] paths taken: {paths}", flush=True)
    assert paths["single"] >= len(FACTORS) and paths["single"] == st["single_lane_completed"]
    assert paths["narrow"] > 0


def test_serving_agrees_with_single_solves(gpu):
    import torch
    reqs = make_requests(9, 11)
    srv = ax.Server(slots=8)
    solvers = {name: ax.Solver(factor(name)[0], factor=factor(name)[1], precision="fp32") for name in FACTORS}
    fid = {name: srv.add(solvers[name]) for name in FACTORS}
    ids = [srv.submit(fid[name], b, tol) for name, b, tol, _ in reqs]
    out = [srv.wait(r, timeout=120) for r in ids]
    srv.close()
    for (name, b, tol, _), (x, r) in zip(reqs, out):
        A, F = factor(name)
        one = ax.Solver(A, factor=F, precision="fp32", tol=tol)
        o = one.solve(torch.tensor(b, device=gpu))
        one.close()
        assert abs(o.iterations - r.iterations) <= 1
        x1 = o.x.cpu().numpy()
        assert np.linalg.norm(x - x1) <= 1e-4 * np.linalg.norm(x1) + 1e-300


def test_serving_memory_budget(gpu):
    A, F = factor("sddm3d_16")
    srv = ax.Server(memory_budget=1 << 20, slots=8)
    with pytest.raises(RuntimeError, match="memory budget"):
        srv.add(ax.Solver(A, factor=F))
    srv.close()


def test_a_serving_solver_takes_no_other_call(gpu):
    """A solver serves one server at a time and takes no other device call
    while it does (the session owns its block buffers, graphs and capture
    stream): a second add -- to the same or another server -- and every
    direct call are refused before any launch; the server keeps serving it,
    and once the server is gone the solver works again."""
    import torch
    A, F = factor("sddm3d_16")
    n = A.shape[0]
    s = ax.Solver(A, factor=F, precision="fp32")
    srv = ax.Server(slots=8)
    fid = srv.add(s)
    other = ax.Server(slots=8)
    for target in (other, srv):
        with pytest.raises(ValueError, match="already serves"):
            target.add(s)
    b = torch.tensor(np.ones(n), device=gpu)
    B = torch.ones((2, n), dtype=torch.float64, device=gpu)
    for call in (lambda: s.solve(b), lambda: s.solve_batch(B), lambda: s.apply(b), lambda: s.apply_batch(B),
                 lambda: s.spmv(b), lambda: s.spmv_batch(B), lambda: s.solve_LLt(b), lambda: s.autotune()):
        with pytest.raises(ValueError, match="serving"):
            call()
    x, r = srv.wait(srv.submit(fid, np.ones(n), 1e-8), timeout=120)
    assert r.status == 1 and np.linalg.norm(np.ones(n) - A @ x) <= 2e-8 * np.sqrt(n)
    srv.close()
    other.close()
    assert s.solve(b).status == 1
    s.close()
