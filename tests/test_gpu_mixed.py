"""The headline mixed-precision path against the CPU oracle, through the C ABI.

The reference's own GPU test holds the GPU sweep pair to the CPU pair on the
same factor (tests/test_sptrsv_drop.cpp:1195-1206, 1479-1530): fp32 storage
with fp32 vectors within kPairTolFp32 = 1e-6 and fp16 storage within
kPairTolFp16 = 1e-3 of the CPU fp32-storage pair (fp64 accumulation), as
max |difference| / max |reference|, and bit-identical run to run. The same
bars hold here for apx_solver_solve_LLt (the pair of cuda_sptrsv::solve_LLt_dev),
for the fp32 preconditioner application (apx_solver_apply, 12 consecutive
applications with the center-k schedule) and for every column of the blocked
fp32 application. The one-shot apx_solve reaches the oracle's tolerance within
+-2 of its iteration count in both precisions."""
import numpy as np
import pytest

import apxchol_amd as ax
import oracle
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu

PAIR_TOL_FP32 = 1e-6
PAIR_TOL_FP16 = 1e-3

CASES = {
    "grid2d_256": lambda: mg.grid_laplacian((256, 256)),
    "sddm3d_16": lambda: mg.add_excess(mg.grid_laplacian((16, 16, 16), "lognormal")),
    "aniso3d_16": lambda: mg.grid_laplacian((16, 16, 16), "lognormal", 0.5, seed=5, axis_weights=(1, 10, 100)),
    "rmat_16": lambda: mg.rmat_laplacian(16),
}
_cache = {}


def case(name):
    if name not in _cache:
        A = CASES[name]()
        F = ax.factorize(A, seed=0)
        fac = F.export()
        _cache[name] = (A, F, fac, oracle.Oracle(A, fac, F.sddm))
    return _cache[name]


def maxrel(a, b):
    """max |a - b| / max |b| (the reference's pair gate)."""
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("schedule", ["hybrid", "levels", "dataflow"])
def test_pair_fp32_vectors_matches_cpu_pair(gpu, name, schedule):
    import torch
    A, F, fac, o = case(name)
    print(f"[case] fp32 pair {name} {schedule}", flush=True)
    s = ax.Solver(A, factor=F, precision="fp32", schedule=schedule)
    m = s.m
    x = np.random.default_rng(11).uniform(-1.0, 1.0, m)
    xd = torch.tensor(x, device=gpu)
    y = s.solve_LLt(xd).cpu().numpy()
    y_cpu = o.back(o.forward(x))
    err = maxrel(y, y_cpu)
    print(f"  fp32 pair vs CPU pair: {err:.3e}", flush=True)
    assert np.all(np.isfinite(y))
    assert err < PAIR_TOL_FP32, err
    assert not np.array_equal(y, y_cpu)          # fp32 accumulation: not the fp64 bits
    y2 = s.solve_LLt(xd).cpu().numpy()
    assert np.array_equal(y.view(np.int64), y2.view(np.int64)), "pair not bit-identical run to run"
    # In place (d_x == d_y), as solve_LLt_dev allows.
    xd2 = xd.clone()
    s.solve_LLt(xd2, out=xd2)
    assert np.array_equal(xd2.cpu().numpy().view(np.int64), y.view(np.int64))
    s.close()


@pytest.mark.parametrize("name", list(CASES))
def test_pair_fp64_vectors_is_the_oracle_pair(gpu, name):
    """With fp64 vectors the same entry point is the CPU pair to 1e-12."""
    import torch
    A, F, fac, o = case(name)
    print(f"[case] fp64 pair {name}", flush=True)
    s = ax.Solver(A, factor=F, precision="fp64")
    x = np.random.default_rng(12).uniform(-1.0, 1.0, s.m)
    y = s.solve_LLt(torch.tensor(x, device=gpu)).cpu().numpy()
    assert maxrel(y, o.back(o.forward(x))) < 1e-12
    s.close()


@pytest.mark.parametrize("name", list(CASES))
def test_pair_fp16_storage_matches_cpu_pair(gpu, name):
    """fp16 per-column-scaled storage with fp32 vectors: within 1e-3 of the
    CPU fp32-storage pair (the reference's gate: a different preconditioner,
    values rounded to 11 bits), and within 1e-5 of the CPU pair on the same
    fp16 storage (only the accumulation width differs)."""
    import torch
    A, F, fac, o = case(name)
    print(f"[case] fp16 pair {name}", flush=True)
    o16 = oracle.Oracle(A, fac, F.sddm, fp16=True)
    s = ax.Solver(A, factor=F, precision="fp32", fp16_storage=True)
    x = np.random.default_rng(13).uniform(-1.0, 1.0, s.m)
    xd = torch.tensor(x, device=gpu)
    y = s.solve_LLt(xd).cpu().numpy()
    e32 = maxrel(y, o.back(o.forward(x)))
    e16 = maxrel(y, o16.back(o16.forward(x)))
    print(f"  fp16 pair vs CPU fp32 pair {e32:.3e}, vs CPU fp16 pair {e16:.3e}", flush=True)
    assert e32 < PAIR_TOL_FP16, e32
    assert e16 < 1e-5, e16
    y2 = s.solve_LLt(xd).cpu().numpy()
    assert np.array_equal(y.view(np.int64), y2.view(np.int64))
    s.close()


@pytest.mark.parametrize("name", list(CASES))
def test_apply_fp32_matches_oracle(gpu, name):
    """cpu_solver::apply in the mixed configuration (fp32 vectors), 12
    consecutive applications -- the 10th centres on a Laplacian -- each within
    1e-6 of the oracle's fp64-accumulated application."""
    import torch
    A, F, fac, o = case(name)
    print(f"[case] fp32 apply {name}", flush=True)
    s = ax.Solver(A, factor=F, precision="fp32")
    o.reset_apply_count()
    s.reset_apply_count()
    rg = np.random.default_rng(21)
    for k in range(12):
        r = rg.standard_normal(A.shape[0])
        z = s.apply(torch.tensor(r, device=gpu)).cpu().numpy()
        zo = o.apply(r)
        assert maxrel(z, zo) < PAIR_TOL_FP32, f"application {k + 1}: {maxrel(z, zo):.3e}"
    s.close()


@pytest.mark.parametrize("name", list(CASES))
def test_apply_batch_fp32_columns_match_oracle(gpu, name):
    """The blocked fp32 application (k columns counted as one application of
    the center-k schedule), 12 rounds: every column within 1e-6 of the oracle's
    application of that column at the same count (one oracle per column)."""
    import torch
    A, F, fac, _ = case(name)
    print(f"[case] fp32 apply_batch {name}", flush=True)
    k = 5
    s = ax.Solver(A, factor=F, precision="fp32")
    s.reset_apply_count()
    ocol = [oracle.Oracle(A, fac, F.sddm) for _ in range(k)]
    rg = np.random.default_rng(22)
    for j in range(12):
        R = rg.standard_normal((k, A.shape[0]))
        Z = s.apply_batch(torch.tensor(R, device=gpu)).cpu().numpy()
        for c in range(k):
            zo = ocol[c].apply(R[c])
            assert maxrel(Z[c], zo) < PAIR_TOL_FP32, (j + 1, c, maxrel(Z[c], zo))
    s.close()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_one_shot_solve_matches_oracle(gpu, name, precision):
    """apxchol::solve (apx_solve): factor at the default seed, GPU PCG to 1e-8,
    within +-2 iterations of the oracle's solve on the same factor, graded by
    the true residual; its timings carry the setup phases and the PCG."""
    A, _, _, _ = case(name)
    print(f"[case] one-shot {name} {precision}", flush=True)
    b = np.random.default_rng(31).standard_normal(A.shape[0])
    F = ax.factorize(A, seed=42)
    if not F.sddm:
        b = mg.project_range(A, b)
    res = ax.solve(A, b, tol=1e-8, precision=precision)
    o = oracle.Oracle(A, F.export(), F.sddm)
    xo, ito, _ = o.solve(b, tol=1e-8)
    true_res = np.linalg.norm(b - A @ res.x) / np.linalg.norm(b)
    assert res.status == 1, res
    assert abs(res.iterations - ito) <= 2, (res.iterations, ito)
    assert true_res < 2e-8, true_res
    assert np.linalg.norm(res.x - xo) / np.linalg.norm(xo) < 1e-4
    t = res.timings
    assert t["factorize_ms"] > 0 and t["sptrsv_setup_ms"] > 0 and t["pcg_setup_ms"] > 0 and t["pcg_ms"] > 0, t


def test_solver_solve_reports_setup_timings(gpu):
    """A solver created from a given factor reports factorize_ms = 0 and its own
    setup phases with every solve."""
    import torch
    A, F, _, _ = case("sddm3d_16")
    s = ax.Solver(A, factor=F)
    r1 = s.solve(torch.tensor(np.ones(A.shape[0]), device=gpu))
    r2 = s.solve(torch.tensor(np.ones(A.shape[0]), device=gpu))
    assert r1.timings["factorize_ms"] == 0.0
    assert r1.timings["sptrsv_setup_ms"] == r2.timings["sptrsv_setup_ms"] > 0
    assert r1.timings["pcg_ms"] > 0 and r2.timings["pcg_ms"] > 0
    s.close()
