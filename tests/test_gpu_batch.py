"""GPU tests of the blocked k-right-hand-side path through the C ABI.

Bars: a non-centring blocked preconditioner application gives every column
the single apply's bits (the blocked sweeps replay the single-vector lane
order); a centring one agrees to 1e-12 relative. A blocked solve stops every
column exactly like its own solve would, within one iteration of the single
solve (their dot products are reduced in another fixed order), with the
solutions within 1e-6 relative and every true residual below 2e-8 at tol 1e-8;
graph replay and eager launches give the same bits."""
import numpy as np
import pytest

import apxchol_amd as ax
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu

CASES = {
    "grid2d_64": lambda: mg.grid_laplacian((64, 64)),
    "grid2d_256": lambda: mg.grid_laplacian((256, 256)),
    "sddm3d_16": lambda: mg.add_excess(mg.grid_laplacian((16, 16, 16), "lognormal")),
    "aniso3d_16": lambda: mg.grid_laplacian((16, 16, 16), "lognormal", 0.5, seed=5, axis_weights=(1, 10, 100)),
    "rmat_12": lambda: mg.rmat_laplacian(12),
    "rmat_14": lambda: mg.rmat_laplacian(14),   # hub rows: overflow cuts in the tail runs
}
_cache = {}


def case(name):
    if name not in _cache:
        A = CASES[name]()
        _cache[name] = (A, ax.factorize(A, seed=0))
    return _cache[name]


def rhs_block(A, sddm, k, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((k, A.shape[0]))
    if not sddm:
        B = np.stack([mg.project_range(A, b) for b in B])
    return B


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("name", ["grid2d_64", "sddm3d_16", "rmat_12", "rmat_14"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("fp16", [False, True])
@pytest.mark.parametrize("col_tail", ["1", "0"])
def test_apply_batch_bits_equal_single_apply(gpu, name, precision, fp16, col_tail, monkeypatch):
    """Every column of a blocked application has the single application's
    bits, with the tail runs walked column by column (one tail workgroup per
    column, the default) or by the blocked one-workgroup tail."""
    import torch
    A, F = case(name)
    print(f"[case] apply_batch {name} {precision} fp16={fp16} col_tail={col_tail}", flush=True)
    monkeypatch.setenv("APXCHOL_AMD_BLOCK_COL_TAIL", col_tail)
    s = ax.Solver(A, factor=F, precision=precision, fp16_storage=fp16)
    k = 6
    R = torch.tensor(rhs_block(A, F.sddm, k, 1), device=gpu)
    s.reset_apply_count()
    Z = s.apply_batch(R).cpu().numpy()          # application 1: no centring
    for c in range(k):
        s.reset_apply_count()
        z = s.apply(R[c].clone()).cpu().numpy()
        assert np.array_equal(bits(z), bits(Z[c])), c
    s.close()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("fmt", ["auto", "csr"])
@pytest.mark.parametrize("k", [5, 70])
def test_spmv_batch_bits_equal_single_spmv(gpu, name, fmt, k, monkeypatch):
    """The blocked SpMM gives every column the single SpMV's bits, over the
    sliced-ELL form (slices, and wave-walked long rows on R-MAT) or the CSR
    lanes alike; k = 70 runs as blocks of 64 and 6."""
    import torch
    A, F = case(name)
    print(f"[case] spmv_batch {name} {fmt} k={k}", flush=True)
    if fmt != "auto":
        monkeypatch.setenv("APXCHOL_AMD_SPMV", fmt)
    s = ax.Solver(A, factor=F, precision="fp32")
    X = torch.tensor(np.random.default_rng(13).standard_normal((k, A.shape[0])), device=gpu)
    Y = s.spmv_batch(X).cpu().numpy()
    for c in range(k):
        y = s.spmv(X[c].clone()).cpu().numpy()
        assert np.array_equal(bits(y), bits(Y[c])), c
    Yr = (A @ X.cpu().numpy().T).T
    assert np.abs(Y - Yr).max() <= 1e-12 * np.abs(Yr).max()
    s.close()


@pytest.mark.parametrize("name", ["grid2d_64", "rmat_12"])
def test_apply_batch_centring_application(gpu, name):
    import torch
    A, F = case(name)
    s = ax.Solver(A, factor=F)
    k = 4
    R = torch.tensor(rhs_block(A, F.sddm, k, 2), device=gpu)
    s.reset_apply_count()
    for _ in range(9):
        s.apply_batch(R)
    Z = s.apply_batch(R).cpu().numpy()         # application 10: centres
    for c in range(k):
        s.reset_apply_count()
        for _ in range(9):
            s.apply(R[c].clone())
        z = s.apply(R[c].clone()).cpu().numpy()
        assert np.linalg.norm(z - Z[c]) <= 1e-12 * np.linalg.norm(z)
        assert abs(Z[c].mean()) <= 1e-12 * np.linalg.norm(Z[c])
    s.close()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("k", [5, 32])
def test_solve_batch_matches_single_solves(gpu, name, precision, k):
    import torch
    A, F = case(name)
    print(f"[case] solve_batch {name} {precision} k={k}", flush=True)
    s = ax.Solver(A, factor=F, precision=precision)
    B = rhs_block(A, F.sddm, k, 3)
    Bd = torch.tensor(B, device=gpu)
    X, res = s.solve_batch(Bd)
    X = X.cpu().numpy()
    for c in range(k):
        one = s.solve(Bd[c].clone())
        x1 = one.x.cpu().numpy()
        assert res[c].status == 1 and one.status == 1
        assert abs(res[c].iterations - one.iterations) <= 1, (c, res[c].iterations, one.iterations)
        true_res = np.linalg.norm(B[c] - A @ X[c]) / np.linalg.norm(B[c])
        assert true_res < 2e-8, (c, true_res)
        assert np.linalg.norm(X[c] - x1) <= 1e-6 * np.linalg.norm(x1), c
    s.close()


def test_solve_batch_zero_column_and_wide_block(gpu):
    """A zero right-hand side ends at once with x = 0 (status 5); k > 64 runs
    as blocks of 64."""
    import torch
    A, F = case("sddm3d_16")
    s = ax.Solver(A, factor=F, precision="fp32")
    k = 70
    B = rhs_block(A, True, k, 4)
    B[3] = 0.0
    X, res = s.solve_batch(torch.tensor(B, device=gpu))
    X = X.cpu().numpy()
    assert res[3].status == 5 and not X[3].any()
    for c in range(k):
        if c == 3:
            continue
        assert res[c].status == 1
        assert np.linalg.norm(B[c] - A @ X[c]) / np.linalg.norm(B[c]) < 2e-8
    s.close()


@pytest.mark.parametrize("name", ["grid2d_256", "rmat_12"])
def test_solve_batch_graph_matches_eager(gpu, name):
    import torch
    A, F = case(name)
    B = torch.tensor(rhs_block(A, F.sddm, 8, 5), device=gpu)
    e = ax.Solver(A, factor=F, precision="fp32", graph=False)
    Xe, re = e.solve_batch(B)
    e.close()
    g = ax.Solver(A, factor=F, precision="fp32", graph=True)
    for _ in range(2):
        Xg, rg = g.solve_batch(B)
        torch.cuda.synchronize()
        assert np.array_equal(bits(Xe.cpu().numpy()), bits(Xg.cpu().numpy()))
        assert [r.iterations for r in re] == [r.iterations for r in rg]
    g.close()


def test_autotune_keeps_results_consistent(gpu):
    """After autotuning (which may change the SpMV lanes), the single and the
    blocked paths still agree: blocked apply bits = single apply bits, and a
    blocked solve matches single solves within one iteration."""
    import torch
    A, F = case("sddm3d_16")
    s = ax.Solver(A, factor=F, precision="fp32")
    r = s.autotune(max_block_width=32)
    assert r["spmv_lanes"] in (1, 2, 4, 8, 16, 32, 64) and r["block_width"] in (8, 16, 32)
    assert s.info["spmv_lanes"] == r["spmv_lanes"]
    B = torch.tensor(rhs_block(A, True, 4, 9), device=gpu)
    s.reset_apply_count()
    Z = s.apply_batch(B).cpu().numpy()
    for c in range(4):
        s.reset_apply_count()
        assert np.array_equal(bits(s.apply(B[c].clone()).cpu().numpy()), bits(Z[c]))
    X, res = s.solve_batch(B)
    for c in range(4):
        one = s.solve(B[c].clone())
        assert abs(one.iterations - res[c].iterations) <= 1
    s.close()
