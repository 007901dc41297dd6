"""The solve path at the benchmark's full size -- the 3-D 128^3 SDDM
(n = 2,097,152, lognormal weights, diagonal excess) and an R-MAT scale-18
graph Laplacian -- checked through properties that do not depend on size,
since the CPU oracle's solve at this size belongs to the benchmark, not to the
test suite:

* the preconditioner with fp64 vectors is linear (M(a r1 + r2) = a M r1 +
  M r2) and symmetric (r1 . M r2 = r2 . M r1, M = (L L^T)^-1 on the
  permuted system), both to 1e-12 relative -- a transposed sweep, a dropped
  term or a wrong permutation breaks the second;
* PCG reaches ||b - A x|| / ||b|| <= 1e-8, recomputed in fp64 on the host,
  in fp64 and mixed precision, with iteration counts within 2 of each other
  (the reference's +-2 bar between precisions);
* a blocked solve of k right-hand sides gives each column the single solve's
  convergence (true residual, iterations within 1);
* the one-launch-per-sweep dataflow schedule gives the hybrid schedule's
  bits (application and PCG solve)."""
import numpy as np
import pytest

import apxchol_amd as ax
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu

_cache = {}


def workload(name):
    if name not in _cache:
        A, b = mg.workload(name)
        F = ax.factorize(A, seed=0)
        _cache[name] = (A, b, F)
    return _cache[name]


def true_res(A, b, x):
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


@pytest.mark.parametrize("name", ["sddm3d_128", "rmat_18"])
def test_fp64_preconditioner_is_linear_and_symmetric(gpu, name):
    import torch
    A, b, F = workload(name)
    print(f"[case] full-size apply properties {name} n={A.shape[0]}", flush=True)
    s = ax.Solver(A, factor=F, precision="fp64")
    rg = np.random.default_rng(21)
    r1, r2 = rg.standard_normal(A.shape[0]), rg.standard_normal(A.shape[0])
    if not F.sddm:   # the range of a Laplacian: orthogonal to the constants
        r1 -= r1.mean()
        r2 -= r2.mean()
    a = 0.37
    zs = []
    for r in (r1, r2, a * r1 + r2):
        s.reset_apply_count()   # the same application index: no centring difference
        zs.append(s.apply(torch.tensor(r, device=gpu)).cpu().numpy())
    z1, z2, z12 = zs
    lin = np.linalg.norm(z12 - (a * z1 + z2)) / np.linalg.norm(z12)
    assert lin <= 1e-12, lin
    sym = abs(r1 @ z2 - r2 @ z1) / (np.linalg.norm(r1) * np.linalg.norm(z2))
    assert sym <= 1e-12, sym
    # Positive definite on the range: r . M r > 0.
    assert r1 @ z1 > 0 and r2 @ z2 > 0
    s.close()


@pytest.mark.parametrize("name", ["sddm3d_128", "rmat_18"])
def test_full_size_pcg_meets_tolerance_in_both_precisions(gpu, name):
    import torch
    A, b, F = workload(name)
    print(f"[case] full-size PCG {name} n={A.shape[0]}", flush=True)
    its = {}
    xs = {}
    for prec in ("fp64", "fp32"):
        s = ax.Solver(A, factor=F, precision=prec)
        bd = torch.tensor(b, device=gpu)
        r = s.solve(bd)
        x = r.x.cpu().numpy()
        assert r.status == 1, (prec, r.status)
        assert true_res(A, b, x) <= 1e-8 * 1.05, (prec, true_res(A, b, x))
        its[prec] = r.iterations
        xs[prec] = x
        s.close()
    assert abs(its["fp64"] - its["fp32"]) <= 2, its
    # Two solutions to 1e-8 relative residual agree to the conditioning's
    # share of that (loose: the smallest eigenvalues set it); a Laplacian's
    # solutions are compared up to the constant vector.
    x64, x32 = xs["fp64"], xs["fp32"]
    if not F.sddm:
        x64, x32 = x64 - x64.mean(), x32 - x32.mean()
    assert np.linalg.norm(x64 - x32) / np.linalg.norm(x64) < 1e-4


def test_full_size_blocked_columns_converge_like_single_solves(gpu):
    import torch
    A, b, F = workload("sddm3d_128")
    print("[case] full-size blocked sddm3d_128 k=4", flush=True)
    s = ax.Solver(A, factor=F, precision="fp32")
    rg = np.random.default_rng(5)
    B = rg.standard_normal((4, A.shape[0]))
    X, res = s.solve_batch(torch.tensor(B, device=gpu))
    X = X.cpu().numpy()
    for c in range(4):
        assert res[c].status == 1, (c, res[c].status)
        assert true_res(A, B[c], X[c]) <= 1e-8 * 1.05, (c, true_res(A, B[c], X[c]))
        single = s.solve(torch.tensor(B[c], device=gpu))
        assert abs(single.iterations - res[c].iterations) <= 1, (c, single.iterations, res[c].iterations)
    s.close()


@pytest.mark.parametrize("name", ["sddm3d_128", "rmat_18"])
def test_full_size_dataflow_schedule_gives_the_hybrid_bits(gpu, name):
    """The one-launch-per-sweep dataflow schedule at full size, on its default
    resident grid (2 workgroups per CU, the largest validated; larger grids
    are refused): the mixed-precision preconditioner application and a mixed
    PCG solve are the bits of the hybrid schedule's, since every schedule
    runs the same per-row arithmetic."""
    import torch
    A, b, F = workload(name)
    print(f"[case] full-size dataflow schedule {name}", flush=True)
    out = {}
    for sched in ("hybrid", "dataflow"):
        s = ax.Solver(A, factor=F, precision="fp32", schedule=sched)
        bd = torch.tensor(b, device=gpu)
        z = s.apply(bd)
        r = s.solve(bd)
        torch.cuda.synchronize()
        out[sched] = (z.view(torch.int64).cpu(), r.x.view(torch.int64).cpu(), r.iterations, r.status)
        s.close()
    (zh, xh, ih, sh), (zd, xd, idf, sd) = out["hybrid"], out["dataflow"]
    assert sh == sd == 1
    assert ih == idf
    assert torch.equal(zh, zd)
    assert torch.equal(xh, xd)
