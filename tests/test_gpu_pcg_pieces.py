"""The stand-alone PCG pieces of the C ABI (the reference's device boundary,
include/apxchol/solver/pcg_cuda_kernels.h:43-93) on the GPU: each against
scipy / numpy in fp64, bit-identical run to run, and a whole PCG driven from
Python with them and apx_solver_solve_LLt as the preconditioner, against the
library's own solve."""
import numpy as np
import pytest
import scipy.sparse as sp

import apxchol_amd as ax
from apxchol_amd import matgen as mg
from apxchol_amd import pcg

pytestmark = pytest.mark.gpu


def csr_on(gpu, A, fp32=False):
    import torch
    C = sp.csr_matrix(A)
    C.sort_indices()
    vals = C.data.astype(np.float32 if fp32 else np.float64)
    return (torch.tensor(C.indptr.astype(np.int32), device=gpu), torch.tensor(C.indices.astype(np.int32), device=gpu),
            torch.tensor(vals, device=gpu), C)


@pytest.mark.parametrize("name,fp32", [("grid", False), ("grid", True), ("rmat", False)])
@pytest.mark.parametrize("lanes", [1, 4, 32])
def test_spmv_pAp(gpu, name, fp32, lanes):
    import torch
    print(f"[case] pieces spmv {name} fp32={fp32} lanes={lanes}", flush=True)
    A = mg.grid_laplacian((40, 50)) if name == "grid" else mg.rmat_laplacian(11)
    rp, ci, cv, C = csr_on(gpu, A, fp32)
    n = A.shape[0]
    x = np.random.default_rng(1).standard_normal(n)
    xd = torch.tensor(x, device=gpu)
    yd = torch.empty_like(xd)
    part = pcg.partials(gpu)
    nb = pcg.spmv_pAp(rp, ci, cv, xd, yd, part, lanes)
    assert nb == pcg.blocks(n * lanes)
    out = torch.zeros(1, dtype=torch.float64, device=gpu)
    pcg.reduce_partials(part, nb, out)
    y = yd.cpu().numpy()
    Aref = C.astype(np.float64)
    bound = abs(Aref) @ abs(x)
    assert np.all(np.abs(y - Aref @ x) <= 1e-13 * bound + 1e-300)
    assert abs(out.item() - x @ y) <= 1e-12 * (np.abs(x) @ np.abs(y))
    # Run to run: the same bits (fixed grid, fixed reduction order).
    y2 = torch.empty_like(xd)
    part2 = pcg.partials(gpu)
    pcg.spmv_pAp(rp, ci, cv, xd, y2, part2, lanes)
    out2 = torch.zeros(1, dtype=torch.float64, device=gpu)
    pcg.reduce_partials(part2, nb, out2)
    assert torch.equal(yd.view(torch.int64), y2.view(torch.int64))
    assert out.view(torch.int64).item() == out2.view(torch.int64).item()


@pytest.mark.parametrize("n", [1, 1000, 256 * 2048 + 17])
def test_vector_pieces(gpu, n):
    import torch
    print(f"[case] pieces vectors n={n}", flush=True)
    rg = np.random.default_rng(n)
    x, p, r, ap, z = (rg.standard_normal(n) for _ in range(5))
    alpha, beta = 0.37, -1.25
    xd, pd, rd, apd, zd = (torch.tensor(v, device=gpu) for v in (x, p, r, ap, z))
    part = pcg.partials(gpu)
    out = torch.zeros(1, dtype=torch.float64, device=gpu)
    nb = pcg.update_xr(xd, pd, rd, apd, alpha, part)
    assert nb == pcg.blocks(n)
    pcg.reduce_partials(part, nb, out)
    x1, r1 = x + alpha * p, r - alpha * ap
    assert np.allclose(xd.cpu().numpy(), x1, rtol=1e-15, atol=1e-15)
    assert np.allclose(rd.cpu().numpy(), r1, rtol=1e-15, atol=1e-15)
    rn = rd.cpu().numpy()
    assert abs(out.item() - rn @ rn) <= 1e-13 * (rn @ rn)
    nb = pcg.dot(rd, zd, part)
    pcg.reduce_partials(part, nb, out)
    assert abs(out.item() - rn @ z) <= 1e-12 * (np.abs(rn) @ np.abs(z))
    pcg.update_p(pd, zd, beta)
    assert np.allclose(pd.cpu().numpy(), z + beta * p, rtol=1e-15, atol=1e-15)


def test_reduce_partials_order_and_empty(gpu):
    """The reduction is the fixed-order sum of exactly `blocks` partials:
    entries past it are ignored, zero partials give 0."""
    import torch
    part = pcg.partials(gpu)
    vals = np.random.default_rng(4).standard_normal(pcg.MAX_BLOCKS)
    part.copy_(torch.tensor(vals))
    out = torch.zeros(1, dtype=torch.float64, device=gpu)
    for nb in (0, 1, 7, 300, pcg.MAX_BLOCKS):
        pcg.reduce_partials(part, nb, out)
        assert abs(out.item() - vals[:nb].sum()) <= 1e-13 * max(np.abs(vals[:nb]).sum(), 1e-300), nb


@pytest.mark.parametrize("name", ["sddm3d_16", "grid2d_64"])
def test_pcg_from_pieces_matches_the_library_solve(gpu, name):
    """A PCG loop written against the pieces alone (SpMV + p.Ap, the x / r
    update + r.r, r.z, the p update; scalars through the host as in
    cuda_pcg::solve, pcg_cuda.h:157-261) with apx_solver_solve_LLt as M^-1
    (scatter to factor order, the pair, gather back; centred every time on a
    Laplacian) converges to 1e-8 within 2 iterations of the library's solve."""
    import torch
    print(f"[case] pieces pcg {name}", flush=True)
    A = mg.add_excess(mg.grid_laplacian((16, 16, 16), "lognormal")) if name == "sddm3d_16" else \
        mg.grid_laplacian((64, 64))
    F = ax.factorize(A, seed=0)
    perm = torch.tensor(F.export()[3].astype(np.int64), device=gpu)
    s = ax.Solver(A, factor=F)
    n, m = A.shape[0], s.m
    b = np.random.default_rng(5).standard_normal(n)
    if not F.sddm:
        b = mg.project_range(A, b)
    rp, ci, cv, _ = csr_on(gpu, A)
    lanes = pcg.spmv_lanes_for(A.nnz / n)
    part = pcg.partials(gpu)
    out = torch.zeros(1, dtype=torch.float64, device=gpu)

    def reduce(nb):
        pcg.reduce_partials(part, nb, out)
        return out.item()

    def precond(r):
        # z[v] = (L^-T L^-1 (P r))[perm v]; the grounded last vertex of a
        # Laplacian gets 0, and the output is centred (min-norm).
        t = torch.zeros(n, dtype=torch.float64, device=gpu)
        t[perm] = r - (0.0 if F.sddm else r.mean())
        y = torch.zeros(n, dtype=torch.float64, device=gpu)
        y[:m] = s.solve_LLt(t[:m].contiguous())
        z = y[perm]
        return z if F.sddm else z - z.mean()

    bd = torch.tensor(b, device=gpu)
    x = torch.zeros_like(bd)
    r = bd.clone()
    bnorm = float(np.linalg.norm(b))
    z = precond(r)
    p = z.clone()
    rz = reduce(pcg.dot(r, z, part))
    ap = torch.empty_like(bd)
    it = 0
    while it < 200:
        pap = reduce(pcg.spmv_pAp(rp, ci, cv, p, ap, part, lanes))
        alpha = rz / pap
        rr = reduce(pcg.update_xr(x, p, r, ap, alpha, part))
        it += 1
        if np.sqrt(rr) / bnorm < 1e-8:
            break
        z = precond(r)
        rz_new = reduce(pcg.dot(r, z, part))
        pcg.update_p(p, z, rz_new / rz)
        rz = rz_new
    ref = s.solve(bd, tol=1e-8)
    xh = x.cpu().numpy()
    if not F.sddm:
        xh = xh - xh.mean()
    assert np.linalg.norm(b - A @ xh) / bnorm < 2e-8
    assert abs(it - ref.iterations) <= 2, (it, ref.iterations)
    assert np.linalg.norm(xh - ref.x.cpu().numpy()) / np.linalg.norm(xh) < 1e-6
    s.close()
