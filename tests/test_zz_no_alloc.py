"""The allocation-free entry point apx_pcg_spmv_pAp_noalloc against the entry
point it varies, apx_pcg_spmv_pAp followed by apx_pcg_reduce_partials (the
reference's device boundary, include/apxchol/solver/pcg_cuda_kernels.h:75-93):
the same bits for y and for x.y, on consecutive calls whose sizes shrink (the
library's scratch array keeps the previous call's partials), at the largest
producer grid, and for an empty system; and y against scipy in fp64."""
import numpy as np
import pytest
import scipy.sparse as sp

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


def two_step(gpu, rp, ci, cv, xd, lanes):
    """The existing entry points: spmv_pAp into a caller partials buffer, then
    the fixed-order reduction."""
    import torch
    y = torch.empty_like(xd)
    part = pcg.partials(gpu)
    nb = pcg.spmv_pAp(rp, ci, cv, xd, y, part, lanes)
    out = torch.zeros(1, dtype=torch.float64, device=gpu)
    pcg.reduce_partials(part, nb, out)
    return y, out


def bits(t):
    import torch
    return t.view(torch.int64)


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_equals_two_step_on_shrinking_calls(gpu, fp32):
    """Three calls in a row, each smaller than the one before (the largest one
    at the full producer grid of 2048 partials): every call's y and x.y are
    the two-step form's bits, and y is A x to 1e-13 of (|A||x|)_i."""
    import torch
    cases = [((600, 600), 4), ((40, 50), 4), ((12, 9), 1)]
    for shape, lanes in cases:
        A = mg.grid_laplacian(shape, "lognormal")
        n = A.shape[0]
        print(f"[case] noalloc grid {shape} n={n} lanes={lanes} fp32={fp32} "
              f"partials={pcg.blocks(n * lanes)}", flush=True)
        rp, ci, cv, C = csr_on(gpu, A, fp32)
        x = np.random.default_rng(n).standard_normal(n)
        xd = torch.tensor(x, device=gpu)
        y_ref, pAp_ref = two_step(gpu, rp, ci, cv, xd, lanes)
        y = torch.full_like(xd, float("nan"))
        pAp = torch.full((1,), float("nan"), dtype=torch.float64, device=gpu)
        pcg.spmv_pAp_noalloc(rp, ci, cv, xd, y, pAp, lanes)
        torch.cuda.synchronize()
        assert torch.equal(bits(y), bits(y_ref))
        assert bits(pAp).item() == bits(pAp_ref).item()
        Aref = C.astype(np.float64)
        if fp32:   # the matrix the kernel multiplies: values rounded to fp32
            Aref.data = C.data.astype(np.float32).astype(np.float64)
        yh = y.cpu().numpy()
        assert np.all(np.abs(yh - Aref @ x) <= 1e-13 * (abs(Aref) @ abs(x)) + 1e-300)
        assert abs(pAp.item() - x @ yh) <= 1e-12 * (np.abs(x) @ np.abs(yh))
    assert pcg.blocks(360000 * 4) == pcg.MAX_BLOCKS


def test_noalloc_power_law_rows_and_empty_system(gpu):
    """An R-MAT Laplacian (rows of very different lengths) at 32 lanes, then a
    system with no rows: x.y = 0 is written."""
    import torch
    A = mg.rmat_laplacian(11)
    n = A.shape[0]
    print(f"[case] noalloc rmat n={n} lanes=32, then n=0", flush=True)
    rp, ci, cv, _ = csr_on(gpu, A)
    xd = torch.tensor(np.random.default_rng(3).standard_normal(n), device=gpu)
    y_ref, pAp_ref = two_step(gpu, rp, ci, cv, xd, 32)
    y = torch.empty_like(xd)
    pAp = torch.full((1,), float("nan"), dtype=torch.float64, device=gpu)
    pcg.spmv_pAp_noalloc(rp, ci, cv, xd, y, pAp, 32)
    torch.cuda.synchronize()
    assert torch.equal(bits(y), bits(y_ref)) and bits(pAp).item() == bits(pAp_ref).item()
    e_rp = torch.zeros(1, dtype=torch.int32, device=gpu)
    e_ci = torch.zeros(0, dtype=torch.int32, device=gpu)
    e_cv = torch.zeros(0, dtype=torch.float64, device=gpu)
    e_x = torch.zeros(0, dtype=torch.float64, device=gpu)
    pcg.spmv_pAp_noalloc(e_rp, e_ci, e_cv, e_x, torch.zeros_like(e_x), pAp, 4)
    torch.cuda.synchronize()
    assert pAp.item() == 0.0


def test_noalloc_refuses_bad_lanes_before_any_launch(gpu):
    import torch
    A = mg.grid_laplacian((8, 8))
    rp, ci, cv, _ = csr_on(gpu, A)
    xd = torch.ones(A.shape[0], dtype=torch.float64, device=gpu)
    pAp = torch.zeros(1, dtype=torch.float64, device=gpu)
    for lanes in (0, 3, 128):
        with pytest.raises(ValueError, match="lanes"):
            pcg.spmv_pAp_noalloc(rp, ci, cv, xd, torch.empty_like(xd), pAp, lanes)
