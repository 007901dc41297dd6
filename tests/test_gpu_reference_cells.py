"""The GPU solve path's factor + PCG against the reference's own published
result cells (tests/refcells.py): unit 3-D grids of side 100 and 150 with the
reference benchmark's right-hand side (mt19937_64 + normal draw, b = L g,
centred, normalised) and factor seed 42. In fp64 sweeps (the reference CPU
cells' arithmetic) and fp32 sweeps (the reference GPU cells'), the solve
reaches the recomputed residual 1e-8 within 2 iterations of the iterations
each cell reports -- the north star's "+-2 iterations of the reference",
measured against the reference itself."""
import numpy as np
import pytest

import apxchol_amd as ax
import refcells
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[100, 150])
def cell(request):
    side = request.param
    print(f"[case] reference cell grid3d_{side}: factorization (seed 42) starts", flush=True)
    A = mg.grid_laplacian((side, side, side))
    b = refcells.make_rhs(A)
    F = ax.factorize(A, seed=refcells.FACTOR_SEED)
    return side, A, b, F


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gpu_iterations_match_the_reference_cells(gpu, cell, precision):
    import torch
    side, A, b, F = cell
    n, it_cpu, _, it_gpu, _ = refcells.CELLS[side]
    print(f"[case] reference cell grid3d_{side} {precision}", flush=True)
    s = ax.Solver(A, factor=F, precision=precision)
    bd = torch.tensor(b, device=gpu)
    res = s.solve(bd, tol=1e-8, maxiter=500)
    x = res.x.cpu().numpy()
    s.close()
    tr = refcells.graded_residual(A, b, x)
    print(f"grid3d_{side} {precision}: {res.iterations} iterations (reference CPU cell {it_cpu}, GPU cell {it_gpu}), "
          f"graded residual {tr:.3e}")
    assert res.status == 1
    assert abs(res.iterations - it_cpu) <= 2 and abs(res.iterations - it_gpu) <= 2
    assert tr <= 1e-8
