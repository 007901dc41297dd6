"""The oracle's factor + PCG against the reference's own published result
cells (tests/refcells.py names them): on the unit 3-D grid of side 100 with
the reference benchmark's right-hand side and factor seed, the one-thread
restatement reaches the cells' tolerance within 2 iterations of the
iterations the reference reported. Also pins the RHS generator's engine."""
import numpy as np
import pytest

import apxchol_amd as ax
import oracle
import refcells
from apxchol_amd import matgen as mg


def test_rhs_engine_is_the_standard_mt19937_64():
    assert refcells.lib().refrhs_mt64_10000th() == 9981545732273789042


def test_rhs_rule():
    A = mg.grid_laplacian((5, 6, 7))
    b = refcells.make_rhs(A)
    assert abs(b.sum()) < 1e-12 and abs(np.linalg.norm(b) - 1.0) < 1e-12
    assert np.array_equal(b, refcells.make_rhs(A))


@pytest.mark.parametrize("side", [100])
def test_oracle_iterations_match_the_reference_cells(side):
    n, it_cpu, res_cpu, it_gpu, _ = refcells.CELLS[side]
    A = mg.grid_laplacian((side, side, side))
    assert A.shape[0] == n and A.nnz == n + 2 * 3 * side * side * (side - 1)
    b = refcells.make_rhs(A)
    F = ax.factorize(A, seed=refcells.FACTOR_SEED)
    assert not F.sddm
    o = oracle.Oracle(A, F.export(), F.sddm)
    x, it, _ = o.solve_omp(b, 4, tol=1e-8, max_iter=500)
    print(f"grid3d_{side}: oracle {it} iterations, reference CPU cell {it_cpu}, GPU cell {it_gpu}")
    assert abs(it - it_cpu) <= 2 and abs(it - it_gpu) <= 2
    assert refcells.graded_residual(A, b, x) < 1e-8
