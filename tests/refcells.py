"""Iteration counts the reference published for its own solve path, and the
reference benchmark's right-hand-side rule, for tests that pin this build's
factor + PCG against them (test infrastructure).

The cells: results/cells/grids/grid3d_{100,150}__apxchol_v1__bg_tree_vec_pool___t16__{cpu,gpu}.json
in the reference tree -- unit-weight 3-D 7-point grid Laplacians built as
benchmarks/src/v0/graphs.cpp:53-73 (grid_graph_3d, vertex (x, y, z) at
(x n + y) n + z) and benchmarks/src/benchmark.cpp:204-223 (laplacian_from_adj),
factor seed 42 (benchmark.cpp:541), block_greedy + vec_pool, tol 1e-8, graded
by the recomputed residual of the centred solution (benchmark.cpp:594-611).
The CPU cells ran the OpenMP solver (fp64 sweeps of the fp32 factor), the GPU
cells the CUDA dataflow sweeps (fp32 vectors), both with 16 host threads.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "golden" / "librefrhs.so"

# grid side -> (n, CPU cell iterations, CPU rel_res, GPU cell iterations, GPU rel_res)
CELLS = {
    100: (1_000_000, 30, 7.129866e-09, 29, 7.815489e-09),
    150: (3_375_000, 31, 7.75537e-09, 30, 9.748837e-09),
}
FACTOR_SEED = 42   # benchmark.cpp:541
RHS_SEED = 42      # benchmark.cpp:852 (--seed default), passed to make_rhs at :2626


def lib():
    L = ctypes.CDLL(str(LIB))
    L.refrhs_normal.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p]
    L.refrhs_mt64_10000th.restype = ctypes.c_uint64
    return L


def make_rhs(A, seed: int = RHS_SEED) -> np.ndarray:
    """benchmark.cpp:440-455: g ~ N(0,1) from mt19937_64(seed), b = A g,
    centred (the cells' operator is a Laplacian), normalised."""
    n = A.shape[0]
    g = np.zeros(n)
    lib().refrhs_normal(int(seed), n, g.ctypes.data)
    b = A @ g
    b -= b.mean()
    return b / np.linalg.norm(b)


def graded_residual(A, b, x) -> float:
    """benchmark.cpp:594-611: centre x, r = b - A x, centre r, ||r|| / ||b||."""
    xc = x - x.mean()
    r = b - A @ xc
    r -= r.mean()
    return float(np.linalg.norm(r) / np.linalg.norm(b))
