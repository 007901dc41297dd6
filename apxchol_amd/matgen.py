"""Synthetic Laplacian / SDDM operators and right-hand sides of the solve-path
workloads (fixed seeds, NumPy PCG64 streams):

* 2-D 5-point and 3-D 7-point grid Laplacians, unit / lognormal / anisotropic
  edge weights;
* SDDM from a Laplacian by a diagonal excess U[0, f) * degree;
* R-MAT power-law graph Laplacians (duplicates merged, self-loops dropped);
* N(0,1) right-hand sides, projected onto the range of a singular Laplacian
  (per connected component, isolated vertices zeroed).

Every operator is returned as a scipy CSC matrix with both triangles stored.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg


def _laplacian_from_edges(n: int, u: np.ndarray, v: np.ndarray, w: np.ndarray) -> sp.csc_matrix:
    W = sp.coo_matrix((w, (u, v)), shape=(n, n)).tocsr()
    W = W + W.T
    deg = np.asarray(W.sum(axis=1)).ravel()
    L = (sp.diags(deg) - W).tocsc()
    L.sum_duplicates()
    L.sort_indices()
    return L


def grid_edges(dims: tuple[int, ...]):
    """Edges of the nearest-neighbour grid of shape dims (C order), by axis."""
    n = int(np.prod(dims))
    idx = np.arange(n).reshape(dims)
    out = []
    for ax in range(len(dims)):
        sl_a = [slice(None)] * len(dims)
        sl_b = [slice(None)] * len(dims)
        sl_a[ax] = slice(0, -1)
        sl_b[ax] = slice(1, None)
        out.append((idx[tuple(sl_a)].ravel(), idx[tuple(sl_b)].ravel()))
    return n, out


def grid_laplacian(dims, weights: str = "unit", sigma: float = 1.0, seed: int = 1,
                   axis_weights=None) -> sp.csc_matrix:
    """Grid Laplacian; weights 'unit' or 'lognormal' (exp(N(0, sigma))), times
    optional per-axis factors (anisotropy)."""
    n, edges = grid_edges(tuple(dims))
    rng = np.random.default_rng(seed)
    us, vs, ws = [], [], []
    for ax, (a, b) in enumerate(edges):
        if weights == "unit":
            w = np.ones(a.size)
        elif weights == "lognormal":
            w = np.exp(rng.normal(0.0, sigma, a.size))
        else:
            raise ValueError(weights)
        if axis_weights is not None:
            w = w * float(axis_weights[ax])
        us.append(a); vs.append(b); ws.append(w)
    return _laplacian_from_edges(n, np.concatenate(us), np.concatenate(vs), np.concatenate(ws))


def add_excess(L: sp.csc_matrix, frac: float = 1e-3, seed: int = 2) -> sp.csc_matrix:
    """SDDM: diagonal += U[0, frac) * weighted degree."""
    rng = np.random.default_rng(seed)
    deg = L.diagonal()
    return (L + sp.diags(rng.uniform(0.0, frac, L.shape[0]) * deg)).tocsc()


def _rmat_endpoints(scale: int, m: int, a: float, b: float, c: float, seed: int, chunk: int = 1 << 20):
    """The m R-MAT samples: bit `bit` of sample i comes from draw bit * m + i of
    one PCG64 stream seeded with `seed` (u gets 1 when the draw is >= a + b, v
    when it lies in [a, a + b) or >= a + b + c). Chunks of samples are drawn
    from the stream advanced to their position, on a few threads; the result
    is that of drawing the whole stream in order."""
    from concurrent.futures import ThreadPoolExecutor
    ab, abc = a + b, a + b + c
    u = np.empty(m, dtype=np.int64)
    v = np.empty(m, dtype=np.int64)

    def work(s0: int) -> None:
        s1 = min(m, s0 + chunk)
        uu = np.zeros(s1 - s0, dtype=np.int64)
        vv = np.zeros(s1 - s0, dtype=np.int64)
        for bit in range(scale):
            bg = np.random.PCG64(seed)
            bg.advance(bit * m + s0)
            r = np.random.Generator(bg).random(s1 - s0)
            uu |= (r >= ab).astype(np.int64) << bit
            vv |= (((r >= a) & (r < ab)) | (r >= abc)).astype(np.int64) << bit
        u[s0:s1] = uu
        v[s0:s1] = vv

    with ThreadPoolExecutor(max(1, min(8, (m + chunk - 1) // chunk))) as ex:
        list(ex.map(work, range(0, m, chunk)))
    return u, v


def rmat_laplacian(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
                   seed: int = 3, wseed: int = 4, wlo: float = 1.0, whi: float = 10.0) -> sp.csc_matrix:
    """R-MAT graph Laplacian: 2^scale vertices, edge_factor * 2^scale samples,
    symmetrized, duplicates merged, self-loops dropped, weights U[wlo, whi)."""
    n = 1 << scale
    m = edge_factor * n
    u, v = _rmat_endpoints(scale, m, a, b, c, seed)
    keep = u != v
    u, v = u[keep], v[keep]
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    key = np.unique(lo * n + hi)
    lo, hi = key // n, key % n
    w = np.random.default_rng(wseed).uniform(wlo, whi, lo.size)
    # Isolated vertices stay empty rows/columns of a singular Laplacian; the
    # elimination gives them a unit factor column (factorization_impl.h:200-206).
    L = _laplacian_from_edges(n, lo, hi, w)
    L.eliminate_zeros()
    return L


def rhs(n: int, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal(n)


def project_range(L: sp.csc_matrix, b: np.ndarray) -> np.ndarray:
    """Project b onto the range of a (possibly disconnected) Laplacian: zero
    on isolated vertices, mean removed per connected component."""
    A = L.copy()
    A.setdiag(0)
    A.eliminate_zeros()
    ncomp, lab = csg.connected_components(A, directed=False)
    out = b.astype(np.float64).copy()
    deg = np.diff(A.indptr)
    sums = np.bincount(lab, weights=out, minlength=ncomp)
    cnt = np.bincount(lab, minlength=ncomp)
    out -= (sums / cnt)[lab]
    out[deg == 0] = 0.0
    return out


def is_singular_laplacian(L: sp.csc_matrix) -> bool:
    return bool(np.abs(np.asarray(L.sum(axis=0)).ravel()).max() <= 1e-12 * np.abs(L.diagonal()).max())


# The named workloads of the solve path.
def workload(name: str):
    """Return (L, b) of a named workload."""
    if name == "grid2d_256":
        L = grid_laplacian((256, 256))
        b = rhs(L.shape[0], 0)
        return L, b - b.mean()
    if name == "sddm3d_128":
        L = add_excess(grid_laplacian((128, 128, 128), "lognormal", 1.0, seed=1), 1e-3, seed=2)
        return L, rhs(L.shape[0], 0)
    if name.startswith("sddm3d_"):
        k = int(name.split("_")[1])
        L = add_excess(grid_laplacian((k, k, k), "lognormal", 1.0, seed=1), 1e-3, seed=2)
        return L, rhs(L.shape[0], 0)
    if name.startswith("aniso3d_"):
        k = int(name.split("_")[1])
        L = grid_laplacian((k, k, k), "lognormal", 0.5, seed=5, axis_weights=(1.0, 10.0, 100.0))
        b = rhs(L.shape[0], 0)
        return L, b - b.mean()
    if name.startswith("rmat_"):
        s = int(name.split("_")[1])
        L = rmat_laplacian(s)
        return L, project_range(L, rhs(L.shape[0], 0))
    if name.startswith("grid2d_"):
        k = int(name.split("_")[1])
        L = grid_laplacian((k, k))
        b = rhs(L.shape[0], 0)
        return L, b - b.mean()
    if name.startswith("grid3d_"):
        k = int(name.split("_")[1])
        L = grid_laplacian((k, k, k))
        b = rhs(L.shape[0], 0)
        return L, b - b.mean()
    raise KeyError(name)


# The serving mix: 8 resident factors -- 2-D grids of 1e4 to 1e6 vertices,
# 3-D grids up to 96^3, R-MAT graphs of scale 16 to 20.
SERVING_FACTORS = ("grid2d_100", "grid2d_316", "grid2d_1000", "grid3d_48", "grid3d_96", "rmat_16", "rmat_18",
                   "rmat_20")


def request_rhs(L: sp.csc_matrix, sddm: bool, seed: int) -> np.ndarray:
    """One N(0,1) right-hand side (seed), projected onto the range of a
    Laplacian per connected component."""
    b = rhs(L.shape[0], seed)
    return b if sddm else project_range(L, b)
