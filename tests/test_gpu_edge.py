"""Degenerate systems through every GPU entry point: a 1x1 SDDM system, 2x2
SDDM and Laplacian systems, a 3-vertex path Laplacian, an SDDM system with
an isolated vertex and a Laplacian of two components (right-hand side
projected per component), each in fp64 and mixed precision.

Bars: every solve converges (status 1) to a true residual ||b - A x|| / ||b||
<= 1e-8 (a zero right-hand side returns status 5 and x = 0); the fp64
preconditioner application equals the oracle's to 1e-12; blocked solves and
serving requests converge the same way; a grouped solve over all of them
gives every member the bits of its own solve."""
import numpy as np
import pytest
import scipy.sparse as sp

import apxchol_amd as ax
import oracle
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu


def _two_components():
    A = sp.lil_matrix((5, 5))
    for i, j, w in ((0, 1, 1.0), (1, 2, 2.0), (3, 4, 0.5)):
        A[i, j] = A[j, i] = -w
    A = A.tocsc()
    d = -np.asarray(A.sum(axis=0)).ravel()
    return (A + sp.diags(d)).tocsc()


CASES = {
    "sddm_n1": lambda: sp.csc_matrix(np.array([[2.0]])),
    "sddm_n2": lambda: sp.csc_matrix(np.array([[2.0, -1.0], [-1.0, 3.0]])),
    "laplacian_n2": lambda: sp.csc_matrix(np.array([[1.0, -1.0], [-1.0, 1.0]])),
    "laplacian_path3": lambda: sp.csc_matrix(np.array([[1.0, -1.0, 0.0], [-1.0, 2.0, -1.0], [0.0, -1.0, 1.0]])),
    "sddm_isolated_vertex": lambda: sp.csc_matrix(np.array([[2.0, -1.0, 0.0], [-1.0, 2.0, 0.0], [0.0, 0.0, 5.0]])),
    "laplacian_two_components": _two_components,
}


def true_res(A, b, x):
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def rhs(A, F, seed):
    return mg.request_rhs(A, F.sddm, seed)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", list(CASES))
def test_degenerate_solve_batch_and_apply(gpu, name, precision):
    import torch
    print(f"[case] degenerate {name} {precision}", flush=True)
    A = CASES[name]()
    F = ax.factorize(A, seed=0)
    n = A.shape[0]
    s = ax.Solver(A, factor=F, precision=precision)
    b = rhs(A, F, 3)
    r = s.solve(torch.tensor(b, device=gpu))
    x = r.x.cpu().numpy()
    assert r.status == 1, r.status
    assert true_res(A, b, x) <= 1.05e-8, true_res(A, b, x)
    z = s.solve(torch.zeros(n, dtype=torch.float64, device=gpu))
    assert z.status == 5 and float(z.x.abs().max()) == 0.0
    B = np.stack([rhs(A, F, 10), np.zeros(n), rhs(A, F, 11)])
    X, res = s.solve_batch(torch.tensor(B, device=gpu))
    X = X.cpu().numpy()
    assert res[1].status == 5 and np.abs(X[1]).max() == 0.0
    for c in (0, 2):
        assert res[c].status == 1, (c, res[c].status)
        assert true_res(A, B[c], X[c]) <= 1.05e-8, (c, true_res(A, B[c], X[c]))
    if precision == "fp64":
        o = oracle.Oracle(A, F.export(), F.sddm)
        s.reset_apply_count()
        rv = np.random.default_rng(5).standard_normal(n)
        za = s.apply(torch.tensor(rv, device=gpu)).cpu().numpy()
        zo = o.apply(rv)
        assert np.max(np.abs(za - zo)) <= 1e-12 * max(np.max(np.abs(zo)), 1e-300)
    s.close()


def test_degenerate_factors_serve(gpu):
    print("[case] degenerate serving", flush=True)
    srv = ax.Server(slots=4)
    items = []
    for i, (name, make) in enumerate(CASES.items()):
        A = make()
        F = ax.factorize(A, seed=0)
        fid = srv.add(ax.Solver(A, factor=F, precision="fp32"))
        for k in range(3):
            b = rhs(A, F, 20 + 3 * i + k)
            items.append((name, A, b, srv.submit(fid, b, 1e-8)))
    for name, A, b, rid in items:
        x, r = srv.wait(rid, timeout=120)
        assert r.status == 1, (name, r.status)
        assert true_res(A, b, x) <= 2e-8, (name, true_res(A, b, x))
    srv.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_degenerate_factors_in_one_group(gpu, precision):
    """A grouped solve over the degenerate systems gives every member the bits
    of its own solve."""
    import torch
    print(f"[case] degenerate group {precision}", flush=True)
    solvers, bs, mats = [], [], []
    for i, make in enumerate(CASES.values()):
        A = make()
        F = ax.factorize(A, seed=0)
        solvers.append(ax.Solver(A, factor=F, precision=precision))
        bs.append(torch.tensor(rhs(A, F, 40 + i), device=gpu))
        mats.append(A)
    g = ax.Group(solvers)
    xs, rs = g.solve(bs, tol=1e-8, maxiter=200)
    torch.cuda.synchronize()
    for i, s in enumerate(solvers):
        own = s.solve(bs[i], tol=1e-8, maxiter=200)
        torch.cuda.synchronize()
        assert (rs[i].status, rs[i].iterations) == (own.status, own.iterations), i
        assert torch.equal(xs[i].view(torch.int64), own.x.view(torch.int64)), i
        assert true_res(mats[i], bs[i].cpu().numpy(), xs[i].cpu().numpy()) <= 1.05e-8, i
    g.close()
    for s in solvers:
        s.close()
