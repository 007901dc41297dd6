"""Many small factors in one launch (apx_group_solve, solver_group): the PCG
solves of several member solvers run together, every kernel kind of an
iteration one launch over all members.

Bars: every member's solution, iteration count, status and residual are the
bits of that member's own solve (same solver, same right-hand side, same
tol / max_iter), in fp64 and mixed precision, with members of different
sizes and kinds (Laplacian grids, an SDDM grid, an R-MAT Laplacian with hub
rows, a zero right-hand side) whose sweep schedules have different numbers
of launched levels and tail runs; repeated group solves give the same bits;
the group needs far fewer launches per iteration than its members' own
solves together; and the group refuses what it cannot run."""
import numpy as np
import pytest

import apxchol_amd as ax
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu

MEMBERS = {
    "grid2d_64": lambda: mg.grid_laplacian((64, 64)),
    "sddm3d_16": lambda: mg.add_excess(mg.grid_laplacian((16, 16, 16), "lognormal")),
    "rmat_12": lambda: mg.rmat_laplacian(12),
    "grid2d_100": lambda: mg.grid_laplacian((100, 100)),
    "aniso3d_12": lambda: mg.grid_laplacian((12, 12, 12), "lognormal", 0.5, seed=5, axis_weights=(1.0, 10.0, 100.0)),
}
_cache = {}


def member(name):
    if name not in _cache:
        A = MEMBERS[name]()
        _cache[name] = (A, ax.factorize(A, seed=0))
    return _cache[name]


def bits(t):
    import torch
    return t.view(torch.int64)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_group_solve_is_every_members_own_solve(gpu, precision):
    import torch
    print(f"[case] group solve {precision}", flush=True)
    names = list(MEMBERS)
    solvers, bs = [], []
    for i, name in enumerate(names):
        A, F = member(name)
        solvers.append(ax.Solver(A, factor=F, precision=precision))
        b = mg.request_rhs(A, F.sddm, 50 + i)
        bs.append(torch.tensor(b, device=gpu))
    # A zero right-hand side stops at once (status 5) inside the group too.
    A0, F0 = member("grid2d_64")
    solvers.append(ax.Solver(A0, factor=F0, precision=precision))
    bs.append(torch.zeros(A0.shape[0], dtype=torch.float64, device=gpu))
    g = ax.Group(solvers)
    # One solve's iteration: SpMV, update_xr, dot_center, update_p and three
    # reductions, one launch per launched level, two per tail run.
    own_launches = sum(7 + sum(s.info["level_launches"]) + 2 * sum(s.info["tail_runs"]) for s in solvers)
    assert g.launches_per_iteration < own_launches
    xs, rs = g.solve(bs, tol=1e-8, maxiter=300)
    xs2, rs2 = g.solve(bs, tol=1e-8, maxiter=300)
    torch.cuda.synchronize()
    for i, s in enumerate(solvers):
        own = s.solve(bs[i], tol=1e-8, maxiter=300)
        torch.cuda.synchronize()
        assert (rs[i].status, rs[i].iterations) == (own.status, own.iterations), (i, rs[i], own)
        assert np.float64(rs[i].residual).view(np.int64) == np.float64(own.residual).view(np.int64), i
        assert torch.equal(bits(xs[i]), bits(own.x)), i
        assert torch.equal(bits(xs2[i]), bits(xs[i])), i
    assert rs[-1].status == 5 and float(xs[-1].abs().max()) == 0.0
    for i, name in enumerate(names):
        A, _ = member(name)
        x = xs[i].cpu().numpy()
        b = bs[i].cpu().numpy()
        assert rs[i].status == 1
        assert np.linalg.norm(b - A @ x) <= 1.05e-8 * np.linalg.norm(b), name
    g.close()
    for s in solvers:
        s.close()


def test_group_refuses_mixed_members_and_serving_solvers(gpu):
    import torch
    A, F = member("grid2d_64")
    a = ax.Solver(A, factor=F, precision="fp64")
    b = ax.Solver(A, factor=F, precision="fp32")
    with pytest.raises(ValueError, match="precision"):
        ax.Group([a, b])
    srv = ax.Server(slots=8)
    srv.add(b)
    with pytest.raises(ValueError, match="serving"):
        ax.Group([a, b])
    srv.close()
    g = ax.Group([a])
    with pytest.raises(ValueError, match="one right-hand side per member"):
        g.solve([])
    with pytest.raises(ValueError):
        g.solve([torch.zeros(5, dtype=torch.float64, device=gpu)])
    g.close()
    a.close()
    b.close()
