"""GPU parity of the HIP solve path against the CPU oracle (one-thread
restatement of the reference's cpu_solver), through the C ABI.

Bars (stated per test): sweeps and apply with fp64 vectors within 1e-12
relative of the oracle and bit-identical run to run and across schedules;
PCG to the same tolerance within +-2 iterations of the oracle; fp32/fp16 sweeps
within their storage's precision and converging to 1e-8."""
import numpy as np
import pytest

import apxchol_amd as ax
import oracle
from apxchol_amd import matgen as mg

pytestmark = pytest.mark.gpu

CASES = {
    "grid2d_64": lambda: mg.grid_laplacian((64, 64)),
    "grid2d_256": lambda: mg.grid_laplacian((256, 256)),          # fat levels (> 1024 rows)
    "sddm3d_16": lambda: mg.add_excess(mg.grid_laplacian((16, 16, 16), "lognormal")),
    "aniso3d_16": lambda: mg.grid_laplacian((16, 16, 16), "lognormal", 0.5, seed=5, axis_weights=(1, 10, 100)),
    "rmat_12": lambda: mg.rmat_laplacian(12),
    "rmat_16": lambda: mg.rmat_laplacian(16),                     # BK residual + 500-column peel chain
}
_cache = {}


def case(name):
    if name not in _cache:
        A = CASES[name]()
        F = ax.factorize(A, seed=0)
        _cache[name] = (A, F, F.export())
    return _cache[name]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("schedule", ["hybrid", "dataflow", "levels"])
def test_sweeps_fp64_match_oracle(gpu, name, schedule):
    import torch
    A, F, fac = case(name)
    print(f"[case] sweeps {name} {schedule}", flush=True)
    s = ax.Solver(A, factor=F, schedule=schedule)
    o = oracle.Oracle(A, fac, F.sddm)
    m = s.info["m"]
    x = np.random.default_rng(7).standard_normal(m)
    xd = torch.tensor(x, device=gpu)
    yd = torch.empty(m, dtype=torch.float64, device=gpu)
    zd = torch.empty(m, dtype=torch.float64, device=gpu)
    s.sweep(xd, yd, backward=False)
    s.sweep(yd, zd, backward=True)
    torch.cuda.synchronize()
    y_ref = o.forward(x)
    z_ref = o.back(y_ref)
    y, z = yd.cpu().numpy(), zd.cpu().numpy()
    assert rel(y, y_ref) < 1e-12
    assert rel(z, o.back(y)) < 1e-12
    assert rel(z, z_ref) < 1e-11
    # Run to run: bit-identical.
    y2 = torch.empty_like(yd)
    s.sweep(xd, y2, backward=False)
    torch.cuda.synchronize()
    assert torch.equal(y2.view(torch.int64), yd.view(torch.int64))
    s.close()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("fmt", ["auto", "csr", "sell"])
def test_spmv_matches_operator(gpu, name, fmt, monkeypatch):
    """The single-vector SpMV in each layout -- CSR lanes, sliced ELL in row
    order, sliced ELL sorted by length with a row map and wave-walked long
    rows (R-MAT) -- against scipy's fp64 A x, row by row within 1e-13 of
    (|A| |x|)_i (the rounding of a fixed-order sum of <= a few hundred terms)."""
    import torch
    A, F, _ = case(name)
    print(f"[case] spmv {name} {fmt}", flush=True)
    if fmt != "auto":
        monkeypatch.setenv("APXCHOL_AMD_SPMV", fmt)
    s = ax.Solver(A, factor=F)
    if fmt == "csr":
        assert s.info["spmv_format"] == 0 and s.info["spmv_sell_slots"] == 0
    if fmt == "sell":
        assert s.info["spmv_format"] == 1 and s.info["spmv_sell_slots"] > 0   # long rows are not in slices
    x = np.random.default_rng(12).standard_normal(A.shape[0])
    y = s.spmv(torch.tensor(x, device=gpu)).cpu().numpy()
    y_ref = A @ x
    bound = abs(A) @ abs(x)
    assert np.all(np.abs(y - y_ref) <= 1e-13 * bound + 1e-300), float(np.max(np.abs(y - y_ref) / (bound + 1e-300)))
    y2 = s.spmv(torch.tensor(x, device=gpu)).cpu().numpy()
    assert np.array_equal(y.view(np.int64), y2.view(np.int64))
    s.close()


SCHEDULE_CONFIGS = [
    dict(schedule="hybrid"),
    dict(schedule="dataflow"),
    dict(schedule="levels"),
    dict(schedule="levels", lane_batches=-1),    # every level from the batch tables
    dict(schedule="hybrid", lane_batches=-1),
    dict(schedule="levels", lane_batches=1),     # lane layout for one-batch levels only
    dict(schedule="hybrid", tail_batches=64, tail_rows=4096),   # tail waves take several batches a level
    dict(schedule="hybrid", tail_cut=-1),        # tail runs end only at launched levels / LDS capacity
    dict(schedule="hybrid", tail_cut=1),         # a run ends before any level with overflow reads
]


@pytest.mark.parametrize("name", ["grid2d_64", "grid2d_256", "sddm3d_16", "rmat_16"])
@pytest.mark.parametrize("precision,fp16", [("fp64", False), ("fp32", False), ("fp32", True)])
def test_schedules_bit_identical(gpu, name, precision, fp16):
    """Every schedule -- the dataflow launch, per-level launches from the batch
    tables or from the static lane layout, the hybrid with its one-workgroup
    tail -- gives the same bits: two sweeps with fp64 vectors, and one
    preconditioner application in the solver's own vector precision."""
    import torch
    A, F, fac = case(name)
    print(f"[case] schedules {name} {precision} fp16={fp16}", flush=True)
    sweeps, applies = [], []
    for cfg in SCHEDULE_CONFIGS:
        s = ax.Solver(A, factor=F, precision=precision, fp16_storage=fp16, **cfg)
        lanes = sum(s.info["lane_levels"])
        if cfg["schedule"] == "dataflow" or cfg.get("lane_batches", 0) < 0:
            assert lanes == 0, cfg
        elif cfg.get("lane_batches", 0) == 0:
            assert lanes > 0 or sum(s.info["level_launches"]) == 0, cfg
        m = s.info["m"]
        x = torch.tensor(np.random.default_rng(3).standard_normal(m), device=gpu)
        y = torch.empty_like(x)
        z = torch.empty_like(x)
        s.sweep(x, y, backward=False)
        s.sweep(y, z, backward=True)
        r = torch.tensor(np.random.default_rng(4).standard_normal(A.shape[0]), device=gpu)
        s.reset_apply_count()
        a = s.apply(r)
        torch.cuda.synchronize()
        sweeps.append(z.cpu().numpy().view(np.int64).copy())
        applies.append(a.cpu().numpy().view(np.int64).copy())
        s.close()
    for k in range(1, len(SCHEDULE_CONFIGS)):
        assert np.array_equal(sweeps[0], sweeps[k]), SCHEDULE_CONFIGS[k]
        assert np.array_equal(applies[0], applies[k]), SCHEDULE_CONFIGS[k]


def test_tail_cut_splits_runs_at_long_rows(gpu):
    """The overflow cut only moves run boundaries: on a power-law factor
    (hub rows of thousands of entries) cutting before every level with
    overflow reads yields more tail runs and no more overflow-heavy levels,
    with the same bits as never cutting (checked by the schedule test)."""
    A, F, fac = case("rmat_16")
    print("[case] tail cut rmat_16", flush=True)
    runs = {}
    for cut in (-1, 0, 1):
        s = ax.Solver(A, factor=F, precision="fp32", tail_cut=cut)
        runs[cut] = s.info["tail_runs"]
        s.close()
    assert runs[1][0] >= runs[0][0] >= runs[-1][0] and runs[1][1] >= runs[0][1] >= runs[-1][1], runs
    assert sum(runs[1]) > sum(runs[-1]), runs


@pytest.mark.parametrize("name", ["sddm3d_16", "rmat_16"])
@pytest.mark.parametrize("rotate,phases", [("0", "1"), ("0", "0"), ("1", "0")])
def test_tail_dealing_and_phases_keep_the_bits(gpu, name, rotate, phases, monkeypatch):
    """How the tail runs deal their batches to the waves (round robin over the
    run, the default, or level by level) and whether chains of one-batch
    levels of one wave share a barrier phase only change which wave computes a
    row and when: both sweeps, in both precisions, keep the default's bits.
    These factors hold long chains of one-batch levels (rmat_16: ~460 per
    sweep), the case the shared phases exist for."""
    import torch
    A, F, fac = case(name)
    print(f"[case] tail dealing {name} rotate={rotate} phases={phases}", flush=True)
    for precision in ("fp64", "fp32"):
        monkeypatch.delenv("APXCHOL_AMD_TAIL_ROTATE", raising=False)
        monkeypatch.delenv("APXCHOL_AMD_TAIL_PHASES", raising=False)
        ref = ax.Solver(A, factor=F, precision=precision)
        monkeypatch.setenv("APXCHOL_AMD_TAIL_ROTATE", rotate)
        monkeypatch.setenv("APXCHOL_AMD_TAIL_PHASES", phases)
        s = ax.Solver(A, factor=F, precision=precision)
        m = s.info["m"]
        x = torch.tensor(np.random.default_rng(17).standard_normal(m), device=gpu)
        for backward in (False, True):
            a, b = torch.empty_like(x), torch.empty_like(x)
            ref.sweep(x, a, backward=backward)
            s.sweep(x, b, backward=backward)
            torch.cuda.synchronize()
            assert np.array_equal(a.cpu().numpy().view(np.int64), b.cpu().numpy().view(np.int64)), (precision, backward)
        ref.close()
        s.close()


@pytest.mark.parametrize("name", ["grid2d_64", "grid2d_256", "sddm3d_16", "aniso3d_16", "rmat_12"])
def test_apply_matches_cpu_apply(gpu, name):
    """cpu_solver::apply parity over 12 consecutive applications (the center-k
    schedule centres the 10th): 1e-12 relative."""
    A, F, fac = case(name)
    print(f"[case] apply {name}", flush=True)
    s = ax.Solver(A, factor=F)
    o = oracle.Oracle(A, fac, F.sddm)
    rg = np.random.default_rng(11)
    for k in range(12):
        r = rg.standard_normal(A.shape[0])
        z = s.apply(r)
        zr = o.apply(r)
        assert rel(z, zr) < 1e-12, f"application {k + 1}"
    s.close()


@pytest.mark.parametrize("name", ["grid2d_64", "grid2d_256", "sddm3d_16", "aniso3d_16", "rmat_12", "rmat_16"])
def test_pcg_fp64_matches_oracle(gpu, name):
    A, F, fac = case(name)
    print(f"[case] pcg {name}", flush=True)
    b = np.random.default_rng(0).standard_normal(A.shape[0])
    if not F.sddm:
        b = mg.project_range(A, b)
    s = ax.Solver(A, factor=F)
    res = s.solve(b, tol=1e-8)
    o = oracle.Oracle(A, fac, F.sddm)
    xo, ito, reso = o.solve(b, tol=1e-8)
    true_res = np.linalg.norm(b - A @ res.x) / np.linalg.norm(b)
    assert res.status == 1, res
    assert abs(res.iterations - ito) <= 2, (res.iterations, ito)
    assert true_res < 2e-8
    assert rel(res.x, xo) < 1e-4, rel(res.x, xo)
    s.close()


@pytest.mark.parametrize("name", ["grid2d_256", "sddm3d_16", "rmat_16"])
def test_pcg_mixed_precision_converges(gpu, name):
    """FP32 sweeps (fp32 vectors, fp32 accumulation) with FP64 CG reach 1e-8
    within a few iterations of the FP64-sweep solve."""
    A, F, fac = case(name)
    print(f"[case] mixed {name}", flush=True)
    b = np.random.default_rng(0).standard_normal(A.shape[0])
    if not F.sddm:
        b = mg.project_range(A, b)
    s64 = ax.Solver(A, factor=F)
    s32 = ax.Solver(A, factor=F, precision="fp32")
    r64 = s64.solve(b, tol=1e-8)
    r32 = s32.solve(b, tol=1e-8)
    assert r32.status == 1
    assert np.linalg.norm(b - A @ r32.x) / np.linalg.norm(b) < 2e-8
    assert r32.iterations <= r64.iterations + 3
    s64.close()
    s32.close()


@pytest.mark.parametrize("name", ["grid2d_64", "sddm3d_16", "rmat_12"])
def test_fp16_storage_apply_matches_oracle(gpu, name):
    """fp16 per-column-scaled storage, fp64 vectors: the stored factor is the
    oracle's bit for bit, so the apply agrees to fp64 rounding."""
    A, F, fac = case(name)
    print(f"[case] fp16 {name}", flush=True)
    s = ax.Solver(A, factor=F, fp16_storage=True)
    o = oracle.Oracle(A, fac, F.sddm, fp16=True)
    r = np.random.default_rng(5).standard_normal(A.shape[0])
    assert rel(s.apply(r), o.apply(r)) < 1e-12
    b = r if F.sddm else mg.project_range(A, r)
    res = s.solve(b, tol=1e-8)
    assert res.status == 1 and np.linalg.norm(b - A @ res.x) / np.linalg.norm(b) < 2e-8
    s.close()


def test_solve_device_vectors_and_warm_start(gpu):
    import torch
    A, F, fac = case("sddm3d_16")
    s = ax.Solver(A, factor=F)
    b = np.random.default_rng(2).standard_normal(A.shape[0])
    bd = torch.tensor(b, device=gpu)
    r1 = s.solve(bd)
    x1 = r1.x.cpu().numpy()
    assert r1.status == 1
    # Warm start from the solution: converged at iteration 0.
    r2 = s.solve(bd, x0=r1.x)
    assert r2.iterations == 0 and r2.status == 1
    # Repeated solves are bit-identical.
    r3 = s.solve(bd)
    assert np.array_equal(x1.view(np.int64), r3.x.cpu().numpy().view(np.int64))
    assert r3.iterations == r1.iterations
    s.close()


@pytest.mark.parametrize("name", ["grid2d_256", "sddm3d_16", "rmat_16"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replay_matches_eager(gpu, name, precision):
    """The captured iteration graph, replayed many times per solve and across
    solves with eager work and synchronizations in between, gives the eager
    launches' bits exactly (x, iterations, residual)."""
    import torch
    A, F, fac = case(name)
    print(f"[case] graph vs eager {name} {precision}", flush=True)
    b = mg.project_range(A, np.random.default_rng(3).standard_normal(A.shape[0])) if not F.sddm else \
        np.random.default_rng(3).standard_normal(A.shape[0])
    bd = torch.tensor(b, device=gpu)
    eager = ax.Solver(A, factor=F, precision=precision, graph=False)
    ref = eager.solve(bd)
    x_ref = ref.x.cpu().numpy()
    eager.close()
    g = ax.Solver(A, factor=F, precision=precision, graph=True)
    for k in range(3):
        r = g.solve(bd)
        torch.cuda.synchronize()
        junk = (bd * 2.0).sum().item()        # eager work between replays
        assert np.isfinite(junk)
        assert r.iterations == ref.iterations and r.status == ref.status
        assert r.residual == ref.residual
        assert np.array_equal(x_ref.view(np.int64), r.x.cpu().numpy().view(np.int64)), f"solve {k}"
    # Timing forces eager launches in the same solver: still the same bits.
    g.set_timing(True)
    r = g.solve(bd)
    g.set_timing(False)
    tm = g.take_timing()
    assert np.array_equal(x_ref.view(np.int64), r.x.cpu().numpy().view(np.int64))
    assert tm["sweep_levels"][1] + tm["sweep_tail"][1] > 0
    g.close()


@pytest.mark.parametrize("name", ["grid2d_256", "sddm3d_16", "rmat_12", "rmat_16"])
@pytest.mark.parametrize("depth", [False, True])
def test_relabel_matches_factor_order(gpu, name, depth):
    """Sorting each level's rows by original vertex -- with the elimination
    rounds as levels, or with levels regrouped by topological depth -- changes
    only the storage order (and so the order of each row's entries): sweeps
    agree with the factor-order solver to 1e-12, solves within one iteration
    and 1e-6. Depth levels are never more than the rounds."""
    import torch
    A, F, fac = case(name)
    print(f"[case] relabel {name} depth={depth}", flush=True)
    s0 = ax.Solver(A, factor=F, relabel=False)
    s1 = ax.Solver(A, factor=F, relabel=True, depth_levels=depth)
    if depth:
        assert s1.info["levels"] <= s0.info["levels"]
    else:
        assert s1.info["levels"] == s0.info["levels"]
    m = s0.info["m"]
    x = torch.tensor(np.random.default_rng(8).standard_normal(m), device=gpu)
    for backward in (False, True):
        y0 = torch.empty_like(x)
        y1 = torch.empty_like(x)
        s0.sweep(x, y0, backward=backward)
        s1.sweep(x, y1, backward=backward)
        torch.cuda.synchronize()
        assert rel(y1.cpu().numpy(), y0.cpu().numpy()) < 1e-12
    b = np.random.default_rng(9).standard_normal(A.shape[0])
    if not F.sddm:
        b = mg.project_range(A, b)
    bd = torch.tensor(b, device=gpu)
    r0, r1 = s0.solve(bd), s1.solve(bd)
    assert abs(r0.iterations - r1.iterations) <= 1
    assert rel(r1.x.cpu().numpy(), r0.x.cpu().numpy()) < 1e-6
    s0.close()
    s1.close()


@pytest.mark.parametrize("name", ["path5_laplacian", "path5_sddm", "path5_nearly_sddm", "path5_pattern"])
def test_fixture_files_solve_end_to_end(gpu, name):
    """The reference CLI's path on its own fixture files (tests/fixtures/*.mtx):
    read, interpret (an adjacency file becomes L = D - A), solve on the GPU to
    tol 1e-8, graded by the true residual."""
    import torch
    from pathlib import Path
    from apxchol_amd import mtx
    print(f"[case] fixture {name}", flush=True)
    L, kind, _ = mtx.load_operator(Path(__file__).resolve().parent / "golden" / "fixtures" / f"{name}.mtx")
    assert kind == ("adjacency" if name == "path5_pattern" else "laplacian")
    b = np.random.default_rng(14).standard_normal(L.shape[0])
    if name in ("path5_laplacian", "path5_pattern"):
        b = mg.project_range(L, b)
    s = ax.Solver(L, seed=0)
    r = s.solve(torch.tensor(b, device=gpu))
    x = r.x.cpu().numpy()
    assert r.status == 1 and np.linalg.norm(b - L @ x) <= 2e-8 * np.linalg.norm(b)
    s.close()


def test_adjacency_graph_through_laplacian_converges(gpu):
    """An adjacency matrix converted with apxchol_amd.laplacian solves (the
    raw adjacency is refused by the operator contract, tested on the CPU)."""
    import torch
    import scipy.sparse as sp
    idx = np.arange(64 * 64).reshape(64, 64)
    a = np.concatenate([idx[:-1, :].ravel(), idx[:, :-1].ravel()])
    c = np.concatenate([idx[1:, :].ravel(), idx[:, 1:].ravel()])
    A = sp.coo_matrix((np.ones(2 * a.size), (np.concatenate([a, c]), np.concatenate([c, a]))), shape=(4096, 4096))
    L = ax.laplacian(A.tocsc())
    b = mg.project_range(L, np.random.default_rng(15).standard_normal(4096))
    s = ax.Solver(L, seed=0, precision="fp32")
    r = s.solve(torch.tensor(b, device=gpu))
    assert r.status == 1 and np.linalg.norm(b - L @ r.x.cpu().numpy()) <= 2e-8 * np.linalg.norm(b)
    s.close()


@pytest.mark.parametrize("name", ["sddm3d_16", "rmat_16"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lane_levels_several_batches_per_wave_same_bits(gpu, name, precision, monkeypatch):
    """Launched lane levels run 1, 2 or 4 batches per wave (thresholds taken
    from APXCHOL_AMD_LANE_BPW2 / _BPW4 when a solver is built; 1 forces the
    wider form on every launched level): both sweeps and an application give
    the same bits as one batch per wave."""
    import torch
    A, F, fac = case(name)
    print(f"[case] lane batches per wave {name} {precision}", flush=True)
    outs = []
    for bpw2, bpw4 in (("0", "0"), ("1", "0"), ("0", "1"), ("2", "5")):
        monkeypatch.setenv("APXCHOL_AMD_LANE_BPW2", bpw2)
        monkeypatch.setenv("APXCHOL_AMD_LANE_BPW4", bpw4)
        s = ax.Solver(A, factor=F, precision=precision, tail_rows=1, tail_batches=1)
        assert sum(s.info["lane_levels"]) > 0
        m = s.info["m"]
        x = torch.tensor(np.random.default_rng(3).standard_normal(m), device=gpu)
        y = torch.empty_like(x)
        z = torch.empty_like(x)
        s.sweep(x, y, backward=False)
        s.sweep(y, z, backward=True)
        r = torch.tensor(np.random.default_rng(4).standard_normal(A.shape[0]), device=gpu)
        s.reset_apply_count()
        a = s.apply(r)
        torch.cuda.synchronize()
        outs.append((z.cpu().numpy().view(np.int64).copy(), a.cpu().numpy().view(np.int64).copy()))
        s.close()
    for k in range(1, len(outs)):
        assert np.array_equal(outs[0][0], outs[k][0]) and np.array_equal(outs[0][1], outs[k][1]), k


@pytest.mark.parametrize("name", ["rmat_16"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_split_couplings_give_the_same_bits(gpu, name, precision, monkeypatch):
    """A tail run's coupling split in two launches (the entries before the
    previous run on the side stream, beside that run's tail; the rest after
    it) gives the one-launch chains: forced on every run that follows a run
    (APXCHOL_AMD_COUPLE_SPLIT_MIN=1), the sweeps, an application and a graph-
    replayed solve are the bits of the unsplit schedule."""
    import torch
    A, F, _ = case(name)
    print(f"[case] split couplings {name} {precision}", flush=True)
    out = {}
    for mode in ("off", "forced"):
        if mode == "off":
            monkeypatch.setenv("APXCHOL_AMD_COUPLE_SPLIT", "0")
        else:
            monkeypatch.delenv("APXCHOL_AMD_COUPLE_SPLIT", raising=False)
            monkeypatch.setenv("APXCHOL_AMD_COUPLE_SPLIT_MIN", "1")
        s = ax.Solver(A, factor=F, precision=precision)
        splits = sum(s.info["split_couplings"])
        assert splits == 0 if mode == "off" else splits > 0, (mode, s.info["tail_runs"])
        m = s.info["m"]
        x = torch.tensor(np.random.default_rng(3).standard_normal(m), device=gpu)
        y = torch.empty_like(x)
        z = torch.empty_like(x)
        s.sweep(x, y, backward=False)
        s.sweep(y, z, backward=True)
        b = torch.tensor(mg.request_rhs(A, F.sddm, 11), device=gpu)
        s.reset_apply_count()
        a = s.apply(b)
        r = s.solve(b)
        torch.cuda.synchronize()
        out[mode] = (z.cpu().numpy().view(np.int64).copy(), a.cpu().numpy().view(np.int64).copy(),
                     r.x.cpu().numpy().view(np.int64).copy(), r.iterations)
        s.close()
    for k in range(3):
        assert np.array_equal(out["off"][k], out["forced"][k]), k
    assert out["off"][3] == out["forced"][3]
