"""CPU tests (no GPU): the oracle and the product's host code against the
reference's golden fixtures, closed forms and the reference's own test
properties; the C ABI's exports.

Fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
reference built in place) hold, per case, the reference factorization at seed
0 (one thread), omp_sptrsv's stored factor and sweeps, and cuda_host's device
preparation. Bars: bit-identical for the factor, the stored arrays and the
oracle's sweeps; 1e-12 relative for the textbook substitution; the reference's
fp16 known answers (tests/test_lowprec.cpp:134-190) exactly."""
from __future__ import annotations

import math
import re
import struct
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import apxchol_amd as ax
import oracle
from apxchol_amd import matgen as mg

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def golden(name):
    g = dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))
    n = len(g["A_indptr"]) - 1
    A = sp.csc_matrix((g["A_data"], g["A_indices"], g["A_indptr"]), shape=(n, n))
    return A, g


def golden_factor(g):
    return (g["outer"].astype(np.uint64), g["inner"], g["vals"], g["perm"], g["round_is"])


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view({2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def test_golden_fixtures_present():
    assert {"path5_laplacian", "path5_sddm", "grid2d_12", "sddm3d_6", "aniso3d_6", "rmat_9"} <= set(CASES)


# ---- the factorization ------------------------------------------------------

@pytest.mark.parametrize("threads", ["1", "8"])
@pytest.mark.parametrize("name", CASES)
def test_product_factor_bit_identical_to_reference(name, threads, monkeypatch):
    """The reference's one-thread factor (fixtures written from oracle/_ref),
    bit for bit -- also when the rounds run on 8 host threads (every round
    and selection pass forced onto the team)."""
    monkeypatch.setenv("APXCHOL_AMD_FACTOR_THREADS", threads)
    monkeypatch.setenv("APXCHOL_AMD_FACTOR_PAR_MIN", "1")
    A, g = golden(name)
    F = ax.factorize(A, seed=0)
    assert F.threads == int(threads)
    outer, inner, vals, perm, ris = F.export()
    assert F.sddm == bool(g["sddm"])
    assert np.array_equal(outer, g["outer"].astype(np.uint64))
    assert np.array_equal(inner, g["inner"])
    assert np.array_equal(bits(vals), bits(g["vals"]))
    assert np.array_equal(perm, g["perm"])
    assert np.array_equal(ris, g["round_is"])


@pytest.mark.parametrize("gen", ["sddm3d_40", "rmat_14", "aniso3d_30"])
def test_threaded_factor_equals_one_thread_factor(gen, monkeypatch):
    """Larger graphs than the fixtures (main-loop rounds of thousands of
    vertices, Baumann-Kyng rounds over hubs): the factor on 1, 3 and 8 host
    threads, with the default and with a forced-parallel round threshold, is
    the same bits."""
    if gen == "sddm3d_40":
        A = mg.add_excess(mg.grid_laplacian((40, 40, 40), "lognormal"), 1e-3)
    elif gen == "rmat_14":
        A = mg.rmat_laplacian(14)
    else:
        A = mg.grid_laplacian((30, 30, 30), "lognormal", sigma=0.5, axis_weights=(1.0, 10.0, 100.0))
    ref = None
    for threads, par_min in (("1", "0"), ("3", "1"), ("8", "0"), ("8", "1")):
        monkeypatch.setenv("APXCHOL_AMD_FACTOR_THREADS", threads)
        monkeypatch.setenv("APXCHOL_AMD_FACTOR_PAR_MIN", par_min)
        e = ax.factorize(A, seed=3).export()
        if ref is None:
            ref = e
            continue
        for a, b in zip(ref, e):
            assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)), (threads, par_min)


@pytest.mark.parametrize("name", CASES)
def test_reference_build_reproduces_fixture(name):
    refdrv = pytest.importorskip("refdrv")
    if not refdrv.available():
        pytest.skip("oracle/_ref not built (no reference tree here)")
    A, g = golden(name)
    outer, inner, vals, perm, ris, _, sddm = refdrv.factorize(A, 0)
    assert np.array_equal(inner, g["inner"]) and np.array_equal(bits(vals), bits(g["vals"]))
    assert np.array_equal(perm, g["perm"]) and sddm == bool(g["sddm"])


def _dense_factor(F_or_export, n):
    outer, inner, vals = F_or_export[:3]
    return sp.csc_matrix((vals.astype(np.float64), inner.astype(np.int64), outer.astype(np.int64)), shape=(n, n))


@pytest.mark.parametrize("graph", ["path", "cycle", "star", "tree"])
def test_factor_exact_on_degree_two_eliminations(graph):
    """Eliminating a vertex of degree <= 2 is exact (its clique is one edge of
    weight w1 w2 / (w1 + w2)), so on paths, cycles, stars (leaves first) and
    trees L L^T = P A P^T up to the fp32 storage of L (but the singular
    Laplacian's last pivot)."""
    rng = np.random.default_rng(11)
    n = 40
    if graph == "path":
        edges = [(i, i + 1) for i in range(n - 1)]
    elif graph == "cycle":
        edges = [(i, (i + 1) % n) for i in range(n)]
    elif graph == "star":
        edges = [(0, i) for i in range(1, n)]
    else:
        edges = [(int(rng.integers(0, i)), i) for i in range(1, n)]
    w = rng.uniform(1.0, 10.0, len(edges))
    r = [a for a, _ in edges] + [b for _, b in edges]
    c = [b for _, b in edges] + [a for a, _ in edges]
    W = sp.csc_matrix((np.concatenate([w, w]), (r, c)), shape=(n, n))
    A = sp.csc_matrix(sp.diags(np.asarray(W.sum(axis=1)).ravel()) - W)
    F = ax.factorize(A, seed=0)
    ex = F.export()
    L = _dense_factor(ex, n).toarray()
    P = np.zeros((n, n))
    P[ex[3].astype(np.int64), np.arange(n)] = 1.0        # row perm[v] of P A P^T is vertex v
    PAP = P @ A.toarray() @ P.T
    D = L @ L.T - PAP
    # A Laplacian's last pivot is zero; the factor stores 1 there
    # (a vertex with no neighbours and no excess, factorization_impl.h:199-204),
    # outside L11.
    assert not F.sddm and L[n - 1, n - 1] == 1.0
    D[n - 1, n - 1] = 0.0
    err = np.abs(D).max() / np.abs(PAP).max()
    assert err < 1e-6, err


@pytest.mark.parametrize("name", ["grid2d_12", "sddm3d_6", "rmat_9"])
def test_factor_properties_of_the_reference_tests(name):
    """tests/test_factorize.cpp:193-254, 365-402, 548-560: dimensions, valid
    permutation, lower triangular with the diagonal first, positive diagonal
    (but the Laplacian's last), deterministic, seed-dependent, sddm flag."""
    A, g = golden(name)
    n = A.shape[0]
    F = ax.factorize(A, seed=0)
    outer, inner, vals, perm, ris = F.export()
    assert F.n == n and len(outer) == n + 1
    assert np.array_equal(np.sort(perm), np.arange(n))
    for j in range(n):
        col = inner[outer[j]:outer[j + 1]]
        assert col[0] == j and np.all(col[1:] > j) and np.all(np.diff(col[1:].astype(np.int64)) > 0)
        if j < n - 1 or F.sddm:
            assert vals[outer[j]] > 0
    F2 = ax.factorize(A, seed=0)
    assert np.array_equal(bits(F2.export()[2]), bits(vals))
    F3 = ax.factorize(A, seed=7)
    assert not np.array_equal(bits(F3.export()[2]), bits(vals)) or len(vals) == n
    assert F.sddm == (name == "sddm3d_6")
    assert int(np.sum(ris)) <= n


# ---- the stored factor and the sweeps ----------------------------------------

@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("storage", ["f", "h"])
def test_oracle_sweeps_bit_identical_to_reference(name, storage):
    """The oracle (one-thread restatement of omp_sptrsv) against the
    reference's own sweeps of x on the reference factor."""
    A, g = golden(name)
    O = oracle.Oracle(A, golden_factor(g), bool(g["sddm"]), fp16=(storage == "h"))
    a = O.arrays()
    assert np.array_equal(a["csr_ptr"].astype(np.uint32), g[f"{storage}_csr_ptr"])
    assert np.array_equal(a["csr_idx"], g[f"{storage}_csr_idx"])
    assert np.array_equal(a["csc_ptr"].astype(np.uint32), g[f"{storage}_csc_ptr"])
    assert np.array_equal(a["csc_idx"], g[f"{storage}_csc_idx"])
    if storage == "h":
        assert np.array_equal(a["csr_h"], g["h_csr_v16"]) and np.array_equal(a["csc_h"], g["h_csc_v16"])
    else:
        assert np.array_equal(bits(a["csr_v"]), bits(g["f_csr_v32"]))
        assert np.array_equal(bits(a["csc_v"]), bits(g["f_csc_v32"]))
    y = O.forward(g["x"])
    assert np.array_equal(bits(y), bits(g[f"{storage}_fwd"]))
    z = O.back(y)
    assert np.array_equal(bits(z), bits(g[f"{storage}_bck"]))


@pytest.mark.parametrize("name", CASES)
def test_reference_sweeps_are_the_textbook_substitution(name):
    """Closed form: forward / back substitution with the stored fp32 factor in
    fp64 (scipy's triangular solve) equals the fixture's sweeps to 1e-12."""
    A, g = golden(name)
    ptr, idx, v = g["f_csc_ptr"].astype(np.int64), g["f_csc_idx"].astype(np.int64), g["f_csc_v32"].astype(np.float64)
    m = len(ptr) - 1
    L = sp.csr_matrix(sp.csc_matrix((v, idx, ptr), shape=(m, m)))
    x = g["x"]
    y = spla.spsolve_triangular(L, x, lower=True)
    z = spla.spsolve_triangular(sp.csr_matrix(L.T), g["f_fwd"], lower=False)
    assert np.linalg.norm(y - g["f_fwd"]) <= 1e-12 * np.linalg.norm(y)
    assert np.linalg.norm(z - g["f_bck"]) <= 1e-12 * np.linalg.norm(z)


def _columns(ptr, idx, vals, j):
    return idx[ptr[j]:ptr[j + 1]], vals[ptr[j]:ptr[j + 1]]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("fp16", [False, True])
def test_product_device_layout_matches_reference_prep(name, fp16):
    """The product's host preparation (L11, drop, fp16 narrowing) against
    cuda_host's: the same columns, values, diagonal and inverse scales. The
    product stores each back-sweep column without its diagonal and rows
    descending, and the forward rows as their transpose."""
    A, g = golden(name)
    F = ax.factorize(A, seed=0)
    T = F.trsv(fp16=fp16)
    tag = "h" if fp16 else "f"
    gp, gi = g[f"{tag}_gpu_ptr"].astype(np.int64), g[f"{tag}_gpu_idx"].astype(np.int64)
    gv = g["h_gpu_v16"] if fp16 else bits(g["f_gpu_v32"])
    m = T["m"]
    assert len(gp) == m + 1
    bp, bi = T["b_ptr"].astype(np.int64), T["b_idx"].astype(np.int64)
    bv = T["b_h"] if fp16 else bits(T["b_val"])
    for j in range(m):
        ri, rv = _columns(gp, gi, gv, j)
        assert ri[0] == j
        pi, pv = _columns(bp, bi, bv, j)
        assert np.array_equal(pi[::-1], ri[1:]) and np.array_equal(pv[::-1], rv[1:]), j
    if fp16:
        assert np.array_equal(bits(T["diag"]), bits(g["h_gpu_diag"]))
        assert np.array_equal(bits(T["inv_scale"]), bits(g["h_gpu_inv_scale"]))
    else:
        assert np.array_equal(bits(T["diag"]), bits(g["f_gpu_v32"])[gp[:-1]])
    # Forward rows = the transpose of the columns, columns ascending.
    C = sp.csc_matrix((np.ones(len(bi)), bi, bp), shape=(m, m))
    R = sp.csr_matrix(C)
    R.sort_indices()
    assert np.array_equal(T["f_ptr"].astype(np.int64), R.indptr)
    assert np.array_equal(T["f_idx"].astype(np.int64), R.indices)
    assert T["nnz_stored"] == int(g[f"{tag}_stored_nnz"])


# ---- the PCG recurrence ---------------------------------------------------

@pytest.mark.parametrize("name", ["path5_sddm", "sddm3d_6", "grid2d_12", "rmat_9"])
def test_oracle_pcg_reaches_the_direct_solution(name):
    """Closed form: the oracle's PCG (tol 1e-12) against a direct solve; for a
    singular Laplacian, the minimum-norm solution (the reference grounds to
    it, tests/test_factorize.cpp:630-653)."""
    A, g = golden(name)
    n = A.shape[0]
    O = oracle.Oracle(A, golden_factor(g), bool(g["sddm"]))
    b = np.random.default_rng(4).standard_normal(n)
    if not g["sddm"]:
        b = mg.project_range(A, b)
        x_ref = np.linalg.lstsq(A.toarray(), b, rcond=None)[0]
    else:
        x_ref = spla.spsolve(sp.csc_matrix(A), b)
    x, it, res = O.solve(b, tol=1e-12, max_iter=500)
    assert res < 1e-12
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
    if g["sddm"]:
        assert np.linalg.norm(x - x_ref) <= 1e-8 * np.linalg.norm(x_ref), (it, res)
    else:
        # Centring removes the global mean: on a connected graph that is the
        # minimum-norm solution; on a disconnected one x - x_ref is constant on
        # each component.
        ncomp, lab = sp.csgraph.connected_components(A, directed=False)
        d = x - x_ref
        for c in range(ncomp):
            dc = d[lab == c]
            assert np.ptp(dc) <= 1e-8 * np.linalg.norm(x_ref), (c, np.ptp(dc))
        if ncomp == 1:
            assert abs(x.mean()) <= 1e-12 * np.linalg.norm(x)


def test_oracle_pcg_repeatable_bits():
    A, g = golden("grid2d_12")
    O = oracle.Oracle(A, golden_factor(g), False)
    b = mg.project_range(A, np.random.default_rng(5).standard_normal(A.shape[0]))
    x1, it1, _ = O.solve(b)
    O.reset_apply_count()
    x2, it2, _ = O.solve(b)
    assert it1 == it2 and np.array_equal(bits(x1), bits(x2))


# ---- fp16 known answers (tests/test_lowprec.cpp:134-190) --------------------

def _f(u):
    return struct.unpack("<f", struct.pack("<I", u))[0]


def test_fp16_known_answers():
    L = ax.load_library()
    r = L.apx_fp16_round
    w = L.apx_fp16_widen
    two_m14, two_m24, two_m25 = math.ldexp(1.0, -14), math.ldexp(1.0, -24), math.ldexp(1.0, -25)
    assert r(two_m14) == 0x0400
    assert r(two_m24) == 0x0001
    assert r(two_m25) == 0x0000
    assert r(-two_m25) == 0x8000
    assert r(float(np.nextafter(np.float32(two_m25), np.float32(1.0)))) == 0x0001
    assert r(float(np.nextafter(np.float32(two_m25), np.float32(0.0)))) == 0x0000
    assert r(1e-9) == 0x0000
    assert r(65504.0) == 0x7BFF and r(65519.0) == 0x7BFF and r(65520.0) == 0x7C00
    assert r(1e6) == 0x7C00 and r(-1e6) == 0xFC00
    assert math.isinf(w(r(float("inf")))) and math.isnan(w(r(float("nan"))))
    assert w(r(1.0 + math.ldexp(1.0, -11))) == 1.0
    assert w(r(1.0 + 3.0 * math.ldexp(1.0, -11))) == 1.0 + math.ldexp(1.0, -9)
    assert r(_f(0x3FFFFFFF)) == 0x4000
    for x in (0.0, -0.0, 1.0, -1.0, 0.5, 1.5, 255.0, 0.75, -3.25, 2048.0):
        assert struct.pack("<f", w(r(x))) == struct.pack("<f", x)


def test_fp16_round_trip_exhaustive():
    """Every non-NaN binary16 pattern widens and narrows back to itself, and
    the narrowing agrees with numpy's IEEE binary16 on random fp32 inputs."""
    L = ax.load_library()
    for h in range(0, 1 << 16):
        if (h & 0x7C00) == 0x7C00 and (h & 0x03FF):
            continue
        assert L.apx_fp16_round(L.apx_fp16_widen(h)) == h, hex(h)
    rng = np.random.default_rng(9)
    xs = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 12, 20000))).astype(np.float32)
    with np.errstate(over="ignore"):   # overflow to infinity is part of what is compared
        ref = xs.astype(np.float16).view(np.uint16)
    got = np.array([L.apx_fp16_round(float(x)) for x in xs], np.uint16)
    assert np.array_equal(got, ref)


# ---- generators ---------------------------------------------------------------

def test_matgen_laplacians_and_projection():
    A = mg.grid_laplacian((8, 8, 8), "lognormal")
    assert abs(A - A.T).max() == 0
    assert np.abs(np.asarray(A.sum(axis=1)).ravel()).max() < 1e-12
    S = mg.add_excess(A)
    assert np.all(np.asarray(S.sum(axis=1)).ravel() > 0)
    R = mg.rmat_laplacian(10)
    assert abs(R - R.T).max() == 0 and np.all(R.diagonal() >= 0)
    b = mg.project_range(R, np.random.default_rng(0).standard_normal(R.shape[0]))
    ncomp, lab = sp.csgraph.connected_components(R, directed=False)
    for c in range(ncomp):
        assert abs(b[lab == c].sum()) < 1e-9


# ---- the C ABI ---------------------------------------------------------------

def test_c_abi_exports_every_declared_symbol():
    """The library loads (no GPU needed) and exports every function that
    include/apxchol_amd.h declares."""
    text = (ROOT / "include" / "apxchol_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(apx_[a-z0-9_]+)\s*\(", text))
    assert len(names) > 20
    L = ax.load_library()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_host_factorization_through_c_abi_without_gpu():
    A, g = golden("sddm3d_6")
    F = ax.factorize(A, seed=0)
    assert F.n == A.shape[0] and F.nnz == len(g["inner"]) and F.nrounds == len(g["round_is"])


# ---- the center-k centring of a Laplacian apply (preconditioner.h:335-396) ----

def _tree_laplacian(n, seed):
    """A random weighted spanning tree: the approximate Cholesky of a tree is
    exact (every elimination has at most one neighbour left to sample)."""
    rng = np.random.default_rng(seed)
    parent = np.array([rng.integers(0, v) for v in range(1, n)])
    child = np.arange(1, n)
    w = np.exp(rng.standard_normal(n - 1))
    Adj = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([parent, child]), np.concatenate([child, parent]))),
                        shape=(n, n)).tocsc()
    return ax.laplacian(Adj)


@pytest.mark.parametrize("src", ["grid2d_12", "rmat_9", "tree_40"])
def test_center_k_schedule_and_projection(src):
    """Which applications centre, and what centring is, fixed without the
    oracle's own formula: with K = 10 the 10th application of a run (after a
    reset) is the centred operator C M C r (C = I - 1 1^T / n, M the
    uncentred apply, taken from an oracle with K so large it never centres),
    and every other one is M r bit for bit; K = 1 centres every application;
    M grounds the last-eliminated vertex (its output entry is 0) and is
    symmetric. A dropped mean, a centring of the input only or the output
    only, or an off-by-one schedule fails one of these."""
    if src.startswith("tree"):
        A = _tree_laplacian(int(src.split("_")[1]), 3)
        F = ax.factorize(A, seed=0)
        fac = F.export()
    else:
        A, g = golden(src)
        fac = golden_factor(g)
    n = A.shape[0]
    perm = np.asarray(fac[3])
    O10 = oracle.Oracle(A, fac, False, center_k=10)
    O1 = oracle.Oracle(A, fac, False, center_k=1)
    Ou = oracle.Oracle(A, fac, False, center_k=1 << 30)
    rng = np.random.default_rng(11)
    r = rng.standard_normal(n)
    C = lambda v: v - v.mean()
    Mr = Ou.apply(r)
    assert Mr[int(np.flatnonzero(perm == n - 1)[0])] == 0.0
    cmc = C(Ou.apply(C(r)))
    for k in range(1, 13):
        z = O10.apply(r)
        if k == 10:
            assert np.linalg.norm(z - cmc) <= 1e-13 * np.linalg.norm(cmc), k
            assert abs(z.mean()) <= 1e-14 * np.linalg.norm(z)
        else:
            assert np.array_equal(bits(z), bits(Mr)), k
    for _ in range(2):
        z = O1.apply(r)
        assert np.linalg.norm(z - cmc) <= 1e-13 * np.linalg.norm(cmc)
    r2 = rng.standard_normal(n)
    assert abs(r @ Ou.apply(r2) - r2 @ Mr) <= 1e-12 * np.linalg.norm(r) * np.linalg.norm(Mr)


@pytest.mark.parametrize("n,seed", [(30, 1), (200, 2)])
def test_centred_apply_of_a_tree_is_the_pseudo_inverse(n, seed):
    """Closed form: on a connected tree the factor is exact (up to its fp32
    storage), so a centring application C M C equals the Moore-Penrose
    pseudo-inverse of the Laplacian (numpy's pinv)."""
    A = _tree_laplacian(n, seed)
    F = ax.factorize(A, seed=0)
    O1 = oracle.Oracle(A, F.export(), False, center_k=1)
    Ap = np.linalg.pinv(A.toarray())
    R = np.random.default_rng(seed).standard_normal((3, n))
    for r in R:
        z = O1.apply(r)
        ref = Ap @ r
        assert np.linalg.norm(z - ref) <= 1e-5 * np.linalg.norm(ref)


# ---- the C ABI's option block ---------------------------------------------------

def test_options_default_fills_exactly_the_python_struct():
    """apx_options_default writes the documented defaults and nothing past
    the ctypes mirror of apx_options (a field added on one side only would
    write past it or leave a field unset)."""
    import ctypes
    L = ax.load_library()
    n = ctypes.sizeof(ax.Options)
    buf = (ctypes.c_ubyte * (n + 64))(*([0xAB] * (n + 64)))
    L.apx_options_default(ctypes.cast(buf, ctypes.c_void_p))
    assert bytes(buf[n:]) == b"\xab" * 64
    o = ax.Options.from_buffer(buf)
    assert (o.tol, o.max_iter, o.stagnation_window, o.seed) == (1e-8, 200, 50, 42)
    assert (o.precision, o.fp16_storage, o.schedule, o.drop_rel, o.keep_factor) == (0, 0, 2, -1.0, 0)
    assert (o.iters_per_check, o.graph, o.relabel, o.depth_levels) == (2, 1, 1, 1)
    assert (o.sweep_chunk, o.sweep_blocks, o.tail_rows, o.lane_batches, o.tail_batches, o.tail_cut) == (0,) * 6


# ---- the OpenMP CPU path (bench.py's timed baseline) ---------------------------

@pytest.mark.parametrize("name", ["sddm3d_6", "grid2d_12", "rmat_9"])
def test_openmp_cpu_path_matches_the_one_thread_oracle(name):
    """The OpenMP form of the CPU restatement: its sweeps give the one-thread
    bits at any thread count (thin levels on one thread, fat levels split, the
    per-row arithmetic of omp.h:1343-1385 either way); at one thread its whole
    solve is the oracle's bit for bit; at four threads the reductions are
    summed per thread (preconditioner.h:44-127), so the solve repeats its own
    bits and stays within 1e-10 and one iteration of the one-thread solve."""
    A, g = golden(name)
    F = ax.factorize(A, seed=0)
    o = oracle.Oracle(A, F.export(), F.sddm)
    r = mg.request_rhs(A, F.sddm, 7)   # N(0,1), projected per component onto the range of a Laplacian
    y = o.forward(r[:o.m])
    assert np.array_equal(o.forward_omp(r[:o.m], 4), y)
    assert np.array_equal(o.back_omp(y, 4), o.back(y))
    x1, i1, _ = o.solve(r, tol=1e-10)
    xa, ia, _ = o.solve_omp(r, 1, tol=1e-10)
    assert ia == i1 and np.array_equal(xa, x1)
    x4, i4, res4 = o.solve_omp(r, 4, tol=1e-10)
    x4b, _, _ = o.solve_omp(r, 4, tol=1e-10)
    assert np.array_equal(x4, x4b)
    assert abs(i4 - i1) <= 1 and res4 < 1e-10
    assert np.linalg.norm(x4 - x1) <= 1e-8 * np.linalg.norm(x1)


def test_cpu_baseline_threads_follow_the_cores_and_the_omp_cap(monkeypatch):
    from oracle.cpu_baseline import baseline_threads, physical_cores
    phys, logical = physical_cores()
    assert 1 <= phys <= logical
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert baseline_threads()[0] == 1
    monkeypatch.delenv("OMP_NUM_THREADS")
    t, why = baseline_threads()
    assert t == phys and "physical cores" in why
