"""MatrixMarket I/O and the input interpretation of the reference CLI
(src/mtx_input.{h,cpp}, src/main.cpp:14-98), through the C ABI, on the CPU.

Pins: the reference's own fixture files (tests/fixtures/*.mtx, copied under
tests/golden/fixtures) and the expectations of its test_mtx_input.cpp
(scan counts, the automatic reading of each matrix class, the refusals, the
L = D - A identities); file round trips are checked bit for bit, and the
reading of the fixtures against scipy's own MatrixMarket reader."""
from pathlib import Path

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import apxchol_amd as ax
from apxchol_amd import mtx

FIX = Path(__file__).resolve().parent / "golden" / "fixtures"


def grid_adjacency(rows, cols, w=1.0):
    """4-neighbour grid adjacency, both triangles (test_mtx_input.cpp:36-55)."""
    idx = np.arange(rows * cols).reshape(rows, cols)
    a = np.concatenate([idx[:-1, :].ravel(), idx[:, :-1].ravel()])
    b = np.concatenate([idx[1:, :].ravel(), idx[:, 1:].ravel()])
    n = rows * cols
    A = sp.coo_matrix((np.full(2 * a.size, w), (np.concatenate([a, b]), np.concatenate([b, a]))), shape=(n, n))
    return A.tocsc()


def grid_laplacian(rows, cols, w=1.0):
    return ax.laplacian(grid_adjacency(rows, cols, w))


def dense(A):
    return np.asarray(sp.csc_matrix(A).todense())


# ---- files ------------------------------------------------------------------

@pytest.mark.parametrize("name", ["path5_laplacian", "path5_sddm", "path5_nearly_sddm", "path5_pattern",
                                  "ambiguous"])
def test_reader_matches_scipy_on_the_reference_fixtures(name):
    A, pattern = mtx.read_mtx(FIX / f"{name}.mtx")
    ref = scipy.io.mmread(str(FIX / f"{name}.mtx"))
    assert pattern == (name == "path5_pattern")
    assert np.array_equal(dense(A), np.asarray(sp.csc_matrix(ref).todense(), dtype=np.float64))
    # CSC with ascending rows in each column, no duplicates.
    for j in range(A.shape[1]):
        rows = A.indices[A.indptr[j]:A.indptr[j + 1]]
        assert np.all(np.diff(rows) > 0)


def test_file_round_trips_are_bit_exact(tmp_path):
    rng = np.random.default_rng(3)
    A = sp.random(40, 40, density=0.1, random_state=4, format="csc")
    A = (A + A.T).tocsc()
    A.data = rng.standard_normal(A.nnz) * 1e-7 + rng.standard_normal(A.nnz)
    A = ((A + A.T) * 0.5).tocsc()
    A.sort_indices()
    for symmetric in (False, True):
        p = tmp_path / f"m{int(symmetric)}.mtx"
        mtx.write_mtx(p, A, symmetric=symmetric)
        B, pattern = mtx.read_mtx(p)
        assert not pattern
        assert np.array_equal(dense(A).view(np.int64), dense(B).view(np.int64))
    x = rng.standard_normal(57) * np.exp(rng.standard_normal(57) * 20)
    mtx.write_vector(tmp_path / "x.mtx", x)
    assert np.array_equal(mtx.read_vector(tmp_path / "x.mtx").view(np.int64), x.view(np.int64))
    assert np.array_equal(scipy.io.mmread(str(tmp_path / "x.mtx")).ravel(), x)


def test_reader_sums_duplicates_and_rejects_malformed_files(tmp_path):
    p = tmp_path / "dup.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.5\n1 1 2.5\n2 2 1\n")
    A, _ = mtx.read_mtx(p)
    assert dense(A)[0, 0] == 4.0 and A.nnz == 2
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n3 1 1\n")
    with pytest.raises(ax.ApxError, match="out of range"):
        mtx.read_mtx(bad)
    short = tmp_path / "short.mtx"
    short.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n")
    with pytest.raises(ax.ApxError, match="expected 3 entries"):
        mtx.read_mtx(short)
    cplx = tmp_path / "c.mtx"
    cplx.write_text("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n")
    with pytest.raises(ax.ApxError, match="unsupported field"):
        mtx.read_mtx(cplx)


# ---- the CLI's reading of the reference fixtures ---------------------------

@pytest.mark.parametrize("name,kind", [("path5_laplacian", "laplacian"), ("path5_sddm", "laplacian"),
                                       ("path5_nearly_sddm", "laplacian"), ("path5_pattern", "adjacency")])
def test_fixture_readings(name, kind):
    L, got, msg = mtx.load_operator(FIX / f"{name}.mtx")
    assert got == kind
    # The pattern file is the path graph whose Laplacian is path5_laplacian.mtx.
    if name == "path5_pattern":
        assert "ADJACENCY" in msg and "L = D - A" in msg
        ref, _ = mtx.read_mtx(FIX / "path5_laplacian.mtx")
        assert np.array_equal(dense(L), dense(ref))
    else:
        assert "LAPLACIAN/SDDM" in msg


def test_ambiguous_fixture_is_refused():
    with pytest.raises(ax.ApxError) as e:
        mtx.load_operator(FIX / "ambiguous.mtx")
    assert "cannot tell" in str(e.value) and "--input-kind" in str(e.value)


# ---- scan_input (test_mtx_input.cpp:74-117) ---------------------------------

def test_scan_assembled_laplacian():
    s = mtx.scan_input(grid_laplacian(6, 6))
    assert s.n == 36 and s.offdiag_pos == 0 and s.offdiag_neg == s.offdiag_nnz
    assert s.diag_pos == 36 and s.deficient_rows == 0 and s.excess_rows == 0
    assert s.upper_nnz == s.lower_nnz


def test_scan_adjacency():
    s = mtx.scan_input(grid_adjacency(6, 6))
    assert s.n == 36 and s.offdiag_neg == 0 and s.offdiag_pos == s.offdiag_nnz
    assert s.diag_pos == 0 and s.deficient_rows == 0 and s.excess_rows == 36


def test_scan_counts_dominance_violations():
    L = grid_laplacian(4, 4).tolil()
    L[0, 15] = -1.0
    L[15, 0] = -1.0
    s = mtx.scan_input(L.tocsc())
    assert s.deficient_rows == 2 and abs(s.worst_deficit + 1.0) < 1e-12 and s.worst_row >= 0


# ---- resolve_input_kind (test_mtx_input.cpp:119-218) ------------------------

def test_resolve_automatic_readings():
    kind, why = mtx.resolve_input_kind("automatic", mtx.scan_input(grid_laplacian(6, 6)))
    assert kind == "laplacian" and "non-positive" in why
    kind, why = mtx.resolve_input_kind("automatic", mtx.scan_input(grid_adjacency(6, 6)), pattern_field=True)
    assert kind == "adjacency" and "pattern" in why
    A = grid_adjacency(6, 6, 2.5).tolil()
    A[0, 0] = 3.0
    A[7, 7] = 1.0
    kind, why = mtx.resolve_input_kind("automatic", mtx.scan_input(A.tocsc()))
    assert kind == "adjacency" and "positive" in why
    L = grid_laplacian(6, 6).tolil()
    L[0, 1] = 0.25
    L[1, 0] = 0.25
    kind, why = mtx.resolve_input_kind("automatic", mtx.scan_input(L.tocsc()))
    assert kind == "laplacian" and "mixed-sign" in why
    kind, _ = mtx.resolve_input_kind("automatic", mtx.scan_input(sp.identity(3, format="csc")))
    assert kind == "laplacian"


def test_resolve_refusals():
    M = sp.csc_matrix(np.array([[0, 1, 0], [1, 0, -1], [0, -1, 0]], dtype=float))
    with pytest.raises(ax.ApxError, match="cannot tell"):
        mtx.resolve_input_kind("automatic", mtx.scan_input(M))
    one_sided = sp.csc_matrix(np.array([[1, 0, 0], [-1, 2, 0], [0, -1, 1]], dtype=float))
    for kind in ("automatic", "laplacian", "adjacency"):
        with pytest.raises(ax.ApxError, match="not symmetric"):
            mtx.resolve_input_kind(kind, mtx.scan_input(one_sided))


def test_resolve_explicit_kind_wins_and_warns():
    adj, lap = mtx.scan_input(grid_adjacency(6, 6)), mtx.scan_input(grid_laplacian(6, 6))
    assert mtx.resolve_input_kind("laplacian", adj)[0] == "laplacian"
    assert mtx.resolve_input_kind("adjacency", lap)[0] == "adjacency"
    _, why = mtx.resolve_input_kind("laplacian", adj, pattern_field=True)
    assert "WARNING" in why and "--input-kind adjacency" in why


def test_describe_and_diagnose():
    d = mtx.describe_input("adjacency", mtx.scan_input(grid_adjacency(4, 4)), "r")
    assert "ADJACENCY" in d and "L = D - A" in d
    d = mtx.describe_input("laplacian", mtx.scan_input(grid_laplacian(4, 4)), "r")
    assert "LAPLACIAN/SDDM" in d and "graph Laplacian: all row sums are zero" in d
    assert mtx.diagnose_failure("laplacian", mtx.scan_input(grid_laplacian(4, 4))) == ""
    assert "--input-kind adjacency" in mtx.diagnose_failure("laplacian", mtx.scan_input(grid_adjacency(4, 4)))


# ---- adjacency_to_laplacian (test_mtx_input.cpp:220-290) --------------------

def test_laplacian_is_degree_minus_adjacency():
    L = grid_laplacian(4, 5, 2.0)
    D = dense(L)
    assert L.shape == (20, 20)
    off = D - np.diag(np.diag(D))
    assert np.all(np.diag(D) > 0) and set(np.unique(off[off != 0])) == {-2.0}
    assert np.abs(D.sum(axis=0)).max() < 1e-12 and np.array_equal(D, D.T)


def test_laplacian_drops_self_loops_and_takes_absolute_values():
    A = grid_adjacency(1, 5).tolil()
    A[2, 2] = 7.0
    assert dense(ax.laplacian(A.tocsc()))[2, 2] == 2.0
    P = grid_adjacency(1, 4)
    assert np.array_equal(dense(ax.laplacian(-P)), dense(ax.laplacian(P)))


def test_laplacian_is_idempotent_on_a_pure_laplacian():
    L = grid_laplacian(5, 6, 1.5)
    assert np.abs(dense(ax.laplacian(L)) - dense(L)).max() < 1e-12


def test_laplacian_keeps_isolated_vertices():
    A = sp.csc_matrix(np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], dtype=float))
    L = ax.laplacian(A)
    assert L.nnz == 5 and dense(L)[2, 2] == 0.0          # explicit zero diagonal
    F = ax.factorize(L, seed=0)                        # an isolated vertex is no contract violation
    assert F.nnz >= 3


# ---- the operator contract (test_mtx_input.cpp:292-330) ---------------------

def test_raw_adjacency_is_refused_with_the_conversion_named():
    with pytest.raises(ax.ApxInvalid) as e:
        ax.factorize(grid_adjacency(20, 20), seed=0)
    msg = str(e.value)
    assert "positive diagonal" in msg and "ADJACENCY" in msg and "laplacian(A)" in msg


def test_converted_adjacency_factorizes_with_fill():
    edges = 2 * 20 * 19
    F = ax.factorize(ax.laplacian(grid_adjacency(20, 20)), seed=0)
    offdiag = F.nnz - 400
    assert offdiag > edges
