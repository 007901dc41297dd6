"""MatrixMarket files and the reading of an input matrix, over the C ABI.

The reference's CLI input layer (src/main.cpp:14-98, src/mtx_input.{h,cpp}):
read a ``.mtx`` file, decide whether it holds an assembled Laplacian/SDDM
operator or a graph adjacency matrix (``resolve_input_kind``), and assemble
``L = D - A`` for the latter (``adjacency_to_laplacian``; also exported as
``apxchol_amd.laplacian``, the reference Python package's name for it).
Host-only: nothing here touches the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from . import _check, _csc, load_library

KINDS = {"automatic": 0, "laplacian": 1, "adjacency": 2}
_KIND_NAMES = {v: k for k, v in KINDS.items()}


class InputScan(ctypes.Structure):
    """apx_input_scan (src/mtx_input.h:46-73)."""
    _fields_ = [(k, ctypes.c_int64) for k in ("n", "offdiag_nnz", "offdiag_pos", "offdiag_neg", "upper_nnz",
                                              "lower_nnz", "diag_pos", "deficient_rows")] + \
               [("worst_deficit", ctypes.c_double), ("worst_row", ctypes.c_int64), ("excess_rows", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def _path(p) -> bytes:
    return str(p).encode()


class _Matrix:
    """Owning wrapper of an apx_matrix_t."""

    def __init__(self, h):
        self.h = h

    @classmethod
    def from_scipy(cls, A) -> "_Matrix":
        n, cp, ri, v = _csc(A)
        h = ctypes.c_void_p()
        _check(load_library().apx_matrix_create(n, cp.ctypes.data, ri.ctypes.data, v.ctypes.data, ctypes.byref(h)))
        m = cls(h)
        m._keep = (cp, ri, v)
        return m

    def to_scipy(self) -> sp.csc_matrix:
        lib = load_library()
        rows, cols, nnz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib.apx_matrix_size(self.h, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz)))
        cp = np.zeros(cols.value + 1, np.int64)
        ri = np.zeros(nnz.value, np.int32)
        v = np.zeros(nnz.value, np.float64)
        _check(lib.apx_matrix_export(self.h, cp.ctypes.data, ri.ctypes.data if nnz.value else None,
                                     v.ctypes.data if nnz.value else None))
        return sp.csc_matrix((v, ri, cp), shape=(rows.value, cols.value))

    def __del__(self):
        if getattr(self, "h", None):
            load_library().apx_matrix_destroy(self.h)
            self.h = None


def read_mtx(path) -> tuple[sp.csc_matrix, bool]:
    """A ``matrix coordinate`` file as a CSC matrix (symmetric files mirrored,
    duplicates summed) and whether its field is ``pattern``."""
    h = ctypes.c_void_p()
    pat = ctypes.c_int32()
    _check(load_library().apx_mtx_read(_path(path), ctypes.byref(h), ctypes.byref(pat)))
    return _Matrix(h).to_scipy(), bool(pat.value)


def write_mtx(path, A, symmetric: bool = False) -> None:
    m = _Matrix.from_scipy(A)
    _check(load_library().apx_mtx_write(_path(path), m.h, int(symmetric)))


def read_vector(path) -> np.ndarray:
    lib = load_library()
    n = ctypes.c_int64()
    _check(lib.apx_mtx_read_vector(_path(path), None, 0, ctypes.byref(n)))
    x = np.zeros(n.value)
    _check(lib.apx_mtx_read_vector(_path(path), x.ctypes.data if n.value else None, n.value, ctypes.byref(n)))
    return x


def write_vector(path, x) -> None:
    x = np.ascontiguousarray(x, dtype=np.float64)
    _check(load_library().apx_mtx_write_vector(_path(path), x.ctypes.data if x.size else None, x.size))


def scan_input(A) -> InputScan:
    """scan_input (mtx_input.cpp:23-60): one pass of structural facts."""
    m = _Matrix.from_scipy(A)
    s = InputScan()
    _check(load_library().apx_input_scan_get(m.h, ctypes.byref(s)))
    return s


def resolve_input_kind(requested: str, scan: InputScan, pattern_field: bool = False) -> tuple[str, str]:
    """resolve_input_kind (mtx_input.cpp:62-147): ("laplacian" | "adjacency",
    reason); raises ApxError on a one-sided or ambiguous matrix."""
    kind = ctypes.c_int32()
    buf = ctypes.create_string_buffer(1024)
    _check(load_library().apx_input_resolve(KINDS[requested], ctypes.byref(scan), int(pattern_field),
                                            ctypes.byref(kind), buf, len(buf)))
    return _KIND_NAMES[kind.value], buf.value.decode()


def adjacency_to_laplacian(A) -> sp.csc_matrix:
    """L = D - |A| (mtx_input.cpp:149-175): self-loops dropped, every vertex
    keeps an explicit diagonal entry (0 for an isolated one)."""
    m = _Matrix.from_scipy(A)
    h = ctypes.c_void_p()
    _check(load_library().apx_adjacency_to_laplacian(m.h, ctypes.byref(h)))
    return _Matrix(h).to_scipy()


def describe_input(kind: str, scan: InputScan, reason: str) -> str:
    buf = ctypes.create_string_buffer(2048)
    _check(load_library().apx_input_describe(KINDS[kind], ctypes.byref(scan), reason.encode(), buf, len(buf)))
    return buf.value.decode()


def diagnose_failure(kind: str, scan: InputScan) -> str:
    buf = ctypes.create_string_buffer(2048)
    _check(load_library().apx_input_diagnose(KINDS[kind], ctypes.byref(scan), buf, len(buf)))
    return buf.value.decode()


def load_operator(path, input_kind: str = "automatic") -> tuple[sp.csc_matrix, str, str]:
    """The reference CLI's reading of a matrix file (src/main.cpp:14-45):
    read, scan, resolve the kind, convert an adjacency matrix; returns
    (operator, kind, the one-line description)."""
    A, pattern = read_mtx(path)
    if A.shape[0] != A.shape[1]:
        raise ValueError(f"matrix must be square, got {A.shape[0]}x{A.shape[1]}")
    s = scan_input(A)
    kind, reason = resolve_input_kind(input_kind, s, pattern)
    msg = describe_input(kind, s, reason)
    if kind == "adjacency":
        A = adjacency_to_laplacian(A)
    return A, kind, msg
