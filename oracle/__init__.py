"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py). A ctypes wrapper over oracle/liboracle.so, the
one-thread CPU restatement of the reference's solve path (see oracle.cpp for
the reference lines each function follows), and the OpenMP form of the same
path (the *_omp methods: the timed CPU baseline, not the checker). The
product library never loads it."""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
P = ctypes.c_void_p
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        L.or_create.restype = P
        L.or_create.argtypes = [ctypes.c_uint32, P, P, P, P, P, P, P, P, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.or_destroy.argtypes = [P]
        L.or_info.argtypes = [P, P]
        L.or_export.argtypes = [P] * 11
        L.or_level_sizes.argtypes = [P, ctypes.c_int, P]
        L.or_forward.argtypes = [P, P, P]
        L.or_back.argtypes = [P, P, P]
        L.or_apply.argtypes = [P, P, P]
        L.or_reset_apply_count.argtypes = [P]
        L.or_spmv.restype = ctypes.c_double
        L.or_spmv.argtypes = [P, P, P]
        L.or_solve.argtypes = [P, P, P, P, ctypes.c_double, ctypes.c_int, ctypes.c_int, P]
        L.or_forward_omp.argtypes = [P, P, P, ctypes.c_int]
        L.or_back_omp.argtypes = [P, P, P, ctypes.c_int]
        L.or_apply_omp.argtypes = [P, P, P, ctypes.c_int]
        L.or_solve_omp.argtypes = [P, P, P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
        _lib = L
    return _lib


class Oracle:
    """CPU restatement of cpu_solver over the operator A and a given factor
    (outer, inner, vals, perm, round_is, sddm) -- e.g. Factor.export()."""

    def __init__(self, A, factor, sddm: bool, fp16: bool = False, drop_rel: float = 1e-4, use_rounds: bool = True,
                 center_k: int = 10):
        import scipy.sparse as sp
        C = sp.csc_matrix(A)
        C.sum_duplicates()
        C.sort_indices()
        self.n = C.shape[0]
        self._a = (np.ascontiguousarray(C.indptr, np.int64), np.ascontiguousarray(C.indices, np.int32),
                   np.ascontiguousarray(C.data, np.float64))
        outer, inner, vals, perm, rounds = factor
        self._f = (np.ascontiguousarray(outer, np.uint64), np.ascontiguousarray(inner, np.uint32),
                   np.ascontiguousarray(vals, np.float32), np.ascontiguousarray(perm, np.uint32),
                   np.ascontiguousarray(rounds, np.uint64))
        self.sddm = bool(sddm)
        self.h = lib().or_create(self.n, *[a.ctypes.data for a in self._a], *[a.ctypes.data for a in self._f],
                                 len(self._f[4]), int(sddm), int(fp16), float(drop_rel), int(use_rounds),
                                 int(center_k))
        info = np.zeros(6, np.uint64)
        lib().or_info(self.h, info.ctypes.data)
        self.m, self.nnz_factor, self.nnz_stored, self.nfwd, self.nbck, self.op32 = (int(v) for v in info)
        self.fp16 = bool(fp16)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_destroy(self.h)
            self.h = None

    def arrays(self):
        m, nnz = self.m, self.nnz_stored
        a = dict(csr_ptr=np.zeros(m + 1, np.uint64), csr_idx=np.zeros(nnz, np.uint32),
                 csr_v=np.zeros(0 if self.fp16 else nnz, np.float32), csr_h=np.zeros(nnz if self.fp16 else 0, np.uint16),
                 csc_ptr=np.zeros(m + 1, np.uint64), csc_idx=np.zeros(nnz, np.uint32),
                 csc_v=np.zeros(0 if self.fp16 else nnz, np.float32), csc_h=np.zeros(nnz if self.fp16 else 0, np.uint16),
                 diag16=np.zeros(m if self.fp16 else 0, np.float32), inv_scale16=np.zeros(m if self.fp16 else 0, np.float32))
        lib().or_export(self.h, *[a[k].ctypes.data if a[k].size else None for k in
                                  ["csr_ptr", "csr_idx", "csr_v", "csr_h", "csc_ptr", "csc_idx", "csc_v", "csc_h",
                                   "diag16", "inv_scale16"]])
        return a

    def level_sizes(self, fwd: bool):
        out = np.zeros(self.nfwd if fwd else self.nbck, np.int64)
        lib().or_level_sizes(self.h, int(fwd), out.ctypes.data)
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.m)
        lib().or_forward(self.h, x.ctypes.data, y.ctypes.data)
        return y

    def back(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.m)
        lib().or_back(self.h, x.ctypes.data, y.ctypes.data)
        return y

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros(self.n)
        lib().or_apply(self.h, r.ctypes.data, z.ctypes.data)
        return z

    def reset_apply_count(self):
        lib().or_reset_apply_count(self.h)

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n)
        d = lib().or_spmv(self.h, x.ctypes.data, y.ctypes.data)
        return y, d

    def solve(self, b, tol=1e-8, max_iter=200, x0=None, stagnation_window=50):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n)
        out = np.zeros(2)
        x0a = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        lib().or_solve(self.h, b.ctypes.data, x.ctypes.data, None if x0a is None else x0a.ctypes.data, float(tol),
                       int(max_iter), int(stagnation_window), out.ctypes.data)
        return x, int(out[0]), float(out[1])

    # ---- the OpenMP CPU path (timed baseline; the one-thread methods above are the checker) ----

    def forward_omp(self, x, threads: int):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.m)
        lib().or_forward_omp(self.h, x.ctypes.data, y.ctypes.data, int(threads))
        return y

    def back_omp(self, x, threads: int):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.m)
        lib().or_back_omp(self.h, x.ctypes.data, y.ctypes.data, int(threads))
        return y

    def apply_omp(self, r, threads: int):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros(self.n)
        lib().or_apply_omp(self.h, r.ctypes.data, z.ctypes.data, int(threads))
        return z

    def solve_omp(self, b, threads: int, tol=1e-8, max_iter=200, stagnation_window=50):
        """cpu_solver::solve on `threads` OpenMP threads (x0 = 0): x, iterations, residual."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n)
        out = np.zeros(2)
        lib().or_solve_omp(self.h, b.ctypes.data, x.ctypes.data, float(tol), int(max_iter), int(stagnation_window),
                           int(threads), out.ctypes.data)
        return x, int(out[0]), float(out[1])
