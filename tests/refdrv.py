"""ctypes wrapper of oracle/_ref/libapxref.so -- the reference's own Eigen-free
headers compiled in place (see oracle/ref_driver.cpp). Test infrastructure
only; absent on machines without the reference tree, where callers skip."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

PATH = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "libapxref.so"
P = ctypes.c_void_p
_lib = None


def available() -> bool:
    return PATH.exists()


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(str(PATH))
        L.ref_factorize.restype = P
        L.ref_factorize.argtypes = [ctypes.c_uint32, P, P, P, ctypes.c_uint]
        L.ref_factor_sizes.argtypes = [P, P]
        L.ref_factor_export.argtypes = [P] * 7
        L.ref_factor_free.argtypes = [P]
        L.ref_sample_clique.restype = ctypes.c_int64
        L.ref_sample_clique.argtypes = [ctypes.c_int64, P, P, ctypes.c_double, ctypes.c_uint64, P, P, P]
        L.ref_fp16_round.restype = ctypes.c_uint16
        L.ref_fp16_round.argtypes = [ctypes.c_float]
        L.ref_fp16_widen.restype = ctypes.c_float
        L.ref_fp16_widen.argtypes = [ctypes.c_uint16]
        L.ref_trsv_setup.restype = P
        L.ref_trsv_setup.argtypes = [ctypes.c_uint32, P, P, P, ctypes.c_uint32, P, ctypes.c_int64]
        L.ref_trsv_free.argtypes = [P]
        L.ref_trsv_info.argtypes = [P, P]
        L.ref_trsv_export.argtypes = [P] * 11
        L.ref_trsv_level_stats.argtypes = [P, ctypes.c_int, P, P]
        L.ref_trsv_forward.argtypes = [P, P, P]
        L.ref_trsv_back.argtypes = [P, P, P]
        L.ref_gpu_prep.restype = ctypes.c_int64
        L.ref_gpu_prep.argtypes = [ctypes.c_uint32, P, P, P, ctypes.c_int64, ctypes.c_int, ctypes.c_double] + [P] * 6
        _lib = L
    return _lib


def csc_arrays(A):
    import scipy.sparse as sp
    C = sp.csc_matrix(A)
    C.sum_duplicates()
    C.sort_indices()
    return (C.shape[0], np.ascontiguousarray(C.indptr, np.int64), np.ascontiguousarray(C.indices, np.int32),
            np.ascontiguousarray(C.data, np.float64))


def factorize(A, seed: int):
    """The reference factorization at one thread: (outer, inner, vals, perm,
    round_is, round_active, sddm)."""
    L = lib()
    n, cp, ri, v = csc_arrays(A)
    h = L.ref_factorize(n, cp.ctypes.data, ri.ctypes.data, v.ctypes.data, seed)
    s = np.zeros(4, np.uint64)
    L.ref_factor_sizes(h, s.ctypes.data)
    n, nnz, nr, sd = (int(x) for x in s)
    outer = np.zeros(n + 1, np.uint32)
    inner = np.zeros(nnz, np.uint32)
    vals = np.zeros(nnz, np.float32)
    perm = np.zeros(n, np.uint32)
    ris = np.zeros(max(nr, 1), np.uint64)
    ra = np.zeros(max(nr, 1), np.uint64)
    L.ref_factor_export(h, outer.ctypes.data, inner.ctypes.data, vals.ctypes.data, perm.ctypes.data, ris.ctypes.data,
                        ra.ctypes.data)
    L.ref_factor_free(h)
    return outer, inner, vals, perm, ris[:nr], ra[:nr], bool(sd)


def sample_clique(verts, weights, deg, seed):
    L = lib()
    d = len(verts)
    vv = np.ascontiguousarray(verts, np.uint32)
    ww = np.ascontiguousarray(weights, np.float64)
    ou = np.zeros(max(d, 1), np.uint32)
    ov = np.zeros(max(d, 1), np.uint32)
    ow = np.zeros(max(d, 1), np.float64)
    k = L.ref_sample_clique(d, vv.ctypes.data, ww.ctypes.data, float(deg), int(seed), ou.ctypes.data,
                            ov.ctypes.data, ow.ctypes.data)
    return ou[:k], ov[:k], ow[:k]


class Trsv:
    """The reference's omp_sptrsv on a factor (storage and drop from the
    APXCHOL_SPTRSV_FP16 / APXCHOL_FACTOR_DROP environment at construction)."""

    def __init__(self, outer, inner, vals, n, m, bounds=None):
        L = lib()
        self.outer = np.ascontiguousarray(outer, np.uint32)
        self.inner = np.ascontiguousarray(inner, np.uint32)
        self.vals = np.ascontiguousarray(vals, np.float32)
        b = None if bounds is None else np.ascontiguousarray(bounds, np.uint32)
        self.h = L.ref_trsv_setup(n, self.outer.ctypes.data, self.inner.ctypes.data, self.vals.ctypes.data, m,
                                  None if b is None else b.ctypes.data, 0 if b is None else b.size)
        info = np.zeros(5, np.uint64)
        L.ref_trsv_info(self.h, info.ctypes.data)
        self.stored_nnz, self.fp16, self.nfwd, self.nbck, self.dropped = (int(x) for x in info)
        self.m = m

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_trsv_free(self.h)
            self.h = None

    def arrays(self):
        m, nnz = self.m, self.stored_nnz
        a = dict(csr_ptr=np.zeros(m + 1, np.uint32), csr_idx=np.zeros(nnz, np.uint32),
                 csr_v32=np.zeros(nnz, np.float32), csr_v16=np.zeros(nnz, np.uint16),
                 csc_ptr=np.zeros(m + 1, np.uint32), csc_idx=np.zeros(nnz, np.uint32),
                 csc_v32=np.zeros(nnz, np.float32), csc_v16=np.zeros(nnz, np.uint16),
                 diag=np.zeros(m, np.float32), scale=np.zeros(m, np.float32))
        lib().ref_trsv_export(self.h, *[a[k].ctypes.data for k in
                                        ["csr_ptr", "csr_idx", "csr_v32", "csr_v16", "csc_ptr", "csc_idx",
                                         "csc_v32", "csc_v16", "diag", "scale"]])
        return a

    def level_sizes(self, fwd: bool):
        nl = self.nfwd if fwd else self.nbck
        sz = np.zeros(nl, np.int64)
        wk = np.zeros(nl, np.int64)
        lib().ref_trsv_level_stats(self.h, 1 if fwd else 0, sz.ctypes.data, wk.ctypes.data)
        return sz, wk

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.m)
        lib().ref_trsv_forward(self.h, x.ctypes.data, y.ctypes.data)
        return y

    def back(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.m)
        lib().ref_trsv_back(self.h, x.ctypes.data, y.ctypes.data)
        return y


def gpu_prep(outer, inner, vals, n, m, fp16, rel):
    """cuda_host.h's L11 + drop (+ fp16 narrowing): (ptr, idx, v32, v16, diag, inv_scale)."""
    L = lib()
    o = np.ascontiguousarray(outer, np.uint32)
    i = np.ascontiguousarray(inner, np.uint32)
    v = np.ascontiguousarray(vals, np.float32)
    nnz = L.ref_gpu_prep(n, o.ctypes.data, i.ctypes.data, v.ctypes.data, m, int(fp16), rel, None, None, None, None,
                         None, None)
    ptr = np.zeros(m + 1, np.int32)
    idx = np.zeros(nnz, np.int32)
    v32 = np.zeros(nnz, np.float32)
    v16 = np.zeros(nnz, np.uint16)
    dg = np.zeros(m, np.float32)
    isc = np.zeros(m, np.float32)
    L.ref_gpu_prep(n, o.ctypes.data, i.ctypes.data, v.ctypes.data, m, int(fp16), rel, ptr.ctypes.data,
                   idx.ctypes.data, v32.ctypes.data, v16.ctypes.data, dg.ctypes.data, isc.ctypes.data)
    return ptr, idx, v32, v16, dg, isc
