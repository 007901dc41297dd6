"""apxchol_amd -- the MI355X (gfx950) solve path of apxchol, from Python.

A thin ctypes layer over the C ABI in include/apxchol_amd.h (the library is
apxchol_amd/libapxchol_amd.so, built in-tree by apxchol_amd.build). It mirrors
the reference's Python surface (python/apxchol/__init__.py:111-338: ``Solver``
with ``solve``/``apply``, ``factorize``, ``solve``) on top of the GPU solver.

Vectors may be NumPy arrays (host; copied through the library's host entry
points) or torch CUDA tensors of float64 (device; passed by pointer, no copy).
There is no CPU fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = ["Solver", "Factor", "factorize", "laplacian", "solve", "SolveResult", "load_library", "ApxError"]

_LIB_PATH = Path(__file__).resolve().parent / "libapxchol_amd.so"
_lib = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int32
U32 = ctypes.c_uint32
F64 = ctypes.c_double


class ApxError(RuntimeError):
    """A failure reported through the C ABI (status + apx_last_error())."""


class ApxInvalid(ApxError, ValueError):
    pass


class Options(ctypes.Structure):
    _fields_ = [("tol", F64), ("max_iter", I32), ("stagnation_window", I32), ("seed", U32),
                ("precision", I32), ("fp16_storage", I32), ("schedule", I32), ("drop_rel", F64),
                ("keep_factor", I32), ("iters_per_check", I32), ("sweep_chunk", I32), ("sweep_blocks", I32),
                ("tail_rows", I32), ("graph", I32), ("relabel", I32),
                ("lane_batches", I32), ("depth_levels", I32), ("tail_batches", I32),
                ("tail_cut", I32)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", I64), ("residual", F64), ("lumped_offdiag", I64), ("solve_vram_mb", F64),
                ("status", I32), ("factorize_ms", F64), ("sptrsv_setup_ms", F64), ("pcg_setup_ms", F64),
                ("pcg_ms", F64)]


class SolverInfo(ctypes.Structure):
    _fields_ = [("n", I64), ("m", I64), ("sddm", I32), ("factor_nnz", ctypes.c_uint64),
                ("stored_nnz", ctypes.c_uint64), ("offdiag", ctypes.c_uint64), ("dropped", ctypes.c_uint64),
                ("rounds", I64), ("levels", I64), ("fwd_batches", I64), ("bck_batches", I64), ("op_nnz", I64),
                ("op_fp32", I32), ("spmv_lanes", I32), ("dataflow_grid", I32), ("fp16_storage", I32),
                ("device_bytes", ctypes.c_uint64), ("level_bytes", ctypes.c_uint64 * 2),
                ("tail_bytes", ctypes.c_uint64 * 2), ("level_launches", I64 * 2), ("tail_runs", I64 * 2),
                ("lane_levels", I64 * 2), ("spmv_format", I32), ("spmv_sell_slots", ctypes.c_uint64),
                ("split_couplings", I64 * 2)]


_SIGS = {
    "apx_last_error": ([], ctypes.c_char_p),
    "apx_options_default": ([P], None),
    "apx_solver_solve_batch": ([P, I32, P, P, F64, I32, P, P], ctypes.c_int),
    "apx_solver_apply_batch": ([P, I32, P, P, P], ctypes.c_int),
    "apx_server_create": ([ctypes.c_uint64, I32, I32, P], ctypes.c_int),
    "apx_group_create": ([P, I32, P], ctypes.c_int),
    "apx_group_destroy": ([P], None),
    "apx_group_info": ([P, P, P], ctypes.c_int),
    "apx_group_solve": ([P, P, P, F64, I32, P, P], ctypes.c_int),
    "apx_solver_autotune": ([P, I32, P, P], ctypes.c_int),
    "apx_server_destroy": ([P], None),
    "apx_server_set_single_lane": ([P, I32], ctypes.c_int),
    "apx_server_set_narrow_block": ([P, I32], ctypes.c_int),
    "apx_server_add": ([P, P, P], ctypes.c_int),
    "apx_server_submit": ([P, I32, P, F64, I32, P], ctypes.c_int),
    "apx_server_wait": ([P, I64, P, P, F64, P], ctypes.c_int),
    "apx_server_get_stats": ([P, P], ctypes.c_int),
    "apx_solver_solve_batch_host": ([P, I32, P, P, F64, I32, P], ctypes.c_int),
    "apx_fp16_round": ([ctypes.c_float], ctypes.c_uint16),
    "apx_fp16_widen": ([ctypes.c_uint16], ctypes.c_float),
    "apx_factorize": ([I64, P, P, P, U32, P], ctypes.c_int),
    "apx_factor_info": ([P, P], ctypes.c_int),
    "apx_factor_export": ([P, P, P, P, P, P], ctypes.c_int),
    "apx_factor_destroy": ([P], None),
    "apx_trsv_sizes": ([P, I32, F64, P], ctypes.c_int),
    "apx_trsv_export": ([P, I32, F64] + [P] * 11, ctypes.c_int),
    "apx_solver_create": ([I64, P, P, P, P, P], ctypes.c_int),
    "apx_solver_create_with_factor": ([I64, P, P, P, P, P, P], ctypes.c_int),
    "apx_solver_destroy": ([P], None),
    "apx_solver_info_get": ([P, P], ctypes.c_int),
    "apx_solver_solve": ([P, P, P, P, F64, I32, P, P], ctypes.c_int),
    "apx_solver_solve_host": ([P, P, P, P, F64, I32, P], ctypes.c_int),
    "apx_solver_apply": ([P, P, P, P], ctypes.c_int),
    "apx_solver_apply_host": ([P, P, P], ctypes.c_int),
    "apx_solver_reset_apply_count": ([P], ctypes.c_int),
    "apx_solver_sweep": ([P, I32, P, P, P], ctypes.c_int),
    "apx_solver_spmv": ([P, P, P, P], ctypes.c_int),
    "apx_solver_solve_LLt": ([P, P, P, P], ctypes.c_int),
    "apx_pcg_blocks": ([I64], I32),
    "apx_pcg_spmv_lanes_for": ([F64], I32),
    "apx_pcg_spmv_pAp": ([I32, P, P, P, I32, P, P, P, I32, P], ctypes.c_int),
    "apx_pcg_spmv_pAp_noalloc": ([I32, P, P, P, I32, P, P, P, I32, P], ctypes.c_int),
    "apx_pcg_update_xr": ([I64, P, P, P, P, F64, P, P], ctypes.c_int),
    "apx_pcg_dot": ([I64, P, P, P, P], ctypes.c_int),
    "apx_pcg_update_p": ([I64, P, P, F64, P], ctypes.c_int),
    "apx_pcg_reduce_partials": ([P, I32, P, P], ctypes.c_int),
    "apx_solver_spmv_batch": ([P, I32, P, P, P], ctypes.c_int),
    "apx_solver_export_factor": ([P, P, P, P, P], ctypes.c_int),
    "apx_solver_set_timing": ([P, I32], ctypes.c_int),
    "apx_solver_take_timing": ([P, P, P], ctypes.c_int),
    "apx_solve": ([I64, P, P, P, P, P, P, P], ctypes.c_int),
    "apx_mtx_read": ([ctypes.c_char_p, P, P], ctypes.c_int),
    "apx_matrix_create": ([I64, P, P, P, P], ctypes.c_int),
    "apx_matrix_size": ([P, P, P, P], ctypes.c_int),
    "apx_matrix_export": ([P, P, P, P], ctypes.c_int),
    "apx_matrix_destroy": ([P], None),
    "apx_mtx_write": ([ctypes.c_char_p, P, I32], ctypes.c_int),
    "apx_mtx_read_vector": ([ctypes.c_char_p, P, I64, P], ctypes.c_int),
    "apx_mtx_write_vector": ([ctypes.c_char_p, P, I64], ctypes.c_int),
    "apx_input_scan_get": ([P, P], ctypes.c_int),
    "apx_input_resolve": ([I32, P, I32, P, ctypes.c_char_p, I64], ctypes.c_int),
    "apx_adjacency_to_laplacian": ([P, P], ctypes.c_int),
    "apx_input_describe": ([I32, P, ctypes.c_char_p, ctypes.c_char_p, I64], ctypes.c_int),
    "apx_input_diagnose": ([I32, P, ctypes.c_char_p, I64], ctypes.c_int),
}

EXPORTED = tuple(_SIGS)


def load_library(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load (once) the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else _LIB_PATH
    # torch bundles its own HIP runtime under the same SONAME
    # (libamdhip64.so.7); loading torch first makes this library bind to that
    # copy, so a process never holds two HIP runtimes.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not p.exists():
        raise ApxError(f"{p} is missing: build it with `python -m apxchol_amd.build`")
    lib = ctypes.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        msg = load_library().apx_last_error().decode(errors="replace")
        raise (ApxInvalid if status == 1 else ApxError)(msg)


def _ptr(a) -> int:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


def _host_vec(a, numel: int, what: str, *, writable: bool = False) -> np.ndarray:
    """A host vector handed to the library: float64, C-contiguous, exactly
    `numel` elements. Inputs are converted (a copy when needed); an output
    array must already have that form, since the library writes through its
    pointer."""
    if writable:
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous and a.flags.writeable):
            raise ValueError(f"{what} must be a writable C-contiguous float64 numpy array")
    else:
        a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != numel:
        raise ValueError(f"{what} holds {a.size} values, expected {numel}")
    return a


def _dev_vec(t, numel: int, device, what: str):
    """A device vector handed to the library by pointer: a torch float64
    tensor on the solver's HIP device, contiguous, with exactly `numel`
    elements. Anything else raises ValueError before any launch: the kernels
    index n (or k * n) doubles from the pointer."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what} must be a torch tensor on the GPU or a numpy array")
    if t.dtype != torch.float64:
        raise ValueError(f"{what} must be float64, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous (a strided view would be read as the wrong elements)")
    if t.numel() != numel:
        raise ValueError(f"{what} holds {t.numel()} values, expected {numel}")
    if t.device.type != "cuda" or (device is not None and t.device.index != device):
        raise ValueError(f"{what} must live on cuda:{device}, got {t.device}")
    return t


def _csc(A):
    """Full-symmetric CSC arrays (int64 colptr, int32 rowidx, float64 vals)."""
    import scipy.sparse as sp
    if not sp.issparse(A):
        raise TypeError("A must be a scipy.sparse matrix")
    C = sp.csc_matrix(A)
    C.sum_duplicates()
    C.sort_indices()
    n = C.shape[0]
    if C.shape[0] != C.shape[1]:
        raise ValueError("A must be square")
    return (n, np.ascontiguousarray(C.indptr, dtype=np.int64), np.ascontiguousarray(C.indices, dtype=np.int32),
            np.ascontiguousarray(C.data, dtype=np.float64))


def _options(tol=1e-8, maxiter=200, stagnation_window=50, seed=42, precision="fp64", fp16_storage=False,
             schedule="hybrid", drop_rel=-1.0, keep_factor=False, iters_per_check=2, sweep_chunk=0,
             sweep_blocks=0, tail_rows=0, graph=True, relabel=True, lane_batches=0,
             depth_levels=True, tail_batches=0, tail_cut=0) -> Options:
    o = Options()
    load_library().apx_options_default(ctypes.byref(o))
    o.tol = float(tol)
    o.max_iter = int(maxiter)
    o.stagnation_window = int(stagnation_window)
    o.seed = int(seed)
    o.precision = {"fp64": 0, "fp32": 1}[precision]
    o.fp16_storage = 1 if fp16_storage else 0
    o.schedule = {"dataflow": 0, "levels": 1, "hybrid": 2}[schedule]
    o.drop_rel = float(drop_rel)
    o.keep_factor = 1 if keep_factor else 0
    o.iters_per_check = int(iters_per_check)
    o.sweep_chunk = int(sweep_chunk)
    o.sweep_blocks = int(sweep_blocks)
    o.tail_rows = int(tail_rows)
    o.graph = 1 if graph else 0
    o.relabel = 1 if relabel else 0
    o.lane_batches = int(lane_batches)
    o.depth_levels = 1 if depth_levels else 0
    o.tail_batches = int(tail_batches)
    o.tail_cut = int(tail_cut)
    return o


@dataclass
class SolveResult:
    x: object
    iterations: int
    residual: float
    status: int
    lumped_offdiag: int = 0
    solve_vram_mb: float = -1.0
    # solve_result::timings: setup/factorize, setup/sptrsv_setup,
    # setup/gpu_pcg_setup and pcg/gpu_pcg_loop in milliseconds.
    timings: dict | None = None


def _result(x, info) -> SolveResult:
    return SolveResult(x, int(info.iterations), float(info.residual), int(info.status), int(info.lumped_offdiag),
                       float(info.solve_vram_mb),
                       {"factorize_ms": float(info.factorize_ms), "sptrsv_setup_ms": float(info.sptrsv_setup_ms),
                        "pcg_setup_ms": float(info.pcg_setup_ms), "pcg_ms": float(info.pcg_ms)})


def _stream(stream) -> int:
    import torch
    return stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream


def _current_device():
    """Index of the current HIP device as torch sees it (None without torch or a GPU)."""
    try:
        import torch
        return torch.cuda.current_device() if torch.cuda.is_available() else None
    except ImportError:
        return None


class Factor:
    """Host approximate-Cholesky factorization: apxchol::factorize's one-thread
    factor (src/factorization.cpp:125-175), bit for bit, computed on
    ``threads`` host threads."""

    def __init__(self, A, *, seed: int = 42):
        lib = load_library()
        self._arrays = _csc(A)
        n, cp, ri, v = self._arrays
        h = P()
        _check(lib.apx_factorize(n, cp.ctypes.data, ri.ctypes.data, v.ctypes.data, seed, ctypes.byref(h)))
        self._h = h
        info = np.zeros(7, np.int64)
        _check(lib.apx_factor_info(h, info.ctypes.data))
        self.n, self.nnz, self.nrounds, self.sddm, self.lumped = (int(info[0]), int(info[1]), int(info[2]),
                                                                   bool(info[3]), int(info[4]))
        self.threads = int(info[5])  # host threads the factorization ran on

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.apx_factor_destroy(self._h)
            self._h = None

    def export(self):
        """(outer, inner, vals, perm, round_is) of the factor L (CSC, diagonal first)."""
        outer = np.zeros(self.n + 1, np.uint64)
        inner = np.zeros(self.nnz, np.uint32)
        vals = np.zeros(self.nnz, np.float32)
        perm = np.zeros(self.n, np.uint32)
        rounds = np.zeros(max(self.nrounds, 1), np.uint64)
        _check(load_library().apx_factor_export(self._h, outer.ctypes.data, inner.ctypes.data, vals.ctypes.data,
                                                perm.ctypes.data, rounds.ctypes.data))
        return outer, inner, vals, perm, rounds[:self.nrounds]

    def trsv(self, fp16: bool = False, drop_rel: float = -1.0) -> dict:
        """Host SpTRSV preparation: forward/back off-diagonal lists, diag, levels."""
        lib = load_library()
        sz = np.zeros(6, np.int64)
        _check(lib.apx_trsv_sizes(self._h, int(fp16), drop_rel, sz.ctypes.data))
        m, off, nl = int(sz[0]), int(sz[1]), int(sz[2])
        out = dict(m=m, offdiag=off, levels=nl, nnz_factor=int(sz[3]), nnz_stored=int(sz[4]), fp16=bool(sz[5]))
        arr = dict(f_ptr=np.zeros(m + 1, np.uint32), f_idx=np.zeros(off, np.int32),
                   f_val=np.zeros(0 if fp16 else off, np.float32), f_h=np.zeros(off if fp16 else 0, np.uint16),
                   b_ptr=np.zeros(m + 1, np.uint32), b_idx=np.zeros(off, np.int32),
                   b_val=np.zeros(0 if fp16 else off, np.float32), b_h=np.zeros(off if fp16 else 0, np.uint16),
                   diag=np.zeros(m, np.float32), inv_scale=np.zeros(m if fp16 else 0, np.float32),
                   level_ptr=np.zeros(nl + 1, np.uint32))
        order = ["f_ptr", "f_idx", "f_val", "f_h", "b_ptr", "b_idx", "b_val", "b_h", "diag", "inv_scale",
                 "level_ptr"]
        _check(lib.apx_trsv_export(self._h, int(fp16), drop_rel,
                                   *[arr[k].ctypes.data if arr[k].size else None for k in order]))
        out.update(arr)
        return out


def factorize(A, *, seed: int = 42) -> Factor:
    return Factor(A, seed=seed)


def laplacian(A):
    """The graph Laplacian L = D - |A| of a weighted ADJACENCY matrix, as the
    reference package's ``apxchol.laplacian`` (python/apxchol/__init__.py:50-100):
    self-loops and explicit zeros dropped, duplicates summed, every vertex
    with an explicit diagonal entry (0 for an isolated one). Host-only."""
    from .mtx import adjacency_to_laplacian
    return adjacency_to_laplacian(A)


class Solver:
    """GPU approximate-Cholesky PCG solver: factor built once, solve many b.

    Mirrors the reference's ``apxchol.Solver`` (python/apxchol/__init__.py:111-270);
    the solve runs on the current HIP device. Device vectors are torch float64
    tensors on that device, contiguous, of the stated length; host vectors are
    numpy arrays. Anything else raises ValueError before the library is called.
    """

    def __init__(self, A, *, seed: int = 42, factor: Factor | None = None, **opts):
        lib = load_library()
        self._arrays = _csc(A)
        n, cp, ri, v = self._arrays
        self._n = n
        self._opts = _options(seed=seed, **opts)
        h = P()
        if factor is None:
            _check(lib.apx_solver_create(n, cp.ctypes.data, ri.ctypes.data, v.ctypes.data,
                                         ctypes.byref(self._opts), ctypes.byref(h)))
        else:
            _check(lib.apx_solver_create_with_factor(n, cp.ctypes.data, ri.ctypes.data, v.ctypes.data, factor._h,
                                                     ctypes.byref(self._opts), ctypes.byref(h)))
        self._h = h
        self._device = _current_device()
        self._refresh_info()

    def _refresh_info(self):
        inf = SolverInfo()
        _check(load_library().apx_solver_info_get(self._h, ctypes.byref(inf)))
        self.info = {k: (list(v) if isinstance(v, ctypes.Array) else v)
                     for k, v in ((k, getattr(inf, k)) for k, _ in SolverInfo._fields_)}

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.apx_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def shape(self):
        return (self._n, self._n)

    @property
    def sddm(self) -> bool:
        return bool(self.info["sddm"])

    @property
    def m(self) -> int:
        """Rows of the stored triangular factor L11 (n for SDDM, n - 1 for a Laplacian)."""
        return int(self.info["m"])

    def _dev(self, t, numel, what):
        return _dev_vec(t, numel, self._device, what)

    def solve(self, b, tol: float | None = None, maxiter: int | None = None, x0=None, out=None,
              stream=None) -> SolveResult:
        lib = load_library()
        tol = -1.0 if tol is None else float(tol)
        maxiter = -1 if maxiter is None else int(maxiter)
        info = SolveInfo()
        n = self._n
        if isinstance(b, np.ndarray):
            b = _host_vec(b, n, "b")
            x = np.zeros(n) if out is None else _host_vec(out, n, "out", writable=True)
            x0a = None if x0 is None else _host_vec(x0, n, "x0")
            _check(lib.apx_solver_solve_host(self._h, b.ctypes.data, x.ctypes.data, _ptr(x0a), tol, maxiter,
                                             ctypes.byref(info)))
        else:
            import torch
            self._dev(b, n, "b")
            x = torch.empty_like(b) if out is None else self._dev(out, n, "out")
            if x0 is not None:
                self._dev(x0, n, "x0")
            s = _stream(stream)
            _check(lib.apx_solver_solve(self._h, _ptr(b), _ptr(x), _ptr(x0), tol, maxiter, ctypes.byref(info), s))
        return _result(x, info)

    def solve_batch(self, B, tol: float | None = None, maxiter: int | None = None, out=None,
                    stream=None) -> tuple:
        """k right-hand sides (rows of B, shape (k, n)) in one blocked PCG.
        Returns (X, [SolveResult per column]) -- X of B's kind (numpy array on
        the host, or a torch tensor on the device)."""
        lib = load_library()
        tol = -1.0 if tol is None else float(tol)
        maxiter = -1 if maxiter is None else int(maxiter)
        n = self._n
        if isinstance(B, np.ndarray):
            B = np.ascontiguousarray(B, dtype=np.float64)
            if B.ndim != 2 or B.shape[1] != n:
                raise ValueError(f"B must have shape (k, {n}), got {B.shape}")
            k = B.shape[0]
            X = np.zeros_like(B) if out is None else _host_vec(out, k * n, "out", writable=True)
            infos = (SolveInfo * max(k, 1))()
            _check(lib.apx_solver_solve_batch_host(self._h, k, B.ctypes.data, X.ctypes.data, tol, maxiter, infos))
        else:
            import torch
            if B.dim() != 2 or B.shape[1] != n:
                raise ValueError(f"B must have shape (k, {n}), got {tuple(B.shape)}")
            k = B.shape[0]
            self._dev(B, k * n, "B")
            X = torch.empty_like(B) if out is None else self._dev(out, k * n, "out")
            infos = (SolveInfo * max(k, 1))()
            _check(lib.apx_solver_solve_batch(self._h, k, _ptr(B), _ptr(X), tol, maxiter, infos, _stream(stream)))
        return X, [_result(None, i) for i in list(infos)[:k]]

    def autotune(self, max_block_width: int = 64, stream=None) -> dict:
        """Measure and adopt the fastest SpMV lanes; record the best block
        width for blocked solves (see apx_solver_autotune)."""
        r = AutotuneResult()
        _check(load_library().apx_solver_autotune(self._h, int(max_block_width), ctypes.byref(r), _stream(stream)))
        out = {"spmv_lanes": r.spmv_lanes, "block_width": r.block_width, "spmv_us": list(r.spmv_us),
               "sweep_pair_us": list(r.sweep_pair_us), "spmv_sell": r.spmv_sell, "spmv_sell_us": r.spmv_sell_us}
        self._refresh_info()
        return out

    def apply_batch(self, R, out=None, stream=None):
        """Preconditioner applied to the k rows of the device tensor R (k, n)
        at once (one application of the centring schedule)."""
        import torch
        if R.dim() != 2 or R.shape[1] != self._n:
            raise ValueError(f"R must have shape (k, {self._n}), got {tuple(R.shape)}")
        k = R.shape[0]
        self._dev(R, k * self._n, "R")
        Z = torch.empty_like(R) if out is None else self._dev(out, k * self._n, "out")
        _check(load_library().apx_solver_apply_batch(self._h, k, _ptr(R), _ptr(Z), _stream(stream)))
        return Z

    def apply(self, r, out=None, stream=None):
        lib = load_library()
        n = self._n
        if isinstance(r, np.ndarray):
            r = _host_vec(r, n, "r")
            z = np.zeros(n) if out is None else _host_vec(out, n, "out", writable=True)
            _check(lib.apx_solver_apply_host(self._h, r.ctypes.data, z.ctypes.data))
            return z
        import torch
        self._dev(r, n, "r")
        z = torch.empty_like(r) if out is None else self._dev(out, n, "out")
        _check(lib.apx_solver_apply(self._h, _ptr(r), _ptr(z), _stream(stream)))
        return z

    TIMING_CLASSES = ("fwd_sweep", "bck_sweep", "spmv", "vector", "sweep_levels", "sweep_tail", "sweep_couple")

    def set_timing(self, on: bool = True):
        """HIP-event timing on the launching stream of whole sweeps, SpMV and
        vector passes, and of the single sweep kernels (sweep_levels,
        sweep_tail, sweep_couple). Solves launch eagerly while it is on."""
        _check(load_library().apx_solver_set_timing(self._h, int(on)))

    def take_timing(self):
        """{class: (total_ms, launches)} since the last call (then reset)."""
        k = len(self.TIMING_CLASSES)
        ms = np.zeros(k)
        cnt = np.zeros(k, np.int64)
        _check(load_library().apx_solver_take_timing(self._h, ms.ctypes.data, cnt.ctypes.data))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(self.TIMING_CLASSES)}

    def reset_apply_count(self):
        _check(load_library().apx_solver_reset_apply_count(self._h))

    def spmv(self, x, out=None, stream=None):
        """y = A x with the solver's single-vector SpMV (torch float64 CUDA tensors of length n)."""
        import torch
        self._dev(x, self._n, "x")
        out = torch.empty_like(x) if out is None else self._dev(out, self._n, "out")
        _check(load_library().apx_solver_spmv(self._h, _ptr(x), _ptr(out), _stream(stream)))
        return out

    def spmv_batch(self, X, out=None, stream=None):
        """Y = A X for the k rows of X (torch float64 CUDA tensor k x n) with the blocked SpMM."""
        import torch
        if X.dim() != 2 or X.shape[1] != self._n:
            raise ValueError(f"X must have shape (k, {self._n}), got {tuple(X.shape)}")
        k = X.shape[0]
        self._dev(X, k * self._n, "X")
        out = torch.empty_like(X) if out is None else self._dev(out, k * self._n, "out")
        _check(load_library().apx_solver_spmv_batch(self._h, int(k), _ptr(X), _ptr(out), _stream(stream)))
        return out

    def sweep(self, rhs, out, backward: bool, stream=None):
        """One sweep in factor order with fp64 vectors (torch float64 CUDA tensors of length m)."""
        self._dev(rhs, self.m, "rhs")
        self._dev(out, self.m, "out")
        _check(load_library().apx_solver_sweep(self._h, int(backward), _ptr(rhs), _ptr(out), _stream(stream)))
        return out

    def solve_LLt(self, x, out=None, stream=None):
        """y = L^-T L^-1 x on the stored factor in factor order (length m), the
        sweeps in the solver's vector precision -- cuda_sptrsv::solve_LLt_dev."""
        import torch
        self._dev(x, self.m, "x")
        out = torch.empty_like(x) if out is None else self._dev(out, self.m, "out")
        _check(load_library().apx_solver_solve_LLt(self._h, _ptr(x), _ptr(out), _stream(stream)))
        return out


class AutotuneResult(ctypes.Structure):
    _fields_ = [("spmv_lanes", I32), ("block_width", I32), ("spmv_us", F64 * 7), ("sweep_pair_us", F64 * 4),
                ("spmv_sell", I32), ("spmv_sell_us", F64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", I64), ("completed", I64), ("factors", I64), ("p50_ms", F64), ("p99_ms", F64),
                ("mean_ms", F64), ("solves_per_s", F64), ("memory_used", ctypes.c_uint64),
                ("memory_budget", ctypes.c_uint64), ("single_lane_completed", I64)]


class Server:
    """Serving loop over resident factors (continuous batching): every added
    Solver keeps one blocked PCG of `slots` columns running on its own stream
    (and a narrow one of ``narrow_slots`` columns, filled first); submitted
    right-hand sides join free columns between iteration groups and leave
    when they finish. With ``single_lane`` (the default) each factor also
    runs one request at a time as a single-vector solve beside its block (the
    oldest waiting request takes it when it is free; its result is the bits of
    Solver.solve with the same tol and maxiter). A solver added here is used
    by the server only."""

    def __init__(self, memory_budget: int = 0, slots: int = 32, iters_per_check: int = 2, single_lane: bool = True,
                 narrow_slots: int = 4):
        lib = load_library()
        h = P()
        _check(lib.apx_server_create(int(memory_budget), int(slots), int(iters_per_check), ctypes.byref(h)))
        self._h = h
        self._n = {}
        _check(lib.apx_server_set_single_lane(h, int(bool(single_lane))))
        _check(lib.apx_server_set_narrow_block(h, int(narrow_slots)))

    def add(self, solver: "Solver") -> int:
        fid = I32()
        _check(load_library().apx_server_add(self._h, solver._h, ctypes.byref(fid)))
        self._n[int(fid.value)] = solver._n
        return int(fid.value)

    def submit(self, factor: int, b, tol: float = 1e-8, maxiter: int = 200) -> int:
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (self._n[factor],):
            raise ValueError("b has the wrong length")
        rid = I64()
        _check(load_library().apx_server_submit(self._h, int(factor), b.ctypes.data, float(tol), int(maxiter),
                                                ctypes.byref(rid)))
        self._pending = getattr(self, "_pending", {})
        self._pending[int(rid.value)] = factor
        return int(rid.value)

    def wait(self, rid: int, timeout: float | None = None):
        """(x, SolveResult) of a request; raises TimeoutError."""
        factor = self._pending[rid]
        x = np.zeros(self._n[factor])
        info = SolveInfo()
        ready = I32()
        _check(load_library().apx_server_wait(self._h, int(rid), x.ctypes.data, ctypes.byref(info),
                                              -1.0 if timeout is None else float(timeout), ctypes.byref(ready)))
        if not ready.value:
            raise TimeoutError(f"request {rid} not done")
        del self._pending[rid]
        return x, _result(x, info)

    def stats(self) -> dict:
        st = ServerStats()
        _check(load_library().apx_server_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            load_library().apx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()


class Group:
    """Many small factors in one launch: the PCG solves of several Solvers
    (one right-hand side each) run together, every kernel kind of an
    iteration one launch over all members (apx_group_solve). Each member's
    result is the bits of its own Solver.solve(b, tol, maxiter). Members share
    the vector precision and factor storage; the group keeps them alive."""

    def __init__(self, solvers):
        solvers = list(solvers)
        if not solvers:
            raise ValueError("a group needs at least one solver")
        arr = (P * len(solvers))(*[s._h for s in solvers])
        h = P()
        _check(load_library().apx_group_create(arr, len(solvers), ctypes.byref(h)))
        self._h = h
        self._solvers = solvers

    @property
    def launches_per_iteration(self) -> int:
        m, k = I32(), I32()
        _check(load_library().apx_group_info(self._h, ctypes.byref(m), ctypes.byref(k)))
        return int(k.value)

    def solve(self, bs, tol: float | None = None, maxiter: int | None = None, outs=None, stream=None):
        """bs: one float64 CUDA tensor per member (its n values, caller order);
        returns (xs, results)."""
        import torch
        if len(bs) != len(self._solvers):
            raise ValueError("one right-hand side per member")
        bs = [s._dev(b, s._n, "b") for s, b in zip(self._solvers, bs)]
        xs = [torch.empty_like(b) for b in bs] if outs is None else \
            [s._dev(x, s._n, "out") for s, x in zip(self._solvers, outs)]
        k = len(bs)
        pb = (P * k)(*[b.data_ptr() for b in bs])
        px = (P * k)(*[x.data_ptr() for x in xs])
        infos = (SolveInfo * k)()
        _check(load_library().apx_group_solve(self._h, pb, px, -1.0 if tol is None else float(tol),
                                              -1 if maxiter is None else int(maxiter), infos, _stream(stream)))
        return xs, [_result(x, info) for x, info in zip(xs, infos)]

    def close(self):
        if getattr(self, "_h", None):
            load_library().apx_group_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()


def solve(A, b, tol: float = 1e-8, maxiter: int = 200, **opts) -> SolveResult:
    """One-shot apxchol::solve (src/solve.cpp:681-737) through apx_solve:
    factorize on the host, upload, PCG on the current HIP device; b and the
    returned x are host arrays. The result's timings hold the setup phases and
    the PCG, as the reference's solve_result::timings."""
    lib = load_library()
    n, cp, ri, v = _csc(A)
    b = _host_vec(b, n, "b")
    o = _options(tol=tol, maxiter=maxiter, **opts)
    x = np.zeros(n)
    info = SolveInfo()
    _check(lib.apx_solve(n, cp.ctypes.data, ri.ctypes.data, v.ctypes.data, b.ctypes.data, x.ctypes.data,
                         ctypes.byref(o), ctypes.byref(info)))
    return _result(x, info)
