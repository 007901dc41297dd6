"""The stand-alone PCG pieces of the C ABI on torch device tensors.

The reference's GPU PCG device boundary (include/apxchol/solver/pcg_cuda_kernels.h:43-93):
``spmv_pAp``, ``update_xr``, ``dot``, ``update_p`` and ``reduce_partials``,
with a caller-owned partials buffer (``partials()``) and a grid that is a pure
function of the size (``blocks``), so every reduction is bit-identical run to
run. Arguments are torch tensors on the current HIP device: float64 vectors,
int32 CSR index arrays, float64 or float32 CSR values; they are checked before
any launch.
"""
from __future__ import annotations

import ctypes

from . import _check, _current_device, _dev_vec, _ptr, _stream, load_library

MAX_BLOCKS = 2048   # APX_PCG_MAX_BLOCKS


def blocks(units: int) -> int:
    """pcg_blocks(units): workgroups (= partials) of a producer over `units` work items."""
    return int(load_library().apx_pcg_blocks(int(units)))


def spmv_lanes_for(avg_nnz_per_row: float) -> int:
    return int(load_library().apx_pcg_spmv_lanes_for(float(avg_nnz_per_row)))


def partials(device=None):
    """A partials buffer of MAX_BLOCKS doubles on the device."""
    import torch
    return torch.zeros(MAX_BLOCKS, dtype=torch.float64, device=device or "cuda")


def _dev(t, numel, what):
    return _dev_vec(t, numel, _current_device(), what)


def _idx(t, numel, what):
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.int32 or t.device.type != "cuda" or not t.is_contiguous():
        raise ValueError(f"{what} must be a contiguous int32 CUDA tensor")
    if t.numel() != numel:
        raise ValueError(f"{what} holds {t.numel()} values, expected {numel}")
    return t


def spmv_pAp(rowptr, colidx, vals, x, y, part, lanes: int, stream=None) -> int:
    """y = A x and the partials of x.y; returns the partials' count."""
    import torch
    n = x.numel()
    _idx(rowptr, n + 1, "rowptr")
    nnz = int(rowptr[-1].item()) if n > 0 else 0
    _idx(colidx, nnz, "colidx")
    if not isinstance(vals, torch.Tensor) or vals.dtype not in (torch.float64, torch.float32) or \
            vals.device.type != "cuda" or not vals.is_contiguous() or vals.numel() != nnz:
        raise ValueError(f"vals must be a contiguous float64 / float32 CUDA tensor of {nnz} values")
    _dev(x, n, "x")
    _dev(y, n, "y")
    _dev(part, MAX_BLOCKS, "part")
    _check(load_library().apx_pcg_spmv_pAp(n, _ptr(rowptr), _ptr(colidx), _ptr(vals),
                                           int(vals.dtype == torch.float32), _ptr(x), _ptr(y), _ptr(part),
                                           int(lanes), _stream(stream)))
    return blocks(n * lanes)


def update_xr(x, p, r, ap, alpha: float, part, stream=None) -> int:
    n = x.numel()
    for t, w in ((x, "x"), (p, "p"), (r, "r"), (ap, "Ap")):
        _dev(t, n, w)
    _dev(part, MAX_BLOCKS, "part")
    _check(load_library().apx_pcg_update_xr(n, _ptr(x), _ptr(p), _ptr(r), _ptr(ap), float(alpha), _ptr(part),
                                            _stream(stream)))
    return blocks(n)


def dot(a, b, part, stream=None) -> int:
    n = a.numel()
    _dev(a, n, "a")
    _dev(b, n, "b")
    _dev(part, MAX_BLOCKS, "part")
    _check(load_library().apx_pcg_dot(n, _ptr(a), _ptr(b), _ptr(part), _stream(stream)))
    return blocks(n)


def update_p(p, z, beta: float, stream=None) -> None:
    n = p.numel()
    _dev(p, n, "p")
    _dev(z, n, "z")
    _check(load_library().apx_pcg_update_p(n, _ptr(p), _ptr(z), float(beta), _stream(stream)))


def reduce_partials(part, nblocks: int, out, stream=None) -> None:
    """out[0] = the fixed-order sum of part[:nblocks]."""
    _dev(part, MAX_BLOCKS, "part")
    import torch
    if not isinstance(out, torch.Tensor) or out.dtype != torch.float64 or out.device.type != "cuda" or out.numel() < 1:
        raise ValueError("out must be a float64 CUDA tensor with at least one element")
    _check(load_library().apx_pcg_reduce_partials(_ptr(part), int(nblocks), _ptr(out), _stream(stream)))


def spmv_pAp_noalloc(rowptr, colidx, vals, x, y, pAp, lanes: int, stream=None) -> None:
    """y = A x and pAp[0] = x.y with no partials buffer: the library's
    per-device scratch array holds them (apx_pcg_spmv_pAp_noalloc). The same
    bits as spmv_pAp followed by reduce_partials; calls that overlap on one
    GPU race on the scratch array."""
    import torch
    n = x.numel()
    _idx(rowptr, n + 1, "rowptr")
    nnz = int(rowptr[-1].item()) if n > 0 else 0
    _idx(colidx, nnz, "colidx")
    if not isinstance(vals, torch.Tensor) or vals.dtype not in (torch.float64, torch.float32) or \
            vals.device.type != "cuda" or not vals.is_contiguous() or vals.numel() != nnz:
        raise ValueError(f"vals must be a contiguous float64 / float32 CUDA tensor of {nnz} values")
    _dev(x, n, "x")
    _dev(y, n, "y")
    if not isinstance(pAp, torch.Tensor) or pAp.dtype != torch.float64 or pAp.device.type != "cuda" or \
            pAp.numel() < 1:
        raise ValueError("pAp must be a float64 CUDA tensor with at least one element")
    _check(load_library().apx_pcg_spmv_pAp_noalloc(n, _ptr(rowptr), _ptr(colidx), _ptr(vals),
                                                   int(vals.dtype == torch.float32), _ptr(x), _ptr(y), _ptr(pAp),
                                                   int(lanes), _stream(stream)))
