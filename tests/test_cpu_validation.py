"""Input checks of the Python layer, on the CPU: a device vector of the wrong
dtype, layout, length or device, and a host output of the wrong form, raise
ValueError before the library is called (the kernels would index n doubles
from the pointer)."""
import numpy as np
import pytest
import torch

import apxchol_amd as ax
from apxchol_amd import pcg


def fake_solver(n=10, m=None):
    """A Solver shell with no library handle: any call that got past the
    checks would fail on the null handle instead of raising ValueError."""
    s = ax.Solver.__new__(ax.Solver)
    s._n = n
    s._device = 0
    s._h = None
    s.info = {"m": n if m is None else m, "sddm": 1}
    return s


BAD_DEVICE_VECTORS = {
    "float32": lambda n: torch.zeros(n, dtype=torch.float32),
    "strided": lambda n: torch.zeros(2 * n, dtype=torch.float64)[::2],
    "short": lambda n: torch.zeros(n - 1, dtype=torch.float64),
    "long": lambda n: torch.zeros(n + 1, dtype=torch.float64),
    "host tensor": lambda n: torch.zeros(n, dtype=torch.float64),
}


@pytest.mark.parametrize("kind", list(BAD_DEVICE_VECTORS))
@pytest.mark.parametrize("method", ["solve", "apply", "spmv", "solve_LLt"])
def test_device_vector_checks(kind, method):
    s = fake_solver()
    v = BAD_DEVICE_VECTORS[kind](10)
    with pytest.raises(ValueError):
        getattr(s, method)(v)


def test_each_check_names_its_reason():
    with pytest.raises(ValueError, match="float64"):
        ax._dev_vec(torch.zeros(4, dtype=torch.float32), 4, 0, "b")
    with pytest.raises(ValueError, match="contiguous"):
        ax._dev_vec(torch.zeros(8, dtype=torch.float64)[::2], 4, 0, "b")
    with pytest.raises(ValueError, match="expected 4"):
        ax._dev_vec(torch.zeros(3, dtype=torch.float64), 4, 0, "b")
    with pytest.raises(ValueError, match="cuda:0"):
        ax._dev_vec(torch.zeros(4, dtype=torch.float64), 4, 0, "b")


def test_block_shapes_are_checked():
    s = fake_solver()
    with pytest.raises(ValueError):
        s.solve_batch(torch.zeros(3, 9, dtype=torch.float64))
    with pytest.raises(ValueError):
        s.apply_batch(torch.zeros(3, 10, dtype=torch.float32))
    with pytest.raises(ValueError):
        s.spmv_batch(torch.zeros(10, 3, dtype=torch.float64).t())    # strided (k, n) view
    with pytest.raises(ValueError):
        s.solve_batch(np.zeros((3, 11)))


def test_sweep_vectors_have_length_m():
    s = fake_solver(n=10, m=9)     # a Laplacian: L11 has n - 1 rows
    with pytest.raises(ValueError, match="expected 9"):
        s.sweep(torch.zeros(10, dtype=torch.float64), torch.zeros(9, dtype=torch.float64), backward=False)
    with pytest.raises(ValueError, match="expected 9"):
        s.solve_LLt(torch.zeros(10, dtype=torch.float64))


def test_host_arrays_are_checked():
    s = fake_solver()
    with pytest.raises(ValueError, match="expected 10"):
        s.solve(np.zeros(9))
    with pytest.raises(ValueError, match="expected 10"):
        s.solve(np.zeros(10), x0=np.zeros(11))
    with pytest.raises(ValueError, match="writable"):
        s.solve(np.zeros(10), out=np.zeros(10, np.float32))
    with pytest.raises(ValueError, match="expected 10"):
        s.solve(np.zeros(10), out=np.zeros(12))
    with pytest.raises(ValueError, match="expected 10"):
        s.apply(np.zeros(3))
    with pytest.raises(ValueError, match="writable"):
        s.apply(np.zeros(10), out=np.zeros((10, 1))[:, 0][::1].reshape(10)[::-1])


def test_pcg_piece_arguments_are_checked():
    x = torch.zeros(10, dtype=torch.float64)
    part = torch.zeros(pcg.MAX_BLOCKS, dtype=torch.float64)
    with pytest.raises(ValueError):
        pcg.dot(x, x.float(), part)
    with pytest.raises(ValueError):
        pcg.dot(x, x, torch.zeros(16, dtype=torch.float64))     # partials buffer too small
    with pytest.raises(ValueError):
        pcg.update_p(x, x[:9], 0.5)
    with pytest.raises(ValueError):
        pcg.spmv_pAp(torch.zeros(11, dtype=torch.int64), torch.zeros(0, dtype=torch.int32),
                     torch.zeros(0, dtype=torch.float64), x, x, part, 4)


def test_pcg_grid_rules():
    """pcg_blocks and spmv_lanes_for follow pcg_cuda_kernels.h:47-69: at least one
    block, at most 2048, ceil(units / 256) between; lanes the power of two
    nearest the mean row length in log2, at most 32 (grids, ~5 per row, get 4;
    ~15 per row gets 16)."""
    assert [pcg.blocks(u) for u in (0, 1, 256, 257, 2048 * 256, 10**12)] == [1, 1, 1, 2, 2048, 2048]
    assert pcg.spmv_lanes_for(5.0) == 4
    assert pcg.spmv_lanes_for(15.0) == 16
    assert pcg.spmv_lanes_for(1.0) == 1
    assert pcg.spmv_lanes_for(1000.0) == 32
