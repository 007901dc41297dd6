"""Write the golden fixtures of the solve path from the reference itself.

Run in a container that holds the reference tree (it reads
tests/fixtures/*.mtx there and loads oracle/_ref/libapxref.so, the reference's
own Eigen-free headers compiled in place by oracle/Makefile):

    make -C oracle ref && python tests/golden/make_golden.py

For every case it stores, as one .npz (no pickles):
  * the inputs: the operator A (CSC) and a right-hand side x ~ N(0,1), seed 1;
  * the reference factorization at seed 0, one thread (outer, inner, vals,
    perm, round_is, sddm) -- src/factorization.cpp through the components
    oracle/ref_driver.cpp binds;
  * omp_sptrsv's stored factor (CSR rows, CSC columns, diag, scale) and its
    forward / back sweeps of x, for fp32 and fp16 storage at the default drop
    (include/apxchol/solver/sptrsv/omp.h:505-861, 1094-1325);
  * cuda_host's device preparation (L11 + drop [+ fp16 narrowing]):
    columns, diag, inv_scale (include/apxchol/solver/sptrsv/cuda_host.h:62-232).
The fixtures are data: inputs and the reference's outputs.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import scipy.io
import scipy.sparse as sp

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import refdrv  # noqa: E402
from apxchol_amd import matgen as mg  # noqa: E402

REF_FIXTURES = Path("/root/reference/tests/fixtures")


def cases():
    out = {}
    for name in ("path5_laplacian", "path5_sddm"):
        A = sp.csc_matrix(scipy.io.mmread(str(REF_FIXTURES / f"{name}.mtx")))
        out[name] = A
    out["grid2d_12"] = mg.grid_laplacian((12, 12))
    out["sddm3d_6"] = mg.add_excess(mg.grid_laplacian((6, 6, 6), "lognormal"))
    out["aniso3d_6"] = mg.grid_laplacian((6, 6, 6), "lognormal", 0.5, seed=5, axis_weights=(1, 10, 100))
    out["rmat_9"] = mg.rmat_laplacian(9)
    return out


def write_case(name, A):
    A = sp.csc_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    n = A.shape[0]
    outer, inner, vals, perm, ris, ra, sddm = refdrv.factorize(A, 0)
    m = n if sddm else n - 1
    bounds = np.concatenate([[0], np.minimum(np.cumsum(ris), m)]).astype(np.uint32)
    x = np.random.default_rng(1).standard_normal(m)
    rec = dict(A_indptr=A.indptr.astype(np.int64), A_indices=A.indices.astype(np.int32), A_data=A.data,
               x=x, outer=outer, inner=inner, vals=vals, perm=perm, round_is=ris, sddm=np.array(sddm))
    for fp16 in (0, 1):
        tag = "h" if fp16 else "f"
        os.environ["APXCHOL_SPTRSV_FP16"] = str(fp16)
        os.environ.pop("APXCHOL_FACTOR_DROP", None)
        R = refdrv.Trsv(outer, inner, vals, n, m, bounds)
        a = R.arrays()
        y = R.forward(x)
        z = R.back(y)
        rec.update({f"{tag}_csr_ptr": a["csr_ptr"], f"{tag}_csr_idx": a["csr_idx"],
                    f"{tag}_csc_ptr": a["csc_ptr"], f"{tag}_csc_idx": a["csc_idx"],
                    f"{tag}_diag": a["diag"], f"{tag}_scale": a["scale"],
                    f"{tag}_fwd": y, f"{tag}_bck": z,
                    f"{tag}_stored_nnz": np.array(R.stored_nnz), f"{tag}_dropped": np.array(R.dropped)})
        if fp16:
            rec.update(h_csr_v16=a["csr_v16"], h_csc_v16=a["csc_v16"])
        else:
            rec.update(f_csr_v32=a["csr_v32"], f_csc_v32=a["csc_v32"])
        ptr, idx, v32, v16, dg, isc = refdrv.gpu_prep(outer, inner, vals, n, m, fp16, 1e-4)
        rec.update({f"{tag}_gpu_ptr": ptr, f"{tag}_gpu_idx": idx, f"{tag}_gpu_diag": dg})
        if fp16:
            rec.update(h_gpu_v16=v16, h_gpu_inv_scale=isc)
        else:
            rec.update(f_gpu_v32=v32)
        del R
    os.environ.pop("APXCHOL_SPTRSV_FP16", None)
    np.savez_compressed(HERE / f"{name}.npz", **rec)
    print(f"{name}: n={n} nnz(L)={len(inner)} rounds={len(ris)} sddm={sddm}")


def main():
    if not refdrv.available():
        raise SystemExit("oracle/_ref/libapxref.so missing: run `make -C oracle ref` first")
    for name, A in cases().items():
        write_case(name, A)


if __name__ == "__main__":
    main()
