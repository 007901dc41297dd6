"""Build the MI355X library in-tree: apxchol_amd/libapxchol_amd.so.

Host C++ (factorization, SpTRSV prep, solver orchestration, C ABI) is compiled
with amdclang++ under ISO floating-point contraction rules (-ffp-contract=off:
the factor is a bit-exact function of the input); the HIP kernels with hipcc
for gfx950 only, packed-fp32 instructions disabled and no implicit FMA
contraction (the kernels write every fused multiply-add out, so the same
arithmetic runs whatever the schedule).

    python -m apxchol_amd.build        # or apxchol_amd.build.build()
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
CSRC = ROOT / "csrc"
LIB = ROOT / "libapxchol_amd.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

HOST_SOURCES = ["host/factorize.cpp", "host/trsv_prep.cpp", "host/gpu_solver.cpp", "host/server.cpp", "host/group.cpp", "host/mtx.cpp",
                "capi.cpp"]
KERNEL_SOURCES = ["kernels/kernels.hip"]

HOST_FLAGS = ["-std=c++20", "-O3", "-fPIC", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
              f"-I{ROCM}/include", "-Wall", "-Wno-unused-result"]
HIP_FLAGS = ["-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")


def build(verbose: bool = False) -> Path:
    obj_dir = ROOT / "_build"
    obj_dir.mkdir(exist_ok=True)
    clang = str(ROCM / "llvm" / "bin" / "clang++")
    hipcc = str(ROCM / "bin" / "hipcc")
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        o = obj_dir / (Path(src).stem + ".o")
        jobs.append([clang, *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(o)
    for src in KERNEL_SOURCES:
        o = obj_dir / (Path(src).stem + ".hip.o")
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp),
          f"-L{ROCM}/lib", "-lamdhip64", "-pthread", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
