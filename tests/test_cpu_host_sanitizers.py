"""The host code of the solve path (factorization with its thread team,
SpTRSV preparation, relabeling, operator and sliced-ELL builds, MatrixMarket
input) built with AddressSanitizer + UndefinedBehaviorSanitizer and with
ThreadSanitizer, and run over assorted graphs and the fixture files: every
build must finish its cases without a finding. CPU only; skipped when the
compiler has no sanitizer runtime."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "apxchol_amd" / "csrc"
SOURCES = [ROOT / "tests" / "sanitize" / "host_harness.cpp", CSRC / "host" / "factorize.cpp",
           CSRC / "host" / "trsv_prep.cpp", CSRC / "host" / "mtx.cpp"]
CXX = next((c for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++") or "") if c and Path(c).exists()), None)


def build(tmp: Path, flags: list[str]) -> Path:
    exe = tmp / "harness"
    cmd = [CXX, "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", *flags, f"-I{CSRC}",
           *map(str, SOURCES), "-o", str(exe), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        if "sanitizer" in r.stderr.lower() or "cannot find" in r.stderr.lower():
            pytest.skip(f"no sanitizer runtime: {r.stderr.strip()[:200]}")
        raise AssertionError(r.stderr)
    return exe


@pytest.mark.skipif(CXX is None, reason="no clang++")
@pytest.mark.parametrize("flags,env", [
    (["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
     {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1"}),
    (["-fsanitize=thread"], {"TSAN_OPTIONS": "halt_on_error=1"}),
], ids=["asan_ubsan", "tsan"])
def test_host_code_is_clean_under_sanitizers(tmp_path, flags, env):
    exe = build(tmp_path, flags)
    run_env = dict(os.environ, **env)
    # Every round of the factorization on the thread team, however small.
    run_env["APXCHOL_AMD_FACTOR_PAR_MIN"] = "16"
    r = subprocess.run([str(exe), str(ROOT / "tests" / "golden" / "fixtures"), str(tmp_path)], capture_output=True,
                       text=True, timeout=600, env=run_env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "Sanitizer" not in out and "runtime error" not in out, out[-4000:]
    for case in ("grid3d_20 sddm", "grid3d_20 laplacian", "random 4096", "star+matching", "path", "edge",
                 "single sddm", "path5_laplacian.mtx"):
        assert case in out, out[-2000:]
