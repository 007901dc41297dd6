"""bench.py's command line on the CPU: the measurement is one GPU's, so any
other --gpus value exits at once with status 2 and says why, before torch or
the GPU is touched."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in (2, 8):
        r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", str(n)], capture_output=True,
                           text=True, timeout=60)
        assert r.returncode == 2, (n, r.returncode, r.stderr)
        assert "one GPU" in r.stderr and f"--gpus {n}" in r.stderr
        assert r.stdout == ""
