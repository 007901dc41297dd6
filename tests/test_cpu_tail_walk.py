"""The host's tail walk table (gpu_solver.cpp, append_tail_walk), on the CPU.

A tail run's one workgroup walks its batches as each wave's chain of `next`
links, passing one workgroup barrier per phase. A C++ driver built against
the library's exported host function checks, for every combination of the
two options (round-robin dealing over the run or per level; shared phases
for one-wave chains of one-batch levels or not), what the kernel relies on:

* every batch of the run is on exactly one wave's chain, reached from that
  wave's first batch, and the chains visit batches in ascending order;
* the wave a batch is dealt to: run index mod 16 (round robin) or index
  inside its level mod 16 (per level);
* phases never decrease along the run, and level q joins the phase of level
  q - 1 exactly when sharing is on and both levels hold one batch of the same
  wave; the returned phase count is the last phase + 1;
* a second run appended to the same tables keeps indices local to its run.

A dropped `first` flag, a wrong wave, a broken chain or a phase shared
across waves fails one of these; the kernel itself is tested bit for bit on
the GPU."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

DRIVER = r"""
#include "apxchol_amd/csrc/host/gpu_solver.hpp"
#include <cstdio>
#include <vector>
using namespace apxchol_amd;

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { ++fails; std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
    std::printf(__VA_ARGS__); std::printf("\n"); } } while (0)

// Levels with these batch counts form one run (level_first = prefix sums).
static void run_case(const std::vector<int>& counts, bool merge, bool rotate) {
    std::vector<int> lf(1, 0);
    for (int c : counts) lf.push_back(lf.back() + c);
    const int L = static_cast<int>(counts.size()), nb = lf.back();
    tail_tables tt;
    tt.walk.resize(5);   // entries of an earlier run: the new run appends after them
    const int phases = append_tail_walk(lf, 0, L, tt, merge, rotate);
    CHECK(static_cast<int>(tt.walk.size()) == 5 + nb, "size %zu", tt.walk.size());
    std::vector<int> level_of(nb), phase_of(nb), wave_of(nb), seen(nb, 0);
    for (int q = 0; q < L; ++q)
        for (int b = lf[q]; b < lf[q + 1]; ++b) level_of[b] = q;
    int firsts[kTailWaves];
    for (int w = 0; w < kTailWaves; ++w) firsts[w] = -1;
    for (int i = 0; i < nb; ++i) {
        const tail_walk& x = tt.walk[5 + i];
        phase_of[i] = x.info & (kTailWalkMaxLevels - 1);
        wave_of[i] = (x.info >> kTailWalkWaveShift) & (kTailWaves - 1);
        const int want = rotate ? i % kTailWaves : (i - lf[level_of[i]]) % kTailWaves;
        CHECK(wave_of[i] == want, "batch %d wave %d want %d", i, wave_of[i], want);
        if (x.info & kTailWalkFirst) {
            CHECK(firsts[wave_of[i]] < 0, "two firsts for wave %d", wave_of[i]);
            firsts[wave_of[i]] = i;
        }
        CHECK(x.next == -1 || (x.next > i && x.next < nb), "batch %d next %d", i, x.next);
    }
    for (int w = 0; w < kTailWaves; ++w)
        for (int c = firsts[w]; c >= 0; c = tt.walk[5 + c].next) {
            CHECK(wave_of[c] == w, "batch %d on wave %d's chain", c, w);
            ++seen[c];
        }
    for (int i = 0; i < nb; ++i) CHECK(seen[i] == 1, "batch %d seen %d times", i, seen[i]);
    // Phases: one per level, except the shared ones.
    int last = -1;
    for (int q = 0; q < L; ++q) {
        const int b = lf[q], p = phase_of[b];
        for (int k = lf[q]; k < lf[q + 1]; ++k) CHECK(phase_of[k] == p, "level %d mixed phases", q);
        if (q == 0) {
            CHECK(p == 0, "first phase %d", p);
        } else {
            const bool share = merge && counts[q] == 1 && counts[q - 1] == 1 && wave_of[b] == wave_of[lf[q - 1]];
            CHECK(p == last + (share ? 0 : 1), "level %d phase %d after %d (share %d)", q, p, last, share);
        }
        last = p;
    }
    CHECK(phases == last + 1, "returned %d phases, last %d", phases, last);
}

int main() {
    const std::vector<std::vector<int>> runs = {
        {1}, {3, 1, 1, 1, 20, 1, 1, 40, 2, 1, 1},
        {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
        {64, 17, 16, 15, 1, 33, 1, 1, 48}};
    for (const auto& r : runs)
        for (int merge = 0; merge < 2; ++merge)
            for (int rotate = 0; rotate < 2; ++rotate) run_case(r, merge, rotate);
    // Per-level dealing with sharing: a chain of one-batch levels is wave 0's
    // and shares one phase; round robin gives each its own.
    {
        std::vector<int> lf = {0, 1, 2, 3, 4};
        tail_tables a, b;
        CHECK(append_tail_walk(lf, 0, 4, a, true, false) == 1, "per-level chain phases");
        CHECK(append_tail_walk(lf, 0, 4, b, true, true) == 4, "round-robin chain phases");
    }
    std::printf(fails ? "FAILED %d\n" : "OK\n", fails);
    return fails ? 1 : 0;
}
"""


def test_tail_walk_invariants(tmp_path):
    from apxchol_amd import build as b
    lib = b.LIB
    if not lib.exists():
        pytest.fail(f"{lib} is missing: build it first")
    src = tmp_path / "walk.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "walk"
    clang = str(b.ROCM / "llvm" / "bin" / "clang++")
    r = subprocess.run([clang, *b.HOST_FLAGS, f"-I{ROOT}", str(src), "-o", str(exe), f"-L{lib.parent}",
                        "-lapxchol_amd", f"-Wl,-rpath,{lib.parent}", f"-L{b.ROCM}/lib", "-lamdhip64",
                        f"-Wl,-rpath,{b.ROCM}/lib"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout[-3000:] + out.stderr[-2000:]
