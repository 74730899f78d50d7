"""Host helpers of the upload path, built with g++ into a temporary
directory (no GPU):
  - the thread pool (csrc/host_pool.h): every index of a range visited
    exactly once for any grain, also for calls nested inside a task and for
    calls from several threads at once (the busy pool then starts threads
    of its own); a task that throws reaches the caller after the other
    tasks have finished, and the pool keeps working afterwards;
  - the closed-form pair count of a same-event tile (csrc/host_tiles.h)
    against the loop it replaces: pairs (i, j), i in [a, a+ni), j in
    [jb, j1), i < j."""
import shutil
import subprocess
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "hadronic_afterburner_toolkit_amd" / "csrc"

POOL = r"""
#include <atomic>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>
#include "host_pool.h"

static bool cover(long long n, long long grain) {
    std::vector<std::atomic<int>> seen((size_t)n);
    for (auto &x : seen) x = 0;
    hbt::parallel_ranges(n, grain, [&](long long b, long long e) {
        for (long long k = b; k < e; k++) seen[(size_t)k]++;
    });
    for (auto &x : seen)
        if (x != 1) return false;
    return true;
}

int main() {
    for (long long n : {0ll, 1ll, 7ll, 64ll, 1000ll, 100003ll})
        for (long long g : {1ll, 3ll, 64ll, 100000ll})
            if (!cover(n, g)) { std::printf("cover %lld %lld\n", n, g); return 1; }
    // nested: each outer range runs an inner parallel_ranges
    std::atomic<long long> inner{0};
    hbt::parallel_ranges(16, 1, [&](long long b, long long e) {
        for (long long k = b; k < e; k++)
            hbt::parallel_ranges(1000, 10, [&](long long x, long long y) { inner += y - x; });
    });
    if (inner != 16000) { std::printf("nested %lld\n", (long long)inner); return 1; }
    // concurrent callers
    std::vector<std::thread> th;
    std::atomic<int> bad{0};
    for (int t = 0; t < 4; t++)
        th.emplace_back([&] {
            for (int r = 0; r < 20; r++)
                if (!cover(50000, 100)) bad++;
        });
    for (auto &x : th) x.join();
    if (bad) { std::printf("concurrent\n"); return 1; }
    // a throwing task: the exception reaches the caller once the rest ran
    std::atomic<int> ran{0};
    bool caught = false;
    try {
        hbt::parallel_ranges(64, 1, [&](long long b, long long e) {
            ran += (int)(e - b);
            if (b == 0) throw std::runtime_error("task");
        });
    } catch (const std::runtime_error &) {
        caught = true;
    }
    if (!caught || ran != 64) { std::printf("throw %d %d\n", (int)caught, (int)ran); return 1; }
    if (!cover(100003, 7)) { std::printf("after throw\n"); return 1; }
    std::printf("ok %d\n", hbt::HostPool::get().threads());
    return 0;
}
"""

TILES = r"""
#include <cstdio>
#include "host_tiles.h"
int main() {
    long long checked = 0;
    for (long long a = 0; a < 40; a++)
        for (long long ni = 1; ni <= 20; ni++)
            for (long long jb = a + 1; jb < 70; jb++)
                for (long long j1 = jb; j1 < 80; j1++) {
                    unsigned long long want = 0;
                    for (long long i = a; i < a + ni; i++) {
                        const long long lo = jb > i + 1 ? jb : i + 1;
                        if (j1 > lo) want += (unsigned long long)(j1 - lo);
                    }
                    if (hbt::tile_pairs(a, ni, jb, j1) != want) {
                        std::printf("a %lld ni %lld jb %lld j1 %lld\n", a, ni, jb, j1);
                        return 1;
                    }
                    checked++;
                }
    // one large tile: 64 i's against 2^20 j's
    const long long a = 1000, ni = 64, jb = 1010, j1 = 1010 + (1ll << 20);
    unsigned long long want = 0;
    for (long long i = a; i < a + ni; i++) want += (unsigned long long)(j1 - (jb > i + 1 ? jb : i + 1));
    if (hbt::tile_pairs(a, ni, jb, j1) != want) { std::printf("large\n"); return 1; }
    std::printf("ok %lld\n", checked);
    return 0;
}
"""


def build_and_run(tmp_path, name, program, flags=()):
    src = tmp_path / f"{name}.cpp"
    src.write_text(program)
    exe = tmp_path / name
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", *flags, str(src), "-o", str(exe)],
                   check=True, timeout=120)
    return subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pool_visits_every_index_once_and_survives_a_throw(tmp_path):
    r = build_and_run(tmp_path, "pool", POOL, ("-pthread",))
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_tile_pair_count_equals_the_loop(tmp_path):
    r = build_and_run(tmp_path, "tiles", TILES)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
