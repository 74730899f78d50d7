"""Work items of the fast kernel (csrc/host_items.h), built with g++ into a
temporary directory (no GPU).  The kernel evaluates, for a same-event item,
the pairs (i, j) with i in [i0, i0 + ni), j in [j0, j0 + nj) and i < j in
the group's cell-sorted order; for a mixed item, every (i, j) of the tile
against each of its partner windows.  Pairs the items leave out are never
binned, so the check is on the enumeration itself:
  - same-event: every pair of a group with |dpx| <= D and |dpy| <= D is
    enumerated exactly once, no pair is enumerated twice, and each item's
    pair count is the number of pairs it enumerates.  D infinite (the q_inv
    branch): every pair of the group exactly once.  Inputs: random groups
    of 0 to 1500 particles, clusters of equal momenta, coordinates on a grid
    of spacing D (pairs exactly D apart), a px span wide enough to reach
    the 2^16-cell cap, and several tile and chunk shapes;
  - cell_order: a stable sort by (cell, py) with cells of width >= D,
    also through its per-cell parallel sort (lone groups of up to 9000);
  - mixed: for each plan row, every (i, j) with ||pT_i| - |pT_j|| <= D
    exactly once (also |pT| on a grid of spacing D/2, so pairs exactly D
    apart), and the window's row is the plan row.
Each check fails on a planted mistake: a py band, px cell or |pT| window
a hair too narrow, or a tile without its next-cell window."""
import shutil
import subprocess
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "hadronic_afterburner_toolkit_amd" / "csrc"

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "host_items.h"
using namespace hbt;

static int fail(const char *what, long long a, long long b) {
    std::printf("FAIL %s %lld %lld\n", what, a, b);
    return 1;
}

// One trial of same-event items over ngroups groups.
static int same_trial(std::mt19937_64 &rng, int ngroups, int nmax, double D, int mode,
                      ItemShape sh, long long &checked, bool par = false) {
    std::uniform_int_distribution<int> nd(0, nmax);
    std::vector<long long> foff(1, 0);
    std::vector<int> grp_ev(1, 0);
    std::vector<double> px, py;
    std::normal_distribution<double> gauss(0.0, 0.5);
    std::uniform_real_distribution<double> uni(-1.0, 1.0);
    for (int g = 0; g < ngroups; g++) {
        const int n = nd(rng);
        for (int k = 0; k < n; k++) {
            double x, y;
            if (mode == 0) {          // smooth spectrum
                x = gauss(rng), y = gauss(rng);
            } else if (mode == 1) {   // a few clusters of equal momenta
                const int c = (int)(rng() % 5);
                x = 0.1 * c, y = -0.05 * c;
            } else if (mode == 2) {   // grid of spacing D: pairs exactly D apart
                x = D * (double)(int)(rng() % 9) - 4 * D, y = D * (double)(int)(rng() % 9) - 4 * D;
            } else {                  // wide px span: the cell count is capped
                x = 1e4 * uni(rng), y = uni(rng);
            }
            px.push_back(x);
            py.push_back(y);
        }
        foff.push_back(foff.back() + n);
        grp_ev.push_back(g + 1);  // one event per group
    }
    // per group: the cell order, then the sorted coordinates
    const long long nf = foff.back();
    std::vector<int> cell((size_t)nf), order((size_t)nf);
    std::vector<double> spx((size_t)nf), spy((size_t)nf);
    CellOrderScratch sc;
    for (int g = 0; g < ngroups; g++) {
        const long long g0 = foff[g];
        const int n = (int)(foff[g + 1] - g0);
        cell_order(n, [&](int k) { return px[g0 + k]; }, [&](int k) { return py[g0 + k]; }, D,
                   sc, order.data() + g0, cell.data() + g0, par);
        std::vector<char> seen((size_t)n, 0);
        for (int p = 0; p < n; p++) {
            const int o = order[g0 + p];
            if (o < 0 || o >= n || seen[(size_t)o]) return fail("order is not a permutation", g, p);
            seen[(size_t)o] = 1;
            spx[g0 + p] = px[g0 + o];
            spy[g0 + p] = py[g0 + o];
            if (p > 0) {
                const int c0 = cell[g0 + p - 1], c1 = cell[g0 + p];
                if (c1 < c0 || (c1 == c0 && spy[g0 + p] < spy[g0 + p - 1]))
                    return fail("not sorted by (cell, py)", g, p);
                if (c1 == c0 && spy[g0 + p] == spy[g0 + p - 1] && order[g0 + p] < order[g0 + p - 1])
                    return fail("not stable", g, p);
            }
        }
        // cells two or more apart are more than D apart in px
        for (int a = 0; a < n; a++)
            for (int b = 0; b < n; b++)
                if (cell[g0 + b] >= cell[g0 + a] + 2 && !(std::fabs(spx[g0 + b] - spx[g0 + a]) > D))
                    return fail("cells two apart within D", a, b);
    }
    std::vector<WorkItem> items;
    std::vector<uint64_t> pairs;
    uint64_t es = 0;
    items_same(sh, foff, foff, grp_ev, cell.data(), spy.data(), D, 0, items, &pairs, es);
    uint64_t want_es = 0;
    for (int g = 0; g < ngroups; g++) {
        const long long n = foff[g + 1] - foff[g];
        if (n >= 2) want_es += (uint64_t)(n * (n - 1) / 2);
    }
    if (es != want_es) return fail("enum_same", (long long)es, (long long)want_es);
    if (pairs.size() != items.size()) return fail("pairs per item", 0, 0);
    // coverage per group
    std::vector<std::vector<unsigned char>> cnt((size_t)ngroups);
    for (int g = 0; g < ngroups; g++) {
        const long long n = foff[g + 1] - foff[g];
        cnt[(size_t)g].assign((size_t)(n * n), 0);
    }
    for (size_t k = 0; k < items.size(); k++) {
        const WorkItem &w = items[k];
        if (w.kind != 0 || w.ni < 1 || w.ni > sh.TI || w.nj < 1 || w.nj > sh.JC)
            return fail("item shape", (long long)k, w.ni);
        const int g = w.grp;
        const long long g0 = foff[g], n = foff[g + 1] - g0;
        if (w.i0 < g0 || w.i0 + w.ni > g0 + n || w.j0 < g0 || w.j0 + w.nj > g0 + n)
            return fail("item outside its group", (long long)k, g);
        for (long long i = w.i0; i < w.i0 + w.ni; i++)
            if (cell[i] != cell[w.i0]) return fail("tile spans two cells", (long long)k, i);
        uint64_t np = 0;
        for (long long i = w.i0; i < w.i0 + w.ni; i++)
            for (long long j = w.j0; j < w.j0 + w.nj; j++)
                if (j > i) {
                    unsigned char &c = cnt[(size_t)g][(size_t)((i - g0) * n + (j - g0))];
                    if (++c > 1) return fail("pair enumerated twice", i, j);
                    np++;
                }
        if (np != pairs[k]) return fail("item pair count", (long long)np, (long long)pairs[k]);
    }
    for (int g = 0; g < ngroups; g++) {
        const long long g0 = foff[g], n = foff[g + 1] - g0;
        for (long long a = 0; a < n; a++)
            for (long long b = a + 1; b < n; b++) {
                const bool near = !std::isfinite(D) || (std::fabs(spx[g0 + a] - spx[g0 + b]) <= D &&
                                                        std::fabs(spy[g0 + a] - spy[g0 + b]) <= D);
                if (near && cnt[(size_t)g][(size_t)(a * n + b)] != 1)
                    return fail("pair within D left out", g0 + a, g0 + b);
                checked += near;
            }
    }
    return 0;
}

// Mixed items: events with |pT|-sorted lists, plan rows pointing at
// partner events; every (i, j) with ||pT_i| - |pT_j|| <= D once per row.
static int mixed_trial(std::mt19937_64 &rng, double D, bool grid, ItemShape sh,
                       long long &checked) {
    const int nev = 6;
    std::vector<long long> foff(1, 0);
    std::vector<double> pt;
    std::exponential_distribution<double> ex(3.0);
    for (int e = 0; e < nev; e++) {
        const int n = (int)(rng() % 300);
        std::vector<double> v((size_t)n);
        for (double &x : v)  // on a grid of spacing D/2: pairs exactly D apart
            x = grid && std::isfinite(D) ? 0.5 * D * (double)(int)(rng() % 12) : ex(rng);
        std::sort(v.begin(), v.end());
        pt.insert(pt.end(), v.begin(), v.end());
        foff.push_back(foff.back() + n);
    }
    std::vector<int> grp_ev = {0, 3, nev};  // two groups of three events
    std::vector<long long> rowoff(1, 0);
    std::vector<int> ids;
    for (int e = 0; e < nev; e++) {
        const int M = (int)(rng() % 4);
        for (int r = 0; r < M; r++) ids.push_back((int)(rng() % nev));
        rowoff.push_back((long long)ids.size());
    }
    std::vector<WorkItem> items;
    std::vector<JWin> wins;
    std::vector<uint64_t> pairs;
    items_mixed(sh, foff, grp_ev, rowoff, ids, foff, pt.data(), pt.data(), D, 0, items, wins,
                &pairs);
    // count[row][i][j]
    std::vector<std::vector<unsigned char>> cnt(ids.size());
    for (size_t r = 0; r < ids.size(); r++) cnt[r].assign((size_t)(foff.back() * foff.back()), 0);
    for (size_t k = 0; k < items.size(); k++) {
        const WorkItem &w = items[k];
        if (w.kind != 1 || w.nj < 1 || w.nj > sh.KC) return fail("mixed item shape", (long long)k, w.nj);
        uint64_t np = 0;
        for (int s = 0; s < w.nj; s++) {
            const JWin &jw = wins[(size_t)(w.j0 + s)];
            for (long long i = w.i0; i < w.i0 + w.ni; i++) {
                int e = 0;
                while (foff[e + 1] <= i) e++;
                if (jw.row < rowoff[e] || jw.row >= rowoff[e + 1]) return fail("row of another event", (long long)k, jw.row);
                const int p = ids[(size_t)jw.row];
                if (jw.b < foff[p] || jw.e > foff[p + 1]) return fail("window outside the partner", (long long)k, s);
                for (long long j = jw.b; j < jw.e; j++) {
                    if (++cnt[(size_t)jw.row][(size_t)(i * foff.back() + j)] > 1)
                        return fail("mixed pair twice", i, j);
                    np++;
                }
            }
        }
        if (np != pairs[k]) return fail("mixed pair count", (long long)np, (long long)pairs[k]);
    }
    for (int e = 0; e < nev; e++)
        for (long long r = rowoff[e]; r < rowoff[e + 1]; r++) {
            const int p = ids[(size_t)r];
            for (long long i = foff[e]; i < foff[e + 1]; i++)
                for (long long j = foff[p]; j < foff[p + 1]; j++)
                    if (std::fabs(pt[i] - pt[j]) <= D) {
                        if (cnt[(size_t)r][(size_t)(i * foff.back() + j)] != 1)
                            return fail("mixed pair within D left out", i, j);
                        checked++;
                    }
        }
    return 0;
}

int main() {
    std::mt19937_64 rng(2024);
    long long checked = 0;
    const ItemShape shapes[] = {{64, 256, 4}, {128, 64, 1}, {256, 4096, 16}};
    const double Ds[] = {0.2175, 0.05, INFINITY};
    for (const ItemShape &sh : shapes)
        for (double D : Ds)
            for (int mode = 0; mode < 4; mode++)
                for (int t = 0; t < 2; t++)
                    if (same_trial(rng, 3, 1500, D, mode, sh, checked)) return 1;
    // lone groups large enough for the per-cell parallel sort
    for (double D : Ds)
        for (int mode = 0; mode < 4; mode++)
            if (same_trial(rng, 1, 9000, D, mode, shapes[0], checked, true)) return 1;
    for (const ItemShape &sh : shapes)
        for (double D : Ds)
            for (int grid = 0; grid < 2; grid++)
                if (mixed_trial(rng, D, grid == 1, sh, checked)) return 1;
    std::printf("ok %lld\n", checked);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_items_enumerate_every_pair_within_the_windows_once(tmp_path):
    src = tmp_path / "items.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "items"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{CSRC}", str(src), "-o", str(exe)],
                   check=True, timeout=180)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout
    assert int(r.stdout.split()[1]) > 10 ** 6
