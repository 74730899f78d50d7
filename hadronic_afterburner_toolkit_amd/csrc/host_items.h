// host_items.h -- work items of the fast kernel and how the host builds
// them (no device code: the unit tests build it with g++).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

#include "host_pool.h"
#include "host_sort.h"
#include "host_tiles.h"

namespace hbt {

// Partner window of a mixed work item: the j range [b, e) of the mixed
// list (the particles of the partner event whose |pT| can pair with the
// item's i tile) and the plan row that gives the partner's rotation.
struct JWin {
    int b, e;
    int row;
    int pad;
};

// One unit of work for one wave.
struct WorkItem {
    long long i0;   // first i particle (same-event: cell-sorted list, mixed: fast list)
    long long j0;   // same-event: first j particle; mixed: first window (View::mwin)
    int ni;         // number of i particles (<= 64*R)
    int nj;         // same-event: number of j particles; mixed: number of windows
    int kind;       // 0 same-event, 1 mixed
    int grp;        // group (buffer) of the pair, selects psi_ref
    int view;       // analysis the item belongs to (index into FastArgs::views)
};

// Launch shape of the fast kernel's work items.
struct ItemShape {
    int TI;  // i particles per item (64 R)
    int JC;  // same-event j chunk
    int KC;  // mixed partner windows per item
};

// Width of the px cells of the same-event lists: particles two or more
// cells apart differ in px by more than D (the margin covers the rounding
// of the cell index), and a group has at most 2^16 cells.
inline double srt_cell_width(double D, double px_span) {
    if (!std::isfinite(D)) return INFINITY;
    return std::max(D * (1.0 + 1e-6), px_span / 65536.0);
}

// Order of one group's same-event list: by px cell (cells of width
// srt_cell_width(D, px span) from the group's smallest px), then by py;
// stable (input order among equal keys).  px(k), py(k): the momentum of
// input particle k < n.  order[p] receives the input particle at position
// p, cell[p] its cell.  One sort by py, then a stable counting sort by
// cell; or, with par (a lone group: a caller looping over many groups in
// parallel passes false), a stable counting sort by cell, then each cell's
// particles sorted by py on the host pool.  Both give the same order.
struct CellOrderScratch {
    std::vector<std::pair<double, int>> idx;       // (py, input particle)
    std::vector<std::pair<uint64_t, int>> ra, rb;  // radix scratch
    std::vector<int> cell, start;
};

template <typename PX, typename PY>
inline void cell_order(int n, PX px, PY py, double D, CellOrderScratch &s, int *order, int *cell,
                       bool par = false) {
    if (n <= 0) return;
    double lo = INFINITY, hi = -INFINITY;
    for (int k = 0; k < n; k++) {
        lo = std::min(lo, (double)px(k));
        hi = std::max(hi, (double)px(k));
    }
    const double w = srt_cell_width(D, hi - lo);
    s.cell.resize((size_t)n);
    int ncell = 1;
    for (int k = 0; k < n; k++) {
        s.cell[(size_t)k] = std::isfinite(w) ? std::min(65536, (int)(((double)px(k) - lo) / w)) : 0;
        ncell = std::max(ncell, s.cell[(size_t)k] + 1);
    }
    s.start.assign((size_t)ncell + 1, 0);
    for (int c : s.cell) s.start[(size_t)c + 1]++;
    for (int c = 0; c < ncell; c++) s.start[(size_t)c + 1] += s.start[(size_t)c];
    s.idx.resize((size_t)n);
    if (!(par && n > 4096)) {  // one sort by py, then a stable counting sort by cell
        for (int k = 0; k < n; k++) s.idx[(size_t)k] = {(double)py(k), k};
        radix_sort_keys(s.idx, s.ra, s.rb);
        std::vector<int> at(s.start.begin(), s.start.end() - 1);
        for (int k = 0; k < n; k++) {
            const int i = s.idx[(size_t)k].second, c = s.cell[(size_t)i];
            const int q = at[(size_t)c]++;
            order[q] = i;
            cell[q] = c;
        }
        return;
    }
    {
        std::vector<int> at(s.start.begin(), s.start.end() - 1);
        for (int k = 0; k < n; k++) {
            const int c = s.cell[(size_t)k];
            s.idx[(size_t)at[(size_t)c]++] = {(double)py(k), k};
        }
    }
    auto sort_cells = [&](long long c0, long long c1, CellOrderScratch &t) {
        for (long long c = c0; c < c1; c++) {
            const int b = s.start[(size_t)c], e = s.start[(size_t)c + 1];
            if (e - b > 1) {
                t.idx.assign(s.idx.begin() + b, s.idx.begin() + e);
                radix_sort_keys(t.idx, t.ra, t.rb);  // by py
                std::copy(t.idx.begin(), t.idx.end(), s.idx.begin() + b);
            }
            for (int q = b; q < e; q++) {
                order[q] = s.idx[(size_t)q].second;
                cell[q] = (int)c;
            }
        }
    };
    parallel_ranges(ncell, 1, [&](long long c0, long long c1) {
        CellOrderScratch t;
        sort_cells(c0, c1, t);
    });
}

// Work items of consecutive chunks of units (cells, events), built on
// several host threads and concatenated in order (the result is the serial
// one).
struct ItemChunk {
    std::vector<WorkItem> items;
    std::vector<JWin> wins;
    std::vector<uint64_t> pairs;
    uint64_t enum_same = 0;
};

template <typename F>
inline void items_by_group_chunks(int ngroups, std::vector<WorkItem> &items,
                                  std::vector<JWin> *wins, std::vector<uint64_t> *pairs,
                                  uint64_t *enum_same, F build) {
    const int nchunks = std::max(1, std::min(ngroups, 64));
    std::vector<ItemChunk> ch((size_t)nchunks);
    parallel_ranges(nchunks, 1, [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; c++)
            build((int)((long long)ngroups * c / nchunks),
                  (int)((long long)ngroups * (c + 1) / nchunks), ch[(size_t)c]);
    });
    size_t nw = wins ? wins->size() : 0;
    for (const ItemChunk &c : ch) {
        for (WorkItem w : c.items) {
            if (w.kind != 0) w.j0 += (long long)nw;  // window index in the joint table
            items.push_back(w);
        }
        if (wins) {
            wins->insert(wins->end(), c.wins.begin(), c.wins.end());
            nw = wins->size();
        }
        if (pairs) pairs->insert(pairs->end(), c.pairs.begin(), c.pairs.end());
        if (enum_same) *enum_same += c.enum_same;
    }
}

// Same-event items of the groups over the (px cell, py)-sorted fast lists
// (group g is [foff[grp_ev[g]], foff[grp_ev[g+1]]); srt_cell / srt_py: the
// cell and py of each position): i tiles of up to TI consecutive particles
// inside one cell, each with two j windows: the particles after the tile
// start in its own cell whose py is within dpx above the tile's, and the
// particles of the next cell whose py is within dpx of the tile's py range.
// Particles two or more cells apart differ in px by more than dpx, so no
// binnable pair is left out.  Windows are cut into chunks of JC.  pairs
// (optional) receives each item's evaluated pair count.  enum_same receives
// the reference's enumerated pair count sum n(n-1)/2 over the kept lists
// (off).
inline void items_same(const ItemShape &s, const std::vector<long long> &off,
                       const std::vector<long long> &foff, const std::vector<int> &grp_ev,
                       const int *srt_cell, const double *srt_py, double dpx, int view,
                       std::vector<WorkItem> &items, std::vector<uint64_t> *pairs,
                       uint64_t &enum_same) {
    const int ngroups = (int)grp_ev.size() - 1;
    for (int g = 0; g < ngroups; g++) {
        const long long nk = off[grp_ev[g + 1]] - off[grp_ev[g]];
        if (nk >= 2) enum_same += (uint64_t)(nk * (nk - 1) / 2);
    }
    // the work splits by cell: (group, first position of the cell), in order
    std::vector<std::vector<long long>> starts((size_t)std::max(0, ngroups));
    parallel_ranges(ngroups, 1, [&](long long ga, long long gb) {
        for (long long g = ga; g < gb; g++) {
            const long long g0 = foff[grp_ev[g]], n = foff[grp_ev[g + 1]] - g0;
            if (n < 2) continue;
            for (long long q = 0; q < n; q++)
                if (q == 0 || srt_cell[g0 + q] != srt_cell[g0 + q - 1]) starts[(size_t)g].push_back(q);
        }
    });
    std::vector<std::pair<int, long long>> units;
    for (int g = 0; g < ngroups; g++)
        for (long long q : starts[(size_t)g]) units.push_back({g, q});
    items_by_group_chunks((int)units.size(), items, nullptr, pairs, nullptr,
                          [&](int ua, int ub, ItemChunk &out) {
        for (int u = ua; u < ub; u++) {
            const int g = units[(size_t)u].first;
            const long long cb = units[(size_t)u].second;
            const long long g0 = foff[grp_ev[g]], n = foff[grp_ev[g + 1]] - g0;
            const int *cell = srt_cell + g0;
            const double *py = srt_py + g0;
            auto emit = [&](long long a, int ni, long long jb, long long je) {
                for (; jb < je; jb += s.JC) {
                    WorkItem w{};
                    w.i0 = g0 + a;
                    w.j0 = g0 + jb;
                    w.ni = ni;
                    w.nj = (int)std::min<long long>(s.JC, je - jb);
                    w.kind = 0;
                    w.grp = g;
                    w.view = view;
                    out.items.push_back(w);
                    if (pairs) out.pairs.push_back(tile_pairs(a, ni, jb, jb + w.nj));
                }
            };
            const long long ce = std::upper_bound(cell + cb, cell + n, cell[cb]) - cell;
            const bool next = ce < n && cell[ce] == cell[cb] + 1;
            const long long ne = next ? std::upper_bound(cell + ce, cell + n, cell[ce]) - cell : ce;
            for (long long a = cb; a < ce; a += s.TI) {
                const int ni = (int)std::min<long long>(s.TI, ce - a);
                const double ylo = py[a] - dpx, yhi = py[a + ni - 1] + dpx;
                emit(a, ni, a + 1, std::upper_bound(py + a, py + ce, yhi) - py);
                if (next) {
                    const long long lb = std::lower_bound(py + ce, py + ne, ylo) - py;
                    emit(a, ni, lb, std::upper_bound(py + lb, py + ne, yhi) - py);
                }
            }
        }
    });
}

// Mixed items over the fast lists (|pT|-sorted per event): i tile of TI
// particles of event e x up to KC partner windows, one per plan row whose
// partner has particles within dpt of the tile's |pT| range (rows with an
// empty window are skipped).  Windows are appended to wins (absolute
// indices into the mixed fast list, whose event offsets are mfoff); the
// item's j0 / nj give its first window and their number.  plan_ids: the
// partner of each plan row; f_pt / mf_pt: |pT| of the two fast lists.
// pairs (optional): evaluated pairs per item.
inline void items_mixed(const ItemShape &s, const std::vector<long long> &foff,
                        const std::vector<int> &grp_ev, const std::vector<long long> &rowoff,
                        const std::vector<int> &plan_ids, const std::vector<long long> &mfoff,
                        const double *f_pt, const double *mf_pt, double dpt, int view,
                        std::vector<WorkItem> &items, std::vector<JWin> &wins,
                        std::vector<uint64_t> *pairs) {
    const int ngroups = (int)grp_ev.size() - 1;
    const double *pt = mf_pt;
    // the work splits by event
    const int nev = ngroups > 0 ? grp_ev[ngroups] : 0;
    std::vector<int> ev_grp((size_t)nev);
    for (int g = 0; g < ngroups; g++)
        for (int e = grp_ev[g]; e < grp_ev[g + 1]; e++) ev_grp[(size_t)e] = g;
    items_by_group_chunks(nev, items, &wins, pairs, nullptr,
                          [&](int ea, int eb, ItemChunk &out) {
        {
            for (int e = ea; e < eb; e++) {
                const int g = ev_grp[(size_t)e];
                const long long n1 = foff[e + 1] - foff[e];
                for (long long a = 0; a < n1; a += s.TI) {
                    const int ni = (int)std::min<long long>(s.TI, n1 - a);
                    const double lo = f_pt[foff[e] + a] - dpt;
                    const double hi = f_pt[foff[e] + a + ni - 1] + dpt;
                    WorkItem w{};
                    uint64_t cnt = 0;
                    auto flush = [&]() {
                        if (w.nj == 0) return;
                        out.items.push_back(w);
                        if (pairs) out.pairs.push_back(cnt);
                        w.nj = 0;
                        cnt = 0;
                    };
                    for (long long row = rowoff[e]; row < rowoff[e + 1]; row++) {
                        const int p = plan_ids[row];
                        long long jb = mfoff[p], je = mfoff[p + 1];
                        if (std::isfinite(dpt)) {
                            jb = std::lower_bound(pt + jb, pt + je, lo) - pt;
                            je = std::upper_bound(pt + jb, pt + je, hi) - pt;
                        }
                        if (je <= jb) continue;
                        if (w.nj == 0) {
                            w.i0 = foff[e] + a;
                            w.ni = ni;
                            w.j0 = (long long)out.wins.size();  // chunk-local until joined
                            w.kind = 1;
                            w.grp = g;
                            w.view = view;
                        }
                        out.wins.push_back(JWin{(int)jb, (int)je, (int)row, 0});
                        w.nj++;
                        cnt += (uint64_t)ni * (uint64_t)(je - jb);
                        if (w.nj == s.KC) flush();
                    }
                    flush();
                }
            }
        }
    });
}

}  // namespace hbt
