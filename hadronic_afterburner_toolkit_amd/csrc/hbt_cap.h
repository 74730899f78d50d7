// hbt_cap.h -- the reference's order-dependent pair cap on the GPU.
//
// HBT_correlation.cpp:402-406 (same-event), :651-655 (mixed) and the K_phi
// variants :424-428, :673-677: a pair that passes every window is admitted
// while its counter (per K_T bin, or per (K_T, K_phi) bin) is at most
// needed_number_of_pairs, i.e. the first needed+1-c0 such pairs in
// enumeration order are binned (c0 = the counter when the call starts).
// The q_inv branch (:339-346, :596-603) admits the first 50*needed-c0
// pairs inside its window, per K_T bin, independently of the 3D windows.
//
// Enumeration order is the reference's loop order: rows are the particles
// of the same-event list in order (groups in order, i within a group), and
// within a row j runs over i+1.. for same-event pairs, or over the plan's
// partners in order and their particles in order for mixed pairs.
//
// Three passes, all in FP64 with the literal cascade:
//   count  arrivals per (key, row) -- cnt[key * nrows + row]
//   scan   per key: the row where the admitted total is reached and how
//          many arrivals of that row are still admitted
//   pos    per key: the position, in that row's j order, of the last
//          admitted arrival
// and the binning pass admits (row, pos) <= (cut_row, cut_pos) per key.
// A pair whose K_phi angle lies within the guard band of a bin edge is
// recorded by the count pass with its (row, pos); the host names its key
// with glibc's atan2 and adds its arrival before the scan, the pos pass
// finds its key in a sorted table, and the host bins it if admitted.
#pragma once
#include "hbt_device.h"

namespace hbt {

// Per-key admission limit: rows < row are admitted, row == row admits
// positions <= pos, later rows nothing.  row = INT64_MAX admits all.
struct CapCut {
    long long row;
    long long pos;
};

struct CapView {
    unsigned *cnt3;        // [nkey3][nrows] arrivals at the 3D cap
    unsigned *cntq;        // [n_KT][nrows] arrivals at the q_inv cap
    const CapCut *cut3;    // [nkey3]
    const CapCut *cutq;    // [n_KT]
    long long nrows;
};

// A pair whose K_phi angle lies within the guard band of a bin edge: the
// device cannot name its 3D key, the host decides it with glibc's atan2 as
// the reference does.  Sorted by (row, pos) for the position pass.
struct DeferKey {
    long long row, pos;
    int key;  // (K_T, K_phi) key, or -1: the reference skips the pair
    int pad;
};

// Key of the deferred pair at (row, pos), -1 if none (binary search).
__device__ __forceinline__ int defer_key(const DeferKey *dk, int n, long long row,
                                         long long pos) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const DeferKey &d = dk[mid];
        if (d.row == row && d.pos == pos) return d.key;
        if (d.row < row || (d.row == row && d.pos < pos))
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

__device__ __forceinline__ int key3_of(const DevParams &P, const PairBins &b) {
    return b.kt * P.nphi + b.kphi;
}

__device__ __forceinline__ bool admitted(const CapCut &c, long long row, long long pos) {
    return row < c.row || (row == c.row && pos <= c.pos);
}

// Same-event pair (i, j) of row gi in cap mode: COUNT adds arrivals,
// otherwise bins the admitted parts.  Returns the stage mask.
template <bool COUNT>
__device__ __forceinline__ int cap_same_pair(const DevParams &P, const DevHist &H,
                                             const CapView &C, double psi, long long gi,
                                             long long pos, double p1x, double p1y, double p1z,
                                             double p1E, double x1, double y1, double z1,
                                             double t1, double p2x, double p2y, double p2z,
                                             double p2E, double x2, double y2, double z2,
                                             double t2) {
    PairBins b;
    double q_inv;
    int qinv_ok;
    const int st = cascade_f64<false, true>(P, p1x, p1y, p1z, p1E, p2x, p2y, p2z, p2E, psi, b,
                                            q_inv, qinv_ok);
    const bool bin = (st & ST_BIN) != 0;
    if (COUNT) {
        // a pair on a K_phi bin edge: recorded with its place in the order;
        // the host names its key and adds its arrival before the scan
        if (st & ST_DEFER) {
            const double a[8] = {p1x, p1y, p1z, p1E, x1, y1, z1, t1};
            const double c2[8] = {p2x, p2y, p2z, p2E, x2, y2, z2, t2};
            push_defer(H, a, c2, psi, 0, gi, pos);
        }
        if (bin) atomicAdd(&C.cnt3[(long long)key3_of(P, b) * C.nrows + gi], 1u);
        if (qinv_ok) atomicAdd(&C.cntq[(long long)b.kt * C.nrows + gi], 1u);
        return st;
    }
    // (deferred pairs are binned by the host once the cuts are known)
    const bool take3 = bin && admitted(C.cut3[key3_of(P, b)], gi, pos);
    const bool takeq = qinv_ok && admitted(C.cutq[b.kt], gi, pos);
    if (!take3 && !takeq) return st & ~ST_BIN;
    const double cos_qx =
        cos(P.hbarC_inv * (b.q_E * (t1 - t2) - b.q_x * (x1 - x2) - b.q_y * (y1 - y2) -
                           b.q_z * (z1 - z2)));
    if (takeq) {
        const long k = (long)b.kt * P.qnpts + (qinv_ok - 1);
        atomic_add_u64(&H.qinv_count[k], 1ull);
        atomic_add_f64(&H.qinv_sum[k], q_inv);
        atomic_add_f64(&H.qinv_num[k], cos_qx);
    }
    if (!take3) return st & ~ST_BIN;
    SameRec r;
    r.k = bin_index(P, b.kt, b.kphi, b.io, b.is, b.il);
    r.qo = b.q_out;
    r.qs = b.q_side;
    r.ql = b.q_long;
    r.c = cos_qx;
    add_same_direct(H, r);
    return st;
}

template <bool COUNT>
__device__ __forceinline__ int cap_mixed_pair(const DevParams &P, const DevHist &H,
                                              const CapView &C, double psi, long long gi,
                                              long long pos, double p1x, double p1y, double p1z,
                                              double p1E, double p2x, double p2y, double p2z,
                                              double p2E) {
    PairBins b;
    double q_inv;
    int qinv_ok;
    const int st = cascade_f64<true, true>(P, p1x, p1y, p1z, p1E, p2x, p2y, p2z, p2E, psi, b,
                                           q_inv, qinv_ok);
    const bool bin = (st & ST_BIN) != 0;
    if (COUNT) {
        if (st & ST_DEFER) {  // as for same-event pairs
            const double a[8] = {p1x, p1y, p1z, p1E, 0, 0, 0, 0};
            const double c2[8] = {p2x, p2y, p2z, p2E, 0, 0, 0, 0};
            push_defer(H, a, c2, psi, 1, gi, pos);
        }
        if (bin) atomicAdd(&C.cnt3[(long long)key3_of(P, b) * C.nrows + gi], 1u);
        if (qinv_ok) atomicAdd(&C.cntq[(long long)b.kt * C.nrows + gi], 1u);
        return st;
    }
    if (qinv_ok && admitted(C.cutq[b.kt], gi, pos))
        atomic_add_u64(&H.qinv_den[(long)b.kt * P.qnpts + (qinv_ok - 1)], 1ull);
    if (!bin || !admitted(C.cut3[key3_of(P, b)], gi, pos)) return st & ~ST_BIN;
    add_mixed_direct(H, bin_index(P, b.kt, b.kphi, b.io, b.is, b.il));
    return st;
}

// One block per key: finds the row where the arrivals of rows in order
// exceed the remaining capacity cap[key].  Writes cut[key].row (INT64_MAX
// when every arrival fits) and, in cut[key].pos, the number of arrivals of
// that row that are still admitted (the pos pass turns it into a position).
__global__ void k_cap_scan(const unsigned *cnt, long long nrows, const unsigned long long *cap,
                           CapCut *cut) {
    const int key = blockIdx.x;
    const unsigned *c = cnt + (long long)key * nrows;
    __shared__ unsigned long long part[1024 / 64];
    __shared__ unsigned long long base;
    __shared__ long long found;
    if (threadIdx.x == 0) {
        base = 0;
        found = -1;
    }
    __syncthreads();
    const unsigned long long lim = cap[key];
    for (long long r0 = 0; r0 < nrows; r0 += blockDim.x) {
        const long long r = r0 + threadIdx.x;
        const unsigned v = r < nrows ? c[r] : 0u;
        // inclusive scan of v over the block (wave scan, then wave totals)
        unsigned long long x = v;
        const int lane = __lane_id();
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) part[threadIdx.x >> 6] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long acc = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
                const unsigned long long t = part[w];
                part[w] = acc;
                acc += t;
            }
        }
        __syncthreads();
        const unsigned long long incl = base + part[threadIdx.x >> 6] + x;
        const unsigned long long excl = incl - v;
        // the first row whose inclusive total passes the limit
        if (r < nrows && incl > lim && excl <= lim) found = r;
        __syncthreads();
        if (found >= 0) {
            if (threadIdx.x == (unsigned)(found - r0)) {
                cut[key].row = found;
                cut[key].pos = (long long)(lim - excl);  // arrivals admitted in the row
            }
            return;
        }
        if (threadIdx.x == blockDim.x - 1) base = incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cut[key].row = 0x7fffffffffffffffll;
        cut[key].pos = 0;
    }
}

}  // namespace hbt

namespace hbt {

// Index of the segment [off[k], off[k+1]) holding v (off ascending, n+1 entries).
__device__ __forceinline__ int segment_of(const long long *off, int n, long long v) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One thread per key: the position of the last admitted arrival of the cut
// row (cut[key].pos holds the number of admitted arrivals on entry; -1 when
// none).  Same-event rows: j runs over i+1 .. end of the row's group, and
// the position is the particle index j.  gpoff: particle offsets of the
// groups (ngroups+1).
__global__ void k_cap_pos_same(DevParams P, DevSoA A, const long long *gpoff, int ngroups,
                               const double *psi, CapCut *cut, int nkeys, bool qinv_keys,
                               const DeferKey *dk, int ndk) {
    const int key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= nkeys) return;
    CapCut c = cut[key];
    if (c.row == 0x7fffffffffffffffll) return;
    long long want = c.pos;
    if (want <= 0) {
        cut[key].pos = -1;
        return;
    }
    const long long i = c.row;
    const int g = segment_of(gpoff, ngroups + 1, i);
    const long long end = gpoff[g + 1];
    const double p1x = A.px[i], p1y = A.py[i], p1z = A.pz[i], p1E = A.E[i];
    for (long long j = i + 1; j < end; j++) {
        PairBins b;
        double q_inv;
        int qinv_ok;
        const int st = cascade_f64<false, true>(P, p1x, p1y, p1z, p1E, A.px[j], A.py[j],
                                                A.pz[j], A.E[j], psi[g], b, q_inv, qinv_ok);
        const bool hit = qinv_keys ? (qinv_ok && b.kt == key)
                                   : (((st & ST_BIN) && key3_of(P, b) == key) ||
                                      ((st & ST_DEFER) && defer_key(dk, ndk, i, j) == key));
        if (hit && --want == 0) {
            cut[key].pos = j;
            return;
        }
    }
    cut[key].pos = end;  // not reached when the counts are consistent
}

// Mixed rows: j runs over the plan's partners of the row's event in order,
// then over the partner's particles; position = (partner k) << 32 | j in
// the partner.  s_off: event offsets of the same-event list; b_off of the
// mixed list; rowoff / ids / cs: the plan.
__global__ void k_cap_pos_mixed(DevParams P, DevSoA A, DevSoA B, const long long *s_off, int nev,
                                const long long *b_off, const long long *rowoff, const int *ids,
                                const double *cs, const long long *gpoff, int ngroups,
                                const double *psi, CapCut *cut, int nkeys, bool qinv_keys,
                                const DeferKey *dk, int ndk) {
    const int key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= nkeys) return;
    CapCut c = cut[key];
    if (c.row == 0x7fffffffffffffffll) return;
    long long want = c.pos;
    if (want <= 0) {
        cut[key].pos = -1;
        return;
    }
    const long long i = c.row;
    const int e = segment_of(s_off, nev + 1, i);
    const int g = segment_of(gpoff, ngroups + 1, i);
    const double p1x = A.px[i], p1y = A.py[i], p1z = A.pz[i], p1E = A.E[i];
    for (long long row = rowoff[e]; row < rowoff[e + 1]; row++) {
        const int pid = ids[row];
        const double co = cs[2 * row], sn = cs[2 * row + 1];
        const long long k = row - rowoff[e];
        for (long long j = b_off[pid]; j < b_off[pid + 1]; j++) {
            const double px = B.px[j], py = B.py[j];
            PairBins b;
            double q_inv;
            int qinv_ok;
            const int st = cascade_f64<true, true>(P, p1x, p1y, p1z, p1E, px * co - py * sn,
                                                   px * sn + py * co, B.pz[j], B.E[j], psi[g], b,
                                                   q_inv, qinv_ok);
            const long long pos = (k << 32) | (j - b_off[pid]);
            const bool hit = qinv_keys ? (qinv_ok && b.kt == key)
                                       : (((st & ST_BIN) && key3_of(P, b) == key) ||
                                          ((st & ST_DEFER) && defer_key(dk, ndk, i, pos) == key));
            if (hit && --want == 0) {
                cut[key].pos = pos;
                return;
            }
        }
    }
    cut[key].pos = 0x7fffffffffffffffll;  // not reached when the counts are consistent
}

}  // namespace hbt
