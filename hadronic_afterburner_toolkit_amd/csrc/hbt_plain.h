// hbt_plain.h -- the plain FP64 kernels' tiles: the literal cascade of
// HBT_correlation.cpp:291-460 (same-event) and :563-689 (mixed) for one
// 256 x 256 tile of pairs, with the j tile staged in LDS.  Shared by the
// context's plain kernels (hbt_engine.hip) and the allocation-free entry
// point (hbt_noalloc.hip).
#pragma once
#include "hbt_device.h"

namespace hbt {

constexpr int TILE = 256;

__device__ __forceinline__ void wave_add_stage(unsigned long long *dst, unsigned v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (__lane_id() == 0 && x) atomic_add_u64(dst, x);
}

// Same-event pairs i<j of one group [g0, g0+n), tile (ti, tj): one thread
// per i of tile ti against the j of tile tj (tj >= ti), stage counts of
// every pair added to H.stage[1..4].  Block-uniform early exit only.
__device__ __forceinline__ void same_plain_tile(const DevParams &P, const DevHist &H,
                                                const DevSoA &A, long long g0, int n,
                                                double psi, int ti, int tj) {
    if (tj < ti || tj * TILE >= n) return;
    __shared__ double s[8][TILE];
    const int j0 = tj * TILE;
    const int jn = min(TILE, n - j0);
    const int tid = threadIdx.x;
    if (tid < jn) {
        const long long j = g0 + j0 + tid;
        s[0][tid] = A.px[j];
        s[1][tid] = A.py[j];
        s[2][tid] = A.pz[j];
        s[3][tid] = A.E[j];
        s[4][tid] = A.x[j];
        s[5][tid] = A.y[j];
        s[6][tid] = A.z[j];
        s[7][tid] = A.t[j];
    }
    __syncthreads();
    const int i = ti * TILE + tid;
    unsigned c_kt = 0, c_out = 0, c_side = 0, c_bin = 0;
    if (i < n) {
        const long long gi = g0 + i;
        const double p1x = A.px[gi], p1y = A.py[gi], p1z = A.pz[gi], p1E = A.E[gi];
        const double x1 = A.x[gi], y1 = A.y[gi], z1 = A.z[gi], t1 = A.t[gi];
        const int jstart = (ti == tj) ? tid + 1 : 0;
        for (int jj = jstart; jj < jn; jj++) {
            const int st = bin_same_f64(P, H, psi, p1x, p1y, p1z, p1E, x1, y1, z1, t1, s[0][jj],
                                        s[1][jj], s[2][jj], s[3][jj], s[4][jj], s[5][jj],
                                        s[6][jj], s[7][jj]);
            c_kt += st & 1;
            c_out += (st >> 1) & 1;
            c_side += (st >> 2) & 1;
            c_bin += (st >> 3) & 1;
        }
    }
    wave_add_stage(&H.stage[1], c_kt);
    wave_add_stage(&H.stage[2], c_out);
    wave_add_stage(&H.stage[3], c_side);
    wave_add_stage(&H.stage[4], c_bin);
}

// Mixed pairs, tile (ti, tj) of plan row row0 + z: event e0 + z/M of the
// list A (offsets a_off) against its partner ids[row] of the list B
// (offsets b_off), the partner's tile rotated by (c, s) while staged
// (HBT_correlation.cpp:522-523); stage counts to H.stage[6..9].
__device__ __forceinline__ void mixed_plain_tile(const DevParams &P, const DevHist &H,
                                                 const DevSoA &A, const DevSoA &B,
                                                 const long long *a_off, const long long *b_off,
                                                 const int *ids, const double *cs, int M, int e0,
                                                 long long row0, double psi, int ti, int tj,
                                                 int z) {
    const long long row = row0 + z;
    const int e = e0 + z / M;
    const long long ib = a_off[e];
    const int n1 = (int)(a_off[e + 1] - ib);
    const int pid = ids[row];
    const long long jb = b_off[pid];
    const int n2 = (int)(b_off[pid + 1] - jb);
    if (ti * TILE >= n1 || tj * TILE >= n2) return;
    const double c = cs[2 * row], sn = cs[2 * row + 1];
    __shared__ double s[4][TILE];
    const int j0 = tj * TILE;
    const int jn = min(TILE, n2 - j0);
    const int tid = threadIdx.x;
    if (tid < jn) {
        const long long j = jb + j0 + tid;
        const double px = B.px[j], py = B.py[j];
        s[0][tid] = px * c - py * sn;
        s[1][tid] = px * sn + py * c;
        s[2][tid] = B.pz[j];
        s[3][tid] = B.E[j];
    }
    __syncthreads();
    const int i = ti * TILE + tid;
    unsigned c_kt = 0, c_out = 0, c_side = 0, c_bin = 0;
    if (i < n1) {
        const long long gi = ib + i;
        const double p1x = A.px[gi], p1y = A.py[gi], p1z = A.pz[gi], p1E = A.E[gi];
        for (int jj = 0; jj < jn; jj++) {
            const int st = bin_mixed_f64(P, H, psi, p1x, p1y, p1z, p1E, s[0][jj], s[1][jj],
                                         s[2][jj], s[3][jj]);
            c_kt += st & 1;
            c_out += (st >> 1) & 1;
            c_side += (st >> 2) & 1;
            c_bin += (st >> 3) & 1;
        }
    }
    wave_add_stage(&H.stage[6], c_kt);
    wave_add_stage(&H.stage[7], c_out);
    wave_add_stage(&H.stage[8], c_side);
    wave_add_stage(&H.stage[9], c_bin);
}

}  // namespace hbt
