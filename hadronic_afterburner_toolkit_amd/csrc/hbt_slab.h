// hbt_slab.h -- slab binning: the histogram accumulation of the fast kernel
// without device-scope atomics.
//
// Every binned pair is one read-modify-write of a random 64-byte bin.  Done
// as a device-scope atomic, each such access is executed at the memory side
// of this part at about 2.35e10 per second for the whole chip, whatever the
// histogram's size or placement: that rate, not arithmetic, bounded the
// fast kernel.  The histogram (qnpts^3 bins per K_T bin, tens of MB) cannot
// live in LDS, and a block's binned pairs land in random bins, so there is
// nothing to combine inside one block.  Instead:
//
//  1. k_fast writes each binned pair as a record into the region of its
//     (block, slab): a slab is one (K_T, K_phi, q_out) plane of qnpts^2
//     bins; the block's LDS fill counter of the slab gives the record's
//     place, so no global atomic is involved.  Same-event record: q_out,
//     q_side, q_long, cos (FP32 in the precision form, FP64 in the FP64
//     form) + the bin inside the slab (u16); mixed record: the bin (u16).
//     A record whose region is full takes the direct atomic path instead
//     (always correct; the host sizes the regions from the demand of earlier
//     launches so that this is rare).
//  2. k_slab_bin: block (slab, part) adds the records of a range of regions
//     of its slab into LDS (u32 counts, FP64 sums) and stores the part's
//     totals.
//  3. k_slab_reduce: one thread per bin adds the parts of its bin to the
//     histogram (count, sums, den), plain read-modify-write: each bin has
//     one owner.
//
// In slab mode the precision form's per-pair values are FP32 (records and
// the direct path alike), accumulated in FP64; without slab binning they
// stay FP64.
#pragma once
#include <type_traits>

#include "hbt_device.h"

namespace hbt {

constexpr int kMaxSlabs = 512;     // slabs per analysis (LDS fill counters of k_fast)
constexpr int kMaxSlabBins = 2048;  // bins per slab (LDS of k_slab_bin)

struct SlabArgs {
    void *vals;                // same-event values: float4 (precision form) or 2 x double2
    unsigned short *sidx;      // same-event records' bin inside the slab
    unsigned short *midx;      // mixed records' bin inside the slab
    const long long *sbase;    // [S] first record of slab s's regions (block 0)
    const long long *mbase;
    const int *scap;           // [S] records per (block, slab) region
    const int *mcap;
    unsigned *sfill;           // [blocks][S] records each block had for the slab
    unsigned *mfill;           //   (counted past the capacity: the demand)
    unsigned long long *stats; // [4S + 1]: demand per slab (same, mixed), direct adds,
                               // largest demand of one block per slab (same, mixed)
    int S, SB;                 // slabs, bins per slab
};

// Partial totals of one slab part, per bin.
struct SlabPart {
    double cnt, qo, qs, ql, c;
    unsigned long long den;
};

template <typename Prec>
__device__ __forceinline__ bool slab_put_same(const SlabArgs &a, unsigned *fill, long k,
                                              double qo, double qs, double ql, double c) {
    const int s = (int)(k / a.SB);
    const int bi = (int)(k - (long)s * a.SB);
    const unsigned pos = atomicAdd(&fill[s], 1u);
    const int cap = a.scap[s];
    if (pos >= (unsigned)cap) return false;
    const long long e = a.sbase[s] + (long long)blockIdx.x * cap + pos;
    if (std::is_same<Prec, float>::value) {
        ((float4 *)a.vals)[e] = make_float4((float)qo, (float)qs, (float)ql, (float)c);
    } else {
        double2 *d = (double2 *)a.vals + 2 * e;
        d[0] = make_double2(qo, qs);
        d[1] = make_double2(ql, c);
    }
    a.sidx[e] = (unsigned short)bi;
    return true;
}

__device__ __forceinline__ bool slab_put_mixed(const SlabArgs &a, unsigned *fill, long k) {
    const int s = (int)(k / a.SB);
    const int bi = (int)(k - (long)s * a.SB);
    const unsigned pos = atomicAdd(&fill[kMaxSlabs + s], 1u);
    const int cap = a.mcap[s];
    if (pos >= (unsigned)cap) return false;
    a.midx[a.mbase[s] + (long long)blockIdx.x * cap + pos] = (unsigned short)bi;
    return true;
}

// Block (slab s, part p) of 1024 threads: the records of regions
// [p * nblk / parts, (p + 1) * nblk / parts) of slab s into LDS, then the
// part's totals to out[(s * parts + p) * SB + bin].  Each part adds its
// regions' demand (records wanted, also past their capacity) to stats.
template <typename Prec>
__global__ __launch_bounds__(1024) void k_slab_bin(SlabArgs a, int nblk, int parts,
                                                   SlabPart *out) {
    __shared__ unsigned cnt[kMaxSlabBins], den[kMaxSlabBins];
    __shared__ double sum[4 * kMaxSlabBins];
    const int s = blockIdx.x, p = blockIdx.y;
    const int SB = a.SB;
    for (int i = threadIdx.x; i < SB; i += blockDim.x) {
        cnt[i] = 0u;
        den[i] = 0u;
        sum[4 * i] = sum[4 * i + 1] = sum[4 * i + 2] = sum[4 * i + 3] = 0.0;
    }
    __syncthreads();
    const int b0 = (int)((long long)p * nblk / parts), b1 = (int)((long long)(p + 1) * nblk / parts);
    const int scap = a.scap[s], mcap = a.mcap[s];
    unsigned long long ds = 0, dm = 0;
    unsigned xs = 0, xm = 0;
    for (int b = b0; b < b1; b++) {
        const unsigned fs = a.sfill[(long long)b * a.S + s], fm = a.mfill[(long long)b * a.S + s];
        ds += fs;
        dm += fm;
        xs = max(xs, fs);
        xm = max(xm, fm);
        const int ns = (int)min(fs, (unsigned)scap), nm = (int)min(fm, (unsigned)mcap);
        const long long es = a.sbase[s] + (long long)b * scap;
        for (int r = threadIdx.x; r < ns; r += blockDim.x) {
            const int bi = a.sidx[es + r];
            double v0, v1, v2, v3;
            if (std::is_same<Prec, float>::value) {
                const float4 f = ((const float4 *)a.vals)[es + r];
                v0 = f.x, v1 = f.y, v2 = f.z, v3 = f.w;
            } else {
                const double2 *d = (const double2 *)a.vals + 2 * (es + r);
                const double2 x = d[0], y = d[1];
                v0 = x.x, v1 = x.y, v2 = y.x, v3 = y.y;
            }
            atomicAdd(&cnt[bi], 1u);
            atomicAdd(&sum[4 * bi], v0);
            atomicAdd(&sum[4 * bi + 1], v1);
            atomicAdd(&sum[4 * bi + 2], v2);
            atomicAdd(&sum[4 * bi + 3], v3);
        }
        const long long em = a.mbase[s] + (long long)b * mcap;
        for (int r = threadIdx.x; r < nm; r += blockDim.x) atomicAdd(&den[a.midx[em + r]], 1u);
    }
    if (threadIdx.x == 0) {  // demand of this part's regions (thread-uniform values)
        atomicAdd(&a.stats[s], ds);
        atomicAdd(&a.stats[a.S + s], dm);
        atomicMax(&a.stats[2 * a.S + 1 + s], (unsigned long long)xs);
        atomicMax(&a.stats[3 * a.S + 1 + s], (unsigned long long)xm);
    }
    __syncthreads();
    SlabPart *o = out + ((long long)s * parts + p) * SB;
    for (int i = threadIdx.x; i < SB; i += blockDim.x)
        o[i] = SlabPart{(double)cnt[i], sum[4 * i], sum[4 * i + 1], sum[4 * i + 2],
                        sum[4 * i + 3], (unsigned long long)den[i]};
}

// One thread per bin of the slabs: the parts' totals into the histogram.
// Slab s covers bins [s * SB, (s + 1) * SB) of the [kt][kphi][q_out][q_side]
// [q_long] layout.
static __global__ void k_slab_reduce(const SlabPart *parts_out, int S, int SB, int parts, double *bins) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)S * SB) return;
    const int s = (int)(k / SB), bi = (int)(k - (long long)s * SB);
    double c = 0, qo = 0, qs = 0, ql = 0, cs = 0;
    unsigned long long d = 0;
    for (int p = 0; p < parts; p++) {
        const SlabPart &x = parts_out[((long long)s * parts + p) * SB + bi];
        c += x.cnt;
        qo += x.qo;
        qs += x.qs;
        ql += x.ql;
        cs += x.c;
        d += x.den;
    }
    if (c == 0.0 && d == 0) return;
    double *b = bins + k * BIN_STRIDE;
    if (c != 0.0) {
        b[BIN_COUNT] += c;
        b[BIN_QO] += qo;
        b[BIN_QS] += qs;
        b[BIN_QL] += ql;
        b[BIN_NUM] += cs;
    }
    if (d) ((unsigned long long *)b)[BIN_DEN] += d;
}

}  // namespace hbt
