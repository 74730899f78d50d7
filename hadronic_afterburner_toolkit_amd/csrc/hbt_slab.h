// hbt_slab.h -- slab binning: the same-event histogram accumulation of the
// fast kernel without device-scope atomics.
//
// Every binned pair is one read-modify-write of a random 64-byte bin.  Done
// as a device-scope atomic, each such access is executed at the memory side
// of this part at about 2.35e10 per second for the whole chip, whatever the
// histogram's size or placement: that rate, not arithmetic, bounds the fast
// kernel.  The histogram (qnpts^3 bins per K_T bin, tens of MB) cannot live
// in LDS, and a block's binned pairs land in random bins, so there is
// nothing to combine inside one block.  Instead, for same-event pairs:
//
//  1. k_fast writes each binned pair as a record into the region of its
//     (block, slab): a slab is one (K_T, K_phi, q_out) plane of qnpts^2 bins; the
//     block's LDS fill counter of the slab gives the record's place, so no
//     global atomic is involved.  A record is the pair's q_out, q_side,
//     q_long and cos: one 16-byte store (FP32, the precision form) or 32
//     bytes (FP64).  It holds no bin index: k_slab_bin recovers the bin
//     inside the slab from the stored q_side and q_long (slab_coord), and
//     k_fast writes the record only when that gives the pair's bin, so a
//     pair whose stored values round across a bin edge takes the direct
//     atomic path.  So does a record whose region is full (the host sizes
//     the regions from the demand of earlier launches so that this is rare).
//  2. k_slab_bin: one block per unit (slab, range of regions) adds its
//     records into LDS (u32 counts; sums as 64-bit fixed point in the
//     precision form, FP64 in the FP64 form) and stores the unit's totals.  Slabs get units in proportion to their demand on the last
//     launch, so the units carry similar work.
//  3. k_slab_reduce: one thread per bin adds the units of its slab to the
//     histogram, plain read-modify-write: each bin has one owner.
//
// Mixed pairs add one u32 to the denominator directly: their atomic is the
// cheaper pattern, and a 2-byte record store per pair measured costlier than
// it.
//
// In slab mode the precision form's per-pair values are FP32 (records and
// the direct path alike), accumulated in FP64; without slab binning they
// stay FP64.
#pragma once
#include <type_traits>

#include "hbt_device.h"

namespace hbt {

constexpr int kMaxSlabs = 512;     // slabs per analysis (LDS fill counters of k_fast)
// slabs of a K_phi analysis: its instance of k_fast keeps only the fill
// counters in LDS, in the launch's dynamic LDS (S x 4 bytes, at most 16 KB;
// capacity and first record come from SlabArgs), enough for 6 K_T x 16 K_phi
// x 41 q_out planes
constexpr int kMaxSlabsKphi = 4096;
constexpr int kMaxSlabBins = 2048;  // bins per slab (LDS of k_slab_bin)
constexpr int kMaxUnitRecords = (1 << 23) - 1;  // records of one k_slab_bin unit

struct SlabArgs {
    void *vals;                // records: float4 (precision form) or 2 x double2
    const long long *sbase;    // [S] first record of slab s's regions (block 0)
    const int *scap;           // [S] records per (block, slab) region
    unsigned *sfill;           // [blocks][S] records each block had for the slab
                               //   (counted past the capacity: the demand)
    unsigned long long *stats; // [2S + 1]: demand per slab, direct adds,
                               // largest demand of one block per slab
    const int4 *units;         // k_slab_bin blocks: (slab, first region, end region, 0)
    const int *uoff;           // [S + 1] first unit of each slab
    int S, SB, q;              // slabs, bins per slab, q bins per axis
    double q_base, inv_dq;     // q_side / q_long bin of a stored value
};

// Totals of one unit, per bin of its slab.
struct SlabPart {
    double cnt, qo, qs, ql, c;
};

// The q_side or q_long bin of a record's stored value.  k_fast and
// k_slab_bin evaluate it on the same stored value with the same two
// rounded operations (the build contracts nothing), so they agree bit for
// bit; k_fast writes a record only where it equals the pair's bin.
__device__ __forceinline__ int slab_coord(const SlabArgs &a, double v) {
    return (int)((v - a.q_base) * a.inv_dq);
}

// An FP32 record value k_slab_bin sums exactly in its fixed point: 0, or
// 2^-17 <= |v| <= 1 (then v is a multiple of 2^-40 and |v| 2^40 <= 2^40).
__device__ __forceinline__ bool fix_exact(float v) {
    const float m = fabsf(v);
    return m == 0.0f || (m >= 0x1p-17f && m <= 1.0f);
}

// A binned same-event pair r as a record of this block's region of slab
// r.s; false: the pair takes the direct path (also for a precision-form
// value the fixed point would not hold exactly).  fill: the block's LDS fill
// counters; cap and base: the regions' capacity and first record (as in
// SlabArgs), LDS copies or SlabArgs' own arrays.
template <typename Prec>
__device__ __forceinline__ bool slab_put_same(const SlabArgs &a, unsigned *fill, const int *cap,
                                              const long long *base, const SameRec &r) {
    constexpr bool f32 = std::is_same<Prec, float>::value;
    if (slab_coord(a, f32 ? (double)(float)r.qs : r.qs) != r.is ||
        slab_coord(a, f32 ? (double)(float)r.ql : r.ql) != r.il)
        return false;
    if (f32 && !(fix_exact((float)r.qo) && fix_exact((float)r.qs) && fix_exact((float)r.ql) &&
                 fix_exact((float)r.c)))
        return false;
    const unsigned pos = atomicAdd(&fill[r.s], 1u);
    const int c = cap[r.s];
    if (pos >= (unsigned)c) return false;
    const long long e = base[r.s] + (long long)blockIdx.x * c + pos;
    if (f32) {
        ((float4 *)a.vals)[e] = make_float4((float)r.qo, (float)r.qs, (float)r.ql, (float)r.c);
    } else {
        double2 *d = (double2 *)a.vals + 2 * e;
        d[0] = make_double2(r.qo, r.qs);
        d[1] = make_double2(r.ql, r.c);
    }
    return true;
}

template <typename Prec>
__device__ __forceinline__ void slab_load(const SlabArgs &a, long long e, double &v0, double &v1,
                                          double &v2, double &v3) {
    if (std::is_same<Prec, float>::value) {
        const float4 f = ((const float4 *)a.vals)[e];
        v0 = f.x, v1 = f.y, v2 = f.z, v3 = f.w;
    } else {
        const double2 *d = (const double2 *)a.vals + 2 * e;
        const double2 x = d[0], y = d[1];
        v0 = x.x, v1 = x.y, v2 = y.x, v3 = y.y;
    }
}

// Fixed-point scale of the precision form's LDS sums: a record value is 0
// or 2^-17 <= |v| <= 1 (fix_exact; other pairs take the direct path), so
// v * 2^40 is an exact integer, and at most kMaxUnitRecords records per unit
// keep every sum below 2^63.  Integer sums are exact,
// so a unit's totals do not depend on the order of its LDS adds, and the
// 64-bit integer LDS add is cheaper than the FP64 one.
constexpr double kFixScale = 1099511627776.0;  // 2^40

// The LDS sums are four arrays of kMaxSlabBins (field f of bin bi at
// f * kMaxSlabBins + bi): one 8-byte add per lane lands on the bank pair of
// its bin, so 64 random bins spread over all banks.  (Four sums side by side
// per bin, 32 bytes apart, left a quarter of the banks to each add and
// measured 2.9e8 bank-conflict cycles per step.)
template <typename Prec>
__device__ __forceinline__ void slab_add(const SlabArgs &a, unsigned *cnt, double *sum,
                                         double v0, double v1, double v2, double v3) {
    // clamped for safety only: a stored record's coordinates are in range
    const int is = min(max(slab_coord(a, v1), 0), a.q - 1);
    const int il = min(max(slab_coord(a, v2), 0), a.q - 1);
    const int bi = is * a.q + il;
    atomicAdd(&cnt[bi], 1u);
    constexpr int F = kMaxSlabBins;
    if (std::is_same<Prec, float>::value) {
        unsigned long long *fix = (unsigned long long *)sum;
        atomicAdd(&fix[bi], (unsigned long long)__double2ll_rn(v0 * kFixScale));
        atomicAdd(&fix[F + bi], (unsigned long long)__double2ll_rn(v1 * kFixScale));
        atomicAdd(&fix[2 * F + bi], (unsigned long long)__double2ll_rn(v2 * kFixScale));
        atomicAdd(&fix[3 * F + bi], (unsigned long long)__double2ll_rn(v3 * kFixScale));
    } else {
        atomicAdd(&sum[bi], v0);
        atomicAdd(&sum[F + bi], v1);
        atomicAdd(&sum[2 * F + bi], v2);
        atomicAdd(&sum[3 * F + bi], v3);
    }
}

// Block u of 1024 threads: unit (slab s, regions [r0, r1)).  Each wave takes
// every 16th region of the range and its lanes the region's records, two
// loads in flight per lane; the totals go to out[u * SB + bin].  LDS sums:
// fixed point (kFixScale) in the precision form, FP64 in the FP64 form.  Each unit
// adds its regions' demand (records wanted, also past their capacity) to
// stats.
template <typename Prec>
__global__ __launch_bounds__(1024) void k_slab_bin(SlabArgs a, SlabPart *out) {
    __shared__ unsigned cnt[kMaxSlabBins];
    __shared__ double sum[4 * kMaxSlabBins];
    const int4 u = a.units[blockIdx.x];
    const int s = u.x, SB = a.SB;
    for (int i = threadIdx.x; i < SB; i += blockDim.x) {
        cnt[i] = 0u;
        for (int f = 0; f < 4; f++) sum[f * kMaxSlabBins + i] = 0.0;  // also integer 0
    }
    __syncthreads();
    const int scap = a.scap[s];
    const long long base = a.sbase[s];
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    unsigned long long demand = 0;  // wave-uniform
    unsigned most = 0;
    for (int b = u.y + (int)(threadIdx.x >> 6); b < u.z; b += nw) {
        const unsigned f = a.sfill[(long long)b * a.S + s];
        demand += f;
        most = max(most, f);
        const int n = (int)min(f, (unsigned)scap);
        const long long e0 = base + (long long)b * scap;
        for (int r = lane; r < n; r += 128) {
            double x0, x1, x2, x3, y0 = 0, y1 = 0, y2 = 0, y3 = 0;
            const bool two = r + 64 < n;
            slab_load<Prec>(a, e0 + r, x0, x1, x2, x3);
            if (two) slab_load<Prec>(a, e0 + r + 64, y0, y1, y2, y3);
            slab_add<Prec>(a, cnt, sum, x0, x1, x2, x3);
            if (two) slab_add<Prec>(a, cnt, sum, y0, y1, y2, y3);
        }
    }
    if (lane == 0 && demand) {
        atomicAdd(&a.stats[s], demand);
        atomicMax(&a.stats[a.S + 1 + s], (unsigned long long)most);
    }
    __syncthreads();
    SlabPart *o = out + (long long)blockIdx.x * SB;
    for (int i = threadIdx.x; i < SB; i += blockDim.x) {
        double t[4];
        for (int f = 0; f < 4; f++)
            t[f] = std::is_same<Prec, float>::value
                       ? (double)((const long long *)sum)[f * kMaxSlabBins + i] / kFixScale
                       : sum[f * kMaxSlabBins + i];
        o[i] = SlabPart{(double)cnt[i], t[0], t[1], t[2], t[3]};
    }
}

// One thread per bin of the slabs: the units' totals into the histogram.
// Slab s covers bins [s * SB, (s + 1) * SB) of the [kt][q_out][q_side]
// [q_long] layout.
static __global__ void k_slab_reduce(const SlabPart *parts, const int *uoff, int S, int SB,
                                     double *bins) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)S * SB) return;
    const int s = (int)(k / SB), bi = (int)(k - (long long)s * SB);
    double c = 0, qo = 0, qs = 0, ql = 0, cs = 0;
    for (int u = uoff[s]; u < uoff[s + 1]; u++) {
        const SlabPart &x = parts[(long long)u * SB + bi];
        c += x.cnt;
        qo += x.qo;
        qs += x.qs;
        ql += x.ql;
        cs += x.c;
    }
    if (c == 0.0) return;
    double *b = bins + k * BIN_STRIDE;
    b[BIN_COUNT] += c;
    b[BIN_QO] += qo;
    b[BIN_QS] += qs;
    b[BIN_QL] += ql;
    b[BIN_NUM] += cs;
}

}  // namespace hbt
