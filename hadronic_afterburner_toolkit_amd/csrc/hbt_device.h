// hbt_device.h -- device-side parameters, histogram views and the exact
// FP64 pair cascade shared by every HBT kernel.
//
// The cascade restates HBT_correlation.cpp:311-458 (same-event) and
// :574-687 (mixed) operation for operation.  It must be compiled with
// -ffp-contract=off: every decision (K_T cut, K_T index, window cuts, bin
// indices) then uses the same correctly rounded + - * / sqrt sequence as
// the reference's x86-64 build, so bin indices and counts are bit-exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hbt {

struct DevParams {
    int qnpts, n_KT, n_Kphi, nphi;
    int azimuthal, qinv, boost, pad0;
    double KT_min, dKT, dKphi, delta_q;
    double q_base, q_lo_cut, q_hi_cut;
    double KT_min_sq, KT_max_sq, hbarC_inv, two_pi;
    // Conservative prefilter thresholds (a pair the exact cascade accepts
    // always passes).  Per particle the record holds nrm = |pT|^2 - mid/2,
    // so for a pair u = T - mid and v = Q - mid come from one dot product
    // (T = |p1T+p2T|^2 = 4 K_perp^2, Q = |p1T-p2T|^2) and dn = |p1T|^2 -
    // |p2T|^2 from the two records.  As q_out = dn / sqrt(T) and q_side^2 =
    // Q - q_out^2, the windows |q_out|, |q_side| <= H read dn^2 <= H^2 T and
    // (Q - H^2) T <= dn^2; each side carries the rounding margin c.
    double pf_half;   // |T - mid| <= pf_half
    double pf_h2;     // H^2 with margin (0 with the q_inv branch on)
    double pf_midh2;  // mid - pf_h2
    double pf_negc;   // -c (-inf with the q_inv branch on: the q tests pass every pair)
    double pf_mid;
    float f_half, f_h2, f_midh2, f_negc, f_mid, f_pad;
    // precision-form exact-stage constants
    double inv_dKT, inv_dq, inv_two_pi, two_pi_hi, two_pi_lo, inv_dKphi;
    // K_phi angles closer than this (rad) to a bin edge, 0 or 2 pi are
    // decided on the host with glibc's atan2 (the device's may differ from it
    // in the last bits)
    double kphi_guard;
};

// Histogram bins, array of structures: 8 doubles (64 bytes, one aligned
// segment) per bin, so one bin's accumulations share a cache line.
enum { BIN_COUNT = 0, BIN_QO = 1, BIN_QS = 2, BIN_QL = 3, BIN_NUM = 4, BIN_DEN = 5,
       BIN_STRIDE = 8 };

// A pair whose K_phi bin the device leaves to the host: the cascade's
// inputs (same-event: both particles' px py pz E x y z t in the reference's
// orientation; mixed: px py pz E of particle 1 and of the rotated partner),
// psi_ref, and the histogram it belongs to.
struct DeferRec {
    double p1[8], p2[8];
    double psi;
    double *bins;
    int kind;  // 0 same-event, 1 mixed
    int pad;
    // the pair's place in the reference's enumeration order, recorded by the
    // ordered pair-cap path (row: particle of the same-event list; pos: j, or
    // partner << 32 | j for mixed pairs); -1 elsewhere
    long long row, pos;
};

struct DevHist {
    // [nbins][BIN_STRIDE]; COUNT holds the pair count as an FP64 integer
    // (exact below 2^53 pairs per bin), DEN a u64 bit pattern
    double *bins;
    unsigned long long *qinv_count, *qinv_den;
    double *qinv_sum, *qinv_num;
    unsigned long long *stage;  // [0..4] same, [5..9] mixed
    // pairs deferred to the host (K_phi at a bin edge): records and count;
    // a count above defer_cap is an overflow the host reports
    DeferRec *defer;
    unsigned *ndefer;
    unsigned defer_cap, pad;
};

// FP64 particle arrays in HBM (structure of arrays).
struct DevSoA {
    const double *__restrict__ px, *__restrict__ py, *__restrict__ pz, *__restrict__ E;
    const double *__restrict__ x, *__restrict__ y, *__restrict__ z, *__restrict__ t;
};

// One particle for the fast kernel's exact stage, array of structures:
// momentum then space-time in one 64-byte line, so a gathered particle
// costs one line request instead of eight.
struct __align__(64) PartRec {
    double px, py, pz, E, x, y, z, t;
};

// FP32 hi/lo split copies (v = hi + lo to ~2^-48 relative).
struct DevSoA32 {
    const float *__restrict__ pxh, *__restrict__ pyh, *__restrict__ pzh, *__restrict__ Eh;
    const float *__restrict__ pxl, *__restrict__ pyl, *__restrict__ pzl, *__restrict__ El;
    const float *__restrict__ xh, *__restrict__ yh, *__restrict__ zh, *__restrict__ th;
    const float *__restrict__ xl, *__restrict__ yl, *__restrict__ zl, *__restrict__ tl;
};

__device__ __forceinline__ void atomic_add_u64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add_f64(double *p, double v) {
    unsafeAtomicAdd(p, v);
}

__host__ __device__ __forceinline__ long bin_index(const DevParams &P, int kt, int kphi, int io,
                                                   int is, int il) {
    const long q = P.qnpts;
    return ((((long)kt * P.nphi + kphi) * q + io) * q + is) * q + il;
}

// K_phi bin with the reference's wrap loops (HBT_correlation.cpp:408-423).
// Returns -1 when the reference would skip the pair.  On the device, an
// angle within kphi_guard of 0 or of a bin edge (2 pi included) returns -3:
// the reference's glibc atan2 and the device's can differ there in the last
// bits, so the host decides the pair with glibc's atan2 (the host build of
// this function has no guard).
__host__ __device__ __forceinline__ int kphi_index(const DevParams &P, double K_x, double K_y,
                                                   double psi_ref) {
    double local_K_phi = atan2(K_y, K_x);
    double delta_phi = local_K_phi - psi_ref;
#if defined(__HIP_DEVICE_COMPILE__)
    if (fabs(delta_phi) < P.kphi_guard) return -3;
#endif
    while (delta_phi < 0.) delta_phi += P.two_pi;
    while (delta_phi > P.two_pi) delta_phi -= P.two_pi;
#if defined(__HIP_DEVICE_COMPILE__)
    {
        const double w = delta_phi / P.dKphi;
        if (fabs(w - rint(w)) * P.dKphi < P.kphi_guard) return -3;
    }
#endif
    int idx = (int)(delta_phi / P.dKphi);
    if (idx < 0 || idx >= P.n_Kphi) return -1;
    return idx;
}

// Result of the K/q cascade for one pair (shared by same and mixed).
struct PairBins {
    int kt, io, is, il, kphi;
    double q_out, q_side, q_long;
    double q_x, q_y, q_z, q_E;
};

// Stages reached, for the plain kernel's counters; ST_DEFER: every window
// passed and the K_phi bin is left to the host (no ST_BIN).
enum { ST_KT = 1, ST_OUT = 2, ST_SIDE = 4, ST_BIN = 8, ST_DEFER = 16 };

// The K/q cascade (HBT_correlation.cpp:311-398 / :574-647).  MIXED selects
// the mixed-event upper cuts (>=, :613, :622, :641).  Returns the stage
// mask; ST_BIN set means all windows passed and b holds the bins.  The
// q_inv branch is handled by the caller through qinv_* outputs.
// EXT = false compiles the q_inv and K_phi branches out (the caller
// guarantees invariant_radius_flag = azimuthal_flag = 0).
template <bool MIXED, bool EXT = true>
__host__ __device__ __forceinline__ int cascade_f64(const DevParams &P, double p1x, double p1y,
                                           double p1z, double p1E, double p2x, double p2y,
                                           double p2z, double p2E, double psi_ref,
                                           PairBins &b, double &q_inv, int &qinv_ok) {
    qinv_ok = 0;
    double K_z = 0.5 * (p1z + p2z);
    double K_E = 0.5 * (p1E + p2E);
    double K_z_over_K_E = K_z / K_E;
    double K_x = 0.5 * (p1x + p2x);
    double K_y = 0.5 * (p1y + p2y);
    double K_perp_sq = K_x * K_x + K_y * K_y;
    // written so that NaN is rejected (the reference indexes with it: UB)
    if (!(K_perp_sq >= P.KT_min_sq && K_perp_sq <= P.KT_max_sq)) return 0;
    int st = ST_KT;
    double K_perp = sqrt(K_perp_sq);
    b.kt = (int)((K_perp - P.KT_min) / P.dKT);
    if (b.kt < 0 || b.kt >= P.n_KT) return 0;  // cannot happen for KT_min <= K_perp <= KT_max
    b.q_x = p1x - p2x;
    b.q_y = p1y - p2y;
    b.q_z = p1z - p2z;
    b.q_E = p1E - p2E;
    if (EXT && P.qinv) {
        q_inv = sqrt(-(b.q_E * b.q_E - b.q_x * b.q_x - b.q_y * b.q_y - b.q_z * b.q_z));
        if (q_inv > P.q_lo_cut && q_inv < P.q_hi_cut) {
            int qi = (int)((q_inv - P.q_base) / P.delta_q);
            if (qi < P.qnpts) qinv_ok = qi + 1;
        }
    }
    double cos_K_phi = K_x / K_perp;
    double sin_K_phi = K_y / K_perp;
    double q_out = b.q_x * cos_K_phi + b.q_y * sin_K_phi;
    if (MIXED) {
        if (!(q_out >= P.q_lo_cut && q_out < P.q_hi_cut)) return st;
    } else {
        if (!(q_out >= P.q_lo_cut && q_out <= P.q_hi_cut)) return st;
    }
    b.io = (int)((q_out - P.q_base) / P.delta_q);
    if (b.io >= P.qnpts || b.io < 0) return st;
    st |= ST_OUT;
    double q_side = b.q_y * cos_K_phi - b.q_x * sin_K_phi;
    if (MIXED) {
        if (!(q_side >= P.q_lo_cut && q_side < P.q_hi_cut)) return st;
    } else {
        if (!(q_side >= P.q_lo_cut && q_side <= P.q_hi_cut)) return st;
    }
    b.is = (int)((q_side - P.q_base) / P.delta_q);
    if (b.is >= P.qnpts || b.is < 0) return st;
    st |= ST_SIDE;
    double q_long = b.q_z;
    if (P.boost) {
        double Mt = sqrt(K_E * K_E - K_z * K_z);
        double boost_gamma = K_E / Mt;
        double boost_beta = K_z_over_K_E;
        q_long = boost_gamma * (b.q_z - boost_beta * b.q_E);
    }
    if (MIXED) {
        if (!(q_long >= P.q_lo_cut && q_long < P.q_hi_cut)) return st;
    } else {
        if (!(q_long >= P.q_lo_cut && q_long <= P.q_hi_cut)) return st;
    }
    b.il = (int)((q_long - P.q_base) / P.delta_q);
    if (b.il >= P.qnpts || b.il < 0) return st;
    b.kphi = 0;
    b.q_out = q_out;
    b.q_side = q_side;
    b.q_long = q_long;
    if (EXT && P.azimuthal) {
        b.kphi = kphi_index(P, K_x, K_y, psi_ref);
        if (b.kphi == -3) return st | ST_DEFER;
        if (b.kphi < 0) return st;
    }
    return st | ST_BIN;
}

// Records a pair for the host (a count beyond defer_cap keeps counting, so
// the host sees the overflow).
__device__ __forceinline__ void push_defer(const DevHist &H, const double a[8], const double c[8],
                                           double psi, int kind, long long row = -1,
                                           long long pos = -1) {
    const unsigned slot = atomicAdd(H.ndefer, 1u);
    if (slot >= H.defer_cap) return;
    DeferRec &d = H.defer[slot];
    for (int f = 0; f < 8; f++) {
        d.p1[f] = a[f];
        d.p2[f] = c[f];
    }
    d.psi = psi;
    d.bins = H.bins;
    d.kind = kind;
    d.pad = 0;
    d.row = row;
    d.pos = pos;
}

// One binned same-event pair: bin and the four FP64 sums it adds to.
struct SameRec {
    long k;
    double qo, qs, ql, c;
    int s, is, il;  // slab (K_T, K_phi, q_out plane) and q_side, q_long bins of k
};

// Same-event pair: cascade and cos (HBT_correlation.cpp:339-356, :431-446).
// q_inv contributions are accumulated here; for ST_BIN the caller adds rec.
template <bool EXT = true>
__device__ __forceinline__ int eval_same_f64(const DevParams &P, const DevHist &H, double psi_ref,
                                             double p1x, double p1y, double p1z, double p1E,
                                             double x1, double y1, double z1, double t1,
                                             double p2x, double p2y, double p2z, double p2E,
                                             double x2, double y2, double z2, double t2,
                                             SameRec &rec) {
    PairBins b;
    double q_inv;
    int qinv_ok;
    int st = cascade_f64<false, EXT>(P, p1x, p1y, p1z, p1E, p2x, p2y, p2z, p2E, psi_ref, b,
                                     q_inv, qinv_ok);
    if (!(st & (ST_BIN | ST_DEFER)) && !qinv_ok) return st;
    double t_diff = t1 - t2;
    double x_diff = x1 - x2;
    double y_diff = y1 - y2;
    double z_diff = z1 - z2;
    double cos_qx = cos(P.hbarC_inv *
                        (b.q_E * t_diff - b.q_x * x_diff - b.q_y * y_diff - b.q_z * z_diff));
    if (qinv_ok) {
        long k = (long)b.kt * P.qnpts + (qinv_ok - 1);
        atomic_add_u64(&H.qinv_count[k], 1ull);
        atomic_add_f64(&H.qinv_sum[k], q_inv);
        atomic_add_f64(&H.qinv_num[k], cos_qx);
    }
    if (EXT && (st & ST_DEFER)) {
        const double a[8] = {p1x, p1y, p1z, p1E, x1, y1, z1, t1};
        const double c[8] = {p2x, p2y, p2z, p2E, x2, y2, z2, t2};
        push_defer(H, a, c, psi_ref, 0);
    }
    if (st & ST_BIN) {
        rec.k = bin_index(P, b.kt, b.kphi, b.io, b.is, b.il);
        rec.s = (b.kt * P.nphi + b.kphi) * P.qnpts + b.io;
        rec.is = b.is;
        rec.il = b.il;
        rec.qo = b.q_out;
        rec.qs = b.q_side;
        rec.ql = b.q_long;
        rec.c = cos_qx;
    }
    return st;
}

// Mixed pair: cascade (HBT_correlation.cpp:592-607); q_inv denominators are
// accumulated here, for ST_BIN the caller adds 1 to the den of bin k.
template <bool EXT = true>
__device__ __forceinline__ int eval_mixed_f64(const DevParams &P, const DevHist &H, double psi_ref,
                                              double p1x, double p1y, double p1z, double p1E,
                                              double p2x, double p2y, double p2z, double p2E,
                                              long &k) {
    PairBins b;
    double q_inv;
    int qinv_ok;
    int st = cascade_f64<true, EXT>(P, p1x, p1y, p1z, p1E, p2x, p2y, p2z, p2E, psi_ref, b, q_inv,
                                    qinv_ok);
    if (qinv_ok) atomic_add_u64(&H.qinv_den[(long)b.kt * P.qnpts + (qinv_ok - 1)], 1ull);
    if (EXT && (st & ST_DEFER)) {
        const double a[8] = {p1x, p1y, p1z, p1E, 0, 0, 0, 0};
        const double c[8] = {p2x, p2y, p2z, p2E, 0, 0, 0, 0};
        push_defer(H, a, c, psi_ref, 1);
    }
    if (st & ST_BIN) k = bin_index(P, b.kt, b.kphi, b.io, b.is, b.il);
    return st;
}

// Direct accumulation of one pair (HBT_correlation.cpp:437-446, :681-683).
__device__ __forceinline__ void add_same_direct(const DevHist &H, const SameRec &r) {
    double *b = H.bins + r.k * BIN_STRIDE;
    atomic_add_f64(b + BIN_COUNT, 1.0);
    atomic_add_f64(b + BIN_QO, r.qo);
    atomic_add_f64(b + BIN_QS, r.qs);
    atomic_add_f64(b + BIN_QL, r.ql);
    atomic_add_f64(b + BIN_NUM, r.c);
}
// The denominator is a u64 in memory, incremented as a u32 add on its low
// word, with the carry added to the high word by the add that wraps: the
// value stays exact, and 32-bit adds cost ~14 % less at the memory side,
// where histogram atomics are executed.
__device__ __forceinline__ void add_mixed_direct(const DevHist &H, long k) {
    unsigned *lo = (unsigned *)(H.bins + k * BIN_STRIDE + BIN_DEN);  // little-endian low word
    if (__hip_atomic_fetch_add(lo, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0xffffffffu)
        __hip_atomic_fetch_add(lo + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool EXT = true>
__device__ __forceinline__ int bin_same_f64(const DevParams &P, const DevHist &H, double psi_ref,
                                            double p1x, double p1y, double p1z, double p1E,
                                            double x1, double y1, double z1, double t1,
                                            double p2x, double p2y, double p2z, double p2E,
                                            double x2, double y2, double z2, double t2) {
    SameRec r;
    const int st = eval_same_f64<EXT>(P, H, psi_ref, p1x, p1y, p1z, p1E, x1, y1, z1, t1, p2x,
                                      p2y, p2z, p2E, x2, y2, z2, t2, r);
    if (st & ST_BIN) add_same_direct(H, r);
    return st;
}

template <bool EXT = true>
__device__ __forceinline__ int bin_mixed_f64(const DevParams &P, const DevHist &H, double psi_ref,
                                             double p1x, double p1y, double p1z, double p1E,
                                             double p2x, double p2y, double p2z, double p2E) {
    long k = 0;
    const int st =
        eval_mixed_f64<EXT>(P, H, psi_ref, p1x, p1y, p1z, p1E, p2x, p2y, p2z, p2E, k);
    if (st & ST_BIN) add_mixed_direct(H, k);
    return st;
}

}  // namespace hbt
