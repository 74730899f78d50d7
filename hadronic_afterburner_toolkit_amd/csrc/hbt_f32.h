// hbt_f32.h -- exact stage of the precision form (HBT_PREC_FP32C).
//
// On CDNA4 an unpacked FP32 add/mul/fma issues at the same rate as its FP64
// form, so FP32 only pays where FP64 needs a multi-instruction sequence:
// division, square root and cos.  The precision form therefore keeps the
// pair's momentum sums/differences and products in FP64 (one instruction
// each, exact same rounding as the reference), takes reciprocals and
// reciprocal square roots from the FP32 transcendental unit
// (v_rcp_f32 / v_rsq_f32, ~1 ulp) refined by one Newton step in FP64, and
// evaluates cos on the FP32 unit after an FP64 Cody-Waite reduction to
// [-pi, pi].  All survivors of the FP32 prefilter come here.
//
// Bit-identical counts: every decision value (K_T index, window cuts, bin
// indices) is within 1e-13 relative of the reference's correctly rounded value;
// a pair whose value lies within a guard band of a cut or bin edge is
// re-run through the literal FP64 cascade (bin_*_f64), so bins and counts
// equal the reference's exactly.  The accumulated q values carry ~1e-13
// relative error, cos ~1e-7 absolute (FP32 in slab records); sums are
// accumulated in FP64 (atomics) or exact fixed point (slab binning).
#pragma once
#include "hbt_device.h"

namespace hbt {

// 1/sqrt(a) from v_rsq_f32 (relative error e <= 2^-22 for a normal FP32
// input) and one Newton step in FP64: relative error ~1.5 e^2 < 1e-13, far
// inside the guard bands (1e-9 bin units, 1e-12 relative on the cuts).
__device__ __forceinline__ double rsqrt_refined(double a) {
    const double r = (double)__builtin_amdgcn_rsqf((float)a);
    return r * fma(-0.5 * a * r, r, 1.5);
}
// 1/a from v_rcp_f32 and one Newton step in FP64: relative error e^2.
__device__ __forceinline__ double rcp_refined(double a) {
    const double r = (double)__builtin_amdgcn_rcpf((float)a);
    return r * fma(-a, r, 2.0);
}

// cos(x): argument reduced to [-pi, pi] in FP64, then split r = rh + rl
// with rh in FP32; cos and sin of rh on FP32 and the first-order
// compensation cos(rh + rl) = cos(rh) - rl sin(rh) (|rl| < 2^-22, so the
// dropped term rl^2/2 is below 1e-13).  Error ~1 FP32 ulp of the result.
__device__ __forceinline__ double cos_reduced_f32(const DevParams &P, double x) {
    const double k = rint(x * P.inv_two_pi);
    const double r = fma(-k, P.two_pi_lo, fma(-k, P.two_pi_hi, x));
    const float rh = (float)r;
    const float rl = (float)(r - (double)rh);
    float s, c;
    sincosf(rh, &s, &c);
    return (double)c - (double)rl * (double)s;
}

// Index with a guard: returns -2 when (v - base)*inv lies within g of an
// integer boundary (the caller then falls back to the literal cascade).
__device__ __forceinline__ int guarded_index(double w, double g) {
    const double n = rint(w);
    if (fabs(w - n) < g && n != 0.0) return -2;
    return (int)w;
}

// Window test with guard: 1 inside, 0 outside, -1 ambiguous.  upper_open
// selects the mixed-event form (v < hi) instead of (v <= hi).
__device__ __forceinline__ int guarded_window(double v, double lo, double hi, double g) {
    if (fabs(v - lo) < g || fabs(v - hi) < g) return -1;
    return (v >= lo && v <= hi) ? 1 : 0;
}

constexpr double GUARD_IDX = 1e-9;   // in bin-index units
constexpr double GUARD_REL = 1e-12;  // relative to the q scale of the pair
constexpr double GUARD_PHI = 4e-6;   // radians, around K_phi bin edges

// K_phi bin of the precision form (the reference's wrap loops,
// HBT_correlation.cpp:408-423): atan2 on the FP32 unit (error below 1e-6
// rad); an angle within GUARD_PHI of a bin edge, 0 or 2 pi returns -2 and
// the pair is re-run through the literal cascade.
__device__ __forceinline__ int kphi_index_fast(const DevParams &P, double K_x, double K_y,
                                               double psi_ref) {
    double delta = (double)atan2f((float)K_y, (float)K_x) - psi_ref;
    while (delta < 0.) delta += P.two_pi;
    while (delta > P.two_pi) delta -= P.two_pi;
    const double w = delta * P.inv_dKphi;
    const double n = rint(w);
    if (fabs(w - n) * P.dKphi < GUARD_PHI) return -2;
    const int idx = (int)w;
    return (idx < 0 || idx >= P.n_Kphi) ? -2 : idx;
}

// The K/q cascade of the precision form.  Returns the stage mask as
// cascade_f64 does, or -1 when a decision is ambiguous.
// neg: the pair is passed as (particle 2, particle 1) of the reference's
// orientation.  K is symmetric in the two particles and IEEE subtraction
// is antisymmetric (a - b == -(b - a) exactly), so q_out, q_side and q_long
// of the reference's orientation are exactly the negated ones of the pair
// as passed; b.q_x .. b.q_E keep the orientation as passed (the phase,
// a sum of products of two differences, is symmetric).
template <bool MIXED, bool EXT = false>
__device__ __forceinline__ int cascade_fast(const DevParams &P, double p1x, double p1y,
                                            double p1z, double p1E, double p2x, double p2y,
                                            double p2z, double p2E, double psi_ref,
                                            PairBins &b, bool neg = false) {
    const double K_z = 0.5 * (p1z + p2z);
    const double K_E = 0.5 * (p1E + p2E);
    const double K_x = 0.5 * (p1x + p2x);
    const double K_y = 0.5 * (p1y + p2y);
    const double K_perp_sq = K_x * K_x + K_y * K_y;  // same ops as the reference
    if (!(K_perp_sq >= P.KT_min_sq && K_perp_sq <= P.KT_max_sq)) return 0;
    int st = ST_KT;
    if (!(K_perp_sq >= 1e-30)) return -1;  // below FP32's normal range: the literal cascade
    const double rK = rsqrt_refined(K_perp_sq);
    const double K_perp = K_perp_sq * rK;
    b.kt = guarded_index((K_perp - P.KT_min) * P.inv_dKT, GUARD_IDX);
    if (b.kt < -1) return -1;
    if (b.kt < 0 || b.kt >= P.n_KT) return -1;  // out of range: decide with the literal cascade
    b.q_x = p1x - p2x;
    b.q_y = p1y - p2y;
    b.q_z = p1z - p2z;
    b.q_E = p1E - p2E;
    const double gq = GUARD_REL * (fabs(b.q_x) + fabs(b.q_y) + 1e-3);
    const double cK = K_x * rK, sK = K_y * rK;
    double q_out = b.q_x * cK + b.q_y * sK;
    if (neg) q_out = -q_out;
    int w = guarded_window(q_out, P.q_lo_cut, P.q_hi_cut, gq);
    if (w < 0) return -1;
    if (!w) return st;
    b.io = guarded_index((q_out - P.q_base) * P.inv_dq, GUARD_IDX);
    if (b.io < -1) return -1;
    if (b.io >= P.qnpts) return st;
    st |= ST_OUT;
    double q_side = b.q_y * cK - b.q_x * sK;
    if (neg) q_side = -q_side;
    w = guarded_window(q_side, P.q_lo_cut, P.q_hi_cut, gq);
    if (w < 0) return -1;
    if (!w) return st;
    b.is = guarded_index((q_side - P.q_base) * P.inv_dq, GUARD_IDX);
    if (b.is < -1) return -1;
    if (b.is >= P.qnpts) return st;
    st |= ST_SIDE;
    double q_long = b.q_z;
    double gl = GUARD_REL * (fabs(b.q_z) + 1e-3);
    if (P.boost) {
        const double rM = rsqrt_refined(K_E * K_E - K_z * K_z);
        const double gamma = K_E * rM;
        const double beta = K_z * rcp_refined(K_E);
        q_long = gamma * (b.q_z - beta * b.q_E);
        gl = GUARD_REL * gamma * (fabs(b.q_z) + fabs(b.q_E) + 1e-3);
    }
    if (neg) q_long = -q_long;
    w = guarded_window(q_long, P.q_lo_cut, P.q_hi_cut, gl);
    if (w < 0) return -1;
    if (!w) return st;
    b.il = guarded_index((q_long - P.q_base) * P.inv_dq, GUARD_IDX);
    if (b.il < -1) return -1;
    if (b.il >= P.qnpts) return st;
    b.kphi = 0;
    if (EXT && P.azimuthal) {
        b.kphi = kphi_index_fast(P, K_x, K_y, psi_ref);
        if (b.kphi < 0) return -1;
    }
    b.q_out = q_out;
    b.q_side = q_side;
    b.q_long = q_long;
    return st | ST_BIN;
}

// Same-event pair in the precision form (see eval_same_f64 for the
// contract).  The pair is (a, b) in the reference's orientation, or (b, a)
// when neg is set (see cascade_fast).  The q_inv variant and ambiguous
// pairs use the literal FP64 cascade, in the reference's orientation.
template <bool EXT = true>
__device__ __forceinline__ int eval_same_fast(const DevParams &P, const DevHist &H, double psi,
                                              const PartRec &a, const PartRec &b2, bool neg,
                                              SameRec &rec) {
    PairBins b;
    const int st = (EXT && P.qinv)
                       ? -1
                       : cascade_fast<false, EXT>(P, a.px, a.py, a.pz, a.E, b2.px, b2.py,
                                                  b2.pz, b2.E, psi, b, neg);
    if (st < 0) {
        const PartRec &p1 = neg ? b2 : a, &p2 = neg ? a : b2;
        return eval_same_f64<EXT>(P, H, psi, p1.px, p1.py, p1.pz, p1.E, p1.x, p1.y, p1.z, p1.t,
                                  p2.px, p2.py, p2.pz, p2.E, p2.x, p2.y, p2.z, p2.t, rec);
    }
    if (!(st & ST_BIN)) return st;
    const PartRec &p1 = a, &p2 = b2;
    const double t_diff = p1.t - p2.t;
    const double x_diff = p1.x - p2.x;
    const double y_diff = p1.y - p2.y;
    const double z_diff = p1.z - p2.z;
    const double phase =
        P.hbarC_inv * (b.q_E * t_diff - b.q_x * x_diff - b.q_y * y_diff - b.q_z * z_diff);
    rec.k = bin_index(P, b.kt, b.kphi, b.io, b.is, b.il);
    rec.s = (b.kt * P.nphi + b.kphi) * P.qnpts + b.io;
    rec.is = b.is;
    rec.il = b.il;
    rec.qo = b.q_out;
    rec.qs = b.q_side;
    rec.ql = b.q_long;
    rec.c = cos_reduced_f32(P, phase);
    return st;
}

// Mixed pair in the precision form; p2x, p2y are the partner's momentum
// already rotated as the reference does (HBT_correlation.cpp:522-523).
template <bool EXT = true>
__device__ __forceinline__ int eval_mixed_fast(const DevParams &P, const DevHist &H, double psi,
                                               double p1x, double p1y, double p1z, double p1E,
                                               double p2x, double p2y, double p2z, double p2E,
                                               long &k) {
    PairBins b;
    // mixed windows are half-open: v < hi (HBT_correlation.cpp:613); the
    // guard band around the cut sends v == hi to the literal cascade
    const int st = (EXT && P.qinv)
                       ? -1
                       : cascade_fast<true, EXT>(P, p1x, p1y, p1z, p1E, p2x, p2y, p2z, p2E, psi,
                                                 b);
    if (st < 0)
        return eval_mixed_f64<EXT>(P, H, psi, p1x, p1y, p1z, p1E, p2x, p2y, p2z, p2E, k);
    if (st & ST_BIN) k = bin_index(P, b.kt, b.kphi, b.io, b.is, b.il);
    return st;
}

}  // namespace hbt
