// hbt_internal.h -- host-side pieces shared by the per-buffer engine
// (hbt_engine.hip) and the request server (hbt_server.hip): derived
// constants, the particle gather with the rapidity cut, work-item
// construction, result assembly, and the launch of the fast kernel (defined
// in hbt_engine.hip, where its instances are compiled).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hbt_hip.h"
#include "hbt_device.h"
#include "hbt_fast.h"

namespace hbt {

constexpr double kHbarC = 0.197327053;  // parameters.h:4

// Everything the kernels and the host need from an hbt_config.
struct Derived {
    DevParams P;
    double mid = 0.0, H = 0.0, c_beta = 0.0;  // prefilter constants
    uint64_t needed = 0, needed_qinv = 0;     // pair caps
    long nbins = 0, nq = 0;                   // histogram sizes
};

// Validates the configuration and derives the grid constants
// (HBT_correlation.cpp:25-61 and the loop constants of :253-256, :289-290,
// :363-369, each evaluated in the reference's expression order).
inline int derive(const hbt_config &k, Derived &d, std::string &err) {
    if (k.qnpts < 2) return err = "qnpts must be >= 2", HBT_ERR_ARG;
    if (k.n_KT < 2) return err = "n_KT (number of K_T edges) must be >= 2", HBT_ERR_ARG;
    if (k.azimuthal_flag && k.n_Kphi < 1) return err = "n_Kphi must be >= 1", HBT_ERR_ARG;
    if (!(k.q_max > k.q_min)) return err = "q_max must exceed q_min", HBT_ERR_ARG;
    if (!(k.KT_max > k.KT_min)) return err = "KT_max must exceed KT_min", HBT_ERR_ARG;
    if (k.precision != HBT_PREC_FP64 && k.precision != HBT_PREC_FP32C)
        return err = "unknown precision", HBT_ERR_ARG;
    if (k.kernel != HBT_KERNEL_PLAIN && k.kernel != HBT_KERNEL_FAST)
        return err = "unknown kernel", HBT_ERR_ARG;
    if (k.precision == HBT_PREC_FP32C && k.kernel != HBT_KERNEL_FAST)
        return err = "the precision form runs on the fast kernel only", HBT_ERR_ARG;
    DevParams &P = d.P;
    std::memset(&P, 0, sizeof(P));
    P.qnpts = k.qnpts;
    P.n_KT = k.n_KT;
    P.n_Kphi = k.n_Kphi;
    P.nphi = k.azimuthal_flag ? k.n_Kphi : 1;
    P.azimuthal = k.azimuthal_flag == 1;
    P.qinv = k.invariant_radius_flag == 1;
    P.boost = k.long_comoving_boost == 1;
    P.delta_q = (k.q_max - k.q_min) / (k.qnpts - 1);
    P.KT_min = k.KT_min;
    P.dKT = (k.KT_max - k.KT_min) / (k.n_KT - 1);
    P.dKphi = 2 * M_PI / (k.n_Kphi > 0 ? k.n_Kphi : 1);
    P.q_base = k.q_min - P.delta_q / 2.;
    P.q_lo_cut = k.q_min - P.delta_q / 2. + 1e-8;
    P.q_hi_cut = k.q_max + P.delta_q / 2. - 1e-8;
    P.KT_min_sq = k.KT_min * k.KT_min;
    P.KT_max_sq = k.KT_max * k.KT_max;
    P.hbarC_inv = 1. / kHbarC;
    P.two_pi = 2. * M_PI;
    P.inv_dKT = 1.0 / P.dKT;
    P.inv_dq = 1.0 / P.delta_q;
    P.inv_two_pi = 1.0 / (2.0 * M_PI);
    P.two_pi_hi = 6.28318530717958623199592693709e+00;  // 2*pi rounded to double
    P.two_pi_lo = 2.44929359829470635445e-16;            // 2*pi - two_pi_hi
    // Prefilter (hbt_fast.h).  T = 4 K_perp^2 must lie in
    // [4 KT_min_sq, 4 KT_max_sq]; Q <= 2 H^2.  Margins of 1e-9 relative
    // (+1e-12 absolute) dominate the ~1e-15 rounding of either side.
    const double t_lo = 4.0 * P.KT_min_sq, t_hi = 4.0 * P.KT_max_sq;
    d.mid = 0.5 * (t_lo + t_hi);
    P.pf_mid = d.mid;
    P.pf_half = 0.5 * (t_hi - t_lo) * (1.0 + 1e-9) + 1e-9 * t_hi + 1e-12;
    d.H = std::max(std::fabs(P.q_lo_cut), std::fabs(P.q_hi_cut));
    const double Hm = d.H * (1.0 + 1e-9) + 1e-12;
    P.pf_qthr = P.qinv ? INFINITY : (2.0 * Hm * Hm * (1.0 + 1e-9) + 1e-12 - d.mid);
    d.c_beta = 0.5 * Hm;
    P.f_mid = (float)d.mid;
    P.f_half = (float)P.pf_half;
    P.f_qthr = P.qinv ? INFINITY : (float)P.pf_qthr;
    d.needed = (uint64_t)(unsigned long long)k.needed_number_of_pairs;
    d.needed_qinv = (uint64_t)50 * d.needed;
    const long q = k.qnpts;
    d.nbins = (long)k.n_KT * P.nphi * q * q * q;
    d.nq = (long)k.n_KT * q;
    return HBT_OK;
}

// Prefilter margins for the FP32 records depend on the data's momentum
// scale (|pT|^2 up to pmax2): FP32 rounding of px, py and of the dot
// product stays below 8 ulp of that scale; 64 ulp are allowed.
inline void fp32_margins(const Derived &d, double pmax2, DevParams &P) {
    const double pm = std::max(1.0, pmax2);
    const double eps32 = 5.9604644775390625e-8;
    const double m = 64.0 * eps32 * (2.0 * pm + std::fabs(d.mid));
    P.f_half = (float)(d.P.pf_half + m);
    P.f_qthr = std::isfinite(d.P.pf_qthr) ? (float)(d.P.pf_qthr + m) : INFINITY;
}

// Fast-kernel records of one ordering of a particle list.
struct RecordSet {
    std::vector<PartRec> rec;
    std::vector<PfRec64> pf64;
    std::vector<PfRec32> pf32;
    std::vector<double> cw64;
    std::vector<float> cw32;
    void resize(size_t n) {
        rec.resize(n);
        pf64.resize(n);
        pf32.resize(n);
        cw64.resize(n);
        cw32.resize(n);
    }
    void copy_from(const RecordSet &o, size_t dst, size_t src) {
        rec[dst] = o.rec[src];
        pf64[dst] = o.pf64[src];
        pf32[dst] = o.pf32[src];
        cw64[dst] = o.cw64[src];
        cw32[dst] = o.cw32[src];
    }
};

// Particles of one list after the rapidity cut, in the reference's gather
// order (ev), and the same particles sorted by px inside each group (srt)
// for the same-event tiles, whose j range is then limited to the px window
// a binned pair can have: |px1 - px2| <= |p1T - p2T| <= sqrt(2) H.
struct HostParticles {
    long long n = 0, n_in = 0;
    std::vector<long long> keep;  // index into the input of each kept particle
    std::vector<long long> off;   // nev+1 kept-particle offsets
    std::vector<double> psi2;     // psi_ref (n=2) per group, over all input particles
    RecordSet ev;                 // gather order
    RecordSet srt;                // px-sorted inside each group
    std::vector<int> orig;        // srt position -> gather-order index
    std::vector<double> srt_px;   // px in srt order (work-item windows)
    double pmax2 = 0.0;  // max px^2+py^2 of kept particles
};

// Rapidity cut of the gathers (HBT_correlation.cpp:255-266, :468-476),
// psi_ref per group (:181-183, :233-249; over all particles passed in), and
// the prefilter records.  Events keep < 2^18 particles (the fast kernel's
// survivor encoding).
inline int gather_particles(const Derived &d, const hbt_config &k, int ngroups,
                            const int32_t *grp_ev, int nev, const int64_t *ev_off,
                            const double *const soa[8], HostParticles &h, std::string &err) {
    if (nev < 0 || ngroups < 0) return err = "negative size", HBT_ERR_ARG;
    if (nev > 0 && (!ev_off || !soa)) return err = "null particle arrays", HBT_ERR_ARG;
    for (int a = 0; a < 8 && nev > 0; a++)
        if (!soa[a]) return err = "null particle column", HBT_ERR_ARG;
    for (int e = 0; e < nev; e++)
        if (ev_off[e + 1] < ev_off[e]) return err = "event offsets decrease", HBT_ERR_ARG;
    if (grp_ev[0] != 0 || grp_ev[ngroups] != nev)
        return err = "group offsets must run from 0 to nev", HBT_ERR_ARG;
    for (int g = 0; g < ngroups; g++)
        if (grp_ev[g + 1] < grp_ev[g]) return err = "group offsets decrease", HBT_ERR_ARG;
    const DevParams &P = d.P;
    h = HostParticles();
    h.n_in = nev > 0 ? ev_off[nev] - ev_off[0] : 0;
    const double cut_max = tanh(k.HBTrap_max);
    const double cut_min = tanh(k.HBTrap_min);
    h.keep.reserve((size_t)h.n_in);
    h.off.assign(nev + 1, 0);
    h.psi2.assign(ngroups, 0.0);
    for (int g = 0; g < ngroups; g++) {
        double vr = 0.0, vi = 0.0;
        for (int e = grp_ev[g]; e < grp_ev[g + 1]; e++) {
            for (int64_t i = ev_off[e]; i < ev_off[e + 1]; i++) {
                const double ratio = soa[2][i] / soa[3][i];
                if (ratio > cut_min && ratio < cut_max) h.keep.push_back(i);
                if (P.azimuthal) {
                    const double phi = atan2(soa[1][i], soa[0][i]);
                    double sn, cn;
                    sincos(2 * phi, &sn, &cn);  // as the reference's build evaluates it
                    vr += cn;
                    vi += sn;
                }
            }
            h.off[e + 1] = (long long)h.keep.size();
        }
        h.psi2[g] = P.azimuthal ? atan2(vi, vr) / 2 : 0.0;
    }
    const long long n = (long long)h.keep.size();
    for (int e = 0; e < nev; e++)
        if (h.off[e + 1] - h.off[e] >= (1 << 18))
            return err = "an event keeps more than 262143 particles", HBT_ERR_CAPACITY;
    if (n >= (1ll << 31)) return err = "buffer too large", HBT_ERR_CAPACITY;
    h.n = n;
    RecordSet &R = h.ev;
    R.resize((size_t)n);
    double bmax = 0.0;
    for (long long q = 0; q < n; q++) {
        const long long i = h.keep[q];
        const double px = soa[0][i], py = soa[1][i], pz = soa[2][i], E = soa[3][i];
        R.rec[q] = PartRec{px, py, pz, E, soa[4][i], soa[5][i], soa[6][i], soa[7][i]};
        const double n2 = px * px + py * py;
        h.pmax2 = std::max(h.pmax2, n2);
        PfRec64 &r = R.pf64[q];
        r.px = px;
        r.py = py;
        r.nrm = n2 - 0.5 * d.mid;
        r.beta = P.boost ? pz / E : pz;
        bmax = std::max(bmax, std::fabs(r.beta));
        R.cw64[q] = P.qinv ? INFINITY : (P.boost ? d.c_beta / E : d.c_beta);
    }
    const double eps32 = 5.9604644775390625e-8;
    for (long long q = 0; q < n; q++) {
        R.pf32[q].px = (float)R.pf64[q].px;
        R.pf32[q].py = (float)R.pf64[q].py;
        R.pf32[q].nrm = (float)R.pf64[q].nrm;
        R.pf32[q].beta = (float)R.pf64[q].beta;
        // covers the FP32 rounding of beta (|beta| <= bmax) and of cw itself
        R.cw32[q] = P.qinv ? INFINITY
                           : (float)(R.cw64[q] * (1.0 + 8 * eps32) + 4 * eps32 * std::max(1.0, bmax));
    }
    // px-sorted copy per group (stable, so equal px keep gather order)
    h.srt.resize((size_t)n);
    h.orig.resize((size_t)n);
    h.srt_px.resize((size_t)n);
    std::vector<int> idx;
    for (int g = 0; g < ngroups; g++) {
        const long long g0 = h.off[grp_ev[g]], g1 = h.off[grp_ev[g + 1]];
        idx.resize((size_t)(g1 - g0));
        for (long long q = g0; q < g1; q++) idx[q - g0] = (int)q;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int b) { return R.rec[a].px < R.rec[b].px; });
        for (long long q = g0; q < g1; q++) {
            const int src = idx[q - g0];
            h.srt.copy_from(R, (size_t)q, (size_t)src);
            h.orig[q] = src;
            h.srt_px[q] = R.rec[src].px;
        }
    }
    return HBT_OK;
}

// Launch shape of the fast kernel's work items.
struct ItemShape {
    int TI;  // i particles per item (64 R)
    int JC;  // same-event j chunk
    int KC;  // mixed partners per item
};

inline ItemShape item_shape(const hbt_tuning &t) {
    return ItemShape{64 * t.R, std::max(64, t.jchunk), std::max(1, std::min(t.kchunk, 64))};
}

// Half-width of the px window of a binnable same-event pair: a pair inside
// the q_out and q_side windows has |px1 - px2| <= |p1T - p2T| <= sqrt(2) H;
// margins as the prefilter's.  Infinite (no window) when the q_inv branch
// is on, since it bins pairs outside the 3D windows.
inline double px_window(const Derived &d) {
    if (d.P.qinv) return INFINITY;
    const double Hm = d.H * (1.0 + 1e-9) + 1e-12;
    return std::sqrt(2.0) * Hm * (1.0 + 1e-9) + 1e-12;
}

// Same-event items of the groups, in px-sorted order: i tile of TI
// consecutive particles x its j window (particles after the tile start whose
// px lies within dpx of the tile's px range), cut into chunks of JC.
// pairs (optional) receives each item's evaluated pair count.  enum_same
// receives the reference's enumerated pair count sum n(n-1)/2.
inline void items_same(const ItemShape &s, const std::vector<long long> &off,
                       const std::vector<int> &grp_ev, const std::vector<double> &srt_px,
                       double dpx, int view, std::vector<WorkItem> &items,
                       std::vector<uint64_t> *pairs, uint64_t &enum_same) {
    const int ngroups = (int)grp_ev.size() - 1;
    for (int g = 0; g < ngroups; g++) {
        const long long g0 = off[grp_ev[g]], n = off[grp_ev[g + 1]] - g0;
        if (n < 2) continue;
        enum_same += (uint64_t)(n * (n - 1) / 2);
        const double *px = srt_px.data() + g0;
        for (long long a = 0; a + 1 < n; a += s.TI) {
            const int ni = (int)std::min<long long>(s.TI, n - a);
            long long jhi = n;
            if (std::isfinite(dpx)) {
                // px ascending inside the group: the tile spans [px[a], px[a+ni-1]]
                const double top = px[a + ni - 1] + dpx;
                jhi = std::upper_bound(px + a, px + n, top) - px;
            }
            for (long long jb = a + 1; jb < jhi; jb += s.JC) {
                WorkItem w{};
                w.i0 = g0 + a;
                w.j0 = g0 + jb;
                w.ni = ni;
                w.nj = (int)std::min<long long>(s.JC, jhi - jb);
                w.kind = 0;
                w.grp = g;
                w.view = view;
                items.push_back(w);
                if (pairs) {
                    // pairs (i, j) with i in [a, a+ni), j in [jb, jb+nj), i < j
                    uint64_t cnt = 0;
                    const long long j1 = jb + w.nj;
                    for (long long i = a; i < a + ni; i++) {
                        const long long lo = std::max(jb, i + 1);
                        if (j1 > lo) cnt += (uint64_t)(j1 - lo);
                    }
                    pairs->push_back(cnt);
                }
            }
        }
    }
}

// Mixed items: i tile of TI particles of event e x up to KC of its plan
// rows [rowoff[e], rowoff[e+1]).  mix_off are the mixed list's event
// offsets, plan_ids the partner of each row (index into the mixed list).
inline void items_mixed(const ItemShape &s, const std::vector<long long> &off,
                        const std::vector<int> &grp_ev, const std::vector<long long> &rowoff,
                        const std::vector<int> &plan_ids, const std::vector<long long> &mix_off,
                        int view, std::vector<WorkItem> &items, std::vector<uint64_t> *pairs) {
    const int ngroups = (int)grp_ev.size() - 1;
    for (int g = 0; g < ngroups; g++) {
        for (int e = grp_ev[g]; e < grp_ev[g + 1]; e++) {
            const long long n1 = off[e + 1] - off[e];
            const int M = (int)(rowoff[e + 1] - rowoff[e]);
            for (long long a = 0; a < n1; a += s.TI) {
                for (int k0 = 0; k0 < M; k0 += s.KC) {
                    WorkItem w{};
                    w.i0 = off[e] + a;
                    w.ni = (int)std::min<long long>(s.TI, n1 - a);
                    w.kind = 1;
                    w.e = e;
                    w.k0 = k0;
                    w.k1 = std::min(M, k0 + s.KC);
                    w.grp = g;
                    w.view = view;
                    items.push_back(w);
                    if (pairs) {
                        uint64_t nj = 0;
                        for (int k = w.k0; k < w.k1; k++) {
                            const int p = plan_ids[rowoff[e] + k];
                            nj += (uint64_t)(mix_off[p + 1] - mix_off[p]);
                        }
                        pairs->push_back((uint64_t)w.ni * nj);
                    }
                }
            }
        }
    }
}

// hbt_result from the AoS bins and q_inv arrays in host memory; the pair
// counters are sums of the count/denominator fields (every counter
// increment of the reference is paired with one bin increment).
inline void result_from_host(const Derived &d, const hbt_config &k, const double *aos,
                             const unsigned long long *qc, const unsigned long long *qd,
                             const double *qs, const double *qn, hbt_result *o) {
    const DevParams &P = d.P;
    const long per = (long)P.qnpts * P.qnpts * P.qnpts;
    const int nKT = k.n_KT, nphi = P.nphi;
    std::vector<unsigned long long> sn((size_t)nKT * nphi, 0), sd((size_t)nKT * nphi, 0);
    for (long b = 0; b < d.nbins; b++) {
        const double *x = aos + b * BIN_STRIDE;
        unsigned long long cnt, den;
        std::memcpy(&cnt, x + BIN_COUNT, 8);
        std::memcpy(&den, x + BIN_DEN, 8);
        sn[b / per] += cnt;
        sd[b / per] += den;
        if (o->num_count) o->num_count[b] = cnt;
        if (o->den) o->den[b] = den;
        if (o->q_out_sum) o->q_out_sum[b] = x[BIN_QO];
        if (o->q_side_sum) o->q_side_sum[b] = x[BIN_QS];
        if (o->q_long_sum) o->q_long_sum[b] = x[BIN_QL];
        if (o->num) o->num[b] = x[BIN_NUM];
    }
    const size_t q8 = (size_t)d.nq * 8;
    if (o->qinv_count) std::memcpy(o->qinv_count, qc, q8);
    if (o->qinv_den) std::memcpy(o->qinv_den, qd, q8);
    if (o->qinv_sum) std::memcpy(o->qinv_sum, qs, q8);
    if (o->qinv_num) std::memcpy(o->qinv_num, qn, q8);
    for (int kt = 0; kt < nKT; kt++) {
        unsigned long long a = 0, b = 0, c = 0, e = 0;
        for (int p = 0; p < nphi; p++) {
            a += sn[kt * nphi + p];
            b += sd[kt * nphi + p];
        }
        for (int q = 0; q < P.qnpts; q++) {
            c += qc[kt * P.qnpts + q];
            e += qd[kt * P.qnpts + q];
        }
        // with azimuthal_flag=1 the reference leaves the K_T counters at 0
        if (o->cnt_num_kt) o->cnt_num_kt[kt] = P.azimuthal ? 0 : a;
        if (o->cnt_den_kt) o->cnt_den_kt[kt] = P.azimuthal ? 0 : b;
        if (o->cnt_num_qinv) o->cnt_num_qinv[kt] = c;
        if (o->cnt_den_qinv) o->cnt_den_qinv[kt] = e;
    }
    if (P.azimuthal) {
        for (int q = 0; q < nKT * nphi; q++) {
            if (o->cnt_num_ktphi) o->cnt_num_ktphi[q] = sn[q];
            if (o->cnt_den_ktphi) o->cnt_den_ktphi[q] = sd[q];
        }
    }
}

// Defined in hbt_engine.hip.  Zeroes the work counter and launches
// k_fast<Prec, R, EXT> with `blocks` workgroups of 256 on stream st.
hipError_t launch_fast_kernel(const FastArgs &a, bool f32, int R, bool ext, long blocks,
                              hipStream_t st);
// Zeroes n 8-byte words with a fill kernel (safe inside graph capture).
hipError_t launch_zero_words(void *p, size_t n, hipStream_t st);

}  // namespace hbt
