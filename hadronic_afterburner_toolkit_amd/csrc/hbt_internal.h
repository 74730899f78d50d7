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
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/hbt_hip.h"
#include "hbt_device.h"
#include "hbt_fast.h"
#include "host_pool.h"
#include "host_sort.h"
#include "host_items.h"
#include "host_tiles.h"

namespace hbt {

constexpr double kHbarC = 0.197327053;  // parameters.h:4

// Everything the kernels and the host need from an hbt_config.
struct Derived {
    DevParams P;
    double mid = 0.0, H = 0.0, h2 = 0.0, c_beta = 0.0;  // prefilter constants
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
    if (k.binning < 0 || k.binning > 2)
        return err = "binning must be 0 (atomics), 1 (slab) or 2 (chosen by hbt_autotune)",
               HBT_ERR_ARG;
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
    P.inv_dKphi = 1.0 / P.dKphi;
    P.kphi_guard = (k.kphi_guard_prad > 0 ? k.kphi_guard_prad : 1) * 1e-12;
    P.two_pi_hi = 6.28318530717958623199592693709e+00;  // 2*pi rounded to double
    P.two_pi_lo = 2.44929359829470635445e-16;            // 2*pi - two_pi_hi
    // Prefilter (hbt_fast.h).  T = 4 K_perp^2 must lie in
    // [4 KT_min_sq, 4 KT_max_sq]; q_out^2, q_side^2 <= H^2.  Margins of 1e-9
    // relative (+1e-12 absolute) dominate the ~1e-15 rounding of the
    // reference's values; prefilter_margins adds those of the prefilter's
    // own arithmetic, which scale with the data.
    const double t_lo = 4.0 * P.KT_min_sq, t_hi = 4.0 * P.KT_max_sq;
    d.mid = 0.5 * (t_lo + t_hi);
    P.pf_mid = d.mid;
    P.pf_half = 0.5 * (t_hi - t_lo) * (1.0 + 1e-9) + 1e-9 * t_hi + 1e-12;
    d.H = std::max(std::fabs(P.q_lo_cut), std::fabs(P.q_hi_cut));
    const double Hm = d.H * (1.0 + 1e-9) + 1e-12;
    d.h2 = Hm * Hm * (1.0 + 1e-9) + 1e-12;
    P.pf_h2 = P.qinv ? 0.0 : d.h2;
    P.pf_midh2 = d.mid - P.pf_h2;
    P.pf_negc = P.qinv ? -INFINITY : 0.0;
    d.c_beta = 0.5 * Hm;
    P.f_mid = (float)d.mid;
    P.f_half = (float)P.pf_half;
    P.f_h2 = (float)P.pf_h2;
    P.f_midh2 = (float)P.pf_midh2;
    P.f_negc = (float)P.pf_negc;
    d.needed = (uint64_t)(unsigned long long)k.needed_number_of_pairs;
    d.needed_qinv = (uint64_t)50 * d.needed;
    const long q = k.qnpts;
    d.nbins = (long)k.n_KT * P.nphi * q * q * q;
    d.nq = (long)k.n_KT * q;
    return HBT_OK;
}

// Prefilter record of one particle: px, py, nrm = |pT|^2 - mid/2, and its
// beta interval [blo, bhi] = beta -/+ cw with cw = (H/2) / E (or H/2
// without the LCMS boost; infinite with the q_inv branch on): the beta test
// |beta1 - beta2| <= cw1 + cw2 holds exactly when blo1 <= bhi2 and
// blo2 <= bhi1.  The ends are rounded outward (blo down, bhi up), so the
// two comparisons pass every pair the real-valued test passes.  FP32
// records: cw also covers the rounding of beta (|beta| <= bm over the list)
// and of cw itself.
inline void pf_record(const Derived &d, double px, double py, double pz, double E, double bm,
                      PfRec64 &r, double &bhi) {
    const DevParams &P = d.P;
    const double beta = P.boost ? pz / E : pz;
    const double cw = P.qinv ? INFINITY : (P.boost ? d.c_beta / E : d.c_beta);
    r = PfRec64{px, py, px * px + py * py - 0.5 * d.mid, std::nextafter(beta - cw, -INFINITY)};
    bhi = std::nextafter(beta + cw, INFINITY);
}

inline void pf_record(const Derived &d, double px, double py, double pz, double E, double bm,
                      PfRec32 &r, float &bhi) {
    const double eps32 = 5.9604644775390625e-8;
    const DevParams &P = d.P;
    const double beta = P.boost ? pz / E : pz;
    const double cw = P.qinv ? INFINITY
                             : (P.boost ? d.c_beta / E : d.c_beta) * (1.0 + 8 * eps32) +
                                   4 * eps32 * std::max(1.0, bm);
    // to FP32 rounded to nearest, then one FP32 ulp outward
    r = PfRec32{(float)px, (float)py, (float)(px * px + py * py - 0.5 * d.mid),
                std::nextafter((float)(beta - cw), -INFINITY)};
    bhi = std::nextafter((float)(beta + cw), INFINITY);
}

// Prefilter margins that depend on the data's momentum scale (|pT|^2 up to
// pmax2), for the records' precision (FP32 in the precision form).  With
// S = max(1, pmax2) + |mid|: the rounding of the records (px, py, nrm) and
// of u, v, dn, T, Q - H^2 stays below 8 ulp of S, so that of the products
// dn^2, (Q - H^2) T and H^2 T below 150 ulp of max(1, H^2) S^2 (|dn|, T,
// Q <= 4 S); the T window allows 64 ulp of 2 S, each q window a margin c
// of 256 ulp of max(1, H^2) S^2.
inline void prefilter_margins(const Derived &d, double pmax2, DevParams &P) {
    const double pm = std::max(1.0, pmax2);
    const double eps32 = 5.9604644775390625e-8, eps64 = 1.1102230246251565e-16;
    const double m = 64.0 * eps32 * (2.0 * pm + std::fabs(d.mid));
    P.f_half = (float)(d.P.pf_half + m);
    if (P.qinv) return;  // pf_negc = -inf: the q windows pass every pair
    const double S = pm + std::fabs(d.mid);
    const double s2 = std::max(1.0, d.h2) * S * S;
    P.pf_negc = -256.0 * eps64 * s2;
    P.f_negc = (float)(-256.0 * eps32 * s2);
}

// Allocator whose construct() without arguments default-initializes: a
// resize() of plain data leaves the memory untouched, so the pages are
// first written (and faulted in) by the parallel loops that fill them
// rather than by one thread zeroing hundreds of MB.
template <typename T, typename A = std::allocator<T>>
struct default_init_allocator : A {
    using traits = std::allocator_traits<A>;
    template <typename U>
    struct rebind {
        using other = default_init_allocator<U, typename traits::template rebind_alloc<U>>;
    };
    using A::A;
    default_init_allocator() = default;
    template <typename U, typename B>
    default_init_allocator(const default_init_allocator<U, B> &o) : A(o) {}
    template <typename U>
    void construct(U *p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void *>(p)) U;
    }
    template <typename U, typename... Args>
    void construct(U *p, Args &&...args) {
        traits::construct(static_cast<A &>(*this), p, std::forward<Args>(args)...);
    }
};
template <typename T>
using hvec = std::vector<T, default_init_allocator<T>>;

// Fast-kernel records of one ordering of a particle list: the 64-byte
// particle records and the prefilter records of one precision (FP32 for
// the precision form, FP64 otherwise; the other pair stays empty).
struct RecordSet {
    hvec<PartRec> rec;
    hvec<PfRec64> pf64;
    hvec<PfRec32> pf32;
    hvec<double> cw64;
    hvec<float> cw32;
    void resize_prefilter(size_t n, bool f32) {
        pf64.resize(f32 ? 0 : n);
        cw64.resize(f32 ? 0 : n);
        pf32.resize(f32 ? n : 0);
        cw32.resize(f32 ? n : 0);
    }
    void resize(size_t n, bool f32) {
        rec.resize(n);
        resize_prefilter(n, f32);
    }
    void copy_from(const RecordSet &o, size_t dst, size_t src, bool f32) {  // prefilter records
        if (f32) {
            pf32[dst] = o.pf32[src];
            cw32[dst] = o.cw32[src];
        } else {
            pf64[dst] = o.pf64[src];
            cw64[dst] = o.cw64[src];
        }
    }
};

// Particles of one list after the rapidity cut, in the reference's gather
// order (keep, off: the plain kernels and the order of the pair cap), and
// the fast kernel's lists:
//  - fast: per event, the particles whose |pT| a binnable pair allows
//    (|p1T| <= K_T + |q_T|/2 <= KT_max + D/2, >= KT_min - D/2 with
//    D = sqrt(2) H, as |q_out|, |q_side| <= H), sorted by |pT|, so a mixed
//    tile's partners are limited to a |pT| window (|pT1 - pT2| <= D);
//  - srt: the fast list of each group sorted by px cell (cells of width
//    srt_cell_width(D) from the group's smallest px), then by py, so a
//    same-event tile's j's lie in its own cell and the next one (|px1 -
//    px2| <= D) and in a py band (|py1 - py2| <= D).
// With the q_inv branch on (it bins pairs outside the 3D windows) neither
// cut applies and D is infinite.
struct HostParticles {
    long long n = 0, n_in = 0;
    hvec<long long> keep;         // index into the input of each kept particle
    std::vector<long long> off;   // nev+1 kept-particle offsets
    std::vector<double> psi2;     // psi_ref (n=2) per group, over all input particles
    long long nf = 0;             // fast-list particles
    std::vector<long long> foff;  // nev+1 fast-list offsets
    RecordSet fast;               // fast list: records
    hvec<double> f_pt;            // |pT| of the fast list (mixed windows)
    RecordSet srt;                // (px cell, py)-sorted per group (prefilter records only)
    hvec<OrigRank> orr;           // srt position -> fast-list index, gather rank
    hvec<int> srt_cell;           // px cell in srt order (same-event windows)
    hvec<double> srt_py;          // py in srt order
    double pmax2 = 0.0;  // max px^2+py^2 over the fast list (FP32 prefilter margins)
    // Empties the lists for the next gather and keeps their storage, so a
    // caller that reuses one (the request server's pool) writes into pages
    // already faulted in: fresh buffers cost a third of a gather in page
    // faults.
    void reset() {
        n = n_in = nf = 0;
        pmax2 = 0.0;
    }
};

// Half-width D of the px and |pT| windows of a binnable pair: a pair inside
// the q_out and q_side windows has |q_T| <= sqrt(2) H (margins as the
// prefilter's).  Infinite when the q_inv branch is on.
inline double px_window(const Derived &d) {
    if (d.P.qinv) return INFINITY;
    const double Hm = d.H * (1.0 + 1e-9) + 1e-12;
    return std::sqrt(2.0) * Hm * (1.0 + 1e-9) + 1e-12;
}

// Half-width of the mixed partners' |p_T| window: a binnable pair has
// ||p1T| - |p2T|| <= |q_out| <= H, since (|p1T| - |p2T|)(|p1T| + |p2T|) =
// (p1T - p2T).(p1T + p2T) = 2 |K_T| q_out and |p1T| + |p2T| >= 2 |K_T|
// (margins as the prefilter's).  Infinite when the q_inv branch is on.
inline double pt_window(const Derived &d) {
    if (d.P.qinv) return INFINITY;
    return (d.H * (1.0 + 1e-9) + 1e-12) * (1.0 + 1e-9) + 1e-12;
}

// Host-to-device copies through a pinned staging buffer: a pageable copy
// runs at a fraction of the link rate and cannot overlap host work.
struct StagePiece {
    void *dst;        // device destination
    const void *src;  // host source
    size_t bytes;
};

inline size_t stage_round(size_t b) { return (b + 255) / 256 * 256; }

inline size_t stage_bytes(const std::vector<StagePiece> &v) {
    size_t t = 0;
    for (const StagePiece &x : v) t += stage_round(x.bytes);
    return t;
}

// Grows a pinned buffer to at least `bytes` (by a quarter at least). The
// caller has waited for every copy out of the old one.
inline hipError_t ensure_pinned(char **buf, size_t *cap, size_t bytes) {
    if (bytes <= *cap) return hipSuccess;
    const size_t want = std::max(bytes, *cap + *cap / 4);
    if (*buf) {
        const hipError_t e = hipHostFree(*buf);
        if (e != hipSuccess) return e;
    }
    *buf = nullptr;
    *cap = 0;
    const hipError_t e = hipHostMalloc((void **)buf, want, hipHostMallocDefault);
    if (e == hipSuccess) *cap = want;
    return e;
}

// Copies the pieces into stage (back to back, 256-byte aligned) in 1 MiB
// chunks on up to 16 threads, then queues their copies to the device on st.
inline hipError_t stage_pieces(char *stage, const std::vector<StagePiece> &v, hipStream_t st) {
    const size_t chunk = 1 << 20;
    std::vector<size_t> off(v.size());
    std::vector<std::pair<size_t, size_t>> parts;  // (piece, first byte)
    size_t t = 0;
    for (size_t q = 0; q < v.size(); q++) {
        off[q] = t;
        t += stage_round(v[q].bytes);
        for (size_t b = 0; b < v[q].bytes; b += chunk) parts.emplace_back(q, b);
    }
    parallel_ranges((long long)parts.size(), 4, [&](long long p0, long long p1) {
        for (long long p = p0; p < p1; p++) {
            const StagePiece &x = v[parts[p].first];
            const size_t b = parts[p].second, len = std::min(chunk, x.bytes - b);
            std::memcpy(stage + off[parts[p].first] + b, static_cast<const char *>(x.src) + b, len);
        }
    });
    for (size_t q = 0; q < v.size(); q++) {
        if (!v[q].bytes) continue;
        const hipError_t e = hipMemcpyAsync(v[q].dst, stage + off[q], v[q].bytes,
                                            hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// after_keep (optional) runs once the kept particles (keep, off) and the
// list sizes (n, nf) are known, before the px sort; a non-zero return ends
// the gather with that code.
// The rapidity cut of the gathers: tanh(HBTrap_min) < pz/E < tanh(HBTrap_max)
// (HBT_correlation.cpp:255-266, :468-476; cut_* = tanh of the bounds).
inline bool rapidity_kept(double pz, double E, double cut_min, double cut_max) {
    const double ratio = pz / E;
    return ratio > cut_min && ratio < cut_max;
}

// Rapidity cut of the gathers (HBT_correlation.cpp:255-266, :468-476),
// psi_ref per group (:181-183, :233-249; over all particles passed in), and
// the prefilter records.  Events keep < 2^18 particles (the fast kernel's
// survivor encoding).
inline int gather_particles(const Derived &d, const hbt_config &k, int ngroups,
                            const int32_t *grp_ev, int nev, const int64_t *ev_off,
                            const double *const soa[8], HostParticles &h, std::string &err,
                            const std::function<int()> *after_keep = nullptr) {
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
    const bool f32 = k.precision == HBT_PREC_FP32C;
    h.reset();  // every list below is sized and written in full
    h.n_in = nev > 0 ? ev_off[nev] - ev_off[0] : 0;
    const double cut_max = tanh(k.HBTrap_max);
    const double cut_min = tanh(k.HBTrap_min);
    auto kept = [&](int64_t i) { return rapidity_kept(soa[2][i], soa[3][i], cut_min, cut_max); };
    // the fast lists' |pT| range: particles a binnable pair can contain
    const double D = px_window(d);
    const double r_lo = std::isfinite(D) ? (P.KT_min - 0.5 * D) * (1.0 - 1e-9) : -INFINITY;
    const double r_hi =
        std::isfinite(D) ? (std::sqrt(P.KT_max_sq) + 0.5 * D) * (1.0 + 1e-9) : INFINITY;
    auto pt_of = [&](int64_t i) { return std::sqrt(soa[0][i] * soa[0][i] + soa[1][i] * soa[1][i]); };
    auto fast = [&](double pt) { return pt >= r_lo && pt <= r_hi; };
    auto beta_of = [&](int64_t i) { return P.boost ? soa[2][i] / soa[3][i] : soa[2][i]; };
    // pass 1, per event: kept and fast-list counts, and the fast list's
    // momentum and beta scales (FP32 margins)
    h.off.assign(nev + 1, 0);
    h.foff.assign(nev + 1, 0);
    std::vector<double> pmax((size_t)nev + 1, 0.0), bmax((size_t)nev + 1, 0.0);
    // events per host thread: about 2048 particles' worth at least
    const long long ev_grain = std::max(1ll, 2048ll * nev / std::max(1ll, h.n_in));
    parallel_ranges(nev, ev_grain, [&](long long e0, long long e1) {
        for (long long e = e0; e < e1; e++) {
            long long c = 0, cf = 0;
            for (int64_t i = ev_off[e]; i < ev_off[e + 1]; i++) {
                if (!kept(i)) continue;
                c++;
                const double pt = pt_of(i);
                if (!fast(pt)) continue;
                cf++;
                pmax[e] = std::max(pmax[e], soa[0][i] * soa[0][i] + soa[1][i] * soa[1][i]);
                bmax[e] = std::max(bmax[e], std::fabs(beta_of(i)));
            }
            h.off[e + 1] = c;
            h.foff[e + 1] = cf;
        }
    });
    double bm = 0.0;
    for (int e = 0; e < nev; e++) {
        if (h.off[e + 1] >= (1 << 18))
            return err = "an event keeps more than 262143 particles", HBT_ERR_CAPACITY;
        h.off[e + 1] += h.off[e];
        h.foff[e + 1] += h.foff[e];
        h.pmax2 = std::max(h.pmax2, pmax[e]);
        bm = std::max(bm, bmax[e]);
    }
    const long long n = h.off[nev], nf = h.foff[nev];
    if (n >= (1ll << 31)) return err = "buffer too large", HBT_ERR_CAPACITY;
    h.n = n;
    h.nf = nf;
    h.keep.resize((size_t)n);
    RecordSet &F = h.fast;
    F.resize((size_t)nf, f32);
    h.f_pt.resize((size_t)nf);
    hvec<int> frank((size_t)nf);
    // pass 2, per event: gather order, then the fast list sorted by |pT|
    // (stable, so equal |pT| keep gather order)
    parallel_ranges(nev, ev_grain, [&](long long e0, long long e1) {
        std::vector<std::pair<double, long long>> ev;  // |pT|, kept index
        std::vector<std::pair<uint64_t, long long>> ra, rb;  // radix scratch
        for (long long e = e0; e < e1; e++) {
            ev.clear();
            long long q = h.off[e];
            for (int64_t i = ev_off[e]; i < ev_off[e + 1]; i++) {
                if (!kept(i)) continue;
                h.keep[q] = i;
                const double pt = pt_of(i);
                if (fast(pt)) ev.emplace_back(pt, q);
                q++;
            }
            // stable by |pT| (gather order among equal |pT|)
            radix_sort_keys(ev, ra, rb);
            long long f = h.foff[e];
            for (const auto &x : ev) {
                const int64_t i = h.keep[x.second];
                const double px = soa[0][i], py = soa[1][i], pz = soa[2][i], E = soa[3][i];
                F.rec[f] = PartRec{px, py, pz, E, soa[4][i], soa[5][i], soa[6][i], soa[7][i]};
                if (f32)
                    pf_record(d, px, py, pz, E, bm, F.pf32[f], F.cw32[f]);
                else
                    pf_record(d, px, py, pz, E, bm, F.pf64[f], F.cw64[f]);
                h.f_pt[f] = x.first;
                frank[f] = (int)x.second;
                f++;
            }
        }
    });
    if (after_keep) {
        const int rc = (*after_keep)();
        if (rc) return rc;
    }
    // psi_ref per group over all particles passed in, in order (a sum per group)
    h.psi2.assign(ngroups, 0.0);
    if (P.azimuthal) {
        parallel_ranges(ngroups, 1, [&](long long g0, long long g1) {
            for (long long g = g0; g < g1; g++) {
                double vr = 0.0, vi = 0.0;
                for (int64_t i = ev_off[grp_ev[g]]; i < ev_off[grp_ev[g + 1]]; i++) {
                    const double phi = atan2(soa[1][i], soa[0][i]);
                    double sn, cn;
                    sincos(2 * phi, &sn, &cn);  // as the reference's build evaluates it
                    vr += cn;
                    vi += sn;
                }
                h.psi2[g] = atan2(vi, vr) / 2;
            }
        });
    }
    // prefilter records per group sorted by px cell, then py (stable)
    h.srt.resize_prefilter((size_t)nf, f32);
    h.orr.resize((size_t)nf);
    h.srt_cell.resize((size_t)nf);
    h.srt_py.resize((size_t)nf);
    // groups on the host threads, or, with fewer groups than threads (one
    // buffer of the per-buffer path), each group's cells and copies
    const bool lone = ngroups < HostPool::get().threads();
    auto sort_group = [&](long long g, CellOrderScratch &sc, std::vector<int> &order) {
        const long long g0 = h.foff[grp_ev[g]], g1 = h.foff[grp_ev[g + 1]];
        order.resize((size_t)(g1 - g0));
        cell_order((int)(g1 - g0), [&](int k) { return F.rec[g0 + k].px; },
                   [&](int k) { return F.rec[g0 + k].py; }, D, sc, order.data(),
                   h.srt_cell.data() + g0, lone);
        auto copy = [&](long long q0, long long q1) {
            for (long long q = g0 + q0; q < g0 + q1; q++) {
                const int src = (int)g0 + order[(size_t)(q - g0)];
                h.srt.copy_from(F, (size_t)q, (size_t)src, f32);
                h.orr[q] = OrigRank{src, frank[src]};
                h.srt_py[q] = F.rec[src].py;
            }
        };
        if (lone)
            parallel_ranges(g1 - g0, 2048, copy);
        else
            copy(0, g1 - g0);
    };
    if (lone) {
        CellOrderScratch sc;
        std::vector<int> order;
        for (long long g = 0; g < ngroups; g++) sort_group(g, sc, order);
    } else {
        parallel_ranges(ngroups, 1, [&](long long ga, long long gb) {
            CellOrderScratch sc;
            std::vector<int> order;
            for (long long g = ga; g < gb; g++) sort_group(g, sc, order);
        });
    }
    return HBT_OK;
}

// Work items per resident wave below which a launch's items are made smaller.
constexpr size_t kItemsPerWave = 4;
// L2 of one XCD (MI355X: 4 MB).
constexpr size_t kL2BytesPerXcd = (size_t)4 << 20;

inline ItemShape item_shape(const hbt_tuning &t) {
    return ItemShape{64 * t.R, std::max(64, t.jchunk), std::max(1, std::min(t.kchunk, 64))};
}

// Pairs left to the host by the device (DeferRec): room per context or
// request, and one bin addition each.
constexpr unsigned kDeferCap = 1u << 16;     // per context
constexpr unsigned kReqDeferCap = 1u << 12;  // per server request

struct DeferAdd {
    double *bins;  // the histogram (device address)
    long long k;   // bin
    int kind;      // 0: count + q_out, q_side, q_long, cos sums; 1: denominator
    int pad;
    double v[5];
};

// The host's decision for deferred pairs: the same FP64 cascade, whose host
// build takes K_phi from glibc's atan2 as the reference does
// (HBT_correlation.cpp:408-423), and the cos phase of a same-event pair
// (:431-436).
inline void resolve_records(const DevParams &P, const DeferRec *rec, unsigned n,
                            std::vector<DeferAdd> &adds) {
    for (unsigned i = 0; i < n; i++) {
        const DeferRec &r = rec[i];
        PairBins b;
        double q_inv;
        int qinv_ok;
        if (r.kind == 0) {
            const int st = cascade_f64<false, true>(P, r.p1[0], r.p1[1], r.p1[2], r.p1[3], r.p2[0],
                                                    r.p2[1], r.p2[2], r.p2[3], r.psi, b, q_inv,
                                                    qinv_ok);
            if (!(st & ST_BIN)) continue;
            const double t_diff = r.p1[7] - r.p2[7];
            const double x_diff = r.p1[4] - r.p2[4];
            const double y_diff = r.p1[5] - r.p2[5];
            const double z_diff = r.p1[6] - r.p2[6];
            const double cq = std::cos(P.hbarC_inv * (b.q_E * t_diff - b.q_x * x_diff -
                                                      b.q_y * y_diff - b.q_z * z_diff));
            adds.push_back(DeferAdd{r.bins, bin_index(P, b.kt, b.kphi, b.io, b.is, b.il), 0, 0,
                                    {1.0, b.q_out, b.q_side, b.q_long, cq}});
        } else {
            const int st = cascade_f64<true, true>(P, r.p1[0], r.p1[1], r.p1[2], r.p1[3], r.p2[0],
                                                   r.p2[1], r.p2[2], r.p2[3], r.psi, b, q_inv,
                                                   qinv_ok);
            if (!(st & ST_BIN)) continue;
            adds.push_back(DeferAdd{r.bins, bin_index(P, b.kt, b.kphi, b.io, b.is, b.il), 1, 0,
                                    {0, 0, 0, 0, 0}});
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
        unsigned long long den;
        const unsigned long long cnt = (unsigned long long)x[BIN_COUNT];
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
