// hbt_bf.hip -- the balance-function pair histograms on the GPU
// (BalanceFunction.cpp:63-197): for the species a = particle_alpha,
// b = particle_beta and their antiparticles, the (Delta y, Delta phi)
// histograms of same-event pairs (a,b), (abar,bbar), (a,bbar), (abar,b) and
// of mixed pairs against one randomly chosen, randomly rotated event per
// event and histogram.
//
// Host: per-particle phi and y (or eta) with the reference's expressions
// (particleSamples::boostParticles, particleSamples.cpp:442-468), the species
// lists with the pT cut applied (the pair loops skip those particles), built
// on up to 16 threads, and the mixing draws in the reference's order.
//
// Device: work units of (histogram, event, 64 B particles, a run of A rows);
// a wave takes one unit at a time from its histogram's queue, one lane per B
// particle, the A rows read as wave-uniform scalars.  Each block keeps its
// histogram in LDS as u32 counts and adds it to the FP64 histogram in HBM
// once, at its end.  The host splits the work into launches of fewer than
// 2^32 pairs per histogram, so no LDS count can overflow.
//
// Two per-pair forms give identical counts:
//  * exact: the reference's FP64 arithmetic for every pair
//    (BalanceFunction.cpp:136-153), without contraction;
//  * fast (default): the bin coordinates in bin units as FP32 differences of
//    per-particle values, t_phi = U_a - V_b and t_y = A_a - B_b, prepared in
//    FP64 and rounded to FP32 once.  Then |t - T| < G/3 for the value T of
//    the reference's FP64 chain, with G = 2^-21 (max|U| + max|V|) + 1e-6 (the
//    same for y, plus 4e-10/drap for the |Delta y| < 1e-10 test).  A pair
//    whose t lies within G of an integer (or, for odd Bnpts, of a
//    half-integer: Delta y = 0 falls there) takes the exact form; for every
//    other pair floor(t) = floor(T), so its bins are the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/hbt_hip.h"

namespace {

constexpr int kBnphi = 20;        // BalanceFunction.cpp:33
constexpr int kMaxLds = 16384;    // u32 words of one block's LDS histogram (64 KB)
constexpr int kThreads = 256;
constexpr int kFoldCols = 67;     // LDS columns of the unfolded Delta phi index (fast form)
constexpr double kCapY = 65536.;  // |A|, |B| above this (bin units) take the exact form
constexpr unsigned long long kLaunchPairs = 0xffffffffull;  // < pairs per histogram per launch

struct BfRec {
    double phi, y;
};

// Per-particle values of the fast form in bin units (Delta phi, Delta y).
struct BfPrep {
    float p, y;
};

struct BfGeom {
    double dphi, Bphi_min, drap, Brap_min;
    int Bnpts;
};

// One unit of work: event e, B particles [64 c, 64 c + 64), A rows [i0, i1).
struct BfUnit {
    int e, c, i0, i1;
};

struct BfArgs {
    BfGeom g;
    const BfRec *Arec[8], *Brec[8];
    const long long *Aoff[8], *Boff[8];
    const BfPrep *Aprep[8], *Bprep[8];
    const BfUnit *units[8];
    int nunits[8];
    int *queue;               // [8] next unit of each histogram, zero at launch
    const int *mix_ev;        // [4][nev] partner event of the mixed histograms
    const double *mix_rot;    // [4][nev] rotation
    int nev;
    float Hp, Hy, Gy;         // fast form: |frac(t) - 1/2| < H; odd Bnpts also > Gy
    double *hist;             // [8][Bnpts][20]
    unsigned long long *fallback;
};

// Exact bin of one pair, the reference's arithmetic (BalanceFunction.cpp:
// 136-153); returns the Delta y bin or -1 when the pair is not binned.  A
// NaN Delta y is skipped: the reference's int conversion of it gives INT_MIN
// on x86, outside the bins.
__device__ __forceinline__ int bf_bin_exact(const BfGeom &g, double dphi_local, double dy,
                                            int &phi_idx) {
    int p = ((int)floor((dphi_local - g.Bphi_min) / g.dphi)) % kBnphi;
    if (p < 0) p += kBnphi;
    phi_idx = p;
    if (fabs(dy) < 1e-10) return -1;
    if (!(dy >= g.Brap_min)) return -1;
    const int y_idx = (int)((dy - g.Brap_min) / g.drap);
    return (y_idx >= 0 && y_idx < g.Bnpts) ? y_idx : -1;
}

// Read-only, wave-uniform data seen through the constant address space
// is read with scalar loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define BF_CONST_PTR(T) const __attribute__((address_space(4))) T *
#else
#define BF_CONST_PTR(T) const T *
#endif
#define BF_CONST(T, p) ((BF_CONST_PTR(T))(p))

__device__ __forceinline__ int next_unit(int *queue) {
    int u = 0;
    if ((threadIdx.x & 63) == 0) u = atomicAdd(queue, 1);
    return __builtin_amdgcn_readfirstlane(u);
}

// LDS layout: rows 0 .. Bnpts+1 (row y_idx + 1 holds Delta y bin y_idx; rows
// 0 and Bnpts + 1 collect the fast form's pairs outside the Delta y range and
// the lanes that bin nothing; they are never read),
// S columns: the Delta phi bin (S = 20) or, FOLD, the unfolded index
// floor(t_phi) in [5, 65], folded mod 20 when the block adds its histogram.
template <bool EXACT, bool FOLD, bool ODD>
__global__ __launch_bounds__(kThreads) void k_bf(BfArgs a) {
    extern __shared__ unsigned h[];
    constexpr int S = FOLD ? kFoldCols : kBnphi;
    const int k = blockIdx.y;
    const int nwords = (a.g.Bnpts + 2) * S;
    for (int b = threadIdx.x; b < nwords; b += blockDim.x) h[b] = 0u;
    __syncthreads();

    const bool mixed = k >= 4;
    const BfRec *__restrict__ Arec = a.Arec[k];
    const BfRec *__restrict__ Brec = a.Brec[k];
    const long long *__restrict__ Aoff = a.Aoff[k];
    const long long *__restrict__ Boff = a.Boff[k];
    const BfPrep *__restrict__ Aprep = a.Aprep[k];
    const BfPrep *__restrict__ Bprep = a.Bprep[k];
    const BfUnit *__restrict__ units = a.units[k];
    const int nunits = a.nunits[k];
    const int lane = threadIdx.x & 63;
    const float top = (float)(a.g.Bnpts + 1);
    const unsigned junk = (unsigned)(lane % S);  // this lane's word of row 0
    unsigned nfb = 0;

    for (;;) {
        const int u = next_unit(a.queue + k);
        if (u >= nunits) break;
        const BfUnit un = units[u];
        int eb = un.e;
        double rot = 0.0;
        if (mixed) {
            eb = a.mix_ev[(k - 4) * a.nev + un.e];
            rot = a.mix_rot[(k - 4) * a.nev + un.e];
        }
        const long long a0 = Aoff[un.e];
        const long long b0 = Boff[eb];
        const long long nb = Boff[eb + 1] - b0;
        const long long j = (long long)un.c * 64 + lane;
        const bool active = j < nb;
        BfRec rb{0.0, 0.0};
        BfPrep bp{0.f, 0.f};
        if (active) {
            rb = Brec[b0 + j];
            if (!EXACT) bp = Bprep[b0 + j];
        }
        // A rows are wave-uniform: constant-address-space loads become
        // scalar loads, eight rows per load in the unrolled loop
        BF_CONST_PTR(BfPrep) ap_row = BF_CONST(BfPrep, Aprep + a0);
        BF_CONST_PTR(BfRec) ar_row = BF_CONST(BfRec, Arec + a0);
        // Fast form of one pair: bins it and returns whether it needs the
        // exact form.  Branch-free: lanes that bin nothing (inactive, or
        // inside a guard band) add to their own word of LDS row 0, which is
        // never read, so the eight rows of a batch form one basic block the
        // compiler can interleave.
        auto fastpair = [&](BfPrep ap) -> bool {
            const float tp = ap.p - bp.p;
            const float ty = ap.y - bp.y;
            const float ep = fabsf(__builtin_amdgcn_fractf(tp) - 0.5f);
            const float ey = fabsf(__builtin_amdgcn_fractf(ty) - 0.5f);
            bool ok = ep < a.Hp && ey < a.Hy;
            if (ODD) ok = ok && ey > a.Gy;
            const unsigned row = (unsigned)__builtin_amdgcn_fmed3f(ty, 0.f, top);
            const unsigned x = (unsigned)tp;
            const unsigned col = FOLD ? x : x % (unsigned)kBnphi;
            unsigned word = __umul24(row, (unsigned)S) + col;
            asm("" : "+v"(word));  // keep the select below a select
            atomicAdd(&h[active && ok ? word : junk], 1u);
            return active && !ok;
        };
        // Exact form (the reference's FP64 arithmetic) of the pair with row i.
        auto slowpair = [&](int i) {
            const BfRec ra = ar_row[i];
            const double d = ra.phi - rb.phi;
            int p;
            const int y = bf_bin_exact(a.g, mixed ? d + rot : d, ra.y - rb.y, p);
            if (y >= 0) atomicAdd(&h[(y + 1) * S + p], 1u);
            if (!EXACT) nfb++;
        };
        int i = un.i0;
        if (!EXACT && i + 8 <= un.i1) {
            // the next batch of rows is loaded while this one is binned
            // (its start clamped to i1 - 8, so the load stays in the unit);
            // the rare exact-form pairs of a batch follow it
            BfPrep q[8];
#pragma unroll
            for (int r = 0; r < 8; r++) q[r] = ap_row[i + r];
            for (; i + 8 <= un.i1; i += 8) {
                const int nx = min(i + 8, un.i1 - 8);
                BfPrep qn[8];
#pragma unroll
                for (int r = 0; r < 8; r++) qn[r] = ap_row[nx + r];
                bool sl[8];
#pragma unroll
                for (int r = 0; r < 8; r++) sl[r] = fastpair(q[r]);
                if (__builtin_amdgcn_ballot_w64(sl[0] || sl[1] || sl[2] || sl[3] || sl[4] ||
                                                sl[5] || sl[6] || sl[7])) {
#pragma unroll
                    for (int r = 0; r < 8; r++)
                        if (sl[r]) slowpair(i + r);
                }
#pragma unroll
                for (int r = 0; r < 8; r++) q[r] = qn[r];
            }
        }
        for (; i < un.i1; i++) {
            if (EXACT) {
                if (active) slowpair(i);
            } else if (fastpair(ap_row[i])) {
                slowpair(i);
            }
        }
    }
    if (!EXACT && nfb) atomicAdd(a.fallback, (unsigned long long)nfb);

    __syncthreads();
    const int nbins = a.g.Bnpts * kBnphi;
    double *out = a.hist + (size_t)k * nbins;
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
        const int y = b / kBnphi, c = b - y * kBnphi;
        unsigned long long s = 0;
        for (int x = c; x < S; x += kBnphi) s += h[(y + 1) * S + x];
        if (s) atomicAdd(&out[b], (double)s);
    }
}

// Per-particle values of the fast form, in bin units:
//   A rows of histogram k: U = (phi + rot - Bphi_min)/dphi + 20, in [15, 55]
//     for |phi| <= pi and rot in [0, 2 pi]; A = (y - Brap_min)/drap + 1;
//   B particles: V = phi/dphi, in [-10, 10]; B = y/drap;
// so t_phi = U - V lies in [5, 65] and floor(t_y) = y_idx + 1.  Values with
// |A| or |B| above kCapY become NaN, so their pairs take the exact form.
// blockIdx.y < 8: A rows of histogram blockIdx.y (blockIdx.x strides the
// events); blockIdx.y >= 8: B values of list blockIdx.y - 8.
struct BfPrepArgs {
    BfGeom g;
    double inv_dphi, inv_drap;
    const BfRec *Arec[8];
    const long long *Aoff[8];
    BfPrep *Aprep[8];
    const BfRec *Lrec[6];
    long long Ln[6];
    BfPrep *Lprep[6];
    int nlists;
    const double *mix_rot;
    int nev;
};

__device__ __forceinline__ float bf_capped(double v) {
    return fabs(v) <= kCapY ? (float)v : __builtin_nanf("");
}

__global__ __launch_bounds__(kThreads) void k_bf_prep(BfPrepArgs a) {
    const int y = blockIdx.y;
    if (y < 8) {
        const BfRec *rec = a.Arec[y];
        const long long *off = a.Aoff[y];
        BfPrep *out = a.Aprep[y];
        for (int e = blockIdx.x; e < a.nev; e += gridDim.x) {
            const double rot = y >= 4 ? a.mix_rot[(y - 4) * a.nev + e] : 0.0;
            for (long long i = off[e] + threadIdx.x; i < off[e + 1]; i += blockDim.x) {
                const BfRec r = rec[i];
                const double u = ((r.phi + rot) - a.g.Bphi_min) * a.inv_dphi + 20.0;
                const double av = (r.y - a.g.Brap_min) * a.inv_drap + 1.0;
                out[i] = BfPrep{(float)u, bf_capped(av)};
            }
        }
    } else {
        const int l = y - 8;
        if (l >= a.nlists) return;
        const BfRec *rec = a.Lrec[l];
        BfPrep *out = a.Lprep[l];
        for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.Ln[l];
             i += (long long)gridDim.x * blockDim.x) {
            const BfRec r = rec[i];
            out[i] = BfPrep{(float)(r.phi * a.inv_dphi), bf_capped(r.y * a.inv_drap)};
        }
    }
}

// One species list of a buffer: (phi, y or eta) of the particles with the
// id (99999: all) that pass the pT cut, in file order, and event offsets.
// The records live in the context's pinned staging area.
struct HostList {
    BfRec *rec = nullptr;
    size_t n = 0;
    std::vector<long long> off;
    double max_phi = 0, max_y = 0;  // largest finite |phi|, |y|
};

// Charge, baryon number and strangeness of a species.
struct SpQ {
    int32_t charge, baryon, strange;
};

bool is_class(int id) {
    return id == 9999 || id == 9998 || id == -9998 || id == 9997 || id == -9997 || id == 9996 ||
           id == -9996;
}

// particleSamples::decide_to_pick_OSCAR (particleSamples.cpp:625-678).
bool bf_pick(int poi, int mon, const SpQ *q) {
    switch (poi) {
        case 99999: return true;
        case 9999: return q->charge != 0;
        case 9998: return q->charge > 0;
        case -9998: return q->charge < 0;
        case 9997: return q->baryon > 0;
        case -9997: return q->baryon < 0;
        case 9996: return q->strange > 0;
        case -9996: return q->strange < 0;
        default: return mon == poi;
    }
}

// Which of the lists ids[0..nl) a particle belongs to (bit l).  With a
// class among the ids every particle's quantum numbers are looked up, as the
// reference looks them up; a monval missing from the table is recorded in
// bad and belongs to no list.
struct Picker {
    int nl = 0;
    int ids[4] = {0, 0, 0, 0};
    const std::unordered_map<int32_t, SpQ> *sp = nullptr;  // set when a class is used
    std::atomic<long long> *bad = nullptr;
    unsigned mask(int mon) const {
        const SpQ *q = nullptr;
        if (sp) {
            const auto it = sp->find(mon);
            if (it == sp->end()) {
                long long none = kNoBad;
                bad->compare_exchange_strong(none, mon);
                return 0u;
            }
            q = &it->second;
        }
        unsigned m = 0u;
        for (int l = 0; l < nl; l++)
            if (bf_pick(ids[l], mon, q)) m |= 1u << l;
        return m;
    }
    static constexpr long long kNoBad = 1ll << 40;
};

template <typename F>
void par_ranges(long long n, long long min_per_thread, F fn) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const long long nt =
        std::max(1ll, std::min<long long>(hw, n / std::max(1ll, min_per_thread)));
    if (nt <= 1) {
        fn(0ll, n);
        return;
    }
    std::vector<std::thread> th;
    for (long long t = 0; t < nt; t++) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
    for (auto &x : th) x.join();
}

// First pass over one buffer: event offsets and sizes of the lists of
// ids[0..nl), over event ranges on up to 16 threads (the pT cut of
// BalanceFunction.cpp:127-129).
void count_lists(const hbt_bf_config &cfg, int32_t nev, const int64_t *ev_off,
                 const int32_t *monval, const double *const cols[5], const Picker &pk,
                 HostList *L) {
    const int nl = pk.nl;
    const double *px = cols[1], *py = cols[2];
    std::vector<long long> cnt((size_t)nl * std::max(nev, 1), 0);
    par_ranges(nev, 32, [&](long long e0, long long e1) {
        for (long long e = e0; e < e1; e++)
            for (int64_t i = ev_off[e]; i < ev_off[e + 1]; i++) {
                const unsigned m = pk.mask(monval[i]);
                if (!m) continue;
                const double pT = sqrt(px[i] * px[i] + py[i] * py[i]);
                if (pT < cfg.BpT_min || pT > cfg.BpT_max) continue;
                for (int l = 0; l < nl; l++)
                    if (m >> l & 1u) cnt[(size_t)l * nev + e]++;
            }
    });
    for (int l = 0; l < nl; l++) {
        L[l].off.assign((size_t)nev + 1, 0);
        for (int e = 0; e < nev; e++) L[l].off[e + 1] = L[l].off[e] + cnt[(size_t)l * nev + e];
        L[l].n = (size_t)L[l].off[nev];
        L[l].max_phi = L[l].max_y = 0;
    }
}

// Second pass: the records, into L[l].rec (particleSamples.cpp:447-465 with
// rap_shift 0), and the largest finite |phi| and |y| of each list.  (One
// pass into per-thread vectors copied into place afterwards measured slower
// on 16 host threads.)
void fill_lists(const hbt_bf_config &cfg, int32_t nev, const int64_t *ev_off,
                const int32_t *monval, const double *const cols[5], const Picker &pk,
                HostList *L) {
    const int nl = pk.nl;
    const double *mass = cols[0], *px = cols[1], *py = cols[2], *pz = cols[3];
    std::mutex mu;
    par_ranges(nev, 32, [&](long long e0, long long e1) {
        double mphi[4] = {0, 0, 0, 0}, my[4] = {0, 0, 0, 0};
        for (long long e = e0; e < e1; e++) {
            long long pos[4] = {0, 0, 0, 0};
            for (int l = 0; l < nl; l++) pos[l] = L[l].off[e];
            for (int64_t i = ev_off[e]; i < ev_off[e + 1]; i++) {
                const unsigned m = pk.mask(monval[i]);
                if (!m) continue;
                const double pT = sqrt(px[i] * px[i] + py[i] * py[i]);
                if (pT < cfg.BpT_min || pT > cfg.BpT_max) continue;
                const double mT = sqrt(pT * pT + mass[i] * mass[i]);
                const double y = cfg.rap_type == 0 ? asinh(pz[i] / pT) : asinh(pz[i] / mT);
                const BfRec r{atan2(py[i], px[i]), y};
                for (int l = 0; l < nl; l++)
                    if (m >> l & 1u) {
                        L[l].rec[pos[l]++] = r;
                        if (std::isfinite(r.phi)) mphi[l] = std::max(mphi[l], std::fabs(r.phi));
                        if (std::isfinite(r.y)) my[l] = std::max(my[l], std::fabs(r.y));
                    }
            }
        }
        std::lock_guard<std::mutex> g(mu);
        for (int l = 0; l < nl; l++) {
            L[l].max_phi = std::max(L[l].max_phi, mphi[l]);
            L[l].max_y = std::max(L[l].max_y, my[l]);
        }
    });
}

// The distinct ids of ids[0..n) and each id's slot among them.
int unique_ids(const int *ids, int n, int *uniq, int *slot) {
    int nu = 0;
    for (int s = 0; s < n; s++) {
        int u = 0;
        while (u < nu && uniq[u] != ids[s]) u++;
        if (u == nu) uniq[nu++] = ids[s];
        slot[s] = u;
    }
    return nu;
}

}  // namespace

struct hbt_bf_ctx {
    hbt_bf_config cfg;
    BfGeom g;
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {};
    double *d_hist = nullptr;
    size_t nbins = 0;
    int64_t N_b = 0, N_bbar = 0;
    void *d_rec[6] = {};
    size_t rec_cap[6] = {};
    long long *d_off[6] = {};
    size_t off_cap[6] = {};
    BfPrep *d_prep = nullptr;
    size_t prep_cap = 0;
    BfUnit *d_units = nullptr;
    size_t units_cap = 0;
    int *d_queue = nullptr;
    size_t queue_cap = 0;
    int *d_mix_ev = nullptr;
    double *d_mix_rot = nullptr;
    size_t mix_ev_cap = 0, mix_rot_cap = 0;
    unsigned long long *d_fallback = nullptr;
    std::unordered_map<int32_t, SpQ> species;  // for the species classes
    bool has_species = false;
    BfRec *h_pin = nullptr;  // pinned staging of the species lists
    size_t pin_cap = 0;      // records
    int cu_count = 0;
    hbt_bf_info info{};
    std::string err;
};

namespace {

std::string g_bf_create_err;

int bf_fail(hbt_bf_ctx *c, int code, const std::string &msg) {
    if (c)
        c->err = msg;
    else
        g_bf_create_err = msg;
    return code;
}

#define BFCHECK(c, x)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return bf_fail(c, HBT_ERR_HIP, std::string(#x) + ": " +      \
                                                                 hipGetErrorString(e_));   \
    } while (0)

template <typename T>
int bf_ensure(hbt_bf_ctx *c, T **p, size_t &cap, size_t n) {
    n = std::max<size_t>(n, 1);
    if (*p && n <= cap) return HBT_OK;
    if (*p) hipFree(*p);
    *p = nullptr;
    cap = 0;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess)
        return bf_fail(c, HBT_ERR_NOMEM, "hipMalloc failed");
    cap = n;
    return HBT_OK;
}

// LDS bytes of one block of the chosen form.
size_t lds_bytes(int Bnpts, bool fold) {
    return (size_t)(Bnpts + 2) * (fold ? kFoldCols : kBnphi) * sizeof(unsigned);
}

using KernelFn = void (*)(BfArgs);

KernelFn pick_kernel(bool exact, bool fold, bool odd) {
    if (exact) return k_bf<true, false, false>;
    if (fold) return odd ? k_bf<false, true, true> : k_bf<false, true, false>;
    return odd ? k_bf<false, false, true> : k_bf<false, false, false>;
}

}  // namespace

extern "C" {

int hbt_bf_create(const hbt_bf_config *cfg, hbt_bf_ctx **out) {
    if (!cfg || !out) return HBT_ERR_ARG;
    *out = nullptr;
    if (cfg->Bnpts < 2 || (long)(cfg->Bnpts + 2) * kBnphi > kMaxLds)
        return bf_fail(nullptr, HBT_ERR_ARG, "Bnpts must lie in [2, 817]");
    if (cfg->rap_type != 0 && cfg->rap_type != 1)
        return bf_fail(nullptr, HBT_ERR_ARG, "rap_type must be 0 or 1");
    if (cfg->kernel != HBT_BF_KERNEL_FAST && cfg->kernel != HBT_BF_KERNEL_EXACT)
        return bf_fail(nullptr, HBT_ERR_ARG, "kernel must be HBT_BF_KERNEL_FAST or _EXACT");
    hbt_bf_ctx *c = new hbt_bf_ctx();
    c->cfg = *cfg;
    // BalanceFunction.cpp:29-35
    c->g.Bnpts = cfg->Bnpts;
    c->g.drap = 2. * std::fabs(cfg->Brap_max) / (cfg->Bnpts - 1);
    c->g.Brap_min = -std::fabs(cfg->Brap_max) - 0.5 * c->g.drap;
    c->g.dphi = 2. * M_PI / kBnphi;
    c->g.Bphi_min = -M_PI / 2.;
    c->nbins = (size_t)cfg->Bnpts * kBnphi;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
        delete c;
        return bf_fail(nullptr, HBT_ERR_NODEVICE, "no usable device");
    }
    int rc = HBT_OK;
    if (hipSetDevice(cfg->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess ||
        hipEventCreate(&c->ev[2]) != hipSuccess ||
        hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount,
                              cfg->device) != hipSuccess ||
        hipMalloc((void **)&c->d_hist, 8 * c->nbins * sizeof(double)) != hipSuccess ||
        hipMemset(c->d_hist, 0, 8 * c->nbins * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&c->d_fallback, sizeof(unsigned long long)) != hipSuccess)
        rc = HBT_ERR_HIP;
    if (rc) {
        g_bf_create_err = "device set-up failed";
        hbt_bf_destroy(c);
        return rc;
    }
    *out = c;
    return HBT_OK;
}

int hbt_bf_destroy(hbt_bf_ctx *c) {
    if (!c) return HBT_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int l = 0; l < 6; l++) {
        if (c->d_rec[l]) hipFree(c->d_rec[l]);
        if (c->d_off[l]) hipFree(c->d_off[l]);
    }
    for (void *p : {(void *)c->d_prep, (void *)c->d_units, (void *)c->d_queue,
                    (void *)c->d_mix_ev, (void *)c->d_mix_rot, (void *)c->d_fallback,
                    (void *)c->d_hist})
        if (p) hipFree(p);
    if (c->h_pin) hipHostFree(c->h_pin);
    for (hipEvent_t e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return HBT_OK;
}

const char *hbt_bf_last_error(const hbt_bf_ctx *c) {
    return c ? c->err.c_str() : g_bf_create_err.c_str();
}

int hbt_bf_reset(hbt_bf_ctx *c) {
    if (!c) return HBT_ERR_ARG;
    BFCHECK(c, hipSetDevice(c->cfg.device));
    BFCHECK(c, hipMemsetAsync(c->d_hist, 0, 8 * c->nbins * sizeof(double), c->stream));
    BFCHECK(c, hipStreamSynchronize(c->stream));
    c->N_b = c->N_bbar = 0;
    return HBT_OK;
}

}  // extern "C"

namespace {

// One buffer; the mixing draws come from rng or, when rng is null, from
// plan_ev / plan_rot ([4][nev], validated here).
int bf_calculate(hbt_bf_ctx *c, int32_t nev, const int64_t *ev_off, const int32_t *monval,
                 const double *const cols[5], int32_t mix_nev, const int64_t *mix_off,
                 const int32_t *mix_monval, const double *const mix_cols[5], hbt_rng *rng,
                 const int32_t *plan_ev, const double *plan_rot) {
    if (!c || nev < 0) return HBT_ERR_ARG;
    if (nev > 0 && (!ev_off || !monval || !cols)) return bf_fail(c, HBT_ERR_ARG, "null buffer");
    const auto t0 = std::chrono::steady_clock::now();
    const bool alias = mix_off == nullptr;
    if (alias) {
        mix_nev = nev;
        mix_off = ev_off;
        mix_monval = monval;
        mix_cols = cols;
    }
    if (nev > 0 && mix_nev < 1) return bf_fail(c, HBT_ERR_ARG, "no mixed events");
    if (nev > 0 && !alias && (!mix_monval || !mix_cols))
        return bf_fail(c, HBT_ERR_ARG, "null mixed buffer");
    if (!rng) {
        if (nev > 0 && (!plan_ev || !plan_rot)) return bf_fail(c, HBT_ERR_ARG, "null plan");
        for (size_t i = 0; i < (size_t)4 * nev; i++)
            if (plan_ev[i] < 0 || plan_ev[i] >= mix_nev)
                return bf_fail(c, HBT_ERR_ARG, "mixed partner outside the mixed buffer");
    }
    BFCHECK(c, hipSetDevice(c->cfg.device));
    const hbt_bf_config &k = c->cfg;
    c->info = hbt_bf_info{};
    c->info.kernel = k.kernel;

    // species lists: a, abar, b, bbar of this buffer and b, bbar of the
    // mixed buffer, one list per distinct (buffer, id)
    const int sid[4] = {k.particle_alpha, -k.particle_alpha, k.particle_beta, -k.particle_beta};
    int uid[4], sslot[4];
    const int nu = unique_ids(sid, 4, uid, sslot);
    int mslot[2] = {sslot[2], sslot[3]};
    int muid[2], mu_slot[2];
    int nmu = 0;
    if (!alias) {
        nmu = unique_ids(sid + 2, 2, muid, mu_slot);
        mslot[0] = nu + mu_slot[0];
        mslot[1] = nu + mu_slot[1];
    }
    const int nlists = nu + nmu;
    bool classes = false;
    for (int id : sid) classes = classes || is_class(id);
    if (classes && !c->has_species)
        return bf_fail(c, HBT_ERR_STATE, "species classes need hbt_bf_set_species");
    std::atomic<long long> bad{Picker::kNoBad};
    Picker pk_main, pk_mix;
    pk_main.nl = nu;
    std::copy(uid, uid + nu, pk_main.ids);
    pk_mix.nl = nmu;
    std::copy(muid, muid + nmu, pk_mix.ids);
    for (Picker *pk : {&pk_main, &pk_mix}) {
        pk->sp = classes ? &c->species : nullptr;
        pk->bad = &bad;
    }
    HostList L[6];
    count_lists(k, nev, ev_off, monval, cols, pk_main, L);
    if (nmu) count_lists(k, mix_nev, mix_off, mix_monval, mix_cols, pk_mix, L + nu);
    if (bad.load() != Picker::kNoBad)
        return bf_fail(c, HBT_ERR_ARG, "monval " + std::to_string(bad.load()) +
                                           " is not in the species table");
    // the records go to pinned memory (reused across calls: the previous
    // call's copies have finished), so their upload is a DMA at link speed
    size_t nrec = 0;
    for (int l = 0; l < nlists; l++) nrec += L[l].n;
    if (nrec > c->pin_cap) {
        if (c->h_pin) hipHostFree(c->h_pin);
        c->h_pin = nullptr;
        c->pin_cap = 0;
        const size_t want = nrec + nrec / 2;
        if (hipHostMalloc((void **)&c->h_pin, want * sizeof(BfRec), hipHostMallocDefault) !=
            hipSuccess)
            return bf_fail(c, HBT_ERR_NOMEM, "hipHostMalloc failed");
        c->pin_cap = want;
    }
    size_t at = 0;
    for (int l = 0; l < nlists; l++) {
        L[l].rec = c->h_pin + at;
        at += L[l].n;
    }
    fill_lists(k, nev, ev_off, monval, cols, pk_main, L);
    if (nmu) fill_lists(k, mix_nev, mix_off, mix_monval, mix_cols, pk_mix, L + nu);
    // N_b, N_bbar: b and bbar particles passing the pT cut
    // (BalanceFunction.cpp:68-69, :199-209)
    c->N_b += (int64_t)L[sslot[2]].n;
    c->N_bbar += (int64_t)L[sslot[3]].n;
    // mixing draws of the four mixed histograms, in call order, per event a
    // partner then a rotation (BalanceFunction.cpp:165-169)
    std::vector<int> mix_ev((size_t)4 * std::max(nev, 1));
    std::vector<double> mix_rot((size_t)4 * std::max(nev, 1));
    bool rot_in_range = true;  // the fast form's bounds assume rot in [0, 2 pi]
    for (int hh = 0; hh < 4; hh++)
        for (int e = 0; e < nev; e++) {
            const size_t i = (size_t)hh * nev + e;
            if (rng) {
                mix_ev[i] = hbt_rng_int(rng) % mix_nev;
                mix_rot[i] = hbt_rng_uniform(rng) * 2. * M_PI;
            } else {
                mix_ev[i] = plan_ev[i];
                mix_rot[i] = plan_rot[i];
            }
            if (!(mix_rot[i] >= 0.0 && mix_rot[i] <= 2. * M_PI)) rot_in_range = false;
        }
    if (nev == 0) return HBT_OK;

    // lists of each histogram's A rows and B particles
    int Aslot[8], Bslot[8];
    for (int hh = 0; hh < 8; hh++) {
        Aslot[hh] = sslot[hh & 1];
        const int bb = ((hh & 3) == 1 || (hh & 3) == 2) ? 1 : 0;
        Bslot[hh] = hh < 4 ? sslot[2 + bb] : mslot[bb];
    }

    // fast-form guards from the lists' largest |phi| and |y|
    const double inv_dphi = 1.0 / c->g.dphi, inv_drap = 1.0 / c->g.drap;
    double mpa = 0, mpb = 0, mya = 0, myb = 0;
    for (int hh = 0; hh < 8; hh++) {
        mpa = std::max(mpa, L[Aslot[hh]].max_phi);
        mya = std::max(mya, L[Aslot[hh]].max_y);
        mpb = std::max(mpb, L[Bslot[hh]].max_phi);
        myb = std::max(myb, L[Bslot[hh]].max_y);
    }
    const double maxU = (mpa + 2. * M_PI + std::fabs(c->g.Bphi_min)) * inv_dphi + 20.0;
    const double maxV = mpb * inv_dphi;
    const double maxA = std::min(kCapY, (mya + std::fabs(c->g.Brap_min)) * inv_drap + 1.0);
    const double maxB = std::min(kCapY, myb * inv_drap);
    const double Gp = std::ldexp(maxU + maxV, -21) + 1e-6;
    const double Gy = std::ldexp(maxA + maxB, -21) + 1e-6 + 4e-10 * inv_drap;
    const bool exact = k.kernel == HBT_BF_KERNEL_EXACT || mpa > M_PI || mpb > M_PI ||
                       !rot_in_range || Gp > 0.25 || Gy > 0.25;
    const bool odd = (k.Bnpts & 1) != 0;
    const bool fold = !exact && lds_bytes(k.Bnpts, true) <= kMaxLds * sizeof(unsigned);
    c->info.kernel = exact ? HBT_BF_KERNEL_EXACT : HBT_BF_KERNEL_FAST;
    // FP32 thresholds on the safe side: H rounded down, Gy rounded up
    float Hp = (float)(0.5 - Gp), Hy = (float)(0.5 - Gy), Gyf = (float)Gy;
    if ((double)Hp > 0.5 - Gp) Hp = std::nextafter(Hp, 0.f);
    if ((double)Hy > 0.5 - Gy) Hy = std::nextafter(Hy, 0.f);
    if ((double)Gyf < Gy) Gyf = std::nextafter(Gyf, 1.f);

    // work units per histogram: events x 64-particle B chunks x runs of A
    // rows sized for about four units per wave; launches of < 2^32 pairs
    // per histogram
    const int bx = std::max(1, c->cu_count);
    const long long waves = (long long)bx * (kThreads / 64);
    std::vector<BfUnit> units[8];
    std::vector<int> seg[8];  // first unit of each launch
    unsigned long long pairs_total = 0;
    for (int hh = 0; hh < 8; hh++) {
        const std::vector<long long> &Ao = L[Aslot[hh]].off;
        const std::vector<long long> &Bo = L[Bslot[hh]].off;
        long long rows = 0;
        for (int e = 0; e < nev; e++) {
            const int eb = hh < 4 ? e : mix_ev[(size_t)(hh - 4) * nev + e];
            rows += (Ao[e + 1] - Ao[e]) * ((Bo[eb + 1] - Bo[eb] + 63) / 64);
        }
        const long long run =
            std::min(1ll << 24, std::max(32ll, (rows + 4 * waves - 1) / (4 * waves)));
        unsigned long long acc = 0;
        seg[hh].push_back(0);
        for (int e = 0; e < nev; e++) {
            const int eb = hh < 4 ? e : mix_ev[(size_t)(hh - 4) * nev + e];
            const long long na = Ao[e + 1] - Ao[e], nb = Bo[eb + 1] - Bo[eb];
            if (na == 0 || nb == 0) continue;
            for (long long i0 = 0; i0 < na; i0 += run)
                for (long long cc = 0; cc * 64 < nb; cc++) {
                    const long long i1 = std::min(na, i0 + run);
                    const unsigned long long p = (unsigned long long)(i1 - i0) *
                                                 (unsigned long long)std::min(64ll, nb - cc * 64);
                    if (acc + p > kLaunchPairs) {
                        seg[hh].push_back((int)units[hh].size());
                        acc = 0;
                    }
                    acc += p;
                    pairs_total += p;
                    units[hh].push_back(BfUnit{e, (int)cc, (int)i0, (int)i1});
                }
            if (units[hh].size() > (size_t)0x7fffffff)
                return bf_fail(c, HBT_ERR_CAPACITY, "too many work units");
        }
    }
    c->info.pairs = pairs_total;
    size_t nlaunch = 1;
    for (int hh = 0; hh < 8; hh++) nlaunch = std::max(nlaunch, seg[hh].size());

    // device buffers: lists, offsets, draws, units, queues, prepared values
    for (int l = 0; l < nlists; l++) {
        int rc = bf_ensure(c, (BfRec **)&c->d_rec[l], c->rec_cap[l], L[l].n);
        if (!rc) rc = bf_ensure(c, &c->d_off[l], c->off_cap[l], L[l].off.size());
        if (rc) return rc;
        if (L[l].n)
            BFCHECK(c, hipMemcpyAsync(c->d_rec[l], L[l].rec, L[l].n * sizeof(BfRec),
                                      hipMemcpyHostToDevice, c->stream));
        BFCHECK(c, hipMemcpyAsync(c->d_off[l], L[l].off.data(), L[l].off.size() * sizeof(long long),
                                  hipMemcpyHostToDevice, c->stream));
    }
    int rc = bf_ensure(c, &c->d_mix_ev, c->mix_ev_cap, mix_ev.size());
    if (!rc) rc = bf_ensure(c, &c->d_mix_rot, c->mix_rot_cap, mix_rot.size());
    size_t nunits_all = 0, uoff[8];
    for (int hh = 0; hh < 8; hh++) {
        uoff[hh] = nunits_all;
        nunits_all += units[hh].size();
    }
    if (!rc) rc = bf_ensure(c, &c->d_units, c->units_cap, nunits_all);
    if (!rc) rc = bf_ensure(c, &c->d_queue, c->queue_cap, 8 * nlaunch);
    size_t poff_A[8] = {}, poff_L[6] = {}, nprep = 0;
    if (!exact) {
        for (int hh = 0; hh < 8; hh++) {
            poff_A[hh] = nprep;
            nprep += L[Aslot[hh]].n;
        }
        for (int l = 0; l < nlists; l++) {
            poff_L[l] = nprep;
            nprep += L[l].n;
        }
        if (!rc) rc = bf_ensure(c, &c->d_prep, c->prep_cap, nprep);
    }
    if (rc) return rc;
    BFCHECK(c, hipMemcpyAsync(c->d_mix_ev, mix_ev.data(), mix_ev.size() * sizeof(int),
                              hipMemcpyHostToDevice, c->stream));
    BFCHECK(c, hipMemcpyAsync(c->d_mix_rot, mix_rot.data(), mix_rot.size() * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    for (int hh = 0; hh < 8; hh++)
        if (!units[hh].empty())
            BFCHECK(c, hipMemcpyAsync(c->d_units + uoff[hh], units[hh].data(),
                                      units[hh].size() * sizeof(BfUnit), hipMemcpyHostToDevice,
                                      c->stream));
    BFCHECK(c, hipMemsetAsync(c->d_queue, 0, 8 * nlaunch * sizeof(int), c->stream));
    BFCHECK(c, hipMemsetAsync(c->d_fallback, 0, sizeof(unsigned long long), c->stream));
    c->info.host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    BFCHECK(c, hipEventRecord(c->ev[0], c->stream));
    if (!exact) {
        BfPrepArgs pa{};
        pa.g = c->g;
        pa.inv_dphi = inv_dphi;
        pa.inv_drap = inv_drap;
        for (int hh = 0; hh < 8; hh++) {
            pa.Arec[hh] = (const BfRec *)c->d_rec[Aslot[hh]];
            pa.Aoff[hh] = c->d_off[Aslot[hh]];
            pa.Aprep[hh] = c->d_prep + poff_A[hh];
        }
        for (int l = 0; l < nlists; l++) {
            pa.Lrec[l] = (const BfRec *)c->d_rec[l];
            pa.Ln[l] = (long long)L[l].n;
            pa.Lprep[l] = c->d_prep + poff_L[l];
        }
        pa.nlists = nlists;
        pa.mix_rot = c->d_mix_rot;
        pa.nev = nev;
        hipLaunchKernelGGL(k_bf_prep, dim3(std::max(1, std::min(nev, 4 * bx)), 8 + nlists),
                           dim3(kThreads), 0, c->stream, pa);
        BFCHECK(c, hipGetLastError());
    }
    BFCHECK(c, hipEventRecord(c->ev[1], c->stream));
    BfArgs ka{};
    ka.g = c->g;
    for (int hh = 0; hh < 8; hh++) {
        ka.Arec[hh] = (const BfRec *)c->d_rec[Aslot[hh]];
        ka.Brec[hh] = (const BfRec *)c->d_rec[Bslot[hh]];
        ka.Aoff[hh] = c->d_off[Aslot[hh]];
        ka.Boff[hh] = c->d_off[Bslot[hh]];
        ka.Aprep[hh] = exact ? nullptr : c->d_prep + poff_A[hh];
        ka.Bprep[hh] = exact ? nullptr : c->d_prep + poff_L[Bslot[hh]];
    }
    ka.mix_ev = c->d_mix_ev;
    ka.mix_rot = c->d_mix_rot;
    ka.nev = nev;
    ka.Hp = Hp;
    ka.Hy = Hy;
    ka.Gy = Gyf;
    ka.hist = c->d_hist;
    ka.fallback = c->d_fallback;
    const KernelFn fn = pick_kernel(exact, fold, odd);
    const size_t shm = lds_bytes(k.Bnpts, fold);
    for (size_t s = 0; s < nlaunch; s++) {
        for (int hh = 0; hh < 8; hh++) {
            const size_t b = s < seg[hh].size() ? (size_t)seg[hh][s] : units[hh].size();
            const size_t e = s + 1 < seg[hh].size() ? (size_t)seg[hh][s + 1] : units[hh].size();
            ka.units[hh] = c->d_units + uoff[hh] + b;
            ka.nunits[hh] = (int)(e - b);
        }
        ka.queue = c->d_queue + 8 * s;
        hipLaunchKernelGGL(fn, dim3(bx, 8), dim3(kThreads), shm, c->stream, ka);
        BFCHECK(c, hipGetLastError());
    }
    BFCHECK(c, hipEventRecord(c->ev[2], c->stream));
    BFCHECK(c, hipStreamSynchronize(c->stream));  // the host lists are freed on return
    float ms_prep = 0, ms_pairs = 0;
    BFCHECK(c, hipEventElapsedTime(&ms_prep, c->ev[0], c->ev[1]));
    BFCHECK(c, hipEventElapsedTime(&ms_pairs, c->ev[1], c->ev[2]));
    unsigned long long fb = 0;
    BFCHECK(c, hipMemcpy(&fb, c->d_fallback, sizeof(fb), hipMemcpyDeviceToHost));
    c->info.prep_ms = ms_prep;
    c->info.kernel_ms = ms_pairs;
    c->info.fallback = exact ? 0 : fb;
    c->info.launches = (int32_t)nlaunch;
    return HBT_OK;
}

}  // namespace

extern "C" {

int hbt_bf_calculate(hbt_bf_ctx *c, int32_t nev, const int64_t *ev_off, const int32_t *monval,
                     const double *const cols[5], int32_t mix_nev, const int64_t *mix_off,
                     const int32_t *mix_monval, const double *const mix_cols[5], hbt_rng *rng) {
    if (!rng) return HBT_ERR_ARG;
    return bf_calculate(c, nev, ev_off, monval, cols, mix_nev, mix_off, mix_monval, mix_cols, rng,
                        nullptr, nullptr);
}

int hbt_bf_calculate_planned(hbt_bf_ctx *c, int32_t nev, const int64_t *ev_off,
                             const int32_t *monval, const double *const cols[5],
                             int32_t mix_nev, const int64_t *mix_off,
                             const int32_t *mix_monval, const double *const mix_cols[5],
                             const int32_t *mix_ev, const double *mix_rot) {
    return bf_calculate(c, nev, ev_off, monval, cols, mix_nev, mix_off, mix_monval, mix_cols,
                        nullptr, mix_ev, mix_rot);
}

int hbt_bf_set_species(hbt_bf_ctx *c, int32_t n, const int32_t *monval, const int32_t *charge,
                       const int32_t *baryon, const int32_t *strange) {
    if (!c || n < 0 || (n > 0 && (!monval || !charge || !baryon || !strange))) return HBT_ERR_ARG;
    c->species.clear();
    for (int32_t i = 0; i < n; i++) c->species[monval[i]] = SpQ{charge[i], baryon[i], strange[i]};
    c->has_species = true;
    return HBT_OK;
}

int hbt_bf_last_info(hbt_bf_ctx *c, hbt_bf_info *info) {
    if (!c || !info) return HBT_ERR_ARG;
    *info = c->info;
    return HBT_OK;
}

int hbt_bf_read(hbt_bf_ctx *c, double *hist, int64_t *N_b, int64_t *N_bbar) {
    if (!c) return HBT_ERR_ARG;
    BFCHECK(c, hipSetDevice(c->cfg.device));
    BFCHECK(c, hipStreamSynchronize(c->stream));
    if (hist)
        BFCHECK(c, hipMemcpy(hist, c->d_hist, 8 * c->nbins * sizeof(double),
                             hipMemcpyDeviceToHost));
    if (N_b) *N_b = c->N_b;
    if (N_bbar) *N_bbar = c->N_bbar;
    return HBT_OK;
}

}  // extern "C"
