// hbt_fast.h -- the fast pair kernel: a cheap conservative prefilter over
// every enumerated pair, per-wave compaction of the survivors into an LDS
// ring, and the exact cascade on full waves of survivors.
//
// Why: the reference's cascade exits early for ~98% of pairs
// (HBT_correlation.cpp:319-398).  On a 64-wide wave, lanes that exit early
// idle while any lane continues, so a direct port pays the whole cascade
// (sqrt, three to five divisions, cos) for almost every pair.  Here the
// work items' windows leave out most pairs unevaluated, every evaluated pair
// costs 16 VALU operations, and only its survivors (~1.4% of the enumerated
// pairs on the bench workload) reach the exact cascade, 64 at a time with
// all lanes busy.
//
// Prefilter (all conservative: a pair the exact cascade accepts always
// passes; the margins exceed the rounding error of both computations):
//   T = |p1T + p2T|^2 = 4 K_perp^2           in [4 KT_min^2, 4 KT_max^2]
//   q_out^2 = dn^2 / T <= H^2                (q_out = (p1T - p2T).(p1T + p2T)
//        / |p1T + p2T| = dn / sqrt(T), dn = n1 - n2)
//   q_side^2 = Q - q_out^2 <= H^2            (Q = |p1T - p2T|^2)
//   |beta1 - beta2| <= (H/2) (1/E1 + 1/E2)   (from |q_long| <= H, M_t <= K_E,
//        q_long M_t = E2 pz1 - E1 pz2; beta = pz/E)
// with H = max(|q_lo_cut|, |q_hi_cut|), the q windows tested without a
// division as dn^2 <= H^2 T and (Q - H^2) T <= dn^2.  T and Q come from one
// dot product: T = n1 + n2 + 2 p1.p2, Q = n1 + n2 - 2 p1.p2 (n = |pT|^2).
// For mixed pairs the i particle is rotated by -theta instead of rotating
// every j by +theta (T, Q, dn and beta are rotation invariant); the exact
// cascade then uses the reference's rotated j (HBT_correlation.cpp:522-523).
#pragma once
#include "hbt_device.h"
#include "hbt_f32.h"
#include "hbt_slab.h"
#include "host_items.h"

namespace hbt {

// Per-particle prefilter record: nrm = px^2+py^2 - mid/2 (so that
// n1 + n2 + 2 p1.p2 = T - mid), and blo, the lower end of the particle's
// beta interval [beta - cw, beta + cw] (beta = pz/E, or pz without the LCMS
// boost); the upper end bhi is kept in a separate array beside the records.
struct PfRec64 {
    double px, py, nrm, blo;
};
struct PfRec32 {
    float px, py, nrm, blo;
};

// Prefilter constants in the records' precision (FP64 or FP32 set of
// DevParams): the T window's centre and half-width, H^2, mid - H^2, and the
// q windows' margin c as -c and -2c.
template <typename Prec>
struct PfConst {
    Prec half, mid, h2, midh2, negc, neg2c;
};

template <typename Prec>
__host__ __device__ __forceinline__ PfConst<Prec> pf_const(const DevParams &P) {
    constexpr bool f64 = std::is_same<Prec, double>::value;
    PfConst<Prec> k;
    k.half = (Prec)(f64 ? P.pf_half : P.f_half);
    k.mid = (Prec)(f64 ? P.pf_mid : P.f_mid);
    k.h2 = (Prec)(f64 ? P.pf_h2 : P.f_h2);
    k.midh2 = (Prec)(f64 ? P.pf_midh2 : P.f_midh2);
    k.negc = (Prec)(f64 ? P.pf_negc : P.f_negc);
    k.neg2c = k.negc * (Prec)2;
    return k;
}

// The prefilter's tests of one pair (for a mixed pair, i rotated by
// -theta), as five comparisons lhs[k] <= rhs[k]: the K_T window |T - mid|
// <= half; q_out: dn^2 - c <= H^2 T; q_side: (Q - H^2) T - 2c <= dn^2 - c;
// beta: |beta_i - beta_j| <= cw_i + cw_j as the overlap of the two beta
// intervals, blo_i <= bhi_j and blo_j <= bhi_i (no arithmetic).  The kernel
// and the host's check of the margins evaluate this same function.
template <typename Prec>
struct PfTerms {
    Prec lhs[5], rhs[5];
};

template <typename Prec>
__host__ __device__ __forceinline__ PfTerms<Prec> pf_terms(Prec ix, Prec iy, Prec in, Prec ilo,
                                                           Prec ihi, Prec jx, Prec jy, Prec jn,
                                                           Prec jlo, Prec jhi,
                                                           const PfConst<Prec> &k) {
    PfTerms<Prec> t;
    const Prec d = fma(ix, jx, iy * jy);
    const Prec sn = in + jn;
    const Prec u = fma((Prec)2, d, sn);   // T - mid
    const Prec v = fma((Prec)-2, d, sn);  // Q - mid
    const Prec dn = in - jn;
    const Prec T = u + k.mid;
    const Prec ddc = fma(dn, dn, k.negc);
    t.lhs[0] = fabs(u);
    t.rhs[0] = k.half;
    t.lhs[1] = ddc;
    t.rhs[1] = T * k.h2;
    t.lhs[2] = fma(v + k.midh2, T, k.neg2c);
    t.rhs[2] = ddc;
    t.lhs[3] = ilo;
    t.rhs[3] = jhi;
    t.lhs[4] = jlo;
    t.rhs[4] = ihi;
    return t;
}

template <typename Prec>
__host__ __device__ __forceinline__ bool pf_pass(const PfTerms<Prec> &t) {
    return t.lhs[0] <= t.rhs[0] && t.lhs[1] <= t.rhs[1] && t.lhs[2] <= t.rhs[2] &&
           t.lhs[3] <= t.rhs[3] && t.lhs[4] <= t.rhs[4];
}

// Same-event pair orientation: for a position in the cell-sorted list, the
// particle's index in the fast list and its rank in gather order (the
// reference's particle 1 is the earlier of the two in gather order).
struct OrigRank {
    int idx;
    int rank;
};

// One analysis resident in HBM: its particles, mixing plan and histograms.
// A launch may carry items of many analyses (the request server batches
// them); every index in a WorkItem is relative to its own View.  The
// particle lists are the fast lists: per event, the particles whose |pT|
// can reach a K_T bin, sorted by |pT|.
struct View {
    DevHist H;
    const PartRec *A;    // same-event fast list, 64 B per particle
    const PartRec *B;    // mixed fast list (separate, or an alias of A)
    const void *pfA;     // PfRec64 (FP64 kernel) or PfRec32 (precision form) of A
    const void *pfB;     // of B
    const void *pfS;     // of the same-event fast list sorted by (px cell, py) per group
    const void *cwA;     // per-particle upper end bhi of the beta interval of A, like pfA
    const void *cwB;
    const void *cwS;
    const OrigRank *origS;     // cell-sorted position -> fast-list index and gather rank
    const PartRec *AS;         // records of the same-event fast list in the cell-sorted order
    const JWin *mwin;          // partner windows of the mixed items
    const double *plan_cs;     // cos, sin of each plan row's rotation
    const double *psi;         // psi_ref per group
    // the analysis' grid constants, for the exact stage: read from the view
    // table (constant address space) where they are used, so the kernel
    // neither holds them in SGPRs nor reloads spilled copies with VALU
    // instructions
    DevParams P;
};

// The cell-sorted copy of the same-event records, AS[q] = A[origS[q].idx],
// built on the device once the lists are resident.
static __global__ __launch_bounds__(256) void k_gather_srt(const PartRec *__restrict__ A,
                                                           const OrigRank *__restrict__ orr,
                                                           PartRec *__restrict__ AS,
                                                           long long nf) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nf) AS[q] = A[orr[q].idx];
}

static inline hipError_t gather_srt(const PartRec *A, const OrigRank *orr, PartRec *AS,
                                    long long nf, hipStream_t st) {
    if (nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_srt, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, A,
                       orr, AS, nf);
    return hipGetLastError();
}

// Work queues of a launch: at most one per XCD.  Queue q holds items
// [qoff[q], qoff[q+1]); its counter is counter[kQueueStride * q].
constexpr int kMaxQueues = 8;
constexpr int kQueueStride = 16;  // ints: one 64-byte line per counter

struct FastArgs {
    DevParams P;
    const View *views;
    const WorkItem *items;
    int nitems;
    int nqueues;                   // 1, 2, 4 or 8
    int qoff[kMaxQueues + 1];
    int *counter;              // work counters, zeroed before launch
    unsigned long long *binned;  // [2]: binned same, mixed (for stats)
    // slab binning (hbt_slab.h) of the common-case instances; vals == null:
    // every binned pair is added with device-scope atomics
    SlabArgs slab;
};

// The view table does not change during a launch.  Reading it through the
// constant address space marks its pointers as invariant scalar loads, which
// the compiler re-issues where they are used instead of holding ~40 SGPRs
// for the whole work item.
#if defined(__HIP_DEVICE_COMPILE__)
#define HBT_LOAD_VIEW(tab, k) (((const __attribute__((address_space(4))) View *)(tab))[k])
#else
#define HBT_LOAD_VIEW(tab, k) ((tab)[k])
#endif

// The view table entry k, read anew where this is called: the empty asm
// hides the table's address from the optimiser, so the loads that follow
// cannot be hoisted out of the caller's loop.  For the exact stage, whose
// grid constants and pointers would otherwise be loaded once per work item
// and held in (spilled) scalar registers across the whole j loop.
__device__ __forceinline__ View view_fresh(const View *views, int k) {
#if defined(__HIP_DEVICE_COMPILE__)
    const __attribute__((address_space(4))) View *t =
        (const __attribute__((address_space(4))) View *)views + k;
    __asm__ volatile("" : "+s"(t));
    return *t;
#else
    return views[k];
#endif
}

// ------------------------------------------------------- wave helpers

// Mask of lanes where a <= b (ordered: NaN compares false), as one SGPR pair.
__device__ __forceinline__ unsigned long long le_mask(float a, float b) {
    return __builtin_amdgcn_fcmpf(a, b, 5 /* FCMP_OLE */);
}
__device__ __forceinline__ unsigned long long le_mask(double a, double b) {
    return __builtin_amdgcn_fcmp(a, b, 5 /* FCMP_OLE */);
}
__host__ __device__ __forceinline__ float absv(float x) { return __builtin_fabsf(x); }
__host__ __device__ __forceinline__ double absv(double x) { return __builtin_fabs(x); }

// Value of lane t (t uniform) in every lane.
__device__ __forceinline__ float bcast(float v, int t) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), t));
}
__device__ __forceinline__ double bcast(double v, int t) {
    const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, t);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), t);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// Lanes 0..n-1 (n clamped to [0, 64]); 32-bit count.
__device__ __forceinline__ unsigned long long low_lanes32(int n) {
    return n <= 0 ? 0ull : (n >= 64 ? ~0ull : ((1ull << n) - 1ull));
}
// Lanes 0..n-1 (n clamped to [0, 64]).
__device__ __forceinline__ unsigned long long low_lanes(long long n) {
    return n <= 0 ? 0ull : (n >= 64 ? ~0ull : ((1ull << n) - 1ull));
}

// ---------------------------------------------------------------- drain
// Each lane takes one queued survivor (lane < cnt), runs the exact cascade.
template <typename Prec, bool EXT>
struct Exact;

template <bool EXT>
struct Exact<double, EXT> {
    // i, j are positions in the cell-sorted list; origS maps them to the
    // fast list, and the pair is evaluated in the reference's orientation
    // (particle 1 = the earlier in gather order).
    static __device__ __forceinline__ int same(const DevParams &P, const View &v, double psi,
                                               long long i, long long j, SameRec &rec) {
        const OrigRank oi = v.origS[i], oj = v.origS[j];
        const bool fwd = oi.rank < oj.rank;
        const PartRec p = v.A[fwd ? oi.idx : oj.idx], r = v.A[fwd ? oj.idx : oi.idx];
        return eval_same_f64<EXT>(P, v.H, psi, p.px, p.py, p.pz, p.E, p.x, p.y, p.z, p.t, r.px,
                                  r.py, r.pz, r.E, r.x, r.y, r.z, r.t, rec);
    }
    static __device__ __forceinline__ int mixed(const DevParams &P, const View &v, double psi,
                                                long long i, long long j, double c, double s,
                                                long &k) {
        const PartRec &p = v.A[i];
        const PartRec &r = v.B[j];
        const double px = r.px, py = r.py;
        return eval_mixed_f64<EXT>(P, v.H, psi, p.px, p.py, p.pz, p.E, px * c - py * s,
                                   px * s + py * c, r.pz, r.E, k);
    }
};

template <bool EXT>
struct Exact<float, EXT> {
    // The records come from the cell-sorted copy, so their loads do not
    // wait for the gather ranks; the ranks only orient the pair.
    static __device__ __forceinline__ int same(const DevParams &P, const View &v, double psi,
                                               long long i, long long j, SameRec &rec) {
        const int ri = v.origS[i].rank, rj = v.origS[j].rank;
        return eval_same_fast<EXT>(P, v.H, psi, v.AS[i], v.AS[j], rj < ri, rec);
    }
    static __device__ __forceinline__ int mixed(const DevParams &P, const View &v, double psi,
                                                long long i, long long j, double c, double s,
                                                long &k) {
        const PartRec &p = v.A[i];
        const PartRec &r = v.B[j];
        const double px = r.px, py = r.py;
        return eval_mixed_fast<EXT>(P, v.H, psi, p.px, p.py, p.pz, p.E, px * c - py * s,
                                    px * s + py * c, r.pz, r.E, k);
    }
};

// A binned pair's value as the form accumulates it: FP64 as computed, or,
// in the precision form with slab binning (whose records hold FP32 values),
// rounded to FP32 on the direct path too.
template <typename Prec>
__device__ __forceinline__ double fp_pair(double v, bool slab) {
    return std::is_same<Prec, float>::value && slab ? (double)(float)v : v;
}

// Per-wave LDS staging of binned same-event pairs for the cooperative
// scatter: bin index and the four sums of up to 64 pairs.
struct Stage {
    long k[64];
    double v[64][4];
};

// Each lane with lane < cnt takes one queued survivor and runs the exact
// cascade.  Same-event entries: (j - j0) << 8 | slot; mixed entries:
// (window - j0) << 26 | (j - window start) << 8 | slot, slot = r*64 + lane
// of the i particle in the tile.
// Binned same-event pairs are added cooperatively: they are staged in LDS,
// then groups of 5 lanes add one pair's count and four sums, so one atomic
// instruction touches 12 bins (12 cache lines) instead of 64.  Atomic
// throughput on this part is set by (instruction, cache line) accesses,
// about 2.4e10 per second whatever the histogram's size or placement, so
// the count is an FP64 add in the same instruction as the sums: one access
// per binned pair.
template <typename Prec, bool EXT, int QCAP>
__device__ __forceinline__ int drain(const DevParams &P, const View &v, const unsigned *q,
                                     Stage &stg,
                                     unsigned head, int cnt, const WorkItem &w, double psi,
                                     const SlabArgs &sl, unsigned *fill, const int *scap,
                                     const long long *sbase, unsigned &direct_n) {
    const int lane = __lane_id();
    int st = 0;
    SameRec rec;
    long kmix = 0;
    if (lane < cnt) {
        const unsigned ent = q[(head + lane) & (QCAP - 1)];
        const int slot = (int)(ent & 255u);
        if (slot < w.ni) {  // always true: only lanes inside the tile enqueue
            const long long i = w.i0 + slot;
            if (w.kind == 0) {
                st = Exact<Prec, EXT>::same(P, v, psi, i, w.j0 + (ent >> 8), rec);
            } else {
                const JWin jw = v.mwin[w.j0 + (ent >> 26)];
                const long long j = jw.b + ((ent >> 8) & 0x3ffffu);
                st = Exact<Prec, EXT>::mixed(P, v, psi, i, j, v.plan_cs[2 * jw.row],
                                             v.plan_cs[2 * jw.row + 1], kmix);
            }
        }
    }
    const bool binned = (st & ST_BIN) != 0;
    // slab binning: a same-event pair becomes a record of its (block, slab)
    // region; a full region, or stored values that would land in another
    // bin, send it down the direct path.  Mixed pairs are always direct.
    // (The host turns slab binning off for q_inv analyses.)
    const bool slab = sl.vals != nullptr;
    if (w.kind != 0) {
        if (binned) add_mixed_direct(v.H, kmix);
        return binned ? 1 : 0;
    }
    bool direct = binned;
    if (slab && binned) {
        direct = !slab_put_same<Prec>(sl, fill, scap, sbase, rec);
        direct_n += direct ? 1u : 0u;
    }
    const unsigned long long m = __ballot(direct);
    const int nb = __popcll(m);
    if (nb == 0) return binned ? 1 : 0;
    if (direct) {
        const unsigned idx = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        stg.k[idx] = rec.k;
        // with slab binning the precision form's per-pair values are FP32 on
        // either path (its records hold them as FP32), so the path a pair
        // takes does not change what it adds; without it they stay FP64
        stg.v[idx][0] = fp_pair<Prec>(rec.qo, slab);
        stg.v[idx][1] = fp_pair<Prec>(rec.qs, slab);
        stg.v[idx][2] = fp_pair<Prec>(rec.ql, slab);
        stg.v[idx][3] = fp_pair<Prec>(rec.c, slab);
    }
    __builtin_amdgcn_wave_barrier();
    // lane sub of a group adds field sub of its pair's bin: the count (an
    // exact FP64 integer) and the four sums go out as one FP64 atomic
    // instruction, one cache line per pair
    const int grp = lane / 5, sub = lane - 5 * grp;
    for (int base = 0; base < nb; base += 12) {
        const int p = base + grp;
        if (lane < 60 && p < nb)
            atomic_add_f64(v.H.bins + stg.k[p] * BIN_STRIDE + sub,
                           sub == 0 ? 1.0 : stg.v[p][max(sub - 1, 0)]);
    }
    __builtin_amdgcn_wave_barrier();
    return binned ? 1 : 0;
}

// XCD of the executing wave (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ int xcc_id() {
    return (int)(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7u);
}

// j records are read with scalar loads (uniform address, read-only data
// seen through the constant address space), so the j values arrive in SGPRs
// and cost no vector instructions to broadcast.
#if defined(__HIP_DEVICE_COMPILE__)
#define HBT_CONST(T, p) ((const __attribute__((address_space(4))) T *)(p))
#else
#define HBT_CONST(T, p) ((const T *)(p))
#endif

// Survivors queued between drains: fewer than 64 plus one unrolled step of
// JU j's times 64R lanes.
constexpr int JU = 4;
template <int R>
struct Ring {
    static constexpr int NEED = 64 + JU * 64 * R;
    static constexpr int CAP = NEED <= 256 ? 256 : (NEED <= 512 ? 512 : 1024);
};

// ---------------------------------------------------------------- kernel
// Persistent: each wave pulls work items from the queue of its XCD (the host
// puts all items of a group in one queue, so a group's particles stay in one
// XCD's L2), then from the next queues in turn, until all are empty.  Per item the wave holds 64R i particles (R per lane) and walks
// its j segments (same-event: one range with j > i in the cell-sorted list;
// mixed: one partner window per segment with the i particles rotated by
// -theta),
// JU j's per step.  Survivors of the prefilter go to the wave's LDS ring;
// whenever 64 are queued they are binned, one per lane, and the last
// survivors of the item are flushed at its final step (the exact cascade
// has a single call site).
// Occupancy: the exact stage is latency-bound (the particle records' loads
// and its FP64 chain), so the kernels trade a few spilled registers for
// more waves per SIMD: common case 6 at R=1 (80 VGPRs; the LDS of six
// blocks fills the CU's 160 KB) and 4 at R=2 (128 VGPRs).  At R=1, 6 waves
// ran the precision form's bench step 3 % faster than 5, and 5 ran it 5 %
// faster than 4.  The K_phi / q_inv
// instances carry the literal cascade with atan2 and the host-defer path
// inline (168 VGPRs unconstrained): the precision form's, which reaches it
// only for guard-band pairs, runs at 4 (128 VGPRs, spills in that rare
// path), 10 % faster than at 3 on the batched K_phi workload; the FP64
// form's, where it is the main path, at 3.  R=4 keeps the compiler's choice.
template <typename Prec, int R, bool EXT>
constexpr int fast_waves_per_eu() {
    return R == 4 ? 1
                  : (EXT ? (std::is_same<Prec, float>::value ? 4 : 3) : (R == 1 ? 6 : 4));
}

template <typename Prec, int R, bool EXT>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(fast_waves_per_eu<Prec, R, EXT>()))) void k_fast(FastArgs a) {
    using Rec = typename std::conditional<std::is_same<Prec, double>::value, PfRec64,
                                          PfRec32>::type;
    constexpr int QCAP = Ring<R>::CAP;
    __shared__ unsigned ring[4][QCAP];
    __shared__ Stage stage[4];
    // slab binning: records per slab of this block, and (common-case
    // instances) the regions' capacity and first record per slab; the K_phi
    // instance keeps its up to kMaxSlabsKphi fill counters in the launch's
    // dynamic LDS (none when it bins with atomics) and reads capacity and
    // first record from SlabArgs
    extern __shared__ unsigned slab_fill_dyn[];
    __shared__ unsigned slab_fill_st[EXT ? 1 : kMaxSlabs];
    unsigned *const slab_fill = EXT ? slab_fill_dyn : slab_fill_st;
    __shared__ int slab_cap[EXT ? 1 : kMaxSlabs];
    __shared__ long long slab_base[EXT ? 1 : kMaxSlabs];
    const bool slab = a.slab.vals != nullptr;
    if (slab) {
        for (int k = threadIdx.x; k < a.slab.S; k += blockDim.x) {
            slab_fill[k] = 0u;
            if (!EXT) {
                slab_cap[k] = a.slab.scap[k];
                slab_base[k] = a.slab.sbase[k];
            }
        }
        __syncthreads();
    }
    const int *scap = EXT ? a.slab.scap : slab_cap;
    const long long *sbase = EXT ? a.slab.sbase : slab_base;
    unsigned direct_n = 0;
    const int lane = __lane_id();
    unsigned *q = ring[threadIdx.x >> 6];
    Stage &stg = stage[threadIdx.x >> 6];
    const PfConst<Prec> pk = pf_const<Prec>(a.P);
    unsigned long long nb_same = 0, nb_mixed = 0;

    const int nq = a.nqueues;
    int wq = xcc_id() & (nq - 1);  // nq is a power of two
    int queues_left = nq;
    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(a.counter + kQueueStride * wq, 1);
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
        if (it >= a.qoff[wq + 1] - a.qoff[wq]) {  // this queue is empty: the next one
            if (--queues_left == 0) break;
            wq = (wq + 1) & (nq - 1);
            continue;
        }
        const WorkItem w = a.items[a.qoff[wq] + it];
        const View v = HBT_LOAD_VIEW(a.views, w.view);
        const double psi = v.psi[w.grp];
        const bool mixed = w.kind != 0;
        // same-event tiles run over the cell-sorted list, mixed ones over the
        // |pT|-sorted fast lists
        const Rec *__restrict__ pfi = (const Rec *)(mixed ? v.pfA : v.pfS);
        const Prec *__restrict__ hii = (const Prec *)(mixed ? v.cwA : v.cwS);
        const Rec *pfj = (const Rec *)(mixed ? v.pfB : v.pfS);
        const Prec *hij = (const Prec *)(mixed ? v.cwB : v.cwS);
        // i particles, R per lane
        Prec ix[R], iy[R], in[R], ilo[R], ihi[R], rx[R], ry[R];
        unsigned long long vmask[R];
        unsigned slot[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            slot[r] = (unsigned)(r * 64 + lane);
            vmask[r] = low_lanes((long long)w.ni - r * 64);
            const long long i = w.i0 + min((int)slot[r], w.ni - 1);
            const Rec p = pfi[i];
            ix[r] = rx[r] = p.px;
            iy[r] = ry[r] = p.py;
            in[r] = p.nrm;
            ilo[r] = p.blo;
            ihi[r] = hii[i];
        }
        const bool partial = w.ni < 64 * R;
        const int nseg = mixed ? w.nj : 1;
        unsigned head = 0, tail = 0;
        for (int seg = 0; seg < nseg; seg++) {
            long long pbase, jb, je;
            unsigned segbits;
            if (mixed) {
                const JWin jw = v.mwin[w.j0 + seg];
                pbase = jb = jw.b;
                je = jw.e;
                segbits = (unsigned)seg << 26;
                const double c = v.plan_cs[2 * jw.row], s = v.plan_cs[2 * jw.row + 1];
#pragma unroll
                for (int r = 0; r < R; r++) {  // rotate i by -theta
                    rx[r] = (Prec)((double)ix[r] * c + (double)iy[r] * s);
                    ry[r] = (Prec)((double)iy[r] * c - (double)ix[r] * s);
                }
            } else {
                pbase = jb = w.j0;
                je = w.j0 + w.nj;
                segbits = 0;
            }
            const bool last_seg = seg == nseg - 1;
            long long jc = jb;
            do {  // chunks of 64 j's (at least one pass, for the final flush)
                const int cnt = (int)max(0ll, min(64ll, je - jc));
                const bool last = last_seg && jc + 64 >= je;
                // same-event: lanes with i < j matter only where the chunk
                // reaches into the tile itself
                const bool diag = !mixed && jc < w.i0 + 64 * R;
                int tri = (int)min(jc - w.i0, 1ll << 20);
                unsigned ebase = segbits | ((unsigned)(jc - pbase) << 8);
                const auto *cpf = HBT_CONST(Rec, pfj + jc);
                const auto *chi = HBT_CONST(Prec, hij + jc);
                int t = 0;
                for (;;) {
                    if (tail - head >= 64 || (last && t >= cnt && tail != head)) {
                        const int n = (int)min(64u, tail - head);
                        __builtin_amdgcn_wave_barrier();
                        const View vd = view_fresh(a.views, w.view);
                        const int nb = drain<Prec, EXT, QCAP>(vd.P, vd, q, stg, head, n, w, psi,
                                                              a.slab, slab_fill, scap, sbase,
                                                              direct_n);
                        if (mixed)
                            nb_mixed += nb;
                        else
                            nb_same += nb;
                        head += n;
                        continue;
                    }
                    if (t >= cnt) break;
                    // one j against the wave's 64R i particles
                    auto pair_step = [&](int u, const Rec &jr, Prec jhi) {
#pragma unroll
                        for (int r = 0; r < R; r++) {
                            const PfTerms<Prec> pt =
                                pf_terms(rx[r], ry[r], in[r], ilo[r], ihi[r], jr.px, jr.py,
                                         jr.nrm, jr.blo, jhi, pk);
                            unsigned long long m = le_mask(pt.lhs[0], pt.rhs[0]) &
                                                   le_mask(pt.lhs[1], pt.rhs[1]) &
                                                   le_mask(pt.lhs[2], pt.rhs[2]) &
                                                   le_mask(pt.lhs[3], pt.rhs[3]) &
                                                   le_mask(pt.lhs[4], pt.rhs[4]);
                            if (partial) m &= vmask[r];
                            if (diag) m &= low_lanes32(tri + u - r * 64);
                            if (m) {
                                const unsigned pos = __builtin_amdgcn_mbcnt_hi(
                                    (unsigned)(m >> 32),
                                    __builtin_amdgcn_mbcnt_lo((unsigned)m, tail));
                                if (__builtin_amdgcn_inverse_ballot_w64(m))
                                    q[pos & (QCAP - 1)] = (ebase + ((unsigned)u << 8)) | slot[r];
                                tail += (unsigned)__popcll(m);
                            }
                        }
                    };
                    if (t + JU <= cnt) {  // whole step: the JU records load together
                        Rec jr[JU];
                        Prec jw[JU];
#pragma unroll
                        for (int u = 0; u < JU; u++) {
                            jr[u] = cpf[t + u];
                            jw[u] = chi[t + u];
                        }
#pragma unroll
                        for (int u = 0; u < JU; u++) pair_step(u, jr[u], jw[u]);
                    } else {
#pragma unroll
                        for (int u = 0; u < JU; u++)
                            if (t + u < cnt) pair_step(u, cpf[t + u], chi[t + u]);
                    }
                    t += JU;
                    tri += JU;
                    ebase += (unsigned)JU << 8;
                }
                jc += 64;
            } while (jc < je);
        }
        __builtin_amdgcn_wave_barrier();
    }
    // per-wave totals of binned pairs
    for (int off = 32; off > 0; off >>= 1) {
        nb_same += __shfl_down(nb_same, off);
        nb_mixed += __shfl_down(nb_mixed, off);
    }
    if (lane == 0) {
        if (nb_same) atomic_add_u64(&a.binned[0], nb_same);
        if (nb_mixed) atomic_add_u64(&a.binned[1], nb_mixed);
    }
    if (slab) {
        // the block's demand per slab, for k_slab_bin (which clamps it to the
        // regions' capacity) and for sizing later launches
        unsigned long long dn = direct_n;
        for (int off = 32; off > 0; off >>= 1) dn += __shfl_down(dn, off);
        if (lane == 0 && dn) atomic_add_u64(&a.slab.stats[a.slab.S], dn);
        __syncthreads();
        for (int k = threadIdx.x; k < a.slab.S; k += blockDim.x)
            a.slab.sfill[(long long)blockIdx.x * a.slab.S + k] = slab_fill[k];
    }
}

}  // namespace hbt
