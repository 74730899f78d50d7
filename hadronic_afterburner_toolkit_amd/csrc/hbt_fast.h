// hbt_fast.h -- the fast pair kernel: a cheap conservative prefilter over
// every enumerated pair, per-wave compaction of the survivors into an LDS
// ring, and the exact cascade on full waves of survivors.
//
// Why: the reference's cascade exits early for ~98% of pairs
// (HBT_correlation.cpp:319-398).  On a 64-wide wave, lanes that exit early
// idle while any lane continues, so a direct port pays the whole cascade
// (sqrt, three to five divisions, cos) for almost every pair.  Here every
// pair costs ~10 VALU operations; only survivors (~2-4%) reach the exact
// cascade, and they do so 64 at a time with all lanes busy.
//
// Prefilter (all conservative: a pair the exact cascade accepts always
// passes; the margins exceed the rounding error of both computations):
//   T = |p1T + p2T|^2 = 4 K_perp^2           in [4 KT_min^2, 4 KT_max^2]
//   Q = |p1T - p2T|^2 = q_out^2 + q_side^2   <= 2 H^2
//   |beta1 - beta2| <= (H/2) (1/E1 + 1/E2)   (from |q_long| <= H, M_t <= K_E,
//        q_long M_t = E2 pz1 - E1 pz2; beta = pz/E)
// with H = max(|q_lo_cut|, |q_hi_cut|).  T and Q come from one dot product:
// T = n1 + n2 + 2 p1.p2, Q = n1 + n2 - 2 p1.p2 (n = |pT|^2).
// For mixed pairs the i particle is rotated by -theta instead of rotating
// every j by +theta (T, Q and beta are rotation invariant); the exact
// cascade then uses the reference's rotated j (HBT_correlation.cpp:522-523).
#pragma once
#include "hbt_device.h"
#include "hbt_f32.h"

namespace hbt {

// Per-particle prefilter record: nrm = px^2+py^2 - mid/2 (so that
// n1 + n2 + 2 p1.p2 = T - mid), beta = pz/E (or pz without the LCMS boost).
struct PfRec64 {
    double px, py, nrm, beta;
};
struct PfRec32 {
    float px, py, nrm, beta;
};

// One unit of work for one wave.
struct WorkItem {
    long long i0;   // first i particle (index into the slot-0 cut arrays)
    long long j0;   // same-event: first j particle (slot-0 index)
    int ni;         // number of i particles (<= 64*R)
    int nj;         // same-event: number of j particles
    int kind;       // 0 same-event, 1 mixed
    int e;          // mixed: event index (its plan rows start at rowoff[e])
    int k0, k1;     // mixed: partner range [k0,k1) of the event's plan rows
    int grp;        // group (buffer) of the pair, selects psi_ref
    int pad;
};

struct FastArgs {
    DevParams P;
    DevHist H;
    DevSoA A;            // slot 0 FP64
    DevSoA B;            // mixed list FP64 (slot 1 or alias of slot 0)
    const void *pfA;     // PfRec64/PfRec32 of slot 0
    const void *pfB;     // of the mixed list
    const double *cwA, *cwB;  // per-particle c/E bound (FP64)
    const float *cwA32, *cwB32;
    const long long *mix_off;  // event offsets of the mixed list (cut)
    const long long *rowoff;   // plan rows of event e: [rowoff[e], rowoff[e+1])
    const int *plan_ids;       // partner event (global index in the mixed list)
    const double *plan_cs;     // cos, sin of the partner's rotation
    const double *psi;         // psi_ref per group
    const WorkItem *items;
    int nitems;
    int *counter;              // work counter, zeroed before launch
    unsigned long long *binned;  // [2]: binned same, mixed (for stats)
};

// LDS ring entries per wave (power of two).  Between drains the ring holds
// fewer than 64 entries, and one j step appends at most 64*R.
template <int R>
struct Ring {
    static constexpr int CAP = (64 * (R + 1) <= 256) ? 256 : 512;
};

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ---------------------------------------------------------------- drain
// Each lane takes one queued survivor (lane < cnt), runs the exact cascade.
template <typename Prec>
struct Exact;

template <>
struct Exact<double> {
    static __device__ __forceinline__ int same(const FastArgs &a, double psi, long long i,
                                               long long j) {
        const DevSoA &S = a.A;
        return bin_same_f64(a.P, a.H, psi, S.px[i], S.py[i], S.pz[i], S.E[i], S.x[i], S.y[i],
                            S.z[i], S.t[i], S.px[j], S.py[j], S.pz[j], S.E[j], S.x[j],
                            S.y[j], S.z[j], S.t[j]);
    }
    static __device__ __forceinline__ int mixed(const FastArgs &a, double psi, long long i,
                                                long long j, double c, double s) {
        const double px = a.B.px[j], py = a.B.py[j];
        return bin_mixed_f64(a.P, a.H, psi, a.A.px[i], a.A.py[i], a.A.pz[i], a.A.E[i],
                             px * c - py * s, px * s + py * c, a.B.pz[j], a.B.E[j]);
    }
};

template <>
struct Exact<float> {
    static __device__ __forceinline__ int same(const FastArgs &a, double psi, long long i,
                                               long long j) {
        return bin_same_fast(a.P, a.H, psi, a.A, i, j);
    }
    static __device__ __forceinline__ int mixed(const FastArgs &a, double psi, long long i,
                                                long long j, double c, double s) {
        return bin_mixed_fast(a.P, a.H, psi, a.A, a.B, i, j, c, s);
    }
};

template <typename Prec, int QCAP>
__device__ __forceinline__ int drain_same(const FastArgs &a, const unsigned *q, unsigned head,
                                          int cnt, long long i0, int ni, long long j0,
                                          double psi) {
    const int lane = __lane_id();
    int binned = 0;
    if (lane < cnt) {
        const unsigned e = q[(head + lane) & (QCAP - 1)];
        if ((int)(e & 255u) < ni) {  // always true: only lanes inside the tile enqueue
            const long long i = i0 + (e & 255u);
            const long long j = j0 + (e >> 8);
            binned = (Exact<Prec>::same(a, psi, i, j) & ST_BIN) ? 1 : 0;
        }
    }
    return binned;
}

template <typename Prec, int QCAP>
__device__ __forceinline__ int drain_mixed(const FastArgs &a, const unsigned *q, unsigned head,
                                           int cnt, long long i0, int ni, long long row0,
                                           double psi) {
    const int lane = __lane_id();
    int binned = 0;
    if (lane < cnt && (int)(q[(head + lane) & (QCAP - 1)] & 255u) < ni) {
        const unsigned ent = q[(head + lane) & (QCAP - 1)];
        const long long row = row0 + (long long)(ent >> 26);
        const int pid = a.plan_ids[row];
        const long long j = a.mix_off[pid] + ((ent >> 8) & 0x3ffffu);
        const double c = a.plan_cs[2 * row], s = a.plan_cs[2 * row + 1];
        binned = (Exact<Prec>::mixed(a, psi, i0 + (ent & 255u), j, c, s) & ST_BIN) ? 1 : 0;
    }
    return binned;
}

// ---------------------------------------------------------------- kernel
template <typename Prec, int R>
__global__ __launch_bounds__(256) void k_fast(FastArgs a) {
    using Rec = typename std::conditional<std::is_same<Prec, double>::value, PfRec64,
                                          PfRec32>::type;
    constexpr int QCAP = Ring<R>::CAP;
    __shared__ unsigned ring[4][QCAP];
    const int lane = __lane_id();
    const int wid = threadIdx.x >> 6;
    unsigned *q = ring[wid];
    const Rec *__restrict__ pfA = (const Rec *)a.pfA;
    const Rec *__restrict__ pfB = (const Rec *)a.pfB;
    const Prec *__restrict__ cwA;
    const Prec *__restrict__ cwB;
    if constexpr (std::is_same<Prec, double>::value) {
        cwA = a.cwA;
        cwB = a.cwB;
    } else {
        cwA = a.cwA32;
        cwB = a.cwB32;
    }
    const Prec two = (Prec)2;
    const Prec half_t = (Prec)(std::is_same<Prec, double>::value ? a.P.pf_half : a.P.f_half);
    const Prec q_thr = (Prec)(std::is_same<Prec, double>::value ? a.P.pf_qthr : a.P.f_qthr);
    unsigned long long nb_same = 0, nb_mixed = 0;

    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(a.counter, 1);
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
        if (it >= a.nitems) break;
        const WorkItem w = a.items[it];
        const double psi = a.psi[w.grp];
        // i particles held in registers, R per lane
        Prec ix[R], iy[R], in[R], ib[R], icw[R];
        long long ii[R];
        bool iv[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int s = r * 64 + lane;
            iv[r] = s < w.ni;
            ii[r] = w.i0 + (iv[r] ? s : 0);
            const Rec p = pfA[ii[r]];
            ix[r] = p.px;
            iy[r] = p.py;
            in[r] = p.nrm;
            ib[r] = p.beta;  // lanes past the tile (iv false) never pass
            icw[r] = cwA[ii[r]];
        }
        unsigned head = 0, tail = 0;
        if (w.kind == 0) {
            const long long i_last = w.i0 + w.ni - 1;
            for (int jj = 0; jj < w.nj; jj++) {
                const long long j = w.j0 + jj;
                const Rec pj = pfA[j];
                const Prec jcw = cwA[j];
                const bool tri = j <= i_last;  // uniform
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const Prec d = fma(ix[r], pj.px, iy[r] * pj.py);
                    const Prec sn = in[r] + pj.nrm;
                    const Prec u = fma(two, d, sn);
                    const Prec v = fma(-two, d, sn);
                    const Prec db = ib[r] - pj.beta;
                    const Prec rw = icw[r] + jcw;
                    bool pass = iv[r] & (fabs(u) <= half_t) & (v <= q_thr) & (fabs(db) <= rw);
                    if (tri) pass = pass & (j > ii[r]);
                    const unsigned long long m = __ballot(pass);
                    if (m) {
                        const unsigned pos = tail + (unsigned)__popcll(m & lanemask_lt());
                        if (pass) q[pos & (QCAP - 1)] = ((unsigned)jj << 8) | (unsigned)(r * 64 + lane);
                        tail += (unsigned)__popcll(m);
                    }
                }
                while (tail - head >= 64) {
                    __builtin_amdgcn_wave_barrier();
                    nb_same += drain_same<Prec, QCAP>(a, q, head, 64, w.i0, w.ni, w.j0, psi);
                    head += 64;
                }
            }
            __builtin_amdgcn_wave_barrier();
            while (tail != head) {
                const int cnt = (int)min(64u, tail - head);
                nb_same += drain_same<Prec, QCAP>(a, q, head, cnt, w.i0, w.ni, w.j0, psi);
                head += cnt;
            }
        } else {
            const long long row0 = a.rowoff[w.e] + w.k0;
            for (int k = w.k0; k < w.k1; k++) {
                const long long row = a.rowoff[w.e] + k;
                const int pid = a.plan_ids[row];
                const double c = a.plan_cs[2 * row], s = a.plan_cs[2 * row + 1];
                const long long jb = a.mix_off[pid], je = a.mix_off[pid + 1];
                // rotate the lane's i particles by -theta
                Prec rx[R], ry[R];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    rx[r] = (Prec)((double)ix[r] * c + (double)iy[r] * s);
                    ry[r] = (Prec)((double)iy[r] * c - (double)ix[r] * s);
                }
                const unsigned kk = (unsigned)(k - w.k0) << 26;
                for (long long j = jb; j < je; j++) {
                    const Rec pj = pfB[j];
                    const Prec jcw = cwB[j];
                    const unsigned jo = (unsigned)(j - jb) << 8;
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const Prec d = fma(rx[r], pj.px, ry[r] * pj.py);
                        const Prec sn = in[r] + pj.nrm;
                        const Prec u = fma(two, d, sn);
                        const Prec v = fma(-two, d, sn);
                        const Prec db = ib[r] - pj.beta;
                        const Prec rw = icw[r] + jcw;
                        const bool pass =
                            iv[r] & (fabs(u) <= half_t) & (v <= q_thr) & (fabs(db) <= rw);
                        const unsigned long long m = __ballot(pass);
                        if (m) {
                            const unsigned pos = tail + (unsigned)__popcll(m & lanemask_lt());
                            if (pass) q[pos & (QCAP - 1)] = kk | jo | (unsigned)(r * 64 + lane);
                            tail += (unsigned)__popcll(m);
                        }
                    }
                    while (tail - head >= 64) {
                        __builtin_amdgcn_wave_barrier();
                        nb_mixed += drain_mixed<Prec, QCAP>(a, q, head, 64, w.i0, w.ni, row0, psi);
                        head += 64;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            while (tail != head) {
                const int cnt = (int)min(64u, tail - head);
                nb_mixed += drain_mixed<Prec, QCAP>(a, q, head, cnt, w.i0, w.ni, row0, psi);
                head += cnt;
            }
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
}

}  // namespace hbt
