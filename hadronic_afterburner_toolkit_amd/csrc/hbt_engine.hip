// hbt_engine.hip -- the HBT pair-correlation engine for gfx950 and its C ABI
// (include/hbt_hip.h).
//
// Data layout in HBM, per context:
//   slot s (0 = same-event list, 1 = mixed-event list when it is read
//   separately): G groups (the reference's particle buffers) resident at
//   once.  Per slot: the particles kept by the rapidity cut in the
//   reference's gather order, as eight FP64 arrays px,py,pz,E,x,y,z,t
//   (64 B/particle); event offsets (int64, nev+1); group -> first event
//   (int32, G+1); the fast kernel's prefilter records (PfRec64 32 B + cw
//   8 B, PfRec32 16 B + cw 4 B per particle).
//   Mixing plan: per event e, rows [rowoff[e], rowoff[e+1]) each holding the
//   partner event (global index in the mixed list) and cos, sin of its
//   rotation.  Histograms: flat [kt][kphi][iqout][iqside][iqlong] (iqlong
//   fastest), u64 counts (numerator, denominator) and FP64 sums (q_out,
//   q_side, q_long, cos); q_inv histograms [kt][iq].
//
// Kernels:
//   k_same_plain / k_mixed_plain  LDS-tiled direct cascade (plain mode)
//   k_fast<double|float, R>       persistent work-queue kernel: prefilter +
//                                 compaction + exact cascade (hbt_fast.h);
//                                 one launch covers every same-event and
//                                 mixed pair of all resident groups
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/hbt_hip.h"
#include "hbt_device.h"
#include "hbt_cap.h"
#include "hbt_fast.h"
#include "hbt_internal.h"
#include "hbt_plain.h"

using namespace hbt;

namespace {

// ------------------------------------------------------------------ kernels

// Same-event pairs i<j of one group [g0, g0+n): block (tj, ti) of 256x256
// tiles, j tile staged in LDS, one thread per i (HBT_correlation.cpp:291-460).
__global__ __launch_bounds__(TILE) void k_same_plain(DevParams P, DevHist H, DevSoA A,
                                                     long long g0, int n, double psi) {
    same_plain_tile(P, H, A, g0, n, psi, blockIdx.y, blockIdx.x);
}

// Mixed pairs of the events [e0, e0 + nz/M) of one group whose plan has M
// rows per event: block (tj, ti, z) with event e0 + z/M, row row0 + z.  The
// partner's j tile is rotated by (c, s) while staged
// (HBT_correlation.cpp:522-523).
__global__ __launch_bounds__(TILE) void k_mixed_plain(DevParams P, DevHist H, DevSoA A,
                                                      DevSoA B, const long long *a_off,
                                                      const long long *b_off, const int *ids,
                                                      const double *cs, int M, int e0,
                                                      long long row0, double psi) {
    mixed_plain_tile(P, H, A, B, a_off, b_off, ids, cs, M, e0, row0, psi, blockIdx.y,
                     blockIdx.x, blockIdx.z);
}

// Pair-cap passes (hbt_cap.h) with the plain kernels' tiling: COUNT adds
// arrivals per (key, row), otherwise admitted pairs are binned.  Rows are
// particle indices of the same-event list; same-event positions are j.
template <bool COUNT>
__global__ __launch_bounds__(TILE) void k_same_cap(DevParams P, DevHist H, DevSoA A, CapView C,
                                                   long long g0, int n, double psi) {
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj < ti) return;
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
    unsigned c_bin = 0;
    if (i < n) {
        const long long gi = g0 + i;
        const double p1x = A.px[gi], p1y = A.py[gi], p1z = A.pz[gi], p1E = A.E[gi];
        const double x1 = A.x[gi], y1 = A.y[gi], z1 = A.z[gi], t1 = A.t[gi];
        const int jstart = (ti == tj) ? tid + 1 : 0;
        for (int jj = jstart; jj < jn; jj++) {
            const int st = cap_same_pair<COUNT>(P, H, C, psi, gi, g0 + j0 + jj, p1x, p1y, p1z,
                                                p1E, x1, y1, z1, t1, s[0][jj], s[1][jj],
                                                s[2][jj], s[3][jj], s[4][jj], s[5][jj],
                                                s[6][jj], s[7][jj]);
            c_bin += (st >> 3) & 1;
        }
    }
    if (!COUNT) wave_add_stage(&H.stage[4], c_bin);
}

// Mixed positions: (plan row k of the event) << 32 | j in the partner.
template <bool COUNT>
__global__ __launch_bounds__(TILE) void k_mixed_cap(DevParams P, DevHist H, DevSoA A, DevSoA B,
                                                    CapView C, const long long *a_off,
                                                    const long long *b_off, const int *ids,
                                                    const double *cs, int M, int e0,
                                                    long long row0, double psi) {
    const long long row = row0 + blockIdx.z;
    const long long k = blockIdx.z % M;
    const int e = e0 + (int)(blockIdx.z / M);
    const long long ib = a_off[e];
    const int n1 = (int)(a_off[e + 1] - ib);
    const int pid = ids[row];
    const long long jb = b_off[pid];
    const int n2 = (int)(b_off[pid + 1] - jb);
    const int ti = blockIdx.y, tj = blockIdx.x;
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
    unsigned c_bin = 0;
    if (i < n1) {
        const long long gi = ib + i;
        const double p1x = A.px[gi], p1y = A.py[gi], p1z = A.pz[gi], p1E = A.E[gi];
        for (int jj = 0; jj < jn; jj++) {
            const int st = cap_mixed_pair<COUNT>(P, H, C, psi, gi, (k << 32) | (j0 + jj), p1x,
                                                 p1y, p1z, p1E, s[0][jj], s[1][jj], s[2][jj],
                                                 s[3][jj]);
            c_bin += (st >> 3) & 1;
        }
    }
    if (!COUNT) wave_add_stage(&H.stage[9], c_bin);
}

// The work counters of a launch's queues (one per 64-byte line).
__global__ void k_zero_counters(int *p, int nq) {
    if ((int)threadIdx.x < nq) p[hbt::kQueueStride * threadIdx.x] = 0;
}

// Arrivals of host-decided pairs (K_phi at a bin edge) at the 3D cap.
__global__ void k_cap_add_arrivals(unsigned *cnt, const long long *idx, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[idx[i]], 1u);
}

// Adds the host-decided pairs to their histograms (a handful per call).
__global__ void k_apply_defer(const DeferAdd *a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DeferAdd d = a[i];
    DevHist H{};
    H.bins = d.bins;
    if (d.kind == 0) {
        const SameRec r{d.k, d.v[1], d.v[2], d.v[3], d.v[4]};
        add_same_direct(H, r);
    } else {
        add_mixed_direct(H, d.k);
    }
}

__global__ void k_zero_words(unsigned long long *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Sum of a u64 field over the q bins of each (kt, kphi): the reference's
// number_of_pairs_*_KTdiff / _KTKphidiff counters.  Element k of group g
// is h[(g*per + k)*stride].
// T = double for the count field (integers held exactly below 2^53).
template <typename T>
__global__ void k_sum_groups(const T *h, long per, int stride, unsigned long long *out) {
    const int g = blockIdx.x;
    unsigned long long acc = 0;
    for (long k = threadIdx.x; k < per; k += blockDim.x)
        acc += (unsigned long long)h[((long)g * per + k) * stride];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    __shared__ unsigned long long part[16];
    if (__lane_id() == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        out[g] = t;
    }
}

// ------------------------------------------------------------------ host

// Fast-kernel records of one ordering of a slot (RecordSet on the device).
struct DevRecords {
    PartRec *rec = nullptr;
    PfRec64 *pf64 = nullptr;
    PfRec32 *pf32 = nullptr;
    double *cw64 = nullptr;
    float *cw32 = nullptr;
};

struct Slot {
    int ngroups = 0, nev = 0;
    long long n = 0;        // kept particles
    long long n_in = 0;     // particles passed in
    std::vector<int> grp_ev;       // ngroups+1
    std::vector<long long> off;    // nev+1, kept-particle offsets
    std::vector<double> psi2;      // psi_ref (n=2) per group, over all particles passed in
    double *d_soa[8] = {};   // plain kernels (structure of arrays)
    long long *d_off = nullptr;
    // fast kernel: the fast list (particles whose |pT| can reach a K_T bin,
    // |pT|-sorted per event) and its (px cell, py)-sorted copy per group
    long long nf = 0;
    std::vector<long long> foff;  // nev+1 fast-list offsets
    hvec<double> f_pt;            // |pT| of the fast list (mixed windows)
    DevRecords ev, srt;
    OrigRank *d_orig = nullptr;   // srt position -> fast-list index, gather rank
    size_t cap = 0, cap_f = 0, cap_ev = 0;
    double pmax2 = 0.0;     // max px^2+py^2 of kept particles
    hvec<int> srt_cell;           // px cell and py of the sorted copy (work-item windows)
    hvec<double> srt_py;
    bool valid = false;
};

}  // namespace

struct hbt_ctx {
    hbt_config cfg;
    Derived D;  // constants derived from cfg
    DevParams P;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    char *h_stage = nullptr;   // pinned staging of uploads: columns (grown on demand)
    size_t stage_cap = 0;
    char *h_stage2 = nullptr;  // and fast-kernel records
    size_t stage2_cap = 0;
    int cu_count = 0;
    long nbins = 0, nq = 0;
    double *d_bins = nullptr;  // [nbins][BIN_STRIDE]
    unsigned long long *d_qinv_count = nullptr, *d_qinv_den = nullptr;
    double *d_qinv_sum = nullptr, *d_qinv_num = nullptr;
    unsigned long long *d_stage = nullptr;   // [10]
    unsigned long long *d_binned = nullptr;  // [2]
    DeferRec *d_defer = nullptr;             // [kDeferCap] pairs left to the host
    unsigned *d_ndefer = nullptr;
    bool batch_capped = false;               // the last hbt_launch_batch took the cap path
    uint64_t deferred_total = 0;             // pairs decided on the host so far
    unsigned long long *d_sums = nullptr;    // scratch for counters
    Slot slot[2];
    // mixing plan (device), and its host copy for the plain kernels
    int *d_ids = nullptr;
    double *d_cs = nullptr;
    long long *d_rowoff = nullptr;
    size_t plan_cap = 0, rowoff_cap = 0;
    std::vector<long long> rowoff;
    std::vector<int> plan_gids;  // partner of each plan row (index in the mixed list)
    std::vector<int> M_grp;
    bool plan_valid = false;
    double *d_psi = nullptr;
    size_t psi_cap = 0;
    std::vector<double> psi;  // per group (set from the upload or by the caller)
    WorkItem *d_items = nullptr;
    size_t items_cap = 0;
    JWin *d_wins = nullptr;  // partner windows of the mixed items
    size_t wins_cap = 0;
    int nitems = 0;
    int nqueues = 1;                          // work queues of the uploaded items
    std::vector<int> qoff = std::vector<int>(2, 0);
    int *d_counter = nullptr;
    View *d_view = nullptr;  // the context's one analysis view
    // ordered pair-cap path (allocated when a cap binds)
    unsigned *d_cnt = nullptr;  // [nkey3 + n_KT][nrows] arrivals
    size_t cnt_cap = 0;
    CapCut *d_cut = nullptr;    // [nkey3 + n_KT]
    unsigned long long *d_lim = nullptr;  // [nkey3 + n_KT] remaining capacity
    size_t cut_cap = 0, lim_cap = 0;
    long long *d_gpoff = nullptr;  // group particle offsets of slot 0
    size_t gpoff_cap = 0;
    // slab binning of the fast kernel (hbt_slab.h)
    struct {
        bool on = false;            // the next fast launch bins through slabs
        int S = 0, SB = 0, nunits = 0, nblk = 0;
        std::vector<int> scap, uoff;
        std::vector<long long> sbase;
        std::vector<int4> units;    // k_slab_bin blocks: (slab, first region, end region)
        // demand measured on the last launch: per slab total and largest
        // per block, with that launch's same-event pairs and blocks
        std::vector<double> tot_s, max_s;
        double hist_enum_s = 0;
        int hist_nblk = 0;
        bool have_hist = false, pending = false, resize = false;
        uint64_t direct = 0;        // binned same-event pairs of the last launch added directly
        char *d_vals = nullptr;
        long long *d_sbase = nullptr;
        int *d_scap = nullptr, *d_uoff = nullptr;
        int4 *d_units = nullptr;
        unsigned *d_sfill = nullptr;
        unsigned long long *d_stats = nullptr;
        SlabPart *d_parts = nullptr;
        size_t vals_cap = 0, base_cap = 0, fill_cap = 0, stats_cap = 0, parts_cap = 0,
               scap_cap = 0, units_cap = 0, uoff_cap = 0;
    } slab;
    uint64_t launch_enum_same = 0, launch_enum_mixed = 0;  // pairs of the items uploaded last
    FastArgs batch_args{};  // arguments of the prepared batch's fast launch
    DeferKey *d_dkey = nullptr;    // deferred pairs of a capped call, by (row, pos)
    long long *d_darr = nullptr;   // and their arrivals (key * nrows + row)
    size_t dkey_cap = 0, darr_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool batch_ready = false;
    bool batch_capture = true;   // the prepared batch is captured as a graph (launched more than once)
    // histogram privatisation of the fast kernel: slab binning (true) or
    // device-scope atomics; cfg.binning 1 / 0, or hbt_autotune's pick for 2
    bool slab_binning = false;
    std::string err;
    uint64_t needed = 0, needed_qinv = 0;
    uint64_t enum_same_total = 0, enum_mixed_total = 0;  // since reset
    uint64_t batch_enum_same = 0, batch_enum_mixed = 0, plan_enum_mixed = 0;
    float ms_same = 0.f, ms_mixed = 0.f, ms_batch = 0.f;
    double c_beta = 0.0, H = 0.0, mid = 0.0;
    hbt_tuning tune{2, 1024, 8, 8, 0.0};
};

struct hbt_rng {
    std::mt19937 gen;
    std::uniform_real_distribution<double> uni{0.0, 1.0};
    std::uniform_int_distribution<int> uint_{};
};

namespace {

std::string g_create_err;

int fail(hbt_ctx *c, int code, const std::string &msg) {
    if (c)
        c->err = msg;
    else
        g_create_err = msg;
    return code;
}

#define HIPCHECK(ctx, call)                                                              \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(ctx, HBT_ERR_HIP,                                                \
                        std::string(#call) + ": " + hipGetErrorString(e_));              \
    } while (0)

// Host-side copies into the context's device buffers are issued on the
// null stream, which does not wait for the context's (non-blocking) stream:
// every entry point that rewrites such a buffer first waits for the work
// queued there, e.g. a batch launched without hbt_sync.
int quiesce(hbt_ctx *c) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return HBT_OK;
}

template <typename T>
int dev_alloc(hbt_ctx *c, T **p, size_t n) {
    if (*p) hipFree(*p);
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, HBT_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return HBT_OK;
}

void free_slot(Slot &s) {
    for (auto &p : s.d_soa) {
        if (p) hipFree(p);
        p = nullptr;
    }
    if (s.d_off) hipFree(s.d_off);
    for (DevRecords *r : {&s.ev, &s.srt}) {
        void *ptrs[] = {r->rec, r->pf64, r->pf32, r->cw64, r->cw32};
        for (void *q : ptrs)
            if (q) hipFree(q);
    }
    if (s.d_orig) hipFree(s.d_orig);
    s = Slot();
}

void drop_graph(hbt_ctx *c) {
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->batch_ready = false;
}

DevHist hist_view(hbt_ctx *c) {
    DevHist h{};
    h.bins = c->d_bins;
    h.qinv_count = c->d_qinv_count;
    h.qinv_den = c->d_qinv_den;
    h.qinv_sum = c->d_qinv_sum;
    h.qinv_num = c->d_qinv_num;
    h.stage = c->d_stage;
    h.defer = c->d_defer;
    h.ndefer = c->d_ndefer;
    h.defer_cap = kDeferCap;
    return h;
}

DevSoA soa_view(const Slot &s) {
    DevSoA a;
    a.px = s.d_soa[0];
    a.py = s.d_soa[1];
    a.pz = s.d_soa[2];
    a.E = s.d_soa[3];
    a.x = s.d_soa[4];
    a.y = s.d_soa[5];
    a.z = s.d_soa[6];
    a.t = s.d_soa[7];
    return a;
}

const Slot &mixed_slot(const hbt_ctx *c) { return c->slot[1].valid ? c->slot[1] : c->slot[0]; }

int zero_hist(hbt_ctx *c) {
    const size_t b8 = (size_t)c->nbins * 8, q8 = (size_t)c->nq * 8;
    HIPCHECK(c, hipMemsetAsync(c->d_bins, 0, b8 * BIN_STRIDE, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_qinv_count, 0, q8, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_qinv_den, 0, q8, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_qinv_sum, 0, q8, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_qinv_num, 0, q8, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_stage, 0, 10 * 8, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_binned, 0, 2 * 8, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_ndefer, 0, sizeof(unsigned), c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->enum_same_total = c->enum_mixed_total = 0;
    return HBT_OK;
}

// Pairs the device left to the host (K_phi within kphi_guard of a bin edge):
// the host runs the same FP64 cascade with glibc's atan2 (kphi_index's host
// build), as the reference does, and adds each pair it bins to the device
// histogram.  Called after the stream is idle, before results are read.  A
// capped run cannot place such a pair in the cap's order and fails instead.
int resolve_deferred(hbt_ctx *c) {
    if (!c->P.azimuthal) return HBT_OK;  // only K_phi bins defer pairs
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    unsigned n = 0;
    HIPCHECK(c, hipMemcpy(&n, c->d_ndefer, sizeof(n), hipMemcpyDeviceToHost));
    if (n == 0) return HBT_OK;
    if (n > kDeferCap)
        return fail(c, HBT_ERR_CAPACITY, "more K_phi bin-edge pairs than the deferral buffer");
    std::vector<DeferRec> rec(n);
    HIPCHECK(c, hipMemcpy(rec.data(), c->d_defer, n * sizeof(DeferRec), hipMemcpyDeviceToHost));
    c->deferred_total += n;
    std::vector<DeferAdd> adds;
    resolve_records(c->P, rec.data(), n, adds);
    HIPCHECK(c, hipMemsetAsync(c->d_ndefer, 0, sizeof(unsigned), c->stream));
    if (!adds.empty()) {
        DeferAdd *d = nullptr;
        HIPCHECK(c, hipMalloc((void **)&d, adds.size() * sizeof(DeferAdd)));
        hipError_t e = hipMemcpyAsync(d, adds.data(), adds.size() * sizeof(DeferAdd),
                                      hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            const int nb = (int)((adds.size() + 255) / 256);
            hipLaunchKernelGGL(k_apply_defer, dim3(nb), dim3(256), 0, c->stream, d,
                               (int)adds.size());
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        hipFree(d);
        HIPCHECK(c, e);
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return HBT_OK;
}

// Per-(kt,kphi) sums of a u64 histogram field into host memory.
template <typename T>
int group_sums(hbt_ctx *c, const T *h, long per, int groups, int stride,
               std::vector<unsigned long long> &out) {
    out.assign(groups, 0);
    if (groups <= 0) return HBT_OK;
    hipLaunchKernelGGL(k_sum_groups<T>, dim3(groups), dim3(1024), 0, c->stream, h, per, stride,
                       c->d_sums);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(out.data(), c->d_sums, groups * 8, hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return HBT_OK;
}

// The pair cap (HBT_correlation.cpp:402-406, :651-655, q_inv :340-341,
// :600-603) admits pairs in enumeration order while a counter is at most
// needed_number_of_pairs.  The fast and plain kernels bin without ordering,
// which is exact as long as no counter can pass the cap inside this call;
// binds is set when one might, and the call then runs the ordered path
// (run_capped).
int check_cap(hbt_ctx *c, bool mixed, uint64_t enum_call, bool &binds) {
    binds = false;
    const uint64_t total = (mixed ? c->enum_mixed_total : c->enum_same_total) + enum_call;
    const bool qinv = c->cfg.invariant_radius_flag == 1;
    if (total <= c->needed + 1 && (!qinv || total <= c->needed_qinv)) return HBT_OK;
    const int groups = c->cfg.n_KT * c->P.nphi;
    std::vector<unsigned long long> cnt;
    const long per = (long)c->P.qnpts * c->P.qnpts * c->P.qnpts;
    int rc = mixed ? group_sums(c, (const unsigned long long *)(c->d_bins + BIN_DEN), per, groups,
                                BIN_STRIDE, cnt)
                   : group_sums(c, (const double *)(c->d_bins + BIN_COUNT), per, groups,
                                BIN_STRIDE, cnt);
    if (rc) return rc;
    for (auto v : cnt)
        if (v + enum_call > c->needed + 1) binds = true;
    if (qinv) {
        std::vector<unsigned long long> qc;
        rc = group_sums(c, mixed ? c->d_qinv_den : c->d_qinv_count, c->P.qnpts, c->cfg.n_KT, 1,
                        qc);
        if (rc) return rc;
        for (auto v : qc)
            if (v + enum_call > c->needed_qinv) binds = true;
    }
    return HBT_OK;
}

template <typename T>
int ensure(hbt_ctx *c, T **p, size_t &cap, size_t n) {
    if (n <= cap && *p) return HBT_OK;
    int rc = dev_alloc(c, p, n);
    if (rc) return rc;
    cap = n;
    return HBT_OK;
}

void set_prefilter_margins(hbt_ctx *c) {
    prefilter_margins(c->D, std::max(c->slot[0].pmax2, mixed_slot(c).pmax2), c->P);
}

int upload_psi(hbt_ctx *c) {
    const size_t G = std::max<size_t>(1, c->psi.size());
    int rc = ensure(c, &c->d_psi, c->psi_cap, G);
    if (rc) return rc;
    std::vector<double> p(c->psi);
    if (p.empty()) p.push_back(0.0);
    HIPCHECK(c, hipMemcpy(c->d_psi, p.data(), p.size() * 8, hipMemcpyHostToDevice));
    return HBT_OK;
}

View make_view(hbt_ctx *c) {
    View v{};
    v.H = hist_view(c);
    const Slot &s0 = c->slot[0];
    const Slot &s1 = mixed_slot(c);
    v.A = s0.ev.rec;
    v.B = s1.ev.rec;
    const bool f32 = c->cfg.precision == HBT_PREC_FP32C;
    auto pf = [&](const DevRecords &r) {
        return f32 ? (const void *)r.pf32 : (const void *)r.pf64;
    };
    auto cw = [&](const DevRecords &r) {
        return f32 ? (const void *)r.cw32 : (const void *)r.cw64;
    };
    v.pfA = pf(s0.ev);
    v.pfB = pf(s1.ev);
    v.pfS = pf(s0.srt);
    v.cwA = cw(s0.ev);
    v.cwB = cw(s1.ev);
    v.cwS = cw(s0.srt);
    v.origS = s0.d_orig;
    v.AS = s0.srt.rec;
    v.mwin = c->d_wins;
    v.plan_cs = c->d_cs;
    v.psi = c->d_psi;
    v.P = c->P;
    return v;
}

// Kernel arguments of the context's own launches: one view (index 0),
// copied to the device before the launch (and before a graph capture).
long fast_blocks(const hbt_ctx *c, int nitems) {
    return std::max(1l, std::min<long>((long)c->cu_count * c->tune.blocks_per_cu,
                                       ((long)nitems + 3) / 4));
}

// Reads the demand the last slab-binned launch measured (after the stream
// has drained); a launch that had to add more than 0.2 % of its pairs
// directly marks the regions for resizing.
int slab_read_stats(hbt_ctx *c) {
    auto &sl = c->slab;
    if (!sl.pending) return HBT_OK;
    sl.pending = false;
    const int S = sl.S;
    std::vector<unsigned long long> st((size_t)2 * S + 1);
    HIPCHECK(c, hipMemcpy(st.data(), sl.d_stats, st.size() * 8, hipMemcpyDeviceToHost));
    sl.tot_s.assign(S, 0);
    sl.max_s.assign(S, 0);
    double tot = 0;
    for (int k = 0; k < S; k++) {
        sl.tot_s[k] = (double)st[k];
        sl.max_s[k] = (double)st[S + 1 + k];
        tot += sl.tot_s[k];
    }
    sl.direct = st[S];
    sl.hist_enum_s = (double)c->launch_enum_same;
    sl.hist_nblk = sl.nblk;
    sl.have_hist = true;
    if ((double)sl.direct > 0.002 * tot) sl.resize = true;
    return HBT_OK;
}

// Sizes the (block, slab) regions of the next fast launch, on the host and
// outside any graph capture.  Capacity per region: the largest demand one
// block had for the slab on the last launch, scaled to this launch's pairs
// and blocks, with a margin; with no history, a uniform share of a generous
// binned fraction.  The record buffer is capped at kSlabBudget bytes:
// beyond, regions shrink and more pairs take the direct path.
constexpr double kSlabBudget = 16e9;

// Slab binning needs a histogram without the q_inv branch whose slabs (K_T
// x K_phi x q_out planes) fit the fast kernel's LDS fill counters and whose
// q_side x q_long planes fit k_slab_bin's LDS.
bool slab_applies(const hbt_ctx *c) {
    const int q = c->P.qnpts;
    return c->cfg.kernel == HBT_KERNEL_FAST && !c->P.qinv &&
           c->cfg.n_KT * c->P.nphi * q <= (c->P.azimuthal ? kMaxSlabsKphi : kMaxSlabs) &&
           q * q <= kMaxSlabBins;
}

// k_slab_bin's units: about four per resident block (two 72 KB blocks per
// CU), dealt to slabs in proportion to their demand on the last launch (a
// uniform share without one), each slab's regions split evenly among its
// units, and at most kMaxUnitRecords records of capacity in a unit.
void slab_units(hbt_ctx *c) {
    auto &sl = c->slab;
    const int S = sl.S, nblk = sl.nblk;
    const double target = 4.0 * 2 * std::max(1, c->cu_count);
    double tot = 0;
    if (sl.have_hist)
        for (int k = 0; k < S; k++) tot += sl.tot_s[k];
    sl.units.clear();
    sl.uoff.assign(S + 1, 0);
    for (int k = 0; k < S; k++) {
        const double w = tot > 0 ? sl.tot_s[k] / tot : 1.0 / S;
        const double need = std::ceil((double)nblk * sl.scap[k] / kMaxUnitRecords);
        const int n = (int)std::min((double)nblk, std::max({1.0, need, std::round(target * w)}));
        sl.uoff[k] = (int)sl.units.size();
        for (int u = 0; u < n; u++)
            sl.units.push_back(int4{k, (int)((long long)u * nblk / n),
                                    (int)((long long)(u + 1) * nblk / n), 0});
    }
    sl.uoff[S] = (int)sl.units.size();
    sl.nunits = (int)sl.units.size();
}

int slab_prepare(hbt_ctx *c, int nitems) {
    auto &sl = c->slab;
    const int q = c->P.qnpts;
    const int S = c->cfg.n_KT * c->P.nphi * q, SB = q * q;
    sl.on = c->slab_binning && slab_applies(c) && nitems > 0;
    if (!sl.on) return HBT_OK;
    const int nblk = (int)fast_blocks(c, nitems);
    const bool f32 = c->cfg.precision == HBT_PREC_FP32C;
    const size_t vsz = f32 ? 16 : 32;
    if (sl.S != S) sl.have_hist = false;
    sl.S = S;
    sl.SB = SB;
    sl.nblk = nblk;
    sl.resize = false;
    sl.scap.assign(S, 0);
    const double es = (double)c->launch_enum_same;
    for (int k = 0; k < S; k++) {
        double xs;
        if (sl.have_hist) {
            const double bs = (double)sl.hist_nblk / nblk;
            xs = sl.hist_enum_s > 0 ? sl.max_s[k] * es / sl.hist_enum_s * bs : 0.0;
            // a slab the last launch never reached may still get pairs now
            xs = std::max(xs, 0.002 * es / S / nblk);
        } else {
            xs = 3.0 * 0.03 * es / S / nblk;
        }
        sl.scap[k] = (int)std::min((double)kMaxUnitRecords, std::ceil(1.25 * xs) + 32);
    }
    double rs = 0;
    for (int k = 0; k < S; k++) rs += (double)sl.scap[k] * nblk;
    if (rs * vsz > kSlabBudget) {
        const double f = kSlabBudget / (rs * vsz);
        for (int k = 0; k < S; k++) sl.scap[k] = std::max(1, (int)(sl.scap[k] * f));
    }
    sl.sbase.assign(S, 0);
    long long as = 0;
    for (int k = 0; k < S; k++) {
        sl.sbase[k] = as;
        as += (long long)sl.scap[k] * nblk;
    }
    slab_units(c);
    int rc = ensure(c, &sl.d_vals, sl.vals_cap, (size_t)std::max(1ll, as) * vsz);
    if (!rc) rc = ensure(c, &sl.d_sbase, sl.base_cap, (size_t)S);
    if (!rc) rc = ensure(c, &sl.d_scap, sl.scap_cap, (size_t)S);
    if (!rc) rc = ensure(c, &sl.d_uoff, sl.uoff_cap, (size_t)S + 1);
    if (!rc) rc = ensure(c, &sl.d_units, sl.units_cap, (size_t)sl.nunits);
    if (!rc) rc = ensure(c, &sl.d_sfill, sl.fill_cap, (size_t)nblk * S);
    if (!rc) rc = ensure(c, &sl.d_stats, sl.stats_cap, (size_t)2 * S + 1);
    if (!rc) rc = ensure(c, &sl.d_parts, sl.parts_cap, (size_t)sl.nunits * SB);
    if (rc) return rc;
    HIPCHECK(c, hipMemcpy(sl.d_sbase, sl.sbase.data(), S * 8, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(sl.d_scap, sl.scap.data(), S * 4, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(sl.d_uoff, sl.uoff.data(), (S + 1) * 4, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(sl.d_units, sl.units.data(), sl.units.size() * sizeof(int4),
                          hipMemcpyHostToDevice));
    return HBT_OK;
}

int fast_args(hbt_ctx *c, FastArgs &a) {
    set_prefilter_margins(c);
    const View v = make_view(c);
    HIPCHECK(c, hipMemcpy(c->d_view, &v, sizeof(View), hipMemcpyHostToDevice));
    a.P = c->P;
    a.views = c->d_view;
    a.items = c->d_items;
    a.nitems = c->nitems;
    a.nqueues = c->nqueues;
    for (int q = 0; q <= kMaxQueues; q++)
        a.qoff[q] = q <= c->nqueues ? c->qoff[(size_t)q] : c->nitems;
    a.counter = c->d_counter;
    a.binned = c->d_binned;
    a.slab = SlabArgs{};
    int rc = slab_prepare(c, c->nitems);
    if (rc) return rc;
    const auto &sl = c->slab;
    if (sl.on) {
        a.slab.vals = sl.d_vals;
        a.slab.sbase = sl.d_sbase;
        a.slab.scap = sl.d_scap;
        a.slab.sfill = sl.d_sfill;
        a.slab.stats = sl.d_stats;
        a.slab.units = sl.d_units;
        a.slab.uoff = sl.d_uoff;
        a.slab.S = sl.S;
        a.slab.SB = sl.SB;
        a.slab.q = c->P.qnpts;
        a.slab.q_base = c->P.q_base;
        a.slab.inv_dq = 1.0 / c->P.delta_q;
    }
    return HBT_OK;
}

// The fast kernel, and with slab binning its two binning kernels, queued on
// st (kernel launches only: this runs inside a graph capture).
int launch_fast(hbt_ctx *c, FastArgs &a, hipStream_t st) {
    const long blocks = fast_blocks(c, a.nitems);
    const bool f32 = c->cfg.precision == HBT_PREC_FP32C;
    const bool slab = a.slab.vals != nullptr;
    if (slab) {
        if (blocks != c->slab.nblk)
            return fail(c, HBT_ERR_STATE, "slab regions sized for another launch shape");
        HIPCHECK(c, launch_zero_words(a.slab.stats, (size_t)2 * a.slab.S + 1, st));
    }
    HIPCHECK(c, launch_fast_kernel(a, f32, c->tune.R, c->P.qinv || c->P.azimuthal, blocks, st));
    if (slab) {
        const int nu = c->slab.nunits;
        if (f32)
            hipLaunchKernelGGL(k_slab_bin<float>, dim3(nu), dim3(1024), 0, st, a.slab,
                               c->slab.d_parts);
        else
            hipLaunchKernelGGL(k_slab_bin<double>, dim3(nu), dim3(1024), 0, st, a.slab,
                               c->slab.d_parts);
        HIPCHECK(c, hipGetLastError());
        const long long nb = (long long)a.slab.S * a.slab.SB;
        hipLaunchKernelGGL(k_slab_reduce, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st,
                           (const SlabPart *)c->slab.d_parts, a.slab.uoff, a.slab.S, a.slab.SB,
                           c->d_bins);
        HIPCHECK(c, hipGetLastError());
        c->slab.pending = true;
    }
    return HBT_OK;
}

struct ItemList {
    std::vector<WorkItem> items;
    std::vector<JWin> wins;
    int nqueues = 1;                          // work queues (one per XCD at most)
    std::vector<int> qoff = std::vector<int>(2, 0);
};

// Work items: same-event tiles of every group, then mixed tiles of every
// event.  Same: i tile of 64R particles x j chunk of its px window.
// Mixed: i tile x up to kchunk partner windows of the event's plan rows.
// enum_same receives the reference's same-event pair count; the mixed
// count is known from the plan (upload_plan).
// Work queues, one per XCD: every item of a group goes to one queue, so the
// waves of one XCD work on one group at a time, whose particle records stay
// in that XCD's L2.  Groups are dealt to queues largest first, each to the
// least loaded queue; the queues keep the items' size-class order.
// nqueues is a power of two up to max_queues and at most the number of
// groups with items.
void assign_queues(ItemList &L, const std::vector<uint64_t> &pairs, int ngroups,
                   int max_queues) {
    const size_t n = L.items.size();
    std::vector<uint64_t> load((size_t)std::max(1, ngroups), 0);
    for (size_t k = 0; k < n; k++) load[(size_t)L.items[k].grp] += pairs[k] + 1;
    int busy = 0;
    for (uint64_t x : load) busy += x > 0;
    int nq = 1;
    while (nq * 2 <= std::min(max_queues, busy)) nq *= 2;
    std::vector<int> gq(load.size(), 0);
    if (nq > 1) {
        std::vector<int> order(load.size());
        for (size_t g = 0; g < order.size(); g++) order[g] = (int)g;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return load[(size_t)a] > load[(size_t)b]; });
        std::vector<uint64_t> qload((size_t)nq, 0);
        for (int g : order) {
            const int q = (int)(std::min_element(qload.begin(), qload.end()) - qload.begin());
            gq[(size_t)g] = q;
            qload[(size_t)q] += load[(size_t)g];
        }
    }
    L.nqueues = nq;
    L.qoff.assign((size_t)nq + 1, 0);
    for (size_t k = 0; k < n; k++) L.qoff[(size_t)gq[(size_t)L.items[k].grp] + 1]++;
    for (int q = 0; q < nq; q++) L.qoff[(size_t)q + 1] += L.qoff[(size_t)q];
    std::vector<int> at(L.qoff.begin(), L.qoff.end() - 1);
    std::vector<WorkItem> byq(n);
    for (size_t k = 0; k < n; k++) byq[(size_t)at[(size_t)gq[(size_t)L.items[k].grp]]++] = L.items[k];
    L.items.swap(byq);
}

void build_items(hbt_ctx *c, bool same, bool mixed, ItemList &L, uint64_t &enum_same) {
    const Slot &s = c->slot[0];
    ItemShape sh = item_shape(c->tune);
    const double D = px_window(c->D);
    // a launch with fewer items than kItemsPerWave per resident wave (one
    // buffer of the per-buffer path, a small request) leaves most waves idle
    // and ends on its largest item: its j chunks and partner windows per item
    // shrink (to 64 j's and one window at most) until it has enough
    const size_t waves = (size_t)std::max(1, c->cu_count) * (size_t)c->tune.blocks_per_cu * 4;
    std::vector<uint64_t> pairs;
    for (;;) {
        L.items.clear();
        L.wins.clear();
        pairs.clear();
        enum_same = 0;
        if (same)
            items_same(sh, s.off, s.foff, s.grp_ev, s.srt_cell.data(), s.srt_py.data(), D, 0,
                       L.items, &pairs, enum_same);
        if (mixed && c->plan_valid) {
            const Slot &m = mixed_slot(c);
            items_mixed(sh, s.foff, s.grp_ev, c->rowoff, c->plan_gids, m.foff, s.f_pt.data(),
                        m.f_pt.data(), pt_window(c->D), 0, L.items, L.wins, &pairs);
        }
        const size_t want = kItemsPerWave * waves;
        if (L.items.size() >= want || (sh.JC <= 64 && sh.KC <= 1)) break;
        // items grow about inversely with the chunk sizes: one rebuild with
        // both cut by the missing factor (rounded up to a power of two)
        int f = 2;
        while ((size_t)f * L.items.size() < want && f < 64) f *= 2;
        sh.JC = std::max(64, sh.JC / f);
        sh.KC = std::max(1, sh.KC / f);
    }
    // larger items first, by factor-4 size classes (stable inside a class,
    // which keeps a group's items together): the persistent waves pull items
    // in queue order, so the launch ends on small items instead of one wave
    // finishing a large one (uneven multiplicities make items differ by 10x
    // and more)
    // (a counting sort over the 33 classes of a 64-bit count)
    auto size_class = [&](size_t k) {
        int c = 0;
        for (uint64_t v = pairs[k]; v >= 4; v >>= 2) c++;
        return c;
    };
    const size_t n = L.items.size();
    std::vector<uint8_t> cls(n);
    size_t start[34] = {};
    for (size_t k = 0; k < n; k++) {
        cls[k] = (uint8_t)size_class(k);
        start[32 - cls[k] + 1]++;
    }
    for (int c = 1; c < 34; c++) start[c] += start[c - 1];
    std::vector<WorkItem> sorted(n);
    std::vector<uint64_t> spairs(n);
    for (size_t k = 0; k < n; k++) {
        const size_t at = start[32 - cls[k]]++;
        sorted[at] = L.items[k];
        spairs[at] = pairs[k];
    }
    L.items.swap(sorted);
    // per-XCD queues where a group's records take at most 3/4 of one XCD's
    // L2, which also holds the mixed lists and the histogram lines (on the
    // bench workloads: 1 % on the headline's precision form, 3 % on the
    // batched K_phi line, 18 % on its K+ launch; the headline's FP64 form
    // loses 1 %); larger groups (the
    // large-histogram form's 4e5 particles: 7 %) run faster with every XCD
    // on the same group
    long long maxg = 0;
    for (int g = 0; g < s.ngroups; g++)
        maxg = std::max(maxg, s.foff[(size_t)s.grp_ev[g + 1]] - s.foff[(size_t)s.grp_ev[g]]);
    const bool f32 = c->cfg.precision == HBT_PREC_FP32C;
    const size_t rec = sizeof(PartRec) + (f32 ? sizeof(PfRec32) + 4 : sizeof(PfRec64) + 8);
    assign_queues(L, spairs, s.ngroups,
                  (size_t)maxg * rec * 4 <= kL2BytesPerXcd * 3 ? kMaxQueues : 1);
}

// Uploads the plan: ids are relative to the event's own group and are
// turned into global event indices of the mixed list.
int upload_plan(hbt_ctx *c, const std::vector<int> &M_per_event, const int32_t *ids,
                const double *cs, uint64_t &enum_mixed) {
    Slot &s = c->slot[0];
    const Slot &m = mixed_slot(c);
    // the previous plan (and a batch built on it) is gone from here on; the
    // new one is committed only once every partner id has been checked and
    // copied, so a failed call leaves no plan at all
    drop_graph(c);
    c->plan_valid = false;
    std::vector<long long> rowoff(s.nev + 1, 0);
    for (int e = 0; e < s.nev; e++) rowoff[e + 1] = rowoff[e] + M_per_event[e];
    const long long rows = rowoff[s.nev];
    std::vector<int> gids((size_t)std::max<long long>(rows, 1));
    enum_mixed = 0;
    for (int g = 0; g < s.ngroups; g++) {
        const int mb = m.grp_ev[g], mn = m.grp_ev[g + 1] - m.grp_ev[g];
        for (int e = s.grp_ev[g]; e < s.grp_ev[g + 1]; e++) {
            const long long n1 = s.off[e + 1] - s.off[e];
            for (long long r = rowoff[e]; r < rowoff[e + 1]; r++) {
                const int p = ids[r];
                if (p < 0 || p >= mn) return fail(c, HBT_ERR_ARG, "partner id out of range");
                gids[r] = mb + p;
                enum_mixed += (uint64_t)(n1 * (m.off[mb + p + 1] - m.off[mb + p]));
            }
        }
    }
    const size_t need = (size_t)std::max<long long>(rows, 1);
    if (need > c->plan_cap || !c->d_ids || !c->d_cs) {
        int rc = dev_alloc(c, &c->d_ids, need);
        if (!rc) rc = dev_alloc(c, &c->d_cs, 2 * need);
        if (rc) return rc;
        c->plan_cap = need;
    }
    int rc = ensure(c, &c->d_rowoff, c->rowoff_cap, (size_t)s.nev + 1);
    if (rc) return rc;
    HIPCHECK(c, hipMemcpy(c->d_ids, gids.data(), rows * 4, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_cs, cs, rows * 16, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_rowoff, rowoff.data(), (s.nev + 1) * 8, hipMemcpyHostToDevice));
    gids.resize((size_t)rows);
    c->plan_gids.swap(gids);
    c->rowoff.swap(rowoff);
    c->plan_valid = true;
    c->plan_enum_mixed = enum_mixed;
    return HBT_OK;
}

int upload_items(hbt_ctx *c, const ItemList &L) {
    const std::vector<WorkItem> &items = L.items;
    if (items.size() >= (size_t)INT32_MAX) return fail(c, HBT_ERR_CAPACITY, "too many work items");
    int rc = ensure(c, &c->d_items, c->items_cap, std::max<size_t>(1, items.size()));
    if (!rc) rc = ensure(c, &c->d_wins, c->wins_cap, std::max<size_t>(1, L.wins.size()));
    if (rc) return rc;
    if (!items.empty())
        HIPCHECK(c, hipMemcpy(c->d_items, items.data(), items.size() * sizeof(WorkItem),
                              hipMemcpyHostToDevice));
    if (!L.wins.empty())
        HIPCHECK(c, hipMemcpy(c->d_wins, L.wins.data(), L.wins.size() * sizeof(JWin),
                              hipMemcpyHostToDevice));
    c->nitems = (int)items.size();
    c->nqueues = L.nqueues;
    c->qoff = L.qoff;
    return HBT_OK;
}

int timed_fast(hbt_ctx *c, float *ms) {
    FastArgs a;
    int rc0 = fast_args(c, a);
    if (rc0) return rc0;
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    if (a.nitems > 0) {
        int rc = launch_fast(c, a, c->stream);
        if (rc) return rc;
    }
    HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHECK(c, hipEventSynchronize(c->ev[1]));
    HIPCHECK(c, hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
    return slab_read_stats(c);
}

// Plain kernels: one launch per group (same) and per group (mixed).
int run_plain(hbt_ctx *c, bool same, bool mixed) {
    Slot &s = c->slot[0];
    const Slot &m = mixed_slot(c);
    DevHist H = hist_view(c);
    if (same) {
        for (int g = 0; g < s.ngroups; g++) {
            const long long g0 = s.off[s.grp_ev[g]], n = s.off[s.grp_ev[g + 1]] - g0;
            if (n < 2) continue;
            const int T = (int)((n + TILE - 1) / TILE);
            if (T > 65535) return fail(c, HBT_ERR_CAPACITY, "group too large for the plain kernel");
            hipLaunchKernelGGL(k_same_plain, dim3(T, T), dim3(TILE), 0, c->stream, c->P, H,
                               soa_view(s), g0, (int)n, c->psi[g]);
            HIPCHECK(c, hipGetLastError());
        }
    }
    if (mixed) {
        long long max2 = 0;
        for (int e = 0; e < m.nev; e++) max2 = std::max(max2, m.off[e + 1] - m.off[e]);
        const int Tj = (int)((max2 + TILE - 1) / TILE);
        for (int g = 0; g < s.ngroups && Tj > 0; g++) {
            const int e0 = s.grp_ev[g], e1 = s.grp_ev[g + 1];
            if (e1 == e0) continue;
            const int M = (int)(c->rowoff[e0 + 1] - c->rowoff[e0]);
            long long max1 = 0;
            for (int e = e0; e < e1; e++) max1 = std::max(max1, s.off[e + 1] - s.off[e]);
            const int Ti = (int)((max1 + TILE - 1) / TILE);
            if (Ti == 0 || M == 0) continue;
            // events per launch so that gridDim.z stays below 65536
            const int per = std::max(1, 65535 / M);
            for (int eb = e0; eb < e1; eb += per) {
                const int ne = std::min(per, e1 - eb);
                hipLaunchKernelGGL(k_mixed_plain, dim3(Tj, Ti, ne * M), dim3(TILE), 0, c->stream,
                                   c->P, H, soa_view(s), soa_view(m), s.d_off, m.d_off, c->d_ids,
                                   c->d_cs, M, eb, c->rowoff[eb], c->psi[g]);
                HIPCHECK(c, hipGetLastError());
            }
        }
    }
    return HBT_OK;
}

// The ordered path for one side (same-event or mixed) of the current call:
// arrivals per (key, row), the cut per key, then binning of the admitted
// pairs (hbt_cap.h).  FP64 with the literal cascade in every precision.
// Deferred pairs of a capped call (K_phi within the guard band of a bin
// edge): the count pass recorded each with its (row, pos).  The host runs
// the cascade with glibc's atan2 (as the reference computes K_phi) to name
// the pair's (K_T, K_phi) key, or none when the reference skips it; adds the
// arrivals; and uploads the keys sorted by (row, pos) for the pos pass.
struct CapDefer {
    std::vector<DeferRec> rec;
    std::vector<DeferKey> keys;   // sorted by (row, pos)
    std::vector<DeferAdd> adds;   // per record, what it adds if admitted
    std::vector<int> key_of;      // per record, its key or -1
};

int cap_defer_keys(hbt_ctx *c, long long nrows, CapDefer &dk) {
    dk = CapDefer();
    if (!c->P.azimuthal) return HBT_OK;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    unsigned n = 0;
    HIPCHECK(c, hipMemcpy(&n, c->d_ndefer, sizeof(n), hipMemcpyDeviceToHost));
    if (n == 0) return HBT_OK;
    if (n > kDeferCap)
        return fail(c, HBT_ERR_CAPACITY, "more K_phi bin-edge pairs than the deferral buffer");
    dk.rec.resize(n);
    HIPCHECK(c, hipMemcpy(dk.rec.data(), c->d_defer, n * sizeof(DeferRec), hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemset(c->d_ndefer, 0, sizeof(unsigned)));
    c->deferred_total += n;
    std::vector<long long> arr;
    dk.key_of.assign(n, -1);
    dk.adds.resize(n);
    for (unsigned i = 0; i < n; i++) {
        std::vector<DeferAdd> one;
        resolve_records(c->P, &dk.rec[i], 1, one);
        if (one.empty()) continue;  // the reference skips it (K_phi index n_Kphi)
        const long per = (long)c->P.qnpts * c->P.qnpts * c->P.qnpts;
        const int key = (int)(one[0].k / per);  // bins are [kt][kphi][q^3]
        dk.key_of[i] = key;
        dk.adds[i] = one[0];
        arr.push_back((long long)key * nrows + dk.rec[i].row);
    }
    for (unsigned i = 0; i < n; i++)
        dk.keys.push_back(DeferKey{dk.rec[i].row, dk.rec[i].pos, dk.key_of[i], 0});
    std::sort(dk.keys.begin(), dk.keys.end(), [](const DeferKey &a, const DeferKey &b) {
        return a.row < b.row || (a.row == b.row && a.pos < b.pos);
    });
    int rc = ensure(c, &c->d_dkey, c->dkey_cap, dk.keys.size());
    if (!rc) rc = ensure(c, &c->d_darr, c->darr_cap, std::max<size_t>(1, arr.size()));
    if (rc) return rc;
    HIPCHECK(c, hipMemcpy(c->d_dkey, dk.keys.data(), dk.keys.size() * sizeof(DeferKey),
                          hipMemcpyHostToDevice));
    if (!arr.empty()) {
        HIPCHECK(c, hipMemcpy(c->d_darr, arr.data(), arr.size() * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_cap_add_arrivals, dim3((unsigned)((arr.size() + 255) / 256)),
                           dim3(256), 0, c->stream, c->d_cnt, (const long long *)c->d_darr,
                           (int)arr.size());
        HIPCHECK(c, hipGetLastError());
    }
    return HBT_OK;
}

// After the binning pass: each deferred pair whose key's cut admits its
// (row, pos) is added to its bin (HBT_correlation.cpp:402-429, :651-678).
int cap_defer_bin(hbt_ctx *c, int nkey3, const CapDefer &dk) {
    if (dk.rec.empty()) return HBT_OK;
    std::vector<CapCut> cut(nkey3);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(cut.data(), c->d_cut, nkey3 * sizeof(CapCut), hipMemcpyDeviceToHost));
    std::vector<DeferAdd> adds;
    for (size_t i = 0; i < dk.rec.size(); i++) {
        const int key = dk.key_of[i];
        if (key < 0) continue;
        const CapCut &ct = cut[key];
        const long long row = dk.rec[i].row, pos = dk.rec[i].pos;
        if (row < ct.row || (row == ct.row && pos <= ct.pos)) adds.push_back(dk.adds[i]);
    }
    if (adds.empty()) return HBT_OK;
    DeferAdd *d = nullptr;
    HIPCHECK(c, hipMalloc((void **)&d, adds.size() * sizeof(DeferAdd)));
    hipError_t e = hipMemcpy(d, adds.data(), adds.size() * sizeof(DeferAdd), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_apply_defer, dim3((unsigned)((adds.size() + 255) / 256)), dim3(256),
                           0, c->stream, d, (int)adds.size());
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d);
    HIPCHECK(c, e);
    return HBT_OK;
}

int run_capped(hbt_ctx *c, bool mixed) {
    const Slot &s = c->slot[0];
    const Slot &m = mixed_slot(c);
    CapDefer dk;
    // pairs deferred by an uncapped pass queued before this one are binned
    // first, so the records the count pass leaves are this pass's own
    if (int rc0 = resolve_deferred(c)) return rc0;
    const long long nrows = std::max(1ll, s.n);
    const int nkey3 = c->cfg.n_KT * c->P.nphi, nkq = c->cfg.n_KT, nkeys = nkey3 + nkq;
    int rc = ensure(c, &c->d_cnt, c->cnt_cap, (size_t)nkeys * nrows);
    if (!rc) rc = ensure(c, &c->d_cut, c->cut_cap, (size_t)nkeys);
    if (!rc) rc = ensure(c, &c->d_lim, c->lim_cap, (size_t)nkeys);
    if (!rc) rc = ensure(c, &c->d_gpoff, c->gpoff_cap, (size_t)s.ngroups + 1);
    if (rc) return rc;
    HIPCHECK(c, hipMemsetAsync(c->d_cnt, 0, (size_t)nkeys * nrows * 4, c->stream));
    // remaining capacity per key from the counters so far (sums of the bins)
    std::vector<unsigned long long> c3, cq(nkq, 0);
    const long per = (long)c->P.qnpts * c->P.qnpts * c->P.qnpts;
    rc = mixed ? group_sums(c, (const unsigned long long *)(c->d_bins + BIN_DEN), per, nkey3,
                            BIN_STRIDE, c3)
               : group_sums(c, (const double *)(c->d_bins + BIN_COUNT), per, nkey3, BIN_STRIDE,
                            c3);
    if (!rc && c->P.qinv)
        rc = group_sums(c, mixed ? c->d_qinv_den : c->d_qinv_count, c->P.qnpts, nkq, 1, cq);
    if (rc) return rc;
    std::vector<unsigned long long> lim(nkeys);
    for (int k = 0; k < nkey3; k++) lim[k] = c3[k] > c->needed ? 0 : c->needed + 1 - c3[k];
    for (int k = 0; k < nkq; k++)
        lim[nkey3 + k] = cq[k] >= c->needed_qinv ? 0 : c->needed_qinv - cq[k];
    std::vector<long long> gpoff(s.ngroups + 1);
    for (int g = 0; g <= s.ngroups; g++) gpoff[g] = s.off[s.grp_ev[g]];
    HIPCHECK(c, hipMemcpy(c->d_lim, lim.data(), nkeys * 8, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_gpoff, gpoff.data(), gpoff.size() * 8, hipMemcpyHostToDevice));
    CapView C;
    C.cnt3 = c->d_cnt;
    C.cntq = c->d_cnt + (size_t)nkey3 * nrows;
    C.cut3 = c->d_cut;
    C.cutq = c->d_cut + nkey3;
    C.nrows = nrows;
    const DevHist H = hist_view(c);
    const DevSoA A = soa_view(s), B = soa_view(m);
    for (int pass = 0; pass < 2; pass++) {
        const bool count = pass == 0;
        if (!mixed) {
            for (int g = 0; g < s.ngroups; g++) {
                const long long g0 = gpoff[g], n = gpoff[g + 1] - g0;
                if (n < 2) continue;
                const int T = (int)((n + TILE - 1) / TILE);
                if (T > 65535) return fail(c, HBT_ERR_CAPACITY, "group too large for the cap path");
                if (count)
                    hipLaunchKernelGGL(k_same_cap<true>, dim3(T, T), dim3(TILE), 0, c->stream,
                                       c->P, H, A, C, g0, (int)n, c->psi[g]);
                else
                    hipLaunchKernelGGL(k_same_cap<false>, dim3(T, T), dim3(TILE), 0, c->stream,
                                       c->P, H, A, C, g0, (int)n, c->psi[g]);
                HIPCHECK(c, hipGetLastError());
            }
        } else {
            long long max2 = 0;
            for (int e = 0; e < m.nev; e++) max2 = std::max(max2, m.off[e + 1] - m.off[e]);
            const int Tj = (int)((max2 + TILE - 1) / TILE);
            for (int g = 0; g < s.ngroups && Tj > 0; g++) {
                const int e0 = s.grp_ev[g], e1 = s.grp_ev[g + 1];
                if (e1 == e0) continue;
                const int M = (int)(c->rowoff[e0 + 1] - c->rowoff[e0]);
                long long max1 = 0;
                for (int e = e0; e < e1; e++) max1 = std::max(max1, s.off[e + 1] - s.off[e]);
                const int Ti = (int)((max1 + TILE - 1) / TILE);
                if (Ti == 0 || M == 0) continue;
                const int perl = std::max(1, 65535 / M);
                for (int eb = e0; eb < e1; eb += perl) {
                    const int ne = std::min(perl, e1 - eb);
                    if (count)
                        hipLaunchKernelGGL(k_mixed_cap<true>, dim3(Tj, Ti, ne * M), dim3(TILE), 0,
                                           c->stream, c->P, H, A, B, C, s.d_off, m.d_off,
                                           c->d_ids, c->d_cs, M, eb, c->rowoff[eb], c->psi[g]);
                    else
                        hipLaunchKernelGGL(k_mixed_cap<false>, dim3(Tj, Ti, ne * M), dim3(TILE),
                                           0, c->stream, c->P, H, A, B, C, s.d_off, m.d_off,
                                           c->d_ids, c->d_cs, M, eb, c->rowoff[eb], c->psi[g]);
                    HIPCHECK(c, hipGetLastError());
                }
            }
        }
        if (!count) break;
        // pairs on a K_phi bin edge: their keys from the host's cascade, their
        // arrivals added before the scan
        rc = cap_defer_keys(c, nrows, dk);
        if (rc) return rc;
        const int ndk = (int)dk.keys.size();
        // cut row per key, then the position inside it
        hipLaunchKernelGGL(k_cap_scan, dim3(nkeys), dim3(1024), 0, c->stream, c->d_cnt, nrows,
                           c->d_lim, c->d_cut);
        HIPCHECK(c, hipGetLastError());
        const double *psi = c->d_psi;
        if (!mixed) {
            hipLaunchKernelGGL(k_cap_pos_same, dim3((nkey3 + 63) / 64), dim3(64), 0, c->stream,
                               c->P, A, c->d_gpoff, s.ngroups, psi, c->d_cut, nkey3, false,
                               c->d_dkey, ndk);
            hipLaunchKernelGGL(k_cap_pos_same, dim3((nkq + 63) / 64), dim3(64), 0, c->stream,
                               c->P, A, c->d_gpoff, s.ngroups, psi, c->d_cut + nkey3, nkq, true,
                               c->d_dkey, ndk);
        } else {
            hipLaunchKernelGGL(k_cap_pos_mixed, dim3((nkey3 + 63) / 64), dim3(64), 0, c->stream,
                               c->P, A, B, s.d_off, s.nev, m.d_off, c->d_rowoff, c->d_ids,
                               c->d_cs, c->d_gpoff, s.ngroups, psi, c->d_cut, nkey3, false,
                               c->d_dkey, ndk);
            hipLaunchKernelGGL(k_cap_pos_mixed, dim3((nkq + 63) / 64), dim3(64), 0, c->stream,
                               c->P, A, B, s.d_off, s.nev, m.d_off, c->d_rowoff, c->d_ids,
                               c->d_cs, c->d_gpoff, s.ngroups, psi, c->d_cut + nkey3, nkq, true,
                               c->d_dkey, ndk);
        }
        HIPCHECK(c, hipGetLastError());
    }
    // the deferred pairs the cuts admit, binned with the host's values
    return cap_defer_bin(c, nkey3, dk);
}

int prepare_and_capture(hbt_ctx *c) {
    drop_graph(c);
    ItemList items;
    build_items(c, true, true, items, c->batch_enum_same);
    c->batch_enum_mixed = c->plan_valid ? c->plan_enum_mixed : 0;
    int rc = upload_items(c, items);
    if (rc) return rc;
    rc = upload_psi(c);
    if (rc) return rc;
    if (c->cfg.kernel != HBT_KERNEL_FAST) {  // the plain batch runs launch by launch
        c->batch_ready = true;
        return HBT_OK;
    }
    c->launch_enum_same = c->batch_enum_same;
    c->launch_enum_mixed = c->batch_enum_mixed;
    FastArgs a;
    rc = fast_args(c, a);
    if (rc) return rc;
    c->batch_args = a;
    // The batch is one graph (counter fill + persistent kernel).  A runtime
    // that refuses the capture (a counter-collecting profiler does) gets the
    // same launches issued directly by hbt_launch_batch, and so does a batch
    // launched once (the per-buffer path), for which instantiating a graph
    // costs more than it saves.
    hipGraph_t g = nullptr;
    bool captured = c->batch_capture &&
                    hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (captured) {
        rc = launch_fast(c, a, c->stream);
        captured = hipStreamEndCapture(c->stream, &g) == hipSuccess && !rc && g;
    }
    if (captured) captured = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) c->graph = g;
    if (!captured) {
        drop_graph(c);
        (void)hipGetLastError();  // the refused capture's error is not the batch's
    }
    c->batch_ready = true;
    return HBT_OK;
}

// First half of an upload, called from the gather once the kept particles
// are known: waits for queued work, invalidates the slot, sizes its device
// buffers, gathers the kept columns into the pinned staging buffer on
// several threads and queues their copies (and the event offsets') on the
// context's stream.
int upload_columns(hbt_ctx *c, Slot &s, bool f32, int32_t nev, const HostParticles &h,
                   const double *const soa[8]) {
    int rc = quiesce(c);
    if (rc) return rc;
    drop_graph(c);
    c->plan_valid = false;
    s.valid = false;  // until every copy has succeeded
    const long long n = h.n, nf = h.nf;
    if ((size_t)n > s.cap || !s.d_soa[0]) {
        for (auto &p : s.d_soa) {
            rc = dev_alloc(c, &p, (size_t)n);
            if (rc) return rc;
        }
        s.cap = (size_t)n;
    }
    if ((size_t)nf > s.cap_f || !s.d_orig) {
        for (DevRecords *r : {&s.ev, &s.srt}) {
            if (!rc) rc = dev_alloc(c, &r->rec, (size_t)nf);
            if (!rc && !f32) rc = dev_alloc(c, &r->pf64, (size_t)nf);
            if (!rc && f32) rc = dev_alloc(c, &r->pf32, (size_t)nf);
            if (!rc && !f32) rc = dev_alloc(c, &r->cw64, (size_t)nf);
            if (!rc && f32) rc = dev_alloc(c, &r->cw32, (size_t)nf);
        }
        if (!rc) rc = dev_alloc(c, &s.d_orig, (size_t)nf);
        if (rc) return rc;
        s.cap_f = (size_t)nf;
    }
    if ((size_t)nev + 1 > s.cap_ev || !s.d_off) {
        rc = dev_alloc(c, &s.d_off, (size_t)nev + 1);
        if (rc) return rc;
        s.cap_ev = (size_t)nev + 1;
    }
    const size_t col_bytes = stage_round((size_t)n * 8);
    const size_t off_bytes = ((size_t)nev + 1) * 8;
    HIPCHECK(c, ensure_pinned(&c->h_stage, &c->stage_cap, 8 * col_bytes + stage_round(off_bytes)));
    char *stage = c->h_stage;
    parallel_ranges(n, 1 << 15, [&](long long k0, long long k1) {
        for (int a = 0; a < 8; a++) {
            double *dst = reinterpret_cast<double *>(stage + a * col_bytes);
            const double *src = soa[a];
            for (long long k = k0; k < k1; k++) dst[k] = src[h.keep[k]];
        }
    });
    for (int a = 0; a < 8 && n > 0; a++)
        HIPCHECK(c, hipMemcpyAsync(s.d_soa[a], stage + a * col_bytes, (size_t)n * 8,
                                   hipMemcpyHostToDevice, c->stream));
    std::memcpy(stage + 8 * col_bytes, h.off.data(), off_bytes);
    HIPCHECK(c, hipMemcpyAsync(s.d_off, stage + 8 * col_bytes, off_bytes, hipMemcpyHostToDevice,
                               c->stream));
    return HBT_OK;
}

int upload_slot(hbt_ctx *c, int32_t slot, int32_t ngroups, const int32_t *grp_ev, int32_t nev,
                const int64_t *ev_off, const double *const soa[8]) {
    if (slot != 0 && slot != 1) return fail(c, HBT_ERR_ARG, "slot must be 0 or 1");
    if (ngroups < 1 || !grp_ev) return fail(c, HBT_ERR_ARG, "bad buffer");
    HostParticles h;
    std::string err;
    Slot &s = c->slot[slot];
    const bool f32 = c->cfg.precision == HBT_PREC_FP32C;  // prefilter records of one precision
    // Once the kept particles are known: wait for work queued on the stream
    // (a batch launched without hbt_sync may still read the slot and the
    // staging), size the device buffers, and queue the columns' copies, which
    // then run while the host sorts the fast lists by px.
    int hook_rc = HBT_OK;
    const std::function<int()> after_keep = [&]() -> int {
        hook_rc = upload_columns(c, s, f32, nev, h, soa);
        return hook_rc;
    };
    int rc = gather_particles(c->D, c->cfg, ngroups, grp_ev, nev, ev_off, soa, h, err, &after_keep);
    if (hook_rc) return hook_rc;
    if (rc) return fail(c, rc, err);
    const long long nf = h.nf;
    // the records, through the second pinned buffer
    std::vector<StagePiece> pieces;
    auto add = [&](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) pieces.push_back(StagePiece{dst, src, bytes});
    };
    if (nf > 0) {
        const std::pair<DevRecords *, const RecordSet *> sets[] = {{&s.ev, &h.fast},
                                                                   {&s.srt, &h.srt}};
        for (const auto &x : sets) {
            DevRecords &d = *x.first;
            const RecordSet &r = *x.second;
            // the srt set holds prefilter records only: its full records are
            // gathered on the device below
            if (x.first == &s.ev) add(d.rec, r.rec.data(), (size_t)nf * sizeof(PartRec));
            if (f32) {
                add(d.pf32, r.pf32.data(), (size_t)nf * sizeof(PfRec32));
                add(d.cw32, r.cw32.data(), (size_t)nf * 4);
            } else {
                add(d.pf64, r.pf64.data(), (size_t)nf * sizeof(PfRec64));
                add(d.cw64, r.cw64.data(), (size_t)nf * 8);
            }
        }
        add(s.d_orig, h.orr.data(), (size_t)nf * sizeof(OrigRank));
    }
    HIPCHECK(c, ensure_pinned(&c->h_stage2, &c->stage2_cap, stage_bytes(pieces)));
    HIPCHECK(c, stage_pieces(c->h_stage2, pieces, c->stream));
    // the records in the cell-sorted order, for the exact stage of same-event pairs
    HIPCHECK(c, gather_srt(s.ev.rec, s.d_orig, s.srt.rec, nf, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    s.nf = nf;
    s.foff = std::move(h.foff);
    s.f_pt = std::move(h.f_pt);
    s.srt_cell = std::move(h.srt_cell);
    s.srt_py = std::move(h.srt_py);
    s.ngroups = ngroups;
    s.grp_ev.assign(grp_ev, grp_ev + ngroups + 1);
    s.nev = nev;
    s.n = h.n;
    s.n_in = h.n_in;
    s.off = h.off;
    s.pmax2 = h.pmax2;
    s.psi2 = h.psi2;
    s.valid = true;
    if (slot == 0) c->psi = h.psi2;
    return HBT_OK;
}

int check_slots(hbt_ctx *c) {
    if (!c->slot[0].valid) return fail(c, HBT_ERR_STATE, "no buffer uploaded in slot 0");
    if (c->slot[1].valid && c->slot[1].ngroups != c->slot[0].ngroups)
        return fail(c, HBT_ERR_STATE, "the mixed-event list has a different number of groups");
    return HBT_OK;
}

}  // namespace

namespace hbt {

hipError_t launch_fast_kernel(const FastArgs &a, bool f32, int R, bool ext, long blocks,
                              hipStream_t st) {
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, a.counter, a.nqueues);
    // the K_phi instance's slab fill counters live in dynamic LDS
    const size_t dyn = ext && a.slab.vals ? (size_t)a.slab.S * sizeof(unsigned) : 0;
    if (dyn > (size_t)kMaxSlabsKphi * sizeof(unsigned)) return hipErrorInvalidValue;
#define HBT_LAUNCH(T, RR, X) \
    hipLaunchKernelGGL((k_fast<T, RR, X>), dim3(blocks), dim3(256), dyn, st, a)
#define HBT_LAUNCH_R(T, X)                    \
    do {                                      \
        if (R == 1) HBT_LAUNCH(T, 1, X);      \
        else if (R == 4) HBT_LAUNCH(T, 4, X); \
        else HBT_LAUNCH(T, 2, X);             \
    } while (0)
    if (f32) {
        if (ext) HBT_LAUNCH_R(float, true);
        else HBT_LAUNCH_R(float, false);
    } else {
        if (ext) HBT_LAUNCH_R(double, true);
        else HBT_LAUNCH_R(double, false);
    }
#undef HBT_LAUNCH_R
#undef HBT_LAUNCH
    return hipGetLastError();
}

hipError_t launch_zero_words(void *p, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, st,
                       (unsigned long long *)p, n);
    return hipGetLastError();
}

}  // namespace hbt

// =================================================================== C ABI

extern "C" {

const char *hbt_last_error(const hbt_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int hbt_create(const hbt_config *cfg, hbt_ctx **out) {
    if (!cfg || !out) return fail(nullptr, HBT_ERR_ARG, "null argument");
    *out = nullptr;
    const hbt_config &k = *cfg;
    Derived D;
    std::string err;
    int rc0 = derive(k, D, err);
    if (rc0) return fail(nullptr, rc0, err);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, HBT_ERR_NODEVICE, "no HIP device visible");
    if (k.device < 0 || k.device >= ndev) return fail(nullptr, HBT_ERR_NODEVICE, "bad device");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, k.device) != hipSuccess)
        return fail(nullptr, HBT_ERR_NODEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, HBT_ERR_NODEVICE,
                    std::string("built for gfx950, device is ") + prop.gcnArchName);
    if (hipSetDevice(k.device) != hipSuccess)
        return fail(nullptr, HBT_ERR_NODEVICE, "hipSetDevice failed");

    std::unique_ptr<hbt_ctx> c(new hbt_ctx());
    c->cfg = k;
    c->slab_binning = k.binning == 1;  // 2: atomics until hbt_autotune measures both
    c->cu_count = prop.multiProcessorCount;
    c->D = D;
    c->P = D.P;
    c->mid = D.mid;
    c->H = D.H;
    c->c_beta = D.c_beta;
    c->needed = D.needed;
    c->needed_qinv = D.needed_qinv;

    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(nullptr, HBT_ERR_HIP, "hipStreamCreate failed");
    for (auto &ev : c->ev)
        if (hipEventCreate(&ev) != hipSuccess)
            return fail(nullptr, HBT_ERR_HIP, "hipEventCreate failed");
    c->nbins = D.nbins;
    c->nq = D.nq;
    hbt_ctx *cp = c.get();
    int rc = dev_alloc(cp, &c->d_bins, (size_t)c->nbins * BIN_STRIDE);
    if (!rc) rc = dev_alloc(cp, &c->d_qinv_count, c->nq);
    if (!rc) rc = dev_alloc(cp, &c->d_qinv_den, c->nq);
    if (!rc) rc = dev_alloc(cp, &c->d_qinv_sum, c->nq);
    if (!rc) rc = dev_alloc(cp, &c->d_qinv_num, c->nq);
    if (!rc) rc = dev_alloc(cp, &c->d_stage, 10);
    if (!rc) rc = dev_alloc(cp, &c->d_binned, 2);
    if (!rc) rc = dev_alloc(cp, &c->d_sums, (size_t)k.n_KT * std::max(1, D.P.nphi));
    if (!rc) rc = dev_alloc(cp, &c->d_counter, (size_t)kQueueStride * kMaxQueues);
    if (!rc) rc = dev_alloc(cp, &c->d_view, 1);
    if (!rc) rc = dev_alloc(cp, &c->d_defer, kDeferCap);
    if (!rc) rc = dev_alloc(cp, &c->d_ndefer, 1);
    if (!rc) rc = zero_hist(cp);
    if (rc) {
        g_create_err = c->err;
        hbt_destroy(c.release());
        return rc;
    }
    *out = c.release();
    return HBT_OK;
}

int hbt_destroy(hbt_ctx *c) {
    if (!c) return HBT_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_graph(c);
    void *ptrs[] = {c->d_bins,      c->d_qinv_count, c->d_qinv_den,
                    c->d_qinv_sum,  c->d_qinv_num,  c->d_stage,      c->d_binned,
                    c->d_sums,      c->d_ids,       c->d_cs,         c->d_rowoff,
                    c->d_items,     c->d_counter,   c->d_psi,        c->d_view,
                    c->d_cnt,       c->d_cut,       c->d_lim,        c->d_gpoff,
                    c->d_wins,      c->d_defer,     c->d_ndefer,     c->d_dkey,
                    c->d_darr,      c->slab.d_vals, c->slab.d_units, c->slab.d_uoff,
                    c->slab.d_sbase, c->slab.d_scap, c->slab.d_sfill, c->slab.d_stats,
                    c->slab.d_parts};
    for (void *p : ptrs)
        if (p) hipFree(p);
    free_slot(c->slot[0]);
    free_slot(c->slot[1]);
    for (auto &ev : c->ev)
        if (ev) hipEventDestroy(ev);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->h_stage) hipHostFree(c->h_stage);
    if (c->h_stage2) hipHostFree(c->h_stage2);
    delete c;
    return HBT_OK;
}

int hbt_reset(hbt_ctx *c) {
    if (!c) return HBT_ERR_ARG;
    return zero_hist(c);
}

int hbt_set_psi_ref(hbt_ctx *c, double psi_ref) {
    if (!c) return HBT_ERR_ARG;
    if (int rc = quiesce(c)) return rc;
    for (auto &p : c->psi) p = psi_ref;
    if (c->psi.empty()) c->psi.push_back(psi_ref);
    c->batch_ready = false;
    return upload_psi(c);
}

int hbt_upload(hbt_ctx *c, int32_t slot, int32_t nev, const int64_t *ev_off,
               const double *const soa[8]) {
    if (!c) return HBT_ERR_ARG;
    const int32_t grp[2] = {0, nev};
    return upload_slot(c, slot, 1, grp, nev, ev_off, soa);
}

int hbt_upload_groups(hbt_ctx *c, int32_t slot, int32_t ngroups, const int32_t *grp_ev,
                      int32_t nev, const int64_t *ev_off, const double *const soa[8]) {
    if (!c || !grp_ev) return HBT_ERR_ARG;
    return upload_slot(c, slot, ngroups, grp_ev, nev, ev_off, soa);
}

int hbt_slot_info(hbt_ctx *c, int32_t slot, int32_t *nev, int64_t *n_kept) {
    if (!c || (slot != 0 && slot != 1)) return HBT_ERR_ARG;
    if (nev) *nev = c->slot[slot].nev;
    if (n_kept) *n_kept = c->slot[slot].n;
    return HBT_OK;
}

int hbt_run_same(hbt_ctx *c) {
    if (!c) return HBT_ERR_ARG;
    int rc = check_slots(c);
    if (!rc) rc = quiesce(c);
    if (rc) return rc;
    ItemList items;
    uint64_t es = 0;
    build_items(c, true, false, items, es);
    bool binds = false;
    rc = check_cap(c, false, es, binds);
    if (rc) return rc;
    rc = upload_psi(c);
    if (rc) return rc;
    if (c->cfg.kernel == HBT_KERNEL_PLAIN || binds) {
        HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
        rc = binds ? run_capped(c, false) : run_plain(c, true, false);
        if (rc) return rc;
        HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
        HIPCHECK(c, hipEventSynchronize(c->ev[1]));
        HIPCHECK(c, hipEventElapsedTime(&c->ms_same, c->ev[0], c->ev[1]));
    } else {
        c->batch_ready = false;
        c->launch_enum_same = es;
        c->launch_enum_mixed = 0;
        rc = upload_items(c, items);
        if (!rc) rc = timed_fast(c, &c->ms_same);
        if (rc) return rc;
    }
    rc = resolve_deferred(c);
    if (rc) return rc;
    c->enum_same_total += es;
    return HBT_OK;
}

int hbt_run_mixed(hbt_ctx *c, int32_t M, const int32_t *partner_ids, const double *cos_sin) {
    if (!c) return HBT_ERR_ARG;
    int rc = check_slots(c);
    if (!rc) rc = quiesce(c);
    if (rc) return rc;
    Slot &s = c->slot[0];
    if (s.nev == 0) return HBT_OK;
    if (M <= 0 || M > 64 * 1024 || !partner_ids || !cos_sin)
        return fail(c, HBT_ERR_ARG, "bad mixing plan");
    std::vector<int> Me(s.nev, M);
    uint64_t em = 0, es = 0;
    c->batch_ready = false;
    rc = upload_plan(c, Me, partner_ids, cos_sin, em);
    if (rc) return rc;
    bool binds = false;
    rc = check_cap(c, true, em, binds);
    if (rc) return rc;
    rc = upload_psi(c);
    if (rc) return rc;
    if (c->cfg.kernel == HBT_KERNEL_PLAIN || binds) {
        HIPCHECK(c, hipEventRecord(c->ev[2], c->stream));
        rc = binds ? run_capped(c, true) : run_plain(c, false, true);
        if (rc) return rc;
        HIPCHECK(c, hipEventRecord(c->ev[3], c->stream));
        HIPCHECK(c, hipEventSynchronize(c->ev[3]));
        HIPCHECK(c, hipEventElapsedTime(&c->ms_mixed, c->ev[2], c->ev[3]));
    } else {
        ItemList items;
        build_items(c, false, true, items, es);
        c->launch_enum_same = 0;
        c->launch_enum_mixed = em;
        rc = upload_items(c, items);
        if (!rc) rc = timed_fast(c, &c->ms_mixed);
        if (rc) return rc;
    }
    rc = resolve_deferred(c);
    if (rc) return rc;
    c->enum_mixed_total += em;
    return HBT_OK;
}

static int prepare_batch(hbt_ctx *c, const int32_t *M_per_group, const int32_t *partner_ids,
                  const double *cos_sin, bool capture) {
    if (!c || !M_per_group) return HBT_ERR_ARG;
    int rc = check_slots(c);
    if (!rc) rc = quiesce(c);
    if (rc) return rc;
    // a failure below leaves no prepared batch and no plan
    drop_graph(c);
    c->plan_valid = false;
    Slot &s = c->slot[0];
    std::vector<int> Me(s.nev, 0);
    long long rows = 0;
    for (int g = 0; g < s.ngroups; g++) {
        if (M_per_group[g] < 0 || M_per_group[g] > 64 * 1024)
            return fail(c, HBT_ERR_ARG, "bad number of mixed events");
        for (int e = s.grp_ev[g]; e < s.grp_ev[g + 1]; e++) {
            Me[e] = M_per_group[g];
            rows += Me[e];
        }
    }
    if (rows > 0 && (!partner_ids || !cos_sin)) return fail(c, HBT_ERR_ARG, "bad mixing plan");
    uint64_t em = 0;
    rc = upload_plan(c, Me, partner_ids, cos_sin, em);
    if (rc) return rc;
    c->batch_capture = capture;
    return prepare_and_capture(c);
}

int hbt_prepare_batch(hbt_ctx *c, const int32_t *M_per_group, const int32_t *partner_ids,
                      const double *cos_sin) {
    return prepare_batch(c, M_per_group, partner_ids, cos_sin, true);
}

int hbt_launch_batch(hbt_ctx *c) {
    if (!c) return HBT_ERR_ARG;
    if (!c->batch_ready) return fail(c, HBT_ERR_STATE, "no prepared batch (hbt_prepare_batch)");
    if (c->slab.resize && c->cfg.kernel == HBT_KERNEL_FAST) {
        // the last launch outgrew its slab regions: size them again from its
        // demand (and capture the batch again) before this launch
        int rc0 = quiesce(c);
        if (!rc0) rc0 = prepare_and_capture(c);
        if (rc0) return rc0;
    }
    bool bs = false, bm = false;
    int rc = check_cap(c, false, c->batch_enum_same, bs);
    if (!rc) rc = check_cap(c, true, c->batch_enum_mixed, bm);
    if (rc) return rc;
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    c->batch_capped = bs || bm;
    if (bs || bm) {  // a cap may bind: each side in enumeration order
        rc = bs ? run_capped(c, false) : run_plain(c, true, false);
        if (!rc) rc = bm ? run_capped(c, true) : (c->plan_valid ? run_plain(c, false, true) : 0);
        if (rc) return rc;
    } else if (c->cfg.kernel == HBT_KERNEL_FAST) {
        if (c->nitems > 0 && c->graph_exec) {
            HIPCHECK(c, hipGraphLaunch(c->graph_exec, c->stream));
        } else if (c->nitems > 0) {
            rc = launch_fast(c, c->batch_args, c->stream);
            if (rc) return rc;
        }
    } else {
        rc = run_plain(c, true, true);
        if (rc) return rc;
    }
    HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
    c->enum_same_total += c->batch_enum_same;
    c->enum_mixed_total += c->batch_enum_mixed;
    return HBT_OK;
}

int hbt_sync(hbt_ctx *c, double *batch_ms) {
    if (!c) return HBT_ERR_ARG;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess)
        c->ms_batch = ms;
    else
        (void)hipGetLastError();  // no batch timed yet: not an error of a later call
    if (batch_ms) *batch_ms = c->ms_batch;
    if (int rc = slab_read_stats(c)) return rc;
    return resolve_deferred(c);
}

int hbt_batch_info(hbt_ctx *c, uint64_t *enum_same, uint64_t *enum_mixed, int32_t *nitems) {
    if (!c) return HBT_ERR_ARG;
    if (enum_same) *enum_same = c->batch_enum_same;
    if (enum_mixed) *enum_mixed = c->batch_enum_mixed;
    if (nitems) *nitems = c->nitems;
    return HBT_OK;
}

int hbt_batch_graph(hbt_ctx *c, int32_t *graph) {
    if (!c || !graph) return HBT_ERR_ARG;
    if (!c->batch_ready) return fail(c, HBT_ERR_STATE, "no prepared batch (hbt_prepare_batch)");
    *graph = c->graph_exec != nullptr ? 1 : 0;
    return HBT_OK;
}

int hbt_batch_queues(hbt_ctx *c, int32_t *nqueues, int32_t *qoff) {
    if (!c || !nqueues) return HBT_ERR_ARG;
    if (!c->batch_ready) return fail(c, HBT_ERR_STATE, "no prepared batch (hbt_prepare_batch)");
    *nqueues = c->nqueues;
    if (qoff)
        for (int q = 0; q <= c->nqueues; q++) qoff[q] = c->qoff[(size_t)q];
    return HBT_OK;
}

int hbt_slab_info(hbt_ctx *c, int32_t *on, uint64_t *direct, uint64_t *capacity) {
    if (!c) return HBT_ERR_ARG;
    if (int rc = quiesce(c)) return rc;
    if (int rc = slab_read_stats(c)) return rc;
    const auto &sl = c->slab;
    uint64_t cap = 0;
    for (size_t k = 0; k < sl.scap.size(); k++) cap += (uint64_t)sl.scap[k] * (uint64_t)sl.nblk;
    if (on) *on = sl.on ? 1 : 0;
    if (direct) *direct = sl.on ? sl.direct : 0;
    if (capacity) *capacity = sl.on ? cap : 0;
    return HBT_OK;
}

int hbt_get_tuning(hbt_ctx *c, hbt_tuning *t) {
    if (!c || !t) return HBT_ERR_ARG;
    *t = c->tune;
    return HBT_OK;
}

int hbt_set_tuning(hbt_ctx *c, const hbt_tuning *t) {
    if (!c || !t) return HBT_ERR_ARG;
    if ((t->R != 1 && t->R != 2 && t->R != 4) || t->jchunk < 64 || t->jchunk > (1 << 24) ||
        t->kchunk < 1 || t->kchunk > 64 || t->blocks_per_cu < 1 || t->blocks_per_cu > 64)
        return fail(c, HBT_ERR_ARG, "tuning out of range");
    c->tune = *t;
    if (int rc = quiesce(c)) return rc;
    if (c->batch_ready) return prepare_and_capture(c);
    return HBT_OK;
}

// Times each candidate (one untimed and one timed launch of the prepared
// batch) and keeps the fastest; histograms are cleared afterwards.
int hbt_autotune(hbt_ctx *c, hbt_tuning *best_out) {
    if (!c) return HBT_ERR_ARG;
    if (!c->batch_ready) return fail(c, HBT_ERR_STATE, "no prepared batch (hbt_prepare_batch)");
    if (int rc = quiesce(c)) return rc;
    if (c->cfg.kernel != HBT_KERNEL_FAST) {
        if (best_out) *best_out = c->tune;
        return HBT_OK;
    }
    // histogram privatisation (binning 2): every launch shape is timed with
    // device-scope atomics and with slab binning, and the fastest pair wins.
    // Each is launched twice and the second launch counts: slab regions are
    // sized from the demand of the launch before.
    const bool both = c->cfg.binning == 2 && slab_applies(c);
    if (c->cfg.binning == 2) c->slab_binning = false;
    // small j chunks and one window per mixed item give small batches
    // enough items to fill the chip; large ones amortize the per-item setup
    const int Rs[] = {1, 2, 4};
    const int JCs[] = {256, 1024, 4096};
    const int KCs[] = {1, 4, 16};
    const int BPCs[] = {4, 8, 16};
    // One timed launch per candidate is noisy (a few %), so the fastest few
    // candidates of the sweep are timed again, kRetime launches each, and
    // the lowest median wins.
    struct Cand {
        hbt_tuning t;
        bool slab;
    };
    std::vector<Cand> cands;
    auto use = [&](const hbt_tuning &t, bool slab) -> int {
        c->tune = t;
        ItemList items;
        uint64_t es;
        build_items(c, true, true, items, es);
        c->launch_enum_same = es;
        c->launch_enum_mixed = c->plan_valid ? c->plan_enum_mixed : 0;
        c->slab_binning = slab;
        return upload_items(c, items);
    };
    for (int R : Rs)
        for (int jc : JCs)
            for (int kc : KCs)
                for (int bpc : BPCs) {
                    const hbt_tuning t{R, jc, kc, bpc, 0.0};
                    int rc = use(t, c->slab_binning);
                    if (rc) return rc;
                    for (int m = 0; m < (both ? 2 : 1); m++) {
                        if (both) c->slab_binning = m == 1;
                        float ms = 0.f;
                        rc = timed_fast(c, &ms);
                        if (!rc) rc = timed_fast(c, &ms);
                        if (rc) return rc;
                        hbt_tuning tm = t;
                        tm.ms = ms;
                        cands.push_back(Cand{tm, c->slab_binning});
                    }
                }
    std::sort(cands.begin(), cands.end(),
              [](const Cand &x, const Cand &y) { return x.t.ms < y.t.ms; });
    constexpr int kFinalists = 4, kRetime = 5;
    hbt_tuning best = cands.front().t;
    bool best_slab = cands.front().slab;
    double best_med = 1e30;
    for (size_t k = 0; k < cands.size() && k < (size_t)kFinalists; k++) {
        int rc = use(cands[k].t, cands[k].slab);
        float ms = 0.f;
        if (!rc) rc = timed_fast(c, &ms);  // regions sized from this shape's demand
        std::vector<float> v;
        for (int r = 0; r < kRetime && !rc; r++) {
            rc = timed_fast(c, &ms);
            v.push_back(ms);
        }
        if (rc) return rc;
        std::sort(v.begin(), v.end());
        if (v[kRetime / 2] < best_med) {
            best_med = v[kRetime / 2];
            best = cands[k].t;
            best.ms = best_med;
            best_slab = cands[k].slab;
        }
    }
    c->tune = best;
    c->slab_binning = best_slab;
    int rc = prepare_and_capture(c);
    if (!rc) rc = zero_hist(c);
    if (rc) return rc;
    if (best_out) *best_out = best;
    return HBT_OK;
}

int hbt_last_kernel_ms(hbt_ctx *c, double *same_ms, double *mixed_ms) {
    if (!c) return HBT_ERR_ARG;
    if (same_ms) *same_ms = c->ms_same;
    if (mixed_ms) *mixed_ms = c->ms_mixed;
    return HBT_OK;
}

int hbt_read(hbt_ctx *c, hbt_result *o) {
    if (!c || !o) return HBT_ERR_ARG;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    {
        const int rc = resolve_deferred(c);
        if (rc) return rc;
    }
    const size_t q8 = (size_t)c->nq * 8;
    if (o->num_count || o->den || o->q_out_sum || o->q_side_sum || o->q_long_sum || o->num) {
        std::vector<double> aos((size_t)c->nbins * BIN_STRIDE);
        HIPCHECK(c, hipMemcpy(aos.data(), c->d_bins, aos.size() * 8, hipMemcpyDeviceToHost));
        for (long k = 0; k < c->nbins; k++) {
            const double *b = aos.data() + k * BIN_STRIDE;
            if (o->num_count) o->num_count[k] = (uint64_t)b[BIN_COUNT];
            if (o->den) std::memcpy(&o->den[k], b + BIN_DEN, 8);
            if (o->q_out_sum) o->q_out_sum[k] = b[BIN_QO];
            if (o->q_side_sum) o->q_side_sum[k] = b[BIN_QS];
            if (o->q_long_sum) o->q_long_sum[k] = b[BIN_QL];
            if (o->num) o->num[k] = b[BIN_NUM];
        }
    }
    struct {
        void *dst;
        const void *src;
    } cp[] = {{o->qinv_count, c->d_qinv_count}, {o->qinv_den, c->d_qinv_den},
              {o->qinv_sum, c->d_qinv_sum},     {o->qinv_num, c->d_qinv_num}};
    for (auto &x : cp)
        if (x.dst) HIPCHECK(c, hipMemcpy(x.dst, x.src, q8, hipMemcpyDeviceToHost));
    const int nKT = c->cfg.n_KT, nphi = c->P.nphi;
    const long per = (long)c->P.qnpts * c->P.qnpts * c->P.qnpts;
    std::vector<unsigned long long> sn, sd, qn, qd;
    int rc = group_sums(c, (const double *)(c->d_bins + BIN_COUNT), per, nKT * nphi,
                        BIN_STRIDE, sn);
    if (!rc)
        rc = group_sums(c, (const unsigned long long *)(c->d_bins + BIN_DEN), per, nKT * nphi,
                        BIN_STRIDE, sd);
    if (!rc) rc = group_sums(c, c->d_qinv_count, c->P.qnpts, nKT, 1, qn);
    if (!rc) rc = group_sums(c, c->d_qinv_den, c->P.qnpts, nKT, 1, qd);
    if (rc) return rc;
    for (int k = 0; k < nKT; k++) {
        uint64_t a = 0, b = 0;
        for (int p = 0; p < nphi; p++) {
            a += sn[k * nphi + p];
            b += sd[k * nphi + p];
        }
        // with azimuthal_flag=1 the reference leaves the K_T counters at 0
        if (o->cnt_num_kt) o->cnt_num_kt[k] = c->P.azimuthal ? 0 : a;
        if (o->cnt_den_kt) o->cnt_den_kt[k] = c->P.azimuthal ? 0 : b;
        if (o->cnt_num_qinv) o->cnt_num_qinv[k] = qn[k];
        if (o->cnt_den_qinv) o->cnt_den_qinv[k] = qd[k];
    }
    if (c->P.azimuthal) {
        for (int k = 0; k < nKT * nphi; k++) {
            if (o->cnt_num_ktphi) o->cnt_num_ktphi[k] = sn[k];
            if (o->cnt_den_ktphi) o->cnt_den_ktphi[k] = sd[k];
        }
    }
    return HBT_OK;
}

int hbt_deferred_pairs(hbt_ctx *c, uint64_t *n) {
    if (!c || !n) return HBT_ERR_ARG;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (int rc = resolve_deferred(c)) return rc;
    *n = c->deferred_total;
    return HBT_OK;
}

int hbt_stage_counts(hbt_ctx *c, uint64_t same[5], uint64_t mixed[5]) {
    if (!c) return HBT_ERR_ARG;
    unsigned long long st[10], nb[2];
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(st, c->d_stage, sizeof(st), hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(nb, c->d_binned, sizeof(nb), hipMemcpyDeviceToHost));
    const bool fast = c->cfg.kernel == HBT_KERNEL_FAST;
    if (same) {
        same[0] = c->enum_same_total;
        for (int k = 1; k < 5; k++) same[k] = fast ? 0 : st[k];
        if (fast) same[4] = nb[0];
    }
    if (mixed) {
        mixed[0] = c->enum_mixed_total;
        for (int k = 1; k < 5; k++) mixed[k] = fast ? 0 : st[5 + k];
        if (fast) mixed[4] = nb[1];
    }
    return HBT_OK;
}

// ---------------------------------------------------------------- RNG

int hbt_rng_create(int32_t seed, hbt_rng **out) {
    if (!out) return HBT_ERR_ARG;
    hbt_rng *r = new hbt_rng();
    int s = seed;
    if (s == -1) {  // Random.cpp:9-11
        std::random_device dev;
        s = dev();
    }
    r->gen.seed(s);
    *out = r;
    return HBT_OK;
}

int hbt_rng_destroy(hbt_rng *r) {
    delete r;
    return HBT_OK;
}

int32_t hbt_rng_int(hbt_rng *r) { return r->uint_(r->gen); }
double hbt_rng_uniform(hbt_rng *r) { return r->uni(r->gen); }

int32_t hbt_rng_draw_plan(hbt_rng *r, int32_t nev, int32_t mixed_nev,
                          int32_t number_of_mixed_events, int32_t *ids, double *cs) {
    if (!r || nev < 0 || mixed_nev <= 0) return HBT_ERR_ARG;
    const int M = number_of_mixed_events > 0 ? number_of_mixed_events : mixed_nev / 2 + 1;
    for (int iev = 0; iev < nev; iev++) {
        for (int count = 0; count < M; count++) {  // HBT_correlation.cpp:206-215
            int id = r->uint_(r->gen) % mixed_nev;
            while (iev == id && mixed_nev != 1) id = r->uint_(r->gen) % mixed_nev;
            ids[(long long)iev * M + count] = id;
        }
        for (int k = 0; k < M; k++) {  // :495-497
            const double random_rotation = r->uni(r->gen) * 2 * M_PI;
            // The reference's g++ -O3 build evaluates this cos/sin pair with
            // glibc sincos(), which can differ from sin() in the last bit.
            double sn, cn;
            sincos(random_rotation, &sn, &cn);
            cs[2 * ((long long)iev * M + k)] = cn;
            cs[2 * ((long long)iev * M + k) + 1] = sn;
        }
    }
    return M;
}

// Every group of slot 0 is one of the reference's buffers; they are
// processed as consecutive calls of calculate_HBT_correlation_function
// (HBT_correlation.cpp:177-218): the plans are drawn group by group from the
// one stream, then all pairs of all groups run in one batch.
int hbt_calculate_buffer(hbt_ctx *c, hbt_rng *rng) {
    if (!c || !rng) return HBT_ERR_ARG;
    int rc = check_slots(c);
    if (rc) return rc;
    // the previous buffer's launch (already waited for by this buffer's
    // upload): its pairs left to the host are decided before this launch
    // adds its own
    rc = hbt_sync(c, nullptr);
    if (rc) return rc;
    Slot &s = c->slot[0];
    const Slot &m = mixed_slot(c);
    std::vector<int32_t> Mg(s.ngroups, 0);
    std::vector<int32_t> ids;
    std::vector<double> cs;
    for (int g = 0; g < s.ngroups; g++) {
        const int nev = s.grp_ev[g + 1] - s.grp_ev[g];
        const int mixed_nev = m.grp_ev[g + 1] - m.grp_ev[g];
        if (nev == 0) continue;  // no draws (the extra empty buffer at end of file)
        if (mixed_nev <= 0) return fail(c, HBT_ERR_STATE, "mixed-event buffer is empty");
        const int M = c->cfg.number_of_mixed_events > 0 ? c->cfg.number_of_mixed_events
                                                        : mixed_nev / 2 + 1;
        Mg[g] = M;
        const size_t base = ids.size();
        ids.resize(base + (size_t)nev * M);
        cs.resize(2 * (base + (size_t)nev * M));
        hbt_rng_draw_plan(rng, nev, mixed_nev, c->cfg.number_of_mixed_events, ids.data() + base,
                          cs.data() + 2 * base);
    }
    rc = prepare_batch(c, Mg.data(), ids.data(), cs.data(), false);
    // not waited for: the caller's next upload does its host work while the
    // launch runs (hbt_read, hbt_sync and the next upload wait for it)
    if (!rc) rc = hbt_launch_batch(c);
    return rc;
}

}  // extern "C"
