// hbt_bf.hip -- the balance-function pair histograms on the GPU
// (BalanceFunction.cpp:63-197): for the species a = particle_alpha,
// b = particle_beta and their antiparticles, the (Delta y, Delta phi)
// histograms of same-event pairs (a,b), (abar,bbar), (a,bbar), (abar,b) and
// of mixed pairs against one randomly chosen, randomly rotated event per
// event and histogram.
//
// Host: per-particle pT, phi, y (or eta) with the reference's expressions
// (particleSamples::boostParticles, particleSamples.cpp:442-468), the species
// lists with the pT cut applied (the pair loops skip those particles), and
// the mixing draws in the reference's order.  Device: one block per
// (histogram, stride of events) with the whole Bnpts x 20 histogram in LDS
// (u32 counts); the bin arithmetic is the reference's, in FP64 without
// contraction, so the counts equal the reference's exactly.  Blocks add
// their LDS histogram to the FP64 histogram in HBM once at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hbt_hip.h"

namespace {

constexpr int kBnphi = 20;          // BalanceFunction.cpp:33
constexpr int kMaxBins = 16384;     // LDS histogram: 64 KB of u32
constexpr int kThreads = 256;

struct BfRec {
    double phi, y;
};

struct BfGeom {
    double dphi, Bphi_min, drap, Brap_min;
    int Bnpts;
};

// The eight histograms' lists: same-event (A, B of event e) for k < 4,
// mixed (A of event e, B of the drawn event) for k >= 4.
struct BfLists {
    const BfRec *rec[6];        // a, abar, b, bbar, b_mix, bbar_mix
    const long long *off[6];    // event offsets of each list
    const int *mix_ev;          // [4][nev] drawn partner event
    const double *mix_rot;      // [4][nev] rotation
    int nev;
};

__device__ __forceinline__ void bf_bin(const BfGeom &g, unsigned *h, double dphi_local,
                                       double dy) {
    int phi_idx = ((int)floor((dphi_local - g.Bphi_min) / g.dphi)) % kBnphi;
    if (phi_idx < 0) phi_idx += kBnphi;
    if (fabs(dy) < 1e-10) return;
    if (dy < g.Brap_min) return;
    const int y_idx = (int)((dy - g.Brap_min) / g.drap);
    if (y_idx >= 0 && y_idx < g.Bnpts) atomicAdd(&h[y_idx * kBnphi + phi_idx], 1u);
}

__device__ void bf_flush(unsigned *h, int nbins, double *out) {
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
        const unsigned v = h[b];
        if (v) {
            atomicAdd(&out[b], (double)v);
            h[b] = 0u;
        }
    }
    __syncthreads();
}

// blockIdx.y = histogram k (0..7); blocks stride over the events.
__global__ __launch_bounds__(kThreads) void k_bf(BfGeom g, BfLists L, double *hist) {
    __shared__ unsigned h[kMaxBins];
    const int nbins = g.Bnpts * kBnphi;
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    // k: ab, abarbbar, abbar, abarb, then the mixed four; A = a or abar,
    // B = b or bbar of the same buffer (k < 4) or of the mixed list
    const int k = blockIdx.y;
    const int ia = k & 1;
    const int ib = (k < 4 ? 2 : 4) + (((k & 3) == 1 || (k & 3) == 2) ? 1 : 0);
    const BfRec *A = ia ? L.rec[1] : L.rec[0];
    const long long *Aoff = ia ? L.off[1] : L.off[0];
    const BfRec *B = ib == 2 ? L.rec[2] : ib == 3 ? L.rec[3] : ib == 4 ? L.rec[4] : L.rec[5];
    const long long *Boff = ib == 2 ? L.off[2] : ib == 3 ? L.off[3] : ib == 4 ? L.off[4] : L.off[5];
    const bool mixed = k >= 4;
    double *out = hist + (size_t)k * nbins;
    unsigned long long since_flush = 0;
    for (int e = blockIdx.x; e < L.nev; e += gridDim.x) {
        const long long a0 = Aoff[e], na = Aoff[e + 1] - a0;
        int eb = e;
        double rot = 0.0;
        if (mixed) {
            eb = L.mix_ev[(k - 4) * L.nev + e];
            rot = L.mix_rot[(k - 4) * L.nev + e];
        }
        const long long b0 = Boff[eb], nb = Boff[eb + 1] - b0;
        const long long np = na * nb;
        if (np == 0) continue;
        // u32 LDS counts: flush before a bin could pass 2^32 - 1
        if (since_flush + (unsigned long long)np > 0xffffffffull) {
            bf_flush(h, nbins, out);
            since_flush = 0;
        }
        since_flush += (unsigned long long)np;
        if (np <= 0xffffffffll) {  // pairs of the event, flattened (32-bit index)
            const unsigned nbu = (unsigned)nb, npu = (unsigned)np;
            for (unsigned p = threadIdx.x; p < npu; p += blockDim.x) {
                const unsigned i = p / nbu, j = p - i * nbu;
                const BfRec ra = A[a0 + i], rb = B[b0 + j];
                const double d = ra.phi - rb.phi;
                bf_bin(g, h, mixed ? d + rot : d, ra.y - rb.y);
            }
        } else {
            for (long long i = 0; i < na; i++) {
                const BfRec ra = A[a0 + i];
                for (long long j = threadIdx.x; j < nb; j += blockDim.x) {
                    const BfRec rb = B[b0 + j];
                    const double d = ra.phi - rb.phi;
                    bf_bin(g, h, mixed ? d + rot : d, ra.y - rb.y);
                }
            }
        }
    }
    bf_flush(h, nbins, out);
}

}  // namespace

struct hbt_bf_ctx {
    hbt_bf_config cfg;
    BfGeom g;
    hipStream_t stream = nullptr;
    double *d_hist = nullptr;
    size_t nbins = 0;
    int64_t N_b = 0, N_bbar = 0;
    void *d_rec[6] = {};
    size_t rec_cap[6] = {};
    long long *d_off[6] = {};
    size_t off_cap[6] = {};
    int *d_mix_ev = nullptr;
    double *d_mix_rot = nullptr;
    size_t mix_ev_cap = 0, mix_rot_cap = 0;
    int cu_count = 0;
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

// One species list of a buffer: (phi, y or eta) of the particles with the
// id (99999: all) that pass the pT cut, event offsets, and how many.
struct HostList {
    std::vector<BfRec> rec;
    std::vector<long long> off;
};

void build_list(const hbt_bf_config &cfg, int32_t nev, const int64_t *ev_off,
                const int32_t *monval, const double *const cols[5], int poi, HostList &L) {
    L.rec.clear();
    L.off.assign(1, 0);
    for (int e = 0; e < nev; e++) {
        for (int64_t i = ev_off[e]; i < ev_off[e + 1]; i++) {
            if (!(poi == 99999 || monval[i] == poi)) continue;
            const double mass = cols[0][i], px = cols[1][i], py = cols[2][i], pz = cols[3][i];
            // particleSamples.cpp:447-465 (rap_shift 0)
            const double pT = sqrt(px * px + py * py);
            if (pT < cfg.BpT_min || pT > cfg.BpT_max) continue;  // BalanceFunction.cpp:127-129
            const double mT = sqrt(pT * pT + mass * mass);
            const double y = cfg.rap_type == 0 ? asinh(pz / pT) : asinh(pz / mT);
            L.rec.push_back(BfRec{atan2(py, px), y});
        }
        L.off.push_back((long long)L.rec.size());
    }
}

}  // namespace

extern "C" {

int hbt_bf_create(const hbt_bf_config *cfg, hbt_bf_ctx **out) {
    if (!cfg || !out) return HBT_ERR_ARG;
    *out = nullptr;
    if (cfg->Bnpts < 2 || (long)cfg->Bnpts * kBnphi > kMaxBins)
        return bf_fail(nullptr, HBT_ERR_ARG, "Bnpts must lie in [2, 819]");
    if (cfg->rap_type != 0 && cfg->rap_type != 1)
        return bf_fail(nullptr, HBT_ERR_ARG, "rap_type must be 0 or 1");
    for (int code : {cfg->particle_alpha, cfg->particle_beta})
        if (code == 9999 || code == 9998 || code == -9998 || code == 9997 || code == -9997 ||
            code == 9996 || code == -9996)
            return bf_fail(nullptr, HBT_ERR_ARG,
                           "species classes need the PDG tables; give one Monte-Carlo id");
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
        hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount,
                              cfg->device) != hipSuccess ||
        hipMalloc((void **)&c->d_hist, 8 * c->nbins * sizeof(double)) != hipSuccess ||
        hipMemset(c->d_hist, 0, 8 * c->nbins * sizeof(double)) != hipSuccess)
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
    if (c->d_mix_ev) hipFree(c->d_mix_ev);
    if (c->d_mix_rot) hipFree(c->d_mix_rot);
    if (c->d_hist) hipFree(c->d_hist);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return HBT_OK;
}

const char *hbt_bf_last_error(const hbt_bf_ctx *c) {
    return c ? c->err.c_str() : g_bf_create_err.c_str();
}

int hbt_bf_reset(hbt_bf_ctx *c) {
    if (!c) return HBT_ERR_ARG;
    BFCHECK(c, hipMemsetAsync(c->d_hist, 0, 8 * c->nbins * sizeof(double), c->stream));
    BFCHECK(c, hipStreamSynchronize(c->stream));
    c->N_b = c->N_bbar = 0;
    return HBT_OK;
}

int hbt_bf_calculate(hbt_bf_ctx *c, int32_t nev, const int64_t *ev_off, const int32_t *monval,
                     const double *const cols[5], int32_t mix_nev, const int64_t *mix_off,
                     const int32_t *mix_monval, const double *const mix_cols[5], hbt_rng *rng) {
    if (!c || !rng || nev < 0) return HBT_ERR_ARG;
    if (nev > 0 && (!ev_off || !monval || !cols)) return bf_fail(c, HBT_ERR_ARG, "null buffer");
    const bool alias = mix_off == nullptr;
    if (alias) {
        mix_nev = nev;
        mix_off = ev_off;
        mix_monval = monval;
        mix_cols = cols;
    }
    if (nev > 0 && mix_nev < 1) return bf_fail(c, HBT_ERR_ARG, "no mixed events");
    BFCHECK(c, hipSetDevice(c->cfg.device));
    const hbt_bf_config &k = c->cfg;
    HostList L[6];
    build_list(k, nev, ev_off, monval, cols, k.particle_alpha, L[0]);
    build_list(k, nev, ev_off, monval, cols, -k.particle_alpha, L[1]);
    build_list(k, nev, ev_off, monval, cols, k.particle_beta, L[2]);
    build_list(k, nev, ev_off, monval, cols, -k.particle_beta, L[3]);
    build_list(k, mix_nev, mix_off, mix_monval, mix_cols, k.particle_beta, L[4]);
    build_list(k, mix_nev, mix_off, mix_monval, mix_cols, -k.particle_beta, L[5]);
    // N_b, N_bbar: b and bbar particles passing the pT cut
    // (BalanceFunction.cpp:68-69, :199-209)
    c->N_b += (int64_t)L[2].rec.size();
    c->N_bbar += (int64_t)L[3].rec.size();
    // mixing draws of the four mixed histograms, in call order, per event a
    // partner then a rotation (BalanceFunction.cpp:165-169)
    std::vector<int> mix_ev((size_t)4 * std::max(nev, 1));
    std::vector<double> mix_rot((size_t)4 * std::max(nev, 1));
    for (int h = 0; h < 4; h++)
        for (int e = 0; e < nev; e++) {
            mix_ev[(size_t)h * nev + e] = hbt_rng_int(rng) % mix_nev;
            mix_rot[(size_t)h * nev + e] = hbt_rng_uniform(rng) * 2. * M_PI;
        }
    if (nev == 0) return HBT_OK;
    BfLists lists{};
    for (int l = 0; l < 6; l++) {
        int rc = bf_ensure(c, (BfRec **)&c->d_rec[l], c->rec_cap[l], L[l].rec.size());
        if (!rc) rc = bf_ensure(c, &c->d_off[l], c->off_cap[l], L[l].off.size());
        if (rc) return rc;
        if (!L[l].rec.empty())
            BFCHECK(c, hipMemcpyAsync(c->d_rec[l], L[l].rec.data(), L[l].rec.size() * sizeof(BfRec),
                                      hipMemcpyHostToDevice, c->stream));
        BFCHECK(c, hipMemcpyAsync(c->d_off[l], L[l].off.data(), L[l].off.size() * sizeof(long long),
                                  hipMemcpyHostToDevice, c->stream));
        lists.rec[l] = (const BfRec *)c->d_rec[l];
        lists.off[l] = c->d_off[l];
    }
    int rc = bf_ensure(c, &c->d_mix_ev, c->mix_ev_cap, mix_ev.size());
    if (!rc) rc = bf_ensure(c, &c->d_mix_rot, c->mix_rot_cap, mix_rot.size());
    if (rc) return rc;
    BFCHECK(c, hipMemcpyAsync(c->d_mix_ev, mix_ev.data(), mix_ev.size() * sizeof(int),
                              hipMemcpyHostToDevice, c->stream));
    BFCHECK(c, hipMemcpyAsync(c->d_mix_rot, mix_rot.data(), mix_rot.size() * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    lists.mix_ev = c->d_mix_ev;
    lists.mix_rot = c->d_mix_rot;
    lists.nev = nev;
    // blocks per histogram: enough to fill the chip, at most one per event
    const int bx = std::max(1, std::min(nev, std::max(1, c->cu_count * 4 / 8)));
    hipLaunchKernelGGL(k_bf, dim3(bx, 8), dim3(kThreads), 0, c->stream, c->g, lists, c->d_hist);
    BFCHECK(c, hipGetLastError());
    BFCHECK(c, hipStreamSynchronize(c->stream));  // host lists are freed on return
    return HBT_OK;
}

int hbt_bf_read(hbt_bf_ctx *c, double *hist, int64_t *N_b, int64_t *N_bbar) {
    if (!c) return HBT_ERR_ARG;
    BFCHECK(c, hipStreamSynchronize(c->stream));
    if (hist)
        BFCHECK(c, hipMemcpy(hist, c->d_hist, 8 * c->nbins * sizeof(double),
                             hipMemcpyDeviceToHost));
    if (N_b) *N_b = c->N_b;
    if (N_bbar) *N_bbar = c->N_bbar;
    return HBT_OK;
}

}  // extern "C"
