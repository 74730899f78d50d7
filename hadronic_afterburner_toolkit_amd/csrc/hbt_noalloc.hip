// hbt_noalloc.hip -- hbt_calculate_noalloc: the plain FP64 pair loops of one
// buffer (what hbt_upload + hbt_run_same + hbt_run_mixed do with
// HBT_KERNEL_PLAIN) for callers that cannot allocate device memory: code
// under HIP graph capture, tight serving loops.  The caller passes the
// particles, the mixing plan and the histogram in device memory; the
// rapidity cut's kept particles, their event offsets and the stage
// counters live in `scratch`, a file-scope device array of fixed size that
// each GPU holds until the process ends.  No hipMalloc, hipFree or other
// allocating call is made; every launch goes on the caller's stream.
//
// Calls on one GPU share its scratch array: calls that overlap on one GPU
// (two streams, two host threads) race and must be serialized by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/hbt_hip.h"
#include "hbt_device.h"
#include "hbt_internal.h"
#include "hbt_plain.h"

using namespace hbt;

namespace {

// Scratch layout (bytes): stage counters, kept counts and kept offsets per
// event, then the kept particles as eight FP64 columns.
constexpr long long kMaxParticles = HBT_NOALLOC_MAX_PARTICLES;
constexpr int kMaxEvents = HBT_NOALLOC_MAX_EVENTS;
constexpr size_t kStageOff = 0;                                   // 16 x u64
constexpr size_t kCountOff = 256;                                 // kMaxEvents x i64
constexpr size_t kKoffOff = kCountOff + 8 * (size_t)kMaxEvents;   // (kMaxEvents + 1) x i64
constexpr size_t kSoaOff = (kKoffOff + 8 * ((size_t)kMaxEvents + 1) + 255) / 256 * 256;
constexpr size_t kScratchBytes = kSoaOff + 8 * sizeof(double) * (size_t)kMaxParticles;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};  // device address of scratch, per device
std::string g_err;

int fail(int code, const char *msg) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_err = msg;
    return code;
}

// The device address of this GPU's scratch array, looked up once per device
// with that device current.  The symbol's host-side name is only a handle:
// kernels and copies take this address.
int scratch_ptr(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            g_err = "hipGetSymbolAddress(scratch) failed";
            return HBT_ERR_HIP;
        }
        g_ptr[dev] = (unsigned char *)p;
    }
    *out = g_ptr[dev];
    return HBT_OK;
}

__global__ void k_na_zero(unsigned long long *p, int n) {
    const int i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Kept particles per event: tanh(HBTrap_min) < pz/E < tanh(HBTrap_max)
// (HBT_correlation.cpp:255-266).
__global__ __launch_bounds__(256) void k_na_count(const long long *ev_off, const double *pz,
                                                  const double *E, double cut_min,
                                                  double cut_max, long long *cnt) {
    const int e = blockIdx.x;
    const long long b = ev_off[e], n = ev_off[e + 1] - b;
    unsigned long long c = 0;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
        const double r = pz[b + i] / E[b + i];
        c += (r > cut_min && r < cut_max) ? 1ull : 0ull;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    __shared__ unsigned long long part[4];
    if (__lane_id() == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[e] = (long long)(part[0] + part[1] + part[2] + part[3]);
}

// Exclusive scan of the kept counts into event offsets koff[0..nev], one
// block of 1024 threads, each summing a run of ceil(nev/1024) events.
__global__ __launch_bounds__(1024) void k_na_scan(const long long *cnt, int nev,
                                                  long long limit, long long *koff) {
    __shared__ long long s[1024];
    const int t = threadIdx.x;
    const int per = (nev + 1023) / 1024;
    const int a = min(nev, t * per), b = min(nev, a + per);
    long long sum = 0;
    for (int e = a; e < b; e++) sum += cnt[e];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
        const long long v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    // offsets are clamped to the kept columns' capacity, so a caller whose
    // n_particles understates its events cannot make a kernel read past them
    long long run = t > 0 ? s[t - 1] : 0;
    for (int e = a; e < b; e++) {
        koff[e] = min(run, limit);
        run += cnt[e];
    }
    if (t == 1023) koff[nev] = min(s[1023], limit);
}

// Order-preserving compaction of each event's kept particles to its kept
// offset (the reference's gather order: event by event, particle order).
__global__ __launch_bounds__(256) void k_na_compact(const long long *ev_off, DevSoA in,
                                                    double cut_min, double cut_max,
                                                    const long long *koff, double *out,
                                                    long long stride) {
    const int e = blockIdx.x;
    const long long b = ev_off[e], n = ev_off[e + 1] - b;
    __shared__ int wsum[4];
    long long base = koff[e];
    const int w = threadIdx.x >> 6;
    for (long long c0 = 0; c0 < n; c0 += 256) {
        const long long i = c0 + threadIdx.x;
        bool keep = false;
        if (i < n) {
            const double r = in.pz[b + i] / in.E[b + i];
            keep = r > cut_min && r < cut_max;
        }
        const unsigned long long m = __ballot(keep);
        const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (__lane_id() == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
        for (int k = 0; k < 4; k++) {
            before += k < w ? wsum[k] : 0;
            total += wsum[k];
        }
        const long long o = base + before + rank;
        if (keep && o < stride) {  // stride: the kept columns' capacity
            const long long s = b + i;
            out[0 * stride + o] = in.px[s];
            out[1 * stride + o] = in.py[s];
            out[2 * stride + o] = in.pz[s];
            out[3 * stride + o] = in.E[s];
            out[4 * stride + o] = in.x[s];
            out[5 * stride + o] = in.y[s];
            out[6 * stride + o] = in.z[s];
            out[7 * stride + o] = in.t[s];
        }
        base += total;
        __syncthreads();  // wsum is rewritten by the next chunk
    }
}

// Same-event tiles over all kept particles of the buffer; the grid is sized
// from the particles passed in, the kept count comes from koff[nev].
__global__ __launch_bounds__(TILE) void k_na_same(DevParams P, DevHist H, DevSoA A,
                                                  const long long *koff, int nev) {
    const long long n = koff[nev];
    same_plain_tile(P, H, A, 0, (int)n, 0.0, blockIdx.y, blockIdx.x);
}

// Mixed tiles of events [eb, eb + gridDim.z / M); partner ids outside
// [0, nev) form no pairs.
__global__ __launch_bounds__(TILE) void k_na_mixed(DevParams P, DevHist H, DevSoA A,
                                                   const long long *koff, const int *ids,
                                                   const double *cs, int M, int eb, int nev) {
    const long long row = (long long)eb * M + blockIdx.z;
    const int pid = ids[row];
    if (pid < 0 || pid >= nev) return;
    mixed_plain_tile(P, H, A, A, koff, koff, ids, cs, M, eb, (long long)eb * M, 0.0,
                     blockIdx.y, blockIdx.x, blockIdx.z);
}

}  // namespace

extern "C" {

const char *hbt_noalloc_last_error(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_err.c_str();
}

int hbt_calculate_noalloc(const hbt_config *cfg, int32_t nev, int64_t n_particles,
                          int64_t max_event, const int64_t *ev_off, const double *const soa[8],
                          int32_t M, const int32_t *partner_ids, const double *cos_sin,
                          double *bins, void *stream) {
    if (!cfg || !soa || !bins || (nev > 0 && !ev_off)) return fail(HBT_ERR_ARG, "null argument");
    Derived D;
    std::string err;
    if (int rc = derive(*cfg, D, err)) return fail(rc, "invalid configuration");
    if (cfg->kernel != HBT_KERNEL_PLAIN || cfg->precision != HBT_PREC_FP64)
        return fail(HBT_ERR_ARG, "hbt_calculate_noalloc runs the plain FP64 kernel only");
    if (cfg->azimuthal_flag || cfg->invariant_radius_flag)
        return fail(HBT_ERR_ARG, "hbt_calculate_noalloc bins the 3D histogram only "
                                 "(azimuthal_flag = invariant_radius_flag = 0)");
    if (!(cfg->needed_number_of_pairs >= 1e18))
        return fail(HBT_ERR_ARG, "hbt_calculate_noalloc needs the pair cap off "
                                 "(needed_number_of_pairs >= 1e18)");
    if (nev < 0 || nev > kMaxEvents)
        return fail(HBT_ERR_CAPACITY, "more events than the scratch array holds "
                                      "(HBT_NOALLOC_MAX_EVENTS)");
    if (n_particles < 0 || n_particles > kMaxParticles)
        return fail(HBT_ERR_CAPACITY, "more particles than the scratch array holds "
                                      "(HBT_NOALLOC_MAX_PARTICLES)");
    if (max_event < 0 || max_event > n_particles)
        return fail(HBT_ERR_ARG, "max_event must lie in [0, n_particles]");
    if (M < 0 || M > 64 * 1024 || (M > 0 && nev > 0 && (!partner_ids || !cos_sin)))
        return fail(HBT_ERR_ARG, "bad mixing plan");
    for (int a = 0; a < 8; a++)
        if (!soa[a] && n_particles > 0) return fail(HBT_ERR_ARG, "null particle column");
    int ndev = 0, cur = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev ||
        cfg->device >= kMaxDevices)
        return fail(HBT_ERR_NODEVICE, "bad device");
    if (hipGetDevice(&cur) != hipSuccess) return fail(HBT_ERR_HIP, "hipGetDevice failed");
    if (cur != cfg->device && hipSetDevice(cfg->device) != hipSuccess)
        return fail(HBT_ERR_NODEVICE, "hipSetDevice failed");
    int rc = HBT_OK;
    unsigned char *scr = nullptr;
    hipStream_t st = (hipStream_t)stream;
    do {
        rc = scratch_ptr(cfg->device, &scr);
        if (rc) break;
        unsigned long long *stage = (unsigned long long *)(scr + kStageOff);
        long long *cnt = (long long *)(scr + kCountOff);
        long long *koff = (long long *)(scr + kKoffOff);
        double *kept = (double *)(scr + kSoaOff);
        const long long stride = kMaxParticles;
        DevSoA A{kept, kept + stride, kept + 2 * stride, kept + 3 * stride,
                 kept + 4 * stride, kept + 5 * stride, kept + 6 * stride, kept + 7 * stride};
        DevSoA in{soa[0], soa[1], soa[2], soa[3], soa[4], soa[5], soa[6], soa[7]};
        DevHist H{};
        H.bins = bins;
        H.stage = stage;
        const double cut_max = tanh(cfg->HBTrap_max), cut_min = tanh(cfg->HBTrap_min);
        // the counters start at zero on every call (a fill kernel, not a
        // memset node, so the call can be captured in a graph)
        hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(64), 0, st, stage, 16);
        if (nev > 0) {
            hipLaunchKernelGGL(k_na_count, dim3(nev), dim3(256), 0, st, (const long long *)ev_off,
                               soa[2], soa[3], cut_min, cut_max, cnt);
            hipLaunchKernelGGL(k_na_scan, dim3(1), dim3(1024), 0, st, (const long long *)cnt,
                               (int)nev, stride, koff);
            hipLaunchKernelGGL(k_na_compact, dim3(nev), dim3(256), 0, st,
                               (const long long *)ev_off, in, cut_min, cut_max,
                               (const long long *)koff, kept, stride);
        }
        if (n_particles >= 2) {
            const int T = (int)((n_particles + TILE - 1) / TILE);
            hipLaunchKernelGGL(k_na_same, dim3(T, T), dim3(TILE), 0, st, D.P, H, A,
                               (const long long *)koff, (int)nev);
        }
        const int Tm = (int)((max_event + TILE - 1) / TILE);
        if (M > 0 && nev > 0 && Tm > 0) {
            const int per = std::max(1, 65535 / M);  // gridDim.z below 65536
            for (int eb = 0; eb < nev; eb += per) {
                const int ne = std::min(per, (int)nev - eb);
                hipLaunchKernelGGL(k_na_mixed, dim3(Tm, Tm, ne * M), dim3(TILE), 0, st, D.P, H, A,
                                   (const long long *)koff, partner_ids, cos_sin, (int)M, eb,
                                   (int)nev);
            }
        }
        if (hipGetLastError() != hipSuccess) rc = fail(HBT_ERR_HIP, "kernel launch failed");
    } while (false);
    if (cur != cfg->device) hipSetDevice(cur);
    return rc;
}

}  // extern "C"
