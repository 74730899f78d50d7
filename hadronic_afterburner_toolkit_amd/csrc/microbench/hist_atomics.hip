// hist_atomics.hip -- throughput of the histogram accumulation patterns the
// pair kernels can use, on a histogram of the bench workload's size
// (41^3 q bins x 5 K_T bins = 344605 bins).  Every pattern adds the same
// per-pair payload (one u64 count + four FP64 sums, or one u64 for a mixed
// pair) at pseudo-random bins; all indices are hash % nbins, so every access
// is inside the nbins * 8 doubles allocated.
//
//   soa_5_per_lane     5 arrays, one pair per lane, 5 atomics per lane
//   aos64_5_per_lane   64-byte bins, one pair per lane
//   aos64_coop5        64-byte bins, 5 lanes per pair: 12 pairs per instruction
//                      (the fast kernel's same-event scatter)
//   u64_1_per_lane     one u64 per pair (the mixed-event denominator)
//   lds_4_per_lane     four FP64 LDS atomics per pair into a 16K-entry table
//   f64coop5           64-byte bins, 5 lanes per pair, count as an FP64 add:
//                      one instruction per 12 pairs (the fast kernel's scatter)
//   u32_1_per_lane     one u32 per pair
//   u32carry_1_per_lane  a u64 counter as a returning u32 add on its low word,
//                      the carry added to the high word when the add wraps
//   f64coop5_quarter   f64coop5 on a quarter of the workgroups (scaling with
//                      the number of issuing CUs)
//
// Locality of the same-event scatter (aos64_coop5 throughout):
//   coop5_bins_N       histogram of N 64-byte bins (from 4 MB-class L2 size to
//                      8x the workload's histogram)
//   coop5_xcd_private  one histogram copy per XCD, the copy chosen by the
//                      wave's XCC_ID hardware register
//   coop5_xcd_part     one histogram, each XCD touching only the bins
//                      b = xcc (mod 8)
//
// The fast kernel's two patterns at the histogram sizes of the K_phi
// workloads (f64coop5_bins_N, u32carry_bins_N): 41^3 x 5 x 16 bins (353 MB,
// the batched-uneven form) and 61^3 x 9 x 16 bins (2.1 GB, the
// large-histogram form).
//
// Prints one line per pattern: ms per launch and binned pairs/s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#define CK(x)                                                                       \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                               \
        }                                                                           \
    } while (0)

__device__ __forceinline__ unsigned mix32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ void add_u64(double *p) {
    __hip_atomic_fetch_add((unsigned long long *)p, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_soa(double *h, unsigned nbins, unsigned per) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (unsigned it = 0; it < per; it++) {
        const unsigned b = mix32(tid * 977u + it * 131071u) % nbins;
        add_u64(&h[b]);
        for (int v = 1; v < 5; v++) unsafeAtomicAdd(&h[(size_t)v * nbins + b], 0.1 * v);
    }
}

__global__ void k_aos(double *h, unsigned nbins, unsigned per) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (unsigned it = 0; it < per; it++) {
        double *p = h + (size_t)(mix32(tid * 977u + it * 131071u) % nbins) * 8;
        add_u64(p);
        for (int v = 1; v < 5; v++) unsafeAtomicAdd(p + v, 0.1 * v);
    }
}

__global__ void k_coop5(double *h, unsigned nbins, unsigned per) {
    const unsigned lane = threadIdx.x & 63u;
    if (lane >= 60) return;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned pair = wave * 12u + lane / 5u, sub = lane % 5u;
    for (unsigned it = 0; it < per; it++) {
        double *p = h + (size_t)(mix32(pair * 977u + it * 131071u) % nbins) * 8;
        if (sub == 0)
            add_u64(p);
        else
            unsafeAtomicAdd(p + sub, 0.1 * sub);
    }
}

__global__ void k_f64coop5(double *h, unsigned nbins, unsigned per) {
    const unsigned lane = threadIdx.x & 63u;
    if (lane >= 60) return;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned pair = wave * 12u + lane / 5u, sub = lane % 5u;
    for (unsigned it = 0; it < per; it++) {
        double *p = h + (size_t)(mix32(pair * 977u + it * 131071u) % nbins) * 8;
        unsafeAtomicAdd(p + sub, sub == 0 ? 1.0 : 0.1 * sub);
    }
}

__global__ void k_u32(double *h, unsigned nbins, unsigned per) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned *u = (unsigned *)h;
    for (unsigned it = 0; it < per; it++)
        __hip_atomic_fetch_add(u + 2 * (mix32(tid * 977u + it * 131071u) % nbins), 1u,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// u64 counter as a u32 add on its low word, with the carry added to the
// high word by the lane whose add wrapped (returning atomic).
__global__ void k_u32carry(double *h, unsigned nbins, unsigned per) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned *u = (unsigned *)h;
    unsigned wraps = 0;
    for (unsigned it = 0; it < per; it++) {
        unsigned *lo = u + 2 * (mix32(tid * 977u + it * 131071u) % nbins);
        const unsigned old = __hip_atomic_fetch_add(lo, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0xffffffffu) {
            __hip_atomic_fetch_add(lo + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wraps++;
        }
    }
    (void)wraps;
}

// XCD of the executing wave (hardware register XCC_ID, bits 3:0), 0..7.
__device__ __forceinline__ unsigned xcc_id() {
    return (unsigned)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7u;
}

__global__ void k_coop5_private(double *h, unsigned nbins, unsigned per) {
    const unsigned lane = threadIdx.x & 63u;
    if (lane >= 60) return;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned pair = wave * 12u + lane / 5u, sub = lane % 5u;
    double *copy = h + (size_t)xcc_id() * nbins * 8;
    for (unsigned it = 0; it < per; it++) {
        double *p = copy + (size_t)(mix32(pair * 977u + it * 131071u) % nbins) * 8;
        if (sub == 0)
            add_u64(p);
        else
            unsafeAtomicAdd(p + sub, 0.1 * sub);
    }
}

__global__ void k_coop5_part(double *h, unsigned nbins, unsigned per) {
    const unsigned lane = threadIdx.x & 63u;
    if (lane >= 60) return;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned pair = wave * 12u + lane / 5u, sub = lane % 5u;
    const unsigned x = xcc_id(), n8 = nbins / 8u;
    for (unsigned it = 0; it < per; it++) {
        double *p = h + (size_t)((mix32(pair * 977u + it * 131071u) % n8) * 8u + x) * 8;
        if (sub == 0)
            add_u64(p);
        else
            unsafeAtomicAdd(p + sub, 0.1 * sub);
    }
}

__global__ void k_u64(double *h, unsigned nbins, unsigned per) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (unsigned it = 0; it < per; it++) add_u64(&h[mix32(tid * 977u + it * 131071u) % nbins]);
}

__global__ void k_lds(double *h, unsigned nbins, unsigned per) {
    __shared__ double t[4 * 4096];
    for (int i = threadIdx.x; i < 4 * 4096; i += blockDim.x) t[i] = 0.0;
    __syncthreads();
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (unsigned it = 0; it < per; it++) {
        const unsigned b = mix32(tid * 977u + it * 131071u) & 4095u;
        for (int v = 0; v < 4; v++) atomicAdd(&t[v * 4096 + b], 0.1 * (v + 1));
    }
    __syncthreads();
    if (threadIdx.x == 0) unsafeAtomicAdd(&h[blockIdx.x % nbins], t[blockIdx.x & 4095u]);
}

int main() {
    const unsigned nbins = 41u * 41u * 41u * 5u;
    const unsigned nbins_kphi = nbins * 16u, nbins_large = 61u * 61u * 61u * 9u * 16u;
    // room for 8 copies of the bench histogram, and for the largest one
    const size_t bytes = std::max((size_t)nbins * 8 * sizeof(double) * 8,
                                  (size_t)nbins_large * 8 * sizeof(double));
    double *h = nullptr;
    CK(hipMalloc(&h, bytes));
    CK(hipMemset(h, 0, bytes));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const unsigned blocks = 256 * 8, threads = 256, per = 64;
    const double lanes = (double)blocks * threads * per;
    struct Pattern {
        const char *name;
        void (*k)(double *, unsigned, unsigned);
        double pairs_per_lane_iter;
    } pats[] = {{"soa_5_per_lane", k_soa, 1.0},
                {"aos64_5_per_lane", k_aos, 1.0},
                {"aos64_coop5", k_coop5, 12.0 / 64.0},
                {"u64_1_per_lane", k_u64, 1.0},
                {"lds_4_per_lane", k_lds, 1.0}};
    const int reps = 5;
    for (const Pattern &p : pats) {
        hipLaunchKernelGGL(p.k, dim3(blocks), dim3(threads), 0, 0, h, nbins, per);
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(p.k, dim3(blocks), dim3(threads), 0, 0, h, nbins, per);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double per_launch = ms / reps;
        std::printf("%-18s %8.3f ms/launch  %.3e binned pairs/s\n", p.name, per_launch,
                    lanes * p.pairs_per_lane_iter / (per_launch * 1e-3));
        std::fflush(stdout);
    }
    const struct {
        const char *name;
        void (*k)(double *, unsigned, unsigned);
        unsigned nbins;
    } loc[] = {{"coop5_bins_16384", k_coop5, 16384u},
               {"coop5_bins_65536", k_coop5, 65536u},
               {"coop5_bins_344605", k_coop5, nbins},
               {"coop5_bins_2756840", k_coop5, nbins * 8u},
               {"coop5_xcd_private", k_coop5_private, nbins},
               {"coop5_xcd_part", k_coop5_part, nbins}};
    for (const auto &p : loc) {
        hipLaunchKernelGGL(p.k, dim3(blocks), dim3(threads), 0, 0, h, p.nbins, per);
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(p.k, dim3(blocks), dim3(threads), 0, 0, h, p.nbins, per);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double per_launch = ms / reps;
        std::printf("%-18s %8.3f ms/launch  %.3e binned pairs/s\n", p.name, per_launch,
                    lanes * (12.0 / 64.0) / (per_launch * 1e-3));
        std::fflush(stdout);
    }
    const struct {
        const char *name;
        void (*k)(double *, unsigned, unsigned);
        unsigned blocks;
        double pairs_per_lane_iter;
    } more[] = {{"f64coop5", k_f64coop5, blocks, 12.0 / 64.0},
                {"u32_1_per_lane", k_u32, blocks, 1.0},
                {"u32carry_1_per_lane", k_u32carry, blocks, 1.0},
                {"f64coop5_quarter", k_f64coop5, blocks / 4, 12.0 / 64.0}};
    for (const auto &p : more) {
        hipLaunchKernelGGL(p.k, dim3(p.blocks), dim3(threads), 0, 0, h, nbins, per);
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(p.k, dim3(p.blocks), dim3(threads), 0, 0, h, nbins, per);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double per_launch = ms / reps;
        const double n = (double)p.blocks * threads * per * p.pairs_per_lane_iter;
        std::printf("%-18s %8.3f ms/launch  %.3e binned pairs/s\n", p.name, per_launch,
                    n / (per_launch * 1e-3));
        std::fflush(stdout);
    }
    const struct {
        const char *name;
        void (*k)(double *, unsigned, unsigned);
        unsigned nbins;
        double pairs_per_lane_iter;
    } sized[] = {{"f64coop5_bins_5513680", k_f64coop5, nbins_kphi, 12.0 / 64.0},
                 {"u32carry_bins_5513680", k_u32carry, nbins_kphi, 1.0},
                 {"f64coop5_bins_32685264", k_f64coop5, nbins_large, 12.0 / 64.0},
                 {"u32carry_bins_32685264", k_u32carry, nbins_large, 1.0}};
    for (const auto &p : sized) {
        hipLaunchKernelGGL(p.k, dim3(blocks), dim3(threads), 0, 0, h, p.nbins, per);
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(p.k, dim3(blocks), dim3(threads), 0, 0, h, p.nbins, per);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double per_launch = ms / reps;
        std::printf("%-18s %8.3f ms/launch  %.3e binned pairs/s\n", p.name, per_launch,
                    lanes * p.pairs_per_lane_iter / (per_launch * 1e-3));
        std::fflush(stdout);
    }
    CK(hipFree(h));
    return 0;
}
