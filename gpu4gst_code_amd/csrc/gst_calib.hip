// Measurement helper (not part of the solver ABI): a plain device copy of known size, used to
// calibrate the FETCH_SIZE / WRITE_SIZE counters against the algorithmic byte counts.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

__global__ void __launch_bounds__(256) gst_calib_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                             uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Copies `bytes` (rounded down to 16) device-to-device `iters` times; returns 0 on success.
extern "C" int gst_calib_copy(int device, uint64_t bytes, int iters) {
    if (hipSetDevice(device) != hipSuccess) return 1;
    const uint64_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return 2;
    if (hipMalloc(&b, n * 16) != hipSuccess) { (void)hipFree(a); return 2; }
    (void)hipMemset(a, 1, n * 16);
    for (int i = 0; i < iters; ++i) gst_calib_copy_kernel<<<2048, 256>>>(a, b, n);
    const hipError_t e = hipDeviceSynchronize();
    (void)hipFree(a);
    (void)hipFree(b);
    return e == hipSuccess ? 0 : 3;
}

// Copy with four independent 16-byte loads in flight per thread before their stores (the grid
// covers the array once, each workgroup a contiguous 64 KiB block).
typedef unsigned int calib_u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) gst_calib_copy4_kernel(const calib_u32x4* __restrict__ src,
                                                              calib_u32x4* __restrict__ dst, uint64_t n) {
    const uint64_t base = (uint64_t)blockIdx.x * 4 * blockDim.x + threadIdx.x;
    calib_u32x4 x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t i = base + (uint64_t)j * blockDim.x;
        if (i < n) x[j] = __builtin_nontemporal_load(&src[i]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t i = base + (uint64_t)j * blockDim.x;
        if (i < n) __builtin_nontemporal_store(x[j], &dst[i]);
    }
}

// Copy bandwidth of the device: one untimed copy, then `iters` copies of `bytes` timed with HIP
// events on the copy's stream; *gbs = bytes read plus bytes written per second / 1e9 (the better of
// the grid-stride copy and the four-loads-in-flight copy).  Returns 0 on success.
extern "C" int gst_calib_copy_gbs(int device, uint64_t bytes, int iters, double* gbs) {
    if (!gbs || iters < 1 || hipSetDevice(device) != hipSuccess) return 1;
    const uint64_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return 2;
    if (hipMalloc(&b, n * 16) != hipSuccess) { (void)hipFree(a); return 2; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = 3;
    if (rc == 0) {
        (void)hipMemset(a, 1, n * 16);
        *gbs = 0.0;
        const unsigned blocks4 = (unsigned)((n + 4 * 256 - 1) / (4 * 256));
        for (int kind = 0; kind < 2 && rc == 0; ++kind) {
            auto launch = [&]() {
                if (kind == 0) gst_calib_copy_kernel<<<4096, 256>>>(a, b, n);
                else gst_calib_copy4_kernel<<<blocks4, 256>>>((const calib_u32x4*)a, (calib_u32x4*)b, n);
            };
            launch();
            (void)hipEventRecord(e0, nullptr);
            for (int i = 0; i < iters; ++i) launch();
            (void)hipEventRecord(e1, nullptr);
            float ms = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f) rc = 3;
            else *gbs = std::max(*gbs, 2.0 * (double)(n * 16) * iters / (ms * 1e-3) / 1e9);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

// ---------------------------------------------------------------------------------------------
// Random-access rates: the ceiling of kernels whose accesses are scattered 8-byte words (the DP's
// grow rounds).  A table of `table_bytes` (rounded down to a power of two of u64 words) is accessed
// at `n` pseudo-random word indices (splitmix64 of the access number masked to the table size, so
// every index is in the table): 8-byte gathers, then 64-bit atomic-min (device scope).  Each access touches its own
// 64-byte line, so these rates, not bytes per second, bound such a kernel.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t calib_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// words: a power of two
__global__ void __launch_bounds__(256) gst_calib_gather_kernel(const unsigned long long* __restrict__ t, uint64_t words,
                                                               uint64_t n, unsigned long long* sink) {
    unsigned long long acc = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        acc += t[calib_mix(i) & (words - 1)];
    if (acc == 0x1234567ull) sink[0] = acc;   // keeps the loads; never true for the filled table
}
__global__ void __launch_bounds__(256) gst_calib_atomic_kernel(unsigned long long* t, uint64_t words, uint64_t n,
                                                               unsigned long long* sink) {
    unsigned long long acc = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        acc += atomicMin(&t[calib_mix(i) & (words - 1)], (unsigned long long)i);
    if (acc == 0x1234567ull) sink[0] = acc;
}
__global__ void __launch_bounds__(256) gst_calib_fill_kernel(unsigned long long* t, uint64_t words) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        t[i] = ~0ull;
}

// *gather_gops, *atomic_gops: billions of random 8-byte gathers / 64-bit atomic-mins per second
// over a table of table_bytes, best of `reps` timed launches of n accesses each.  Returns 0 on
// success.
extern "C" int gst_calib_random_gops(int device, uint64_t table_bytes, uint64_t n, int reps, double* gather_gops,
                                     double* atomic_gops) {
    if (!gather_gops || !atomic_gops || reps < 1 || n == 0 || table_bytes < 8 || hipSetDevice(device) != hipSuccess)
        return 1;
    uint64_t words = 1;
    while (words * 2 <= table_bytes / 8) words *= 2;
    unsigned long long *t = nullptr, *sink = nullptr;
    if (hipMalloc(&t, words * 8) != hipSuccess) return 2;
    if (hipMalloc(&sink, 64) != hipSuccess) { (void)hipFree(t); return 2; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = 3;
    *gather_gops = *atomic_gops = 0.0;
    for (int kind = 0; kind < 2 && rc == 0; ++kind) {
        for (int r = 0; r <= reps && rc == 0; ++r) {   // the first launch is untimed
            gst_calib_fill_kernel<<<2048, 256>>>(t, words);
            (void)hipEventRecord(e0, nullptr);
            if (kind == 0) gst_calib_gather_kernel<<<4096, 256>>>(t, words, n, sink);
            else gst_calib_atomic_kernel<<<4096, 256>>>(t, words, n, sink);
            (void)hipEventRecord(e1, nullptr);
            float ms = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f) {
                rc = 3;
                break;
            }
            if (r == 0) continue;
            double& best = kind == 0 ? *gather_gops : *atomic_gops;
            best = std::max(best, (double)n / (ms * 1e-3) / 1e9);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(t);
    (void)hipFree(sink);
    return rc;
}
