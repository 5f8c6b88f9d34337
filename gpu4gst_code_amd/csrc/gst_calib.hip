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
