// egs_calib.hip -- plain device copy used to calibrate bandwidth figures: its bytes are
// known exactly, so its measured time gives the achievable HBM rate and its profiler
// counters give the scale factor between counter readings and bytes.
#include <string.h>

#include "../../include/egs.h"
#include "egs_device.h"

namespace {
__global__ __launch_bounds__(256) void k_copy_u4(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                 int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x) {
        dst[i] = src[i];
    }
}
}  // namespace

extern "C" int egs_copy_bandwidth(int device, int64_t bytes, int reps, double *gbps) {
    if (bytes < 16 || reps < 1 || !gbps) return EGS_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return EGS_E_HIP;
    const int64_t n16 = bytes / 16;
    void *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, (size_t)n16 * 16) != hipSuccess) return EGS_E_NOMEM;
    if (hipMalloc(&dst, (size_t)n16 * 16) != hipSuccess) { (void)hipFree(src); return EGS_E_NOMEM; }
    hipStream_t st;
    hipEvent_t a, b;
    (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipMemsetAsync(src, 1, (size_t)n16 * 16, st);
    const int grid = 4096;
    hipLaunchKernelGGL(k_copy_u4, dim3(grid), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n16);
    (void)hipEventRecord(a, st);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy_u4, dim3(grid), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n16);
    (void)hipEventRecord(b, st);
    hipError_t err = hipEventSynchronize(b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    *gbps = (double)(2 * n16 * 16) * reps / (ms * 1e-3) / 1e9;   // read + write bytes
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipStreamDestroy(st);
    (void)hipFree(src);
    (void)hipFree(dst);
    return err == hipSuccess ? EGS_OK : EGS_E_HIP;
}
