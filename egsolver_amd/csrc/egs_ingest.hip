// egs_ingest.hip -- a game in read order turned into the solver's internal game on the device.
//
// The reference's host driver renames vertices so that owner-0 vertices (token MAX) come
// first and owner-1 vertices follow, each block in read order (qsort with the default
// comparator, src/egsolver.c:146-214, stable under glibc), builds the internal CSR
// (src/egsolver.c:253-281) and sums MG = sum_v max(0, max_e -w_e) (src/egsolver.c:283-304).
// Here all three run as kernels after one upload of the read-order arrays:
//   check    offsets from 0 to m with an edge per vertex, successors in range;
//   rename   a scan of the owner-0 flags gives every read index its internal id (owner 0:
//            the owner-0 vertices before it; owner 1: n0 plus the owner-1 vertices before it);
//   rows     a scan of the renamed degrees gives the internal offsets, and each read row is
//            copied to its internal place with renamed successors;
//   MG       a reduction in 64 bits, refused past INT32_MAX as the reference refuses it.
#include "egs_device.h"
#include "egs_internal.h"

namespace egs {

__global__ __launch_bounds__(256) void k_rd_check(const int *__restrict__ rptr, const int *__restrict__ rsucc, int n,
                                                  int64_t m, int *__restrict__ bad) {
    int b = 0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        if (rptr[r + 1] <= rptr[r]) b |= 2;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
        if ((unsigned)rsucc[e] >= (unsigned)n) b |= 4;
    if (blockIdx.x == 0 && threadIdx.x == 0 && (rptr[0] != 0 || rptr[n] != m)) b |= 1;
    for (int d = 32; d >= 1; d >>= 1) b |= __shfl_xor(b, d, 64);
    if ((threadIdx.x & 63) == 0 && b) atomicOr(bad, b);
}

__global__ __launch_bounds__(256) void k_rd_flags(const uint8_t *__restrict__ owner, int n, int *__restrict__ flag0) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        flag0[r] = owner[r] == 0;
}

// pre0[r] = owner-0 vertices before read index r (pre0[n] = n0).
__global__ __launch_bounds__(256) void k_rd_rename(const uint8_t *__restrict__ owner, const int *__restrict__ rptr,
                                                   const int *__restrict__ pre0, int n, int *__restrict__ rev,
                                                   int *__restrict__ ext_of_int, int *__restrict__ deg) {
    const int n0 = pre0[n];
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int i = owner[r] == 0 ? pre0[r] : n0 + (int)r - pre0[r];
        rev[r] = i;
        ext_of_int[i] = (int)r;
        deg[i] = rptr[r + 1] - rptr[r];
    }
}

// Read row r -> internal row rev[r]: a lane per row of up to 64 edges, the wave for longer ones.
__global__ __launch_bounds__(256) void k_rd_rows(const int *__restrict__ rptr, const int *__restrict__ rsucc,
                                                 const int *__restrict__ rw, const int *__restrict__ rev,
                                                 const int *__restrict__ ptr, int n, int2 *__restrict__ edges) {
    const int lane = threadIdx.x & 63;
    for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x; r0 < n; r0 += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(r0 + threadIdx.x);
        int s = 0, e = 0, at = 0;
        if (r < n) {
            s = rptr[r];
            e = rptr[r + 1];
            at = ptr[rev[r]];
        }
        const bool small = e - s <= 64;
        if (small)
            for (int k = s; k < e; k++) edges[at + (k - s)] = make_int2(rev[rsucc[k]], rw[k]);
        for (unsigned long long heavy = __ballot(!small); heavy; heavy &= heavy - 1) {
            const int b = __ffsll(heavy) - 1;
            const int hs = __shfl(s, b, 64), he = __shfl(e, b, 64), ha = __shfl(at, b, 64);
            for (int k = hs + lane; k < he; k += 64) edges[ha + (k - hs)] = make_int2(rev[rsucc[k]], rw[k]);
        }
    }
}

__global__ __launch_bounds__(256) void k_rd_mg(const int *__restrict__ ptr, const int2 *__restrict__ edges, int n,
                                               unsigned long long *__restrict__ out) {
    unsigned long long sum = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        long long worst = 0;
        for (int k = ptr[v]; k < ptr[v + 1]; k++) worst = max(worst, -(long long)edges[k].y);
        sum += (unsigned long long)worst;
    }
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if ((threadIdx.x & 63) == 0 && sum) atomicAdd(out, sum);
}

hipError_t launch_read_order_check(const int *rptr, const int *rsucc, int n, int64_t m, int *bad, hipStream_t st) {
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(int), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rd_check, dim3(grid_for(std::max<int64_t>(n, m), 256, 4096)), dim3(256), 0, st, rptr, rsucc, n,
                       m, bad);
    return hipGetLastError();
}

hipError_t launch_internalise(const ReadOrderBuffers &b, int n, int *ptr, int2 *edges, unsigned long long *mg,
                              hipStream_t st) {
    const int grid = grid_for(n, 256, 4096);
    hipLaunchKernelGGL(k_rd_flags, dim3(grid), dim3(256), 0, st, b.owner, n, b.flag0);
    hipError_t e = launch_exclusive_scan(b.flag0, n, b.scan_tiles, b.pre0, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rd_rename, dim3(grid), dim3(256), 0, st, b.owner, b.rptr, (const int *)b.pre0, n, b.rev,
                       b.ext_of_int, b.deg);
    e = launch_exclusive_scan(b.deg, n, b.scan_tiles, ptr, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rd_rows, dim3(grid), dim3(256), 0, st, b.rptr, b.rsucc, b.rw, (const int *)b.rev,
                       (const int *)ptr, n, edges);
    e = hipMemsetAsync(mg, 0, sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rd_mg, dim3(grid), dim3(256), 0, st, (const int *)ptr, (const int2 *)edges, n, mg);
    return hipGetLastError();
}

}  // namespace egs
