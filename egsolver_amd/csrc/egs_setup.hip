// egs_setup.hip -- per-game preparation on the device: edge packing, the CSR->CSC
// transpose (in-degree count, LDS-tiled exclusive scan, predecessor scatter), static
// high-degree vertex lists, the fixpoint certificate and statistic reductions.
//
// Replaces EG_gpu_traspose_graph + cuSPARSE csr2csc (src/dev_EG_alg.cu:277-290,
// src/csr2csc.cu:54-66).  The predecessor weights the reference also transposes are never
// read by its kernels (src/dev_EG_alg.cu:283) and are not built here.
#include "egs_device.h"
#include "egs_internal.h"

namespace egs {

// ---------------------------------------------------------------------------------------
// setup kernels
// ---------------------------------------------------------------------------------------

__global__ void k_pack_edges(const int *__restrict__ succ, const int *__restrict__ w,
                             int2 *__restrict__ edges, int64_t m) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (int64_t)gridDim.x * blockDim.x) {
        edges[e] = make_int2(succ[e], w[e]);
    }
}

__global__ void k_indegree(const int2 *__restrict__ edges, int64_t m, int *__restrict__ deg) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (int64_t)gridDim.x * blockDim.x) {
        atomicAdd(&deg[edges[e].x], 1);
    }
}

// Exclusive scan of cnt[0..n) into out[0..n], out[n] = total.  Three kernels: per-tile
// sums, a single-block scan of the tile sums, then per-tile scans offset by the tile
// prefix.  A tile is SCAN_TILE = 256 threads x 16 items staged through LDS.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const int *__restrict__ cnt, int n,
                                                             int *__restrict__ tile_sum) {
    __shared__ int lds[SCAN_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t i = base + (int64_t)k * SCAN_THREADS + threadIdx.x;
        if (i < n) s += cnt[i];
    }
    int total;
    egs_block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(int *__restrict__ tile_sum, int tiles) {
    __shared__ int lds[SCAN_THREADS / 64];
    int carry = 0;
    for (int b = 0; b < tiles; b += SCAN_THREADS) {
        int i = b + threadIdx.x;
        int x = (i < tiles) ? tile_sum[i] : 0;
        int total;
        int ex = egs_block_exclusive_scan(x, lds, total);
        if (i < tiles) tile_sum[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int *__restrict__ cnt, int n,
                                                             const int *__restrict__ tile_off,
                                                             int *__restrict__ out) {
    __shared__ int items[SCAN_TILE];
    __shared__ int lds[SCAN_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    // coalesced load of the tile into LDS
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int local = k * SCAN_THREADS + threadIdx.x;
        int64_t i = base + local;
        items[local] = (i < n) ? cnt[i] : 0;
    }
    __syncthreads();
    // each thread scans SCAN_ITEMS consecutive items
    int run = 0;
    int mine[SCAN_ITEMS];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        mine[k] = run;
        run += items[threadIdx.x * SCAN_ITEMS + k];
    }
    int total;
    int off = egs_block_exclusive_scan(run, lds, total) + tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) items[threadIdx.x * SCAN_ITEMS + k] = off + mine[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int local = k * SCAN_THREADS + threadIdx.x;
        int64_t i = base + local;
        if (i < n) out[i] = items[local];
    }
    // the last tile also writes the grand total out[n] (items past n count as 0)
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tile_off[blockIdx.x] + total;
}

// Scatter every edge v->u into u's predecessor list.  A lane walks the rows of short
// vertices; rows longer than 64 are walked by the whole wave so one heavy vertex does
// not serialise on a single lane.
__global__ __launch_bounds__(256) void k_scatter_preds(const int *__restrict__ ptr,
                                                       const int2 *__restrict__ edges, int n,
                                                       int *__restrict__ fill,
                                                       int *__restrict__ cpred) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t v0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; v0 < n;
         v0 += waves * 64) {
        const int v = (int)(v0 + lane);
        int s = 0, e = 0;
        if (v < n) { s = ptr[v]; e = ptr[v + 1]; }
        const bool small = (e - s) <= 64;
        if (small) {
            for (int k = s; k < e; k++) {
                int pos = atomicAdd(&fill[edges[k].x], 1);
                cpred[pos] = v;
            }
        }
        uint64_t heavy = __ballot(!small);
        while (heavy) {
            const int src = __ffsll((unsigned long long)heavy) - 1;
            heavy &= heavy - 1;
            const int hv = (int)v0 + src;
            const int hs = __builtin_amdgcn_readlane(s, src);
            const int he = __builtin_amdgcn_readlane(e, src);
            for (int k = hs + lane; k < he; k += 64) {
                int pos = atomicAdd(&fill[edges[k].x], 1);
                cpred[pos] = hv;
            }
        }
    }
}


// Static lists of high-degree vertices, used by dense sweeps: big_deg < deg <= huge_deg
// go to the wavefront kernel, deg > huge_deg to the workgroup kernel.  One-time
// compaction: a workgroup gathers its vertices of each list in LDS (wave-aggregated) and
// reserves their places with one global atomic per list, since thousands of appends to
// the same two counters wait on each other (60 us on the skewed 2^20 game).  What does not
// fit the LDS buffers goes out directly.
constexpr int kClassifyLds = 1024;   // list entries a workgroup holds per list

__global__ __launch_bounds__(1024) void k_classify(const int *__restrict__ ptr, int n, int big_deg, int huge_deg,
                                                   int *__restrict__ sbig, int *__restrict__ shuge,
                                                   int *__restrict__ counts) {
    __shared__ int lbig[kClassifyLds], lhuge[kClassifyLds];
    __shared__ int cbig, chuge, obig, ohuge;
    if (threadIdx.x == 0) { cbig = 0; chuge = 0; }
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int d_n = 0;                              // the next step's degree loads while this one runs
    if ((int64_t)blockIdx.x * blockDim.x + threadIdx.x < n) {
        const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        d_n = ptr[u + 1] - ptr[u];
    }
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x; v0 < n; v0 += stride) {
        const int v = (int)(v0 + threadIdx.x);
        const int d = d_n;
        if (v0 + stride + threadIdx.x < n) d_n = ptr[v0 + stride + threadIdx.x + 1] - ptr[v0 + stride + threadIdx.x];
        const bool big = v < n && d > big_deg && d <= huge_deg, huge = v < n && d > huge_deg;
        // a place in the workgroup's buffer, or straight to the global list when it is full
        const uint64_t mb = __ballot(big), mh = __ballot(huge);
        const int lane = threadIdx.x & 63;
        int bb = 0, bh = 0;
        if (lane == 0 && mb) bb = atomicAdd(&cbig, __popcll(mb));
        if (lane == 0 && mh) bh = atomicAdd(&chuge, __popcll(mh));
        bb = __shfl(bb, 0, 64) + __popcll(mb & egs_lanemask_lt());
        bh = __shfl(bh, 0, 64) + __popcll(mh & egs_lanemask_lt());
        if (big && bb < kClassifyLds) lbig[bb] = v;
        if (huge && bh < kClassifyLds) lhuge[bh] = v;
        egs_append(sbig, &counts[0], big && bb >= kClassifyLds, v);
        egs_append(shuge, &counts[1], huge && bh >= kClassifyLds, v);
    }
    __syncthreads();
    const int nbig = min(cbig, kClassifyLds), nhuge = min(chuge, kClassifyLds);
    if (threadIdx.x == 0) {
        obig = nbig ? atomicAdd(&counts[0], nbig) : 0;
        ohuge = nhuge ? atomicAdd(&counts[1], nhuge) : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbig; i += blockDim.x) sbig[obig + i] = lbig[i];
    for (int i = threadIdx.x; i < nhuge; i += blockDim.x) shuge[ohuge + i] = lhuge[i];
}

// Fixpoint certificate: count vertices with lift(val)(v) != val(v).  A lane takes its own
// vertex when the out-degree is at most kViolLane; the wave then takes the larger ones
// together, one at a time, so no lane walks a hub's edge list alone.
constexpr int kViolLane = 64;

__global__ __launch_bounds__(256) void k_violations(DevGame g, unsigned long long *out) {
    unsigned long long bad = 0;
    const int lane = threadIdx.x & 63;
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x; v0 < g.n; v0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = v0 + threadIdx.x;
        int s = 0, e = 0;
        if (v < g.n) { s = g.ptr[v]; e = g.ptr[v + 1]; }
        const bool wide = e - s > kViolLane;
        if (v < g.n && !wide) {
            const bool owner1 = v >= g.n0;
            int key = EGS_TOP;
            for (int k = s; k < e; k++) {
                const int2 ew = g.edges[k];
                key = min(key, egs_key(egs_ominus(g.val[ew.x], ew.y, g.mg), owner1));
            }
            bad += egs_unkey(key, owner1) != g.val[v];
        }
        for (uint64_t mask = __ballot(wide); mask; mask &= mask - 1) {
            const int b = __ffsll((unsigned long long)mask) - 1;
            const int vb = __shfl((int)v, b, 64), sb = __shfl(s, b, 64), eb = __shfl(e, b, 64);
            const bool owner1 = vb >= g.n0;
            int key = EGS_TOP;
            for (int k = sb + lane; k < eb; k += 64) {
                const int2 ew = g.edges[k];
                key = min(key, egs_key(egs_ominus(g.val[ew.x], ew.y, g.mg), owner1));
            }
            key = egs_group_min<64>(key);
            if (lane == 0) bad += egs_unkey(key, owner1) != g.val[vb];
        }
    }
    for (int d = 32; d >= 1; d >>= 1) bad += __shfl_xor(bad, d, 64);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(out, bad);
}

// Player 0's winning set as a bitmap: bit v of word v/32 is set when f(v) < TOP.
__global__ __launch_bounds__(256) void k_winning(const int *__restrict__ val, int n, unsigned *__restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v - (threadIdx.x & 63) < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long b = __ballot(v < n && val[v] != EGS_TOP);
        const int lane = threadIdx.x & 63;
        if ((lane == 0 || lane == 32) && v < n) out[v >> 5] = (unsigned)(b >> lane);
    }
}

hipError_t launch_winning(const int *val, int n, unsigned *out, hipStream_t st) {
    hipLaunchKernelGGL(k_winning, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, val, n, out);
    return hipGetLastError();
}

__global__ void k_degree_max(const int *__restrict__ ptr, int n, int *__restrict__ out) {
    int mx = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        mx = max(mx, ptr[v + 1] - ptr[v]);
    }
    for (int d = 32; d >= 1; d >>= 1) mx = max(mx, __shfl_xor(mx, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, mx);
}

// out[0..3]: the summed statistics; out[4]: the last round that ran; out[5]: the size of the
// next round's input list (round ring[14]; 0 = the least fixpoint), so that one read-back
// closes a run.
__global__ __launch_bounds__(1024) void k_sum_stats(DevGame g, const unsigned long long *__restrict__ slots,
                                                    int nslots, unsigned long long *__restrict__ out) {
    __shared__ unsigned long long part[4][16];
    if (threadIdx.x == 0) {
        const int *c = g.lcnt + (g.ring[14] % 3) * kListCounters;
        long long t = c[kSub];
        for (int k = 0; k < kSub; k++) t += min(c[k], g.sub_cap);
        out[4] = (unsigned long long)*g.rounds_done;
        out[5] = (unsigned long long)t;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long s[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nslots; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] += slots[(size_t)b * 4 + k];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        for (int d = 32; d >= 1; d >>= 1) s[k] += __shfl_xor(s[k], d, 64);
        if (lane == 0) part[k][wave] = s[k];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[threadIdx.x][w];
        out[threadIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------

int grid_for(int64_t work, int threads, int cap) {
    int64_t b = (work + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

// 32-bit packed edges ((w << shift) | succ, the batched and serving layout) to int2.
__global__ void k_unpack_edges(const uint32_t *__restrict__ packed, int shift, int2 *__restrict__ edges, int64_t m) {
    const uint32_t mask = (1u << shift) - 1u;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t p = packed[e];
        edges[e] = make_int2((int)(p & mask), ((int)p) >> shift);
    }
}

hipError_t launch_unpack(const uint32_t *packed, int shift, int2 *edges, int64_t m, hipStream_t st) {
    if (m > 0)
        hipLaunchKernelGGL(k_unpack_edges, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, packed, shift, edges, m);
    return hipGetLastError();
}

hipError_t launch_pack(const int *succ, const int *w, int2 *edges, int64_t m, hipStream_t st) {
    if (m > 0)
        hipLaunchKernelGGL(k_pack_edges, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, succ, w, edges, m);
    return hipGetLastError();
}

// Exclusive scan of in[0..n) into out[0..n] (out[n] = total); in == out is allowed (each
// tile is read whole into LDS before it is written).  tile_sum holds ceil(n / SCAN_TILE).
hipError_t launch_exclusive_scan(const int *in, int n, int *tile_sum, int *out, hipStream_t st) {
    const int tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 0) return hipMemsetAsync(out, 0, sizeof(int), st);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_THREADS), 0, st, in, n, tile_sum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, st, tile_sum, tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(SCAN_THREADS), 0, st, in, n, tile_sum, out);
    return hipGetLastError();
}

hipError_t launch_transpose(const DevGame &g, int *tile_sum, int *fill, hipStream_t st) {
    const int n = g.n;
    const int64_t m = g.m;
    hipError_t err;
    if ((err = hipMemsetAsync(fill, 0, sizeof(int) * (size_t)(n + 1), st)) != hipSuccess) return err;
    hipLaunchKernelGGL(k_indegree, dim3(grid_for(m, 256, 4096)), dim3(256), 0, st, g.edges, m, fill);
    const int tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_THREADS), 0, st, fill, n, tile_sum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, st, tile_sum, tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(SCAN_THREADS), 0, st, fill, n, tile_sum,
                       const_cast<int *>(g.cptr));
    if ((err = hipMemcpyAsync(fill, g.cptr, sizeof(int) * (size_t)(n + 1), hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return err;
    hipLaunchKernelGGL(k_scatter_preds, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, g.ptr, g.edges,
                       n, fill, const_cast<int *>(g.cpred));
    return hipGetLastError();
}

// Everything a prepare clears, in one launch (in place of a fill per array): the round
// statistics, the rounds counter, the list and deferral counters, the round modes (queue)
// and the ring slots, with ring[14], the number of the first round, set to 1.
__global__ __launch_bounds__(256) void k_prepare_reset(DevGame g) {
    const int64_t nstats = 4LL * kStatSlots;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nstats; i += (int64_t)gridDim.x * blockDim.x)
        g.stats[i] = 0;
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < 3 * kListCounters; i += blockDim.x) g.lcnt[i] = 0;
        if (threadIdx.x < 4) g.mode[threadIdx.x] = threadIdx.x < 3 ? MODE_QUEUE : 0;
        if (threadIdx.x < 16) g.ring[threadIdx.x] = threadIdx.x == 14 ? 1 : 0;
        if (threadIdx.x == 0) *g.rounds_done = 0;
    }
}

hipError_t launch_prepare_reset(const DevGame &g, hipStream_t st) {
    hipLaunchKernelGGL(k_prepare_reset, dim3(64), dim3(256), 0, st, g);
    return hipGetLastError();
}

// counts[0..1] (ring[9], ring[10]) start at zero (k_prepare_reset).
hipError_t launch_classify(const DevGame &g, int *counts, hipStream_t st) {
    hipLaunchKernelGGL(k_classify, dim3(grid_for(g.n, 1024, 512)), dim3(1024), 0, st, g.ptr, g.n, g.big_deg,
                       g.huge_deg, const_cast<int *>(g.sbig), const_cast<int *>(g.shuge), counts);
    return hipGetLastError();
}

hipError_t launch_violations(const DevGame &g, unsigned long long *out, hipStream_t st) {
    hipError_t err = hipMemsetAsync(out, 0, sizeof(unsigned long long), st);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_violations, dim3(grid_for(g.n, 256, 2048)), dim3(256), 0, st, g, out);
    return hipGetLastError();
}

hipError_t launch_degree_stats(const int *ptr, int n, int *out, hipStream_t st) {
    hipError_t err = hipMemsetAsync(out, 0, sizeof(int), st);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_degree_max, dim3(grid_for(n, 256, 1024)), dim3(256), 0, st, ptr, n, out);
    return hipGetLastError();
}

hipError_t launch_sum_stats(const DevGame &g, int nslots, unsigned long long *out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_stats, dim3(1), dim3(1024), 0, st, g, g.stats, nslots, out);
    return hipGetLastError();
}

}  // namespace egs
