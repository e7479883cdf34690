// egs_transpose.hip -- CSR -> CSC of predecessors as a two-level stable counting sort.
//
// Replaces EG_gpu_traspose_graph + cuSPARSE csr2csc (src/dev_EG_alg.cu:277-290,
// src/csr2csc.cu:54-66).  The result equals the reference's host transpose
// (src/host_csr2csc.c:15-46) entry for entry: column u lists the sources of the edges into
// u in edge order, i.e. ascending source id.  No kernel uses a global atomic per edge:
//
//   count    per tile of kTrTile edges, an LDS histogram of the edges' destination buckets
//            (buckets of 2^bshift consecutive vertices) -> counts[tile][bucket];
//   scan     counts turned into each tile's start inside each bucket (a column scan over
//            the tiles), the bucket totals into bucket starts (one workgroup);
//   scatter  each tile moves its edges, as (source, destination within the bucket) items,
//            into its place in each bucket.  The tile is cut into one contiguous run per
//            wave; a wave's items for one bucket keep their edge order (a lane's rank among
//            the lanes of the same bucket comes from ballots over the bucket bits), so
//            every bucket receives its items in edge order;
//   bucket   one workgroup per bucket counts its destinations in LDS, writes cptr for them
//            and places each item at its destination's next slot, again in order.
// The predecessor weights the reference also transposes are never read by its kernels
// (src/dev_EG_alg.cu:283) and are not built.
#include "egs_device.h"
#include "egs_internal.h"

namespace egs {

constexpr int kTrTile = kTrTileEdges;  // edges per tile of the count and scatter kernels
constexpr int kTrWaves = 4;            // 256 threads
constexpr int kTrMaxBuckets = 4096;    // LDS histograms of the count and scatter kernels
constexpr int kTrMaxWidth = 4096;      // destinations per bucket (LDS counters of the bucket kernel)

// Lanes of this wave whose `key` (of `bits` significant bits) equals this lane's.
__device__ __forceinline__ unsigned long long match_lanes(int key, int bits) {
    unsigned long long eq = __builtin_amdgcn_read_exec();
    for (int b = 0; b < bits; b++) {
        const unsigned long long m = __ballot((key >> b) & 1);
        eq &= ((key >> b) & 1) ? m : ~m;
    }
    return eq;
}

__device__ __forceinline__ unsigned long long lanes_below() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__global__ __launch_bounds__(256) void k_tr_count(const int2 *__restrict__ edges, int64_t m, int bshift, int nb,
                                                  int *__restrict__ counts) {
    __shared__ int hist[kTrMaxBuckets];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * kTrTile, e1 = min(m, e0 + kTrTile);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) atomicAdd(&hist[edges[e].x >> bshift], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) counts[(size_t)blockIdx.x * nb + i] = hist[i];
}

// counts[t][b] -> start of tile t inside bucket b (relative to the bucket); tot[b] = bucket size.
__global__ __launch_bounds__(256) void k_tr_colscan(int *__restrict__ counts, int tiles, int nb, int *__restrict__ tot) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int run = 0;
    for (int t = 0; t < tiles; t++) {
        const int x = counts[(size_t)t * nb + b];
        counts[(size_t)t * nb + b] = run;
        run += x;
    }
    tot[b] = run;
}

// Exclusive scan of tot[0..nb) into bstart[0..nb], one workgroup.
__global__ __launch_bounds__(256) void k_tr_bucket_starts(const int *__restrict__ tot, int nb, int *__restrict__ bstart) {
    __shared__ int lds[4];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += blockDim.x) {
        const int i = b0 + threadIdx.x;
        const int x = i < nb ? tot[i] : 0;
        int total;
        const int ex = egs_block_exclusive_scan(x, lds, total);
        if (i < nb) bstart[i] = carry + ex;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) bstart[nb] = carry;
}

// First vertex whose row holds edge e: the largest v with ptr[v] <= e.
__device__ __forceinline__ int row_of(const int *ptr, int n, int64_t e) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ptr[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_tr_scatter(const int *__restrict__ ptr, const int2 *__restrict__ edges, int n,
                                                    int64_t m, int bshift, int nb, const int *__restrict__ counts,
                                                    const int *__restrict__ bstart, int2 *__restrict__ items) {
    __shared__ int pos[kTrWaves][kTrMaxBuckets];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int bits = 32 - __clz(nb - 1 > 0 ? nb - 1 : 1);
    const int64_t e0 = (int64_t)blockIdx.x * kTrTile, e1 = min(m, e0 + kTrTile);
    const int64_t per = (e1 - e0 + kTrWaves - 1) / kTrWaves;
    const int64_t w0 = min(e1, e0 + wave * per), w1 = min(e1, w0 + per);
    for (int i = threadIdx.x; i < kTrWaves * nb; i += blockDim.x) pos[i / nb][i % nb] = 0;
    __syncthreads();
    for (int64_t e = w0 + lane; e < w1; e += 64) atomicAdd(&pos[wave][edges[e].x >> bshift], 1);
    __syncthreads();
    // each wave's run of a bucket starts after the tile's earlier waves' runs of it
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        int base = bstart[b] + counts[(size_t)blockIdx.x * nb + b];
        for (int w = 0; w < kTrWaves; w++) {
            const int c = pos[w][b];
            pos[w][b] = base;
            base += c;
        }
    }
    __syncthreads();
    if (w0 >= w1) return;
    // rows: rc holds the chunk's first edge; the next 64 row starts mark where rows change
    int rc = 0;
    if (lane == 0) rc = row_of(ptr, n, w0);
    rc = __shfl(rc, 0, 64);
    for (int64_t c0 = w0; c0 < w1; c0 += 64) {
        const int64_t e = c0 + lane;
        const bool valid = e < w1;
        // row starts inside (c0, c0 + 64]: rows rc+1, rc+2, ... (every row holds an edge)
        const int r = rc + 1 + lane;
        const int64_t st = r <= n ? ptr[r] : INT64_MAX;
        const bool inside = st <= c0 + 64;
        unsigned long long mark = 0;
        if (inside && st < c0 + 64) mark = 1ull << (st - c0);
        for (int d = 32; d >= 1; d >>= 1) mark |= __shfl_xor(mark, d, 64);
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
        const int src = rc + __popcll(mark & upto);
        const int2 ew = valid ? edges[e] : make_int2(0, 0);
        const int bk = ew.x >> bshift;
        const unsigned long long eq = match_lanes(valid ? bk : -1, bits + 1);
        const int base = valid ? pos[wave][bk] : 0;
        if (valid) items[base + __popcll(eq & lanes_below())] = make_int2(src, ew.x & ((1 << bshift) - 1));
        const bool leader = valid && (eq & lanes_below()) == 0;
        if (leader) pos[wave][bk] = base + __popcll(eq);
        rc += (int)__popcll(__ballot(inside));
    }
}

// One workgroup per bucket: cptr of its destinations and their predecessor lists.
__global__ __launch_bounds__(256) void k_tr_bucket(const int2 *__restrict__ items, int n, int64_t m, int bshift,
                                                   const int *__restrict__ bstart, int *__restrict__ cptr,
                                                   int *__restrict__ cpred) {
    extern __shared__ int cd[];                   // [kTrWaves][width]
    __shared__ int lds[4];
    const int width = 1 << bshift;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x;
    const int64_t i0 = bstart[b], i1 = bstart[b + 1];
    const int64_t per = (i1 - i0 + kTrWaves - 1) / kTrWaves;
    const int64_t w0 = min(i1, i0 + wave * per), w1 = min(i1, w0 + per);
    for (int i = threadIdx.x; i < kTrWaves * width; i += blockDim.x) cd[i] = 0;
    __syncthreads();
    for (int64_t i = w0 + lane; i < w1; i += 64) atomicAdd(&cd[wave * width + items[i].y], 1);
    __syncthreads();
    // destinations in order: each wave's slots of a destination follow the earlier waves'
    int carry = (int)i0;
    for (int d0 = 0; d0 < width; d0 += blockDim.x) {
        const int d = d0 + threadIdx.x;
        int c[kTrWaves], deg = 0;
#pragma unroll
        for (int w = 0; w < kTrWaves; w++) {
            c[w] = cd[w * width + d];
            deg += c[w];
        }
        int total;
        const int ex = egs_block_exclusive_scan(deg, lds, total);
        int base = carry + ex;
        const int64_t u = ((int64_t)b << bshift) + d;
        if (u < n) cptr[u] = base;
#pragma unroll
        for (int w = 0; w < kTrWaves; w++) {
            cd[w * width + d] = base;
            base += c[w];
        }
        carry += total;
        __syncthreads();
    }
    if (b == gridDim.x - 1 && threadIdx.x == 0) cptr[n] = (int)m;
    const int bits = bshift;
    for (int64_t c0 = w0; c0 < w1; c0 += 64) {
        const int64_t i = c0 + lane;
        const bool valid = i < w1;
        const int2 it = valid ? items[i] : make_int2(0, 0);
        const unsigned long long eq = match_lanes(valid ? it.y : -1, bits + 1);
        const int base = valid ? cd[wave * width + it.y] : 0;
        if (valid) cpred[base + __popcll(eq & lanes_below())] = it.x;
        const bool leader = valid && (eq & lanes_below()) == 0;
        if (leader) cd[wave * width + it.y] = base + __popcll(eq);
    }
}

// Bucket width 2^bshift: at least 2048 destinations and at most kTrMaxBuckets buckets.
int transpose_bshift(int n) {
    int s = 11;
    while (((int64_t)n + (1LL << s) - 1) >> s > kTrMaxBuckets) s++;
    return s;
}

bool transpose_sorted_fits(int n, int64_t m, int64_t counts_cap, int64_t items_cap) {
    const int s = transpose_bshift(n);
    const int64_t nb = ((int64_t)n + (1LL << s) - 1) >> s, tiles = (m + kTrTile - 1) / kTrTile;
    return (1 << s) <= kTrMaxWidth && tiles * nb <= counts_cap && m <= items_cap && n > 0 && m > 0;
}

hipError_t launch_transpose_sorted(const DevGame &g, int *counts, int *tot, int *bstart, int2 *items, hipStream_t st) {
    const int n = g.n;
    const int64_t m = g.m;
    const int bshift = transpose_bshift(n);
    const int nb = (int)(((int64_t)n + (1LL << bshift) - 1) >> bshift);
    const int tiles = (int)((m + kTrTile - 1) / kTrTile);
    hipLaunchKernelGGL(k_tr_count, dim3(tiles), dim3(256), 0, st, g.edges, m, bshift, nb, counts);
    hipLaunchKernelGGL(k_tr_colscan, dim3((nb + 255) / 256), dim3(256), 0, st, counts, tiles, nb, tot);
    hipLaunchKernelGGL(k_tr_bucket_starts, dim3(1), dim3(256), 0, st, tot, nb, bstart);
    hipLaunchKernelGGL(k_tr_scatter, dim3(tiles), dim3(256), 0, st, g.ptr, g.edges, n, m, bshift, nb,
                       (const int *)counts, (const int *)bstart, items);
    const size_t lds = sizeof(int) * (size_t)kTrWaves << bshift;
    hipLaunchKernelGGL(k_tr_bucket, dim3(nb), dim3(256), lds, st, (const int2 *)items, n, m, bshift,
                       (const int *)bstart, const_cast<int *>(g.cptr), const_cast<int *>(g.cpred));
    return hipGetLastError();
}

}  // namespace egs
