// egs_transpose.hip -- CSR -> CSC of predecessors as a two-level stable counting sort.
//
// Replaces EG_gpu_traspose_graph + cuSPARSE csr2csc (src/dev_EG_alg.cu:277-290,
// src/csr2csc.cu:54-66).  The result equals the reference's host transpose
// (src/host_csr2csc.c:15-46) entry for entry: column u lists the sources of the edges into
// u in edge order, i.e. ascending source id.  No kernel uses a global atomic per edge:
//
//   source   the source vertex of every edge (one pass over the rows), kept in cpred's
//            space until the last kernel overwrites it;
//   count    per tile of kTrTile edges, an LDS histogram of the edges' destination buckets
//            (buckets of 2^bshift consecutive vertices), counts[tile][bucket];
//   scan     the counts transposed to [bucket][tile] (LDS tiles, coalesced) and scanned
//            once: every (bucket, tile) gets its start, buckets following each other and
//            within a bucket the tiles in edge order; transposed back for the scatter;
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

// Source of every edge, written into src[m] (cpred's space).  A lane takes a row of up to
// 64 edges; longer rows are written by the whole wave, so a hub does not hold one lane.
__global__ __launch_bounds__(256) void k_tr_src(const int *__restrict__ ptr, int n, int *__restrict__ src) {
    const int lane = threadIdx.x & 63;
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x; v0 < n; v0 += (int64_t)gridDim.x * blockDim.x) {
        const int v = (int)(v0 + threadIdx.x);
        int s = 0, e = 0;
        if (v < n) { s = ptr[v]; e = ptr[v + 1]; }
        const bool small = e - s <= 64;
        if (small)
            for (int k = s; k < e; k++) src[k] = v;
        for (unsigned long long heavy = __ballot(!small); heavy; heavy &= heavy - 1) {
            const int b = __ffsll(heavy) - 1;
            const int hv = __shfl(v, b, 64), hs = __shfl(s, b, 64), he = __shfl(e, b, 64);
            for (int k = hs + lane; k < he; k += 64) src[k] = hv;
        }
    }
}

__global__ __launch_bounds__(256) void k_tr_count(const int2 *__restrict__ edges, int64_t m, int bshift, int nb,
                                                  int tiles, int *__restrict__ counts) {
    __shared__ int hist[kTrMaxBuckets];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * kTrTile, e1 = min(m, e0 + kTrTile);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) atomicAdd(&hist[edges[e].x >> bshift], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) counts[(size_t)blockIdx.x * nb + i] = hist[i];
}

// out[c][r] = in[r][c] for an R x C int matrix, through 64 x 64 LDS tiles (both sides
// coalesced).
__global__ __launch_bounds__(256) void k_tr_matrix(const int *__restrict__ in, int R, int C, int *__restrict__ out) {
    __shared__ int t[64][65];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int j = ty; j < 64; j += 4)
        if (r0 + j < R && c0 + tx < C) t[j][tx] = in[(size_t)(r0 + j) * C + c0 + tx];
    __syncthreads();
    for (int j = ty; j < 64; j += 4)
        if (c0 + j < C && r0 + tx < R) out[(size_t)(c0 + j) * R + r0 + tx] = t[tx][j];
}

constexpr int kTrU = 4;   // chunks of 64 whose loads are in flight together

__global__ __launch_bounds__(256) void k_tr_scatter(const int *__restrict__ src, const int2 *__restrict__ edges,
                                                    int64_t m, int bshift, int nb, int tiles,
                                                    const int *__restrict__ offs, int2 *__restrict__ items) {
    extern __shared__ int pos_lds[];               // [kTrWaves][nb]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int *pos = pos_lds + wave * nb;
    const int bits = 32 - __clz(nb - 1 > 0 ? nb - 1 : 1);
    const int64_t e0 = (int64_t)blockIdx.x * kTrTile, e1 = min(m, e0 + kTrTile);
    const int64_t per = (e1 - e0 + kTrWaves - 1) / kTrWaves;
    const int64_t w0 = min(e1, e0 + wave * per), w1 = min(e1, w0 + per);
    for (int i = threadIdx.x; i < kTrWaves * nb; i += blockDim.x) pos_lds[i] = 0;
    __syncthreads();
    for (int64_t c0 = w0; c0 < w1; c0 += 64 * kTrU) {
        int d[kTrU];
#pragma unroll
        for (int q = 0; q < kTrU; q++) d[q] = c0 + q * 64 + lane < w1 ? edges[c0 + q * 64 + lane].x : -1;
#pragma unroll
        for (int q = 0; q < kTrU; q++)
            if (d[q] >= 0) atomicAdd(&pos[d[q] >> bshift], 1);
    }
    __syncthreads();
    // each wave's run of a bucket starts after the tile's earlier waves' runs of it
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        int base = offs[(size_t)blockIdx.x * nb + b];
        for (int w = 0; w < kTrWaves; w++) {
            const int c = pos_lds[w * nb + b];
            pos_lds[w * nb + b] = base;
            base += c;
        }
    }
    __syncthreads();
    const int dmask = (1 << bshift) - 1;
    for (int64_t c0 = w0; c0 < w1; c0 += 64 * kTrU) {
        int2 ew[kTrU];
        int sv[kTrU];
#pragma unroll
        for (int q = 0; q < kTrU; q++) {
            const int64_t e = c0 + q * 64 + lane;
            ew[q] = e < w1 ? edges[e] : make_int2(0, 0);
            sv[q] = e < w1 ? src[e] : 0;
        }
#pragma unroll
        for (int q = 0; q < kTrU; q++) {
            const bool valid = c0 + q * 64 + lane < w1;
            const int bk = ew[q].x >> bshift;
            const unsigned long long eq = match_lanes(valid ? bk : -1, bits + 1);
            const int base = valid ? pos[bk] : 0;
            if (valid) items[base + __popcll(eq & lanes_below())] = make_int2(sv[q], ew[q].x & dmask);
            if (valid && (eq & lanes_below()) == 0) pos[bk] = base + __popcll(eq);
        }
    }
}

// One workgroup per bucket: cptr of its destinations and their predecessor lists.
__global__ __launch_bounds__(256) void k_tr_bucket(const int2 *__restrict__ items, int n, int64_t m, int bshift,
                                                   int tiles, const int *__restrict__ offs, int *__restrict__ cptr,
                                                   int *__restrict__ cpred) {
    extern __shared__ int cd[];                   // [kTrWaves][width]
    __shared__ int lds[4];
    const int width = 1 << bshift;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x;
    const int64_t i0 = offs[(size_t)b * tiles];
    const int64_t i1 = b + 1 < (int)gridDim.x ? offs[(size_t)(b + 1) * tiles] : m;
    const int64_t per = (i1 - i0 + kTrWaves - 1) / kTrWaves;
    const int64_t w0 = min(i1, i0 + wave * per), w1 = min(i1, w0 + per);
    for (int i = threadIdx.x; i < kTrWaves * width; i += blockDim.x) cd[i] = 0;
    __syncthreads();
    for (int64_t c0 = w0; c0 < w1; c0 += 64 * kTrU) {
        int d[kTrU];
#pragma unroll
        for (int q = 0; q < kTrU; q++) d[q] = c0 + q * 64 + lane < w1 ? items[c0 + q * 64 + lane].y : -1;
#pragma unroll
        for (int q = 0; q < kTrU; q++)
            if (d[q] >= 0) atomicAdd(&cd[wave * width + d[q]], 1);
    }
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
    }
    __syncthreads();
    if (b == gridDim.x - 1 && threadIdx.x == 0) cptr[n] = (int)m;
    for (int64_t c0 = w0; c0 < w1; c0 += 64 * kTrU) {
        int2 it[kTrU];
#pragma unroll
        for (int q = 0; q < kTrU; q++) {
            const int64_t i = c0 + q * 64 + lane;
            it[q] = i < w1 ? items[i] : make_int2(0, 0);
        }
#pragma unroll
        for (int q = 0; q < kTrU; q++) {
            const bool valid = c0 + q * 64 + lane < w1;
            const unsigned long long eq = match_lanes(valid ? it[q].y : -1, bshift + 1);
            const int base = valid ? cd[wave * width + it[q].y] : 0;
            if (valid) cpred[base + __popcll(eq & lanes_below())] = it[q].x;
            if (valid && (eq & lanes_below()) == 0) cd[wave * width + it[q].y] = base + __popcll(eq);
        }
    }
}

// Bucket width 2^bshift: at least 2048 destinations and at most kTrMaxBuckets buckets.
int transpose_bshift(int n) {
    int s = 11;
    while (((int64_t)n + (1LL << s) - 1) >> s > kTrMaxBuckets) s++;
    return s;
}

int64_t transpose_counts_len(int n, int64_t m) {
    const int s = transpose_bshift(n);
    return ((((int64_t)n + (1LL << s) - 1) >> s) * ((m + kTrTile - 1) / kTrTile));
}

bool transpose_sorted_fits(int n, int64_t m, int64_t counts_cap, int64_t items_cap) {
    return (1 << transpose_bshift(n)) <= kTrMaxWidth && transpose_counts_len(n, m) <= counts_cap &&
           m <= items_cap && n > 0 && m > 0 && m < 0x7fffffff;
}

hipError_t launch_transpose_sorted(const DevGame &g, int *counts, int *offs, int *scan_tiles, int2 *items,
                                   hipStream_t st) {
    const int n = g.n;
    const int64_t m = g.m;
    const int bshift = transpose_bshift(n);
    const int nb = (int)(((int64_t)n + (1LL << bshift) - 1) >> bshift);
    const int tiles = (int)((m + kTrTile - 1) / kTrTile);
    int *src = const_cast<int *>(g.cpred);        // sources until k_tr_bucket writes cpred
    hipLaunchKernelGGL(k_tr_src, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, g.ptr, n, src);
    // counts[tile][bucket] -> offs as [bucket][tile] -> scanned into counts ([bucket][tile])
    // -> offs as [tile][bucket], the order the scatter reads
    hipLaunchKernelGGL(k_tr_count, dim3(tiles), dim3(256), 0, st, g.edges, m, bshift, nb, tiles, counts);
    hipLaunchKernelGGL(k_tr_matrix, dim3((nb + 63) / 64, (tiles + 63) / 64), dim3(256), 0, st, (const int *)counts,
                       tiles, nb, offs);
    hipError_t e = launch_exclusive_scan(offs, nb * tiles, scan_tiles, counts, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tr_matrix, dim3((tiles + 63) / 64, (nb + 63) / 64), dim3(256), 0, st, (const int *)counts,
                       nb, tiles, offs);
    hipLaunchKernelGGL(k_tr_scatter, dim3(tiles), dim3(256), sizeof(int) * (size_t)kTrWaves * nb, st, (const int *)src,
                       g.edges, m, bshift, nb, tiles, (const int *)offs, items);
    const size_t lds = sizeof(int) * (size_t)kTrWaves << bshift;
    hipLaunchKernelGGL(k_tr_bucket, dim3(nb), dim3(256), lds, st, (const int2 *)items, n, m, bshift, tiles,
                       (const int *)counts, const_cast<int *>(g.cptr), const_cast<int *>(g.cpred));
    return hipGetLastError();
}

}  // namespace egs
