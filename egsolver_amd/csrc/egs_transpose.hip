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
//   scatter  one workgroup per tile sorts the tile's edges by bucket in LDS (stable: waves
//            take consecutive runs, lanes are ranked by ballots over the bucket bits) and
//            writes each bucket's run of (source, destination within the bucket) items with
//            consecutive lanes on consecutive slots; every bucket receives its items in edge
//            order.  With more than 2048 buckets one wave per tile scatters item by item;
//   bucket   one workgroup per bucket counts its destinations in LDS, writes cptr for them,
//            places each item at its destination's next slot, again in order, in LDS, and
//            writes the bucket's predecessor lists out in one coalesced pass; a bucket too
//            large for LDS is placed in HBM directly by a four-wave kernel.
// The predecessor weights the reference also transposes are never read by its kernels
// (src/dev_EG_alg.cu:283) and are not built.
#include "egs_device.h"
#include "egs_internal.h"

namespace egs {

constexpr int kTrTile = kTrTileEdges;  // edges per tile of the count and scatter kernels
constexpr int kTrWaves = 4;            // 256 threads
constexpr int kTrMaxBuckets = 4096;    // LDS histograms of the count and scatter kernels
constexpr int kTrMaxWidth = 4096;      // destinations per bucket (LDS counters of the bucket kernel)
constexpr int64_t kTrBucketLds = 150 * 1024;   // LDS bytes of the LDS bucket kernel at most

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

// Rows longer than kTrLong edges are written by the count kernel, a tile's share each, so
// that no wave walks a hub's edge list alone (a 70,000-edge row took one wave 1,091 serial
// stores).  The source kernel records, for every kTrLong-edge chunk start, the row that
// holds it: a long row contains at least one chunk start in every tile it touches, so a
// tile finds all of its long rows among the rows of its own chunk starts and the next
// tile's first one.
constexpr int kTrLongShift = 12;
constexpr int kTrLong = 1 << kTrLongShift;
constexpr int kTrTileChunks = kTrTileEdges >> kTrLongShift;      // chunk starts per tile
static_assert(kTrTileChunks * kTrLong == kTrTileEdges, "tiles hold whole chunks");

// Source of every edge, written into src[m] (cpred's space), except the edges of rows
// longer than kTrLong.  A wave takes 64 rows and writes their edges' sources in edge order,
// 64 consecutive edges per store.  A chunk of 64 edges inside one row takes that row
// directly; otherwise each lane finds the row of its edge among the 64 row starts the lanes
// hold (a binary search over shuffles; a row without edges is skipped because the search
// takes the last row starting at or before the edge).  rowof[c] gets the row that holds
// edge c * kTrLong.
__device__ __forceinline__ int row_in_wave(int start, int64_t e) {
    int l = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
        if (__shfl(start, l + step, 64) <= e) l += step;
    return l;
}

__global__ __launch_bounds__(256) void k_tr_src(const int *__restrict__ ptr, int n, int *__restrict__ src,
                                                int *__restrict__ rowof) {
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * 64;
    int64_t v0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64;
    // the next 64 rows' starts are loaded while this wave writes the current ones
    int start_n = 0, hi_n = 0;
    if (v0 < n) {
        start_n = ptr[min(v0 + lane, (int64_t)n)];          // rows past n start at m
        hi_n = ptr[min(v0 + 64, (int64_t)n)];
    }
    for (; v0 < n; v0 += stride) {
        const int start = start_n;
        const int lo = __builtin_amdgcn_readlane(start, 0);
        const int hi = __builtin_amdgcn_readfirstlane(hi_n);
        if (v0 + stride < n) {
            start_n = ptr[min(v0 + stride + lane, (int64_t)n)];
            hi_n = ptr[min(v0 + stride + 64, (int64_t)n)];
        }
        for (int64_t c0 = ((int64_t)lo + kTrLong - 1) >> kTrLongShift; (c0 << kTrLongShift) < hi; c0 += 64) {
            const int64_t e = (c0 + lane) << kTrLongShift;
            const int l = row_in_wave(start, e);
            if (e < hi) rowof[c0 + lane] = (int)v0 + l;
        }
        const int after = __shfl(start, min(lane + 1, 63), 64);
        const unsigned long long longrows = __ballot((lane < 63 ? after : hi) - start > kTrLong);
        int cur = 0;                                          // a row starting at or before the chunk
        for (int e0 = lo; e0 < hi;) {
            const int next = cur < 63 ? __builtin_amdgcn_readlane(start, cur + 1) : hi;
            if ((longrows >> cur) & 1) {                      // the count kernel writes it
                e0 = next;
                cur++;
                continue;
            }
            if (next >= e0 + 64 && hi >= e0 + 64) {           // whole chunks inside row cur
                const int stop = min(next, hi);
                for (; e0 + 64 <= stop; e0 += 64) src[e0 + lane] = (int)v0 + cur;
                continue;
            }
            const int e = e0 + lane;
            const int l = row_in_wave(start, e);
            cur = __builtin_amdgcn_readlane(l, 63);
            if (e < hi && !((longrows >> l) & 1)) src[e] = (int)v0 + l;
            e0 += 64;
        }
    }
}

// Per tile: the histogram of the edges' destination buckets, and the sources of the tile's
// edges in rows longer than kTrLong (see k_tr_src).
__global__ __launch_bounds__(256) void k_tr_count(const int2 *__restrict__ edges, int64_t m, int bshift, int nb,
                                                  int tiles, int *__restrict__ counts, const int *__restrict__ ptr,
                                                  const int *__restrict__ rowof, int *__restrict__ src) {
    __shared__ int hist[kTrMaxBuckets];
    __shared__ int lrow[kTrTileChunks + 1], lbeg[kTrTileChunks + 1], lend[kTrTileChunks + 1];
    const int64_t e0 = (int64_t)blockIdx.x * kTrTile, e1 = min(m, e0 + kTrTile);
    const int64_t nc = (m + kTrLong - 1) >> kTrLongShift;
    if (threadIdx.x <= kTrTileChunks) {
        const int64_t c = (int64_t)blockIdx.x * kTrTileChunks + threadIdx.x;
        int r = -1, b = 0, en = 0;
        if (c < nc) {
            r = rowof[c];
            b = ptr[r];
            en = ptr[r + 1];
            if (en - b <= kTrLong) r = -1;
        }
        lrow[threadIdx.x] = r;
        lbeg[threadIdx.x] = b;
        lend[threadIdx.x] = en;
    }
    for (int i = threadIdx.x; i < nb; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int k = 0; k <= kTrTileChunks; k++) {
        const int r = lrow[k];
        if (r < 0 || (k > 0 && lrow[k - 1] == r)) continue;     // chunk starts of one row are adjacent
        const int64_t a = max((int64_t)lbeg[k], e0), b = min((int64_t)lend[k], e1);
        for (int64_t e = a + threadIdx.x; e < b; e += blockDim.x) src[e] = r;
    }
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

constexpr int kTrU = 8;   // chunks of 64 whose loads are in flight together

// One wave per tile: the tile's run of each bucket is this wave's alone, so it is written
// contiguously and in edge order (a lane's rank among the lanes of the same bucket comes
// from ballots over the bucket bits).  pos[] starts at the run starts the scan gave.
__global__ __launch_bounds__(64) void k_tr_scatter(const int *__restrict__ src, const int2 *__restrict__ edges,
                                                   int64_t m, int bshift, int nb, const int *__restrict__ offs,
                                                   int2 *__restrict__ items) {
    extern __shared__ int pos[];                   // [nb]
    const int lane = threadIdx.x;
    const int bits = 32 - __clz(nb - 1 > 0 ? nb - 1 : 1);
    const int64_t e0 = (int64_t)blockIdx.x * kTrTile, e1 = min(m, e0 + kTrTile);
    for (int b = lane; b < nb; b += 64) pos[b] = offs[(size_t)blockIdx.x * nb + b];
    __syncthreads();
    const int dmask = (1 << bshift) - 1;
    for (int64_t c0 = e0; c0 < e1; c0 += 64 * kTrU) {
        int2 ew[kTrU];
        int sv[kTrU];
#pragma unroll
        for (int q = 0; q < kTrU; q++) {
            const int64_t e = c0 + q * 64 + lane;
            ew[q] = e < e1 ? edges[e] : make_int2(0, 0);
            sv[q] = e < e1 ? src[e] : 0;
        }
#pragma unroll
        for (int q = 0; q < kTrU; q++) {
            const bool valid = c0 + q * 64 + lane < e1;
            const int bk = ew[q].x >> bshift;
            const unsigned long long eq = match_lanes(valid ? bk : -1, bits + 1);
            const int base = valid ? pos[bk] : 0;
            if (valid) items[base + __popcll(eq & lanes_below())] = make_int2(sv[q], ew[q].x & dmask);
            if (valid && (eq & lanes_below()) == 0) pos[bk] = base + __popcll(eq);
        }
    }
}

// One workgroup per bucket, for buckets the LDS kernel below does not take (more than
// `skip_below` items): cptr of its destinations and their predecessor lists, placed in HBM.
// The bucket's items are split into one contiguous run per wave; each wave counts its run
// per destination, the counts are scanned (destinations in order, within a destination the
// waves in order), and each wave places its run in order.
__global__ __launch_bounds__(256) void k_tr_bucket(const int2 *__restrict__ items, int n, int64_t m, int bshift,
                                                   int tiles, const int *__restrict__ offs, int skip_below,
                                                   int *__restrict__ cptr, int *__restrict__ cpred) {
    extern __shared__ int cd[];                   // [kTrWaves][width]
    __shared__ int lds[4];
    const int width = 1 << bshift;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x;
    const int64_t i0 = offs[(size_t)b * tiles];
    const int64_t i1 = b + 1 < (int)gridDim.x ? offs[(size_t)(b + 1) * tiles] : m;
    if (i1 - i0 <= skip_below) return;
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

// Sixteen-wave kernels with their output staged in LDS.  Each wave loads its run once into
// registers (a tile's 1,024 edges, a bucket's at most 1/16 of kTrLBucketMax items) and never
// reads it again: a second pass over HBM or L2 was what these kernels spent their time on
// when 32 CUs' tiles did not fit one XCD's L2.  Per-wave counters are 16 bits, two to a
// word, added to with 32-bit LDS atomics on the word and read and rewritten as halves.
constexpr int kTrLWaves = 16;             // 1024 threads
constexpr int kTrLMaxBuckets = 2048;      // counters of the tile kernel: 16 x 2048 x 2 bytes
constexpr int kTrLChunks = kTrTile / (kTrLWaves * 64);   // 16 chunks of 64 edges per wave and tile
constexpr int kTrLBucketChunks = 22;      // chunks of 64 items per wave in the bucket kernel
constexpr int kTrLBucketMax = kTrLWaves * 64 * kTrLBucketChunks;   // 22,528 items

__device__ __forceinline__ void lds_count16(unsigned *cw, int h) {
    atomicAdd(&cw[h >> 1], 1u << ((h & 1) << 4));            // half h: bits 16(h&1).. of word h/2
}

// Largest b in [0, nb) with start[b] <= i (start nondecreasing, start[0] = 0).
__device__ __forceinline__ int bucket_of_slot(const int *start, int nb, int i) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// One workgroup per tile of kTrTile edges: the tile's items sorted by bucket in LDS (stable:
// waves take consecutive runs, lanes are ranked by ballots), then written out so that
// consecutive lanes write consecutive slots of one bucket's run: the tile's run of each
// bucket leaves as whole lines instead of one partial line per item.  An item waits in LDS
// as 32 bits: (source - the tile's first source) << bshift | destination within the bucket,
// or the edge's offset in the tile in place of the source difference when the tile's
// sources span 2^(32 - bshift) or more (only possible with vertices that have no edges).
__global__ __launch_bounds__(1024) void k_tr_scatter_lds(const int *__restrict__ src, const int2 *__restrict__ edges,
                                                         int64_t m, int bshift, int nb, const int *__restrict__ offs,
                                                         int2 *__restrict__ items) {
    extern __shared__ int ldyn[];
    unsigned *cw = (unsigned *)ldyn;                                    // [16][nb] halves
    unsigned short *c16 = (unsigned short *)ldyn;
    int *lstart = ldyn + kTrLWaves * nb / 2;                            // [nb]: local start of each bucket
    int *delta = lstart + nb;                                           // [nb]: global run start - local start
    unsigned *slot = (unsigned *)(delta + nb);                          // [kTrTile]
    __shared__ int lds[kTrLWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int bits = 32 - __clz(nb - 1 > 0 ? nb - 1 : 1);
    const int dmask = (1 << bshift) - 1;
    const int64_t e0 = (int64_t)blockIdx.x * kTrTile, e1 = min(m, e0 + kTrTile);
    const int cnt = (int)(e1 - e0);
    const int src0 = src[e0];
    const bool by_source = (int64_t)src[e1 - 1] - src0 < (1ll << (32 - bshift));
    const int wbase = wave * kTrLChunks * 64;
    int d[kTrLChunks], sv[kTrLChunks];
#pragma unroll
    for (int q = 0; q < kTrLChunks; q++) {
        const int i = wbase + q * 64 + lane;
        d[q] = i < cnt ? edges[e0 + i].x : -1;
        sv[q] = i < cnt ? src[e0 + i] : 0;
    }
    for (int i = threadIdx.x; i < kTrLWaves * nb / 2; i += blockDim.x) cw[i] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kTrLChunks; q++)
        if (d[q] >= 0) lds_count16(cw, wave * nb + (d[q] >> bshift));
    __syncthreads();
    // local starts: buckets in order, within a bucket the waves in order
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += blockDim.x) {
        const int b = b0 + threadIdx.x;
        int c[kTrLWaves], t = 0;
#pragma unroll
        for (int w = 0; w < kTrLWaves; w++) {
            c[w] = b < nb ? c16[w * nb + b] : 0;
            t += c[w];
        }
        int total;
        const int ex = egs_block_exclusive_scan(t, lds, total);
        int base = carry + ex;
        if (b < nb) {
            lstart[b] = base;
            delta[b] = offs[(size_t)blockIdx.x * nb + b] - base;
#pragma unroll
            for (int w = 0; w < kTrLWaves; w++) {
                c16[w * nb + b] = (unsigned short)base;
                base += c[w];
            }
        }
        carry += total;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kTrLChunks; q++) {
        const bool valid = d[q] >= 0;
        const int bk = d[q] >> bshift;
        const unsigned long long eq = match_lanes(valid ? bk : -1, bits + 1);
        const int base = valid ? c16[wave * nb + bk] : 0;
        const unsigned key = (unsigned)(by_source ? sv[q] - src0 : wbase + q * 64 + lane);
        if (valid) slot[base + __popcll(eq & lanes_below())] = (key << bshift) | (unsigned)(d[q] & dmask);
        if (valid && (eq & lanes_below()) == 0) c16[wave * nb + bk] = (unsigned short)(base + __popcll(eq));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const unsigned x = slot[i];
        const int key = (int)(x >> bshift);
        const int b = bucket_of_slot(lstart, nb, i);
        items[delta[b] + i] = make_int2(by_source ? src0 + key : src[e0 + key], (int)(x & (unsigned)dmask));
    }
}

// One workgroup per bucket of at most `cap` (<= kTrLBucketMax) items: its destinations'
// counts, cptr, and the predecessor lists placed in order in LDS and written out coalesced
// (a random 4-byte store per item reached HBM as a partial line, about 8x the bytes).
__global__ __launch_bounds__(1024) void k_tr_bucket_lds(const int2 *__restrict__ items, int n, int64_t m, int bshift,
                                                        int tiles, const int *__restrict__ offs, int cap,
                                                        int *__restrict__ cptr, int *__restrict__ cpred) {
    extern __shared__ int ldyn[];
    const int width = 1 << bshift;
    unsigned *cw = (unsigned *)ldyn;                                    // [16][width] halves
    unsigned short *c16 = (unsigned short *)ldyn;
    int *out = ldyn + kTrLWaves * width / 2;                            // [cap]
    __shared__ int lds[kTrLWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x;
    const int64_t i0 = offs[(size_t)b * tiles];
    const int64_t i1 = b + 1 < (int)gridDim.x ? offs[(size_t)(b + 1) * tiles] : m;
    if (i1 - i0 > cap) return;                     // k_tr_bucket places it
    const int cnt = (int)(i1 - i0);
    const int per = (cnt + kTrLWaves - 1) / kTrLWaves;
    const int w0 = min(cnt, wave * per), w1 = min(cnt, w0 + per);
    int2 it[kTrLBucketChunks];
#pragma unroll
    for (int q = 0; q < kTrLBucketChunks; q++) {
        const int i = w0 + q * 64 + lane;
        it[q] = i < w1 ? items[i0 + i] : make_int2(0, -1);
    }
    for (int i = threadIdx.x; i < kTrLWaves * width / 2; i += blockDim.x) cw[i] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kTrLBucketChunks; q++)
        if (it[q].y >= 0) lds_count16(cw, wave * width + it[q].y);
    __syncthreads();
    int carry = 0;
    for (int d0 = 0; d0 < width; d0 += blockDim.x) {
        const int d = d0 + threadIdx.x;
        int c[kTrLWaves], deg = 0;
#pragma unroll
        for (int w = 0; w < kTrLWaves; w++) {
            c[w] = c16[w * width + d];
            deg += c[w];
        }
        int total;
        const int ex = egs_block_exclusive_scan(deg, lds, total);
        int base = carry + ex;
        const int64_t u = ((int64_t)b << bshift) + d;
        if (u < n) cptr[u] = (int)i0 + base;
#pragma unroll
        for (int w = 0; w < kTrLWaves; w++) {
            c16[w * width + d] = (unsigned short)base;
            base += c[w];
        }
        carry += total;
    }
    __syncthreads();
    if (b == gridDim.x - 1 && threadIdx.x == 0) cptr[n] = (int)m;
#pragma unroll
    for (int q = 0; q < kTrLBucketChunks; q++) {
        const bool valid = it[q].y >= 0;
        const int d = it[q].y;
        const unsigned long long eq = match_lanes(d, bshift + 1);
        const int base = valid ? c16[wave * width + d] : 0;
        if (valid) out[base + __popcll(eq & lanes_below())] = it[q].x;
        if (valid && (eq & lanes_below()) == 0) c16[wave * width + d] = (unsigned short)(base + __popcll(eq));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) cpred[i0 + i] = out[i];
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
    // the chunk starts' rows sit in the items' space until the scatter writes it (ceil(m /
    // kTrLong) ints of its 2m)
    int *rowof = (int *)items;
    hipLaunchKernelGGL(k_tr_src, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, g.ptr, n, src, rowof);
    // counts[tile][bucket] -> offs as [bucket][tile] -> scanned into counts ([bucket][tile])
    // -> offs as [tile][bucket], the order the scatter reads
    hipLaunchKernelGGL(k_tr_count, dim3(tiles), dim3(256), 0, st, g.edges, m, bshift, nb, tiles, counts, g.ptr,
                       (const int *)rowof, src);
    hipLaunchKernelGGL(k_tr_matrix, dim3((nb + 63) / 64, (tiles + 63) / 64), dim3(256), 0, st, (const int *)counts,
                       tiles, nb, offs);
    hipError_t e = launch_exclusive_scan(offs, nb * tiles, scan_tiles, counts, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tr_matrix, dim3((tiles + 63) / 64, (nb + 63) / 64), dim3(256), 0, st, (const int *)counts,
                       nb, tiles, offs);
    if (nb <= kTrLMaxBuckets) {
        const size_t lds = sizeof(int) * ((size_t)kTrLWaves * nb / 2 + 2 * (size_t)nb + kTrTile);
        hipLaunchKernelGGL(k_tr_scatter_lds, dim3(tiles), dim3(1024), lds, st, (const int *)src, g.edges, m, bshift, nb,
                           (const int *)offs, items);
    } else {
        hipLaunchKernelGGL(k_tr_scatter, dim3(tiles), dim3(64), sizeof(int) * (size_t)nb, st, (const int *)src, g.edges,
                           m, bshift, nb, (const int *)offs, items);
    }
    // buckets of at most `cap` items are placed in LDS (16 x width 16-bit counters and
    // out[cap]); the rest, or every bucket when the counters do not fit, in HBM
    const int64_t counters = (int64_t)kTrLWaves * 2 << bshift;          // bytes
    const int cap = (int)std::min<int64_t>(kTrLBucketMax, (kTrBucketLds - counters) / (int64_t)sizeof(int));
    if (cap > 0) {
        const size_t lds = (size_t)counters + sizeof(int) * (size_t)cap;
        hipLaunchKernelGGL(k_tr_bucket_lds, dim3(nb), dim3(1024), lds, st, (const int2 *)items, n, m, bshift, tiles,
                           (const int *)counts, cap, const_cast<int *>(g.cptr), const_cast<int *>(g.cpred));
    }
    hipLaunchKernelGGL(k_tr_bucket, dim3(nb), dim3(256), sizeof(int) * (size_t)kTrWaves << bshift, st,
                       (const int2 *)items, n, m, bshift, tiles, (const int *)counts, cap > 0 ? cap : -1,
                       const_cast<int *>(g.cptr), const_cast<int *>(g.cpred));
    return hipGetLastError();
}

}  // namespace egs
