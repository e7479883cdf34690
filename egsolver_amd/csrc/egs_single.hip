// egs_single.hip -- one energy game resident on one MI355X: CSR->CSC transpose, initial
// frontier, and the worklist fixpoint rounds.
//
// Replaces the reference GPU path: EG_gpu_traspose_graph + cuSPARSE csr2csc
// (src/dev_EG_alg.cu:277-290, src/csr2csc.cu:54-66), kernel_EG_initialize
// (src/dev_EG_alg.cu:176-205), the lift kernels in 1/K lanes per vertex
// (src/dev_EG_alg.cu:73-162, 397-1480) and the ping-pong driver with thrust compaction
// (src/dev_EG_alg.cu:307-381, src/thrust_wrapper.cu:30-41).
//
// Design (device memory, one game):
//   ptr[n+1], edges[m] = {succ, w} packed as int2, cptr[n+1], cpred[m]   (graph)
//   val[n]                                  energy values, updated in place
//   stamp[n]                                last round a vertex was queued for
//   queue[2][n], big[n], huge[n]            frontier lists (ids)
//   ring[9]                                 3 rotating counters each for queue/big/huge
// Round r reads queue[r&1] (length ring[r%3]) and appends to queue[(r+1)&1]
// (length ring[(r+1)%3]); a predecessor u is appended once per round, by the lane whose
// atomicMax raises stamp[u] to r+1.  Values only grow and every read is of a value
// no larger than the least fixpoint, so in-place reads of neighbours lifted in the same
// round are harmless; any vertex whose successor lifted is queued again for the next
// round, so the run ends exactly when lift(val) == val.
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

__device__ __forceinline__ int block_exclusive_scan(int x, int *lds_warp, int &total) {
    // wave-level inclusive scan with shuffles, then wave totals combined through LDS.
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    int inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) lds_warp[wid] = inc;
    __syncthreads();
    int wave_off = 0;
    total = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
        int s = lds_warp[i];
        if (i < wid) wave_off += s;
        total += s;
    }
    __syncthreads();
    return wave_off + inc - x;
}

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
    block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(int *__restrict__ tile_sum, int tiles) {
    __shared__ int lds[SCAN_THREADS / 64];
    int carry = 0;
    for (int b = 0; b < tiles; b += SCAN_THREADS) {
        int i = b + threadIdx.x;
        int x = (i < tiles) ? tile_sum[i] : 0;
        int total;
        int ex = block_exclusive_scan(x, lds, total);
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
    int off = block_exclusive_scan(run, lds, total) + tile_off[blockIdx.x];
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

// Initial frontier (src/dev_EG_alg.cu:176-205, src/cpu_solver.c:217-249): an owner-0
// vertex whose out-weights are all negative, or an owner-1 vertex with a negative
// out-weight, has lift(0) > 0.  Values start at 0.
__global__ __launch_bounds__(256) void k_init(DevGame g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i - threadIdx.x < g.n;
         i += stride) {
        const int v = (int)i;
        bool act = false;
        if (v < g.n) {
            const int s = g.ptr[v], e = g.ptr[v + 1];
            bool all_neg = true, any_neg = false;
            for (int k = s; k < e; k++) {
                const bool neg = g.edges[k].y < 0;
                all_neg &= neg;
                any_neg |= neg;
            }
            act = (v < g.n0) ? all_neg : any_neg;
            g.val[v] = 0;
            g.stamp[v] = act ? 1 : 0;
        }
        egs_append(g.queue[1], &g.ring[1], act, v);
    }
}

// ---------------------------------------------------------------------------------------
// lift kernels
// ---------------------------------------------------------------------------------------

struct LiftAcc {
    uint32_t processed = 0, lifted = 0;
    uint64_t edges = 0, preds = 0;
};

__device__ __forceinline__ void flush_stats(const DevGame &g, LiftAcc &a) {
    // per-block slot; each kernel of a round adds in stream order, so no atomics.
    __shared__ unsigned long long red[4][4];
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    unsigned long long vals[4] = {a.processed, a.edges, a.lifted, a.preds};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        unsigned long long x = vals[k];
        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
        if (lane == 0) red[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[threadIdx.x][i];
        if (s) g.stats[(size_t)blockIdx.x * 4 + threadIdx.x] += s;
    }
}

// Queue predecessors of a lifted vertex v for round r+1.  The lanes of a group of G
// stride over v's predecessor list.
template <int G>
__device__ __forceinline__ void push_preds(const DevGame &g, int v, bool lifted, int gl, int r,
                                           int *qout, int *cout, LiftAcc &acc) {
    int cs = 0, ce = 0;
    if (lifted) { cs = g.cptr[v]; ce = g.cptr[v + 1]; }
    if (lifted && gl == 0) acc.preds += (uint32_t)(ce - cs);
    const int tag = r + 1;
    // the loop trip count differs between groups; egs_append is divergence-safe.
    for (int k = cs + gl; k < ce; k += G) {
        const int u = g.cpred[k];
        const bool first = atomicMax(&g.stamp[u], tag) < tag;
        egs_append(qout, cout, first, u);
    }
}

// Groups of G lanes (G in 1..64) take one queued vertex each: strided edge loop, DPP
// min/max reduction inside the group, one store, shared predecessor loop.  Vertices
// with out-degree above big_deg are handed to the wavefront kernel (G=64 lanes per
// vertex) and above huge_deg to the workgroup kernel, within the same round.
template <int G>
__global__ __launch_bounds__(256) void k_lift_groups(DevGame g, int r, int big_deg, int huge_deg) {
    constexpr int PER_WAVE = 64 / G;
    const int *qin = g.queue[r & 1];
    int *qout = g.queue[(r + 1) & 1];
    int *cout = &g.ring[(r + 1) % 3];
    const int cnt = g.ring[r % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g.ring[(r + 2) % 3] = 0;          // output counter of round r+1
        g.ring[3 + (r + 1) % 3] = 0;      // big-list counter of round r+1
        g.ring[6 + (r + 1) % 3] = 0;      // huge-list counter of round r+1
        if (cnt > 0) *g.rounds_done = r;
    }
    LiftAcc acc;
    const int lane = threadIdx.x & 63;
    const int gl = lane % G;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i0 = wave * PER_WAVE; i0 < cnt; i0 += waves * PER_WAVE) {
        const int64_t item = i0 + lane / G;
        int v = 0, s = 0, e = 0, old = 0;
        bool valid = item < cnt;
        if (valid) {
            v = qin[item];
            s = g.ptr[v];
            e = g.ptr[v + 1];
            if (e - s > big_deg) {
                if (gl == 0) {
                    if (e - s > huge_deg) {
                        int pos = atomicAdd(&g.ring[6 + r % 3], 1);
                        g.huge[pos] = v;
                    } else {
                        int pos = atomicAdd(&g.ring[3 + r % 3], 1);
                        g.big[pos] = v;
                    }
                }
                valid = false;
                s = e = 0;
            } else {
                old = g.val[v];
            }
        }
        const bool owner1 = v >= g.n0;
        int key = EGS_TOP;
        for (int k = s + gl; k < e; k += G) {
            const int2 ew = g.edges[k];
            key = min(key, egs_key(egs_ominus(g.val[ew.x], ew.y, g.mg), owner1));
        }
        key = egs_group_min<G>(key);
        const int t = egs_unkey(key, owner1);
        const bool lifted = valid && old < t;
        if (valid && gl == 0) {
            acc.processed++;
            acc.edges += (uint32_t)(e - s);
            if (lifted) { g.val[v] = t; acc.lifted++; }
        }
        push_preds<G>(g, v, lifted, gl, r, qout, cout, acc);
    }
    flush_stats(g, acc);
}

// One wavefront per vertex from the big list.
__global__ __launch_bounds__(256) void k_lift_wave(DevGame g, int r) {
    const int cnt = g.ring[3 + r % 3];
    if (cnt == 0) return;
    int *qout = g.queue[(r + 1) & 1];
    int *cout = &g.ring[(r + 1) % 3];
    LiftAcc acc;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = wave; i < cnt; i += waves) {
        const int v = g.big[i];
        const int s = g.ptr[v], e = g.ptr[v + 1];
        const int old = g.val[v];
        const bool owner1 = v >= g.n0;
        int key = EGS_TOP;
        for (int k = s + lane; k < e; k += 64) {
            const int2 ew = g.edges[k];
            key = min(key, egs_key(egs_ominus(g.val[ew.x], ew.y, g.mg), owner1));
        }
        key = egs_group_min<64>(key);
        const int t = egs_unkey(key, owner1);
        const bool lifted = old < t;
        if (lane == 0) {
            acc.processed++;
            acc.edges += (uint32_t)(e - s);
            if (lifted) { g.val[v] = t; acc.lifted++; }
        }
        push_preds<64>(g, v, lifted, lane, r, qout, cout, acc);
    }
    flush_stats(g, acc);
}

// One workgroup (256 lanes) per vertex from the huge list: LDS combine of 4 wave minima.
__global__ __launch_bounds__(256) void k_lift_block(DevGame g, int r) {
    const int cnt = g.ring[6 + r % 3];
    if (cnt == 0) return;
    __shared__ int wmin[4];
    int *qout = g.queue[(r + 1) & 1];
    int *cout = &g.ring[(r + 1) % 3];
    LiftAcc acc;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    for (int64_t i = blockIdx.x; i < cnt; i += gridDim.x) {
        const int v = g.huge[i];
        const int s = g.ptr[v], e = g.ptr[v + 1];
        const int old = g.val[v];
        const bool owner1 = v >= g.n0;
        int key = EGS_TOP;
        for (int k = s + (int)threadIdx.x; k < e; k += 256) {
            const int2 ew = g.edges[k];
            key = min(key, egs_key(egs_ominus(g.val[ew.x], ew.y, g.mg), owner1));
        }
        key = egs_group_min<64>(key);
        if (lane == 0) wmin[wid] = key;
        __syncthreads();
        key = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
        __syncthreads();
        const int t = egs_unkey(key, owner1);
        const bool lifted = old < t;
        if (threadIdx.x == 0) {
            acc.processed++;
            acc.edges += (uint32_t)(e - s);
            if (lifted) { g.val[v] = t; acc.lifted++; }
        }
        push_preds<256>(g, v, lifted, (int)threadIdx.x, r, qout, cout, acc);
    }
    flush_stats(g, acc);
}

// Fixpoint certificate: count vertices with lift(val)(v) != val(v).
__global__ __launch_bounds__(256) void k_violations(DevGame g, unsigned long long *out) {
    unsigned long long bad = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const bool owner1 = v >= g.n0;
        int key = EGS_TOP;
        for (int k = g.ptr[v]; k < g.ptr[v + 1]; k++) {
            const int2 ew = g.edges[k];
            key = min(key, egs_key(egs_ominus(g.val[ew.x], ew.y, g.mg), owner1));
        }
        bad += egs_unkey(key, owner1) != g.val[v];
    }
    for (int d = 32; d >= 1; d >>= 1) bad += __shfl_xor(bad, d, 64);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(out, bad);
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

__global__ void k_sum_stats(const unsigned long long *__restrict__ slots, int nslots,
                            unsigned long long *__restrict__ out) {
    __shared__ unsigned long long part[4][256];
    unsigned long long s[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nslots; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] += slots[(size_t)b * 4 + k];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) part[k][threadIdx.x] = s[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long t = 0;
        for (int i = 0; i < (int)blockDim.x; i++) t += part[threadIdx.x][i];
        out[threadIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------

static int grid_for(int64_t work, int threads, int cap) {
    int64_t b = (work + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
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
    // fill <- cptr (running insert positions)
    if ((err = hipMemcpyAsync(fill, g.cptr, sizeof(int) * (size_t)(n + 1), hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return err;
    hipLaunchKernelGGL(k_scatter_preds, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, g.ptr, g.edges,
                       n, fill, const_cast<int *>(g.cpred));
    return hipGetLastError();
}

hipError_t launch_pack(const int *succ, const int *w, int2 *edges, int64_t m, hipStream_t st) {
    if (m > 0)
        hipLaunchKernelGGL(k_pack_edges, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, succ, w, edges, m);
    return hipGetLastError();
}

hipError_t launch_init(const DevGame &g, hipStream_t st) {
    hipError_t err = hipMemsetAsync(g.ring, 0, sizeof(int) * 9, st);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_init, dim3(grid_for(g.n, 256, 4096)), dim3(256), 0, st, g);
    return hipGetLastError();
}

hipError_t launch_round(const DevGame &g, const LaunchShape &sh, int r, hipStream_t st) {
    switch (sh.group) {
    case 1: hipLaunchKernelGGL(k_lift_groups<1>, dim3(sh.grid), dim3(256), 0, st, g, r, sh.big_deg, sh.huge_deg); break;
    case 2: hipLaunchKernelGGL(k_lift_groups<2>, dim3(sh.grid), dim3(256), 0, st, g, r, sh.big_deg, sh.huge_deg); break;
    case 4: hipLaunchKernelGGL(k_lift_groups<4>, dim3(sh.grid), dim3(256), 0, st, g, r, sh.big_deg, sh.huge_deg); break;
    case 8: hipLaunchKernelGGL(k_lift_groups<8>, dim3(sh.grid), dim3(256), 0, st, g, r, sh.big_deg, sh.huge_deg); break;
    case 16: hipLaunchKernelGGL(k_lift_groups<16>, dim3(sh.grid), dim3(256), 0, st, g, r, sh.big_deg, sh.huge_deg); break;
    case 32: hipLaunchKernelGGL(k_lift_groups<32>, dim3(sh.grid), dim3(256), 0, st, g, r, sh.big_deg, sh.huge_deg); break;
    default: hipLaunchKernelGGL(k_lift_groups<64>, dim3(sh.grid), dim3(256), 0, st, g, r, sh.big_deg, sh.huge_deg); break;
    }
    if (sh.use_wave) hipLaunchKernelGGL(k_lift_wave, dim3(sh.wave_grid), dim3(256), 0, st, g, r);
    if (sh.use_block) hipLaunchKernelGGL(k_lift_block, dim3(sh.block_grid), dim3(256), 0, st, g, r);
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
    hipLaunchKernelGGL(k_sum_stats, dim3(1), dim3(256), 0, st, g.stats, nslots, out);
    return hipGetLastError();
}

}  // namespace egs
