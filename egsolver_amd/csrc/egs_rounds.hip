// egs_rounds.hip -- initial frontier and the fixpoint rounds of one resident game.
//
// Replaces kernel_EG_initialize (src/dev_EG_alg.cu:176-205), the lift kernels with 1..32
// lanes per vertex (src/dev_EG_alg.cu:73-162, 397-1480) and the ping-pong round driver
// with its thrust compaction and memsets (src/dev_EG_alg.cu:307-381,
// src/thrust_wrapper.cu:30-41).
//
// Every round r runs one of three bodies, chosen on the device from round r-1's output:
//   DENSE   sweep all vertices in id order (edges stream, values update in place) and
//           list the vertices that rose.  Used while most vertices are active.
//   QUEUE   process the worklist built by round r-1; every vertex that rises queues its
//           predecessors for round r+1 (deduplicated by stamp[u] = r+1).
//   LIFTED  after a dense round: claim each predecessor of the vertices listed as risen
//           (stamp[p] = r) and process it; risers queue their predecessors as in QUEUE.
// All bodies read neighbour values in place.  Values only rise and never pass the least
// fixpoint, and any vertex with a successor that rose during round r is in the work of
// round r+1 (every vertex for DENSE; its predecessors for QUEUE/LIFTED), so an empty
// output ends the run exactly at lift(val) == val.
//
// Vertices of out-degree above big_deg are taken out of the main kernel: a wavefront per
// vertex (k_lift_wave) and, above huge_deg, a workgroup per vertex (k_lift_block), in the
// same round.  Appends are staged in LDS per workgroup and flushed with one atomic per
// sub-queue counter.
#include "egs_device.h"
#include "egs_internal.h"

namespace egs {

// ---------------------------------------------------------------------------------------
// list helpers
// ---------------------------------------------------------------------------------------

// Reserve k slots of an output list for sub-queue `sub`: the first n_sub go to the
// sub-queue at base_sub, the rest to the overflow region at base_ovf.
__device__ __forceinline__ void reserve_split(const DevGame &g, int *cnts, int sub, int k,
                                              int &base_sub, int &n_sub, int &base_ovf) {
    const int p = atomicAdd(&cnts[sub], k);
    const int room = max(0, g.sub_cap - p);
    n_sub = min(k, room);
    base_sub = p;
    base_ovf = (k > n_sub) ? atomicAdd(&cnts[kSub], k - n_sub) : 0;
}

__device__ __forceinline__ int list_slot(const DevGame &g, int sub, int base_sub, int n_sub,
                                         int base_ovf, int i) {
    return i < n_sub ? sub * g.sub_cap + base_sub + i : kSub * g.sub_cap + base_ovf + (i - n_sub);
}

// Wave-aggregated append straight to the output list (LDS spill and rare appends).
__device__ __forceinline__ void push_wave(const DevGame &g, int *list, int *cnts, int sub, bool want,
                                          int item) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return;
    const int leader = __ffsll((unsigned long long)__builtin_amdgcn_read_exec()) - 1;
    int bs = 0, ns = 0, bo = 0;
    if ((int)egs_lane() == leader) reserve_split(g, cnts, sub, __popcll(mask), bs, ns, bo);
    bs = __builtin_amdgcn_readfirstlane(bs);
    ns = __builtin_amdgcn_readfirstlane(ns);
    bo = __builtin_amdgcn_readfirstlane(bo);
    if (want) list[list_slot(g, sub, bs, ns, bo, __popcll(mask & egs_lanemask_lt()))] = item;
}

// Per-workgroup LDS staging of appends.
struct Stage {
    int *buf;   // kLdsCap ints in LDS
    int *cnt;   // LDS counter (may pass kLdsCap; the excess went to global directly)
    int *list;  // output list
    int *cnts;  // its counters
    int sub;
};

__device__ __forceinline__ void stage_push(const DevGame &g, const Stage &s, bool want, int item) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return;
    const int leader = __ffsll((unsigned long long)__builtin_amdgcn_read_exec()) - 1;
    int base = 0;
    if ((int)egs_lane() == leader) base = atomicAdd(s.cnt, __popcll(mask));
    base = __builtin_amdgcn_readfirstlane(base);
    const int pos = base + __popcll(mask & egs_lanemask_lt());
    const bool spill = want && pos >= kLdsCap;
    if (want && !spill) s.buf[pos] = item;
    push_wave(g, s.list, s.cnts, s.sub, spill, item);
}

// Block-uniform: move the staged items to the output list when at least `threshold`.
__device__ __forceinline__ void stage_flush(const DevGame &g, const Stage &s, int threshold) {
    __shared__ int fb[3];
    __syncthreads();
    const int c = min(*s.cnt, kLdsCap);
    __syncthreads();    // every thread has read the count before anyone changes it
    if (c == 0 || c < threshold) return;
    if (threadIdx.x == 0) reserve_split(g, s.cnts, s.sub, c, fb[0], fb[1], fb[2]);
    __syncthreads();
    const int bs = fb[0], ns = fb[1], bo = fb[2];
    for (int i = threadIdx.x; i < c; i += blockDim.x) s.list[list_slot(g, s.sub, bs, ns, bo, i)] = s.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) *s.cnt = 0;
    __syncthreads();
}

// Input list layout: pre[k] = items before sub-queue k (k = 0..kSub, the last being the
// overflow region), pre[kSub+1] = total.  Wave 0 builds it; callers sync.
__device__ __forceinline__ void load_layout(const DevGame &g, int slot, int *pre) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const int *c = g.lcnt + slot * kListCounters;
        int x = 0;
        if (lane < kSub) x = min(c[lane], g.sub_cap);
        int ovf = c[kSub];
        const int inc = egs_wave_incl_scan(x);
        pre[lane] = inc - x;
        if (lane == 63) {
            pre[kSub] = inc;
            pre[kSub + 1] = inc + ovf;
        }
    }
}

// Position in the list of the i-th item of the layout.
__device__ __forceinline__ int layout_slot(const DevGame &g, const int *pre, int i) {
    int lo = 0, hi = kSub;        // largest k in [0, kSub] with pre[k] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo * g.sub_cap + (i - pre[lo]);
}

// Round mode from round r-1's mode and output size (identical in every workgroup).
__device__ __forceinline__ int decide_mode(const DevGame &g, int r, int total) {
    if (g.ring[11] == r) return MODE_DENSE;     // values were raised outside the rounds
    const int prev = g.mode[(r + 2) % 3];
    if (prev == MODE_DONE || total == 0) return MODE_DONE;
    int mode;
    if (prev == MODE_DENSE) {
        const int64_t est = (int64_t)total * g.m / g.n;   // vertices the risers' predecessors cover
        mode = est >= g.dense_min ? MODE_DENSE : MODE_LIFTED;
    } else {
        mode = total >= g.dense_min ? MODE_DENSE : MODE_QUEUE;
    }
    return mode;
}

// The sparse modes walk predecessor lists: until the transpose exists (ring[12], raised by
// the host) such a round sweeps densely instead and counts itself in ring[13].  They also
// read the input list's items: a round whose input holds only the count (mode[3] == r) sweeps
// densely too.
__device__ __forceinline__ int gate_mode(const DevGame &g, int mode, int r) {
    return (mode == MODE_QUEUE || mode == MODE_LIFTED) && (g.ring[12] == 0 || g.mode[3] == r) ? MODE_DENSE : mode;
}

// ---------------------------------------------------------------------------------------
// statistics
// ---------------------------------------------------------------------------------------

struct LiftAcc {
    uint32_t processed = 0, lifted = 0;
    uint64_t edges = 0, preds = 0;
};

__device__ __forceinline__ void flush_stats(const DevGame &g, LiftAcc &a) {
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
        // a slot per workgroup; grids past kStatSlots workgroups share slots, so they add atomically
        unsigned long long *slot = g.stats + (size_t)(blockIdx.x % kStatSlots) * 4 + threadIdx.x;
        if (s) {
            if (gridDim.x <= kStatSlots) *slot += s;
            else atomicAdd(slot, s);
        }
    }
}

// ---------------------------------------------------------------------------------------
// pieces of a lift
// ---------------------------------------------------------------------------------------

// min/max over v's out-edges [s, e) of OMINUS(val[succ], w), lanes of a group of G
// striding; returns the group result in every lane of the group.
// Min-key over the edges [s, e) taken by lane l of `stride` lanes, U edges in flight per
// lane: the vertices given to the wavefront and workgroup kernels have hundreds to tens of
// thousands of edges, and one gather at a time leaves that loop latency-bound.  Slots past
// e load vertex 0's value and are not used.
template <int U>
__device__ __forceinline__ int key_over_edges(const DevGame &g, int s, int e, int l, int stride, bool owner1) {
    int key = EGS_TOP;
    for (int k0 = s + l; k0 < e; k0 += stride * U) {
        int2 ew[U];
        int a[U];
#pragma unroll
        for (int j = 0; j < U; j++) ew[j] = k0 + j * stride < e ? g.edges[k0 + j * stride] : make_int2(0, 0);
#pragma unroll
        for (int j = 0; j < U; j++) a[j] = g.val[ew[j].x];
#pragma unroll
        for (int j = 0; j < U; j++)
            if (k0 + j * stride < e) key = min(key, egs_key(egs_ominus(a[j], ew[j].y, g.mg), owner1));
    }
    return key;
}

template <int G>
__device__ __forceinline__ int lift_value(const DevGame &g, int s, int e, int gl, bool owner1) {
    int key = EGS_TOP;
    if constexpr (G == 64) {
        key = key_over_edges<8>(g, s, e, gl, 64, owner1);
    } else {
        for (int k = s + gl; k < e; k += G) {
            const int2 ew = g.edges[k];
            key = min(key, egs_key(egs_ominus(g.val[ew.x], ew.y, g.mg), owner1));
        }
    }
    key = egs_group_min<G>(key);
    return egs_unkey(key, owner1);
}

// Queue the predecessors of a risen vertex v for round r+1 (stamp dedupe), lanes of a
// group of G striding over v's CSC list.
template <int G>
__device__ __forceinline__ void push_preds(const DevGame &g, const Stage &st, int v, bool lifted, int gl,
                                           int r, LiftAcc &acc) {
    int cs = 0, ce = 0;
    if (lifted) { cs = g.cptr[v]; ce = g.cptr[v + 1]; }
    if (lifted && gl == 0) acc.preds += (uint32_t)(ce - cs);
    const int tag = r + 1;
    for (int k = cs + gl; k < ce; k += G) {
        const int u = g.cpred[k];
        bool first = false;
        if (g.stamp[u] < tag) first = atomicMax(&g.stamp[u], tag) < tag;
        stage_push(g, st, first, u);
    }
}

// Hand a vertex to the wavefront / workgroup kernel of this round (sparse modes).
__device__ __forceinline__ void defer(const DevGame &g, int r, int v, int deg) {
    if (deg > g.huge_deg) {
        const int pos = atomicAdd(&g.ring[6 + r % 3], 1);
        g.dhuge[pos] = v;
    } else {
        const int pos = atomicAdd(&g.ring[3 + r % 3], 1);
        g.dbig[pos] = v;
    }
}

// ---------------------------------------------------------------------------------------
// initial frontier
// ---------------------------------------------------------------------------------------

// lift(0)(v) > 0 exactly for an owner-0 vertex with all out-weights negative or an
// owner-1 vertex with a negative out-weight (src/dev_EG_alg.cu:176-205,
// src/cpu_solver.c:217-249).  Values start at 0; the frontier becomes round 1's queue.
__global__ __launch_bounds__(256) void k_init(DevGame g) {
    __shared__ int sbuf[kLdsCap];
    __shared__ int scnt;
    if (threadIdx.x == 0) scnt = 0;
    __syncthreads();
    const Stage st{sbuf, &scnt, g.list[1], g.lcnt + 1 * kListCounters, (int)(blockIdx.x % kSub)};
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x; v0 < g.n; v0 += (int64_t)gridDim.x * blockDim.x) {
        const int v = (int)(v0 + threadIdx.x);
        bool act = false;
        if (v < g.n) {
            const int s = g.ptr[v], e = g.ptr[v + 1];
            bool all_neg = true, any_neg = false;
            if (e - s <= g.big_deg && g.phases > 1) {
                act = g.stamp[v] != 0;            // decided by k_slice_build
            } else if (e - s <= g.big_deg) {      // above big_deg: k_init_big
                for (int k0 = s; k0 < e; k0 += 4) {
                    int w[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) w[j] = k0 + j < e ? g.edges[k0 + j].y : 0;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        if (k0 + j >= e) break;
                        all_neg &= w[j] < 0;
                        any_neg |= w[j] < 0;
                    }
                }
                act = (v < g.n0) ? all_neg : any_neg;
            }
            g.val[v] = 0;
            g.stamp[v] = act ? 1 : 0;
            if (g.rise) g.rise[v] = 0;
        }
        stage_push(g, st, act, v);
        stage_flush(g, st, kLdsCap / 2);
    }
    stage_flush(g, st, 1);
}

// The vertices above big_deg, one workgroup each, after k_init: the same activity test
// (owner 0: every edge negative, owner 1: some edge negative) and the same append.
__global__ __launch_bounds__(256) void k_init_big(DevGame g) {
    const int nb = g.ring[9], items = nb + g.ring[10];
    for (int64_t i = blockIdx.x; i < items; i += gridDim.x) {
        const int v = egs_big_item(g.sbig, g.shuge, nb, i);
        const int s = g.ptr[v], e = g.ptr[v + 1];
        // one edge decides: a non-negative weight for owner 0 (not all negative), a negative
        // one for owner 1; the workgroup stops at the first block of edges holding one
        const bool owner1 = v >= g.n0;
        bool found = false;
        for (int k0 = s; k0 < e; k0 += blockDim.x) {
            const int k = k0 + (int)threadIdx.x;
            bool hit = false;
            if (k < e) hit = owner1 ? g.edges[k].y < 0 : g.edges[k].y >= 0;
            if (__syncthreads_or(hit)) {
                found = true;
                break;
            }
        }
        const bool act = owner1 ? found : !found;
        if (threadIdx.x < 64) {
            if (threadIdx.x == 0 && act) g.stamp[v] = 1;
            push_wave(g, g.list[1], g.lcnt + 1 * kListCounters, (int)(blockIdx.x % kSub), act && threadIdx.x == 0, v);
        }
    }
}

// ---------------------------------------------------------------------------------------
// round kernels
// ---------------------------------------------------------------------------------------

// The bodies of k_round stride over the low-degree part of its grid, g.round_grid workgroups
// (the grid continues with hub workgroups in fused rounds).
template <int G>
__device__ __forceinline__ void body_dense(const DevGame &g, const Stage &st, int r, LiftAcc &acc) {
    constexpr int ITEMS = 256 / G;
    const int gl = threadIdx.x % G;
    const int64_t stride = (int64_t)g.round_grid * ITEMS;
    // the next vertex's row bounds and value load while this one folds (its value cannot
    // change meanwhile: only its own group writes it in a dense round)
    int s_n = 0, e_n = 0, o_n = 0;
    auto fetch = [&](int64_t t) {
        const int64_t u = t + threadIdx.x / G;
        if (u < g.n) {
            s_n = g.ptr[u];
            e_n = g.ptr[u + 1];
            o_n = g.val[u];
        }
    };
    int it = 0;
    if ((int64_t)blockIdx.x * ITEMS < g.n) fetch((int64_t)blockIdx.x * ITEMS);
    for (int64_t t0 = (int64_t)blockIdx.x * ITEMS; t0 < g.n; t0 += stride, it++) {
        const int v = (int)(t0 + threadIdx.x / G);
        bool valid = v < g.n;
        int s = s_n, e = e_n, old = o_n;
        if (t0 + stride < g.n) fetch(t0 + stride);
        if (valid) {
            if (e - s > g.big_deg) { valid = false; s = e = 0; old = 0; }   // static wave/block lists
            if (old == EGS_TOP) e = s;          // TOP is final: nothing to gather
        } else {
            s = e = old = 0;
        }
        const int t = lift_value<G>(g, s, e, gl, v >= g.n0);
        const bool lifted = valid && old < t;
        if (valid && gl == 0) {
            acc.processed++;
            acc.edges += (uint32_t)(e - s);
            if (lifted) { g.val[v] = t; acc.lifted++; if (g.rise) g.rise[v] = r; if (t == EGS_TOP) g.ring[15] = 1; }
        }
        stage_push(g, st, lifted && gl == 0, v);
        if ((it & 7) == 7) stage_flush(g, st, kLdsCap / 2);
    }
}

template <int G>
__device__ __forceinline__ void body_queue(const DevGame &g, const Stage &st, const int *in_list,
                                           const int *pre, int r, LiftAcc &acc) {
    constexpr int ITEMS = 256 / G;
    const int gl = threadIdx.x % G;
    const int total = pre[kSub + 1];
    int it = 0;
    for (int64_t t0 = (int64_t)blockIdx.x * ITEMS; t0 < total; t0 += (int64_t)g.round_grid * ITEMS, it++) {
        const int64_t item = t0 + threadIdx.x / G;
        bool valid = item < total;
        int v = 0, s = 0, e = 0, old = 0;
        if (valid) {
            v = in_list[layout_slot(g, pre, (int)item)];
            s = g.ptr[v];
            e = g.ptr[v + 1];
            if (e - s > g.big_deg) {
                if (gl == 0) defer(g, r, v, e - s);
                valid = false;
                s = e = 0;
            } else {
                old = g.val[v];
            }
        }
        const int t = lift_value<G>(g, s, e, gl, v >= g.n0);
        const bool lifted = valid && old < t;
        if (valid && gl == 0) {
            acc.processed++;
            acc.edges += (uint32_t)(e - s);
            if (lifted) { g.val[v] = t; acc.lifted++; if (g.rise) g.rise[v] = r; if (t == EGS_TOP) g.ring[15] = 1; }
        }
        push_preds<G>(g, st, v, lifted, gl, r, acc);
        if ((it & 3) == 3) stage_flush(g, st, kLdsCap / 2);
    }
}

// Items are vertices that rose in the previous (dense) round.  Each predecessor p of an
// item is claimed once (stamp[p] -> r) and lifted by the claiming lane.
template <int G>
__device__ __forceinline__ void body_lifted(const DevGame &g, const Stage &st, const int *in_list,
                                            const int *pre, int r, LiftAcc &acc) {
    constexpr int ITEMS = 256 / G;
    const int gl = threadIdx.x % G;
    const int total = pre[kSub + 1];
    int it = 0;
    for (int64_t t0 = (int64_t)blockIdx.x * ITEMS; t0 < total; t0 += (int64_t)g.round_grid * ITEMS, it++) {
        const int64_t item = t0 + threadIdx.x / G;
        int cs = 0, ce = 0;
        if (item < total) {
            const int u = in_list[layout_slot(g, pre, (int)item)];
            cs = g.cptr[u];
            ce = g.cptr[u + 1];
        }
        for (int k = cs + gl; k < ce; k += G) {
            const int p = g.cpred[k];
            bool mine = false;
            if (g.stamp[p] < r) mine = atomicMax(&g.stamp[p], r) < r;
            if (mine) {
                const int s = g.ptr[p], e = g.ptr[p + 1];
                if (e - s > g.big_deg) {
                    defer(g, r, p, e - s);
                } else {
                    const int old = g.val[p];
                    const int t = lift_value<1>(g, s, e, 0, p >= g.n0);
                    acc.processed++;
                    acc.edges += (uint32_t)(e - s);
                    if (old < t) {
                        g.val[p] = t;
                        if (t == EGS_TOP) g.ring[15] = 1;
                        acc.lifted++;
                        if (g.rise) g.rise[p] = r;
                        const int qs = g.cptr[p], qe = g.cptr[p + 1];
                        acc.preds += (uint32_t)(qe - qs);
                        for (int q = qs; q < qe; q++) {
                            const int u2 = g.cpred[q];
                            bool first = false;
                            if (g.stamp[u2] < r + 1) first = atomicMax(&g.stamp[u2], r + 1) < r + 1;
                            stage_push(g, st, first, u2);
                        }
                    }
                }
            }
        }
        if ((it & 3) == 3) stage_flush(g, st, kLdsCap / 2);
    }
}

__device__ __forceinline__ void hub_wave(const DevGame &g, int r, int mode, int bid, int nblk, int *sbuf, int *scnt);
__device__ __forceinline__ void hub_block(const DevGame &g, int r, int mode, int bid, int nblk, int *sbuf, int *scnt,
                                          int *wmin, int *s_key, int *s_last);

// A round.  Workgroups [0, round_grid) sweep or walk the low-degree vertices; with hub_fused
// (one-sweep games with vertices above big_deg) the workgroups past them lift the static
// big and huge lists in dense rounds at the same time, [round_grid, + wave_grid) a wave per
// vertex and the rest split over the huge ones; in sparse rounds those return and the wave
// and workgroup kernels lift the deferred lists after this kernel.
template <int G, bool HUBS>
__global__ __launch_bounds__(256) void k_round(DevGame g, int j) {
    const int r = g.ring[14] + j;
    __shared__ int pre[kSub + 2];
    __shared__ int sbuf[kLdsCap];
    __shared__ int scnt, smode;
    __shared__ int wmin[4];
    __shared__ int s_key, s_last;
    const int in_slot = r % 3;
    if (threadIdx.x == 0) scnt = 0;
    load_layout(g, in_slot, pre);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int want = decide_mode(g, r, pre[kSub + 1]);
        smode = gate_mode(g, want, r);
        if (blockIdx.x == 0 && smode != want && g.ring[12] == 0) g.ring[13]++;   // rounds that wanted the transpose
    }
    __syncthreads();
    const int mode = smode;
    if (blockIdx.x == 0) {
        // prepare round r+1: its output counters and its deferral counters
        int *z = g.lcnt + ((r + 2) % 3) * kListCounters;
        for (int i = threadIdx.x; i < kListCounters; i += blockDim.x) z[i] = 0;
        if (threadIdx.x == 0) {
            g.mode[in_slot] = mode;
            g.ring[3 + (r + 1) % 3] = 0;
            g.ring[6 + (r + 1) % 3] = 0;
            if (mode != MODE_DONE) *g.rounds_done = r;
        }
    }
    if (mode == MODE_DONE) return;
    if (HUBS && (int)blockIdx.x >= g.round_grid) {       // hub part of a fused round
        if (mode != MODE_DENSE) return;
        const int hb = (int)blockIdx.x - g.round_grid;
        if (hb < g.wave_grid) hub_wave(g, r, mode, hb, g.wave_grid, sbuf, &scnt);
        else hub_block(g, r, mode, hb - g.wave_grid, g.block_grid, sbuf, &scnt, wmin, &s_key, &s_last);
        return;
    }
    if (mode == MODE_DENSE && g.phases > 1) return;    // the phase kernels sweep this round
    const Stage st{sbuf, &scnt, g.list[(r + 1) & 1], g.lcnt + ((r + 1) % 3) * kListCounters,
                   (int)(blockIdx.x % kSub)};
    LiftAcc acc;
    if (mode == MODE_DENSE) body_dense<G>(g, st, r, acc);
    else if (mode == MODE_QUEUE) body_queue<G>(g, st, g.list[r & 1], pre, r, acc);
    else body_lifted<G>(g, st, g.list[r & 1], pre, r, acc);
    stage_flush(g, st, 1);
    flush_stats(g, acc);
}

// Exclusive prefix sum over the 64 lanes of a wave (every lane active).
__device__ __forceinline__ int wave_excl_scan(int x) { return egs_wave_incl_scan(x) - x; }

// A phased dense round in one launch.  Workgroup b takes the tiles b, b + W, b + 2W, ...
// (W = the grid) and walks the phases s = 0 .. phases-1 in order, folding each vertex's
// edges into slice s into its partial key, kept in LDS for the whole round (no per-vertex
// array between phases).  The grid is about what the device holds at once, so its
// workgroups move through the slices together and the value gathers of the moment come
// from one slice, which stays in L2.  After the last phase each low-degree vertex applies
// its lift, stamps its rise and is staged for the next round like body_dense.  Values are
// read in place: a workgroup still in an early phase may see values another has already
// raised in this round, which only helps (values never pass the least fixpoint, and every
// vertex is swept again next round while anything rose).  Vertices above big_deg have no
// edges in the slices; the static wave/block kernels lift them after this kernel.
constexpr int kPhaseU = 2;   // edges in flight per lane

template <bool PACK>
__global__ __launch_bounds__(256) void k_dense_fused(DevGame g, int j) {
    const int r = g.ring[14] + j;
    extern __shared__ int keys[];                 // [tile of this workgroup][kSliceTile]
    __shared__ int sbuf[kLdsCap];
    __shared__ int scnt;
    if (g.mode[r % 3] != MODE_DENSE) return;
    if (threadIdx.x == 0) scnt = 0;
    __syncthreads();
    const Stage st{sbuf, &scnt, g.list[(r + 1) & 1], g.lcnt + ((r + 1) % 3) * kListCounters,
                   (int)(blockIdx.x % kSub)};
    // Until the transpose exists the next round can only sweep densely, which reads the
    // output list's count and not its items: then the risers are counted, not written.
    const bool items = g.ring[12] != 0;
    LiftAcc acc;
    const int P = g.phases;
    const int64_t tiles = ((int64_t)g.n + kSliceTile - 1) / kSliceTile;
    const int64_t W = gridDim.x;
    const int K = blockIdx.x < tiles ? (int)((tiles - blockIdx.x + W - 1) / W) : 0;
    const int wave = threadIdx.x >> 6;
    // vertices already at TOP (final) skip their gathers; at most 32 tiles per workgroup.
    // ring[15] tells whether any vertex has reached TOP (until then the mask is not read)
    uint32_t settled = 0;
    if (g.ring[15])
        for (int k = 0; k < K; k++) {
            const int64_t v = (blockIdx.x + k * W) * kSliceTile + threadIdx.x;
            if (v < g.n && g.val[v] == EGS_TOP) settled |= 1u << k;
        }
    uint64_t edges = 0;
    // the next (slice, tile)'s length byte and wave base load while this one folds
    int len_n = 0, base_n = 0;
    if (K > 0) {
        const size_t seg = (size_t)blockIdx.x * P;
        len_n = g.slen[seg * kSliceTile + threadIdx.x];
        base_n = g.sbase[seg * kSliceWaves + wave];
    }
    for (int s = 0; s < P; s++) {
        for (int k = 0; k < K; k++) {
            const int64_t T = blockIdx.x + k * W;
            const int len = len_n;
            const int b = base_n + wave_excl_scan(len);
            {
                const int kn = k + 1 < K ? k + 1 : 0, sn = k + 1 < K ? s : s + 1;
                if (sn < P) {
                    const size_t seg = (size_t)(blockIdx.x + kn * W) * P + sn;
                    len_n = g.slen[seg * kSliceTile + threadIdx.x];
                    base_n = g.sbase[seg * kSliceWaves + wave];
                }
            }
            const int e = ((settled >> k) & 1u) ? b : b + len;   // the prefix counts every lane's edges
            const bool owner1 = T * kSliceTile + threadIdx.x >= g.n0;
            int key = EGS_TOP;
            // kPhaseU edges per lane at a time: their loads, then their gathers, are all in
            // flight together (a slot past e re-reads edge k0, a valid index, and is masked)
            for (int k0 = b; k0 < e; k0 += kPhaseU) {
                int u[kPhaseU], w[kPhaseU], x[kPhaseU];
#pragma unroll
                for (int j = 0; j < kPhaseU; j++) slice_edge<PACK>(g, k0 + j < e ? k0 + j : k0, u[j], w[j]);
#pragma unroll
                for (int j = 0; j < kPhaseU; j++) x[j] = k0 + j < e ? g.val[u[j]] : EGS_TOP;
#pragma unroll
                for (int j = 0; j < kPhaseU; j++)
                    if (k0 + j < e) key = min(key, egs_key(egs_ominus(x[j], w[j], g.mg), owner1));
            }
            edges += (uint32_t)(e - b);
            int &slot = keys[k * kSliceTile + threadIdx.x];
            slot = s == 0 ? key : min(slot, key);
        }
    }
    for (int k = 0; k < K; k++) {
        const int64_t T = blockIdx.x + k * W;
        const int v = (int)(T * kSliceTile + threadIdx.x);
        bool lifted = false;
        if (v < g.n && !((settled >> k) & 1u) && !((g.bigmask[v >> 6] >> (v & 63)) & 1ull)) {
            const int t = egs_unkey(keys[k * kSliceTile + threadIdx.x], v >= g.n0);
            const int old = g.val[v];
            acc.processed++;
            if (old < t) {
                g.val[v] = t;
                if (t == EGS_TOP) g.ring[15] = 1;
                if (g.rise) g.rise[v] = r;
                acc.lifted++;
                lifted = true;
            }
        }
        if (items) {
            stage_push(g, st, lifted, v);
            if ((k & 7) == 7) stage_flush(g, st, kLdsCap / 2);
        } else {
            const uint64_t mask = __ballot(lifted);
            if (egs_lane() == 0 && mask) atomicAdd(&scnt, __popcll(mask));
        }
    }
    if (items) {
        stage_flush(g, st, 1);
    } else {
        __syncthreads();
        if (threadIdx.x == 0) {
            int bs, ns, bo;
            if (scnt > 0) reserve_split(g, st.cnts, st.sub, scnt, bs, ns, bo);
            if (blockIdx.x == 0) g.mode[3] = r + 1;      // round r+1's input: the count only
        }
    }
    acc.edges = edges;
    flush_stats(g, acc);
}

// Slice layout of the low-degree vertices' edges in one pass, one tile of kSliceTile
// vertices per workgroup iteration and no global scan.  G lanes read a vertex's edges (a
// wave reads 64 consecutive edges of 64 / G consecutive vertices per load) and count them
// per slice with LDS atomics; the workgroup sums the counts per (slice, wave) and scans the
// sums in that order from ptr of the tile's first vertex (the wave bases); the same lanes
// then place the edges, each vertex's after those of the lower vertices of its wave.  A
// vertex's first G edges stay in registers between the two steps, so a game of average
// degree up to G reads its edges once.  Vertices past n and those above big_deg get empty
// segments.  The same pass marks the vertices above big_deg in bigmask (a wave ballot) and
// decides k_init's initial activity of the others into stamp.
constexpr int kSliceQ = kSliceTile + 8;   // counter row per slice: rows 8 banks apart

// the slice set-up's reads of the CSR edges: nontemporal, so they do not push the slice
// edges just written out of the memory-side cache before the first round reads them
__device__ __forceinline__ int2 edge_nt(const int2 *e, int64_t k) {
    const long long x = __builtin_nontemporal_load((const long long *)e + k);
    return make_int2((int)(x & 0xffffffffll), (int)(x >> 32));
}

template <int G>
__global__ __launch_bounds__(kSliceTile) void k_slice_build(DevGame g, uint8_t *__restrict__ slen,
                                                           int *__restrict__ sbase, void *__restrict__ sedges,
                                                           unsigned long long *__restrict__ bigmask, int init) {
    constexpr int kGroups = kSliceTile / G;     // vertices one step of the workgroup covers
    __shared__ int s_cnt[kMaxPhases * kSliceQ];  // [slice][vertex of the tile]
    __shared__ int s_beg[kSliceTile], s_end[kSliceTile];
    __shared__ int s_wt[kMaxPhases * kSliceWaves];
    const int P = g.phases;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sub = threadIdx.x & (G - 1), grp = threadIdx.x / G;
    const int64_t tiles = ((int64_t)g.n + kSliceTile - 1) / kSliceTile;
    int active = 0;                               // init: this workgroup's initially active vertices
    for (int64_t T = blockIdx.x; T < tiles; T += gridDim.x) {
        const int64_t first = T * kSliceTile;
        const int64_t v = first + threadIdx.x;
        const bool valid = v < g.n;
        int vs = 0, ve = 0;
        if (valid) {
            vs = g.ptr[v];
            ve = g.ptr[v + 1];
        }
        const bool big = valid && ve - vs > g.big_deg;
        if (init && valid) {                      // k_init's stores (k_init_big decides the big ones)
            g.val[v] = 0;
            if (g.rise) g.rise[v] = 0;
            if (big) g.stamp[v] = 0;
        }
        const unsigned long long word = __ballot(big);
        if (lane == 0 && valid) bigmask[v >> 6] = word;
        s_beg[threadIdx.x] = vs;
        s_end[threadIdx.x] = big ? vs - 1 : ve;   // big vertices (end before start): the hub kernels' edges
        for (int q = 0; q < P; q++) s_cnt[q * kSliceQ + threadIdx.x] = 0;
        __syncthreads();
        // count: every vertex's first G edges (kept), then the rest
        int2 ew[G];
        unsigned neg = 0, nonneg = 0;             // bit i: vertex i * kGroups + grp
#pragma unroll
        for (int i = 0; i < G; i++) {
            const int vl = i * kGroups + grp;
            const int k = s_beg[vl] + sub;
            ew[i] = k < s_end[vl] ? edge_nt(g.edges, k) : make_int2(-1, 0);
        }
#pragma unroll
        for (int i = 0; i < G; i++) {
            if (ew[i].x < 0) continue;
            atomicAdd(&s_cnt[(ew[i].x >> g.slice_shift) * kSliceQ + i * kGroups + grp], 1);
            (ew[i].y < 0 ? neg : nonneg) |= 1u << i;
        }
        for (int i = 0; i < G; i++) {
            const int vl = i * kGroups + grp;
            const int e = s_end[vl];
            for (int k = s_beg[vl] + G + sub; k < e; k += G) {
                const int2 x = edge_nt(g.edges, k);
                atomicAdd(&s_cnt[(x.x >> g.slice_shift) * kSliceQ + vl], 1);
                (x.y < 0 ? neg : nonneg) |= 1u << i;
            }
        }
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
            neg |= __shfl_xor(neg, d, 64);
            nonneg |= __shfl_xor(nonneg, d, 64);
        }
        {   // lane sub of each group decides vertex sub * kGroups + grp
            const int vl = sub * kGroups + grp;
            const int64_t u = first + vl;
            // lift(0)(u) > 0: owner 0 with every edge negative, owner 1 with a negative edge
            bool act = false;
            if (u < g.n && s_end[vl] >= s_beg[vl]) {
                act = (u < g.n0) ? !((nonneg >> sub) & 1) : ((neg >> sub) & 1);
                g.stamp[u] = act ? 1 : 0;
            }
            if (init) active += __syncthreads_count(act);
        }
        __syncthreads();
        uint8_t *len_out = slen + (size_t)T * P * kSliceTile;
        for (int q = 0; q < P; q++) {
            const int c = s_cnt[q * kSliceQ + threadIdx.x];
            len_out[q * kSliceTile + threadIdx.x] = (uint8_t)c;
            const int inc = egs_wave_incl_scan(c);
            s_cnt[q * kSliceQ + threadIdx.x] = inc - c;         // the wave's exclusive prefix
            if (lane == 63) s_wt[q * kSliceWaves + wave] = inc;
        }
        __syncthreads();
        // (slice, wave) sums scanned in that order: at most kMaxPhases * kSliceWaves = 64 items
        if (threadIdx.x < 64) {
            const int x = threadIdx.x < P * kSliceWaves ? s_wt[threadIdx.x] : 0;
            const int ex = wave_excl_scan(x);
            if (threadIdx.x < P * kSliceWaves) {
                const int b = g.ptr[first] + ex;
                s_wt[threadIdx.x] = b;
                sbase[(size_t)T * P * kSliceWaves + threadIdx.x] = b;
            }
        }
        __syncthreads();
        for (int q = 0; q < P; q++) s_cnt[q * kSliceQ + threadIdx.x] += s_wt[q * kSliceWaves + wave];
        __syncthreads();
        // place: each edge at its (slice, vertex) segment's next slot
        auto put = [&](const int2 x, int vl) {
            const int p = atomicAdd(&s_cnt[(x.x >> g.slice_shift) * kSliceQ + vl], 1);
            if (g.sshift) ((uint32_t *)sedges)[p] = ((uint32_t)x.y << g.sshift) | (uint32_t)x.x;
            else ((int2 *)sedges)[p] = x;
        };
#pragma unroll
        for (int i = 0; i < G; i++)
            if (ew[i].x >= 0) put(ew[i], i * kGroups + grp);
        for (int i = 0; i < G; i++) {
            const int vl = i * kGroups + grp;
            const int e = s_end[vl];
            for (int k = s_beg[vl] + G + sub; k < e; k += G) put(edge_nt(g.edges, k), vl);
        }
        __syncthreads();    // s_wt, s_beg, s_end and s_cnt are rewritten for the next tile
    }
    // init: round 1's input is only counted (without the transpose round 1 sweeps densely)
    if (init && threadIdx.x == 0) {
        int bs, ns, bo;
        if (active > 0) reserve_split(g, g.lcnt + 1 * kListCounters, (int)(blockIdx.x % kSub), active, bs, ns, bo);
        if (blockIdx.x == 0) g.mode[3] = 1;
    }
}

// One wavefront per vertex: the static big list in dense rounds, the deferred list in
// sparse rounds.  Workgroup `bid` of `nblk` (the wave kernel's grid, or the hub part of a
// round kernel's grid).
__device__ __forceinline__ void hub_wave(const DevGame &g, int r, int mode, int bid, int nblk, int *sbuf, int *scnt) {
    const bool dense = mode == MODE_DENSE;
    const int *items = dense ? g.sbig : g.dbig;
    const int cnt = dense ? g.ring[9] : g.ring[3 + r % 3];
    if (cnt == 0) return;
    if (threadIdx.x == 0) *scnt = 0;
    __syncthreads();
    const Stage st{sbuf, scnt, g.list[(r + 1) & 1], g.lcnt + ((r + 1) % 3) * kListCounters, bid % kSub};
    LiftAcc acc;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)bid * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t waves = (int64_t)nblk * (blockDim.x >> 6);
    for (int64_t i = wave; i < cnt; i += waves) {
        const int v = items[i];
        const int s = g.ptr[v], e = g.ptr[v + 1];
        const int old = g.val[v];
        if (old == EGS_TOP) continue;           // final; the whole wave skips it
        const int t = lift_value<64>(g, s, e, lane, v >= g.n0);
        const bool lifted = old < t;
        if (lane == 0) {
            acc.processed++;
            acc.edges += (uint32_t)(e - s);
            if (lifted) { g.val[v] = t; acc.lifted++; if (g.rise) g.rise[v] = r; if (t == EGS_TOP) g.ring[15] = 1; }
        }
        if (dense) stage_push(g, st, lifted && lane == 0, v);
        else push_preds<64>(g, st, v, lifted, lane, r, acc);
    }
    stage_flush(g, st, 1);
    flush_stats(g, acc);
}

__global__ __launch_bounds__(256) void k_lift_wave(DevGame g, int j) {
    const int r = g.ring[14] + j;
    __shared__ int sbuf[kLdsCap];
    __shared__ int scnt;
    const int mode = g.mode[r % 3];
    if (mode == MODE_DONE) return;
    if (mode == MODE_DENSE && g.hub_fused) return;     // the round kernel's hub part did it
    hub_wave(g, r, mode, (int)blockIdx.x, (int)gridDim.x, sbuf, &scnt);
}

// Vertices above huge_deg: the static huge list in dense rounds, the deferred list in
// sparse rounds.  A vertex is shared by up to `per` workgroups (the grid over the items):
// each takes every per-th chunk of kHugeChunk edges, folds its minimum key into hacc[v]
// with an atomic, and the workgroup that completes hcnt[v] applies the lift (value, rise
// stamp, list or predecessor push) and resets both for the next round.  One workgroup
// walking a hub of 70K edges alone made this kernel most of a skewed game's round.
constexpr int kHugeChunk = 2048;

__device__ __forceinline__ void hub_block(const DevGame &g, int r, int mode, int bid, int nblk, int *sbuf, int *scnt,
                                          int *wmin, int *s_key, int *s_last) {
    const bool dense = mode == MODE_DENSE;
    const int *items = dense ? g.shuge : g.dhuge;
    const int cnt = dense ? g.ring[10] : g.ring[6 + r % 3];
    if (cnt == 0) return;
    if (threadIdx.x == 0) *scnt = 0;
    __syncthreads();
    const Stage st{sbuf, scnt, g.list[(r + 1) & 1], g.lcnt + ((r + 1) % 3) * kListCounters, bid % kSub};
    LiftAcc acc;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int per = max(1, nblk / cnt);                  // workgroups per vertex
    const int64_t units = (int64_t)cnt * per;
    for (int64_t u = bid; u < units; u += nblk) {
        const int i = (int)(u % cnt), part = (int)(u / cnt);
        const int v = items[i];
        const int s = g.ptr[v], e = g.ptr[v + 1];
        const int chunks = (e - s + kHugeChunk - 1) / kHugeChunk;
        const int contrib = min(per, chunks);
        if (part >= contrib) continue;                   // the whole workgroup: u is shared
        // a TOP value is final; every contributor sees it (only v's own last contributor
        // changes v's value in a round, after all others have counted in)
        if (g.val[v] == EGS_TOP) continue;
        const bool owner1 = v >= g.n0;
        int key = EGS_TOP;
        for (int c = part; c < chunks; c += contrib) {
            const int cs = s + c * kHugeChunk, ce = min(e, cs + kHugeChunk);
            key = min(key, key_over_edges<8>(g, cs, ce, (int)threadIdx.x, 256, owner1));
        }
        key = egs_group_min<64>(key);
        if (lane == 0) wmin[wid] = key;
        __syncthreads();
        if (threadIdx.x == 0) {
            key = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
            for (int c = part; c < chunks; c += contrib) acc.edges += (uint32_t)(min(e, s + (c + 1) * kHugeChunk) - (s + c * kHugeChunk));
            bool last = true;
            if (contrib > 1) {
                atomicMin(&g.hacc[v], key);
                __threadfence();
                last = atomicAdd(&g.hcnt[v], 1) == contrib - 1;
                if (last) {
                    __threadfence();
                    key = atomicExch(&g.hacc[v], EGS_TOP);   // every contribution, and the reset
                    g.hcnt[v] = 0;
                }
            }
            *s_key = key;
            *s_last = last;
        }
        __syncthreads();
        if (*s_last) {
            const int t = egs_unkey(*s_key, owner1);
            const int old = g.val[v];
            const bool lifted = old < t;
            if (threadIdx.x == 0) {
                acc.processed++;
                if (lifted) { g.val[v] = t; acc.lifted++; if (g.rise) g.rise[v] = r; if (t == EGS_TOP) g.ring[15] = 1; }
            }
            if (dense) stage_push(g, st, lifted && threadIdx.x == 0, v);
            else push_preds<256>(g, st, v, lifted, (int)threadIdx.x, r, acc);
        }
        stage_flush(g, st, kLdsCap / 2);
        __syncthreads();
    }
    stage_flush(g, st, 1);
    flush_stats(g, acc);
}

__global__ __launch_bounds__(256) void k_lift_block(DevGame g, int j) {
    const int r = g.ring[14] + j;
    __shared__ int sbuf[kLdsCap];
    __shared__ int scnt;
    __shared__ int wmin[4];
    __shared__ int s_key, s_last;
    const int mode = g.mode[r % 3];
    if (mode == MODE_DONE) return;
    if (mode == MODE_DENSE && g.hub_fused) return;     // the round kernel's hub part did it
    hub_block(g, r, mode, (int)blockIdx.x, (int)gridDim.x, sbuf, &scnt, wmin, &s_key, &s_last);
}

// Both hub lifts of a round in one launch, for games with both lists: workgroups
// [0, wave_grid) run k_lift_wave's body, the next block_grid k_lift_block's.  They lift
// disjoint vertices (as the hub part of a fused round kernel does), so one launch serves
// both and a round whose hubs were lifted by the round kernel pays for one empty launch.
__global__ __launch_bounds__(256) void k_lift_hubs(DevGame g, int j, int wave_grid, int block_grid) {
    const int r = g.ring[14] + j;
    __shared__ int sbuf[kLdsCap];
    __shared__ int scnt;
    __shared__ int wmin[4];
    __shared__ int s_key, s_last;
    const int mode = g.mode[r % 3];
    if (mode == MODE_DONE) return;
    if (mode == MODE_DENSE && g.hub_fused) return;     // the round kernel's hub part did it
    if ((int)blockIdx.x < wave_grid) hub_wave(g, r, mode, (int)blockIdx.x, wave_grid, sbuf, &scnt);
    else hub_block(g, r, mode, (int)blockIdx.x - wave_grid, block_grid, sbuf, &scnt, wmin, &s_key, &s_last);
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------

// The counters, modes and ring slots start cleared (k_prepare_reset).
hipError_t launch_init(const DevGame &g, hipStream_t st, bool in_slices) {
    if (!in_slices) hipLaunchKernelGGL(k_init, dim3(grid_for(g.n, 256, 2048)), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_init_big, dim3(1024), dim3(256), 0, st, g);
    return hipGetLastError();
}

// The round kernels of a batch take their offset j in the batch; the absolute round number
// is ring[14] + j, and k_advance moves ring[14] past the batch after its last kernel.  A
// batch's launches are then the same every time, so a batch can be captured as a graph once
// and replayed.
__global__ void k_advance(DevGame g, int rounds) {
    if (blockIdx.x == 0 && threadIdx.x == 0) g.ring[14] += rounds;
}

hipError_t launch_advance(const DevGame &g, int rounds, hipStream_t st) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, g, rounds);
    return hipGetLastError();
}

hipError_t launch_round(const DevGame &g, const LaunchShape &sh, int j, hipStream_t st) {
    const dim3 grid(g.round_grid + (g.hub_fused ? g.wave_grid + g.block_grid : 0));
#define EGS_ROUND(GG)                                                                            \
    if (g.hub_fused) hipLaunchKernelGGL((k_round<GG, true>), grid, dim3(256), 0, st, g, j);     \
    else hipLaunchKernelGGL((k_round<GG, false>), grid, dim3(256), 0, st, g, j);
    switch (sh.group) {
    case 1: EGS_ROUND(1) break;
    case 2: EGS_ROUND(2) break;
    case 4: EGS_ROUND(4) break;
    case 8: EGS_ROUND(8) break;
    case 16: EGS_ROUND(16) break;
    case 32: EGS_ROUND(32) break;
    default: EGS_ROUND(64) break;
    }
#undef EGS_ROUND
    if (g.phases > 1) {
        const size_t lds = (size_t)g.phase_tiles * kSliceTile * sizeof(int);
        if (g.sshift)
            hipLaunchKernelGGL(k_dense_fused<true>, dim3(g.phase_grid), dim3(256), lds, st, g, j);
        else
            hipLaunchKernelGGL(k_dense_fused<false>, dim3(g.phase_grid), dim3(256), lds, st, g, j);
    }
    if (sh.use_wave && sh.use_block)
        hipLaunchKernelGGL(k_lift_hubs, dim3(sh.wave_grid + sh.block_grid), dim3(256), 0, st, g, j, sh.wave_grid,
                           sh.block_grid);
    else if (sh.use_wave)
        hipLaunchKernelGGL(k_lift_wave, dim3(sh.wave_grid), dim3(256), 0, st, g, j);
    else if (sh.use_block)
        hipLaunchKernelGGL(k_lift_block, dim3(sh.block_grid), dim3(256), 0, st, g, j);
    return hipGetLastError();
}

hipError_t launch_slices(const DevGame &g, hipStream_t st, bool init) {
    if (g.phases <= 1) return hipSuccess;
    const int in = init ? 1 : 0;
    // lanes per vertex: the average degree rounded up to 4, 8 or 16
    const int64_t avg = (g.m + g.n - 1) / g.n;
    const dim3 grid(grid_for(g.n, kSliceTile, 4096));
    uint8_t *slen = const_cast<uint8_t *>(g.slen);
    int *sbase = const_cast<int *>(g.sbase);
    void *sedges = const_cast<void *>(g.sedges);
    unsigned long long *bigmask = const_cast<unsigned long long *>(g.bigmask);
    if (avg <= 4)
        hipLaunchKernelGGL(k_slice_build<4>, grid, dim3(kSliceTile), 0, st, g, slen, sbase, sedges, bigmask, in);
    else if (avg <= 8)
        hipLaunchKernelGGL(k_slice_build<8>, grid, dim3(kSliceTile), 0, st, g, slen, sbase, sedges, bigmask, in);
    else
        hipLaunchKernelGGL(k_slice_build<16>, grid, dim3(kSliceTile), 0, st, g, slen, sbase, sedges, bigmask, in);
    return hipGetLastError();
}

}  // namespace egs
