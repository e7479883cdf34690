// egs_games.h -- the many-games kernel shared by the batched solver (egs_batch.hip) and
// the serving loop (egs_serve.hip).
//
// A launch runs over a list of game descriptors.  A persistent workgroup claims one whole
// game at a time from a device counter and runs in-place (Gauss-Seidel) sweeps over its
// vertices until a complete sweep changes nothing (a block-wide vote), the launch's sweep
// cap for that game is reached, or the workgroup's edge budget for the launch is spent.
// A sweep may stop part-way at a chunk boundary: the game's GameState keeps the vertex
// cursor and whether the sweep has changed anything so far, so a later launch resumes it.
// A sweep that changed nothing proves the fixpoint even when it spanned launches: no value
// moved while it ran, so every vertex was checked against the final values.
// The first values of a game (as many as the workgroup's LDS share holds) sit in LDS, the
// rest in global memory.
// Every `every` sweeps (doubling while proofs fail) the workgroup tries the certified TOP
// proof of egs_accel.hip on its game.
#pragma once

#include <climits>

#include "egs_device.h"

namespace egs {

struct GameState {
    int64_t sweeps;        // sweeps done so far
    int64_t edges;         // edges relaxed so far
    int64_t lifted;
    int64_t cert_passes;
    int64_t next_cert;     // sweep count of the next proof attempt
    int32_t gap;           // current attempt spacing
    int32_t converged;     // 1 once a sweep changed nothing
    int32_t certificates;
    int32_t started;       // values initialised
    int32_t bf_cap;        // longest-path passes per candidate threshold (doubles while attempts run out)
    int32_t cursor;        // first vertex of the current sweep not yet visited
    int32_t dirty;         // the current sweep has changed a value
    int32_t reserved;
};

struct GameDesc {
    const int *ptr;        // n+1 local CSR offsets
    const void *edges;     // m edges: uint32 (w << shift) | succ, or int2 {succ, w} when shift == 0
    int *values;           // n
    int *rise;             // n: number of the sweep in which each vertex last rose (proof candidates)
    int *h;                // n scratch (proofs: potentials above the warm start)
    GameState *state;
    int64_t m;
    int n, n0, mg, wmax;
    int shift, group;
};

constexpr int kGameFirstCap = 32;      // longest-path passes per candidate at a game's first attempt
constexpr int kGameCandidates = 8;     // gap thresholds tried per attempt (plus the last-rise set)
constexpr int kGameCertifyEvery = 4;   // default sweeps before the first proof attempt

struct CertArgs {
    int every;             // sweeps between proof attempts; <= 0 = no proofs
    int max_passes;        // ceiling of the longest-path passes per candidate threshold
    int first_cap;         // passes per candidate at a game's first attempt
    int candidates;        // thresholds tried per attempt
};

// A game's values: the first `split` in LDS (lo), the rest in the game's global array (hi,
// indexed by vertex, so hi[v] for v >= split).  A game that fits the workgroup's LDS share
// has split == n; a larger one keeps most of its gathers in LDS and the remainder in L2.
struct GameVals {
    int *lo;
    int *hi;
    int split;
    __device__ __forceinline__ int get(int v) const { return v < split ? lo[v] : hi[v]; }
    __device__ __forceinline__ void set(int v, int x) const {
        if (v < split) lo[v] = x;
        else hi[v] = x;
    }
};

template <bool SPLIT>
__device__ __forceinline__ int gv(const GameVals &V, int v) { return SPLIT ? V.get(v) : V.lo[v]; }

__device__ __forceinline__ void game_edge(const GameDesc &b, int64_t k, int &succ, int &w) {
    if (b.shift) {
        const uint32_t e = ((const uint32_t *)b.edges)[k];
        succ = (int)(e & ((1u << b.shift) - 1u));
        w = ((int)e) >> b.shift;
    } else {
        const int2 e = ((const int2 *)b.edges)[k];
        succ = e.x;
        w = e.y;
    }
}

// Edge loads in flight per lane: a lane's edges are read one after another from global
// memory, and one load at a time leaves the loops latency-bound.  Slots past a vertex's
// last edge read vertex 0 and are not used.
constexpr int kGameU = 4;

// Part of a sweep by groups of G lanes, from vertex `cursor`, chunk by chunk (a chunk is
// one vertex per group) while `spent` < `budget`; returns the next cursor.  `spent` adds
// the chunk's edge count, read from the offsets by every lane, so it stays block-uniform
// and the loop keeps waves whole for the DPP reductions.
template <int G, bool SPLIT>
__device__ __forceinline__ int game_sweep_part(const GameDesc &b, const GameVals &V, int cursor, int64_t budget, int sweep_no,
                                               int64_t &spent, int64_t &edges_done, int &lifted, bool &changed) {
    const int gl = threadIdx.x % G;
    const int gid = threadIdx.x / G;
    const int ngroups = blockDim.x / G;
    int base = cursor;
    // the next chunk's offsets load while this chunk runs (they never change)
    int s_n = 0, e_n = 0, c0_n = 0, c1_n = 0;
    auto fetch = [&](int at) {
        const int u = at + gid;
        if (u < b.n) {
            s_n = b.ptr[u];
            e_n = b.ptr[u + 1];
        }
        c0_n = b.ptr[at];
        c1_n = b.ptr[min(at + ngroups, b.n)];
    };
    if (base < b.n) fetch(base);
    while (base < b.n && spent < budget) {
        spent += c1_n - c0_n;
        const int v = base + gid;
        const bool valid = v < b.n;
        int s = 0, e = 0, old = 0;
        if (valid) {
            s = s_n;
            e = e_n;
        }
        if (base + ngroups < b.n) fetch(base + ngroups);
        if (valid) {
            old = gv<SPLIT>(V, v);
            if (old == EGS_TOP) e = s;          // TOP is final: nothing to gather
        }
        const bool owner1 = v >= b.n0;
        int key = EGS_TOP;
        for (int k0 = s + gl; k0 < e; k0 += G * kGameU) {
            int su[kGameU], wv[kGameU], a[kGameU];
#pragma unroll
            for (int j = 0; j < kGameU; j++) {
                su[j] = 0;
                wv[j] = 0;
                if (k0 + j * G < e) game_edge(b, k0 + j * G, su[j], wv[j]);
            }
#pragma unroll
            for (int j = 0; j < kGameU; j++) a[j] = gv<SPLIT>(V, su[j]);
#pragma unroll
            for (int j = 0; j < kGameU; j++)
                if (k0 + j * G < e) key = min(key, egs_key(egs_ominus(a[j], wv[j], b.mg), owner1));
        }
        key = egs_group_min<G>(key);
        const int t = egs_unkey(key, owner1);
        if (valid && gl == 0) {
            edges_done += e - s;
            if (old < t) {
                if (SPLIT) V.set(v, t);
                else V.lo[v] = t;
                b.rise[v] = sweep_no;
                changed = true;
                lifted++;
            }
        }
        base += ngroups;
    }
    return min(base, b.n);
}

// The certified TOP proof inside the workgroup (the steps of egs_accel.hip).  Candidate
// sets D, tried in turn until one yields a non-empty proven set:
//   * every admissible gap threshold B, smallest first: D = {v : B < f(v) < TOP};
//   * the vertices that rose in the last sweep (`last`): D = {v : rise(v) == last, f(v) < TOP}.
// When no window is empty, the smallest threshold whose window holds only a few values
// takes the place of the gap thresholds.
// Check 1 (closure): an owner-0 vertex of D has every successor in D or TOP, an owner-1
// vertex at least one; vertices that break it are removed until it holds (the greatest
// closed subset), during the passes of check 2: the longest-path game on D with w' = (n+1) w + 1,
//   h(v) <- max(h(v), max_{u in D} w'(v,u) + h(u))           (owner 0)
//   h(v) <- max(h(v), min_{u in D or TOP} w'(v,u) + h(u))    (owner 1; a TOP successor ends the path)
// reaches a fixpoint within `cap` passes.  At a fixpoint the minimising successors form a
// strategy sigma of player 1 under which every cycle inside D has weight <= -1, which is
// the proof; any fixpoint will do, so the iteration starts from h = (n+1) (MG - f): the
// lifting keeps f(u) - f(v) >= w(v,u) along such edges, so this start is already close.
// Returns the vertices set to TOP; `cap_hit` tells whether some candidate ran out of passes.
// Potentials are kept as d(v) = h(v) - (n+1) (MG - f(v)) >= 0, 4 bytes next to the values;
// kGameDOut marks a vertex outside D (TOP is read off the value).  A d past INT_MAX ends the
// attempt as if it diverged: it only gives up a proof.
constexpr int kGameDOut = -1;

__device__ __forceinline__ bool game_in_d_x(const GameDesc &b, int x, int v, int B, int last) {
    return x != EGS_TOP && (B >= 0 ? x > B : b.rise[v] == last);
}

__device__ int game_certify(const GameDesc &b, const GameVals &V, int *dp, int last, int cap, int max_cands,
                            int64_t &passes_done,
                            bool &cap_hit) {
    __shared__ int win[64], flo[64];
    __shared__ unsigned long long s_cand;
    __shared__ int s_marked;
    const int wmax = b.wmax;
    if (wmax <= 0) return 0;
    const long long scale = (long long)b.n + 1;
    // divergence bound: at a fixpoint h <= max start + (n - 1) max w'; 2x headroom in int64
    const double bound_d = (double)scale * b.mg + (double)b.n * ((double)scale * wmax + 1);
    if (bound_d > 4.0e18) return 0;
    const long long bound = (long long)scale * b.mg + (long long)b.n * (scale * wmax + 1);
    for (int i = threadIdx.x; i < 64; i += blockDim.x) { win[i] = 0; flo[i] = 0; }
    __syncthreads();
    for (int v = threadIdx.x; v < b.n; v += blockDim.x) {
        const int x = V.get(v);
        if (x == EGS_TOP || x <= 0) continue;
        const int64_t q = ((int64_t)x + wmax - 1) / wmax;
        if ((q & (q - 1)) == 0) atomicAdd(&win[63 - __clzll((unsigned long long)q)], 1);
        if (x > wmax) atomicAdd(&flo[63 - __clzll((unsigned long long)(((int64_t)x - 1) / wmax))], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // bit j set: B_j = W * 2^j has an empty window, non-TOP values above it, B_j <= MG - W
        unsigned long long mask = 0, near = 0;
        int64_t above = 0;                    // #{non-TOP values > W * 2^j} while j descends
        for (int k = 63; k >= 40; k--) above += flo[k];
        for (int j = 39; j >= 1; j--) {
            above += flo[j];
            if (((int64_t)wmax << j) > (int64_t)b.mg - wmax || above == 0) continue;
            if (win[j] == 0) mask |= 1ull << j;
            else if (win[j] <= max((int64_t)64, above / 256)) near = 1ull << j;   // ends as the smallest
        }
        // no empty window: the smallest threshold whose window holds a few stragglers (the
        // members whose edges reach them are pruned)
        if (mask == 0) mask = near;
        unsigned long long keep = 0;          // the lowest max_cands thresholds
        for (int k = 0; k < max_cands && mask; k++) {
            keep |= mask & (~mask + 1);
            mask &= mask - 1;
        }
        s_cand = keep | 1ull;                 // bit 0: the vertices that rose in the last sweep
        s_marked = 0;
    }
    __syncthreads();
    unsigned long long cand = s_cand;
    while (cand) {
        // gap thresholds first, smallest first; the last-rise set after them
        const bool recent = (cand & ~1ull) == 0;
        int B;
        if (recent) {
            B = -1;
            cand = 0;
        } else {
            const int j = __ffsll(cand & ~1ull) - 1;
            cand &= ~(1ull << j);
            B = (int)((int64_t)wmax << j);
        }
        // pass 0: membership, closure and the first longest-path step, from the values alone
        // (the warm start of a successor is (n+1) (MG - f)).  A member that breaks closure
        // leaves D (h = kGameHOut); later passes check it again against the removals, so the
        // passes compute the greatest closed subset of the candidate along with h.
        bool ch = false, over = false;
        int members = 0;
        for (int v = threadIdx.x; v < b.n; v += blockDim.x) {
            const int x = V.get(v);
            if (!game_in_d_x(b, x, v, B, last)) { dp[v] = kGameDOut; continue; }
            members++;
            const long long h0 = scale * ((long long)b.mg - x);
            const bool owner1 = v >= b.n0;
            long long best = owner1 ? LLONG_MAX : LLONG_MIN;
            bool any = false, all = true;
            const int s = b.ptr[v], e = b.ptr[v + 1];
            for (int k0 = s; k0 < e; k0 += kGameU) {
                int uu[kGameU], ww[kGameU], xu[kGameU];
#pragma unroll
                for (int j = 0; j < kGameU; j++) {
                    uu[j] = 0;
                    ww[j] = 0;
                    if (k0 + j < e) game_edge(b, k0 + j, uu[j], ww[j]);
                }
#pragma unroll
                for (int j = 0; j < kGameU; j++) xu[j] = V.get(uu[j]);
#pragma unroll
                for (int j = 0; j < kGameU; j++) {
                    if (k0 + j >= e) break;
                    if (xu[j] == EGS_TOP) {
                        any = true;
                        if (owner1) best = LLONG_MIN;
                    } else if (game_in_d_x(b, xu[j], uu[j], B, last)) {
                        any = true;
                        const long long c = scale * ww[j] + 1 + scale * ((long long)b.mg - xu[j]);
                        best = owner1 ? min(best, c) : max(best, c);
                    } else {
                        all = false;
                    }
                }
            }
            if (owner1 ? !any : !all) {
                dp[v] = kGameDOut;
                ch = true;
            } else if (best > h0) {
                dp[v] = (int)min(best - h0, (long long)INT_MAX);
                ch = true;
                over |= best > bound || best - h0 >= INT_MAX;
            } else {
                dp[v] = 0;
            }
        }
        if (__syncthreads_count(members > 0) == 0) continue;
        passes_done++;
        bool converged = __syncthreads_count(ch) == 0, diverged = __syncthreads_or(over) != 0;
        for (int p = 1; p < cap && !converged && !diverged; p++) {
            ch = false;
            over = false;
            for (int v = threadIdx.x; v < b.n; v += blockDim.x) {
                const int dv = dp[v];
                if (dv < 0) continue;
                const long long h0 = scale * ((long long)b.mg - V.get(v));
                const long long hv = h0 + dv;
                const bool owner1 = v >= b.n0;
                long long best = owner1 ? LLONG_MAX : LLONG_MIN;
                bool out = false;
                const int s = b.ptr[v], e = b.ptr[v + 1];
                for (int k0 = s; k0 < e; k0 += kGameU) {
                    int uu[kGameU], ww[kGameU], xu[kGameU], du[kGameU];
#pragma unroll
                    for (int j = 0; j < kGameU; j++) {
                        uu[j] = 0;
                        ww[j] = 0;
                        if (k0 + j < e) game_edge(b, k0 + j, uu[j], ww[j]);
                    }
#pragma unroll
                    for (int j = 0; j < kGameU; j++) {
                        xu[j] = V.get(uu[j]);
                        du[j] = dp[uu[j]];
                    }
#pragma unroll
                    for (int j = 0; j < kGameU; j++) {
                        if (k0 + j >= e) break;
                        if (xu[j] == EGS_TOP) {       // ends the path: -infinity for the game
                            if (owner1) best = LLONG_MIN;
                            continue;
                        }
                        if (du[j] < 0) { out = true; continue; }
                        const long long c = scale * ww[j] + 1 + scale * ((long long)b.mg - xu[j]) + du[j];
                        best = owner1 ? min(best, c) : max(best, c);
                    }
                }
                // owner 1 breaks closure when no edge reaches D or TOP, owner 0 when any leaves
                if (owner1 ? best == LLONG_MAX : out) {
                    dp[v] = kGameDOut;
                    ch = true;
                } else if (best > hv) {
                    dp[v] = (int)min(best - h0, (long long)INT_MAX);
                    ch = true;
                    over |= best > bound || best - h0 >= INT_MAX;
                }
            }
            passes_done++;
            converged = __syncthreads_count(ch) == 0;
            diverged = __syncthreads_or(over) != 0;
        }
        if (diverged || !converged) {
            if (!diverged) cap_hit = true;
            continue;
        }
        int mine = 0;
        for (int v = threadIdx.x; v < b.n; v += blockDim.x)
            if (dp[v] >= 0) { V.set(v, EGS_TOP); mine++; }
        if (mine) atomicAdd(&s_marked, mine);
        __syncthreads();
        if (s_marked > 0) return s_marked;
        __syncthreads();                      // pruning emptied the candidate: try the next
    }
    return 0;
}

// Advance one game until its fixpoint, a due proof attempt, `max_sweeps` more complete
// sweeps, or `budget` edges (checked per chunk).  `sst` is the workgroup's shared copy of
// the game's state; the proof itself runs in game_prove, outside the lane-width variants,
// so the kernel carries one copy of it.  Adds the edges charged to the budget to `spent`
// and the complete sweeps to `sweeps` (both block-uniform); returns whether it stopped
// because a proof attempt is due.
template <int G>
__device__ bool game_run(const GameDesc &b, const GameVals &V, GameState &sst, int64_t budget, int64_t max_sweeps,
                         const CertArgs &ca, int64_t &spent, int64_t &sweeps) {
    // the fields the loop changes, as scalars (a whole-struct copy ends up in scratch)
    int cursor = sst.cursor;
    bool converged = sst.converged != 0, dirty = sst.dirty != 0;
    const int64_t sweeps0 = sst.sweeps, next_cert = sst.next_cert;
    __syncthreads();
    int64_t edges_done = 0, used = 0, done = 0;
    int lifted = 0;
    bool proof_due = false;
    while (!converged && used < budget && done < max_sweeps) {
        bool ch = false;
        if (V.split < b.n)
            cursor = game_sweep_part<G, true>(b, V, cursor, budget, (int)(sweeps0 + done + 1), used, edges_done, lifted, ch);
        else
            cursor = game_sweep_part<G, false>(b, V, cursor, budget, (int)(sweeps0 + done + 1), used, edges_done, lifted, ch);
        dirty = __syncthreads_or(ch || dirty) != 0;
        if (cursor < b.n) break;                    // budget spent part-way through the sweep
        done++;
        cursor = 0;
        if (!dirty) {
            converged = true;
            break;
        }
        dirty = false;
        if (ca.every > 0 && sweeps0 + done >= next_cert) {
            proof_due = true;
            break;
        }
    }
    spent += used;
    sweeps += done;
    __shared__ unsigned long long red[2];
    if (threadIdx.x == 0) { red[0] = 0; red[1] = 0; }
    __syncthreads();
    if (edges_done) atomicAdd(&red[0], (unsigned long long)edges_done);
    if (lifted) atomicAdd(&red[1], (unsigned long long)lifted);
    __syncthreads();
    if (threadIdx.x == 0) {
        sst.edges += (int64_t)red[0];
        sst.lifted += (int64_t)red[1];
        sst.sweeps = sweeps0 + done;
        sst.cursor = cursor;
        sst.converged = converged ? 1 : 0;
        sst.dirty = dirty ? 1 : 0;
    }
    __syncthreads();
    return proof_due;
}

// One proof attempt at a sweep boundary; schedules the next one (spacing back to `every`
// after a proof, doubled after a failure, pass cap doubled when a candidate ran out).
// `dp` holds the potentials: LDS after the values when both fit, else the game's array.
__device__ void game_prove(const GameDesc &b, const GameVals &V, int *dp, GameState &sst, const CertArgs &ca) {
    const int cap = sst.bf_cap;
    __syncthreads();
    int64_t passes = 0;
    bool cap_hit = false;
    const int marked = game_certify(b, V, dp, (int)sst.sweeps, cap, ca.candidates, passes, cap_hit);
    const bool hit = __syncthreads_or(cap_hit) != 0;
    if (threadIdx.x == 0) {
        sst.cert_passes += passes;
        if (marked > 0) {
            sst.certificates++;
            sst.gap = ca.every;
        } else {
            sst.gap = min(sst.gap * 2, 1 << 16);
            if (hit) sst.bf_cap = min(sst.bf_cap * 2, ca.max_passes);
        }
        sst.next_cert = sst.sweeps + sst.gap;
    }
    __syncthreads();
}

// The persistent many-games kernel (internal linkage: each including unit keeps a copy).
// Each workgroup claims games in `order` until they run out or its `wg_budget` edges for
// this launch are spent; a game gets at most `max_sweeps` complete sweeps per launch.
static __global__ __launch_bounds__(1024) void k_games(const GameDesc *__restrict__ games, const int *__restrict__ order,
                                                int ngames, int *counter, int lds_cap, int64_t wg_budget,
                                                int64_t max_sweeps, CertArgs ca) {
    extern __shared__ __align__(16) int lds_val[];
    __shared__ int s_idx;
    __shared__ GameState s_st;
    int64_t left = wg_budget;
    while (left > 0) {
        if (threadIdx.x == 0) s_idx = atomicAdd(counter, 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= ngames) break;          // every workgroup leaves once the games run out
        const GameDesc b = games[order[idx]];
        // the first lds_cap values in LDS, the rest (if any) in the game's global array
        const GameVals V{lds_val, b.values, min(b.n, lds_cap)};
        // proof potentials (4 bytes each) in LDS after the values when the game is small enough
        const int d_at = (b.n + 3) & ~3;
        int *dp = d_at + b.n <= lds_cap ? lds_val + d_at : b.h;
        if (threadIdx.x == 0) {
            s_st = *b.state;
            if (!s_st.started) {
                s_st.started = 1;
                s_st.gap = ca.every;
                s_st.next_cert = ca.every > 0 ? ca.every : INT64_MAX;
                s_st.bf_cap = min(ca.first_cap, ca.max_passes);
                s_st.cursor = 0;
                s_st.dirty = 0;
                s_idx = -1;                // marks a fresh game for the value set-up below
            }
        }
        __syncthreads();
        const bool fresh = s_idx < 0;
        if (fresh) {
            for (int v = threadIdx.x; v < b.n; v += blockDim.x) {
                V.set(v, 0);
                b.rise[v] = 0;
            }
        } else {
            for (int v = threadIdx.x; v < V.split; v += blockDim.x) lds_val[v] = b.values[v];
        }
        __syncthreads();
        int64_t sweeps_left = max_sweeps;
        for (;;) {
            int64_t spent = 0, sweeps = 0;
            bool due;
            switch (b.group) {
            case 1: due = game_run<1>(b, V, s_st, left, sweeps_left, ca, spent, sweeps); break;
            case 2: due = game_run<2>(b, V, s_st, left, sweeps_left, ca, spent, sweeps); break;
            case 4: due = game_run<4>(b, V, s_st, left, sweeps_left, ca, spent, sweeps); break;
            case 8: due = game_run<8>(b, V, s_st, left, sweeps_left, ca, spent, sweeps); break;
            case 16: due = game_run<16>(b, V, s_st, left, sweeps_left, ca, spent, sweeps); break;
            case 32: due = game_run<32>(b, V, s_st, left, sweeps_left, ca, spent, sweeps); break;
            default: due = game_run<64>(b, V, s_st, left, sweeps_left, ca, spent, sweeps); break;
            }
            left -= spent;
            sweeps_left -= sweeps;
            if (!due) break;
            game_prove(b, V, dp, s_st, ca);
            if (left <= 0 || sweeps_left <= 0) break;
        }
        for (int v = threadIdx.x; v < V.split; v += blockDim.x) b.values[v] = lds_val[v];
        if (threadIdx.x == 0) *b.state = s_st;
        __syncthreads();
    }
}

}  // namespace egs
