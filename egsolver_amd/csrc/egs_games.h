// egs_games.h -- the many-games kernel shared by the batched solver (egs_batch.hip) and
// the serving loop (egs_serve.hip).
//
// A launch runs over a list of game descriptors.  A persistent workgroup claims one whole
// game at a time from a device counter and runs in-place sweeps over its vertices until a
// sweep changes nothing (a block-wide vote) or the launch's sweep budget for that game is
// spent; the game's progress lives in its GameState, so a later launch resumes it.
// Values sit in LDS when the game fits the workgroup's LDS share, else in global memory.
// Every `every` sweeps (doubling while proofs fail) the workgroup tries the certified TOP
// proof of egs_accel.hip on its game.
#pragma once

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
};

struct GameDesc {
    const int *ptr;        // n+1 local CSR offsets
    const void *edges;     // m edges: uint32 (w << shift) | succ, or int2 {succ, w} when shift == 0
    int *values;           // n
    int *sigma;            // n scratch (proofs)
    long long *h;          // n scratch (proofs)
    GameState *state;
    int64_t m;
    int n, n0, mg, wmax;
    int shift, group;
};

struct CertArgs {
    int every;             // sweeps between proof attempts; <= 0 = no proofs
    int max_passes;        // longest-path passes per attempt
};

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

// One sweep by groups of G lanes; block-uniform `base` loop keeps waves whole for DPP.
template <int G>
__device__ __forceinline__ bool game_sweep(const GameDesc &b, int *val, int64_t &edges_done, int &lifted) {
    const int gl = threadIdx.x % G;
    const int gid = threadIdx.x / G;
    const int ngroups = blockDim.x / G;
    bool changed = false;
    for (int base = 0; base < b.n; base += ngroups) {
        const int v = base + gid;
        const bool valid = v < b.n;
        int s = 0, e = 0, old = 0;
        if (valid) {
            s = b.ptr[v];
            e = b.ptr[v + 1];
            old = val[v];
        }
        const bool owner1 = v >= b.n0;
        int key = EGS_TOP;
        for (int k = s + gl; k < e; k += G) {
            int succ, w;
            game_edge(b, k, succ, w);
            key = min(key, egs_key(egs_ominus(val[succ], w, b.mg), owner1));
        }
        key = egs_group_min<G>(key);
        const int t = egs_unkey(key, owner1);
        if (valid && gl == 0) {
            edges_done += e - s;
            if (old < t) {
                val[v] = t;
                changed = true;
                lifted++;
            }
        }
    }
    return changed;
}

__device__ __forceinline__ bool game_in_d(int x, int B) { return x > B && x != EGS_TOP; }

// The certified TOP proof inside the workgroup; returns the vertices set to TOP.
__device__ int game_certify(const GameDesc &b, int *val, int max_passes, int64_t &passes_done) {
    __shared__ int win[64], flo[64];
    __shared__ int s_B, s_marked;
    const int wmax = b.wmax;
    if (wmax <= 0) return 0;
    for (int i = threadIdx.x; i < 64; i += blockDim.x) { win[i] = 0; flo[i] = 0; }
    __syncthreads();
    for (int v = threadIdx.x; v < b.n; v += blockDim.x) {
        const int x = val[v];
        if (x == EGS_TOP || x <= 0) continue;
        const int64_t q = ((int64_t)x + wmax - 1) / wmax;
        if ((q & (q - 1)) == 0) atomicAdd(&win[63 - __clzll((unsigned long long)q)], 1);
        if (x > wmax) atomicAdd(&flo[63 - __clzll((unsigned long long)(((int64_t)x - 1) / wmax))], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_B = -1;
        int64_t above[40];       // above[j] = #{non-TOP values > W * 2^j}
        int64_t acc = 0;
        for (int k = 63; k >= 40; k--) acc += flo[k];
        for (int j = 39; j >= 0; j--) { acc += flo[j]; above[j] = acc; }
        for (int j = 1; j < 40; j++) {
            const int64_t B = (int64_t)wmax << j;
            if (B > (int64_t)b.mg - wmax) break;
            if (win[j] == 0 && above[j] > 0) { s_B = (int)B; break; }
        }
        s_marked = 0;
    }
    __syncthreads();
    const int B = s_B;
    if (B < 0) return 0;
    bool bad = false;
    for (int v = threadIdx.x; v < b.n; v += blockDim.x) {
        b.h[v] = 0;
        if (!game_in_d(val[v], B)) continue;
        const int s = b.ptr[v], e = b.ptr[v + 1];
        if (v >= b.n0) {
            int best = -1, arg = s;
            for (int k = s; k < e; k++) {
                int u, w;
                game_edge(b, k, u, w);
                const int c = egs_ominus(val[u], w, b.mg);
                if (c > best) { best = c; arg = k; }
            }
            b.sigma[v] = arg;
            int u, w;
            game_edge(b, arg, u, w);
            if (val[u] <= B) bad = true;
        } else {
            for (int k = s; k < e; k++) {
                int u, w;
                game_edge(b, k, u, w);
                if (val[u] <= B) bad = true;
            }
        }
    }
    if (__syncthreads_or(bad)) return 0;
    const long long scale = (long long)b.n + 1;
    const long long bound = (long long)b.n * (scale * wmax + 1);
    bool converged = false;
    for (int p = 0; p < max_passes; p++) {
        bool ch = false, over = false;
        for (int v = threadIdx.x; v < b.n; v += blockDim.x) {
            if (!game_in_d(val[v], B)) continue;
            long long best = 0;
            const int s = (v >= b.n0) ? b.sigma[v] : b.ptr[v];
            const int e = (v >= b.n0) ? b.sigma[v] + 1 : b.ptr[v + 1];
            for (int k = s; k < e; k++) {
                int u, w;
                game_edge(b, k, u, w);
                if (game_in_d(val[u], B)) {
                    const long long c = scale * w + 1 + b.h[u];
                    if (c > best) best = c;
                }
            }
            if (best > b.h[v]) {
                b.h[v] = best;
                ch = true;
                over |= best > bound;
            }
        }
        passes_done++;
        const int raised = __syncthreads_count(ch);
        if (__syncthreads_or(over)) break;
        if (raised == 0) { converged = true; break; }
    }
    if (!converged) return 0;
    int mine = 0;
    for (int v = threadIdx.x; v < b.n; v += blockDim.x)
        if (game_in_d(val[v], B)) { val[v] = EGS_TOP; mine++; }
    if (mine) atomicAdd(&s_marked, mine);
    __syncthreads();
    return s_marked;
}

// Advance one game by at most `budget` sweeps (the state carries it across launches).
template <int G>
__device__ void game_run(const GameDesc &b, int *val, int64_t budget, const CertArgs &ca) {
    __shared__ GameState s_st;
    if (threadIdx.x == 0) s_st = *b.state;
    __syncthreads();
    GameState st = s_st;
    if (!st.started) {
        for (int v = threadIdx.x; v < b.n; v += blockDim.x) val[v] = 0;
        st.started = 1;
        st.gap = ca.every;
        st.next_cert = ca.every > 0 ? ca.every : INT64_MAX;
    }
    __syncthreads();
    int64_t edges_done = 0, passes = 0;
    int lifted = 0;
    int64_t done_here = 0;
    bool go = !st.converged;
    while (go && done_here < budget) {
        const bool ch = game_sweep<G>(b, val, edges_done, lifted);
        go = __syncthreads_or(ch) != 0;
        st.sweeps++;
        done_here++;
        if (go && ca.every > 0 && st.sweeps >= st.next_cert) {
            const int marked = game_certify(b, val, ca.max_passes, passes);
            if (marked > 0) {
                st.certificates++;
                st.gap = ca.every;
            } else {
                st.gap = min(st.gap * 2, 1 << 16);
            }
            st.next_cert = st.sweeps + st.gap;
            __syncthreads();
        }
    }
    __shared__ unsigned long long red[2];
    if (threadIdx.x == 0) { red[0] = 0; red[1] = 0; }
    __syncthreads();
    if (edges_done) atomicAdd(&red[0], (unsigned long long)edges_done);
    if (lifted) atomicAdd(&red[1], (unsigned long long)lifted);
    __syncthreads();
    if (threadIdx.x == 0) {
        st.edges += (int64_t)red[0];
        st.lifted += (int64_t)red[1];
        st.cert_passes += passes;
        st.converged = go ? 0 : 1;
        *b.state = st;
    }
}

// The persistent many-games kernel (internal linkage: each including unit keeps a copy).
static __global__ __launch_bounds__(1024) void k_games(const GameDesc *__restrict__ games, const int *__restrict__ order,
                                                int ngames, int *counter, int lds_cap, int64_t budget,
                                                CertArgs ca) {
    extern __shared__ int lds_val[];
    __shared__ int s_idx;
    for (;;) {
        if (threadIdx.x == 0) s_idx = atomicAdd(counter, 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= ngames) break;          // every workgroup leaves once the games run out
        const GameDesc b = games[order[idx]];
        const bool in_lds = b.n <= lds_cap;
        int *val = in_lds ? lds_val : b.values;
        if (in_lds) {
            // resumed games bring their values in; new ones are zeroed by game_run
            for (int v = threadIdx.x; v < b.n; v += blockDim.x) lds_val[v] = b.values[v];
            __syncthreads();
        }
        switch (b.group) {
        case 1: game_run<1>(b, val, budget, ca); break;
        case 2: game_run<2>(b, val, budget, ca); break;
        case 4: game_run<4>(b, val, budget, ca); break;
        case 8: game_run<8>(b, val, budget, ca); break;
        case 16: game_run<16>(b, val, budget, ca); break;
        case 32: game_run<32>(b, val, budget, ca); break;
        default: game_run<64>(b, val, budget, ca); break;
        }
        if (in_lds) {
            for (int v = threadIdx.x; v < b.n; v += blockDim.x) b.values[v] = lds_val[v];
        }
        __syncthreads();
    }
}

}  // namespace egs
