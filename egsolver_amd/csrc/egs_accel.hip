// egs_accel.hip -- certified early detection of TOP vertices (exact).
//
// Value iteration needs about MG / (climb per round) rounds to push player-1-winning
// vertices past MG (the reference's own run: MG 1992 -> 2011 sweeps, script/sol.txt:31-32).
// This pass proves, during the run, that a set D of climbing vertices is won by player 1,
// and sets them to TOP at once.  The proof:
//   1. Candidate D.  Either a gap threshold B (no vertex value in (B - W, B], W = max |w|,
//      some non-TOP value above B, B <= MG - W) with D = {v : B < f(v) < TOP}, or the
//      vertices that rose in the last round.  The candidate only decides what is tried.
//   2. Closure, checked: every successor of an owner-0 vertex of D lies in D or is TOP;
//      some successor of an owner-1 vertex of D does.
//   3. The longest-path game on D with w' = (n+1) w + 1:
//        h(v) <- max(h(v), max_{u in D} w'(v,u) + h(u))                 (owner 0)
//        h(v) <- max(h(v), min_{u in D or TOP} w'(v,u) + h(u))          (owner 1)
//      with a TOP successor ending the path.  If it reaches a fixpoint, the minimising
//      successors give a player-1 strategy sigma under which every cycle inside D has
//      h(v) >= w'(v,u) + h(u) on each edge, so its w'-weight is <= 0 and its w-weight
//      (n+1) W_c + L <= 0 with L <= n gives W_c <= -1.  Any fixpoint proves it, so the
//      iteration starts from h = (n+1) (MG - f): the lifting keeps f(u) - f(v) >= w(v,u)
//      along the owner-0 edges and the best owner-1 edges inside D, which leaves only
//      chains of tight edges to resolve.  A value above (n+1) MG + n max w' proves that
//      no fixpoint exists (divergence), and a pass cap bounds the attempt.
//   4. Then every play from D under sigma either reaches a TOP vertex or loses energy
//      without bound, so lfp = TOP on D; raising f to TOP there keeps f <= lfp, and the
//      ordinary rounds continue to the same least fixpoint.
#include <algorithm>
#include <climits>

#include "egs_device.h"
#include "egs_internal.h"

namespace egs {

constexpr int kBins = 64;

// Window and floor histograms of the values against the candidates B_j = W * 2^j:
// win[j]  counts f in (B_j - W, B_j]    (the value with ceil(f/W) == 2^j)
// flo[k]  counts f with B_k < f <= B_{k+1}, so #{f > B_j, f < TOP} = sum_{k >= j} flo[k]
__global__ __launch_bounds__(256) void k_gap_hist(DevGame g, int wmax, int *__restrict__ out) {
    __shared__ int win[kBins], flo[kBins];
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) { win[i] = 0; flo[i] = 0; }
    __syncthreads();
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = g.val[v];
        if (x == EGS_TOP || x <= 0) continue;
        const int64_t q = ((int64_t)x + wmax - 1) / wmax;          // ceil(x / W) >= 1
        if ((q & (q - 1)) == 0) atomicAdd(&win[63 - __clzll((unsigned long long)q)], 1);
        if (x > wmax) {
            const int64_t r = ((int64_t)x - 1) / wmax;             // >= 1: largest k with W 2^k < x
            atomicAdd(&flo[63 - __clzll((unsigned long long)r)], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) {
        if (win[i]) atomicAdd(&out[i], win[i]);
        if (flo[i]) atomicAdd(&out[kBins + i], flo[i]);
    }
}

// D membership: a gap threshold (B >= 0) or the vertices that rose in round `last` (B < 0).
__device__ __forceinline__ bool in_d(const DevGame &g, int v, int B, int last) {
    const int x = g.val[v];
    return x != EGS_TOP && (B >= 0 ? x > B : g.rise[v] == last);
}

__device__ __forceinline__ int wave_sum(int x) {
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Potentials carry D membership for the passes: h >= 0 on D, and two sentinels elsewhere,
// so a pass gathers one 8-byte word per edge.
constexpr long long kHOut = LLONG_MIN;        // not in D, not TOP
constexpr long long kHTop = LLONG_MIN + 1;    // TOP: a successor that ends the path

// Closure check (owner-0 members: every successor in D or TOP; owner-1: at least one),
// member count, and h for every vertex: the warm start (n+1) (MG - f) on D, else a sentinel.
__global__ __launch_bounds__(256) void k_closure(DevGame g, int B, int last, long long *__restrict__ h,
                                                 int *__restrict__ flags) {
    const long long scale = (long long)g.n + 1;
    int members = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = g.val[v];
        if (x == EGS_TOP) { h[v] = kHTop; continue; }
        if (!in_d(g, (int)v, B, last)) { h[v] = kHOut; continue; }
        members++;
        h[v] = scale * ((long long)g.mg - x);
        bool any = false, all = true;
        for (int k = g.ptr[v]; k < g.ptr[v + 1]; k++) {
            const int u = g.edges[k].x;
            const bool ok = g.val[u] == EGS_TOP || in_d(g, u, B, last);
            any |= ok;
            all &= ok;
        }
        if (v >= g.n0 ? !any : !all) atomicOr(&flags[0], 1);
    }
    members = wave_sum(members);
    if ((threadIdx.x & 63) == 0 && members) atomicAdd(&flags[3], members);
}

// One in-place pass of the longest-path game on D, G lanes per vertex: owner 0 takes the
// max over its edges into D, owner 1 the min over its edges into D or TOP (a TOP successor
// ends the path); counts raised entries.
template <int G>
__global__ __launch_bounds__(256) void k_bf_pass(DevGame g, long long *__restrict__ h, long long bound,
                                                 int *__restrict__ changed, int *__restrict__ flags) {
    constexpr int kPer = 256 / G;
    const int gl = threadIdx.x % G;
    const long long scale = (long long)g.n + 1;
    int local = 0;
    for (int64_t base = (int64_t)blockIdx.x * kPer; base < g.n; base += (int64_t)gridDim.x * kPer) {
        const int v = (int)base + (int)threadIdx.x / G;
        const long long hv = v < g.n ? h[v] : kHOut;
        const bool member = hv >= 0;
        const bool owner1 = v >= g.n0;
        long long best = owner1 ? LLONG_MAX : LLONG_MIN;
        if (member) {
            const int e = g.ptr[v + 1];
            for (int k = g.ptr[v] + gl; k < e; k += G) {
                const int2 ew = g.edges[k];
                const long long hu = h[ew.x];
                if (hu == kHOut) continue;
                const long long c = hu == kHTop ? LLONG_MIN : scale * ew.y + 1 + hu;
                best = owner1 ? min(best, c) : max(best, c);
            }
        }
#pragma unroll
        for (int d = G / 2; d >= 1; d >>= 1) {
            const long long o = __shfl_xor(best, d, 64);
            best = owner1 ? min(best, o) : max(best, o);
        }
        if (member && gl == 0 && best != LLONG_MAX && best > hv) {
            h[v] = best;
            local++;
            if (best > bound) atomicOr(&flags[1], 1);
        }
    }
    local = wave_sum(local);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(changed, local);
}

// D is where h >= 0 (set by k_closure for this candidate).
__global__ __launch_bounds__(256) void k_mark_top(DevGame g, const long long *__restrict__ h, int *__restrict__ count) {
    int local = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        if (h[v] >= 0) { g.val[v] = EGS_TOP; local++; }
    }
    local = wave_sum(local);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(count, local);
}

static void launch_bf_pass(int G, int grid, hipStream_t st, const DevGame &g, long long *h, long long bound,
                           int *changed, int *flags) {
    switch (G) {
    case 1: hipLaunchKernelGGL(k_bf_pass<1>, dim3(grid), dim3(256), 0, st, g, h, bound, changed, flags); break;
    case 2: hipLaunchKernelGGL(k_bf_pass<2>, dim3(grid), dim3(256), 0, st, g, h, bound, changed, flags); break;
    case 4: hipLaunchKernelGGL(k_bf_pass<4>, dim3(grid), dim3(256), 0, st, g, h, bound, changed, flags); break;
    case 8: hipLaunchKernelGGL(k_bf_pass<8>, dim3(grid), dim3(256), 0, st, g, h, bound, changed, flags); break;
    default: hipLaunchKernelGGL(k_bf_pass<16>, dim3(grid), dim3(256), 0, st, g, h, bound, changed, flags); break;
    }
}

__global__ void k_abs_wmax(const int2 *__restrict__ edges, int64_t m, int *__restrict__ out) {
    int mx = 0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const int w = edges[e].y;
        mx = max(mx, w < 0 ? -w : w);
    }
    for (int d = 32; d >= 1; d >>= 1) mx = max(mx, __shfl_xor(mx, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, mx);
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------

hipError_t launch_abs_wmax(const DevGame &g, int *out, hipStream_t st) {
    hipError_t err = hipMemsetAsync(out, 0, sizeof(int), st);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_abs_wmax, dim3(grid_for(g.m, 256, 2048)), dim3(256), 0, st, g.edges, g.m, out);
    return hipGetLastError();
}

// Admissible thresholds, smallest first (host side): B_j = W * 2^j with an empty window,
// some non-TOP value above it and B_j <= MG - W.  Returns how many were written.
int gap_candidates(const int *win, const int *flo, int wmax, int mg, int64_t *out, int max_out) {
    int64_t suffix[kBins + 1];
    suffix[kBins] = 0;
    for (int k = kBins - 1; k >= 0; k--) suffix[k] = suffix[k + 1] + flo[k];
    int cnt = 0;
    for (int j = 1; j < 40 && cnt < max_out; j++) {
        const int64_t B = (int64_t)wmax << j;
        if (B > (int64_t)mg - wmax) break;
        if (win[j] == 0 && suffix[j] > 0) out[cnt++] = B;
    }
    return cnt;
}

// One certification attempt on a paused run after round `last`.  Candidates, in turn:
// every admissible gap threshold (smallest first, at most max_cands), then the vertices
// that rose in round `last`.  Each gets the closure check and at most `cap` passes of the
// longest-path game; the first that converges is set to TOP.  Returns 1 when vertices were
// set to TOP (count in res.marked), 0 when no candidate gave a proof, < 0 on a HIP error.
int certify_top(const DevGame &g, const CertBuffers &cb, int wmax, int last, int cap, int max_cands,
                hipStream_t st, CertResult &res) {
    res = CertResult{};
    if (wmax <= 0 || !g.rise) return 0;
    const long long scale = (long long)g.n + 1;
    // divergence bound: at a fixpoint h <= max start + (n - 1) max w'; keep 2x headroom
    if ((double)scale * g.mg + (double)g.n * ((double)scale * wmax + 1) > 4.0e18) return 0;
    const long long bound = scale * g.mg + (long long)g.n * (scale * wmax + 1);
    const int grid = grid_for(g.n, 256, 2048);
    if (hipMemsetAsync(cb.hist, 0, sizeof(int) * (2 * kBins + 8), st) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_gap_hist, dim3(grid), dim3(256), 0, st, g, wmax, cb.hist);
    if (hipMemcpyAsync(cb.pinned, cb.hist, sizeof(int) * 2 * kBins, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    int64_t cands[41];
    int ncand = gap_candidates(cb.pinned, cb.pinned + kBins, wmax, g.mg, cands, std::min(max_cands, 40));
    cands[ncand++] = -1;                    // the vertices that rose in round `last`
    int *flags = cb.hist + 2 * kBins;       // [0] closure failed, [1] bound exceeded, [2] marked, [3] members
    int *cnt = cb.pinned + 2 * kBins;       // up to 8 pass counters
    int *fl = cnt + 8;                      // flags readback
    // convergence is read back every `chunk` passes: often on large games, whose passes
    // cost far more than the read-back
    const int chunk = g.n >= (1 << 20) ? 2 : 8;
    int G = 1;
    while (G < 16 && G < (double)g.m / g.n) G <<= 1;
    const int bf_grid = grid_for((int64_t)g.n * G, 256, 4096);
    for (int c = 0; c < ncand; c++) {
        const int B = (int)cands[c];
        res.threshold = B;
        res.tried++;
        if (hipMemsetAsync(flags, 0, sizeof(int) * 4, st) != hipSuccess) return -1;
        hipLaunchKernelGGL(k_closure, dim3(grid), dim3(256), 0, st, g, B, last, cb.h, flags);
        if (hipMemcpyAsync(fl, flags, sizeof(int) * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        if (fl[3] == 0) continue;
        if (fl[0]) { res.closure_failed++; continue; }
        int passes = 0;
        bool converged = false, diverged = false;
        while (!converged && !diverged && passes < cap) {
            if (hipMemsetAsync(cb.bfcnt, 0, sizeof(int) * chunk, st) != hipSuccess) return -1;
            for (int k = 0; k < chunk; k++) launch_bf_pass(G, bf_grid, st, g, cb.h, bound, cb.bfcnt + k, flags);
            if (hipMemcpyAsync(cnt, cb.bfcnt, sizeof(int) * chunk, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
            if (hipMemcpyAsync(fl, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
            if (hipStreamSynchronize(st) != hipSuccess) return -1;
            if (fl[1]) diverged = true;
            for (int k = 0; k < chunk && !diverged; k++) {
                passes++;
                if (cnt[k] == 0) { converged = true; break; }
            }
        }
        res.passes += passes;
        if (!converged) {
            if (!diverged) res.cap_hit = 1;
            continue;
        }
        hipLaunchKernelGGL(k_mark_top, dim3(grid), dim3(256), 0, st, g, (const long long *)cb.h, flags + 2);
        if (hipMemcpyAsync(fl + 2, flags + 2, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        res.marked = fl[2];
        return res.marked > 0 ? 1 : 0;
    }
    return 0;
}

}  // namespace egs
