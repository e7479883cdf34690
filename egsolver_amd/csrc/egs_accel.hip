// egs_accel.hip -- certified early detection of TOP vertices (exact).
//
// Value iteration needs about MG / (climb per round) rounds to push player-1-winning
// vertices past MG (the reference's own run: MG 1992 -> 2011 sweeps, script/sol.txt:31-32).
// This pass proves, during the run, that a set D of climbing vertices is won by player 1,
// and sets them to TOP at once.  The proof:
//   1. Threshold B with a gap: no vertex value lies in (B - W, B] (W = max |w|), some
//      non-TOP value lies above B, and B <= MG - W.  D = {v : B < f(v) < TOP}.
//   2. Closure.  Every stored value is a lift of values no larger than the current ones,
//      so an owner-0 v in D has f(u) >= f(v) - W > B - W for every successor u, hence
//      f(u) > B by the gap; an owner-1 v in D keeps its best successor sigma(v), which
//      satisfies the same bound.  The pass checks this instead of assuming it.
//   3. Player 1 plays sigma on D.  The graph H on D \ TOP (owner-0 vertices with all their
//      edges, owner-1 vertices with the sigma edge) has every cycle of weight <= -1 iff
//      no cycle is positive under w' = (n+1) w + 1 (a cycle of length L <= n has weight
//      (n+1) W_c + L).  That holds iff the longest-path iteration
//      h(v) = max(0, max_{v->u in H} w'(v,u) + h(u)), from h = 0, reaches a fixpoint;
//      any value above n * max(w', 0) proves a positive cycle.
//   4. Then every play from D under sigma either reaches a TOP vertex or loses energy
//      without bound, so lfp = TOP on D; raising f to TOP there keeps f <= lfp, and the
//      ordinary rounds continue to the same least fixpoint.
// Steps 2-3 are checked, not assumed, so the choice of B only decides which D is tried:
// an attempt walks the admissible thresholds upwards until one D passes both checks.
#include <algorithm>

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

__device__ __forceinline__ bool in_d(int x, int B) { return x > B && x != EGS_TOP; }

// sigma for owner-1 vertices of D (first best edge), closure check, h = 0.
__global__ __launch_bounds__(256) void k_sigma(DevGame g, int B, int *__restrict__ sigma,
                                               long long *__restrict__ h, int *__restrict__ flags) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        h[v] = 0;
        const int x = g.val[v];
        if (!in_d(x, B)) continue;
        const int s = g.ptr[v], e = g.ptr[v + 1];
        if (v >= g.n0) {
            int best = -1, arg = s;
            for (int k = s; k < e; k++) {
                const int2 ew = g.edges[k];
                const int c = egs_ominus(g.val[ew.x], ew.y, g.mg);
                if (c > best) { best = c; arg = k; }
            }
            sigma[v] = arg;
            if (g.val[g.edges[arg].x] <= B) atomicOr(&flags[0], 1);
        } else {
            for (int k = s; k < e; k++)
                if (g.val[g.edges[k].x] <= B) { atomicOr(&flags[0], 1); break; }
        }
    }
}

// One in-place pass of the longest-path iteration over H; counts raised entries.
__global__ __launch_bounds__(256) void k_bf_pass(DevGame g, int B, const int *__restrict__ sigma,
                                                 long long *__restrict__ h, long long bound,
                                                 int *__restrict__ changed, int *__restrict__ flags) {
    int local = 0;
    const long long scale = (long long)g.n + 1;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        if (!in_d(g.val[v], B)) continue;
        long long best = 0;
        if (v >= g.n0) {
            const int2 ew = g.edges[sigma[v]];
            if (in_d(g.val[ew.x], B)) { const long long c = scale * ew.y + 1 + h[ew.x]; if (c > best) best = c; }
        } else {
            for (int k = g.ptr[v]; k < g.ptr[v + 1]; k++) {
                const int2 ew = g.edges[k];
                if (in_d(g.val[ew.x], B)) { const long long c = scale * ew.y + 1 + h[ew.x]; if (c > best) best = c; }
            }
        }
        if (best > h[v]) {
            h[v] = best;
            local++;
            if (best > bound) atomicOr(&flags[1], 1);
        }
    }
    for (int d = 32; d >= 1; d >>= 1) local += __shfl_xor(local, d, 64);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(changed, local);
}

__global__ __launch_bounds__(256) void k_mark_top(DevGame g, int B, int *__restrict__ count) {
    int local = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        if (in_d(g.val[v], B)) { g.val[v] = EGS_TOP; local++; }
    }
    for (int d = 32; d >= 1; d >>= 1) local += __shfl_xor(local, d, 64);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(count, local);
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

// One certification attempt on a paused run.  Each admissible threshold, smallest first,
// gets the closure check and at most `cap` longest-path passes; a smaller-D candidate is
// tried when a larger D fails (a finite vertex parked above the lowest gap keeps that D
// from ever closing).  Returns 1 when vertices were set to TOP (count in res.marked), 0
// when no candidate gave a proof, < 0 on a HIP error.
int certify_top(const DevGame &g, const CertBuffers &cb, int wmax, int cap, int max_cands, hipStream_t st,
                CertResult &res) {
    res = CertResult{};
    if (wmax <= 0) return 0;
    const int grid = grid_for(g.n, 256, 2048);
    if (hipMemsetAsync(cb.hist, 0, sizeof(int) * (2 * kBins + 8), st) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_gap_hist, dim3(grid), dim3(256), 0, st, g, wmax, cb.hist);
    if (hipMemcpyAsync(cb.pinned, cb.hist, sizeof(int) * 2 * kBins, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    int64_t cands[40];
    const int ncand = gap_candidates(cb.pinned, cb.pinned + kBins, wmax, g.mg, cands, std::min(max_cands, 40));
    int *flags = cb.hist + 2 * kBins;       // [0] closure failed, [1] bound exceeded, [2] marked
    int *cnt = cb.pinned + 2 * kBins;       // kChunk pass counters
    int *fl = cnt + 8;                      // flags readback
    const long long bound = (long long)g.n * ((long long)(g.n + 1) * wmax + 1);
    constexpr int kChunk = 8;
    for (int c = 0; c < ncand; c++) {
        const int B = (int)cands[c];
        res.threshold = B;
        res.tried++;
        if (hipMemsetAsync(flags, 0, sizeof(int) * 3, st) != hipSuccess) return -1;
        hipLaunchKernelGGL(k_sigma, dim3(grid), dim3(256), 0, st, g, B, cb.sigma, cb.h, flags);
        if (hipMemcpyAsync(fl, flags, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        if (fl[0]) { res.closure_failed++; continue; }
        int passes = 0;
        bool converged = false, diverged = false;
        while (!converged && !diverged && passes < cap) {
            if (hipMemsetAsync(cb.bfcnt, 0, sizeof(int) * kChunk, st) != hipSuccess) return -1;
            for (int k = 0; k < kChunk; k++)
                hipLaunchKernelGGL(k_bf_pass, dim3(grid), dim3(256), 0, st, g, B, cb.sigma, cb.h, bound,
                                   cb.bfcnt + k, flags);
            if (hipMemcpyAsync(cnt, cb.bfcnt, sizeof(int) * kChunk, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
            if (hipMemcpyAsync(fl, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
            if (hipStreamSynchronize(st) != hipSuccess) return -1;
            if (fl[1]) diverged = true;
            for (int k = 0; k < kChunk && !diverged; k++) {
                passes++;
                if (cnt[k] == 0) { converged = true; break; }
            }
        }
        res.passes += passes;
        if (!converged) {
            if (!diverged) res.cap_hit = 1;
            continue;
        }
        hipLaunchKernelGGL(k_mark_top, dim3(grid), dim3(256), 0, st, g, B, flags + 2);
        if (hipMemcpyAsync(fl + 2, flags + 2, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        res.marked = fl[2];
        return res.marked > 0 ? 1 : 0;
    }
    return 0;
}

}  // namespace egs
