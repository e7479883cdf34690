// egs_accel.hip -- certified early detection of TOP vertices (exact).
//
// Value iteration needs about MG / (climb per round) rounds to push player-1-winning
// vertices past MG (the reference's own run: MG 1992 -> 2011 sweeps, script/sol.txt:31-32).
// This pass proves, during the run, that a set D of climbing vertices is won by player 1,
// and sets them to TOP at once.  The proof:
//   1. Candidate D.  Either a gap threshold B (no vertex value in (B - W, B], W = max |w|,
//      some non-TOP value above B, B <= MG - W) with D = {v : B < f(v) < TOP}, or the
//      vertices that rose in the last round.  The candidate only decides what is tried.
//   2. Closure: every successor of an owner-0 vertex of D lies in D or is TOP; some
//      successor of an owner-1 vertex of D does.  Vertices that break it are removed until
//      it holds (the greatest closed subset of the candidate), during the passes of step 3.
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
// A lane's count added to its bin once per wave: the lanes whose bins match (ballots over the
// bin's six bits) add through their lowest lane, so the few bins the values crowd into do not
// serialise 64 LDS atomics per wave.  bin < 0: nothing to add.
__device__ __forceinline__ void hist_add_wave(int *bins, int bin) {
    unsigned long long eq = __builtin_amdgcn_read_exec();
#pragma unroll
    for (int b = 0; b < 7; b++) {
        const bool bit = ((unsigned)bin >> b) & 1u;    // bit 6: the sign of "nothing"
        const unsigned long long m = __ballot(bit);
        eq &= bit ? m : ~m;
    }
    const int lane = threadIdx.x & 63;
    if (bin >= 0 && (eq & ((1ull << lane) - 1)) == 0) atomicAdd(&bins[bin], __popcll(eq));
}

constexpr int kHistU = 8;   // values in flight per thread

__global__ __launch_bounds__(256) void k_gap_hist(DevGame g, int wmax, int *__restrict__ out) {
    {   // the next round's input counters (round ring[14]): an empty frontier ends the run at
        // this pause before any candidate is read, so no histogram is needed.  Every wave
        // reads the same counters and leaves together.
        const int *c = g.lcnt + (g.ring[14] % 3) * kListCounters;
        const int lane = threadIdx.x & 63;
        const bool any = c[lane] != 0 || (lane == 0 && c[kSub] != 0);
        if (__ballot(any) == 0) return;
    }
    __shared__ int win[kBins], flo[kBins];
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) { win[i] = 0; flo[i] = 0; }
    __syncthreads();
    const int64_t threads = (int64_t)gridDim.x * blockDim.x;
    const unsigned w = (unsigned)wmax;
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v0 < g.n; v0 += kHistU * threads) {
        int x[kHistU];
#pragma unroll
        for (int j = 0; j < kHistU; j++) x[j] = v0 + j * threads < g.n ? g.val[v0 + j * threads] : 0;
#pragma unroll
        for (int j = 0; j < kHistU; j++) {
            int bw = -1, bf = -1;
            if (x[j] != EGS_TOP && x[j] > 0) {
                // k = the smallest k >= 0 with W 2^k >= x (no division: a leading-zero estimate
                // corrected by comparisons).  ceil(x / W) = 2^k exactly when x > (2^k - 1) W, and
                // for x > W the largest k' with W 2^k' < x is k - 1.
                const int64_t xx = x[j], ww = w;
                int k = (32 - __clz((unsigned)x[j] - 1u)) - (32 - __clz(w));
                k = max(k, 0);
                while (k > 0 && (ww << (k - 1)) >= xx) k--;
                while ((ww << k) < xx) k++;
                if (xx > ((1ll << k) - 1) * ww) bw = k;
                if (xx > ww) bf = k - 1;
            }
            hist_add_wave(win, bw);
            hist_add_wave(flo, bf);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) {
        if (win[i]) atomicAdd(&out[i], win[i]);
        if (flo[i]) atomicAdd(&out[kBins + i], flo[i]);
    }
}

// D membership of a vertex whose value x is already loaded: a gap threshold (B >= 0) or
// the vertices that rose in round `last` (B < 0).
__device__ __forceinline__ bool in_d_x(const DevGame &g, int v, int x, int B, int last) {
    return x != EGS_TOP && (B >= 0 ? x > B : g.rise[v] == last);
}

__device__ __forceinline__ int wave_sum(int x) {
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Workgroup sum of x added to *dst with one atomic per workgroup (one per wave on a single
// address serialises in L2 when thousands of waves finish together).  Every thread calls.
__device__ __forceinline__ void block_add(int x, int *dst) {
    __shared__ int part[16];
    x = wave_sum(x);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();
}

// Potentials carry D membership for the passes: h >= 0 on D, and two sentinels elsewhere,
// so a pass gathers one 8-byte word per edge.
constexpr long long kHOut = LLONG_MIN;        // not in D, not TOP
constexpr long long kHTop = LLONG_MIN + 1;    // TOP: a successor that ends the path

// Changed-vertex bitmaps have two levels: ceil(n/32) words with a bit per vertex, then a
// summary with a bit per word (set when the word may be non-zero).  A test reads the
// summary first: it is small enough to stay in the L1 caches, and when few vertices
// changed almost every test ends there.
__device__ __forceinline__ int64_t bm_words(int n) { return ((int64_t)n + 31) >> 5; }
__device__ __forceinline__ int64_t bm_total(int n) { return bm_words(n) + ((bm_words(n) + 31) >> 5); }
__device__ __forceinline__ void bm_set(unsigned *bm, int n, int v) {
    atomicOr(&bm[v >> 5], 1u << (v & 31));
    atomicOr(&bm[bm_words(n) + (v >> 10)], 1u << ((v >> 5) & 31));
}
__device__ __forceinline__ bool bm_test(const unsigned *bm, int n, int u) {
    if (!((bm[bm_words(n) + (u >> 10)] >> ((u >> 5) & 31)) & 1u)) return false;
    return (bm[u >> 5] >> (u & 31)) & 1u;
}

// Pass flags live in a ring of kCntSlots ints: pass p sets slot p % kCntSlots when it changes
// a potential and zeroes slot (p + 1) % kCntSlots for the pass after it.  Only zero or not
// is ever read, so a workgroup stores 1 instead of adding its count: thousands of
// workgroups adding to one address as they finish together queue on it (a 2,048-workgroup
// launch spent about 20 us on such adds).
__device__ __forceinline__ void block_flag(bool any, int *dst) {
    if (__syncthreads_or(any) && threadIdx.x == 0) *dst = 1;
}
constexpr int kCntSlots = 64;

constexpr int kBigU = 4;   // edges in flight per lane in the workgroup-per-vertex kernels
constexpr int kClosureU = 4;   // edges in flight per thread in k_closure (one thread per vertex)
constexpr int kBfU = 2;        // edges in flight per lane in k_bf_pass

// Closure check, member count and the first pass of the longest-path game, in one sweep
// over the values.  A member that breaks closure (an owner-0 edge leaving D and not to TOP,
// or an owner-1 vertex with no edge into D or TOP) is removed from D; the removal is a
// change like a raise, so its predecessors are checked again in the next pass, and the
// passes together compute the greatest closed subset of the candidate.  The first pass needs
// no potentials: the warm start (n+1) (MG - f) of a successor follows from its value, so it
// is computed here from f (a Jacobi pass from the warm start).  Writes h for every vertex
// (the first pass's result on D, else a sentinel), the changed vertices' bits into `bits`
// (every word of the vertex range is stored; the summary must be zero on entry), zeroes
// `clear` (the bitmap of the next pass),
// sets cnt[0] when anything changed, counts members into flags[3] and removals into flags[0].
__global__ __launch_bounds__(256) void k_closure(DevGame g, int B, int last, long long bound,
                                                 long long *__restrict__ h, unsigned *__restrict__ bits,
                                                 unsigned *__restrict__ clear, int *__restrict__ cnt,
                                                 int *__restrict__ flags) {
    const long long scale = (long long)g.n + 1;
    const int lane = threadIdx.x & 63;
    int members = 0, changed = 0, removed = 0;
    // the stride is a multiple of 64, so a wave always covers 64 aligned vertices (two words)
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = g.val[v];
        long long hv;
        bool up = false;
        if (x == EGS_TOP) {
            hv = kHTop;
        } else if (!in_d_x(g, (int)v, x, B, last)) {
            hv = kHOut;
        } else if (g.ptr[v + 1] - g.ptr[v] > g.big_deg) {
            members++;                               // decided by k_closure_big
            hv = scale * ((long long)g.mg - x);
        } else {
            members++;
            const long long h0 = scale * ((long long)g.mg - x);
            const bool owner1 = v >= g.n0;
            long long best = owner1 ? LLONG_MAX : LLONG_MIN;
            bool any = false, all = true;
            const int ke = g.ptr[v + 1];
            for (int k0 = g.ptr[v]; k0 < ke; k0 += kClosureU) {   // kClosureU gathers in flight
                int2 ew[kClosureU];
                int xu[kClosureU];
#pragma unroll
                for (int j = 0; j < kClosureU; j++) ew[j] = k0 + j < ke ? g.edges[k0 + j] : make_int2(0, 0);
#pragma unroll
                for (int j = 0; j < kClosureU; j++) xu[j] = g.val[ew[j].x];
#pragma unroll
                for (int j = 0; j < kClosureU; j++) {
                    if (k0 + j >= ke) break;
                    if (xu[j] == EGS_TOP) {              // ends the path: -infinity for the game
                        any = true;
                        if (owner1) best = LLONG_MIN;
                    } else if (in_d_x(g, ew[j].x, xu[j], B, last)) {
                        any = true;
                        const long long c = scale * ew[j].y + 1 + scale * ((long long)g.mg - xu[j]);
                        best = owner1 ? min(best, c) : max(best, c);
                    } else {
                        all = false;
                    }
                }
            }
            if (owner1 ? !any : !all) {
                hv = kHOut;
                up = true;
                removed++;
            } else if (best > h0) {
                hv = best;
                up = true;
                if (best > bound) atomicOr(&flags[1], 1);
            } else {
                hv = h0;
            }
        }
        h[v] = hv;
        changed += up;
        const unsigned long long bal = __ballot(up);
        if (lane == 0 || lane == 32) {
            const int64_t word = v >> 5;
            const unsigned bw = (unsigned)(bal >> lane);
            bits[word] = bw;
            if (bw) atomicOr(&bits[bm_words(g.n) + (word >> 5)], 1u << (word & 31));
            clear[word] = 0u;
        }
    }
    for (int64_t i = bm_words(g.n) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bm_total(g.n);
         i += (int64_t)gridDim.x * blockDim.x)
        clear[i] = 0u;
    block_add(members, &flags[3]);
    block_add(removed, &flags[0]);
    block_flag(changed > 0, &cnt[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[1] = 0;
}


// k_closure over the phased layout (games whose dense rounds run in successor-range
// phases), in one launch shaped like k_dense_fused: workgroup b takes the tiles b, b + W,
// ... and walks the slices in order, so the value gathers of the moment come from one slice
// that stays in L2.  The fold is the one of k_closure (an OR, an AND and a max or min, so
// the order of the edges does not matter); a member's partial best and its flags (bit 0:
// an edge reaches D or TOP, bit 1: an edge leaves D and does not reach TOP; bit 2: a member
// whose edges are in the slices) stay in LDS for the whole launch.  Membership depends only
// on values and rise stamps, which this kernel does not change.  After the last slice it
// does what k_closure does after its edge loop, for every vertex.
constexpr int kClosurePhaseU = 2;   // edges in flight per thread

__device__ __forceinline__ int wave_excl_scan_a(int x) { return egs_wave_incl_scan(x) - x; }

template <bool PACK>
__global__ __launch_bounds__(256) void k_closure_fused(DevGame g, int B, int last, long long bound,
                                                       long long *__restrict__ h, unsigned *__restrict__ bits,
                                                       unsigned *__restrict__ clear, int *__restrict__ cnt,
                                                       int *__restrict__ flags) {
    extern __shared__ long long lbest[];          // [tile][kSliceTile] partial bests, then the flags
    const long long scale = (long long)g.n + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = g.phases;
    const int64_t tiles = ((int64_t)g.n + kSliceTile - 1) / kSliceTile;
    const int64_t W = gridDim.x;
    const int K = blockIdx.x < tiles ? (int)((tiles - blockIdx.x + W - 1) / W) : 0;
    int *lfl = (int *)(lbest + (size_t)K * kSliceTile);
    for (int k = 0; k < K; k++) {
        const int64_t v = (blockIdx.x + k * W) * kSliceTile + threadIdx.x;
        bool member = false;
        if (v < g.n) member = in_d_x(g, (int)v, g.val[v], B, last) && g.ptr[v + 1] - g.ptr[v] <= g.big_deg;
        lfl[k * kSliceTile + threadIdx.x] = member ? 4 : 0;
        lbest[k * kSliceTile + threadIdx.x] = v >= g.n0 ? LLONG_MAX : LLONG_MIN;
    }
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
            const int i = k * kSliceTile + threadIdx.x;
            int fl = lfl[i];
            // every lane's length counts in the prefix: the segments are laid out for all vertices
            const int len = len_n;
            const int b = base_n + wave_excl_scan_a(len);
            {
                const int kn = k + 1 < K ? k + 1 : 0, sn = k + 1 < K ? s : s + 1;
                if (sn < P) {
                    const size_t seg = (size_t)(blockIdx.x + kn * W) * P + sn;
                    len_n = g.slen[seg * kSliceTile + threadIdx.x];
                    base_n = g.sbase[seg * kSliceWaves + wave];
                }
            }
            const int e = b + len;
            if (!(fl & 4) || len == 0) continue;
            const bool owner1 = T * kSliceTile + threadIdx.x >= g.n0;
            long long best = lbest[i];
            for (int k0 = b; k0 < e; k0 += kClosurePhaseU) {
                int u[kClosurePhaseU], w[kClosurePhaseU], xu[kClosurePhaseU];
#pragma unroll
                for (int j = 0; j < kClosurePhaseU; j++) slice_edge<PACK>(g, k0 + j < e ? k0 + j : k0, u[j], w[j]);
#pragma unroll
                for (int j = 0; j < kClosurePhaseU; j++) xu[j] = k0 + j < e ? g.val[u[j]] : EGS_TOP;
#pragma unroll
                for (int j = 0; j < kClosurePhaseU; j++) {
                    if (k0 + j >= e) break;
                    if (xu[j] == EGS_TOP) {              // ends the path: -infinity for the game
                        fl |= 1;
                        if (owner1) best = LLONG_MIN;
                    } else if (in_d_x(g, u[j], xu[j], B, last)) {
                        fl |= 1;
                        const long long c = scale * w[j] + 1 + scale * ((long long)g.mg - xu[j]);
                        best = owner1 ? min(best, c) : max(best, c);
                    } else {
                        fl |= 2;
                    }
                }
            }
            lbest[i] = best;
            lfl[i] = fl;
        }
    }
    int members = 0, changed = 0, removed = 0;
    for (int k = 0; k < K; k++) {
        const int64_t v = (blockIdx.x + k * W) * kSliceTile + threadIdx.x;
        const int i = k * kSliceTile + threadIdx.x;
        bool up = false;
        if (v < g.n) {
            const int x = g.val[v];
            const bool owner1 = v >= g.n0;
            const int fl = lfl[i];
            long long hv;
            if (x == EGS_TOP) {
                hv = kHTop;
            } else if (!in_d_x(g, (int)v, x, B, last)) {
                hv = kHOut;
            } else if (!(fl & 4)) {
                members++;                               // above big_deg: decided by k_closure_big
                hv = scale * ((long long)g.mg - x);
            } else {
                members++;
                const long long h0 = scale * ((long long)g.mg - x);
                const long long best = lbest[i];
                if (owner1 ? !(fl & 1) : (fl & 2)) {
                    hv = kHOut;
                    up = true;
                    removed++;
                } else if (best > h0) {
                    hv = best;
                    up = true;
                    if (best > bound) atomicOr(&flags[1], 1);
                } else {
                    hv = h0;
                }
            }
            h[v] = hv;
        }
        changed += up;
        // a wave covers 64 aligned vertices (two bitmap words); words wholly past n stay as they are
        const unsigned long long bal = __ballot(up);
        if ((lane == 0 || lane == 32) && v < g.n) {
            const int64_t word = v >> 5;
            const unsigned bw = (unsigned)(bal >> lane);
            bits[word] = bw;
            if (bw) atomicOr(&bits[bm_words(g.n) + (word >> 5)], 1u << (word & 31));
            clear[word] = 0u;
        }
    }
    for (int64_t i = bm_words(g.n) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bm_total(g.n);
         i += (int64_t)gridDim.x * blockDim.x)
        clear[i] = 0u;
    block_add(members, &flags[3]);
    block_add(removed, &flags[0]);
    block_flag(changed > 0, &cnt[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[1] = 0;
}

// k_closure for the vertices above big_deg (launched after it): the same closure test and
// first pass; changes set their bit with an atomic OR, since k_closure has stored every
// word already.  Workgroups [0, big_grid) take the big list, one vertex each.  The vertices
// above huge_deg are shared by up to `per` of the last huge_grid workgroups, each folding
// every per-th chunk of kHugeChunkCl edges into the vertex's accumulators (hflag bit 0: an
// edge reaches D or TOP, bit 1: an edge leaves D and does not reach TOP; hbest: the min or
// max candidate); the workgroup that completes hcnt[v] applies the update and resets them.
// One workgroup walking a 70,000-edge hub alone took 135-265 us.
constexpr int kHugeChunkCl = 2048;

__device__ __forceinline__ void closure_fold(const DevGame &g, int k0, int e, int B, int last, bool owner1,
                                             long long scale, long long &best, bool &any, bool &all) {
    for (int k = k0 + (int)threadIdx.x; k < e; k += kBigU * 256) {
        int2 ew[kBigU];
        int xu[kBigU];
#pragma unroll
        for (int j = 0; j < kBigU; j++) ew[j] = k + 256 * j < e ? g.edges[k + 256 * j] : make_int2(0, 0);
#pragma unroll
        for (int j = 0; j < kBigU; j++) xu[j] = g.val[ew[j].x];
#pragma unroll
        for (int j = 0; j < kBigU; j++) {
            if (k + 256 * j >= e) break;
            if (xu[j] == EGS_TOP) {
                any = true;
                if (owner1) best = LLONG_MIN;
            } else if (in_d_x(g, ew[j].x, xu[j], B, last)) {
                any = true;
                const long long c = scale * ew[j].y + 1 + scale * ((long long)g.mg - xu[j]);
                best = owner1 ? min(best, c) : max(best, c);
            } else {
                all = false;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_closure_big(DevGame g, int B, int last, long long bound,
                                                     long long *__restrict__ h, unsigned *__restrict__ bits,
                                                     int *__restrict__ cnt, int *__restrict__ flags, int big_grid,
                                                     int huge_grid) {
    __shared__ long long sm[4];
    __shared__ long long s_best;
    __shared__ int s_last, s_flag;
    const long long scale = (long long)g.n + 1;
    const bool big_part = (int)blockIdx.x < big_grid;
    const int nb = g.ring[9], nh = g.ring[10];
    const int nblk = big_part ? big_grid : huge_grid;
    const int bid = big_part ? (int)blockIdx.x : (int)blockIdx.x - big_grid;
    const int per = big_part ? 1 : max(1, huge_grid / max(nh, 1));
    const int64_t units = big_part ? (int64_t)nb : (int64_t)nh * per;
    for (int64_t u = bid; u < units; u += nblk) {
        const int i = big_part ? (int)u : (int)(u % nh), part = big_part ? 0 : (int)(u / nh);
        const int v = big_part ? g.sbig[i] : g.shuge[nh - 1 - i];          // the largest hubs first
        const int x = g.val[v];
        if (!in_d_x(g, v, x, B, last)) continue;        // the whole workgroup: v is shared
        const int kb = g.ptr[v], ke = g.ptr[v + 1];
        const int chunks = big_part ? 1 : (ke - kb + kHugeChunkCl - 1) / kHugeChunkCl;
        const int contrib = min(per, max(chunks, 1));
        if (part >= contrib) continue;
        const bool owner1 = v >= g.n0;
        long long best = owner1 ? LLONG_MAX : LLONG_MIN;
        bool any = false, all = true;
        if (big_part) {
            closure_fold(g, kb, ke, B, last, owner1, scale, best, any, all);
        } else {
            for (int c = part; c < chunks; c += contrib) {
                const int cs = kb + c * kHugeChunkCl;
                closure_fold(g, cs, min(ke, cs + kHugeChunkCl), B, last, owner1, scale, best, any, all);
            }
        }
        any = __syncthreads_or(any);
        all = __syncthreads_and(all);
        best = egs_block_minmax_ll(best, owner1, sm);
        if (threadIdx.x == 0) {
            bool done = true;
            int f = (any ? 1 : 0) | (all ? 0 : 2);
            if (contrib > 1) {
                if (f) atomicOr(&g.hflag[v], f);
                if (owner1) atomicMin(&g.hbest[v], best);
                else atomicMax(&g.hbest[v], best);
                __threadfence();
                done = atomicAdd(&g.hcnt[v], 1) == contrib - 1;
                if (done) {
                    __threadfence();
                    f = atomicExch(&g.hflag[v], 0);
                    best = atomicExch((unsigned long long *)&g.hbest[v],
                                      (unsigned long long)(owner1 ? LLONG_MAX : LLONG_MIN));
                    g.hcnt[v] = 0;
                }
            }
            s_last = done;
            s_flag = f;
            s_best = best;
        }
        __syncthreads();
        if (threadIdx.x == 0 && s_last) {
            const long long h0 = scale * ((long long)g.mg - x);
            const long long b2 = s_best;
            if (owner1 ? !(s_flag & 1) : (s_flag & 2)) {
                h[v] = kHOut;
                bm_set(bits, g.n, v);
                cnt[0] = 1;
                atomicAdd(&flags[0], 1);
            } else if (b2 > h0) {
                h[v] = b2;
                bm_set(bits, g.n, v);
                cnt[0] = 1;
                if (b2 > bound) atomicOr(&flags[1], 1);
            }
        }
        __syncthreads();
    }
}

// Pass p >= 1 of the longest-path game on D, in place, G lanes per vertex: owner 0 takes the
// max over its edges into D, owner 1 the min over its edges into D or TOP (a TOP successor
// ends the path); a member that no longer meets closure (its successors left D) is removed
// instead.  Only vertices with a successor changed (raised or removed) in pass p-1, bit set
// in `prev`, are evaluated: a vertex none of whose successors changed since its last
// evaluation keeps its potential and its closure, since every change in pass q is seen by
// the evaluations of pass q+1 (a kernel boundary lies between).  So a pass that changes
// nothing leaves a closed D and a fixpoint on it.  Changed vertices set their bit in `cur`;
// `clear` (the bitmap of pass p+1, last read in pass p-1) is zeroed.  The pass returns at
// once when pass p-1 changed nothing or the potentials have passed the divergence bound.
template <int G>
__device__ __forceinline__ void bf_pass_part(const DevGame &g, long long *__restrict__ h, long long bound,
                                             const unsigned *__restrict__ prev, unsigned *__restrict__ cur,
                                             unsigned *__restrict__ clear, int p, int *__restrict__ cnt,
                                             int *__restrict__ flags, int bid, int nblk) {
    if (cnt[(p - 1) % kCntSlots] == 0 || flags[1]) return;
    const int64_t words = bm_total(g.n);
    for (int64_t i = (int64_t)bid * blockDim.x + threadIdx.x; i < words; i += (int64_t)nblk * blockDim.x)
        clear[i] = 0u;
    if (bid == 0 && threadIdx.x == 0) cnt[(p + 1) % kCntSlots] = 0;
    constexpr int kPer = 256 / G;
    constexpr unsigned long long kGroupMask = (G == 64) ? ~0ull : ((1ull << G) - 1);
    const int gl = threadIdx.x % G;
    const int gbase = (threadIdx.x & 63) & ~(G - 1);
    const long long scale = (long long)g.n + 1;
    int local = 0, removed = 0;
    const int64_t stride = (int64_t)nblk * kPer;
    // the next vertex's potential and row bounds load while this one is evaluated (only its
    // own group writes its potential in this pass)
    long long h_n = kHOut;
    int kb_n = 0, ke_n = 0;
    auto fetch = [&](int64_t b) {
        const int64_t u = b + threadIdx.x / G;
        if (u < g.n) {
            h_n = h[u];
            kb_n = g.ptr[u];
            ke_n = g.ptr[u + 1];
        }
    };
    if ((int64_t)bid * kPer < g.n) fetch((int64_t)bid * kPer);
    for (int64_t base = (int64_t)bid * kPer; base < g.n; base += stride) {
        const int v = (int)base + (int)threadIdx.x / G;
        const long long hv = v < g.n ? h_n : kHOut;
        const bool member = hv >= 0;
        const bool owner1 = v >= g.n0;
        int kb = 0, ke = 0;
        const int kb_c = kb_n, ke_c = ke_n;
        if (base + stride < g.n) fetch(base + stride);
        bool need = false;
        // the lane's first kBfU edges, read once for the test and the evaluation
        int2 ew0[kBfU];
#pragma unroll
        for (int j = 0; j < kBfU; j++) ew0[j] = make_int2(0, 0);
        if (member) {
            kb = kb_c;
            ke = ke_c;
            if (ke - kb > g.big_deg) ke = kb;         // k_bf_big evaluates it: no need here
            // kBfU edges per lane at a time (an edge past ke re-tests edge k0, a valid index)
            const int k0 = kb + gl;
            if (k0 < ke) {
#pragma unroll
                for (int j = 0; j < kBfU; j++) ew0[j] = g.edges[k0 + j * G < ke ? k0 + j * G : k0];
#pragma unroll
                for (int j = 0; j < kBfU; j++) need |= bm_test(prev, g.n, ew0[j].x);
            }
            for (int k1 = k0 + G * kBfU; k1 < ke && !need; k1 += G * kBfU) {
                int u[kBfU];
#pragma unroll
                for (int j = 0; j < kBfU; j++) u[j] = g.edges[k1 + j * G < ke ? k1 + j * G : k1].x;
#pragma unroll
                for (int j = 0; j < kBfU; j++) need |= bm_test(prev, g.n, u[j]);
            }
        }
        need = ((__ballot(need) >> gbase) & kGroupMask) != 0;
        long long best = owner1 ? LLONG_MAX : LLONG_MIN;
        bool out = false;                    // an edge leaves D and does not reach TOP
        if (member && need && kb + gl < ke) {
            const int k0 = kb + gl;
            long long hu[kBfU];
#pragma unroll
            for (int j = 0; j < kBfU; j++) hu[j] = k0 + j * G < ke ? h[ew0[j].x] : kHTop;
#pragma unroll
            for (int j = 0; j < kBfU; j++) {
                if (k0 + j * G >= ke) break;
                if (hu[j] == kHOut) { out = true; continue; }
                const long long c = hu[j] == kHTop ? LLONG_MIN : scale * ew0[j].y + 1 + hu[j];
                best = owner1 ? min(best, c) : max(best, c);
            }
        }
        if (member && need) {
            for (int k0 = kb + gl + G * kBfU; k0 < ke; k0 += G * kBfU) {   // loads, then gathers, in flight together
                int2 ew[kBfU];
                long long hu[kBfU];
#pragma unroll
                for (int j = 0; j < kBfU; j++) ew[j] = g.edges[k0 + j * G < ke ? k0 + j * G : k0];
#pragma unroll
                for (int j = 0; j < kBfU; j++) hu[j] = k0 + j * G < ke ? h[ew[j].x] : kHTop;
#pragma unroll
                for (int j = 0; j < kBfU; j++) {
                    if (k0 + j * G >= ke) break;
                    if (hu[j] == kHOut) { out = true; continue; }
                    const long long c = hu[j] == kHTop ? LLONG_MIN : scale * ew[j].y + 1 + hu[j];
                    best = owner1 ? min(best, c) : max(best, c);
                }
            }
        }
        out = ((__ballot(out) >> gbase) & kGroupMask) != 0;
#pragma unroll
        for (int d = G / 2; d >= 1; d >>= 1) {
            const long long o = __shfl_xor(best, d, 64);
            best = owner1 ? min(best, o) : max(best, o);
        }
        if (member && need && gl == 0) {
            // owner 1 breaks closure when no edge reaches D or TOP, owner 0 when any leaves
            if (owner1 ? best == LLONG_MAX : out) {
                h[v] = kHOut;
                bm_set(cur, g.n, v);
                local++;
                removed++;
            } else if (best > hv) {
                h[v] = best;
                bm_set(cur, g.n, v);
                local++;
                if (best > bound) atomicOr(&flags[1], 1);
            }
        }
    }
    block_flag(local > 0, &cnt[p % kCntSlots]);
    block_add(removed, &flags[0]);
}

template <int G>
__global__ __launch_bounds__(256) void k_bf_pass(DevGame g, long long *__restrict__ h, long long bound,
                                                 const unsigned *__restrict__ prev, unsigned *__restrict__ cur,
                                                 unsigned *__restrict__ clear, int p, int *__restrict__ cnt,
                                                 int *__restrict__ flags) {
    bf_pass_part<G>(g, h, bound, prev, cur, clear, p, cnt, flags, blockIdx.x, gridDim.x);
}

// Pass p for the vertices of the big list (big_deg < deg <= huge_deg; launched after
// k_bf_pass for the same p), one workgroup each, with the same rule and updates.
__device__ __forceinline__ void bf_big_part(const DevGame &g, long long *__restrict__ h, long long bound,
                                            const unsigned *__restrict__ prev, unsigned *__restrict__ cur, int p,
                                            int *__restrict__ cnt, int *__restrict__ flags, int bid, int nblk) {
    __shared__ long long sm[4];
    if (cnt[(p - 1) % kCntSlots] == 0 || flags[1]) return;
    const long long scale = (long long)g.n + 1;
    const int nb = g.ring[9], items = nb;               // the huge list: k_bf_huge
    int local = 0;
    for (int64_t i = bid; i < items; i += nblk) {
        const int v = egs_big_item(g.sbig, g.shuge, nb, items - 1 - i);   // huge vertices first
        const long long hv = h[v];
        if (hv < 0) continue;                           // not in D: the whole workgroup
        const int kb = g.ptr[v], ke = g.ptr[v + 1];
        bool need = false;
        for (int k0 = kb + (int)threadIdx.x; k0 < ke && !need; k0 += kBigU * 256) {
            unsigned wd[kBigU];
            int u[kBigU];
#pragma unroll
            for (int j = 0; j < kBigU; j++) u[j] = k0 + 256 * j < ke ? g.edges[k0 + 256 * j].x : 0;
#pragma unroll
            for (int j = 0; j < kBigU; j++) wd[j] = prev[bm_words(g.n) + (u[j] >> 10)];
#pragma unroll
            for (int j = 0; j < kBigU; j++)
                need |= k0 + 256 * j < ke && ((wd[j] >> ((u[j] >> 5) & 31)) & 1u) && ((prev[u[j] >> 5] >> (u[j] & 31)) & 1u);
        }
        if (!__syncthreads_or(need)) continue;
        const bool owner1 = v >= g.n0;
        long long best = owner1 ? LLONG_MAX : LLONG_MIN;
        bool out = false;
        for (int k0 = kb + (int)threadIdx.x; k0 < ke; k0 += kBigU * 256) {
            int2 ew[kBigU];
            long long hu[kBigU];
#pragma unroll
            for (int j = 0; j < kBigU; j++) ew[j] = k0 + 256 * j < ke ? g.edges[k0 + 256 * j] : make_int2(0, 0);
#pragma unroll
            for (int j = 0; j < kBigU; j++) hu[j] = h[ew[j].x];
#pragma unroll
            for (int j = 0; j < kBigU; j++) {
                if (k0 + 256 * j >= ke) break;
                if (hu[j] == kHOut) { out = true; continue; }
                const long long c = hu[j] == kHTop ? LLONG_MIN : scale * ew[j].y + 1 + hu[j];
                best = owner1 ? min(best, c) : max(best, c);
            }
        }
        out = __syncthreads_or(out);
        best = egs_block_minmax_ll(best, owner1, sm);
        if (threadIdx.x == 0) {
            if (owner1 ? best == LLONG_MAX : out) {
                h[v] = kHOut;
                bm_set(cur, g.n, v);
                local++;
                atomicAdd(&flags[0], 1);
            } else if (best > hv) {
                h[v] = best;
                bm_set(cur, g.n, v);
                local++;
                if (best > bound) atomicOr(&flags[1], 1);
            }
        }
    }
    if (threadIdx.x == 0 && local) cnt[p % kCntSlots] = 1;
}

__global__ __launch_bounds__(256) void k_bf_big(DevGame g, long long *__restrict__ h, long long bound,
                                                const unsigned *__restrict__ prev, unsigned *__restrict__ cur,
                                                int p, int *__restrict__ cnt, int *__restrict__ flags) {
    bf_big_part(g, h, bound, prev, cur, p, cnt, flags, blockIdx.x, gridDim.x);
}

// Pass p for the vertices above huge_deg, each shared by up to `per` workgroups: a
// workgroup takes every per-th chunk of kHugeChunkBf edges and folds into the vertex's
// accumulators whether a successor changed, whether an edge leaves D, and the min (owner 1)
// or max (owner 0) candidate; the workgroup that completes hcnt[v] applies the update as
// k_bf_pass does (only when a successor changed) and resets the accumulators.  All
// contributors read h[v] before any of them counts in, so they agree on membership.
constexpr int kHugeChunkBf = 2048;

__device__ __forceinline__ void bf_huge_part(const DevGame &g, long long *__restrict__ h, long long bound,
                                             const unsigned *__restrict__ prev, unsigned *__restrict__ cur, int p,
                                             int *__restrict__ cnt, int *__restrict__ flags, int bid, int nblk) {
    __shared__ long long sm[4];
    __shared__ long long s_best;
    __shared__ int s_last, s_flag;
    if (cnt[(p - 1) % kCntSlots] == 0 || flags[1]) return;
    const int nh = g.ring[10];
    if (nh == 0) return;
    const long long scale = (long long)g.n + 1;
    const int per = max(1, nblk / nh);
    const int64_t units = (int64_t)nh * per;
    int local = 0, removed = 0;
    for (int64_t u = bid; u < units; u += nblk) {
        const int i = (int)(u % nh), part = (int)(u / nh);
        const int v = g.shuge[nh - 1 - i];                 // largest first
        const long long hv = h[v];
        if (hv < 0) continue;                               // the whole workgroup
        const int kb = g.ptr[v], ke = g.ptr[v + 1];
        const int chunks = (ke - kb + kHugeChunkBf - 1) / kHugeChunkBf;
        const int contrib = min(per, chunks);
        if (part >= contrib) continue;
        const bool owner1 = v >= g.n0;
        long long best = owner1 ? LLONG_MAX : LLONG_MIN;
        bool need = false, out = false;
        for (int c = part; c < chunks; c += contrib) {
            const int cs = kb + c * kHugeChunkBf, ce = min(ke, cs + kHugeChunkBf);
            for (int k0 = cs + (int)threadIdx.x; k0 < ce; k0 += kBigU * 256) {
                int2 ew[kBigU];
                long long hu[kBigU];
#pragma unroll
                for (int j = 0; j < kBigU; j++) ew[j] = k0 + 256 * j < ce ? g.edges[k0 + 256 * j] : make_int2(0, 0);
#pragma unroll
                for (int j = 0; j < kBigU; j++) hu[j] = h[ew[j].x];
#pragma unroll
                for (int j = 0; j < kBigU; j++) {
                    if (k0 + 256 * j >= ce) break;
                    need |= bm_test(prev, g.n, ew[j].x);
                    if (hu[j] == kHOut) { out = true; continue; }
                    const long long c2 = hu[j] == kHTop ? LLONG_MIN : scale * ew[j].y + 1 + hu[j];
                    best = owner1 ? min(best, c2) : max(best, c2);
                }
            }
        }
        need = __syncthreads_or(need);
        out = __syncthreads_or(out);
        best = egs_block_minmax_ll(best, owner1, sm);
        if (threadIdx.x == 0) {
            bool last = true;
            int f = (need ? 1 : 0) | (out ? 2 : 0);
            if (contrib > 1) {
                if (f) atomicOr(&g.hflag[v], f);
                if (owner1) atomicMin(&g.hbest[v], best);
                else atomicMax(&g.hbest[v], best);
                __threadfence();
                last = atomicAdd(&g.hcnt[v], 1) == contrib - 1;
                if (last) {
                    __threadfence();
                    f = atomicExch(&g.hflag[v], 0);
                    best = atomicExch((unsigned long long *)&g.hbest[v],
                                      (unsigned long long)(owner1 ? LLONG_MAX : LLONG_MIN));
                    g.hcnt[v] = 0;
                }
            }
            s_last = last;
            s_flag = f;
            s_best = best;
        }
        __syncthreads();
        if (threadIdx.x == 0 && s_last && (s_flag & 1)) {
            const long long b = s_best;
            if (owner1 ? b == LLONG_MAX : (s_flag & 2) != 0) {
                h[v] = kHOut;
                bm_set(cur, g.n, v);
                local++;
                removed++;
            } else if (b > hv) {
                h[v] = b;
                bm_set(cur, g.n, v);
                local++;
                if (b > bound) atomicOr(&flags[1], 1);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && local) cnt[p % kCntSlots] = 1;
    if (threadIdx.x == 0 && removed) atomicAdd(&flags[0], removed);
}

__global__ __launch_bounds__(256) void k_bf_huge(DevGame g, long long *__restrict__ h, long long bound,
                                                 const unsigned *__restrict__ prev, unsigned *__restrict__ cur,
                                                 int p, int *__restrict__ cnt, int *__restrict__ flags) {
    bf_huge_part(g, h, bound, prev, cur, p, cnt, flags, blockIdx.x, gridDim.x);
}

// One pass over every member in one launch: workgroups [0, grid) run the pass over the
// vertices up to big_deg, the next big_grid the big list and the last huge_grid the huge
// list.  The three touch disjoint potentials, add to the same pass counters and write
// their changes to the same bitmap, so they need no order among themselves: each part
// strides over its own workgroups only.
template <int G>
__global__ __launch_bounds__(256) void k_bf_all(DevGame g, long long *__restrict__ h, long long bound,
                                                const unsigned *__restrict__ prev, unsigned *__restrict__ cur,
                                                unsigned *__restrict__ clear, int p, int *__restrict__ cnt,
                                                int *__restrict__ flags, int grid, int big_grid, int huge_grid) {
    const int b = blockIdx.x;
    if (b < grid) bf_pass_part<G>(g, h, bound, prev, cur, clear, p, cnt, flags, b, grid);
    else if (b < grid + big_grid) bf_big_part(g, h, bound, prev, cur, p, cnt, flags, b - grid, big_grid);
    else bf_huge_part(g, h, bound, prev, cur, p, cnt, flags, b - grid - big_grid, huge_grid);
}

// Starting state of the vertices above huge_deg, in one pass: the shared lifts' partial keys
// and chunk counts (TOP and 0), and the accumulators of k_bf_huge (the identity of each
// owner's fold) with their flags.
__global__ __launch_bounds__(256) void k_hbest_reset(DevGame g) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        g.hacc[v] = EGS_TOP;
        g.hcnt[v] = 0;
        g.hbest[v] = v >= g.n0 ? LLONG_MAX : LLONG_MIN;
        g.hflag[v] = 0;
    }
}

hipError_t launch_hbest_reset(const DevGame &g, hipStream_t st) {
    hipLaunchKernelGGL(k_hbest_reset, dim3(grid_for(g.n, 256, 2048)), dim3(256), 0, st, g);
    return hipGetLastError();
}

// D is where h >= 0 (set by k_closure for this candidate).  The values were raised outside
// the rounds, so the next round (ring[14]) sweeps every vertex: ring[11] names it.
__global__ __launch_bounds__(256) void k_mark_top(DevGame g, const long long *__restrict__ h) {
    if (blockIdx.x == 0 && threadIdx.x == 0) g.ring[11] = g.ring[14];
    int local = 0;
    const int64_t threads = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v0 < g.n; v0 += 4 * threads) {
        long long hv[4];                         // four potentials in flight per thread
#pragma unroll
        for (int k = 0; k < 4; k++) hv[k] = v0 + k * threads < g.n ? h[v0 + k * threads] : -1;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (hv[k] >= 0) { g.val[v0 + k * threads] = EGS_TOP; local++; }
    }
    // one store of the TOP flag per workgroup that marked a vertex (nothing counts the marks:
    // the members left after the passes are the count, and a per-workgroup atomic on one
    // counter serialises when the workgroups finish together)
    if (__syncthreads_or(local > 0) && threadIdx.x == 0) g.ring[15] = 1;
}

static void launch_bf_all(int G, int grid, int big_grid, int huge_grid, hipStream_t st, const DevGame &g,
                          long long *h, long long bound, const unsigned *prev, unsigned *cur, unsigned *clear, int p,
                          int *cnt, int *flags) {
    const dim3 all(grid + big_grid + huge_grid), t(256);
    switch (G) {
    case 1: hipLaunchKernelGGL(k_bf_all<1>, all, t, 0, st, g, h, bound, prev, cur, clear, p, cnt, flags, grid, big_grid, huge_grid); break;
    case 2: hipLaunchKernelGGL(k_bf_all<2>, all, t, 0, st, g, h, bound, prev, cur, clear, p, cnt, flags, grid, big_grid, huge_grid); break;
    case 4: hipLaunchKernelGGL(k_bf_all<4>, all, t, 0, st, g, h, bound, prev, cur, clear, p, cnt, flags, grid, big_grid, huge_grid); break;
    case 8: hipLaunchKernelGGL(k_bf_all<8>, all, t, 0, st, g, h, bound, prev, cur, clear, p, cnt, flags, grid, big_grid, huge_grid); break;
    default: hipLaunchKernelGGL(k_bf_all<16>, all, t, 0, st, g, h, bound, prev, cur, clear, p, cnt, flags, grid, big_grid, huge_grid); break;
    }
}

static void launch_bf_pass(int G, int grid, hipStream_t st, const DevGame &g, long long *h, long long bound,
                           const unsigned *prev, unsigned *cur, unsigned *clear, int p, int *cnt, int *flags) {
    switch (G) {
    case 1: hipLaunchKernelGGL(k_bf_pass<1>, dim3(grid), dim3(256), 0, st, g, h, bound, prev, cur, clear, p, cnt, flags); break;
    case 2: hipLaunchKernelGGL(k_bf_pass<2>, dim3(grid), dim3(256), 0, st, g, h, bound, prev, cur, clear, p, cnt, flags); break;
    case 4: hipLaunchKernelGGL(k_bf_pass<4>, dim3(grid), dim3(256), 0, st, g, h, bound, prev, cur, clear, p, cnt, flags); break;
    case 8: hipLaunchKernelGGL(k_bf_pass<8>, dim3(grid), dim3(256), 0, st, g, h, bound, prev, cur, clear, p, cnt, flags); break;
    default: hipLaunchKernelGGL(k_bf_pass<16>, dim3(grid), dim3(256), 0, st, g, h, bound, prev, cur, clear, p, cnt, flags); break;
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

// Admissible thresholds (host side): B_j = W * 2^j with some non-TOP value above it and
// B_j <= MG - W; first those with an empty window, smallest first, then the smallest whose
// window holds at most max(64, 1/256 of the values above).  Returns how many were written.
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
    // Then the smallest threshold whose window holds only a few stragglers: the members
    // whose edges reach them break closure and are pruned, and the rest can still be proved.
    for (int j = 1; j < 40 && cnt < max_out; j++) {
        const int64_t B = (int64_t)wmax << j;
        if (B > (int64_t)mg - wmax) break;
        if (win[j] > 0 && suffix[j] > 0 && win[j] <= std::max<int64_t>(64, suffix[j] / 256)) {
            out[cnt++] = B;
            break;
        }
    }
    return cnt;
}

// One certification attempt on a paused run after round `last`.  Candidates, in turn:
// every admissible gap threshold (smallest first, at most max_cands), then the vertices
// that rose in round `last`.  Each is pruned to its greatest closed subset while at most
// `cap` passes of the longest-path game run on it; the first that converges with a
// non-empty set is set to TOP (queued, not waited for).  has_big: some vertex is above
// big_deg (its own kernels).  hist_ready: certify_prefetch was queued and has completed.  Returns 1 when vertices were
// set to TOP (count in res.marked), 0 when no candidate gave a proof, < 0 on a HIP error.
// A candidate's starting state in one launch: its flags, the pass counters and the upper
// level of the first pass bitmap (the lower level is stored whole by the closure check).
__global__ __launch_bounds__(256) void k_cert_clear(int *__restrict__ flags, int *__restrict__ cnt,
                                                    unsigned *__restrict__ words, int64_t nwords) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4) flags[i] = 0;
    if (i < kCntSlots) cnt[i] = 0;
    for (int64_t k = i; k < nwords; k += (int64_t)gridDim.x * blockDim.x) words[k] = 0u;
}

hipError_t certify_prefetch(const DevGame &g, const CertBuffers &cb, int wmax, hipStream_t st) {
    hipError_t err = hipMemsetAsync(cb.hist, 0, sizeof(int) * (2 * kBins + 8), st);
    if (err != hipSuccess) return err;
    // each workgroup adds its non-empty bins to the global histogram with atomics (about 60
    // addresses); 1024 workgroups keep four waves per SIMD for the per-value arithmetic
    hipLaunchKernelGGL(k_gap_hist, dim3(std::min(grid_for(g.n, 256 * kHistU, 2048), 1024)), dim3(256), 0, st, g,
                       std::max(wmax, 1), cb.hist);
    return hipMemcpyAsync(cb.pinned, cb.hist, sizeof(int) * 2 * kBins, hipMemcpyDeviceToHost, st);
}

int certify_top(const DevGame &g, const CertBuffers &cb, int wmax, int last, int cap, int max_cands, bool has_big,
                bool hist_ready, const volatile uint32_t *stop, int pass_hint, hipStream_t st, CertResult &res) {
    constexpr int kBigGrid = 512;           // workgroups of the kernels for vertices above big_deg
    res = CertResult{};
    if (wmax <= 0 || !g.rise) return 0;
    const long long scale = (long long)g.n + 1;
    // divergence bound: at a fixpoint h <= max start + (n - 1) max w'; keep 2x headroom
    if ((double)scale * g.mg + (double)g.n * ((double)scale * wmax + 1) > 4.0e18) return 0;
    const long long bound = scale * g.mg + (long long)g.n * (scale * wmax + 1);
    const int grid = grid_for(g.n, 256, 2048);
    if (!hist_ready) {
        if (certify_prefetch(g, cb, wmax, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
    }
    int64_t cands[41];
    int ncand = gap_candidates(cb.pinned, cb.pinned + kBins, wmax, g.mg, cands, std::min(max_cands, 40));
    cands[ncand++] = -1;                    // the vertices that rose in round `last`
    int *flags = cb.hist + 2 * kBins;       // [0] removed, [1] bound exceeded, [2] unused, [3] members
    int *fl = cb.pinned + 2 * kBins;        // flags readback (4 ints), then the counter ring
    int *cnt = fl + 4;
    const int64_t words = ((int64_t)g.n + 31) >> 5, bm_len = words + (words + 31) / 32;
    unsigned *bm[3] = {cb.bits, cb.bits + bm_len, cb.bits + 2 * bm_len};
    // Passes are queued in chunks and read back once per chunk; a pass after the converged
    // one returns at once, so a chunk costs little beyond its real passes.  The first
    // chunk (closure and first pass, then more) is sized for a few passes, or for what an
    // earlier attempt of the run needed plus two, so that an attempt usually ends after one
    // read-back; later chunks are longer.
    const int first_chunk = std::max(6, std::min(pass_hint + 2, 32)), chunk = 8;
    // lanes per vertex: about two edges each (on the 2^22-vertex game, with two edges in
    // flight per lane, the solve took 4.41 ms with 4 lanes, 4.43 with 2 and 4.68 with 8)
    int G = 1;
    while (G < 16 && 2 * G < (double)g.m / g.n) G <<= 1;
    const int bf_grid = grid_for((int64_t)g.n * G, 256, 4096);
    for (int c = 0; c < ncand; c++) {
        if (stop && *stop) return 0;        // the run is being stopped: no new candidate
        const int B = (int)cands[c];
        res.threshold = B;
        res.tried++;
        hipLaunchKernelGGL(k_cert_clear, dim3(grid_for(std::max<int64_t>(bm_len - words, kCntSlots), 256, 64)),
                           dim3(256), 0, st, flags, cb.bfcnt, bm[0] + words, bm_len - words);
        if (g.phases > 1) {
            const size_t lds = (size_t)g.phase_tiles * kSliceTile * (sizeof(long long) + sizeof(int));
            if (g.sshift)
                hipLaunchKernelGGL(k_closure_fused<true>, dim3(g.phase_grid), dim3(256), lds, st, g, B, last, bound,
                                   cb.h, bm[0], bm[1], cb.bfcnt, flags);
            else
                hipLaunchKernelGGL(k_closure_fused<false>, dim3(g.phase_grid), dim3(256), lds, st, g, B, last, bound,
                                   cb.h, bm[0], bm[1], cb.bfcnt, flags);
        } else {
            hipLaunchKernelGGL(k_closure, dim3(grid), dim3(256), 0, st, g, B, last, bound, cb.h, bm[0], bm[1], cb.bfcnt,
                               flags);
        }
        if (has_big)
            hipLaunchKernelGGL(k_closure_big, dim3(3 * kBigGrid), dim3(256), 0, st, g, B, last, bound, cb.h, bm[0],
                               cb.bfcnt, flags, kBigGrid, 2 * kBigGrid);
        int queued = 1;                     // passes queued so far: pass 0 is the closure's
        int passes = 0;
        bool converged = false, diverged = false, rejected = false;
        while (!converged && !diverged && !rejected && passes < cap) {
            if (stop && *stop) return 0;    // abandon the candidate: nothing was marked
            const int upto = std::min(cap, queued == 1 ? first_chunk : queued + chunk);
            for (int p = queued; p < upto; p++) {
                // with vertices above big_deg, their parts of the pass run in the same launch
                if (has_big)
                    launch_bf_all(G, bf_grid, kBigGrid, 2 * kBigGrid, st, g, cb.h, bound, bm[(p - 1) % 3], bm[p % 3],
                                  bm[(p + 1) % 3], p, cb.bfcnt, flags);
                else
                    launch_bf_pass(G, bf_grid, st, g, cb.h, bound, bm[(p - 1) % 3], bm[p % 3], bm[(p + 1) % 3], p,
                                   cb.bfcnt, flags);
            }
            if (hipMemcpyAsync(fl, flags, sizeof(int) * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
            if (hipMemcpyAsync(cnt, cb.bfcnt, sizeof(int) * kCntSlots, hipMemcpyDeviceToHost, st) != hipSuccess)
                return -1;
            if (wait_queued(cb.ev, st) != hipSuccess) return -1;
            if (fl[3] == 0) { rejected = true; break; }
            if (fl[1]) diverged = true;
            for (int p = queued - (queued > 1 ? 0 : 1); p < upto && !diverged; p++) {
                passes = p + 1;
                if (cnt[p % kCntSlots] == 0) { converged = true; res.last_passes = passes; break; }
            }
            queued = upto;
        }
        if (rejected) continue;
        if (fl[0]) res.closure_failed++;
        res.passes += passes;
        if (!converged) {
            if (!diverged) res.cap_hit = 1;
            continue;
        }
        res.marked = fl[3] - fl[0];         // members left after the removals
        if (res.marked <= 0) continue;      // pruning emptied the candidate: try the next
        hipLaunchKernelGGL(k_mark_top, dim3(grid), dim3(256), 0, st, g, (const long long *)cb.h);
        return hipGetLastError() == hipSuccess ? 1 : -1;
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// Certificate of the whole TOP set of the current values
// ---------------------------------------------------------------------------------------
//
// The proofs above mark sets of vertices TOP during a run; ordinary lifting marks others.
// After a run, the longest-path game of step 3 is run once more with D = T, the set of all
// TOP vertices, from h = 0, with no TOP sentinel (every TOP vertex is a member, so every
// edge into T is constrained and none ends a path).  A fixpoint with nothing removed gives
// potentials under which every owner-0 edge inside T and some edge into T of every owner-1
// vertex of T satisfies h(v) >= (n+1) w + 1 + h(u), and no owner-0 vertex of T has an edge
// leaving T: a strategy of player 1 that keeps every play from T inside T with every cycle
// of w-weight <= -1, so T is won by player 1.  A host-side check of exactly these
// inequalities proves T is contained in the least fixpoint's TOP set without trusting any of
// this code; a fixpoint check of the values gives the converse.

// Seeds pass 0: h = 0 on T and kHOut elsewhere, pass 1 evaluates every member (all bits of
// `prev`), the bitmap pass 1 writes starts empty, cnt[0] = members (so pass 1 runs) and
// flags[3] counts the members.
__global__ __launch_bounds__(256) void k_cert_all_seed(DevGame g, long long *__restrict__ h,
                                                       unsigned *__restrict__ prev, unsigned *__restrict__ cur,
                                                       int *__restrict__ cnt, int *__restrict__ flags) {
    int members = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (int64_t)gridDim.x * blockDim.x) {
        const bool top = g.val[v] == EGS_TOP;
        h[v] = top ? 0 : kHOut;
        members += top;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bm_total(g.n);
         i += (int64_t)gridDim.x * blockDim.x) {
        prev[i] = ~0u;
        cur[i] = 0u;
    }
    block_add(members, &flags[3]);
    block_flag(members > 0, &cnt[0]);
}

// Potentials for export: -1 outside the certified set.
__global__ __launch_bounds__(256) void k_cert_export(int n, long long *__restrict__ h) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        if (h[v] < 0) h[v] = -1;
}

int certify_all_top(const DevGame &g, const CertBuffers &cb, int wmax, int max_passes, bool has_big,
                    hipStream_t st, CertResult &res) {
    constexpr int kBigGrid = 512;
    res = CertResult{};
    const long long scale = (long long)g.n + 1;
    // at a fixpoint from h = 0, h <= (n - 1) max w' along a path; twice that is divergence
    const double wp = (double)scale * std::max(wmax, 0) + 1;
    if (2.0 * g.n * wp > 4.0e18) return -2;
    const long long bound = 2LL * g.n * ((long long)scale * std::max(wmax, 0) + 1);
    const int grid = grid_for(g.n, 256, 2048);
    int *flags = cb.hist + 2 * kBins;
    int *fl = cb.pinned + 2 * kBins;
    int *cnt = fl + 4;
    const int64_t words = ((int64_t)g.n + 31) >> 5, bm_len = words + (words + 31) / 32;
    unsigned *bm[3] = {cb.bits, cb.bits + bm_len, cb.bits + 2 * bm_len};
    int G = 1;
    while (G < 16 && 2 * G < (double)g.m / g.n) G <<= 1;
    const int bf_grid = grid_for((int64_t)g.n * G, 256, 4096);
    if (hipMemsetAsync(flags, 0, sizeof(int) * 4, st) != hipSuccess) return -1;
    if (hipMemsetAsync(cb.bfcnt, 0, sizeof(int) * kCntSlots, st) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_cert_all_seed, dim3(grid), dim3(256), 0, st, g, cb.h, bm[0], bm[1], cb.bfcnt, flags);
    int queued = 1, chunk = 8;
    bool converged = false, diverged = false;
    while (!converged && !diverged && queued <= max_passes) {
        const int upto = std::min(max_passes + 1, queued + chunk);
        for (int p = queued; p < upto; p++) {
            launch_bf_pass(G, bf_grid, st, g, cb.h, bound, bm[(p - 1) % 3], bm[p % 3], bm[(p + 1) % 3], p, cb.bfcnt,
                           flags);
            if (has_big) {
                hipLaunchKernelGGL(k_bf_big, dim3(kBigGrid), dim3(256), 0, st, g, cb.h, bound, bm[(p - 1) % 3],
                                   bm[p % 3], p, cb.bfcnt, flags);
                hipLaunchKernelGGL(k_bf_huge, dim3(2 * kBigGrid), dim3(256), 0, st, g, cb.h, bound, bm[(p - 1) % 3],
                                   bm[p % 3], p, cb.bfcnt, flags);
            }
        }
        if (hipMemcpyAsync(fl, flags, sizeof(int) * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipMemcpyAsync(cnt, cb.bfcnt, sizeof(int) * kCntSlots, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        if (fl[1]) diverged = true;
        for (int p = queued; p < upto && !diverged; p++) {
            res.passes = p;
            if (cnt[p % kCntSlots] == 0) { converged = true; break; }
        }
        queued = upto;
        // a chunk must stay below the counter ring: pass p writes slot p and zeroes slot p + 1
        chunk = std::min(chunk * 2, kCntSlots / 2);
    }
    res.marked = fl[3];                     // TOP vertices
    res.closure_failed = fl[0];             // TOP vertices that left the certified set
    res.cap_hit = !converged && !diverged;
    hipLaunchKernelGGL(k_cert_export, dim3(grid), dim3(256), 0, st, g.n, cb.h);
    if (hipGetLastError() != hipSuccess) return -1;
    return converged && !diverged && fl[0] == 0 ? 1 : 0;
}

}  // namespace egs
