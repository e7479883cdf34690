// egs_device.h -- device-side building blocks shared by the single-game, batched and
// serving kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define EGS_TOP 2147483647

// OMINUS(a, w) under threshold mg (src/common.h:270-272): (a<TOP && a-w<=mg) ?
// max(0, a-w) : TOP.  The a<TOP test comes first so a-w never overflows.
__device__ __forceinline__ int egs_ominus(int a, int w, int mg) {
    const int d = a - w;
    return (a < EGS_TOP && d <= mg) ? (d > 0 ? d : 0) : EGS_TOP;
}

// Owner 0 takes the min over out-edges, owner 1 the max (src/cpu_solver.c:300-305).
// Both are folded into one min-reduction over a key: owner-1 candidates are negated
// (candidates are >= 0, so -x never overflows).  EGS_TOP is the identity.
__device__ __forceinline__ int egs_key(int x, bool owner1) { return owner1 ? -x : x; }
__device__ __forceinline__ int egs_unkey(int k, bool owner1) { return owner1 ? -k : k; }

// ---- DPP / swizzle reductions ----------------------------------------------------------
// DPP control words (GCN3+/CDNA): quad_perm(1,0,3,2) = 0xB1, quad_perm(2,3,0,1) = 0x4E,
// row_half_mirror = 0x141, row_mirror = 0x140.  After the steps for group size G every
// lane of an aligned group of G lanes holds the group minimum.  Callers keep the whole
// wave active around these calls.
template <int CTRL>
__device__ __forceinline__ int egs_dpp(int x) {
    return __builtin_amdgcn_update_dpp(EGS_TOP, x, CTRL, 0xF, 0xF, false);
}

template <int G>
__device__ __forceinline__ int egs_group_min(int x) {
    if constexpr (G >= 2) x = min(x, egs_dpp<0xB1>(x));
    if constexpr (G >= 4) x = min(x, egs_dpp<0x4E>(x));
    if constexpr (G >= 8) x = min(x, egs_dpp<0x141>(x));
    if constexpr (G >= 16) x = min(x, egs_dpp<0x140>(x));
    if constexpr (G >= 32) {
        // ds_swizzle bit-mask mode: and=0x1F, or=0, xor=0x10 -> lane ^ 16 inside each 32.
        x = min(x, __builtin_amdgcn_ds_swizzle(x, 0x401F));
    }
    if constexpr (G >= 64) {
        x = min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 32));
    }
    return x;
}

__device__ __forceinline__ int egs_wave_sum(int x) {
    // Plain butterfly over the wave through ds_swizzle/readlane; used off the hot path.
    x += __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false);
    x += __builtin_amdgcn_ds_swizzle(x, 0x401F);
    return __builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 32);
}

// Inclusive prefix sum over the 64 lanes of a wave, every lane active: six DPP adds (VALU
// only, no LDS round trips).  Shifts by 1, 2, 4 and 8 lanes within each row of 16 (lanes
// shifted in from below the row read 0), then the rows' totals: lane 15's sum into rows 1
// and 3 (row_bcast:15), lane 31's into rows 2 and 3 (row_bcast:31).  Rows outside the
// row mask add the 0 passed as the old value.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int egs_dpp_src(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWMASK, 0xf, false);
}

__device__ __forceinline__ int egs_wave_incl_scan(int x) {
    x += egs_dpp_src<0x111, 0xf>(x);   // row_shr:1
    x += egs_dpp_src<0x112, 0xf>(x);   // row_shr:2
    x += egs_dpp_src<0x114, 0xf>(x);   // row_shr:4
    x += egs_dpp_src<0x118, 0xf>(x);   // row_shr:8
    x += egs_dpp_src<0x142, 0xa>(x);   // row_bcast:15
    x += egs_dpp_src<0x143, 0xc>(x);   // row_bcast:31
    return x;
}

// Exclusive scan over a workgroup (a multiple of 64 threads): wave scans, wave totals
// combined through lds_warp (one int per wave).  Returns this thread's exclusive prefix;
// total = the sum over the workgroup.  Every thread must call it.
__device__ __forceinline__ int egs_block_exclusive_scan(int x, int *lds_warp, int &total) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int inc = egs_wave_incl_scan(x);
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

__device__ __forceinline__ unsigned egs_lane() { return __lane_id(); }

// Lanes below this one in the wave, as a mask.
__device__ __forceinline__ uint64_t egs_lanemask_lt() {
    const unsigned l = egs_lane();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Wave-aggregated append: every active lane with want != 0 gets a distinct slot of
// list[0 .. *count) through one atomic per wave.  Safe under divergence: only active
// lanes take part.
__device__ __forceinline__ void egs_append(int *list, int *count, bool want, int item) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return;
    const uint64_t active = __builtin_amdgcn_read_exec();
    const int leader = __ffsll((unsigned long long)active) - 1;
    int base = 0;
    if ((int)egs_lane() == leader) base = atomicAdd(count, __popcll(mask));
    base = __builtin_amdgcn_readfirstlane(base);
    if (want) list[base + __popcll(mask & egs_lanemask_lt())] = item;
}

// ---- workgroup-per-vertex helpers (vertices above big_deg) -----------------------------
// The static lists of big (big_deg < deg <= huge_deg) and huge (deg > huge_deg) vertices,
// read as one sequence of nb + nh items.
__device__ __forceinline__ int egs_big_item(const int *sbig, const int *shuge, int nb, int64_t i) {
    return i < nb ? sbig[i] : shuge[i - nb];
}

// Workgroup-wide min (take_min) or max of a 64-bit value, returned to every thread.  `sm`
// holds one slot per wave; every thread of the workgroup must call.
__device__ __forceinline__ long long egs_block_minmax_ll(long long x, bool take_min, long long *sm) {
    for (int d = 32; d >= 1; d >>= 1) {
        const long long o = __shfl_xor(x, d, 64);
        x = take_min ? (o < x ? o : x) : (o > x ? o : x);
    }
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = x;
    __syncthreads();
    long long r = sm[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = take_min ? (sm[w] < r ? sm[w] : r) : (sm[w] > r ? sm[w] : r);
    __syncthreads();
    return r;
}
