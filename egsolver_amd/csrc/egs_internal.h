// egs_internal.h -- structures shared between the HIP translation units of libegs.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace egs {

// One game resident on the device, passed by value to every round kernel.
struct DevGame {
    int n, n0, mg;
    int64_t m;
    const int *ptr;       // n+1 CSR offsets (internal numbering, owner-0 first)
    const int2 *edges;    // m packed {successor, weight}
    const int *cptr;      // n+1 CSC offsets
    const int *cpred;     // m predecessors
    int *val;             // n values
    int *stamp;           // n: last round a vertex was queued for
    int *queue[2];        // frontier ping-pong, n each
    int *big;             // n: vertices deferred to the wavefront kernel this round
    int *huge;            // n: vertices deferred to the workgroup kernel this round
    int *ring;            // 9 counters: [0..2] queue, [3..5] big, [6..8] huge
    unsigned long long *stats;  // per-workgroup slots of {processed, edges, lifted, preds}
    int *rounds_done;     // last round that found a non-empty frontier
};

// Launch shape of one round.
struct LaunchShape {
    int group;       // lanes per vertex in k_lift_groups (1..64)
    int grid;        // workgroups of k_lift_groups
    int big_deg;     // out-degree above which a vertex goes to k_lift_wave
    int huge_deg;    // out-degree above which a vertex goes to k_lift_block
    int use_wave, use_block;
    int wave_grid, block_grid;
};

constexpr int kStatSlots = 8192;   // upper bound on any round kernel's grid

hipError_t launch_pack(const int *succ, const int *w, int2 *edges, int64_t m, hipStream_t st);
hipError_t launch_transpose(const DevGame &g, int *tile_sum, int *fill, hipStream_t st);
hipError_t launch_init(const DevGame &g, hipStream_t st);
hipError_t launch_round(const DevGame &g, const LaunchShape &sh, int r, hipStream_t st);
hipError_t launch_sum_stats(const DevGame &g, int nslots, unsigned long long *out, hipStream_t st);
hipError_t launch_degree_stats(const int *ptr, int n, int *out2, hipStream_t st);
hipError_t launch_violations(const DevGame &g, unsigned long long *out, hipStream_t st);

}  // namespace egs
