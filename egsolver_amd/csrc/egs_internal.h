// egs_internal.h -- structures shared between the HIP translation units of libegs.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace egs {

// Frontier lists are split into kSub sub-queues of sub_cap slots plus one overflow region
// of n slots; a workgroup appends to sub-queue (blockIdx.x % kSub), so no counter sees
// more than a few dozen workgroups' atomics per round.
constexpr int kSub = 64;
constexpr int kListCounters = kSub + 1;     // kSub sub-queue counters + overflow counter
constexpr int kLdsCap = 2048;               // per-workgroup LDS staging of appends
constexpr int kStatSlots = 8192;            // upper bound on any round kernel's grid

// Round modes, decided on the device at the start of every round from the previous
// round's mode and output-list size (see decide_mode in egs_rounds.hip).
enum : int { MODE_QUEUE = 0, MODE_LIFTED = 1, MODE_DENSE = 2, MODE_DONE = 3 };

// One game resident on the device, passed by value to every kernel.
struct DevGame {
    int n, n0, mg;
    int64_t m;
    const int *ptr;       // n+1 CSR offsets (internal numbering, owner-0 first)
    const int2 *edges;    // m packed {successor, weight}
    const int *cptr;      // n+1 CSC offsets
    const int *cpred;     // m predecessors
    int *val;             // n values
    int *stamp;           // n: last round a vertex was claimed for / queued for
    int *list[2];         // output lists, ping-pong: kSub*sub_cap + n slots each
    int sub_cap;
    int *lcnt;            // 3 slots x kListCounters
    int *mode;            // [0..2] mode of round r at r % 3; [3] a round whose input list
                          // holds its count but no items (a phased dense round before the transpose)
    int *dbig, *dhuge;    // per-round deferral lists (sparse modes)
    int *ring;            // [3..5] dbig counters, [6..8] dhuge counters, [9] nsbig, [10] nshuge,
                          // [11] a round that must sweep all vertices (after a TOP proof),
                          // [12] 1 once the transpose exists (sparse rounds allowed),
                          // [13] rounds that would have gone sparse without it,
                          // [14] the absolute number of the next batch's first round,
                          // [15] nonzero once any vertex has been set to TOP
    const int *sbig;      // static: vertices with big_deg < deg <= huge_deg
    const int *shuge;     // static: vertices with deg > huge_deg
    int big_deg, huge_deg;
    int64_t dense_min;    // estimated work (vertices) at or above which a round sweeps all
    unsigned long long *stats;  // per-workgroup slots of {processed, edges, lifted, preds}
    int *rounds_done;     // last round that was not MODE_DONE
    int *rise;            // n: round in which each vertex last rose (TOP proof candidates);
                          // null when proofs are off
    // Dense rounds in successor-range phases (phases > 1): in phase s a vertex folds its
    // edges whose successor lies in slice s = [s << slice_shift, (s + 1) << slice_shift), so
    // the value gathers of the whole grid come from one slice small enough to stay in L2.
    int phases;           // 1 = one in-place sweep (body_dense)
    int slice_shift;
    // Slice layout in tiles of kSliceTile vertices: tile T's low-degree edges lie in its own
    // CSR range from ptr[T * kSliceTile], grouped by (slice, vertex).  For tile T and slice
    // s, slen[(T * phases + s) * kSliceTile + j] is the number of edges of vertex
    // T * kSliceTile + j into slice s (0 above big_deg and past n) and
    // sbase[(T * phases + s) * kSliceWaves + w] is where the edges of wave w's 64 vertices
    // start; a vertex's edges follow those of the lower lanes of its wave.
    const uint8_t *slen;
    const int *sbase;
    const void *sedges;   // uint32 (w << sshift) | succ when sshift > 0, else int2 {succ, w}
    int sshift;
    const unsigned long long *bigmask;   // bit v: vertex v is above big_deg (lifted by the wave/block kernels)
    int round_grid;       // workgroups of k_round's low-degree part
    int hub_fused;        // 1: k_round's grid continues with the hub part in dense rounds
    int wave_grid, block_grid;   // workgroups of the wave and workgroup-per-vertex hub bodies
    int phase_grid;       // workgroups of the phased kernels (about what the device holds at once)
    int phase_tiles;      // tiles of the slice layout per workgroup, at most 32 (96 KB of LDS
                          // at 12 bytes per vertex; the dense kernel keeps a 32-bit mask of them)
    int *hacc, *hcnt;     // n each: partial keys and finished chunks of the vertices above
                          // huge_deg (TOP and 0 between rounds)
    long long *hbest;     // n: partial longest-path candidates of the vertices above huge_deg
                          // in proof passes (each owner's identity between passes)
    int *hflag;           // n: bit 0 a successor changed, bit 1 an edge leaves D (0 between passes)
};

constexpr int kMaxPhases = 16;             // successor slices of a phased dense round, at most
constexpr int kDefaultPhases = 8;          // slices of 2 MB of values on the 2^22-vertex game
constexpr int kPhaseMinN = 1 << 20;        // up to this many vertices the values (4 MB) fit one XCD's L2
constexpr int kSliceTile = 256;            // vertices per tile of the slice layout
constexpr int kSliceWaves = kSliceTile / 64;
constexpr int kMaxSliceDeg = 255;          // a vertex's edges into one slice fit a byte

// Edge k of the slice layout: successor and weight.
template <bool PACK>
__device__ __forceinline__ void slice_edge(const DevGame &g, int k, int &succ, int &w) {
    if (PACK) {
        const uint32_t e = ((const uint32_t *)g.sedges)[k];
        succ = (int)(e & ((1u << g.sshift) - 1u));
        w = ((int)e) >> g.sshift;
    } else {
        const int2 e = ((const int2 *)g.sedges)[k];
        succ = e.x;
        w = e.y;
    }
}

// Launch shape of one round.
struct LaunchShape {
    int group;       // lanes per vertex in k_round (1..64)
    int grid;        // workgroups of k_round
    int use_wave, use_block;
    int wave_grid, block_grid;
};

hipError_t launch_pack(const int *succ, const int *w, int2 *edges, int64_t m, hipStream_t st);
hipError_t launch_unpack(const uint32_t *packed, int shift, int2 *edges, int64_t m, hipStream_t st);
hipError_t launch_transpose(const DevGame &g, int *tile_sum, int *fill, hipStream_t st);
// The transpose as a two-level stable counting sort (egs_transpose.hip): cpred in ascending
// source order per column.  Buffers: counts and offs of transpose_counts_len + 1 ints each,
// scan_tiles of transpose_counts_len / 4096 + 2 ints, items[m].  transpose_sorted_fits
// says whether a game's sizes fit the buffers.
int transpose_bshift(int n);
int64_t transpose_counts_len(int n, int64_t m);
bool transpose_sorted_fits(int n, int64_t m, int64_t counts_cap, int64_t items_cap);
hipError_t launch_transpose_sorted(const DevGame &g, int *counts, int *offs, int *scan_tiles, int2 *items,
                                   hipStream_t st);
hipError_t launch_exclusive_scan(const int *in, int n, int *tile_sum, int *out, hipStream_t st);
constexpr int kTrTileEdges = 1 << 14;        // edges per tile of the sorted transpose

// A game in read order on the device and the work arrays of its renaming (egs_ingest.hip).
struct ReadOrderBuffers {
    const uint8_t *owner;                      // n: 0 = owner 0 (token MAX), else owner 1
    const int *rptr, *rsucc, *rw;              // read-order CSR, successors as read indices
    int *flag0, *pre0, *rev, *ext_of_int, *deg, *scan_tiles;   // n (+1) each; scan_tiles n/4096 + 1
};
hipError_t launch_read_order_check(const int *rptr, const int *rsucc, int n, int64_t m, int *bad, hipStream_t st);
hipError_t launch_internalise(const ReadOrderBuffers &b, int n, int *ptr, int2 *edges, unsigned long long *mg,
                              hipStream_t st);
hipError_t launch_classify(const DevGame &g, int *counts, hipStream_t st);
// in_slices: the slice set-up (init) already stored the initial state of the vertices up to
// big_deg and counted, without listing, the active ones (no transpose: round 1 sweeps densely)
hipError_t launch_init(const DevGame &g, hipStream_t st, bool in_slices);
hipError_t launch_slices(const DevGame &g, hipStream_t st, bool init);
int pack_shift(int n, int wmax);
// Round j of a batch (absolute round ring[14] + j); launch_advance ends the batch.
hipError_t launch_round(const DevGame &g, const LaunchShape &sh, int j, hipStream_t st);
hipError_t launch_advance(const DevGame &g, int rounds, hipStream_t st);
hipError_t launch_prepare_reset(const DevGame &g, hipStream_t st);
hipError_t launch_sum_stats(const DevGame &g, int nslots, unsigned long long *out, hipStream_t st);
hipError_t launch_degree_stats(const int *ptr, int n, int *out, hipStream_t st);
hipError_t launch_violations(const DevGame &g, unsigned long long *out, hipStream_t st);
hipError_t launch_winning(const int *val, int n, unsigned *out, hipStream_t st);
hipError_t launch_hbest_reset(const DevGame &g, hipStream_t st);
int grid_for(int64_t work, int threads, int cap);

// Certified TOP detection (egs_accel.hip).
struct CertBuffers {
    int *hist;        // device: 2 x 64 histogram bins + 8 flags
    long long *h;     // device: n longest-path potentials
    int *bfcnt;       // device: ring of 64 pass counters
    unsigned *bits;   // device: 3 two-level bitmaps (ceil(n/32) words + a bit-per-word summary)
    int *pinned;      // host pinned: 2 x 64 + 4 + 64 ints
    hipEvent_t ev = nullptr;   // waits on read-backs go through this event when set (else the stream)
};
// Wait for everything queued on `st` so far: through `ev` (recorded, then waited on) when
// given, else the whole stream.
inline hipError_t wait_queued(hipEvent_t ev, hipStream_t st) {
    if (!ev) return hipStreamSynchronize(st);
    hipError_t e = hipEventRecord(ev, st);
    return e == hipSuccess ? hipEventSynchronize(ev) : e;
}
struct CertResult {
    int threshold;       // last threshold tried
    int tried;           // candidates tried
    int passes;          // longest-path passes over all candidates
    int marked;          // vertices set to TOP
    int closure_failed;  // candidates that lost vertices to the closure check
    int cap_hit;         // some candidate ran out of passes (neither converged nor diverged)
    int last_passes;     // passes of the last candidate that converged (0: none did)
};
hipError_t launch_abs_wmax(const DevGame &g, int *out, hipStream_t st);
int gap_candidates(const int *win, const int *flo, int wmax, int mg, int64_t *out, int max_out);
// The value histogram of the next attempt, queued with its read-back into the pinned area.
hipError_t certify_prefetch(const DevGame &g, const CertBuffers &cb, int wmax, hipStream_t st);
// stop (may be null): polled before each candidate and each chunk of passes; a set flag
// abandons the attempt with nothing marked.  pass_hint: passes an earlier attempt of the
// same run needed (0: none yet), which sizes the first chunk queued before a read-back.
int certify_top(const DevGame &g, const CertBuffers &cb, int wmax, int last, int cap, int max_cands, bool has_big,
                bool hist_ready, const volatile uint32_t *stop, int pass_hint, hipStream_t st, CertResult &res);
// Potentials certifying that every TOP vertex of the current values is won by player 1
// (cb.h: >= 0 on the certified set, -1 elsewhere).  Returns 1 when the whole TOP set is
// certified, 0 when not (closure failed, divergence or the pass cap; see res), -1 on a HIP
// error, -2 when the potentials could overflow 64 bits.
int certify_all_top(const DevGame &g, const CertBuffers &cb, int wmax, int max_passes, bool has_big,
                    hipStream_t st, CertResult &res);

}  // namespace egs
