/*
 * egs.h -- C-ABI of the MI355X energy-game solver library (libegs.so).
 *
 * Plain pointers and sizes only.  A game is given in the reference's internal numbering:
 * owner-0 vertices (input token MAX) occupy [0, n0) and take the minimum over their
 * out-edges, owner-1 vertices (token MIN) occupy [n0, n) and take the maximum
 * (src/egsolver.c:185-188,253-281; src/cpu_solver.c:300-305).  CSR arrays:
 * ptr[n+1], succ[m], w[m], int32.  Values are int32; 2147483647 is TOP (player 1 wins).
 * mg is the reference's MG_pesi threshold (src/egsolver.c:283-304).
 *
 * Every function returns 0 on success or a negative EGS_E* code; egs_last_error()
 * gives the message of the calling thread's last failure.  Nothing here falls back to
 * the CPU: without a usable gfx950 device the calls fail.
 */
#ifndef EGS_H
#define EGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    EGS_OK = 0,
    EGS_E_ARG = -1,      /* bad argument (sizes, pointers, options) */
    EGS_E_HIP = -2,      /* HIP runtime error */
    EGS_E_NOMEM = -3,    /* device or pinned allocation failed */
    EGS_E_STATE = -4,    /* call out of order (e.g. iterate before prepare) */
    EGS_E_CAPACITY = -5  /* game larger than the solver or slot was sized for */
};

enum { EGS_MEM_HOST = 0, EGS_MEM_DEVICE = 1 };

/* Tuning knobs of one solve; zero-initialise for the defaults.  The choice between one
 * lane per vertex and one wavefront per vertex is made per vertex by out-degree:
 * vertices with degree <= big_deg are lifted by `group` lanes each, larger ones by a
 * whole wavefront (64 lanes), and those above huge_deg by a whole workgroup. */
typedef struct egs_options {
    int32_t group;          /* lanes per low-degree vertex: 1,2,4,...,64; 0 = auto */
    int32_t big_deg;        /* 0 = auto */
    int32_t huge_deg;       /* 0 = auto */
    int32_t grid;           /* workgroups of the group kernel; 0 = auto */
    int32_t batch_rounds;   /* rounds enqueued between host convergence checks; 0 = auto */
    int32_t dense_pct;      /* a round sweeps every vertex when its estimated work is at
                               least this percent of n (else it walks the worklist);
                               0 = auto (12), > 100 = never sweep */
    int32_t certify_every;  /* rounds between attempts to prove climbing vertices TOP
                               (exact; attempts back off while they fail);
                               0 = auto (7), < 0 = never */
    int32_t cert_max_passes;/* ceiling of the longest-path passes per proof candidate
                               (the cap starts at 64 and doubles while candidates run
                               out of passes); 0 = auto (1024) */
    int64_t max_phases;     /* phase cap; < 0 = the reference bound n*m*MG+1
                               (src/egsolver.c:64-90) */
    const volatile uint32_t *stop_flag;  /* polled before every round and before each step
                               of a proof attempt; non-zero stops the run early with
                               partial values (the reference's timeout_expired,
                               src/sig_handling.c:57-64, polled once per round at
                               src/dev_EG_alg.cu:376).  With a stop flag at most 4 rounds
                               are queued ahead of the host. */
    int32_t slices;         /* dense rounds in successor-range phases (their value gathers
                               stay in L2): 0 = auto (8 phases when n > 2^20 and the solver
                               was created with n_cap >= 2^20), 1 = one sweep, 2..16 = at most
                               that many phases over slices of a power-of-two size.  Phased
                               games lift vertices above 255 edges with the wavefront kernels
                               (big_deg is capped at 255). */
    int32_t phase_wgs_per_cu; /* workgroups per compute unit of the phased kernels (their
                               grid is about what the device holds at once); 0 = auto (8) */
    int32_t graphs;         /* batches of rounds replayed as HIP graphs captured once per
                               prepare (batches of 1,2,4,...,32 rounds): 1 = on, -1 = off
                               (every kernel launched), 0 = auto.  Off while a stop flag
                               is set (it is polled before each round) and in
                               egs_solve_noalloc (instantiating a graph allocates). */
} egs_options;

typedef struct egs_stats {
    int64_t phases;         /* rounds with a non-empty frontier */
    int64_t processed;      /* vertices lifted-or-checked, the reference's processedNodes */
    int64_t edges_relaxed;  /* sum of out-degree over processed vertices */
    int64_t lifted;         /* vertices whose value increased */
    int64_t pred_visits;    /* sum of in-degree over lifted vertices */
    int32_t converged;      /* 1 when the frontier emptied (values are the fixpoint) */
    int32_t initial_active; /* size of the initial frontier */
    double ms_prepare;      /* pack + transpose + initial frontier (device time) */
    double ms_iterate;      /* fixpoint rounds and proofs (device time) */
    int64_t cert_attempts;  /* TOP proofs attempted */
    int64_t certificates;   /* TOP proofs that succeeded */
    int64_t cert_passes;    /* longest-path passes spent on proofs */
    int64_t top_marked;     /* vertices set to TOP by proofs */
    int64_t cert_pruned;    /* proof candidates that lost vertices to the closure check */
    int32_t csc_built;      /* 1 when this solve built the transpose (only sparse rounds
                               need it; dense-only solves never build it) */
    int32_t reserved;
} egs_stats;

/* ---- device ---------------------------------------------------------------------- */

/* Device facts for the driver's banner (replaces checkDevice(), src/utils.cu:51-73,
 * and cudaDriverGetVersion/cudaRuntimeGetVersion, src/egsolver.c:372-373). */
typedef struct egs_device_info {
    char name[256];
    char arch[64];
    int32_t device_count;
    int32_t warp_size;
    int32_t max_threads_per_block;
    int32_t multiprocessors;
    int64_t total_mem;
    int64_t shared_mem_per_block;
    int32_t clock_khz;
    int32_t ecc;
    int32_t driver_version;
    int32_t runtime_version;
    int32_t cap_major;
    int32_t cap_minor;
} egs_device_info;

int egs_device_query(int device, egs_device_info *out);
int egs_device_reset(int device);
const char *egs_last_error(void);

/* Pinned (page-locked) host memory, as cudaMallocHost/cudaFreeHost in the reference
 * (src/dev_solver.cu:156,244). */
int egs_host_alloc(void **p, size_t bytes);
int egs_host_free(void *p);
/* Pops the HIP runtime's sticky error: 0 when none, else the code with its text in buf
 * (replaces cudaGetLastError in checkCUDAError, src/utils.cu:38-45). */
int egs_runtime_error(char *buf, int len);

/* ---- one game per solver --------------------------------------------------------- */

typedef struct egs_solver egs_solver;

/* Allocate device workspace for games up to n_cap vertices / m_cap edges. */
int egs_solver_create(int device, int32_t n_cap, int64_t m_cap, egs_solver **out);
int egs_solver_destroy(egs_solver *s);

/* Copy a game in (host arrays or device arrays per `mem`).  Leaves it unprepared.  Host
 * arrays are checked (offsets from 0 to m, an out-edge per vertex, successors in range);
 * device arrays are trusted to satisfy the same rules, as the reference's parser requires
 * (every vertex lists at least one successor). */
int egs_solver_load(egs_solver *s, int mem, int32_t n, int32_t n0, int64_t m, int32_t mg,
                    const int32_t *ptr, const int32_t *succ, const int32_t *w);

/* Transpose + initial frontier; values reset to 0.  Replaces EG_gpu_traspose_graph and
 * kernel_EG_initialize (src/dev_EG_alg.cu:277-290,176-205). */
int egs_solver_prepare(egs_solver *s, const egs_options *opt);

/* Run up to `max_phases` further rounds (< 0: up to the option's cap).  Returns 1 when
 * the fixpoint was reached, 0 when stopped by the cap or the stop flag. */
int egs_solver_iterate(egs_solver *s, int64_t max_phases, egs_stats *st);

int egs_solver_stats(egs_solver *s, egs_stats *st);
/* No reference counterpart.  Load a game in read order, as a reader or generator produces
 * it (external id = read index): owner[n] (0 = token MAX = owner 0, else owner 1), rptr[n+1],
 * rsucc[m] (read indices), rw[m]; host or device arrays per `mem`.  The owner partition
 * (owner-0 vertices first, each block in read order: src/egsolver.c:146-214 with the
 * default comparator under glibc's stable qsort), the internal CSR (src/egsolver.c:253-281)
 * and MG (src/egsolver.c:283-304) are computed on the device after one upload.  The input is
 * checked on the device (offsets from 0 to m, an edge per vertex, successors in range;
 * EGS_E_ARG otherwise, also when MG passes INT32_MAX).  *n0_out and *mg_out get the
 * internal owner-0 count and MG; ext_of_int[n] (may be NULL) the read index of each
 * internal vertex. */
int egs_solver_load_read_order(egs_solver *s, int mem, int32_t n, int64_t m, const uint8_t *owner,
                               const int32_t *rptr, const int32_t *rsucc, const int32_t *rw, int32_t *n0_out,
                               int32_t *mg_out, int32_t *ext_of_int);
/* No reference counterpart.  The loaded game's internal CSR back on the host: ptr[n+1],
 * succ[m], w[m]. */
int egs_solver_read_game(egs_solver *s, int32_t *ptr, int32_t *succ, int32_t *w);
/* No reference counterpart.  Load a game already resident on the device in the 32-bit
 * packed edge layout of the batched and serving paths (edge e = (w << shift) | succ,
 * 1 <= shift <= 30), with its max |w| and max out-degree known; the CSR is trusted to
 * have been checked (every vertex has an edge, successors in range). */
int egs_solver_load_packed(egs_solver *s, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *d_ptr,
                           const uint32_t *d_edges, int32_t shift, int32_t wmax, int32_t max_outdeg);
int egs_solver_read_values(egs_solver *s, int32_t *host_values);
/* No reference counterpart.  Player 0's winning set {v : value < TOP} as ceil(n/32) words,
 * bit v%32 of word v/32, internal numbering (the reference derives it from the printed
 * values). */
int egs_solver_read_winning(egs_solver *s, uint32_t *host_bits);
const int32_t *egs_solver_device_values(egs_solver *s);
/* Device arrays of the transposed graph (for inspection): cptr[n+1], cpred[m]. */
int egs_solver_read_csc(egs_solver *s, int32_t *cptr, int32_t *cpred);
/* Stream the solver enqueues on (hipStream_t). */
void *egs_solver_stream(egs_solver *s);

/* One-shot: create, load from host, prepare, iterate to the cap, read values back. */
int egs_solve(int device, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
              const int32_t *succ, const int32_t *w, int32_t *values, const egs_options *opt,
              egs_stats *st);

/* Fixpoint certificate on the device: number of vertices with lift(val)(v) != val(v).
 * Zero, for values reached by monotone lifting from 0, certifies the least fixpoint. */
int egs_solver_count_violations(egs_solver *s, int64_t *violations);

/* No reference counterpart.  Replace the prepared solver's values with host_values[n]
 * (for example the result of the batched or serving path, to certify it below). */
int egs_solver_set_values(egs_solver *s, const int32_t *host_values);

/* No reference counterpart.  Certificate that every TOP vertex of the current values is
 * won by player 1: potentials host_h[n] (int64) with host_h[v] >= 0 exactly on
 * T = {v : value == TOP} and -1 elsewhere, such that, with w' = (n+1) w + 1,
 *   every owner-0 vertex v of T has all successors u in T and h(v) >= w'(v,u) + h(u),
 *   every owner-1 vertex v of T has some successor u in T with h(v) >= w'(v,u) + h(u).
 * Such potentials exist exactly when T lies in player 1's winning region (the least
 * fixpoint's TOP set): the chosen owner-1 edges keep every play in T, and every cycle
 * under them has w-weight <= -1.  egs_check_top_certificate (egs_host.h) verifies them
 * on the host in O(m).  Computed on the device by longest-path passes from h = 0, at most
 * max_passes (<= 0: 65,536).  Returns 1 when T is certified, 0 when not (host_h then holds
 * the partial potentials), or an error code; *passes (may be NULL) gets the passes run. */
int egs_solver_top_certificate(egs_solver *s, int32_t max_passes, int64_t *host_h, int32_t *passes);

/* No reference counterpart.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): egs_solve for a
 * game already on the device, with no allocation of any kind.  The solver's workspace is
 * a fixed array in the library's device memory (26 MiB per device that runs it,
 * held until the process ends), so games are limited to 65,536 vertices and 1,048,576
 * edges (EGS_E_CAPACITY beyond).  d_ptr[n+1], d_succ[m], d_w[m] and d_values[n] are device
 * arrays of the current device; work runs on `stream` (a hipStream_t; NULL = the null
 * stream), which is synchronised before the call returns; host-side timing.  Every call on
 * one device shares that device's workspace, which keeps the last call's contents: calls
 * that overlap on one device race (serialise them).  Returns 1 at the fixpoint, 0 when
 * stopped by the options' cap or stop flag, or an error code. */
int egs_solve_noalloc(int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *d_ptr, const int32_t *d_succ,
                      const int32_t *d_w, int32_t *d_values, const egs_options *opt, egs_stats *st, void *stream);

/* ---- many games in one launch ---------------------------------------------------- */

/* No reference counterpart (the reference solves one game per process).  Games are
 * staged with egs_batch_add (internal numbering, as above), copied with
 * egs_batch_upload, and solved together by one persistent kernel launch in which each
 * workgroup takes whole games from a device counter, largest first. */
typedef struct egs_batch egs_batch;

typedef struct egs_batch_options {
    int32_t threads;            /* workgroup size: 256, 512 or 1024; 0 = 1024 */
    int32_t blocks_per_cu;      /* persistent workgroups per compute unit; 0 = 1 */
    int32_t lds_ints;           /* LDS values per workgroup: games up to this many vertices
                                   keep their values in LDS; 0 = all a workgroup may hold */
    int32_t lanes_per_edge_x16; /* lanes per vertex = smallest power of two >= average
                                   degree * this / 16; 0 = 2 */
    int32_t group_min, group_max; /* clamp of the lanes per vertex; 0 = 1 / 64 */
    int64_t max_sweeps;         /* per-game sweep cap; <= 0 = none */
    int32_t certify_every;      /* sweeps between TOP proofs per game; 0 = 4, < 0 = never */
    int32_t cert_max_passes;    /* ceiling of the longest-path passes per proof candidate; 0 = 1024 */
} egs_batch_options;

typedef struct egs_batch_stats {
    int64_t games;
    int64_t sweeps;             /* summed over games */
    int64_t edges_relaxed;
    int64_t lifted;
    int64_t converged;          /* games that reached their fixpoint */
    double ms;                  /* device time of the launch */
    int32_t packed_edges;       /* 1 when edges travelled as 32-bit words */
    int32_t grid, threads, lds_ints;
    int64_t certificates;       /* TOP proofs that succeeded, summed over games */
    int64_t cert_passes;        /* longest-path passes spent on proofs */
} egs_batch_stats;

int egs_batch_create(int device, egs_batch **out);
int egs_batch_destroy(egs_batch *b);
int egs_batch_clear(egs_batch *b);
/* Returns the game's index in the batch (>= 0) or an error code. */
int egs_batch_add(egs_batch *b, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
                  const int32_t *succ, const int32_t *w);
int egs_batch_size(egs_batch *b);
int egs_batch_upload(egs_batch *b);
int egs_batch_solve(egs_batch *b, const egs_batch_options *opt, egs_batch_stats *st);
int egs_batch_read_values(egs_batch *b, int game, int32_t *out);
int egs_batch_read_all(egs_batch *b, int32_t *out);
/* No reference counterpart.  Every game's values, concatenated in game order, in a pinned
 * buffer the batch owns: one device-to-host copy with no host copy after it.  *values
 * stays valid until the next read or solve of this batch and its destruction. */
int egs_batch_read_all_pinned(egs_batch *b, const int32_t **values);
int egs_batch_game_stats(egs_batch *b, int game, int64_t *sweeps, int64_t *edges, int32_t *converged);
const char *egs_batch_last_error(void);

/* ---- serving ------------------------------------------------------------------------ */

/* No reference counterpart.  A server owns a device arena.  submit() admits a game when
 * its footprint fits (else EGS_E_CAPACITY) and uploads it on a copy stream that
 * overlaps any running slice; launch() runs one time slice of the many-games kernel over
 * every resident unfinished game (graph, values and proof state stay resident between
 * slices); wait() ends the slice and lists the games that reached their fixpoint;
 * fetch() returns a finished game's values and frees its memory.  One thread drives a
 * server. */
typedef struct egs_server egs_server;

typedef struct egs_serve_options {
    int32_t threads;            /* 256, 512 or 1024; 0 = 1024 */
    int32_t blocks_per_cu;      /* 0 = 1 */
    int32_t lds_ints;           /* 0 = all a workgroup may hold */
    int32_t lanes_per_edge_x16; /* 0 = 2 */
    int32_t group_min, group_max;
    int32_t slice_sweeps;       /* cap of complete sweeps per game per slice; 0 = none */
    int32_t certify_every;      /* sweeps (small games) or rounds (large route) before proof
                                   attempts; 0 = 4 sweeps / 10 rounds, < 0 = never */
    int32_t cert_max_passes;    /* 0 = 1024 */
    int32_t slice_kedges;       /* edge budget of each workgroup per slice, in thousands of
                                   edges (bounds the slice time; sweeps resume part-way);
                                   0 = 2048 */
    int32_t large_cus;          /* compute units a slice leaves free while large-route games
                                   are in flight (the slice runs one workgroup per remaining
                                   CU); 0 = 64, < 0 = none */
} egs_serve_options;

int egs_server_create(int device, int64_t arena_bytes, int32_t max_games, egs_server **out);
int egs_server_destroy(egs_server *s);
/* Device bytes a game of n vertices, m edges and max |w| occupies in the arena. */
int64_t egs_server_footprint(int32_t n, int64_t m, int32_t wmax);
int64_t egs_server_free_bytes(egs_server *s);
/* Games with at least large_n vertices take the large route: wait() solves them with the
 * device-wide single-game solver, large_workers at a time (host threads, each with its
 * own solver and stream), beside the running slice, instead of one workgroup each.
 * large_n 0 = default (16384); large_workers 0 = default (2), at most 16. */
int egs_server_set_route(egs_server *s, int32_t large_n, int32_t large_workers);
int egs_server_free_slots(egs_server *s);
/* Returns a ticket >= 0, or EGS_E_CAPACITY when the game does not fit right now. */
int egs_server_submit(egs_server *s, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
                      const int32_t *succ, const int32_t *w);
/* No reference counterpart.  Admit a game in read order, as a .mpg file lists it:
 * owner[n] (0 = player 0, token MAX), rptr[n+1], rsucc[m] (successors as read indices),
 * rw[m].  The owner partition, internal CSR and MG are made on the device (the reference's
 * host driver does them at src/egsolver.c:107-220,253-304); the host only validates the
 * arrays while staging them.  A ticket like egs_server_submit's; egs_server_fetch returns
 * its values in read order. */
int egs_server_submit_read_order(egs_server *s, int32_t n, int64_t m, const uint8_t *owner, const int32_t *rptr,
                                 const int32_t *rsucc, const int32_t *rw);
/* Returns the number of games in the slice started (0: nothing resident). */
int egs_server_launch(egs_server *s, const egs_serve_options *opt);
/* Returns how many games finished in the slice; their tickets go to done[0..cap). */
int egs_server_wait(egs_server *s, int32_t *done, int32_t cap);
/* 1 when wait() would return without blocking, else 0 (lets a caller keep admitting
 * requests while a slice runs). */
int egs_server_ready(egs_server *s);
double egs_server_last_slice_ms(egs_server *s);
int egs_server_status(egs_server *s, int ticket);
int egs_server_fetch(egs_server *s, int ticket, int32_t *values, int64_t *sweeps, int64_t *edges);
const char *egs_server_last_error(void);

/* ---- calibration ------------------------------------------------------------------ */

/* Time `reps` launches of a plain 16-byte-vector device copy of `bytes` bytes and report
 * (read + write bytes) / time in GB/s.  Used to state the achievable HBM rate beside the
 * vendor peak and to scale profiler byte counters. */
int egs_copy_bandwidth(int device, int64_t bytes, int reps, double *gbps);

#ifdef __cplusplus
}
#endif
#endif /* EGS_H */
