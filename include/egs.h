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
    int64_t max_phases;     /* phase cap; < 0 = the reference bound n*m*MG+1
                               (src/egsolver.c:64-90) */
    const volatile uint32_t *stop_flag;  /* polled between batches; non-zero stops the
                               run early with partial values (the reference's
                               timeout_expired, src/sig_handling.c:57-64) */
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
    double ms_iterate;      /* fixpoint rounds (device time) */
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

/* Copy a game in (host arrays or device arrays per `mem`).  Leaves it unprepared. */
int egs_solver_load(egs_solver *s, int mem, int32_t n, int32_t n0, int64_t m, int32_t mg,
                    const int32_t *ptr, const int32_t *succ, const int32_t *w);

/* Transpose + initial frontier; values reset to 0.  Replaces EG_gpu_traspose_graph and
 * kernel_EG_initialize (src/dev_EG_alg.cu:277-290,176-205). */
int egs_solver_prepare(egs_solver *s, const egs_options *opt);

/* Run up to `max_phases` further rounds (< 0: up to the option's cap).  Returns 1 when
 * the fixpoint was reached, 0 when stopped by the cap or the stop flag. */
int egs_solver_iterate(egs_solver *s, int64_t max_phases, egs_stats *st);

int egs_solver_stats(egs_solver *s, egs_stats *st);
int egs_solver_read_values(egs_solver *s, int32_t *host_values);
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

/* ---- calibration ------------------------------------------------------------------ */

/* Time `reps` launches of a plain 16-byte-vector device copy of `bytes` bytes and report
 * (read + write bytes) / time in GB/s.  Used to state the achievable HBM rate beside the
 * vendor peak and to scale profiler byte counters. */
int egs_copy_bandwidth(int device, int64_t bytes, int reps, double *gbps);

#ifdef __cplusplus
}
#endif
#endif /* EGS_H */
