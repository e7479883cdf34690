/*
 * egs_oracle.c -- CPU oracle for the energy-game fixpoint (TEST INFRASTRUCTURE ONLY).
 *
 * This file restates, in plain C, the CPU algorithms of fvella/EGSolver that the GPU
 * library must match bit for bit.  It is the checker: only the test suite, the smoke
 * entry and the CPU-baseline leg of bench.py load it.  No product code links it.
 *
 * Each function names the reference lines it follows.  Internal vertex numbering is the
 * reference's: owner-0 vertices occupy [0, n0), owner-1 vertices [n0, n).  Owner 0 takes
 * the MIN over its out-edges, owner 1 the MAX (src/cpu_solver.c:300-305).
 *
 * Pinning (what checks each function, beyond retyping it):
 *   ora_ominus, ora_eg, ora_eg0 : script/sol.txt golden (2000 values, 2011 sweeps),
 *       the compiled reference cpu_solver.c (oracle/_ref), eg_fig2 hand solution,
 *       ladder/clique closed forms, brute-force Kleene iteration on tiny games.
 *   ora_mg                      : closed forms on hand-built games, overflow case.
 *   ora_csr2csc                 : the 4x5 known-answer case of src/host_csr2csc.c:49-133
 *                                 and the compiled reference host_csr2csc.c.
 *   ora_max_loop                : closed forms incl. the LONG_MAX-1 saturation.
 *   ora_fixpoint_violations     : planted non-fixpoints must be rejected.
 *   ora_lift_all                : one step of the brute-force Kleene iteration (oracle.py),
 *                                 lift(lfp) == lfp, and lift(0) against one EG0 sweep.
 *   ora_initial_active          : equals lift(0)(v) > 0 (the closed form of
 *                                 src/cpu_solver.c:217-249) on games of mixed-sign weights
 *                                 for both owners, and hand-built single-vertex cases.
 *   ora_eg0_from                : from 0 it is EG0; seeded with TOP on the lfp's TOP set it
 *                                 returns the lfp in fewer sweeps; seeded above the lfp it
 *                                 keeps the planted value (the seed is used).
 */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORA_TOP INT_MAX

/* OMINUS(a, w) with threshold mg: src/common.h:270-272.  The short-circuit on a<TOP
 * keeps a-w from overflowing when a is TOP. */
static inline int ora_ominus_i(int a, int w, int mg) {
    if (a < ORA_TOP && (a - w) <= mg) {
        int d = a - w;
        return d > 0 ? d : 0;
    }
    return ORA_TOP;
}

int ora_ominus(int a, int w, int mg) { return ora_ominus_i(a, w, mg); }

/* Loop bound: src/egsolver.c:64-90.  A negative request means "theoretical bound"
 * n*m*MG+1, saturated to LONG_MAX-1 when it would overflow. */
long ora_max_loop(long m, long n, long mg, long maxloop) {
    long res = maxloop;
    if (res < 0L) {
        if (mg > (LONG_MAX - 1) / (n * m)) {
            res = LONG_MAX - 1;
        } else {
            res = n * m * mg + 1;
        }
    }
    return res;
}

/* MG threshold: src/egsolver.c:283-304.  MG = sum over vertices of max(0, max_e -w_e).
 * Returns 0 and stores MG, or -1 when the running sum would pass INT_MAX (the
 * reference exits with "Error too high weights"). */
int ora_mg(int n, const int *ptr, const int *w, int *mg_out) {
    int mg = 0;
    for (int v = 0; v < n; v++) {
        int worst = 0;
        for (int e = ptr[v]; e < ptr[v + 1]; e++) {
            int neg = -w[e];
            if (neg > worst) worst = neg;
        }
        if (mg > INT_MAX - worst) return -1;
        mg += worst;
    }
    *mg_out = mg;
    return 0;
}

/* CSR -> CSC of predecessors: src/host_csr2csc.c:15-46.  Counting sort by successor,
 * stable in the source-row order, so pred lists are ascending by source id.  The
 * weight payload is moved along when cw != NULL. */
void ora_csr2csc(int n, int m, const int *ptr, const int *succ, const int *w,
                 int *cptr, int *cpred, int *cw) {
    for (int i = 0; i <= n; i++) cptr[i] = 0;
    for (int e = 0; e < m; e++) cptr[succ[e] + 1]++;
    for (int i = 0; i < n; i++) cptr[i + 1] += cptr[i];
    int *fill = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    memcpy(fill, cptr, sizeof(int) * (size_t)(n + 1));
    for (int v = 0; v < n; v++) {
        for (int e = ptr[v]; e < ptr[v + 1]; e++) {
            int slot = fill[succ[e]]++;
            cpred[slot] = v;
            if (cw) cw[slot] = w[e];
        }
    }
    free(fill);
}

/* Initial worklist membership: src/cpu_solver.c:217-249.  An owner-0 vertex starts
 * active when every out-weight is negative, an owner-1 vertex when at least one is.
 * Equivalently: lift(0)(v) > 0.  Writes flags[v] in {0,1}; returns the count. */
int ora_initial_active(int n, int n0, const int *ptr, const int *w, unsigned char *flags) {
    int cnt = 0;
    for (int v = 0; v < n; v++) {
        int all_neg = 1, any_neg = 0;
        for (int e = ptr[v]; e < ptr[v + 1]; e++) {
            if (w[e] >= 0) all_neg = 0; else any_neg = 1;
        }
        int act = (v < n0) ? all_neg : any_neg;
        flags[v] = (unsigned char)act;
        cnt += act;
    }
    return cnt;
}

static inline int ora_lift_vertex(int v, int n0, int mg, const int *ptr, const int *succ,
                                  const int *w, const int *val) {
    int e = ptr[v];
    int best = ora_ominus_i(val[succ[e]], w[e], mg);
    for (e = e + 1; e < ptr[v + 1]; e++) {
        int x = ora_ominus_i(val[succ[e]], w[e], mg);
        if (v < n0) { if (x < best) best = x; }
        else        { if (x > best) best = x; }
    }
    return best;
}

/* EG worklist solver: src/cpu_solver.c:190-333 (the reference default, --cpu --eg).
 * A FIFO of capacity n seeded in vertex order with the initial active set; pop, lift,
 * and on increase append every not-yet-queued predecessor in CSC order.  `val` must
 * hold zeros on entry.  Returns the reference's processed count (loop-1).  When
 * stop_flag is non-NULL and becomes non-zero the loop ends after the current vertex
 * (the reference polls timeout_expired the same way). */
long ora_eg(int n, int n0, int m, int mg, const int *ptr, const int *succ, const int *w,
            int *val, long max_loop, const volatile int *stop_flag) {
    int *cptr = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *cpred = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    unsigned char *inq = (unsigned char *)calloc((size_t)n + 1, 1);
    /* One slot beyond n: see the wrap note below. */
    int *queue = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    long loop_cap = ora_max_loop((long)m, (long)n, (long)mg, max_loop);

    int tail = 0;
    for (int v = 0; v < n; v++) {
        int all_neg = 1, any_neg = 0;
        for (int e = ptr[v]; e < ptr[v + 1]; e++) {
            if (w[e] >= 0) all_neg = 0; else any_neg = 1;
        }
        if ((v < n0) ? all_neg : any_neg) { inq[v] = 1; queue[tail++] = v; }
    }
    ora_csr2csc(n, m, ptr, succ, w, cptr, cpred, NULL);

    /* Wrap note: the reference seeds the write index with the initial count without
     * reducing it modulo n (cpu_solver.c:247), then advances it modulo n
     * (cpu_solver.c:319).  When all n vertices start active, the first append lands
     * in slot n, which the reader never visits; the reader later sees the stale slot 0
     * instead.  The restatement keeps that behaviour so its output equals the
     * reference's on such inputs too. */
    int head = 0;
    long len = tail;
    long loop = 1;
    while (len > 0 && loop <= loop_cap) {
        loop++;
        int changed = 0;
        int v = queue[head];
        head = (head + 1) % n;
        len--;
        inq[v] = 0;
        int t = ora_lift_vertex(v, n0, mg, ptr, succ, w, val);
        if (val[v] < t) {
            val[v] = t;
            changed = 1;
            for (int e = cptr[v]; e < cptr[v + 1]; e++) {
                int u = cpred[e];
                if (!inq[u]) {
                    queue[tail] = u;
                    tail = (tail + 1) % n;
                    inq[u] = 1;
                    len++;
                }
            }
        }
        if (!changed && len == 0) break;
        if (stop_flag && *stop_flag) break;
    }
    free(cptr); free(cpred); free(inq); free(queue);
    return loop - 1;
}

/* EG0 Jacobi solver: src/cpu_solver.c:109-185 (--cpu --eg0).  Two buffers; each sweep
 * writes buf2[v] = max(buf2[v], lift(buf1)(v)) for all v in parallel (OpenMP, guided
 * schedule, chunk 4 as cpu_solver.c:12,146), then swaps.  Stops on a sweep with no
 * change, on the loop cap, or on stop_flag.  The result buffer choice mirrors
 * cpu_solver.c:178-182: when the loop cap is odd the final answer is read from the
 * second buffer.  `out` receives that buffer; returns the sweep count (loop-1). */
long ora_eg0(int n, int n0, int m, int mg, const int *ptr, const int *succ, const int *w,
             int *out, long max_loop, int threads, const volatile int *stop_flag) {
    int *b1 = (int *)calloc((size_t)n + 1, sizeof(int));
    int *b2 = (int *)calloc((size_t)n + 1, sizeof(int));
    int *first = b1, *second = b2;
    long loop_cap = ora_max_loop((long)m, (long)n, (long)mg, max_loop);
    if (threads > 0) omp_set_num_threads(threads);
    long loop;
    for (loop = 1; loop <= loop_cap; loop++) {
        int changed = 0;
#pragma omp parallel for schedule(guided, 4) reduction(| : changed)
        for (int v = 0; v < n; v++) {
            int t = ora_lift_vertex(v, n0, mg, ptr, succ, w, b1);
            if (b2[v] < t) { b2[v] = t; changed = 1; }
        }
        int *tmp = b1; b1 = b2; b2 = tmp;
        if (!changed) break;
        if (stop_flag && *stop_flag) break;
    }
    /* cpu_solver.c:178-182 swaps the two global result pointers when the loop cap is
     * odd; the caller then prints host_ResNodeValues1.  Track which physical buffer
     * that is: `first` unless the cap is odd. */
    const int *res = ((loop_cap % 2) != 0) ? second : first;
    memcpy(out, res, sizeof(int) * (size_t)n);
    free(first); free(second);
    return loop - 1;
}

/* EG0 from a start vector: the Jacobi sweeps of ora_eg0 (src/cpu_solver.c:109-185) with
 * both buffers seeded from `start` instead of zeros (the reference always starts at 0).
 * From any start f0 <= lfp the sweeps rise to the least fixpoint, since lifting is monotone
 * and lift(lfp) = lfp; the tests seed TOP on a set already proved to be player 1's and 0
 * elsewhere, which leaves only the finite values to compute.  Stops on a sweep with no
 * change or after max_sweeps (< 0: none); `out` gets the last buffer written, the return
 * value is the sweep count. */
long ora_eg0_from(int n, int n0, int mg, const int *ptr, const int *succ, const int *w,
                  const int *start, int *out, long max_sweeps, int threads) {
    int *b1 = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *b2 = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(b1, start, sizeof(int) * (size_t)n);
    memcpy(b2, start, sizeof(int) * (size_t)n);
    if (threads > 0) omp_set_num_threads(threads);
    long sweeps = 0;
    for (;;) {
        if (max_sweeps >= 0 && sweeps >= max_sweeps) break;
        int changed = 0;
#pragma omp parallel for schedule(guided, 4) reduction(| : changed)
        for (int v = 0; v < n; v++) {
            int t = ora_lift_vertex(v, n0, mg, ptr, succ, w, b1);
            if (b2[v] < t) { b2[v] = t; changed = 1; }
        }
        int *tmp = b1; b1 = b2; b2 = tmp;
        sweeps++;
        if (!changed) break;
    }
    memcpy(out, b1, sizeof(int) * (size_t)n);
    free(b1); free(b2);
    return sweeps;
}

/* Fixpoint certificate: returns the number of vertices with lift(f)(v) != f(v).  Zero
 * together with "f was reached by monotone lifting from 0" certifies f = lfp. */
long ora_fixpoint_violations(int n, int n0, int mg, const int *ptr, const int *succ,
                             const int *w, const int *val) {
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int v = 0; v < n; v++) {
        if (ora_lift_vertex(v, n0, mg, ptr, succ, w, val) != val[v]) bad++;
    }
    return bad;
}

/* One synchronous lift of every vertex: out[v] = lift(val)(v). */
void ora_lift_all(int n, int n0, int mg, const int *ptr, const int *succ, const int *w,
                  const int *val, int *out) {
    for (int v = 0; v < n; v++) out[v] = ora_lift_vertex(v, n0, mg, ptr, succ, w, val);
}
