/*
 * ref_harness.c -- drives the reference's own CPU solver (TEST INFRASTRUCTURE ONLY).
 *
 * Linked with the unmodified reference sources src/cpu_solver.c, src/host_csr2csc.c and
 * src/errori.c (compiled where they lie; see oracle/Makefile).  The reference keeps its
 * game and configuration in globals that src/egsolver.h defines and src/egsolver.c
 * fills; egsolver.c itself needs a bison-generated parser.h and cannot be built here,
 * so this harness defines those globals, fills them from plain arrays the way
 * postparsing()/alloca_memoria_host() do (src/egsolver.c:236-304,
 * src/dev_solver.cu:152-187), runs EG_cpu_solver() or EG0_cpu_solver(), and copies
 * host_ResNodeValues1 out.
 */
#include <fcntl.h>
#include <omp.h>
#include <sys/time.h>
#include "cpu_solver.h"

/* Globals consumed by cpu_solver.c (declared extern at cpu_solver.c:18-48). */
char *str;
int counter_nodi, counter_nodi0, num_nodi, num_archi, max_pesi, MG_pesi;
char *nodeFlags;
int *host_csrPtrInSuccLists, *host_csrSuccLists, *host_csrPesiArchi;
int *host_cscPtrInPredLists, *host_cscPredLists, *host_cscPesiArchiPred;
int numLowInDegree[4];
int numAllInDegree[LEN_DEGREEHIST];
int numAllOutDegree[LEN_DEGREEHIST];
int *host_ResNodeValues1, *host_ResNodeValues2, *host_ResNodeValuesAux;
int *host_csrDataArchiAux;
uint timeout_expired;
config configuration;
stat statistics;

void EG_cpu_solver(void);
void EG0_cpu_solver(void);

/* The loop bound helper lives in src/egsolver.c:64-90, which cannot be compiled here;
 * same arithmetic, warning text omitted. */
long aggiorna_max_loop(long Narchi, long Nnodi, long MGpesi, long maxloop) {
    long res = maxloop;
    if (res < 0L) {
        if (MGpesi > (LONG_MAX - 1) / (Nnodi * Narchi)) res = LONG_MAX - 1;
        else res = Nnodi * Narchi * MGpesi + 1;
    }
    return res;
}

/* algo: 0 = EG (worklist, cpu_solver.c:190), 1 = EG0 (Jacobi, cpu_solver.c:109).
 * Returns statistics.processedNodes; *seconds gets the wall time of the solver call
 * (transpose included for EG, as cpu_solver() times it). quiet != 0 sends the
 * reference's progress prints to /dev/null. */
long ref_solve(int algo, int n, int n0, int m, int mg, const int *ptr, const int *succ,
               const int *w, int *out, long max_loop, int threads, int quiet,
               double *seconds) {
    num_nodi = counter_nodi = n;
    counter_nodi0 = n0;
    num_archi = m;
    MG_pesi = mg;
    max_pesi = 0;
    for (int e = 0; e < m; e++) { int a = w[e] < 0 ? -w[e] : w[e]; if (a > max_pesi) max_pesi = a; }
    size_t aux = (size_t)(m > n + 1 ? m : n + 1);
    int *block = (int *)calloc((size_t)(n + 1) + 2 * (size_t)m + 2 * (size_t)n + 1 + aux, sizeof(int));
    host_csrPtrInSuccLists = block;
    host_csrSuccLists = block + n + 1;
    host_csrPesiArchi = host_csrSuccLists + m;
    host_ResNodeValues1 = host_csrPesiArchi + m;
    host_ResNodeValues2 = (algo == 1) ? host_ResNodeValues1 + n : host_ResNodeValues1;
    host_csrDataArchiAux = host_ResNodeValues1 + 2 * n + 1;
    memcpy(host_csrPtrInSuccLists, ptr, sizeof(int) * (size_t)(n + 1));
    memcpy(host_csrSuccLists, succ, sizeof(int) * (size_t)m);
    memcpy(host_csrPesiArchi, w, sizeof(int) * (size_t)m);
    int *trans = (int *)calloc((size_t)(n + 1) + 2 * (size_t)m, sizeof(int));
    host_cscPtrInPredLists = trans;
    host_cscPredLists = trans + n + 1;
    host_cscPesiArchiPred = host_cscPredLists + m;
    nodeFlags = (char *)calloc((size_t)n + 1, 1);
    memset(numLowInDegree, 0, sizeof(numLowInDegree));
    memset(numAllInDegree, 0, sizeof(numAllInDegree));
    memset(numAllOutDegree, 0, sizeof(numAllOutDegree));
    memset(&configuration, 0, sizeof(configuration));
    configuration.algoritmo = (algo == 1) ? ALGOR_EG0 : ALGOR_EG;
    configuration.max_loop_val = max_loop;
    configuration.print_degree_hist = NO_PRINT_DEGREEHIST;
    configuration.computationType = CPU_COMPUTATION;
    memset(&statistics, 0, sizeof(statistics));
    timeout_expired = 0;
    if (threads > 0) omp_set_num_threads(threads);

    int saved = -1;
    if (quiet) {
        fflush(stdout);
        saved = dup(1);
        int dn = open("/dev/null", O_WRONLY);
        dup2(dn, 1);
        close(dn);
    }
    struct timeval t0, t1;
    gettimeofday(&t0, NULL);
    if (algo == 1) EG0_cpu_solver(); else EG_cpu_solver();
    gettimeofday(&t1, NULL);
    if (quiet) { fflush(stdout); dup2(saved, 1); close(saved); }
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-6 * (double)(t1.tv_usec - t0.tv_usec);

    memcpy(out, host_ResNodeValues1, sizeof(int) * (size_t)n);
    free(block); free(trans); free(nodeFlags);
    return statistics.processedNodes;
}

/* The reference transpose, for the known-answer and cross checks. */
void ref_csr2csc(int n, int m, int *ptr, int *succ, int *w, int *cptr, int *cpred, int *cw) {
    host_csr2csc(n, n, m, w, succ, ptr, cw, cpred, cptr);
}
