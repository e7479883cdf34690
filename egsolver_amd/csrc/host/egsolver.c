/*
 * egsolver.c -- command-line driver of the MI355X energy-game solver.
 *
 * Keeps the reference driver's surface (src/egsolver.c): the same options
 * (src/egsolver.c:494-528, 536-883), the same owner partition and CSR build
 * (src/egsolver.c:107-220, 236-315), the same timing block and solution formats
 * (src/egsolver.c:94-104, 432-469).  It reaches the GPU only through the reference's
 * entry points (include/egs_dev_solver.h), which read and write the globals defined
 * below, named and typed as in src/egsolver.h:24-119.
 *
 * Differences: the input is read by this repository's reader (libegs_host.so) instead
 * of a flex/bison parser; the solver is always the GPU one, so --gpu is the default and
 * --cpu stops with an error (the CPU solvers of the reference are not part of this
 * build); --tb is validated and accepted but the kernels choose their own workgroup
 * size; --deviceid selects the device.
 */
#define _GNU_SOURCE 1
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>
#include <unistd.h>

#include "egs.h"
#include "egs_dev_solver.h"
#include "egs_host.h"

/* ---- globals shared with the entry points (src/egsolver.h:24-119) ------------------- */
int counter_nodi, counter_nodi0, num_nodi, num_archi, max_pesi, MG_pesi, max_nodo;
int min_outdegree, max_outdegree;
double avg_outdegree, stddev_outdegree;
int numLowOutDegree[4];
int *host_allData, *host_csrPtrInSuccLists, *host_csrSuccLists, *host_csrPesiArchi;
int *host_ResNodeValues1, *host_ResNodeValues2, *host_ResNodeValuesAux, *host_flag;
int *host_csrDataArchiAux, *host_revSuccLists;
int *host_transData, *host_cscPtrInPredLists, *host_cscPredLists, *host_cscPesiArchiPred;
int *hdev_allData, *hdev_ResNodeValues1;
unsigned int timeout_expired;
egs_ref_config configuration;
egs_ref_stat statistics;
int *csrPtrInSuccLists, *nodePriority, *csrPesiArchi, *csrSuccLists;
int *nomeExt_of_nomeInt, *nomeInt_of_nomeExt, *mapping, *revmapping;
char *nodeOwner, *nodeFlags;
char **nodeName;

/* option values (src/common.h:139-237) */
enum { CPU_COMPUTATION = 0, GPU_COMPUTATION = 1 };
enum { NO_OUTPUT = 0, YES_ONELINEOUT = 1, NO_ONELINEOUT = 2 };
enum { SORT_N = 0, SORT_I = 1, SORT_O = 2, SORT_A = 3, SORT_OI = 4 };
enum { UNSET_TIMEOUT_OPT = 0, SET_TIMEOUT_OPT = 1 };
#define DEFAULT_ADD_RND_WEIGHTS_VAL 1
#define MAX_RAND_NUM (RAND_MAX / 2)
#define NO_USER_RND_SEED (-1)
#define DEFAULT_TIMEOUT_SEC 60

static const char *prog = "egsolver";

static double now_s(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

/* Loop bound: src/egsolver.c:64-90. */
long aggiorna_max_loop(long Narchi, long Nnodi, long MGpesi, long maxloop) {
    long res = maxloop;
    if (res < 0L) {
        if (MGpesi > (LONG_MAX - 1) / (Nnodi * Narchi)) {
            res = LONG_MAX - 1;
            fprintf(stderr, "WARNING too many loops: %ld * %ld * %ld > %ld --> using LONG_MAX-1:%ld\n", Nnodi,
                    Narchi, MGpesi, LONG_MAX, res);
        } else {
            res = Nnodi * Narchi * MGpesi + 1;
        }
    }
    return res;
}

static void short_usage(void) { fprintf(stderr, "   For usage, type:   %s  --help\n", prog); }

static void usage(void) {
    fprintf(stderr, " Usage:   %s [options]\n", prog);
    fprintf(stderr, " EG solving on an AMD MI355X GPU: reads an energy game and computes, for every vertex,\n"
                    " the least initial credit of player 0 (2147483647 where player 1 wins).\n");
    fprintf(stderr, " Options:\n");
    fprintf(stderr, "  --help  -h  -?\n\tShow this message.\n");
    fprintf(stderr, "  --input FILENAME\n\tReads from FILENAME instead of stdin (gzipped input accepted).\n");
    fprintf(stderr, "\n Options to assign/modify edge weights:\n");
    fprintf(stderr, "  --unit-weights\n\tSet every weight to 1.\n");
    fprintf(stderr, "  --exp-weights\n\tSet every weight to 0 (as the reference does).\n");
    fprintf(stderr, "  --rnd-weights [[L] U]\n\tWeights drawn with rand() in [L..U] (default L=-U, U=%d).\n",
            DEFAULT_ADD_RND_WEIGHTS_VAL);
    fprintf(stderr, "  --rnd-seed S\n\tSeed for --rnd-weights (default: time based).\n");
    fprintf(stderr, "  --no-weights\n\tKeep the weights of the input (default).\n");
    fprintf(stderr, "\n Solving options:\n");
    fprintf(stderr, "  --gpu\n\tSolve on the GPU (default; the only solver of this build).\n");
    fprintf(stderr, "  --deviceid I\n\tUse GPU I (default 0).\n");
    fprintf(stderr, "  --tb T\n\tAccepted for compatibility (power of 2); workgroup sizes are chosen per kernel.\n");
    fprintf(stderr, "  --eg [N]\n\tWorklist EG algorithm; at most N rounds (default N=|V||E|MG+1).\n");
    fprintf(stderr, "  --eg0 [N]\n\tSame as --eg on the GPU, as in the reference.\n");
    fprintf(stderr, "  --shuffling N\n\tLanes per low-degree vertex, N=1,2,4,8,16,32 (default: chosen from the degrees).\n");
    fprintf(stderr, "  --timeout [S]\n\tStop after about S seconds with partial values (default S=%d).\n",
            DEFAULT_TIMEOUT_SEC);
    fprintf(stderr, "  --no-proofs\n\tPlain value iteration: no proofs that climbing vertices are TOP (same values, many more rounds).\n");
    fprintf(stderr, "  --save-csr FILE\n\tAfter reading, save the game in its internal numbering as a binary cache.\n");
    fprintf(stderr, "  --load-csr FILE\n\tRead the game from a binary cache made by --save-csr instead of the input.\n");
    fprintf(stderr, "\n Output:\n");
    fprintf(stderr, "  --onelineout\n\tSolution in a single text line.\n");
    fprintf(stderr, "  --noout\n\tDo not print the solution.\n");
    fprintf(stderr, "  --printdegrees\n\tAccepted; degree histograms are a CPU-solver feature of the reference.\n");
    fprintf(stderr, "\n Vertex orders (change only the internal numbering and output order):\n");
    fprintf(stderr, "  --sort_n --sort_o --sort_i --sort_oi --sort_a\n\towner; then out-degree; in-degree; both; their sum.\n");
    fprintf(stderr, "\nInput format:\n\t[arena] MAXNODE [NUMEDGES] ;\n\tnode  MAX|MIN  node:weight, ..., node:weight  [\"string\"] ;\n\n");
}

/* src/egsolver.c:886-903 */
static int has_param(int i, int argc, char **argv) {
    if (i >= argc) return 0;
    const char *a = argv[i];
    if ((a[0] == '-' && a[1] == '-') || a[0] == '|' || (a[0] == '-' && (a[1] == 'h' || a[1] == '?'))) return 0;
    return 1;
}

/* src/egsolver.c:907-944 */
static int to_int(const char *s, const char *opt) {
    char *end;
    errno = 0;
    long v = strtol(s, &end, 10);
    if ((errno == ERANGE && (v == LONG_MAX || v == LONG_MIN)) || (errno != 0 && v == 0) || v < INT_MIN ||
        v > INT_MAX || end == s || *end != '\0') {
        fprintf(stderr, "\nCONVERSION ERROR illegal or out-of-range parameter for option: %s\n\n", opt);
        short_usage();
        exit(EXIT_FAILURE);
    }
    return (int)v;
}

static void bad(const char *opt, const char *val) {
    fprintf(stderr, "\nERROR illegal parameter of option: %s %s\n\n", opt, val ? val : "");
    short_usage();
    exit(1);
}

static const char *input_path = NULL;
static int weight_mode = EGS_W_KEEP;
static const char *save_csr_path = NULL, *load_csr_path = NULL;   /* binary game cache */
static double postt = 0;   /* post-parsing seconds (0 when the game comes from the cache) */

static void setconfig(int argc, char **argv) {
    memset(&configuration, 0, sizeof(configuration));
    configuration.computationType = GPU_COMPUTATION;
    configuration.algoritmo = EGS_REF_ALGOR_EG;
    configuration.stdinputsource = 1;
    configuration.onelineout = NO_ONELINEOUT;
    configuration.rndWeightLow = -DEFAULT_ADD_RND_WEIGHTS_VAL;
    configuration.rndWeightHigh = DEFAULT_ADD_RND_WEIGHTS_VAL;
    configuration.rndSeed = NO_USER_RND_SEED;
    configuration.threadsPerBlock = 256;
    configuration.kinfOfParallelism = 0;    /* 0: chosen from the degrees */
    configuration.max_loop_val = -1;
    configuration.timeoutOpt = UNSET_TIMEOUT_OPT;
    configuration.nodesorting = SORT_N;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "--help") || !strcmp(a, "-?") || !strcmp(a, "-h")) {
            usage();
            exit(EXIT_SUCCESS);
        } else if (!strcmp(a, "--cpu")) {
            configuration.computationType = CPU_COMPUTATION;
        } else if (!strcmp(a, "--gpu")) {
            configuration.computationType = GPU_COMPUTATION;
        } else if (!strcmp(a, "--no-proofs")) {
            egs_ref_certify_every = -1;
        } else if (!strcmp(a, "--printdegrees")) {
        } else if (!strcmp(a, "--onelineout")) {
            configuration.onelineout = YES_ONELINEOUT;
        } else if (!strcmp(a, "--noout")) {
            configuration.onelineout = NO_OUTPUT;
        } else if (!strcmp(a, "--deviceid")) {
            if (!has_param(i + 1, argc, argv)) bad("--deviceid", "(missing)");
            configuration.deviceid = to_int(argv[++i], "--deviceid");
            if (configuration.deviceid < 0) bad("--deviceid", argv[i]);
        } else if (!strcmp(a, "--tb")) {
            if (!has_param(i + 1, argc, argv)) bad("--tb", "(missing)");
            int t = to_int(argv[++i], "--tb");
            if (t < 1 || __builtin_popcount((unsigned)t) != 1) bad("--tb", argv[i]);
            configuration.threadsPerBlock = (unsigned)t;
        } else if (!strcmp(a, "--input")) {
            if (!has_param(i + 1, argc, argv)) bad("--input", "(missing)");
            input_path = argv[++i];
            snprintf(configuration.filename, sizeof(configuration.filename), "%s", input_path);
            configuration.stdinputsource = 0;
            FILE *f = fopen(input_path, "rb");
            if (!f) {
                fprintf(stderr, "\nERROR in opening file %s :", input_path);
                perror("");
                short_usage();
                exit(1);
            }
            fclose(f);
        } else if (!strcmp(a, "--save-csr") || !strcmp(a, "--load-csr")) {
            if (!has_param(i + 1, argc, argv)) bad(a, "(missing)");
            if (a[2] == 's') save_csr_path = argv[++i];
            else load_csr_path = argv[++i];
        } else if (!strcmp(a, "--shuffling")) {
            if (has_param(i + 1, argc, argv)) {
                int k = to_int(argv[++i], "--shuffling");
                if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16 && k != 32) bad("--shuffling", argv[i]);
                configuration.kinfOfParallelism = k;
            }
        } else if (!strcmp(a, "--eg") || !strcmp(a, "--eg0")) {
            if (has_param(i + 1, argc, argv)) {
                int k = to_int(argv[++i], a);
                if (k < 1 || k >= INT_MAX) bad(a, argv[i]);
                configuration.max_loop_val = k;
            }
            configuration.algoritmo = !strcmp(a, "--eg") ? EGS_REF_ALGOR_EG : EGS_REF_ALGOR_EG0;
        } else if (!strcmp(a, "--timeout")) {
            configuration.timeoutSeconds = DEFAULT_TIMEOUT_SEC;
            if (has_param(i + 1, argc, argv)) {
                int k = to_int(argv[++i], "--timeout");
                if (k <= 1 || k >= INT_MAX) bad("--timeout", argv[i]);
                configuration.timeoutSeconds = (unsigned)k;
            }
            configuration.timeoutOpt = SET_TIMEOUT_OPT;
        } else if (!strcmp(a, "--unit-weights")) {
            weight_mode = EGS_W_UNIT;
        } else if (!strcmp(a, "--rnd-weights")) {
            if (has_param(i + 1, argc, argv)) {
                int lo = to_int(argv[++i], "--rnd-weights");
                if (has_param(i + 1, argc, argv)) {
                    int hi = to_int(argv[++i], "--rnd-weights");
                    if (hi < lo || hi >= MAX_RAND_NUM || lo < -MAX_RAND_NUM) bad("--rnd-weights", argv[i]);
                    configuration.rndWeightLow = lo;
                    configuration.rndWeightHigh = hi;
                } else {
                    if (lo >= MAX_RAND_NUM || lo < 0) bad("--rnd-weights", argv[i]);
                    configuration.rndWeightLow = -lo;
                    configuration.rndWeightHigh = lo;
                }
            }
            weight_mode = EGS_W_RND;
        } else if (!strcmp(a, "--rnd-seed")) {
            if (!has_param(i + 1, argc, argv)) bad("--rnd-seed", "(missing)");
            configuration.rndSeed = to_int(argv[++i], "--rnd-seed");
            if (configuration.rndSeed < 0 || configuration.rndSeed >= INT_MAX) bad("--rnd-seed", argv[i]);
        } else if (!strcmp(a, "--exp-weights")) {
            weight_mode = EGS_W_EXP;
        } else if (!strcmp(a, "--no-weights")) {
            weight_mode = EGS_W_KEEP;
        } else if (!strcmp(a, "--sort_n")) {
            configuration.nodesorting = SORT_N;
        } else if (!strcmp(a, "--sort_i")) {
            configuration.nodesorting = SORT_I;
        } else if (!strcmp(a, "--sort_o")) {
            configuration.nodesorting = SORT_O;
        } else if (!strcmp(a, "--sort_a")) {
            configuration.nodesorting = SORT_A;
        } else if (!strcmp(a, "--sort_oi")) {
            configuration.nodesorting = SORT_OI;
        } else {
            fprintf(stderr, "\nERROR unknown option: %s\n\n", a);
            short_usage();
            exit(1);
        }
    }
    if (configuration.computationType == CPU_COMPUTATION) {
        fprintf(stderr, "\nERROR --cpu: the reference's CPU solvers are not part of this build; "
                        "this egsolver solves on the GPU (--gpu).\n\n");
        exit(1);
    }
    if (configuration.rndSeed == NO_USER_RND_SEED) {
        struct timeval tv;
        gettimeofday(&tv, NULL);
        configuration.rndSeed = (int)((tv.tv_usec / 1000) % SHRT_MAX) + (int)((tv.tv_sec / 2) % INT_MAX);
    }
    srand((unsigned)configuration.rndSeed);
}

/* ---- vertex order (src/egsolver.c:146-214, src/sorting_criteria.h:6-102) ----------- */
static const uint8_t *s_owner;
static const int *s_out, *s_in;

static int vkey_cmp(int a, int b) {
    if (s_owner[a] != s_owner[b]) return s_owner[a] < s_owner[b] ? -1 : 1;
    int ka = 0, kb = 0, ka2 = 0, kb2 = 0;
    switch (configuration.nodesorting) {
    case SORT_O: ka = s_out[a]; kb = s_out[b]; break;
    case SORT_I: ka = s_in[a]; kb = s_in[b]; break;
    case SORT_OI: ka = s_out[a]; kb = s_out[b]; ka2 = s_in[a]; kb2 = s_in[b]; break;
    case SORT_A: ka = s_out[a] + s_in[a]; kb = s_out[b] + s_in[b]; break;
    default: break;
    }
    if (ka != kb) return ka < kb ? -1 : 1;
    if (ka2 != kb2) return ka2 < kb2 ? -1 : 1;
    return 0;
}

/* Stable merge sort: ties keep read order, as glibc's merge-sort qsort does. */
static void stable_sort(int *a, int *tmp, int n) {
    if (n < 2) return;
    int h = n / 2;
    stable_sort(a, tmp, h);
    stable_sort(a + h, tmp, n - h);
    int i = 0, j = h, k = 0;
    while (i < h && j < n) tmp[k++] = (vkey_cmp(a[j], a[i]) < 0) ? a[j++] : a[i++];
    while (i < h) tmp[k++] = a[i++];
    while (j < n) tmp[k++] = a[j++];
    memcpy(a, tmp, sizeof(int) * (size_t)n);
}

static void on_alarm(int signo) {
    if (signo == SIGALRM) {
        timeout_expired = 1;
        fprintf(stderr, "\nTIMEOUT EXPIRED: forced exit\n");
    }
}

/* Read the .mpg input and build the internal CSR (src/egsolver.c:107-315). */
static void load_from_mpg(void) {
    double t0 = now_s();
    printf("Parsing input (reading %s)...\n", configuration.stdinputsource ? "stdin" : configuration.filename);
    fflush(stdout);
    egs_mpg g;
    char err[512];
    if (egs_mpg_read(input_path, weight_mode, configuration.rndWeightLow, configuration.rndWeightHigh, &g, err,
                     sizeof(err)) != 0) {
        fprintf(stderr, "%s\n", err);
        exit(EXIT_FAILURE);
    }
    const int n = g.n;
    const long m = (long)g.m;
    if (m > INT_MAX) {
        fprintf(stderr, "Error in input. too many edges (%ld)\n", m);
        exit(EXIT_FAILURE);
    }
    num_nodi = counter_nodi = n;
    num_archi = (int)m;
    max_nodo = n - 1;
    max_pesi = g.max_abs_w;
    min_outdegree = g.min_outdeg;
    max_outdegree = g.max_outdeg;
    memcpy(numLowOutDegree, g.low_outdeg, sizeof(numLowOutDegree));
    int *outdeg = (int *)malloc(sizeof(int) * (size_t)n);
    int *indeg = (int *)calloc((size_t)n, sizeof(int));
    for (int r = 0; r < n; r++) outdeg[r] = g.rptr[r + 1] - g.rptr[r];
    for (long e = 0; e < m; e++) indeg[g.read_of_ext[g.rsucc[e]]]++;
    mapping = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    revmapping = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *tmp = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    for (int i = 0; i < n; i++) mapping[i] = i;
    s_owner = g.owner;
    s_out = outdeg;
    s_in = indeg;
    static const char *order_msg[] = {"owner", "owner, then w.r.t. indegree", "owner, then w.r.t. outdegree",
                                      "owner, then w.r.t. outdegree+indegree",
                                      "owner, then w.r.t. outdegree, then w.r.t. indegree"};
    printf("Sorting nodes w.r.t. %s\n", order_msg[configuration.nodesorting]);
    stable_sort(mapping, tmp, n);
    for (int i = 0; i < n; i++) revmapping[mapping[i]] = i;
    free(tmp);
    counter_nodi0 = 0;
    for (int i = 0; i < n; i++) counter_nodi0 += g.owner[i] == 0;
    statistics.inputtime = now_s() - t0;
    printf("Parsing completed (parsing time %lf).\n", statistics.inputtime);

    /* postparsing (src/egsolver.c:236-315) */
    printf("\nStart post-parsing...\n");
    fflush(stdout);
    double t1 = now_s();
    alloca_memoria_host();
    alloca_memoria_device();
    {
        long e = 0;
        for (int i = 0; i < n; i++) {
            host_csrPtrInSuccLists[i] = (int)e;
            const int r = mapping[i];
            for (int j = g.rptr[r]; j < g.rptr[r + 1]; j++) {
                host_csrSuccLists[e] = revmapping[g.read_of_ext[g.rsucc[j]]];
                host_csrPesiArchi[e] = g.rw[j];
                e++;
            }
        }
        host_csrPtrInSuccLists[n] = (int)e;
    }
    {
        long mg = 0;
        for (int i = 0; i < n; i++) {
            int worst = 0;
            for (int e = host_csrPtrInSuccLists[i]; e < host_csrPtrInSuccLists[i + 1]; e++)
                if (-host_csrPesiArchi[e] > worst) worst = -host_csrPesiArchi[e];
            if (mg > INT_MAX - worst) {
                fprintf(stderr, "Error too high weights: %d + %ld --> overflow\n", worst, mg);
                exit(EXIT_FAILURE);
            }
            mg += worst;
        }
        MG_pesi = (int)mg;
    }
    avg_outdegree = (double)m / (double)n;
    double sq = 0;
    for (int i = 0; i < n; i++) sq += ((double)outdeg[i] - avg_outdegree) * ((double)outdeg[i] - avg_outdegree);
    stddev_outdegree = sqrt(sq / n);
    /* the reader's arrays are released here; the entry points free what they own */
    nomeExt_of_nomeInt = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    for (int i = 0; i < n; i++) nomeExt_of_nomeInt[i] = g.ext_of_read[i];
    egs_mpg_free(&g);
    free(outdeg);
    free(indeg);
    postt = now_s() - t1;
    printf("Post-parsing completed (postparsing time %lf).\n", postt);
}

/* The game from a binary cache made by --save-csr: internal numbering, MG and external ids
 * as saved, so the solve and the output are those of the run that saved it. */
static void load_from_cache(void) {
    double t0 = now_s();
    printf("Reading the game cache %s...\n", load_csr_path);
    fflush(stdout);
    egs_csr c;
    char err[512];
    if (egs_csr_load(load_csr_path, &c, err, sizeof(err)) != 0) {
        fprintf(stderr, "%s\n", err);
        exit(EXIT_FAILURE);
    }
    const int n = c.n;
    const long m = (long)c.m;
    num_nodi = counter_nodi = n;
    num_archi = (int)m;
    max_nodo = n - 1;
    max_pesi = c.max_abs_w;
    counter_nodi0 = c.n0;
    min_outdegree = INT_MAX;
    max_outdegree = 0;
    memset(numLowOutDegree, 0, sizeof(numLowOutDegree));
    for (int i = 0; i < n; i++) {
        const int d = c.ptr[i + 1] - c.ptr[i];
        if (d < min_outdegree) min_outdegree = d;
        if (d > max_outdegree) max_outdegree = d;
        if (d < 4) numLowOutDegree[d]++;
    }
    mapping = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    revmapping = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    nomeExt_of_nomeInt = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    for (int i = 0; i < n; i++) {
        mapping[i] = revmapping[i] = i;
        nomeExt_of_nomeInt[i] = c.ext_of_int[i];
    }
    alloca_memoria_host();
    alloca_memoria_device();
    memcpy(host_csrPtrInSuccLists, c.ptr, sizeof(int) * (size_t)(n + 1));
    memcpy(host_csrSuccLists, c.succ, sizeof(int) * (size_t)m);
    memcpy(host_csrPesiArchi, c.w, sizeof(int) * (size_t)m);
    MG_pesi = c.mg;
    avg_outdegree = (double)m / (double)n;
    double sq = 0;
    for (int i = 0; i < n; i++) {
        const double d = (double)(c.ptr[i + 1] - c.ptr[i]) - avg_outdegree;
        sq += d * d;
    }
    stddev_outdegree = sqrt(sq / n);
    egs_csr_free(&c);
    statistics.inputtime = now_s() - t0;
    printf("Cache read (reading time %lf).\n", statistics.inputtime);
}

int main(int argc, char **argv) {
    prog = argv[0];
    setconfig(argc, argv);
    memset(&statistics, 0, sizeof(statistics));

    char host[256] = "";
    gethostname(host, sizeof(host));
    printf("Process stats:\n\tProcess: PID=%d  running on %s\n", getpid(), host);
    printf("\tCommand line: ");
    for (int i = 0; i < argc; i++) printf("%s ", argv[i]);
    printf("\n");
    fflush(stdout);

    int device_count = 0;
    checkDevice(&device_count, &configuration);
    configuration.deviceCount = (unsigned)device_count;
    {
        egs_device_info info;
        egs_device_query(configuration.deviceid, &info);
        printf("GPU stats:\n\t%d HIP devices detected.\n", device_count);
        printf("\tSelect device number %d  name: %s\n", configuration.deviceid, configuration.devicename);
        printf("\tWavefront size: %d\n", configuration.warpSize);
        printf("\tGPU clock: %.0f MHz\n", (float)configuration.clockRate * 1e-3f);
        printf("\tDevice memory: %.0f MB\n", (double)configuration.deviceProp_totalGlobalMem / 1048576.0);
        printf("\tECC enabled: %s\n", configuration.ECCEnabled ? "yes" : "no");
        printf("\tHIP Driver Version / Runtime Version   %d / %d\n\n", info.driver_version, info.runtime_version);
    }

    if (load_csr_path) {
        load_from_cache();
    } else {
        load_from_mpg();
    }
    if (save_csr_path) {
        int *ext = (int *)malloc(sizeof(int) * (size_t)num_nodi);
        for (int i = 0; i < num_nodi; i++) ext[i] = nomeExt_of_nomeInt[mapping[i]];
        char cerrbuf[512];
        if (egs_csr_save(save_csr_path, num_nodi, counter_nodi0, num_archi, MG_pesi, host_csrPtrInSuccLists,
                         host_csrSuccLists, host_csrPesiArchi, ext, cerrbuf, sizeof(cerrbuf)) != 0) {
            fprintf(stderr, "%s\n", cerrbuf);
            exit(EXIT_FAILURE);
        }
        free(ext);
        printf("Game saved to %s\n", save_csr_path);
    }

    printf("\nInput stats:\n\tNodes: %d (%d+%d)\n\tEdges: %d\n\tMax node Id: %d\n", num_nodi, counter_nodi0,
           counter_nodi - counter_nodi0, num_archi, max_nodo);

    printf("\nInput stats:\n\tNodes: %d (%d+%d)\n\tEdges: %d\n\tMax node Id: %d\n", num_nodi, counter_nodi0,
           counter_nodi - counter_nodi0, num_archi, max_nodo);
    printf("\tMin out-degree: %d\n", min_outdegree);
    printf("\tMax out-degree: %d\n\tMax weight(abs): %d\n", max_outdegree, max_pesi);
    printf("\tAvg out-degree: %.3lf\n\tStd dev: %.3lf\n\tRel std dev: %.2lf\n", avg_outdegree, stddev_outdegree,
           100 * stddev_outdegree / fabs(avg_outdegree));
    printf("\tLow out-degree: 0: %d \t1: %d \t2: %d \t3: %d\n", numLowOutDegree[0], numLowOutDegree[1],
           numLowOutDegree[2], numLowOutDegree[3]);

    printf("\nPreparing to solve...\n");
    timeout_expired = 0;
    if (configuration.timeoutOpt == SET_TIMEOUT_OPT) {
        struct sigaction act;
        memset(&act, 0, sizeof(act));
        act.sa_handler = on_alarm;
        sigfillset(&act.sa_mask);
        sigaction(SIGALRM, &act, NULL);
        printf("Setting %u seconds timeout...\n", configuration.timeoutSeconds);
        alarm(configuration.timeoutSeconds);
    }
    fflush(stdout);

    copia_dati_su_device();
    gpu_solver();
    printf("------------------------\nTiming:\n");
    printf("Parsing time: %lf sec \n", statistics.inputtime);
    printf("Postparsing time: %lf sec \n", postt);
    printf("Solving time: %lf sec \n", statistics.solvingtime / 1000);
    printf("Total time: %lf sec \n", postt + statistics.inputtime + statistics.solvingtime / 1000);
    printf("Nodes per second: %lf \n", (double)statistics.processedNodes / (statistics.solvingtime / 1000));
    printf("------------------------\n");

    if (configuration.onelineout != NO_OUTPUT) {
        printf("\nSolution:\n");
        if (configuration.onelineout == YES_ONELINEOUT) {
            for (int i = 0; i < num_nodi; i++) printf("V(%d)=%d\t", nomeExt_of_nomeInt[mapping[i]], host_ResNodeValues1[i]);
            printf("\n");
        } else {
            for (int i = 0; i < num_nodi; i++) printf("%d)\t%d\n", nomeExt_of_nomeInt[mapping[i]], host_ResNodeValues1[i]);
        }
    } else {
        printf("\nSolution output omitted.\n");
    }
    dealloca_memoria_host();
    dealloca_memoria_device();
    egs_device_reset(configuration.deviceid);
    fflush(stdout);
    return EXIT_SUCCESS;
}
