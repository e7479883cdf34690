/*
 * egs_dev_solver.h -- the reference's GPU entry points, provided by libegs_dev_solver.so.
 *
 * fvella/EGSolver's host driver reaches its GPU path only through these functions
 * (src/dev_solver.h:9-22, src/utils.cu:38-73) and shares the game with them through
 * globals (src/egsolver.h:24-119, declared extern at src/dev_solver.cu:17-75).  This
 * library implements the same functions over libegs.so (include/egs.h), reading and
 * writing the same globals, so a driver built against the reference header links
 * against it unchanged.  The globals are defined by the driver (this repository's
 * egsolver_amd/csrc/host/egsolver.c, or the reference's src/egsolver.c).
 *
 * The two structs below restate the layouts of the reference's `config` and `stat`
 * (src/common.h:314-371) field for field: the library reads configuration.algoritmo,
 * .kinfOfParallelism, .max_loop_val and .deviceid, and writes statistics.solvingtime
 * (milliseconds, as src/dev_solver.cu:132 does) and statistics.processedNodes.
 */
#ifndef EGS_DEV_SOLVER_H
#define EGS_DEV_SOLVER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EGS_REF_FILENAME_LEN 256

/* src/common.h:314-321 */
typedef struct egs_ref_stat {
    double inputtime;
    double alloctime;
    double solvingtime;
    long processedNodes;
    unsigned int numkernelscalls;
    unsigned long device_usedGlobalMem;
} egs_ref_stat;

/* src/common.h:325-371 */
typedef struct egs_ref_config {
    int computationType;
    int algoritmo;
    int stdinputsource;
    char filename[EGS_REF_FILENAME_LEN];
    int onelineout;
    int add_weights_mode;
    int rndWeightLow;
    int rndWeightHigh;
    int rndSeed;
    int deviceid;
    char devicename[256];
    unsigned short warpSize;
    int capabilityMajor;
    int capabilityMinor;
    size_t deviceProp_totalGlobalMem;
    size_t sharedMemPerBlock;
    unsigned int maxThreadsPerBlock;
    int clockRate;
    int ECCEnabled;
    unsigned int threadsPerBlock;
    unsigned int deviceCount;
    int kinfOfParallelism;
    int shuffleThreshold;
    int shuffleSplit_val;
    int shuffleSplit_low;
    int shuffleSplit_up;
    int shuffleSplit_index;
    int shuffleSplit_soglia;
    int shuffleSplit_double;
    long max_loop_val;
    unsigned int nodesorting;
    unsigned int timeoutOpt;
    unsigned int timeoutSeconds;
    int print_degree_hist;
} egs_ref_config;

/* Constants of src/common.h used by the entry points. */
#define EGS_REF_ALGOR_EG 6                 /* common.h:161 */
#define EGS_REF_ALGOR_EG0 7                /* common.h:162 */
#define EGS_REF_KINDPARALLELISM_VERTEXPAR 1 /* common.h:195-200: 1,2,4,...,32 lanes */

/* ---- the reference's seven entry points (src/dev_solver.h:9-22) ------------------- */

/* Pinned host block [ptr n+1 | succ m | w m | val1 n | val2 (EG0: n, else alias) |
 * flag 1 | aux m] and the host transpose block; zeroes the values
 * (src/dev_solver.cu:152-187).  Called before the driver fills the CSR. */
int alloca_memoria_host(void);
/* Device workspace for the game (src/dev_solver.cu:193-239). */
int alloca_memoria_device(void);
/* Frees the pinned block, the transpose block and the parser arrays
 * (src/dev_solver.cu:241-260). */
int dealloca_memoria_host(void);
int dealloca_memoria_device(void);
/* Upload of the CSR (src/dev_solver.cu:142-149). */
int copia_dati_su_device(void);
/* Transpose + fixpoint on the GPU; fills host_ResNodeValues1[0..num_nodi),
 * statistics.solvingtime (ms) and statistics.processedNodes (src/dev_solver.cu:81-139).
 * configuration.kinfOfParallelism = 1,2,...,32 selects the lanes per low-degree vertex
 * (the reference's --shuffling K, src/egsolver.c:699-731); timeout_expired stops the
 * run with partial values. */
void gpu_solver(void);
/* No reference counterpart.  Rounds between attempts to prove climbing vertices TOP,
 * passed to egs_options.certify_every by gpu_solver: 0 (the default) = the library's
 * default, < 0 = never (plain value iteration, as the reference runs).  The driver's
 * --no-proofs sets it to -1. */
extern int egs_ref_certify_every;
/* Defined by the driver (src/egsolver.c:64-90); used here for the phase cap. */
long aggiorna_max_loop(long Narchi, long Nnodi, long MGpesi, long maxloop);

/* ---- device helpers the driver calls (src/utils.cu:38-73) ------------------------- */

/* Fills device name, warp size (64), limits, clock and ECC into *conf; exits with an
 * error when conf->deviceid is not an available device. */
void checkDevice(int *deviceCount, egs_ref_config *conf);
/* Exits with the message when the HIP runtime holds an error. */
void checkCUDAError(const char *msg);

#ifdef __cplusplus
}
#endif
#endif /* EGS_DEV_SOLVER_H */
