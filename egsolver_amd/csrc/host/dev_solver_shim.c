/*
 * dev_solver_shim.c -- the reference's GPU entry points over libegs.so
 * (libegs_dev_solver.so; declarations in include/egs_dev_solver.h).
 *
 * Same calling contract as src/dev_solver.cu: no arguments, the game and the options
 * come through the driver's globals, results go back into host_ResNodeValues1 and
 * `statistics`.  Errors on the device path print a message and exit, as the reference
 * does under CHECKCUDASAFE (src/common.h:79-85).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "egs.h"
#include "egs_dev_solver.h"

/* ---- globals owned by the driver (src/egsolver.h:24-119) ---------------------------- */
extern int num_nodi, num_archi, counter_nodi0, MG_pesi, max_pesi;
extern int *host_allData, *host_csrPtrInSuccLists, *host_csrSuccLists, *host_csrPesiArchi;
extern int *host_ResNodeValues1, *host_ResNodeValues2, *host_ResNodeValuesAux, *host_flag;
extern int *host_csrDataArchiAux, *host_revSuccLists;
extern int *host_transData, *host_cscPtrInPredLists, *host_cscPredLists, *host_cscPesiArchiPred;
extern int *hdev_allData, *hdev_ResNodeValues1;
extern unsigned int timeout_expired;
extern egs_ref_config configuration;
extern egs_ref_stat statistics;
/* parser arrays released by dealloca_memoria_host (src/dev_solver.cu:245-257) */
extern int *csrPtrInSuccLists, *nodePriority, *csrPesiArchi, *csrSuccLists;
extern int *nomeExt_of_nomeInt, *nomeInt_of_nomeExt, *mapping, *revmapping;
extern char *nodeOwner, *nodeFlags;
extern char **nodeName;

static egs_solver *g_solver = NULL;
int egs_ref_certify_every = 0;

static void die(const char *what) {
    fprintf(stderr, "ERROR: %s : %s\n", what, egs_last_error());
    exit(-1);
}

int alloca_memoria_host(void) {
    const long n = num_nodi, m = num_archi;
    const long eg0 = configuration.algoritmo == EGS_REF_ALGOR_EG0;
    const long total = (n + 1) + m + m + n + (eg0 ? n : 0) + 1 + m;
    void *p = NULL;
    if (egs_host_alloc(&p, sizeof(int) * (size_t)total) != EGS_OK) die("cudaMallocHost: host_allData[]");
    host_allData = (int *)p;
    host_csrPtrInSuccLists = host_allData;
    host_csrSuccLists = host_allData + n + 1;
    host_csrPesiArchi = host_csrSuccLists + m;
    host_ResNodeValues1 = host_csrPesiArchi + m;
    host_ResNodeValues2 = eg0 ? host_ResNodeValues1 + n : host_ResNodeValues1;
    host_flag = host_ResNodeValues2 + n;
    host_csrDataArchiAux = host_flag + 1;
    memset(host_ResNodeValues1, 0, sizeof(int) * (size_t)n);
    host_ResNodeValuesAux = NULL;
    host_transData = (int *)malloc(sizeof(int) * (size_t)(n + 1 + 2 * m));
    if (!host_transData) die("allocazione host_transData");
    host_cscPtrInPredLists = host_transData;
    host_cscPredLists = host_transData + n + 1;
    host_cscPesiArchiPred = host_cscPredLists + m;
    return 0;
}

int alloca_memoria_device(void) {
    if (g_solver) { egs_solver_destroy(g_solver); g_solver = NULL; }
    if (egs_solver_create(configuration.deviceid, num_nodi, num_archi, &g_solver) != EGS_OK)
        die("cudaMalloc: &hdev_allData[]");
    hdev_allData = NULL;
    hdev_ResNodeValues1 = (int *)egs_solver_device_values(g_solver);
    return 0;
}

int copia_dati_su_device(void) {
    if (!g_solver) die("copia_dati_su_device: no device workspace");
    if (egs_solver_load(g_solver, EGS_MEM_HOST, num_nodi, counter_nodi0, num_archi, MG_pesi,
                        host_csrPtrInSuccLists, host_csrSuccLists, host_csrPesiArchi) != EGS_OK)
        die("cudaMemcpyHostToDevice dev_allData");
    return 0;
}

void gpu_solver(void) {
    if (!g_solver) die("gpu_solver: no device workspace");
    egs_options opt;
    memset(&opt, 0, sizeof(opt));
    /* --shuffling K (1,2,4,...,32): lanes per vertex; EG0 on the GPU runs the vertex-
     * parallel EG as the reference does (src/dev_solver.cu:90-92). */
    int k = configuration.kinfOfParallelism;
    if (configuration.algoritmo == EGS_REF_ALGOR_EG0 || k < 1 || k > 64 || (k & (k - 1))) k = 0;
    opt.group = k;
    opt.max_phases = aggiorna_max_loop((long)num_archi, (long)num_nodi, (long)MG_pesi,
                                       configuration.max_loop_val);
    opt.stop_flag = (const volatile uint32_t *)&timeout_expired;
    opt.certify_every = egs_ref_certify_every;
    printf("Running EG on GPU. (MG_pesi=%d max_loop=%ld num nodes=%d max weight=%d)\n", MG_pesi,
           (long)opt.max_phases, num_nodi, max_pesi);
    fflush(stdout);
    if (egs_solver_prepare(g_solver, &opt) != EGS_OK) die("gpu_solver: prepare");
    egs_stats st;
    if (egs_solver_iterate(g_solver, -1, &st) < 0) die("gpu_solver: iterate");
    printf("End EG on GPU after %ld loops (each loop involves one or more active nodes). Processed nodes %ld\n",
           (long)st.phases, (long)st.processed);
    statistics.solvingtime = st.ms_prepare + st.ms_iterate;   /* ms, src/dev_solver.cu:132 */
    statistics.processedNodes = (long)st.processed;
    if (egs_solver_read_values(g_solver, host_ResNodeValues1) != EGS_OK)
        die("cudaMemcpyDeviceToHost host_ResNodeValues1");
}

int dealloca_memoria_host(void) {
    egs_host_free(host_allData);
    host_allData = NULL;
    free(host_transData);
    host_transData = NULL;
    free(csrPtrInSuccLists);
    free(nodePriority);
    free(nodeOwner);
    free(nodeFlags);
    free(csrPesiArchi);
    free(csrSuccLists);
    free(nomeExt_of_nomeInt);
    free(nomeInt_of_nomeExt);
    free(mapping);
    free(revmapping);
    if (nodeName) {
        for (int i = 0; i < num_nodi; i++) free(nodeName[i]);
        free(nodeName);
    }
    csrPtrInSuccLists = nodePriority = csrPesiArchi = csrSuccLists = NULL;
    nomeExt_of_nomeInt = nomeInt_of_nomeExt = mapping = revmapping = NULL;
    nodeOwner = nodeFlags = NULL;
    nodeName = NULL;
    return 0;
}

int dealloca_memoria_device(void) {
    if (g_solver) egs_solver_destroy(g_solver);
    g_solver = NULL;
    hdev_ResNodeValues1 = NULL;
    return 0;
}

void checkDevice(int *deviceCount, egs_ref_config *conf) {
    egs_device_info info;
    int rc = egs_device_query(conf->deviceid, &info);
    *deviceCount = info.device_count;
    if (rc != EGS_OK) {
        fprintf(stderr, "checkDevice returned: >%s<\n", egs_last_error());
        exit(EXIT_FAILURE);
    }
    conf->deviceProp_totalGlobalMem = (size_t)info.total_mem;
    conf->maxThreadsPerBlock = (unsigned int)info.max_threads_per_block;
    conf->sharedMemPerBlock = (size_t)info.shared_mem_per_block;
    conf->warpSize = (unsigned short)info.warp_size;
    conf->capabilityMajor = info.cap_major;
    conf->capabilityMinor = info.cap_minor;
    conf->clockRate = info.clock_khz;
    conf->ECCEnabled = info.ecc;
    snprintf(conf->devicename, sizeof(conf->devicename), "%.200s (%.40s)", info.name, info.arch);
}

void checkCUDAError(const char *msg) {
    char buf[256];
    if (egs_runtime_error(buf, sizeof(buf)) != 0) {
        fprintf(stderr, "HIP error: >%s<: >%s<. Exiting\n", msg, buf);
        exit(-1);
    }
}
