// egs_api.hip -- C-ABI of libegs.so: device queries and the one-game solver object.
//
// The solver object owns one device workspace.  The round loop keeps the frontier size
// on the device: rounds are enqueued in batches and the host reads one counter per
// batch through pinned memory, one batch behind, so the GPU never waits for the host
// between rounds (the reference blocks on thrust::remove's count twice per round,
// src/dev_EG_alg.cu:346-373).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>

#include "../../include/egs.h"
#include "egs_device.h"
#include "egs_internal.h"
#include "egs_pack.h"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(EGS_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                                \
    } while (0)

int check_gfx950(int device) {
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, device));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(EGS_E_HIP, "device %d is %s; libegs.so is built for gfx950 only", device,
                    p.gcnArchName);
    return EGS_OK;
}

// Loop bound of src/egsolver.c:64-90: a negative request means n*m*MG+1, saturated to
// LONG_MAX-1 on overflow.
int64_t egs_max_loop_c(int64_t m, int64_t n, int64_t mg, int64_t maxloop) {
    if (maxloop >= 0) return maxloop;
    if (mg > (INT64_MAX - 1) / (n * m)) return INT64_MAX - 1;
    return n * m * mg + 1;
}

// Offsets (in ints) inside the solver's pinned readback block.
constexpr int kPinProbe = 2048;   // 2 x kProbeStride: per-batch counter probes
constexpr int kProbeStride = 512;
constexpr int kPinDeg = 256;
constexpr int kPinInit = 320;     // kListCounters ints
constexpr int kPinConv = 448;     // kListCounters ints
constexpr int kPinStats = 640;    // 6 x u64
constexpr int kPinViol = 704;     // 1 x u64
constexpr int kPinCert = 1024;    // certification readbacks (2 x 64 + 4 + 64 ints)
constexpr int kPinOne = 1281;     // the constant 1 (raises ring[12])
constexpr int kProbeWant = 100;   // offset in a probe slot of ring[13] when it needs a copy of its own
// The transpose costs about ten dense rounds to build, and a sparse round saves at most one
// dense round: build it once this many rounds have wanted it (a proof usually leaves only a
// round or two, which do not pay for it).
constexpr int kSparseWantRounds = 6;
constexpr int kCertFirstCap = 64;     // longest-path passes per candidate at the first attempt
constexpr int kCertCandidates = 8;    // gap thresholds tried per attempt (plus the last-round rise set)
constexpr int kCertEvery = 7;         // default rounds before the first proof attempt
constexpr int kGraphSizes = 6;        // captured batches of 1, 2, 4, ..., 32 rounds
constexpr int kStopBatch = 4;         // rounds queued ahead of the host when a stop flag is polled

double egs_now_ms() {
    return 1e-6 * (double)std::chrono::duration_cast<std::chrono::nanoseconds>(
                      std::chrono::steady_clock::now().time_since_epoch()).count();
}

int pow2ceil(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Byte offsets of every array in a solver's device workspace (one block), for games up to
// n_cap vertices and m_cap edges.  constexpr so that the size of the no-allocation
// workspace below is known at compile time.
struct Layout {
    size_t ptr, edges, cptr, cpred, val, stamp, l0, l1, dbig, dhuge, sbig, shuge, lcnt, mode, ring, rounds, misc,
        tile, fill, stats, statsum, sigma, h, hist, bfcnt, bits, hacc, hcnt, hbest, hflag, tcounts, toffs, tscan,
        titems, slen, sbase, sedges, bigmask, total;
    int64_t tcounts_cap;
    int sub_cap;
    bool phased;
};

constexpr size_t align256c(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr Layout make_layout(int32_t n_cap, int64_t m_cap) {
    Layout L{};
    const size_t n1 = (size_t)n_cap + 1, m = (size_t)m_cap;
    const size_t tiles = (size_t)(n_cap + 4095) / 4096 + 1;
    const int64_t sc = (2 * (int64_t)n_cap + egs::kSub - 1) / egs::kSub;
    L.sub_cap = sc > 256 ? (int)sc : 256;
    const size_t list_ints = (size_t)egs::kSub * (size_t)L.sub_cap + n1;
    size_t off = 0;
    auto take = [&off](size_t bytes) {
        const size_t o = off;
        off += align256c(bytes);
        return o;
    };
    L.ptr = take(4 * n1);
    L.edges = take(8 * m);
    L.cptr = take(4 * n1);
    L.cpred = take(4 * m);
    L.val = take(4 * n1);
    L.stamp = take(4 * n1);
    L.l0 = take(4 * list_ints);
    L.l1 = take(4 * list_ints);
    L.dbig = take(4 * n1);
    L.dhuge = take(4 * n1);
    L.sbig = take(4 * n1);
    L.shuge = take(4 * n1);
    L.lcnt = take(4 * 3 * egs::kListCounters);
    L.mode = take(4 * 4);
    L.ring = take(4 * 16);
    L.rounds = take(4 * 16);
    L.misc = take(8 * 16);
    L.tile = take(4 * tiles);
    L.fill = take(4 * n1);
    L.stats = take(8 * 4 * (size_t)egs::kStatSlots);
    L.statsum = take(8 * 8);
    L.sigma = take(4 * n1);
    L.h = take(8 * n1);
    L.hist = take(4 * 256);
    L.bfcnt = take(4 * 64);
    L.bits = take(4 * 3 * (n1 / 32 + n1 / 1024 + 4));
    L.hacc = take(4 * n1);
    L.hcnt = take(4 * n1);
    L.hbest = take(8 * n1);
    L.hflag = take(4 * n1);
    // sorted transpose: per (edge tile, destination bucket) counts, bucket sizes and starts,
    // and the edges as (source, destination) items grouped by bucket
    const int64_t nb_cap = std::min<int64_t>(4096, ((int64_t)n_cap + 2047) / 2048);
    L.tcounts_cap = ((m_cap + egs::kTrTileEdges - 1) / egs::kTrTileEdges) * nb_cap;
    L.tcounts = take(4 * ((size_t)L.tcounts_cap + 1));
    L.toffs = take(4 * ((size_t)L.tcounts_cap + 1));
    L.tscan = take(4 * ((size_t)L.tcounts_cap / 4096 + 2));
    L.titems = take(8 * m);
    // phased dense rounds (large games): per tile and slice the vertices' edge counts (a
    // byte each) and the wave bases, and the edges in (tile, slice, vertex) order
    L.phased = n_cap >= egs::kPhaseMinN;
    const size_t ptiles = (size_t)n_cap / egs::kSliceTile + 1;
    L.slen = L.phased ? take(ptiles * egs::kMaxPhases * egs::kSliceTile) : 0;
    L.sbase = L.phased ? take(4 * ptiles * egs::kMaxPhases * egs::kSliceWaves) : 0;
    L.sedges = L.phased ? take(8 * m) : 0;
    L.bigmask = L.phased ? take(8 * ((size_t)n_cap / 64 + 2)) : 0;
    L.total = off;
    return L;
}

}  // namespace

struct egs_solver {
    int device = 0;
    int cus = 0;                      // compute units of the device
    hipStream_t stream = nullptr;
    bool csc_built = false;           // the transpose exists for the loaded game (ring[12] raised)
    int32_t n_cap = 0;
    int64_t m_cap = 0;
    void *block = nullptr;
    size_t block_bytes = 0;
    int *d_ptr = nullptr, *d_cptr = nullptr, *d_cpred = nullptr, *d_val = nullptr;
    int *d_stamp = nullptr, *d_list[2] = {nullptr, nullptr}, *d_dbig = nullptr, *d_dhuge = nullptr;
    int *d_sbig = nullptr, *d_shuge = nullptr, *d_lcnt = nullptr, *d_mode = nullptr;
    int *d_ring = nullptr, *d_rounds = nullptr, *d_misc = nullptr, *d_tile = nullptr, *d_fill = nullptr;
    int sub_cap = 0;
    int *d_sigma = nullptr, *d_hist = nullptr, *d_bfcnt = nullptr;   // d_sigma: rise stamps
    unsigned *d_bits = nullptr;     // three bitmaps of raised potentials (proof passes)
    int *d_hacc = nullptr, *d_hcnt = nullptr;   // shared lifts of the vertices above huge_deg
    long long *d_hbest = nullptr;               // shared proof passes of the same vertices
    int *d_hflag = nullptr;
    int *d_tcounts = nullptr, *d_toffs = nullptr, *d_tscan = nullptr;   // sorted transpose
    int2 *d_titems = nullptr;
    int64_t tcounts_cap = 0;
    uint8_t *d_slen = nullptr;                  // phased dense rounds (large n_cap)
    int *d_sbase = nullptr;
    int2 *d_sedges = nullptr;
    unsigned long long *d_bigmask = nullptr;
    long long *d_h = nullptr;
    int wmax = 0;
    int cert_every = 32, cert_gap = 32, cert_max_passes = 1024;
    int cert_cap = 64;        // longest-path passes per candidate; doubles while attempts run out of passes
    int cert_pass_hint = 0;   // passes the run's last converged candidate needed
    int64_t next_cert = 0;
    int2 *d_edges = nullptr;
    unsigned long long *d_stats = nullptr, *d_statsum = nullptr;
    int *h_probe = nullptr;   // pinned: per-batch counter probes + misc readbacks
    int probe_want[2] = {kProbeWant, kProbeWant};   // where each probe slot holds ring[13]
    hipEvent_t ev_probe[2] = {nullptr, nullptr};
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    hipEvent_t ev_pa = nullptr, ev_pb = nullptr;   // prepare's span, read after a later wait
    bool prep_open = false;   // prepare's time and initial frontier not read yet

    egs::DevGame g{};
    egs::LaunchShape sh{};
    // batches of 2^k rounds (k < kGraphSizes) captured once per prepare and replayed
    hipGraphExec_t graph[kGraphSizes] = {};
    egs::DevGame graph_g{};           // the game and shape the graphs were captured with
    egs::LaunchShape graph_sh{};
    bool use_graphs = false;
    bool no_alloc = false;            // the no-allocation entry point's solver: never graphs
    bool loaded = false, prepared = false;
    int max_outdeg = 0;
    int64_t next_round = 1;
    int64_t phase_cap = 0;
    int batch_rounds = 32;
    const volatile uint32_t *stop_flag = nullptr;
    egs_stats st{};
};

// Point the solver's arrays into a workspace laid out by make_layout.
static void bind_layout(egs_solver *s, char *b, const Layout &L) {
    s->sub_cap = L.sub_cap;
    s->d_ptr = (int *)(b + L.ptr);
    s->d_edges = (int2 *)(b + L.edges);
    s->d_cptr = (int *)(b + L.cptr);
    s->d_cpred = (int *)(b + L.cpred);
    s->d_val = (int *)(b + L.val);
    s->d_stamp = (int *)(b + L.stamp);
    s->d_list[0] = (int *)(b + L.l0);
    s->d_list[1] = (int *)(b + L.l1);
    s->d_dbig = (int *)(b + L.dbig);
    s->d_dhuge = (int *)(b + L.dhuge);
    s->d_sbig = (int *)(b + L.sbig);
    s->d_shuge = (int *)(b + L.shuge);
    s->d_lcnt = (int *)(b + L.lcnt);
    s->d_mode = (int *)(b + L.mode);
    s->d_ring = (int *)(b + L.ring);
    s->d_rounds = (int *)(b + L.rounds);
    s->d_misc = (int *)(b + L.misc);
    s->d_tile = (int *)(b + L.tile);
    s->d_fill = (int *)(b + L.fill);
    s->d_stats = (unsigned long long *)(b + L.stats);
    s->d_statsum = (unsigned long long *)(b + L.statsum);
    s->d_sigma = (int *)(b + L.sigma);
    s->d_h = (long long *)(b + L.h);
    s->d_hist = (int *)(b + L.hist);
    s->d_bfcnt = (int *)(b + L.bfcnt);
    s->d_bits = (unsigned *)(b + L.bits);
    s->d_hacc = (int *)(b + L.hacc);
    s->d_hcnt = (int *)(b + L.hcnt);
    s->d_hbest = (long long *)(b + L.hbest);
    s->d_hflag = (int *)(b + L.hflag);
    s->d_tcounts = (int *)(b + L.tcounts);
    s->d_toffs = (int *)(b + L.toffs);
    s->d_tscan = (int *)(b + L.tscan);
    s->d_titems = (int2 *)(b + L.titems);
    s->tcounts_cap = L.tcounts_cap;
    if (L.phased) {
        s->d_slen = (uint8_t *)(b + L.slen);
        s->d_sbase = (int *)(b + L.sbase);
        s->d_sedges = (int2 *)(b + L.sedges);
        s->d_bigmask = (unsigned long long *)(b + L.bigmask);
    }
}

// Timing and probes use events when the solver owns them; a solver without events (the
// no-allocation entry point) times on the host and waits on the whole stream instead.
static hipError_t mark_start(egs_solver *s, double *host_t0) {
    if (s->ev_a) return hipEventRecord(s->ev_a, s->stream);
    *host_t0 = egs_now_ms();
    return hipSuccess;
}
static hipError_t mark_stop(egs_solver *s, double host_t0, float *ms) {
    if (s->ev_b) {
        hipError_t e = hipEventRecord(s->ev_b, s->stream);
        if (e == hipSuccess) e = hipEventSynchronize(s->ev_b);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, s->ev_a, s->ev_b);
        return e;
    }
    hipError_t e = hipStreamSynchronize(s->stream);
    *ms = (float)(egs_now_ms() - host_t0);
    return e;
}
static hipError_t probe_mark(egs_solver *s, int slot) {
    return s->ev_probe[slot] ? hipEventRecord(s->ev_probe[slot], s->stream) : hipSuccess;
}
static hipError_t probe_wait(egs_solver *s, int slot) {
    return s->ev_probe[slot] ? hipEventSynchronize(s->ev_probe[slot]) : hipStreamSynchronize(s->stream);
}

// Items in a list from its counters (sub-queues clamp at sub_cap; the rest overflowed).
static int64_t list_total(const egs_solver *s, const int *c) {
    int64_t t = 0;
    for (int k = 0; k < egs::kSub; k++) t += std::min(c[k], s->sub_cap);
    return t + c[egs::kSub];
}

// prepare's time and initial frontier size, once its closing event has passed
static hipError_t close_prepare(egs_solver *s) {
    if (!s->prep_open) return hipSuccess;
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(s->ev_pb);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, s->ev_pa, s->ev_pb);
    if (e != hipSuccess) return e;
    s->st.ms_prepare = ms;
    s->st.initial_active = (int32_t)list_total(s, s->h_probe + kPinInit);
    s->prep_open = false;
    return hipSuccess;
}

static void drop_graphs(egs_solver *s) {
    for (auto &x : s->graph)
        if (x) {
            (void)hipGraphExecDestroy(x);
            x = nullptr;
        }
}

// One batch of `rounds` rounds (k_advance last) on the solver stream.
static hipError_t launch_batch(egs_solver *s, int rounds) {
    for (int j = 0; j < rounds; j++) {
        hipError_t e = egs::launch_round(s->g, s->sh, j, s->stream);
        if (e != hipSuccess) return e;
    }
    return egs::launch_advance(s->g, rounds, s->stream);
}

// The same batch as a graph of 2^k rounds, captured on first use after a prepare.
static hipError_t replay_batch(egs_solver *s, int k) {
    if (!s->graph[k]) {
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        const hipError_t le = launch_batch(s, 1 << k);
        e = hipStreamEndCapture(s->stream, &gr);
        if (le != hipSuccess) e = le;
        if (e == hipSuccess) e = hipGraphInstantiate(&s->graph[k], gr, nullptr, nullptr, 0);
        if (gr) (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
            s->graph[k] = nullptr;
            return e;
        }
    }
    return hipGraphLaunch(s->graph[k], s->stream);
}

extern "C" {

const char *egs_last_error(void) { return g_err; }

int egs_device_query(int device, egs_device_info *out) {
    if (!out) return fail(EGS_E_ARG, "egs_device_query: null output");
    memset(out, 0, sizeof(*out));
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    out->device_count = count;
    if (device < 0 || device >= count)
        return fail(EGS_E_ARG, "device %d not available (%d devices)", device, count);
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, device));
    snprintf(out->name, sizeof(out->name), "%s", p.name);
    snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
    out->warp_size = p.warpSize;
    out->max_threads_per_block = p.maxThreadsPerBlock;
    out->multiprocessors = p.multiProcessorCount;
    out->total_mem = (int64_t)p.totalGlobalMem;
    out->shared_mem_per_block = (int64_t)p.sharedMemPerBlock;
    out->clock_khz = p.clockRate;
    out->ecc = p.ECCEnabled;
    int dv = 0, rv = 0;
    HIP_TRY(hipDriverGetVersion(&dv));
    HIP_TRY(hipRuntimeGetVersion(&rv));
    out->driver_version = dv;
    out->runtime_version = rv;
    out->cap_major = p.major;
    out->cap_minor = p.minor;
    return EGS_OK;
}

int egs_host_alloc(void **p, size_t bytes) {
    if (!p) return fail(EGS_E_ARG, "egs_host_alloc: null output");
    *p = nullptr;
    if (hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess)
        return fail(EGS_E_NOMEM, "hipHostMalloc of %zu bytes failed", bytes);
    return EGS_OK;
}

int egs_host_free(void *p) {
    if (p) HIP_TRY(hipHostFree(p));
    return EGS_OK;
}

int egs_runtime_error(char *buf, int len) {
    hipError_t e = hipGetLastError();
    if (buf && len > 0) snprintf(buf, (size_t)len, "%s", e == hipSuccess ? "" : hipGetErrorString(e));
    return (int)e;
}

int egs_device_reset(int device) {
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceReset());
    return EGS_OK;
}

int egs_solver_destroy(egs_solver *s) {
    if (!s) return EGS_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    if (s->block) (void)hipFree(s->block);
    if (s->h_probe) (void)hipHostFree(s->h_probe);
    for (auto &e : s->ev_probe) if (e) (void)hipEventDestroy(e);
    if (s->ev_a) (void)hipEventDestroy(s->ev_a);
    if (s->ev_b) (void)hipEventDestroy(s->ev_b);
    if (s->ev_pa) (void)hipEventDestroy(s->ev_pa);
    if (s->ev_pb) (void)hipEventDestroy(s->ev_pb);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return EGS_OK;
}

int egs_solver_create(int device, int32_t n_cap, int64_t m_cap, egs_solver **out) {
    if (!out) return fail(EGS_E_ARG, "egs_solver_create: null output");
    *out = nullptr;
    if (n_cap < 1 || m_cap < 1 || m_cap > 0x7fffffffLL)
        return fail(EGS_E_ARG, "egs_solver_create: need 1 <= n_cap and 1 <= m_cap < 2^31 (got %d, %lld)",
                    n_cap, (long long)m_cap);
    int rc = check_gfx950(device);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(device));
    egs_solver *s = new egs_solver();
    s->device = device;
    if (hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) s->cus = 256;
    s->n_cap = n_cap;
    s->m_cap = m_cap;
    const Layout L = make_layout(n_cap, m_cap);
    s->block_bytes = L.total;
    if (hipMalloc(&s->block, L.total) != hipSuccess) {
        egs_solver_destroy(s);
        return fail(EGS_E_NOMEM, "hipMalloc of %zu bytes failed", L.total);
    }
    bind_layout(s, (char *)s->block, L);
    if (hipHostMalloc((void **)&s->h_probe, 4 * (kPinProbe + 2 * kProbeStride), hipHostMallocDefault) != hipSuccess) {
        egs_solver_destroy(s);
        return fail(EGS_E_NOMEM, "hipHostMalloc failed");
    }
    s->h_probe[kPinOne] = 1;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->ev_probe[0]) != hipSuccess || hipEventCreate(&s->ev_probe[1]) != hipSuccess ||
        hipEventCreate(&s->ev_a) != hipSuccess || hipEventCreate(&s->ev_b) != hipSuccess ||
        hipEventCreate(&s->ev_pa) != hipSuccess || hipEventCreate(&s->ev_pb) != hipSuccess) {
        egs_solver_destroy(s);
        return fail(EGS_E_HIP, "stream/event creation failed");
    }
    *out = s;
    return EGS_OK;
}

static void bind_game(egs_solver *s, int32_t n, int32_t n0, int64_t m, int32_t mg);

int egs_solver_load(egs_solver *s, int mem, int32_t n, int32_t n0, int64_t m, int32_t mg,
                    const int32_t *ptr, const int32_t *succ, const int32_t *w) {
    if (!s) return fail(EGS_E_ARG, "null solver");
    if (n < 1 || n0 < 0 || n0 > n || m < 1 || mg < 0 || !ptr || !succ || !w)
        return fail(EGS_E_ARG, "egs_solver_load: bad game (n=%d n0=%d m=%lld mg=%d)", n, n0,
                    (long long)m, mg);
    if (n > s->n_cap || m > s->m_cap)
        return fail(EGS_E_CAPACITY, "game (n=%d, m=%lld) exceeds solver capacity (%d, %lld)", n,
                    (long long)m, s->n_cap, (long long)s->m_cap);
    HIP_TRY(hipSetDevice(s->device));
    if (mem == EGS_MEM_HOST) {
        // validate the CSR on the host (fused with the int2 packing) before anything reaches
        // the device
        int2 *packed = (int2 *)malloc(sizeof(int2) * (size_t)m);
        if (!packed) return fail(EGS_E_NOMEM, "host staging allocation failed");
        int32_t wm = 0, mx = 0;
        int64_t bad = -1;
        const int chk = egs_pk_check_pack64(n, m, ptr, succ, w, (int32_t *)packed, &wm, &mx, &bad);
        if (chk) free(packed);
        switch (chk) {
        case EGS_PK_BAD_OFFSETS:
            return fail(EGS_E_ARG, "CSR offsets must run from 0 to m (got %d..%d)", ptr[0], ptr[n]);
        case EGS_PK_NO_EDGE: return fail(EGS_E_ARG, "vertex %lld has no out-edge", (long long)bad);
        case EGS_PK_BAD_SUCC:
            return fail(EGS_E_ARG, "edge %lld: successor %d out of range", (long long)bad, succ[bad]);
        default: break;
        }
        s->max_outdeg = mx;
        s->wmax = wm;
        hipError_t e1 = hipMemcpy(s->d_ptr, ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice);
        hipError_t e2 = hipMemcpy(s->d_edges, packed, sizeof(int2) * (size_t)m, hipMemcpyHostToDevice);
        free(packed);
        HIP_TRY(e1);
        HIP_TRY(e2);
    } else if (mem == EGS_MEM_DEVICE) {
        // device inputs are trusted to be a valid CSR of this shape (the caller owns them)
        HIP_TRY(hipMemcpyAsync(s->d_ptr, ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice, s->stream));
        HIP_TRY(egs::launch_pack(succ, w, s->d_edges, m, s->stream));
        HIP_TRY(egs::launch_degree_stats(s->d_ptr, n, s->d_misc, s->stream));
        HIP_TRY(hipMemcpyAsync(s->h_probe + kPinDeg, s->d_misc, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        s->max_outdeg = s->h_probe[kPinDeg];
        egs::DevGame tmp{};
        tmp.edges = s->d_edges;
        tmp.m = m;
        HIP_TRY(egs::launch_abs_wmax(tmp, s->d_misc, s->stream));
        HIP_TRY(hipMemcpyAsync(s->h_probe + kPinDeg, s->d_misc, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        s->wmax = s->h_probe[kPinDeg];
    } else {
        return fail(EGS_E_ARG, "mem must be EGS_MEM_HOST or EGS_MEM_DEVICE");
    }
    bind_game(s, n, n0, m, mg);
    return EGS_OK;
}

int egs_solver_load_packed(egs_solver *s, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *d_ptr,
                           const uint32_t *d_edges, int32_t shift, int32_t wmax, int32_t max_outdeg) {
    if (!s) return fail(EGS_E_ARG, "null solver");
    if (n < 1 || n0 < 0 || n0 > n || m < 1 || mg < 0 || !d_ptr || !d_edges || shift < 1 || shift > 30 || wmax < 0 ||
        max_outdeg < 1)
        return fail(EGS_E_ARG, "egs_solver_load_packed: bad game (n=%d n0=%d m=%lld mg=%d shift=%d)", n, n0,
                    (long long)m, mg, shift);
    if (n > s->n_cap || m > s->m_cap)
        return fail(EGS_E_CAPACITY, "game (n=%d, m=%lld) exceeds solver capacity (%d, %lld)", n,
                    (long long)m, s->n_cap, (long long)s->m_cap);
    HIP_TRY(hipSetDevice(s->device));
    // device inputs are trusted to be a checked CSR of this shape (the caller owns them)
    HIP_TRY(hipMemcpyAsync(s->d_ptr, d_ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(egs::launch_unpack(d_edges, shift, s->d_edges, m, s->stream));
    s->max_outdeg = max_outdeg;
    s->wmax = wmax;
    bind_game(s, n, n0, m, mg);
    return EGS_OK;
}

int egs_solver_load_read_order(egs_solver *s, int mem, int32_t n, int64_t m, const uint8_t *owner,
                               const int32_t *rptr, const int32_t *rsucc, const int32_t *rw, int32_t *n0_out,
                               int32_t *mg_out, int32_t *ext_of_int) {
    if (!s) return fail(EGS_E_ARG, "null solver");
    if (n < 1 || m < 1 || !owner || !rptr || !rsucc || !rw || (mem != EGS_MEM_HOST && mem != EGS_MEM_DEVICE))
        return fail(EGS_E_ARG, "egs_solver_load_read_order: bad game (n=%d m=%lld)", n, (long long)m);
    if (n > s->n_cap || m > s->m_cap)
        return fail(EGS_E_CAPACITY, "game (n=%d, m=%lld) exceeds solver capacity (%d, %lld)", n, (long long)m,
                    s->n_cap, (long long)s->m_cap);
    HIP_TRY(hipSetDevice(s->device));
    egs::ReadOrderBuffers b{};
    if (mem == EGS_MEM_HOST) {
        // one upload into work space that prepare and the rounds rewrite later
        uint8_t *d_owner = (uint8_t *)s->d_stamp;
        int *d_rsucc = (int *)s->d_titems, *d_rw = (int *)s->d_titems + m;
        HIP_TRY(hipMemcpyAsync(d_owner, owner, (size_t)n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->d_cptr, rptr, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(d_rsucc, rsucc, sizeof(int) * (size_t)m, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(d_rw, rw, sizeof(int) * (size_t)m, hipMemcpyHostToDevice, s->stream));
        b.owner = d_owner;
        b.rptr = s->d_cptr;
        b.rsucc = d_rsucc;
        b.rw = d_rw;
    } else {
        b.owner = owner;
        b.rptr = rptr;
        b.rsucc = rsucc;
        b.rw = rw;
    }
    HIP_TRY(egs::launch_read_order_check(b.rptr, b.rsucc, n, m, s->d_misc + 8, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_probe + kPinDeg, s->d_misc + 8, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int bad = s->h_probe[kPinDeg];
    if (bad & 1) return fail(EGS_E_ARG, "read-order offsets must run from 0 to m");
    if (bad & 2) return fail(EGS_E_ARG, "a vertex has no out-edge");
    if (bad & 4) return fail(EGS_E_ARG, "a successor is out of range");
    b.flag0 = s->d_sigma;
    b.pre0 = s->d_fill;
    b.rev = s->d_dbig;
    b.ext_of_int = s->d_sbig;
    b.deg = s->d_dhuge;
    b.scan_tiles = s->d_tile;
    unsigned long long *d_mg = (unsigned long long *)s->d_misc;
    HIP_TRY(egs::launch_internalise(b, n, s->d_ptr, s->d_edges, d_mg, s->stream));
    unsigned long long *h_mg = (unsigned long long *)(s->h_probe + kPinViol);
    HIP_TRY(hipMemcpyAsync(h_mg, d_mg, sizeof(unsigned long long), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_probe + kPinDeg + 1, s->d_fill + n, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    if (ext_of_int)
        HIP_TRY(hipMemcpyAsync(ext_of_int, s->d_sbig, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(egs::launch_degree_stats(s->d_ptr, n, s->d_misc + 2, s->stream));
    egs::DevGame tmp{};
    tmp.edges = s->d_edges;
    tmp.m = m;
    HIP_TRY(egs::launch_abs_wmax(tmp, s->d_misc + 3, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_probe + kPinDeg + 2, s->d_misc + 2, 2 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (*h_mg > 0x7fffffffULL) return fail(EGS_E_ARG, "Error too high weights --> overflow (MG past INT32_MAX)");
    const int32_t n0 = s->h_probe[kPinDeg + 1], mg = (int32_t)*h_mg;
    s->max_outdeg = s->h_probe[kPinDeg + 2];
    s->wmax = s->h_probe[kPinDeg + 3];
    bind_game(s, n, n0, m, mg);
    if (n0_out) *n0_out = n0;
    if (mg_out) *mg_out = mg;
    return EGS_OK;
}

int egs_solver_read_game(egs_solver *s, int32_t *ptr, int32_t *succ, int32_t *w) {
    if (!s || !s->loaded || !ptr || !succ || !w) return fail(EGS_E_STATE, "egs_solver_read_game: nothing loaded");
    HIP_TRY(hipSetDevice(s->device));
    const int64_t m = s->g.m;
    int2 *tmp = (int2 *)malloc(sizeof(int2) * (size_t)m);
    if (!tmp) return fail(EGS_E_NOMEM, "host buffer of %lld edges", (long long)m);
    hipError_t e1 = hipMemcpyAsync(ptr, s->d_ptr, sizeof(int) * ((size_t)s->g.n + 1), hipMemcpyDeviceToHost, s->stream);
    hipError_t e2 = hipMemcpyAsync(tmp, s->d_edges, sizeof(int2) * (size_t)m, hipMemcpyDeviceToHost, s->stream);
    hipError_t e3 = hipStreamSynchronize(s->stream);
    if (e1 == hipSuccess && e2 == hipSuccess && e3 == hipSuccess)
        for (int64_t k = 0; k < m; k++) {
            succ[k] = tmp[k].x;
            w[k] = tmp[k].y;
        }
    free(tmp);
    HIP_TRY(e1);
    HIP_TRY(e2);
    HIP_TRY(e3);
    return EGS_OK;
}

// Point the device-game view at the solver's arrays for a freshly loaded game.
static void bind_game(egs_solver *s, int32_t n, int32_t n0, int64_t m, int32_t mg) {
    egs::DevGame &g = s->g;
    g.n = n;
    g.n0 = n0;
    g.mg = mg;
    g.m = m;
    g.ptr = s->d_ptr;
    g.edges = s->d_edges;
    g.cptr = s->d_cptr;
    g.cpred = s->d_cpred;
    g.val = s->d_val;
    g.stamp = s->d_stamp;
    g.list[0] = s->d_list[0];
    g.list[1] = s->d_list[1];
    g.sub_cap = s->sub_cap;
    g.lcnt = s->d_lcnt;
    g.mode = s->d_mode;
    g.dbig = s->d_dbig;
    g.dhuge = s->d_dhuge;
    g.sbig = s->d_sbig;
    g.shuge = s->d_shuge;
    g.ring = s->d_ring;
    g.stats = s->d_stats;
    g.rounds_done = s->d_rounds;
    s->loaded = true;
    s->prepared = false;
}

static int pick_shape(egs_solver *s, const egs_options *o) {
    egs::DevGame &g = s->g;
    egs::LaunchShape &sh = s->sh;
    egs_options def{};
    if (!o) o = &def;
    const double avg = (double)g.m / (double)g.n;
    int G = o->group;
    if (G == 0) G = std::min(16, pow2ceil((int)(avg + 0.999)));
    if (G < 1 || G > 64 || (G & (G - 1)))
        return fail(EGS_E_ARG, "group must be a power of two in 1..64 (got %d)", G);
    sh.group = G;
    g.big_deg = o->big_deg > 0 ? o->big_deg : std::max(64, 8 * G);
    g.huge_deg = o->huge_deg > 0 ? o->huge_deg : 4096;
    if (g.huge_deg < g.big_deg) g.huge_deg = g.big_deg;
    const int pct = o->dense_pct > 0 ? o->dense_pct : 12;
    g.dense_min = pct > 100 ? INT64_MAX : std::max<int64_t>(1, (int64_t)g.n * pct / 100);
    sh.grid = o->grid > 0 ? o->grid : 1024;
    if (sh.grid > egs::kStatSlots) sh.grid = egs::kStatSlots;
    sh.use_wave = s->max_outdeg > g.big_deg;
    sh.use_block = s->max_outdeg > g.huge_deg;
    sh.wave_grid = 1024;
    sh.block_grid = 1024;
    s->batch_rounds = o->batch_rounds > 0 ? o->batch_rounds : 32;
    // with a stop flag at most kStopBatch rounds are queued ahead of the host, so a stop
    // takes effect within a few rounds
    if (o->stop_flag) s->batch_rounds = std::min(s->batch_rounds, kStopBatch);
    if (o->graphs < -1 || o->graphs > 1) return fail(EGS_E_ARG, "graphs must be -1, 0 or 1 (got %d)", o->graphs);
    s->use_graphs = o->graphs > 0 && !o->stop_flag && !s->no_alloc;
    s->phase_cap = egs_max_loop_c(g.m, g.n, g.mg, o->max_phases);
    s->stop_flag = o->stop_flag;
    s->cert_every = o->certify_every == 0 ? kCertEvery : (o->certify_every < 0 ? 0 : o->certify_every);
    s->cert_max_passes = o->cert_max_passes > 0 ? o->cert_max_passes : 1024;
    g.rise = s->cert_every > 0 ? s->d_sigma : nullptr;    // rise stamps only when proofs run
    // phased dense rounds where the value array outgrows the L2 (and the buffers exist):
    // slices of 2^slice_shift successors, at most kMaxPhases of them
    // auto: phases once the value array outgrows 4 MB (at 2^20 vertices one in-place sweep
    // wins: the skewed 2^20 game solved in 13.6 ms in 81 rounds, against 15.8-16.3 ms in 98-99
    // rounds with 2-8 phases)
    int want = o->slices == 0 ? (g.n > egs::kPhaseMinN ? egs::kDefaultPhases : 1) : o->slices;
    if (want < 1 || want > egs::kMaxPhases)
        return fail(EGS_E_ARG, "slices must be 0 (auto) or 1..%d (got %d)", egs::kMaxPhases, o->slices);
    g.phases = 1;
    g.slice_shift = 31;
    if (want > 1 && s->d_slen) {
        int shift = 0;
        while (((int64_t)(want) << shift) < g.n) shift++;          // want slices of 2^shift cover n
        g.slice_shift = shift;
        g.phases = (int)(((int64_t)g.n + (1LL << shift) - 1) >> shift);
        if (g.big_deg > egs::kMaxSliceDeg) g.big_deg = egs::kMaxSliceDeg;   // lengths are bytes
        if (g.huge_deg < g.big_deg) g.huge_deg = g.big_deg;
        sh.use_wave = s->max_outdeg > g.big_deg;
        sh.use_block = s->max_outdeg > g.huge_deg;
    }
    g.slen = s->d_slen;
    g.sbase = s->d_sbase;
    g.sedges = s->d_sedges;
    g.sshift = g.phases > 1 ? egs::pack_shift(g.n, s->wmax) : 0;   // 32-bit slice edges when they fit
    g.hacc = s->d_hacc;
    g.hcnt = s->d_hcnt;
    g.hbest = s->d_hbest;
    g.hflag = s->d_hflag;
    g.bigmask = s->d_bigmask;
    // The phased kernels run about as many workgroups as the device holds at once (eight of
    // 256 threads per CU), each over a fixed set of tiles whose partial results sit in LDS.
    const int64_t tiles = ((int64_t)g.n + egs::kSliceTile - 1) / egs::kSliceTile;
    const int wpc = o->phase_wgs_per_cu > 0 ? std::min(o->phase_wgs_per_cu, 32) : 8;
    g.phase_grid = (int)std::min<int64_t>(tiles, wpc * (int64_t)std::max(1, s->cus));
    g.phase_tiles = (int)((tiles + g.phase_grid - 1) / g.phase_grid);
    if (g.phases > 1 && g.phase_tiles * egs::kSliceTile * 12 > 96 * 1024) {
        // keep the closure kernel's partial state (12 bytes a vertex) within 96 KB of LDS
        g.phase_tiles = 96 * 1024 / (egs::kSliceTile * 12);
        g.phase_grid = (int)((tiles + g.phase_tiles - 1) / g.phase_tiles);
    }
    // one-sweep games lift their static hub lists in the same launch as the low-degree sweep
    g.round_grid = sh.grid;
    g.wave_grid = sh.use_wave ? sh.wave_grid : 0;
    g.block_grid = sh.use_block ? sh.block_grid : 0;
    g.hub_fused = g.phases == 1 && (sh.use_wave || sh.use_block) &&
                  g.round_grid + g.wave_grid + g.block_grid <= egs::kStatSlots;     // a statistics slot per workgroup
    return EGS_OK;
}

// Transpose on the solver stream, then open the sparse round modes (ring[12] = 1).
static hipError_t build_csc(egs_solver *s) {
    hipError_t e = egs::transpose_sorted_fits(s->g.n, s->g.m, s->tcounts_cap, s->m_cap)
                       ? egs::launch_transpose_sorted(s->g, s->d_tcounts, s->d_toffs, s->d_tscan, s->d_titems,
                                                      s->stream)
                       : egs::launch_transpose(s->g, s->d_tile, s->d_fill, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_ring + 12, s->h_probe + kPinOne, sizeof(int), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) s->csc_built = true;
    return e;
}

int egs_solver_prepare(egs_solver *s, const egs_options *opt) {
    if (!s || !s->loaded) return fail(EGS_E_STATE, "egs_solver_prepare: no game loaded");
    HIP_TRY(hipSetDevice(s->device));
    int rc = pick_shape(s, opt);
    if (rc) return rc;
    // graphs hold the launch arguments they were captured with: kept while those stay the same
    if (memcmp(&s->graph_g, &s->g, sizeof(s->g)) != 0 || memcmp(&s->graph_sh, &s->sh, sizeof(s->sh)) != 0) {
        drop_graphs(s);
        memcpy(&s->graph_g, &s->g, sizeof(s->g));
        memcpy(&s->graph_sh, &s->sh, sizeof(s->sh));
    }
    // The set-up is timed by its own events and not waited for: the first round queues
    // behind it at once, and its time and the initial frontier are read after a later wait.
    // (The no-allocation solver has no events: it times on the host and waits here.)
    HIP_TRY(close_prepare(s));
    const bool lazy = s->ev_pa && s->ev_pb;
    double t0 = 0;
    if (lazy) HIP_TRY(hipEventRecord(s->ev_pa, s->stream));
    else HIP_TRY(mark_start(s, &t0));
    HIP_TRY(egs::launch_prepare_reset(s->g, s->stream));   // statistics, counters, modes, ring slots
    // The transpose only feeds the sparse round modes, and many solves never need them
    // (config-2-like games sweep densely until a proof ends them).  Rounds stay dense while
    // it is missing; a round that would have gone sparse raises ring[13], and the host
    // builds the transpose at its next probe and opens sparse rounds (ring[12]).
    s->csc_built = false;
    if (s->g.dense_min > s->g.n) HIP_TRY(build_csc(s));     // dense sweeps switched off
    if (s->sh.use_block)             // shared lifts of the vertices above huge_deg start empty
        HIP_TRY(egs::launch_hbest_reset(s->g, s->stream));
    HIP_TRY(egs::launch_classify(s->g, s->d_ring + 9, s->stream));
    // with slices and no transpose the slice set-up also stores the initial state
    const bool init_in_slices = s->g.phases > 1 && !s->csc_built;
    HIP_TRY(egs::launch_slices(s->g, s->stream, init_in_slices));
    HIP_TRY(egs::launch_init(s->g, s->stream, init_in_slices));
    // the initial frontier's counters ride on the same wait as the timing
    HIP_TRY(hipMemcpyAsync(s->h_probe + kPinInit, s->d_lcnt + 1 * egs::kListCounters,
                           sizeof(int) * egs::kListCounters, hipMemcpyDeviceToHost, s->stream));
    if (lazy) {
        HIP_TRY(hipEventRecord(s->ev_pb, s->stream));
        memset(&s->st, 0, sizeof(s->st));
        s->prep_open = true;
    } else {
        float ms = 0.f;
        HIP_TRY(mark_stop(s, t0, &ms));
        memset(&s->st, 0, sizeof(s->st));
        s->st.ms_prepare = ms;
        s->st.initial_active = (int32_t)list_total(s, s->h_probe + kPinInit);
    }
    s->next_round = 1;
    s->cert_gap = s->cert_every;
    s->next_cert = s->cert_every;
    s->cert_cap = std::min(kCertFirstCap, s->cert_max_passes);
    s->cert_pass_hint = 0;
    s->prepared = true;
    return EGS_OK;
}

// The run's statistics: queued on the stream before the run's closing wait, read after it.
static int queue_stats(egs_solver *s) {
    HIP_TRY(egs::launch_sum_stats(s->g, egs::kStatSlots, s->d_statsum, s->stream));
    unsigned long long *h = (unsigned long long *)(s->h_probe + kPinStats);
    // statistics, last round and the next round's input size in one read-back
    HIP_TRY(hipMemcpyAsync(h, s->d_statsum, sizeof(unsigned long long) * 6, hipMemcpyDeviceToHost, s->stream));
    return EGS_OK;
}

static int collect_stats(egs_solver *s) {
    const unsigned long long *h = (const unsigned long long *)(s->h_probe + kPinStats);
    s->st.processed = (int64_t)h[0];
    s->st.edges_relaxed = (int64_t)h[1];
    s->st.lifted = (int64_t)h[2];
    s->st.pred_visits = (int64_t)h[3];
    s->st.phases = (int32_t)h[4];
    s->st.csc_built = s->csc_built ? 1 : 0;
    s->st.converged = h[5] == 0;
    return EGS_OK;
}

int egs_solver_iterate(egs_solver *s, int64_t max_phases, egs_stats *st) {
    if (!s || !s->prepared) return fail(EGS_E_STATE, "egs_solver_iterate: call egs_solver_prepare first");
    HIP_TRY(hipSetDevice(s->device));
    int64_t limit = s->phase_cap - (s->next_round - 1);
    if (max_phases >= 0 && max_phases < limit) limit = max_phases;
    if (s->next_round + limit > 0x7ffffff0LL) limit = 0x7ffffff0LL - s->next_round;
    double t0 = 0;
    HIP_TRY(mark_start(s, &t0));
    int64_t launched = 0;
    int pending = -1;     // probe slot in flight
    bool done = false;
    int batch = 4;        // grow towards batch_rounds so short solves do not overshoot much
    while (!done && launched < limit) {
        int nb = (int)std::min<int64_t>(batch, limit - launched);
        const bool accel = s->cert_every > 0;
        if (accel && s->next_round - 1 >= s->next_cert) {
            // pause: converged already?  else try to prove a climbing set is TOP
            HIP_TRY(hipMemcpyAsync(s->h_probe + kPinConv, s->d_lcnt + (s->next_round % 3) * egs::kListCounters,
                                   sizeof(int) * egs::kListCounters, hipMemcpyDeviceToHost, s->stream));
            // the attempt's value histogram rides on the same wait
            // waits through an event: a busy wait on it wakes the host sooner than a stream
            // synchronisation (small games solved 4-8% faster)
            egs::CertBuffers cb{s->d_hist, s->d_h, s->d_bfcnt, s->d_bits, s->h_probe + kPinCert, s->ev_b};
            const bool hist = s->wmax > 0;
            if (hist) HIP_TRY(egs::certify_prefetch(s->g, cb, s->wmax, s->stream));
            // from the run's second attempt on, one more round runs while the host waits for
            // these read-backs (a no-op when the frontier was already empty) and the attempt
            // starts after it; any candidate set is proved or refused exactly, so the
            // histogram may be a round old.  Not before the first attempt: its proof usually
            // marks many vertices TOP, after which the same round costs less.
            const bool ahead = s->st.cert_attempts > 0 && launched < limit && !(s->stop_flag && *s->stop_flag) &&
                               s->ev_b;
            if (ahead) {
                HIP_TRY(hipEventRecord(s->ev_b, s->stream));
                HIP_TRY(egs::launch_round(s->g, s->sh, 0, s->stream));
                HIP_TRY(egs::launch_advance(s->g, 1, s->stream));
                s->next_round++;
                launched++;
                HIP_TRY(hipEventSynchronize(s->ev_b));
            } else {
                HIP_TRY(egs::wait_queued(cb.ev, s->stream));
            }
            if (list_total(s, s->h_probe + kPinConv) == 0) break;
            if (s->stop_flag && *s->stop_flag) break;
            egs::CertResult cr;
            const int got = egs::certify_top(s->g, cb, s->wmax, (int)(s->next_round - 1), s->cert_cap,
                                             kCertCandidates, s->sh.use_wave != 0, hist, s->stop_flag,
                                             s->cert_pass_hint, s->stream, cr);
            if (got < 0) return fail(EGS_E_HIP, "certification: %s", hipGetErrorString(hipGetLastError()));
            if (cr.last_passes > 0) s->cert_pass_hint = cr.last_passes;
            s->st.cert_attempts++;
            s->st.cert_passes += cr.passes;
            s->st.cert_pruned += cr.closure_failed;
            if (got > 0) {
                s->st.certificates++;
                s->st.top_marked += cr.marked;
                // the marking kernel has named the next round (ring[11]) a full sweep
                // the spacing only grows (doubled by failures): a long run that keeps
                // proving small sets does not pay an attempt every few rounds
                batch = 1;        // the fixpoint often follows at once: probe again early
                nb = 1;
            } else {
                s->cert_gap = std::min(s->cert_gap * 2, 1 << 16);
                if (cr.cap_hit) s->cert_cap = std::min(s->cert_cap * 2, s->cert_max_passes);
            }
            s->next_cert = s->next_round - 1 + s->cert_gap;
            pending = -1;
        }
        if (accel) nb = (int)std::max<int64_t>(1, std::min<int64_t>(nb, s->next_cert - (s->next_round - 1)));
        int k = 0;
        if (s->use_graphs) {
            for (int b = kGraphSizes - 1; b >= 0; b--)
                while (nb - k >= (1 << b)) {
                    HIP_TRY(replay_batch(s, b));
                    k += 1 << b;
                }
        } else {
            for (; k < nb; k++) {
                // the stop flag is polled before every round, as the reference polls
                // timeout_expired once per round (src/dev_EG_alg.cu:376)
                if (s->stop_flag && *s->stop_flag) break;
                HIP_TRY(egs::launch_round(s->g, s->sh, k, s->stream));
            }
            if (k > 0) HIP_TRY(egs::launch_advance(s->g, k, s->stream));
        }
        if (k == 0) break;
        s->next_round += k;
        launched += k;
        const int slot = (pending + 1) & 1;
        int *hp = s->h_probe + kPinProbe + slot * kProbeStride;
        const int *src = s->d_lcnt + (s->next_round % 3) * egs::kListCounters;
        // Until the transpose exists the probe also reads ring[13].  The list counters, the
        // round modes and the ring lie in that order in the device arena, so one copy of the
        // span from this slot's counters to ring[13] reads both.
        const ptrdiff_t span = s->d_ring + 14 - src;
        if (s->csc_built) {
            HIP_TRY(hipMemcpyAsync(hp, src, sizeof(int) * egs::kListCounters, hipMemcpyDeviceToHost, s->stream));
        } else if (span > egs::kListCounters && span <= kProbeStride) {
            HIP_TRY(hipMemcpyAsync(hp, src, sizeof(int) * (size_t)span, hipMemcpyDeviceToHost, s->stream));
            s->probe_want[slot] = (int)span - 1;
        } else {
            HIP_TRY(hipMemcpyAsync(hp, src, sizeof(int) * egs::kListCounters, hipMemcpyDeviceToHost, s->stream));
            HIP_TRY(hipMemcpyAsync(hp + kProbeWant, s->d_ring + 13, sizeof(int), hipMemcpyDeviceToHost, s->stream));
            s->probe_want[slot] = kProbeWant;
        }
        HIP_TRY(probe_mark(s, slot));
        if (pending >= 0) {
            HIP_TRY(probe_wait(s, pending));
            if (list_total(s, s->h_probe + kPinProbe + pending * kProbeStride) == 0) done = true;
            // enough rounds wanted to go sparse: build the transpose now
            if (!done && !s->csc_built && s->h_probe[kPinProbe + pending * kProbeStride + s->probe_want[pending]] >= kSparseWantRounds)
                HIP_TRY(build_csc(s));
        }
        pending = slot;
        if (s->stop_flag && *s->stop_flag) break;
        if (batch < s->batch_rounds) batch = std::min(s->batch_rounds, batch * 2);
    }
    int rc = queue_stats(s);
    if (rc) return rc;
    float ms = 0.f;
    HIP_TRY(mark_stop(s, t0, &ms));                  // also the wait for the statistics
    HIP_TRY(close_prepare(s));
    rc = collect_stats(s);
    if (rc) return rc;
    s->st.ms_iterate += ms;
    if (st) *st = s->st;
    return s->st.converged ? 1 : 0;
}

int egs_solver_stats(egs_solver *s, egs_stats *st) {
    if (!s || !st) return fail(EGS_E_ARG, "null argument");
    HIP_TRY(close_prepare(s));
    *st = s->st;
    return EGS_OK;
}

int egs_solver_read_values(egs_solver *s, int32_t *host_values) {
    if (!s || !s->loaded || !host_values) return fail(EGS_E_STATE, "egs_solver_read_values: nothing to read");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(host_values, s->d_val, sizeof(int) * (size_t)s->g.n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return EGS_OK;
}

int egs_solver_read_winning(egs_solver *s, uint32_t *host_bits) {
    if (!s || !s->loaded || !host_bits) return fail(EGS_E_STATE, "egs_solver_read_winning: nothing to read");
    HIP_TRY(hipSetDevice(s->device));
    // the proof bitmaps are idle outside an attempt: the first holds ceil(n/32) words
    HIP_TRY(egs::launch_winning(s->d_val, s->g.n, s->d_bits, s->stream));
    HIP_TRY(hipMemcpyAsync(host_bits, s->d_bits, sizeof(uint32_t) * (((size_t)s->g.n + 31) / 32), hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return EGS_OK;
}

const int32_t *egs_solver_device_values(egs_solver *s) { return s ? s->d_val : nullptr; }

void *egs_solver_stream(egs_solver *s) { return s ? (void *)s->stream : nullptr; }

int egs_solver_read_csc(egs_solver *s, int32_t *cptr, int32_t *cpred) {
    if (!s || !s->prepared) return fail(EGS_E_STATE, "egs_solver_read_csc: not prepared");
    HIP_TRY(hipSetDevice(s->device));
    if (!s->csc_built) HIP_TRY(build_csc(s));
    HIP_TRY(hipMemcpyAsync(cptr, s->d_cptr, sizeof(int) * ((size_t)s->g.n + 1), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(cpred, s->d_cpred, sizeof(int) * (size_t)s->g.m, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return EGS_OK;
}

int egs_solver_count_violations(egs_solver *s, int64_t *violations) {
    if (!s || !s->loaded || !violations) return fail(EGS_E_STATE, "egs_solver_count_violations: nothing loaded");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(egs::launch_violations(s->g, (unsigned long long *)s->d_misc, s->stream));
    unsigned long long *h = (unsigned long long *)(s->h_probe + kPinViol);
    HIP_TRY(hipMemcpyAsync(h, s->d_misc, sizeof(unsigned long long), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *violations = (int64_t)*h;
    return EGS_OK;
}

int egs_solver_set_values(egs_solver *s, const int32_t *host_values) {
    if (!s || !s->prepared || !host_values) return fail(EGS_E_STATE, "egs_solver_set_values: call egs_solver_prepare first");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(s->d_val, host_values, sizeof(int) * (size_t)s->g.n, hipMemcpyHostToDevice, s->stream));
    // the values may hold TOP: the phased rounds then read their TOP masks (ring[15])
    HIP_TRY(hipMemcpyAsync(s->d_ring + 15, s->h_probe + kPinOne, sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return EGS_OK;
}

int egs_solver_top_certificate(egs_solver *s, int32_t max_passes, int64_t *host_h, int32_t *passes) {
    if (!s || !s->prepared || !host_h) return fail(EGS_E_STATE, "egs_solver_top_certificate: call egs_solver_prepare first");
    HIP_TRY(hipSetDevice(s->device));
    egs::CertBuffers cb{s->d_hist, s->d_h, s->d_bfcnt, s->d_bits, s->h_probe + kPinCert};
    egs::CertResult cr;
    const int got = egs::certify_all_top(s->g, cb, s->wmax, max_passes > 0 ? max_passes : (1 << 16),
                                         s->sh.use_wave != 0, s->stream, cr);
    if (got == -1) return fail(EGS_E_HIP, "certificate: %s", hipGetErrorString(hipGetLastError()));
    if (got == -2) return fail(EGS_E_ARG, "certificate: potentials of this game could overflow 64 bits");
    HIP_TRY(hipMemcpyAsync(host_h, s->d_h, sizeof(int64_t) * (size_t)s->g.n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (passes) *passes = cr.passes;
    return got;
}

int egs_solve(int device, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
              const int32_t *succ, const int32_t *w, int32_t *values, const egs_options *opt,
              egs_stats *st) {
    egs_solver *s = nullptr;
    int rc = egs_solver_create(device, n, m, &s);
    if (rc) return rc;
    rc = egs_solver_load(s, EGS_MEM_HOST, n, n0, m, mg, ptr, succ, w);
    if (!rc) rc = egs_solver_prepare(s, opt);
    if (!rc) {
        rc = egs_solver_iterate(s, -1, st);
        if (rc >= 0) rc = egs_solver_read_values(s, values) ? EGS_E_HIP : EGS_OK;
    }
    egs_solver_destroy(s);
    return rc;
}

// ---------------------------------------------------------------------------------------
// Solve without allocating: for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).  The workspace of
// the one-game solver for games up to kNoAllocN vertices and kNoAllocM edges is a
// file-scope device array; each device that runs it holds its own copy until the process
// ends, found once per device with hipGetSymbolAddress (the array's name is not a device
// address in host code).  The host side of the solver is a static object per device.
// Calls that overlap on one device share both and race.
// ---------------------------------------------------------------------------------------

namespace {

constexpr int32_t kNoAllocN = 1 << 16;
constexpr int64_t kNoAllocM = 1 << 20;
constexpr size_t kScratchBytes = make_layout(kNoAllocN, kNoAllocM).total;
constexpr int kNoAllocDevices = 64;

struct NoAllocSlot {
    egs_solver s;
    int probe[kPinProbe + 2 * kProbeStride];   // the solver's read-back area (pageable; every read waits on the stream)
    bool ready = false;
};
NoAllocSlot g_noalloc[kNoAllocDevices];

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static int noalloc_setup(int dev) {
    NoAllocSlot &sl = g_noalloc[dev];
    if (sl.ready) return EGS_OK;
    int rc = check_gfx950(dev);
    if (rc) return rc;
    char *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(EGS_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    egs_solver &s = sl.s;
    s.device = dev;
    s.no_alloc = true;
    if (hipDeviceGetAttribute(&s.cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) s.cus = 256;
    s.n_cap = kNoAllocN;
    s.m_cap = kNoAllocM;
    s.block = nullptr;                 // not owned: never freed
    s.block_bytes = kScratchBytes;
    bind_layout(&s, ptr, make_layout(kNoAllocN, kNoAllocM));
    s.h_probe = sl.probe;
    s.h_probe[kPinOne] = 1;
    sl.ready = true;
    return EGS_OK;
}

int egs_solve_noalloc(int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *d_ptr, const int32_t *d_succ,
                      const int32_t *d_w, int32_t *d_values, const egs_options *opt, egs_stats *st, void *stream) {
    if (n < 1 || n0 < 0 || n0 > n || m < 1 || mg < 0 || !d_ptr || !d_succ || !d_w || !d_values)
        return fail(EGS_E_ARG, "egs_solve_noalloc: bad game (n=%d n0=%d m=%lld mg=%d)", n, n0, (long long)m, mg);
    if (n > kNoAllocN || m > kNoAllocM)
        return fail(EGS_E_CAPACITY,
                    "egs_solve_noalloc: game (n=%d, m=%lld) exceeds the fixed workspace (%d vertices, %lld edges)", n,
                    (long long)m, kNoAllocN, (long long)kNoAllocM);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) return fail(EGS_E_ARG, "egs_solve_noalloc: device %d out of range", dev);
    int rc = noalloc_setup(dev);
    if (rc) return rc;
    egs_solver *s = &g_noalloc[dev].s;
    s->stream = (hipStream_t)stream;
    s->stop_flag = nullptr;
    rc = egs_solver_load(s, EGS_MEM_DEVICE, n, n0, m, mg, d_ptr, d_succ, d_w);
    if (!rc) rc = egs_solver_prepare(s, opt);
    if (rc) return rc;
    rc = egs_solver_iterate(s, -1, st);
    if (rc < 0) return rc;
    HIP_TRY(hipMemcpyAsync(d_values, s->d_val, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return rc;
}

}  // extern "C"
