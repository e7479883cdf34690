// egs_serve.hip -- serving loop: games arrive over time, are admitted into a device
// arena by memory footprint, stay resident (graph, values, proof state) across time
// slices of the many-games kernel, and leave when their fixpoint is reached.
//
// No reference counterpart.  Each game's values are the least fixpoint the single-game
// path gives (the slices resume the same computation; see GameState in egs_games.h).
//
// Two routes, decided at admission by vertex count:
//   * small games (values fit a workgroup's LDS share) run in time slices of the
//     many-games kernel (egs_games.h), a workgroup per game, edge budget per slice;
//   * large games are solved one after another by the device-wide single-game solver
//     (egs_api.hip) from their resident arena copy, while a slice runs beside them.
// Host protocol: submit() packs a game into pinned staging and copies it into its arena
// chunk on the copy stream (overlapping a running slice); launch() starts one slice over
// every resident unfinished small game on the compute stream after the pending uploads;
// wait() solves the queued large games (up to a per-call limit), ends the slice and
// reports the games that finished; fetch() copies a result out and frees the chunk.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/egs.h"
#include "egs_games.h"
#include "egs_internal.h"
#include "egs_pack.h"

namespace egs {
int lanes_for(double avg_deg, const egs_batch_options &o);

// Owner count and MG of a game internalised on the device (submit_read_order), per slot.
struct ReadOrderMeta {
    int valid, n0, mg, pad;
};

// The internal int2 edges of a game internalised on the device, moved into its arena
// chunk in the route's layout: 32-bit (w << shift) | succ when shift > 0, else int2
// {succ, w} (small route) or successors and weights apart (large route, `split`).
__global__ __launch_bounds__(256) void k_ro_place_edges(const int2 *__restrict__ in, int64_t m, int shift, int split,
                                                        void *__restrict__ out, int *__restrict__ out_w) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const int2 x = in[e];
        if (shift) ((uint32_t *)out)[e] = ((uint32_t)x.y << shift) | (uint32_t)x.x;
        else if (split) { ((int *)out)[e] = x.x; out_w[e] = x.y; }
        else ((int2 *)out)[e] = x;
    }
}

// meta = {1, n0 = owner-0 count (pre0[n]), MG}; MG fits 31 bits (checked on the host).
__global__ void k_ro_meta(const int *__restrict__ pre0, int n, const unsigned long long *__restrict__ mg,
                          ReadOrderMeta *__restrict__ meta) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *meta = ReadOrderMeta{1, pre0[n], (int)*mg, 0};
}

// The slice's descriptors of games internalised on the device get their owner count and MG.
__global__ void k_ro_patch_descs(GameDesc *__restrict__ descs, const ReadOrderMeta *__restrict__ meta, int slots) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < slots; t += gridDim.x * blockDim.x)
        if (meta[t].valid) {
            descs[t].n0 = meta[t].n0;
            descs[t].mg = meta[t].mg;
        }
}
}  // namespace egs

namespace {
thread_local char s_err[512];
int sfail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(s_err, sizeof(s_err), fmt, ap);
    va_end(ap);
    return code;
}
#define SHIP(expr)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return sfail(EGS_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// First-fit allocator over [0, size) with coalescing of neighbouring free ranges.
struct Arena {
    std::map<size_t, size_t> free_;   // offset -> length
    size_t total = 0, in_use = 0;
    void init(size_t n) { free_.clear(); free_[0] = n; total = n; in_use = 0; }
    bool fits(size_t n) const {
        for (const auto &f : free_)
            if (f.second >= n) return true;
        return false;
    }
    bool alloc(size_t n, size_t &at) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                at = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[at + n] = rest;
                in_use += n;
                return true;
            }
        }
        return false;
    }
    void release(size_t at, size_t n) {
        in_use -= n;
        auto next = free_.lower_bound(at);
        if (next != free_.end() && at + n == next->first) { n += next->second; next = free_.erase(next); }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == at) { prev->second += n; return; }
        }
        free_[at] = n;
    }
};

struct Slot {
    bool used = false, finished = false, in_slice = false;
    bool large = false;   // device-wide route
    size_t at = 0, bytes = 0;
    int n = 0, n0 = 0, mg = 0;
    int64_t m = 0;
    int64_t seq = 0;      // admission order
    int64_t sweeps = 0, edges = 0;   // large route: rounds and edges relaxed
    int shift = 0, wmax = 0, maxdeg = 0;   // large route: edge packing (0 = separate arrays), max |w|, max degree
    int rc = 0;           // large route: the solve's return code
    std::string err;      // and its message
    bool read_order = false;   // internalised on the device: n0/MG in the device meta,
    size_t ext_at = 0;         // ext_of_int[n] at this arena offset (values go back in read order)
};

// Arena range of a device-side internalisation, freed once its kernels are done.
struct Transient {
    size_t at = 0, bytes = 0;
    hipEvent_t done = nullptr;
};

// Default large-route threads, each with its own solver and stream: two, so the server's
// streams (slices, copies and the two workers) fit the four hardware queues a process gets.
constexpr int kLargeWorkers = 2;
constexpr int kLargeCus = 64;        // default compute units a slice leaves to the large route
// Default route threshold: a workgroup per game pays off while the game's values stay in
// LDS or close to it; from this many vertices the device-wide solver serves it sooner.
constexpr int kLargeN = 16384;
constexpr size_t kStagingInit = 96u << 20;   // largest pinned staging buffer made at creation
// Large-route solvers are created for at least this many vertices and edges (one creation
// covers games up to this size; larger ones grow the worker's solver).
constexpr int kLargeSolverN = 1 << 19;
constexpr int64_t kLargeSolverM = (int64_t)1 << 23;

// One device-wide solver of the large route (grown to the largest game it has seen).
struct LargeWorker {
    egs_solver *solver = nullptr;
    int n = 0;
    int64_t m = 0;
};

constexpr int kSliceKedges = 2048;   // default edge budget per workgroup per slice, in thousands
constexpr int kLargeCertifyEvery = 10;  // rounds before the first proof attempt on the large route

// Arena bytes of a small-route game: offsets, packed edges, values, rise stamps, potentials.
size_t footprint(int n, int64_t m, int wmax) {
    const int shift = egs::pack_shift(n, wmax);
    return al256(4 * ((size_t)n + 1)) + al256((shift ? 4 : 8) * (size_t)m) + al256(4 * (size_t)n) +
           al256(4 * (size_t)n) + al256(4 * (size_t)n);
}

// Arena bytes of a large-route game: offsets, edges (32-bit packed when `packed`, else
// successors and weights), values; never more than the small layout, so footprint() bounds
// both.
size_t footprint_large(int n, int64_t m, bool packed) {
    return al256(4 * ((size_t)n + 1)) + (packed ? 1 : 2) * al256(4 * (size_t)m) + al256(4 * (size_t)n);
}
}  // namespace

struct egs_server {
    int device = 0;
    hipStream_t compute = nullptr, copy = nullptr;
    hipEvent_t ev_upload = nullptr, ev_slice = nullptr, ev_a = nullptr;
    void *arena = nullptr;
    Arena alloc;
    int max_games = 0;
    void *d_small = nullptr;
    egs::GameDesc *d_descs = nullptr;
    egs::GameState *d_states = nullptr;
    int *d_order = nullptr, *d_counter = nullptr;
    std::vector<Slot> slots;
    egs::GameDesc *descs = nullptr;     // pinned host mirrors, so the slice's copies stay
    egs::GameState *states = nullptr;   // asynchronous
    int *order = nullptr;
    std::vector<int> active;
    // Pinned staging ring: a submit packs into the next buffer after that buffer's previous
    // upload has finished, so consecutive submits do not wait for each other's copies.
    static constexpr int kStaging = 6;
    void *staging[kStaging] = {};
    size_t staging_bytes[kStaging] = {};
    hipEvent_t staging_ev[kStaging] = {};
    int staging_next = 0;
    int lds_limit = 0, cus = 0;
    bool slice_running = false;
    double last_slice_ms = 0;
    int64_t admitted = 0;     // admissions so far (slot sequence numbers)
    int hi_slot = -1;         // highest slot index ever used (descriptor/state copies cover 0..hi_slot)
    bool slice_kernel = false;  // the running slice launched the many-games kernel
    int large_n = 0;          // games with n >= large_n take the device-wide route
    // Large route: persistent worker threads take games from large_queue, solve them on
    // their own solver and stream, and hand them back through large_done.  mu guards
    // large_queue, large_done, busy, stopping and large_opt; a worker writes only its
    // game's slot result fields, before the hand-off under mu.
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int> large_queue;        // admitted large games, oldest first
    std::vector<int> large_done;        // solved, not yet reported by wait()
    int busy = 0;                       // large games being solved
    bool stopping = false;
    egs_serve_options large_opt{};      // proof options of the latest launch()
    std::vector<LargeWorker> workers;   // one per thread
    std::vector<std::thread> threads;
    int large_workers = kLargeWorkers;
    egs::ReadOrderMeta *d_meta = nullptr;   // per slot (games submitted in read order)
    std::vector<Transient> transients;      // internalisation work space still in use
    std::vector<hipEvent_t> spare_events;
};

namespace {

// Take the next staging buffer with at least `need` bytes, once its previous upload is done;
// *buf receives its index (record staging_ev[*buf] after enqueueing the upload from it).
int staging_reserve(egs_server *s, size_t need, int *buf) {
    const int i = s->staging_next;
    s->staging_next = (i + 1) % egs_server::kStaging;
    SHIP(hipEventSynchronize(s->staging_ev[i]));
    *buf = i;
    if (s->staging_bytes[i] >= need) return EGS_OK;
    if (s->staging[i]) (void)hipHostFree(s->staging[i]);
    s->staging_bytes[i] = need + need / 2;
    if (hipHostMalloc(&s->staging[i], s->staging_bytes[i], hipHostMallocDefault) != hipSuccess) {
        s->staging[i] = nullptr;
        s->staging_bytes[i] = 0;
        return sfail(EGS_E_NOMEM, "pinned staging allocation failed");
    }
    return EGS_OK;
}

int solve_large(egs_server *s, LargeWorker &wk, int ticket, const egs_serve_options *o);

// (Re)create a worker's solver for at least n vertices and m edges.
int grow_solver(egs_server *s, LargeWorker &wk, int n, int64_t m) {
    if (wk.solver) egs_solver_destroy(wk.solver);
    wk.solver = nullptr;
    int nc = kLargeSolverN;
    while (nc < n) nc *= 2;
    int64_t mc = kLargeSolverM;
    while (mc < m) mc *= 2;
    int rc = egs_solver_create(s->device, nc, mc, &wk.solver);
    if (rc) return sfail(rc, "large-route solver: %s", egs_last_error());
    wk.n = nc;
    wk.m = mc;
    return EGS_OK;
}

// Worker thread of the large route: until the server stops, take the oldest queued game,
// wait for the uploads, solve it, hand it back.
void large_worker_main(egs_server *s, int wi) {
    (void)hipSetDevice(s->device);
    LargeWorker &wk = s->workers[wi];
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] { return s->stopping || !s->large_queue.empty(); });
        if (s->stopping) break;
        const int t = s->large_queue.front();
        s->large_queue.pop_front();
        const egs_serve_options o = s->large_opt;
        s->busy++;
        lk.unlock();
        int rc = hipStreamSynchronize(s->copy) == hipSuccess ? EGS_OK : sfail(EGS_E_HIP, "upload wait failed");
        if (!rc && s->slots[t].read_order) {        // owner count and MG from the device
            egs::ReadOrderMeta meta{};
            if (hipMemcpy(&meta, s->d_meta + t, sizeof(meta), hipMemcpyDeviceToHost) != hipSuccess)
                rc = sfail(EGS_E_HIP, "read-order meta read failed");
            s->slots[t].n0 = meta.n0;
            s->slots[t].mg = meta.mg;
        }
        if (!rc) rc = solve_large(s, wk, t, &o);
        Slot &sl = s->slots[t];
        sl.rc = rc;
        sl.err = rc ? s_err : "";
        lk.lock();
        s->busy--;
        s->large_done.push_back(t);
        s->cv_done.notify_all();
    }
}

int start_workers(egs_server *s) {
    if (!s->threads.empty()) return EGS_OK;
    s->workers.assign(s->large_workers, LargeWorker{});
    for (auto &wk : s->workers) {
        int rc = grow_solver(s, wk, kLargeSolverN, kLargeSolverM);
        if (rc) return rc;
    }
    for (int wi = 0; wi < s->large_workers; wi++) s->threads.emplace_back(large_worker_main, s, wi);
    return EGS_OK;
}

void stop_workers(egs_server *s) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads) t.join();
    s->threads.clear();
    for (auto &w : s->workers)
        if (w.solver) egs_solver_destroy(w.solver);
    s->workers.clear();
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = false;
}

// Large route: the checked CSR goes into an arena chunk (edges packed to 32 bits when the
// weights allow, like the small route); the game waits in large_queue.
int submit_large(egs_server *s, int ticket, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
                 const int32_t *succ, const int32_t *w) {
    const size_t ptr_b = 4 * ((size_t)n + 1), e_b = 4 * (size_t)m;
    const size_t o_e = al256(ptr_b), o_w = o_e + al256(e_b);
    int rc = start_workers(s);
    if (rc) return rc;
    int sb = 0;
    rc = staging_reserve(s, o_w + al256(e_b), &sb);
    if (rc) return rc;
    char *stg = (char *)s->staging[sb];
    int32_t wm = 0, dm = 0;
    int64_t bad = -1;
    const int shift_n = egs::pack_shift(n, 0);
    int chk = shift_n ? egs_pk_check_pack32(n, m, ptr, succ, w, shift_n, (uint32_t *)(stg + o_e), &wm, &dm, &bad)
                      : egs_pk_check(n, m, ptr, succ, w, &wm, &dm, &bad);
    switch (chk) {
    case EGS_PK_BAD_OFFSETS: return sfail(EGS_E_ARG, "egs_server_submit: CSR offsets must run 0..m");
    case EGS_PK_NO_EDGE: return sfail(EGS_E_ARG, "egs_server_submit: vertex %lld has no out-edge", (long long)bad);
    case EGS_PK_BAD_SUCC:
        return sfail(EGS_E_ARG, "egs_server_submit: successor out of range at edge %lld", (long long)bad);
    default: break;
    }
    const int shift = egs::pack_shift(n, wm);
    if (!shift) {                                  // weights too wide: successors and weights apart
        egs_pk_copy(stg + o_e, succ, e_b);
        egs_pk_copy(stg + o_w, w, e_b);
    }
    egs_pk_copy(stg, ptr, ptr_b);
    const size_t up = shift ? o_e + e_b : o_w + e_b;
    const size_t bytes = footprint_large(n, m, shift > 0);
    size_t at = 0;
    if (!s->alloc.alloc(bytes, at)) return sfail(EGS_E_CAPACITY, "arena full (%zu bytes wanted)", bytes);
    char *base = (char *)s->arena + at;
    SHIP(hipMemcpyAsync(base, stg, up, hipMemcpyHostToDevice, s->copy));
    SHIP(hipEventRecord(s->staging_ev[sb], s->copy));
    SHIP(hipEventRecord(s->ev_upload, s->copy));
    Slot &sl = s->slots[ticket];
    sl = Slot{};
    sl.used = true;
    sl.large = true;
    sl.at = at;
    sl.bytes = bytes;
    sl.n = n;
    sl.n0 = n0;
    sl.mg = mg;
    sl.m = m;
    sl.shift = shift;
    sl.wmax = wm;
    sl.maxdeg = dm;
    sl.seq = s->admitted++;
    s->hi_slot = std::max(s->hi_slot, ticket);
    egs::GameDesc d{};
    d.ptr = (const int *)base;
    d.values = (int *)(base + (shift ? o_w : o_w + al256(e_b)));
    d.n = n;
    d.m = m;
    s->descs[ticket] = d;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->large_queue.push_back(ticket);
    }
    s->cv_work.notify_one();
    return ticket;
}

// Free the internalisation work space whose kernels have finished.
void reclaim_transients(egs_server *s) {
    for (size_t i = 0; i < s->transients.size();) {
        Transient &t = s->transients[i];
        if (hipEventQuery(t.done) == hipSuccess) {
            s->alloc.release(t.at, t.bytes);
            s->spare_events.push_back(t.done);
            t = s->transients.back();
            s->transients.pop_back();
        } else {
            i++;
        }
    }
}

// Solve one queued large game with a worker's device-wide solver (on the worker's own
// stream and host thread); its values go to its arena chunk.  Uploads are complete (the
// caller synchronised the copy stream).
int solve_large(egs_server *s, LargeWorker &wk, int ticket, const egs_serve_options *o) {
    Slot &sl = s->slots[ticket];
    if (!wk.solver || wk.n < sl.n || wk.m < sl.m) {
        int rc = grow_solver(s, wk, sl.n, sl.m);
        if (rc) return rc;
    }
    const char *base = (const char *)s->arena + sl.at;
    const size_t o_e = al256(4 * ((size_t)sl.n + 1)), o_w = o_e + al256(4 * (size_t)sl.m);
    int rc = sl.shift ? egs_solver_load_packed(wk.solver, sl.n, sl.n0, sl.m, sl.mg, (const int32_t *)base,
                                               (const uint32_t *)(base + o_e), sl.shift, sl.wmax, sl.maxdeg)
                      : egs_solver_load(wk.solver, EGS_MEM_DEVICE, sl.n, sl.n0, sl.m, sl.mg, (const int32_t *)base,
                                        (const int32_t *)(base + o_e), (const int32_t *)(base + o_w));
    egs_options eo{};
    // auto: 10 rounds for these mid-size games (65K..500K vertices of the serving mix: 0.668 ms
    // per solve against 0.684 with the single-game default of 8)
    eo.certify_every = o->certify_every != 0 ? o->certify_every : kLargeCertifyEvery;
    eo.cert_max_passes = o->cert_max_passes;
    eo.max_phases = -1;
    if (!rc) rc = egs_solver_prepare(wk.solver, &eo);
    egs_stats st{};
    if (!rc) rc = egs_solver_iterate(wk.solver, -1, &st);
    if (rc < 0) return sfail(rc, "large-route solve: %s", egs_last_error());
    hipStream_t ss = (hipStream_t)egs_solver_stream(wk.solver);
    SHIP(hipMemcpyAsync(s->descs[ticket].values, egs_solver_device_values(wk.solver), 4 * (size_t)sl.n,
                        hipMemcpyDeviceToDevice, ss));
    SHIP(hipStreamSynchronize(ss));
    sl.sweeps = st.phases;
    sl.edges = st.edges_relaxed;
    return EGS_OK;
}

}  // namespace

extern "C" {

const char *egs_server_last_error(void) { return s_err; }

int64_t egs_server_footprint(int32_t n, int64_t m, int32_t wmax) { return (int64_t)footprint(n, m, wmax); }

int egs_server_create(int device, int64_t arena_bytes, int32_t max_games, egs_server **out) {
    if (!out || arena_bytes < 4096 || max_games < 1) return sfail(EGS_E_ARG, "egs_server_create: bad arguments");
    *out = nullptr;
    hipDeviceProp_t p;
    SHIP(hipGetDeviceProperties(&p, device));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return sfail(EGS_E_HIP, "device %d is %s; libegs.so is built for gfx950 only", device, p.gcnArchName);
    SHIP(hipSetDevice(device));
    egs_server *s = new egs_server();
    s->device = device;
    s->lds_limit = (int)p.sharedMemPerBlock;
    s->cus = p.multiProcessorCount;
    s->large_n = kLargeN;
    s->max_games = max_games;
    if (hipMalloc(&s->arena, (size_t)arena_bytes) != hipSuccess) {
        delete s;
        return sfail(EGS_E_NOMEM, "arena hipMalloc of %lld bytes failed", (long long)arena_bytes);
    }
    s->alloc.init((size_t)arena_bytes);
    const size_t small = al256(sizeof(egs::GameDesc) * max_games) + al256(sizeof(egs::GameState) * max_games) +
                         al256(4 * (size_t)max_games) + 256 + al256(sizeof(egs::ReadOrderMeta) * max_games);
    if (hipMalloc(&s->d_small, small) != hipSuccess) {
        (void)hipFree(s->arena);
        delete s;
        return sfail(EGS_E_NOMEM, "hipMalloc failed");
    }
    char *b = (char *)s->d_small;
    s->d_descs = (egs::GameDesc *)b;
    s->d_states = (egs::GameState *)(b + al256(sizeof(egs::GameDesc) * max_games));
    s->d_order = (int *)(b + al256(sizeof(egs::GameDesc) * max_games) + al256(sizeof(egs::GameState) * max_games));
    s->d_counter = (int *)((char *)s->d_order + al256(4 * (size_t)max_games));
    s->d_meta = (egs::ReadOrderMeta *)((char *)s->d_counter + 256);
    SHIP(hipMemset(s->d_meta, 0, sizeof(egs::ReadOrderMeta) * max_games));
    s->slots.resize(max_games);
    if (hipHostMalloc((void **)&s->descs, sizeof(egs::GameDesc) * max_games, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->states, sizeof(egs::GameState) * max_games, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->order, sizeof(int) * max_games, hipHostMallocDefault) != hipSuccess) {
        return sfail(EGS_E_NOMEM, "pinned host allocation failed");
    }
    memset(s->descs, 0, sizeof(egs::GameDesc) * max_games);
    memset(s->states, 0, sizeof(egs::GameState) * max_games);
    if (hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev_upload, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&s->ev_slice) != hipSuccess || hipEventCreate(&s->ev_a) != hipSuccess) {
        return sfail(EGS_E_HIP, "stream/event creation failed");
    }
    for (int i = 0; i < egs_server::kStaging; i++) {
        if (hipEventCreateWithFlags(&s->staging_ev[i], hipEventDisableTiming) != hipSuccess)
            return sfail(EGS_E_HIP, "event creation failed");
        // staging sized for typical requests up front, so admissions do not pay for growth
        const size_t init = std::min(kStagingInit, std::max((size_t)4 << 20, (size_t)arena_bytes / 64));
        if (hipHostMalloc(&s->staging[i], init, hipHostMallocDefault) != hipSuccess)
            return sfail(EGS_E_NOMEM, "pinned staging allocation failed");
        s->staging_bytes[i] = init;
    }
    *out = s;
    return EGS_OK;
}

int egs_server_destroy(egs_server *s) {
    if (!s) return EGS_OK;
    (void)hipSetDevice(s->device);
    if (s->compute) (void)hipStreamSynchronize(s->compute);
    if (s->copy) (void)hipStreamSynchronize(s->copy);
    stop_workers(s);
    if (s->arena) (void)hipFree(s->arena);
    if (s->d_small) (void)hipFree(s->d_small);
    for (int i = 0; i < egs_server::kStaging; i++) {
        if (s->staging[i]) (void)hipHostFree(s->staging[i]);
        if (s->staging_ev[i]) (void)hipEventDestroy(s->staging_ev[i]);
    }
    if (s->descs) (void)hipHostFree(s->descs);
    if (s->states) (void)hipHostFree(s->states);
    if (s->order) (void)hipHostFree(s->order);
    for (auto &t : s->transients) (void)hipEventDestroy(t.done);
    for (auto e : s->spare_events) (void)hipEventDestroy(e);
    if (s->ev_upload) (void)hipEventDestroy(s->ev_upload);
    if (s->ev_slice) (void)hipEventDestroy(s->ev_slice);
    if (s->ev_a) (void)hipEventDestroy(s->ev_a);
    if (s->compute) (void)hipStreamDestroy(s->compute);
    if (s->copy) (void)hipStreamDestroy(s->copy);
    delete s;
    return EGS_OK;
}

int egs_server_free_slots(egs_server *s) {
    int k = 0;
    for (const auto &sl : s->slots) k += !sl.used;
    return k;
}

int64_t egs_server_free_bytes(egs_server *s) { return s ? (int64_t)(s->alloc.total - s->alloc.in_use) : 0; }

int egs_server_set_route(egs_server *s, int32_t large_n, int32_t large_workers) {
    if (!s || large_n < 0 || large_workers < 0 || large_workers > 16)
        return sfail(EGS_E_ARG, "egs_server_set_route: bad arguments");
    const int nw = large_workers > 0 ? large_workers : kLargeWorkers;
    if (nw != s->large_workers && !s->threads.empty()) {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->large_queue.empty() || s->busy)
            return sfail(EGS_E_STATE, "egs_server_set_route: large games are in flight");
    }
    s->large_n = large_n > 0 ? large_n : kLargeN;
    if (nw != s->large_workers) {
        stop_workers(s);
        s->large_workers = nw;
    }
    SHIP(hipSetDevice(s->device));
    return start_workers(s);                   // solvers and threads exist before requests arrive
}

// Admit a game: returns its ticket, or EGS_E_CAPACITY when the arena or the slot table is
// full right now (the caller retries after fetching finished games).
int egs_server_submit(egs_server *s, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
                      const int32_t *succ, const int32_t *w) {
    if (!s || n < 1 || n0 < 0 || n0 > n || m < 1 || mg < 0 || !ptr || !succ || !w)
        return sfail(EGS_E_ARG, "egs_server_submit: bad game");
    int ticket = -1;
    for (int i = 0; i < s->max_games; i++)
        if (!s->slots[i].used) { ticket = i; break; }
    if (ticket < 0) return sfail(EGS_E_CAPACITY, "no free game slot");
    const bool large = n >= s->large_n;
    // cheap refusal before any packing: not even the smallest layout of this game fits now
    const size_t least = large ? footprint_large(n, m, egs::pack_shift(n, 0) > 0) : footprint(n, m, 0);
    if (!s->alloc.fits(least)) return sfail(EGS_E_CAPACITY, "arena full (%zu bytes wanted)", least);
    SHIP(hipSetDevice(s->device));
    if (large) return submit_large(s, ticket, n, n0, m, mg, ptr, succ, w);
    // one pass over the input: checks, max |w|, 32-bit packing (int2 when w does not fit)
    const size_t ptr_b = 4 * ((size_t)n + 1);
    int sb = 0;
    int rc = staging_reserve(s, al256(ptr_b) + 8 * (size_t)m, &sb);
    if (rc) return rc;
    char *stg = (char *)s->staging[sb];
    const int shift_n = egs::pack_shift(n, 0);
    int32_t wm = 0, dm = 0;
    int64_t bad = -1;
    uint32_t *pe = (uint32_t *)(stg + al256(ptr_b));
    const int chk = shift_n ? egs_pk_check_pack32(n, m, ptr, succ, w, shift_n, pe, &wm, &dm, &bad)
                            : egs_pk_check(n, m, ptr, succ, w, &wm, &dm, &bad);
    switch (chk) {
    case EGS_PK_BAD_OFFSETS: return sfail(EGS_E_ARG, "egs_server_submit: CSR offsets must run 0..m");
    case EGS_PK_NO_EDGE: return sfail(EGS_E_ARG, "egs_server_submit: vertex %lld has no out-edge", (long long)bad);
    case EGS_PK_BAD_SUCC:
        return sfail(EGS_E_ARG, "egs_server_submit: successor out of range at edge %lld", (long long)bad);
    default: break;
    }
    const int shift = egs::pack_shift(n, wm);
    if (!shift) egs_pk_pack64(m, succ, w, (int32_t *)pe);
    egs_pk_copy(stg, ptr, ptr_b);
    const size_t edge_b = (shift ? 4 : 8) * (size_t)m;
    const size_t bytes = footprint(n, m, wm);
    size_t at = 0;
    if (!s->alloc.alloc(bytes, at)) return sfail(EGS_E_CAPACITY, "arena full (%zu bytes wanted)", bytes);
    char *base = (char *)s->arena + at;
    egs::GameDesc d{};
    d.ptr = (const int *)base;
    d.edges = base + al256(ptr_b);
    d.values = (int *)(base + al256(ptr_b) + al256(edge_b));
    d.rise = (int *)((char *)d.values + al256(4 * (size_t)n));
    d.h = (int *)((char *)d.rise + al256(4 * (size_t)n));
    d.state = s->d_states + ticket;
    d.m = m;
    d.n = n;
    d.n0 = n0;
    d.mg = mg;
    d.wmax = wm;
    d.shift = shift;
    d.group = 1;
    SHIP(hipMemcpyAsync((void *)d.ptr, stg, ptr_b, hipMemcpyHostToDevice, s->copy));
    SHIP(hipMemcpyAsync((void *)d.edges, pe, edge_b, hipMemcpyHostToDevice, s->copy));
    SHIP(hipEventRecord(s->staging_ev[sb], s->copy));
    SHIP(hipMemsetAsync(s->d_states + ticket, 0, sizeof(egs::GameState), s->copy));
    SHIP(hipEventRecord(s->ev_upload, s->copy));
    Slot &sl = s->slots[ticket];
    sl = Slot{};
    sl.used = true;
    sl.at = at;
    sl.bytes = bytes;
    sl.n = n;
    sl.m = m;
    sl.seq = s->admitted++;
    s->hi_slot = std::max(s->hi_slot, ticket);
    s->descs[ticket] = d;
    memset(&s->states[ticket], 0, sizeof(egs::GameState));
    return ticket;
}

// Admit a game in read order: owner[n] (0 = owner 0), rptr[n+1], rsucc[m] (successors as
// read indices) and rw[m], as a .mpg file lists them.  The owner partition (owner-0 vertices
// first, each block in read order), the internal CSR and MG are made on the device by
// launch_internalise, where the reference's host driver makes them
// (src/egsolver.c:107-220,253-304); the host only validates the arrays on their way into
// staging.  Returns a ticket like egs_server_submit; fetch() returns the values in read
// order.
int egs_server_submit_read_order(egs_server *s, int32_t n, int64_t m, const uint8_t *owner, const int32_t *rptr,
                                 const int32_t *rsucc, const int32_t *rw) {
    if (!s || n < 1 || m < 1 || m >= 0x7fffffffLL || !owner || !rptr || !rsucc || !rw)
        return sfail(EGS_E_ARG, "egs_server_submit_read_order: bad game");
    reclaim_transients(s);
    int ticket = -1;
    for (int i = 0; i < s->max_games; i++)
        if (!s->slots[i].used) { ticket = i; break; }
    if (ticket < 0) return sfail(EGS_E_CAPACITY, "no free game slot");
    int32_t wm = 0, dm = 0;
    int64_t bad = -1;
    switch (egs_pk_check(n, m, rptr, rsucc, rw, &wm, &dm, &bad)) {
    case EGS_PK_BAD_OFFSETS: return sfail(EGS_E_ARG, "egs_server_submit_read_order: CSR offsets must run 0..m");
    case EGS_PK_NO_EDGE:
        return sfail(EGS_E_ARG, "egs_server_submit_read_order: vertex %lld has no out-edge", (long long)bad);
    case EGS_PK_BAD_SUCC:
        return sfail(EGS_E_ARG, "egs_server_submit_read_order: successor out of range at edge %lld", (long long)bad);
    default: break;
    }
    // MG <= n * max |w|; only past INT32_MAX can it overflow, and then it is summed here to
    // refuse the game as the reference does (src/egsolver.c:283-304)
    if ((int64_t)n * wm > 0x7fffffffLL) {
        int64_t mg = 0;
        for (int32_t v = 0; v < n; v++) {
            int64_t worst = 0;
            for (int32_t e = rptr[v]; e < rptr[v + 1]; e++) worst = std::max<int64_t>(worst, -(int64_t)rw[e]);
            mg += worst;
        }
        if (mg > 0x7fffffffLL) return sfail(EGS_E_ARG, "Error too high weights --> overflow (MG past INT32_MAX)");
    }
    SHIP(hipSetDevice(s->device));
    const bool large = n >= s->large_n;
    if (large) {
        int rc = start_workers(s);
        if (rc) return rc;
    }
    const int shift = egs::pack_shift(n, wm);
    const size_t ptr_b = 4 * ((size_t)n + 1), n1 = al256(ptr_b);
    const size_t body = large ? footprint_large(n, m, shift > 0) : footprint(n, m, wm);
    const size_t bytes = body + al256(4 * (size_t)n);            // + ext_of_int
    // work space: the read-order arrays, the partition's scans, the internal int2 edges, MG
    const size_t o_rptr = al256((size_t)n), o_rsucc = o_rptr + n1, o_rw = o_rsucc + al256(4 * (size_t)m);
    const size_t in_end = o_rw + al256(4 * (size_t)m);
    const size_t o_flag0 = in_end, o_pre0 = o_flag0 + n1, o_rev = o_pre0 + n1, o_deg = o_rev + n1;
    const size_t o_tiles = o_deg + n1, o_edges = o_tiles + al256(4 * ((size_t)n / 4096 + 4));
    const size_t o_mg = o_edges + al256(8 * (size_t)m), tbytes = o_mg + 256;
    if (!s->alloc.fits(bytes)) return sfail(EGS_E_CAPACITY, "arena full (%zu bytes wanted)", bytes);
    int sb = 0;
    int rc = staging_reserve(s, in_end, &sb);
    if (rc) return rc;
    char *stg = (char *)s->staging[sb];
    egs_pk_copy(stg, owner, (size_t)n);
    egs_pk_copy(stg + o_rptr, rptr, ptr_b);
    egs_pk_copy(stg + o_rsucc, rsucc, 4 * (size_t)m);
    egs_pk_copy(stg + o_rw, rw, 4 * (size_t)m);
    size_t at = 0, tat = 0;
    if (!s->alloc.alloc(bytes, at)) return sfail(EGS_E_CAPACITY, "arena full (%zu bytes wanted)", bytes);
    if (!s->alloc.alloc(tbytes, tat)) {
        s->alloc.release(at, bytes);
        return sfail(EGS_E_CAPACITY, "arena full (%zu bytes of work space wanted)", tbytes);
    }
    char *base = (char *)s->arena + at, *T = (char *)s->arena + tat;
    int *ext = (int *)(base + body);
    SHIP(hipMemcpyAsync(T, stg, in_end, hipMemcpyHostToDevice, s->copy));
    SHIP(hipEventRecord(s->staging_ev[sb], s->copy));
    egs::ReadOrderBuffers b{};
    b.owner = (const uint8_t *)T;
    b.rptr = (const int *)(T + o_rptr);
    b.rsucc = (const int *)(T + o_rsucc);
    b.rw = (const int *)(T + o_rw);
    b.flag0 = (int *)(T + o_flag0);
    b.pre0 = (int *)(T + o_pre0);
    b.rev = (int *)(T + o_rev);
    b.ext_of_int = ext;
    b.deg = (int *)(T + o_deg);
    b.scan_tiles = (int *)(T + o_tiles);
    unsigned long long *d_mg = (unsigned long long *)(T + o_mg);
    SHIP(egs::launch_internalise(b, n, (int *)base, (int2 *)(T + o_edges), d_mg, s->copy));
    const size_t o_e = n1, o_w = o_e + al256(4 * (size_t)m);
    hipLaunchKernelGGL(egs::k_ro_place_edges, dim3(egs::grid_for(m, 256, 4096)), dim3(256), 0, s->copy,
                       (const int2 *)(T + o_edges), m, shift, large && !shift ? 1 : 0, (void *)(base + o_e),
                       (int *)(base + o_w));
    hipLaunchKernelGGL(egs::k_ro_meta, dim3(1), dim3(64), 0, s->copy, (const int *)b.pre0, n,
                       (const unsigned long long *)d_mg, s->d_meta + ticket);
    SHIP(hipGetLastError());
    Transient tr{tat, tbytes, nullptr};
    if (!s->spare_events.empty()) {
        tr.done = s->spare_events.back();
        s->spare_events.pop_back();
    } else {
        SHIP(hipEventCreateWithFlags(&tr.done, hipEventDisableTiming));
    }
    SHIP(hipEventRecord(tr.done, s->copy));
    s->transients.push_back(tr);
    if (!large) SHIP(hipMemsetAsync(s->d_states + ticket, 0, sizeof(egs::GameState), s->copy));
    SHIP(hipEventRecord(s->ev_upload, s->copy));
    Slot &sl = s->slots[ticket];
    sl = Slot{};
    sl.used = true;
    sl.large = large;
    sl.at = at;
    sl.bytes = bytes;
    sl.n = n;
    sl.m = m;
    sl.shift = shift;
    sl.wmax = wm;
    sl.maxdeg = dm;
    sl.read_order = true;
    sl.ext_at = at + body;
    sl.seq = s->admitted++;
    s->hi_slot = std::max(s->hi_slot, ticket);
    egs::GameDesc d{};
    d.ptr = (const int *)base;
    d.m = m;
    d.n = n;
    if (large) {
        d.values = (int *)(base + (shift ? o_w : o_w + al256(4 * (size_t)m)));
        s->descs[ticket] = d;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->large_queue.push_back(ticket);
        }
        s->cv_work.notify_one();
        return ticket;
    }
    const size_t edge_b = (shift ? 4 : 8) * (size_t)m;
    d.edges = base + o_e;
    d.values = (int *)(base + o_e + al256(edge_b));
    d.rise = (int *)((char *)d.values + al256(4 * (size_t)n));
    d.h = (int *)((char *)d.rise + al256(4 * (size_t)n));
    d.state = s->d_states + ticket;
    d.wmax = wm;
    d.shift = shift;
    d.group = 1;                                   // n0 and MG: patched on the device at launch
    s->descs[ticket] = d;
    memset(&s->states[ticket], 0, sizeof(egs::GameState));
    return ticket;
}

// Start one slice over every resident unfinished game; returns how many are in it.
int egs_server_launch(egs_server *s, const egs_serve_options *o) {
    if (!s) return sfail(EGS_E_ARG, "null server");
    if (s->slice_running) return sfail(EGS_E_STATE, "a slice is already running");
    egs_serve_options def{};
    if (!o) o = &def;
    SHIP(hipSetDevice(s->device));
    egs_batch_options bo{};
    bo.lanes_per_edge_x16 = o->lanes_per_edge_x16;
    bo.group_min = o->group_min;
    bo.group_max = o->group_max;
    s->active.clear();
    for (int i = 0; i <= s->hi_slot; i++) {
        Slot &sl = s->slots[i];
        sl.in_slice = sl.used && !sl.finished && !sl.large;
        if (sl.in_slice) {
            s->descs[i].group = egs::lanes_for((double)sl.m / sl.n, bo);
            s->active.push_back(i);
        }
    }
    int large_in_flight;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->large_opt = *o;
        large_in_flight = (int)s->large_queue.size() + s->busy + (int)s->large_done.size();
    }
    if (s->active.empty()) {
        // nothing for the many-games kernel; wait() still reports the large route
        s->slice_kernel = false;
        s->slice_running = large_in_flight > 0;
        return large_in_flight;
    }
    // oldest admission first: every workgroup's edge budget bounds the slice, so the games
    // at the front of the order progress every slice and none waits behind a large one
    std::stable_sort(s->active.begin(), s->active.end(),
                     [&](int a, int c) { return s->slots[a].seq < s->slots[c].seq; });
    const int threads = o->threads > 0 ? o->threads : 1024;
    if (threads != 256 && threads != 512 && threads != 1024) return sfail(EGS_E_ARG, "threads must be 256, 512 or 1024");
    int lds_ints = o->lds_ints > 0 ? o->lds_ints : 1 << 30;   // default: all the LDS a workgroup may hold
    lds_ints = std::min(lds_ints, (s->lds_limit - 2048) / 4);
    // while large-route games are in flight the slice leaves compute units to their kernels
    const int free_cus = large_in_flight == 0 ? 0 : o->large_cus < 0 ? 0 : o->large_cus > 0 ? o->large_cus : kLargeCus;
    const int grid = std::max(1, std::max(1, s->cus - free_cus) * (o->blocks_per_cu > 0 ? o->blocks_per_cu : 1));
    const int64_t max_sweeps = o->slice_sweeps > 0 ? o->slice_sweeps : INT64_MAX;
    const int64_t wg_budget = 1000 * (int64_t)(o->slice_kedges > 0 ? o->slice_kedges : kSliceKedges);
    egs::CertArgs ca{o->certify_every == 0 ? egs::kGameCertifyEvery : (o->certify_every < 0 ? 0 : o->certify_every),
                     o->cert_max_passes > 0 ? o->cert_max_passes : 1024, egs::kGameFirstCap,
                     egs::kGameCandidates};
    // descriptors of the slice's games, then the uploads, then the kernel
    // (a previous slice's copies have completed: wait() synchronised on its event)
    std::copy(s->active.begin(), s->active.end(), s->order);
    SHIP(hipMemcpyAsync(s->d_descs, s->descs, sizeof(egs::GameDesc) * (size_t)(s->hi_slot + 1),
                        hipMemcpyHostToDevice, s->compute));
    SHIP(hipMemcpyAsync(s->d_order, s->order, 4 * s->active.size(), hipMemcpyHostToDevice, s->compute));
    SHIP(hipMemsetAsync(s->d_counter, 0, 4, s->compute));
    SHIP(hipStreamWaitEvent(s->compute, s->ev_upload, 0));
    bool any_ro = false;
    for (int i : s->active) any_ro |= s->slots[i].read_order;
    if (any_ro)   // owner counts and MGs of games internalised on the device
        hipLaunchKernelGGL(egs::k_ro_patch_descs, dim3(std::max(1, (s->hi_slot + 256) / 256)), dim3(256), 0,
                           s->compute, s->d_descs, (const egs::ReadOrderMeta *)s->d_meta, s->hi_slot + 1);
    SHIP(hipEventRecord(s->ev_a, s->compute));
    hipLaunchKernelGGL(egs::k_games, dim3(grid), dim3(threads), sizeof(int) * (size_t)lds_ints, s->compute,
                       (const egs::GameDesc *)s->d_descs, (const int *)s->d_order, (int)s->active.size(),
                       s->d_counter, lds_ints, wg_budget, max_sweeps, ca);
    SHIP(hipGetLastError());
    SHIP(hipMemcpyAsync(s->states, s->d_states, sizeof(egs::GameState) * (size_t)(s->hi_slot + 1),
                        hipMemcpyDeviceToHost, s->compute));
    SHIP(hipEventRecord(s->ev_slice, s->compute));
    s->slice_running = true;
    s->slice_kernel = true;
    return (int)s->active.size() + large_in_flight;
}

// Finish the running slice; writes the tickets that reached their fixpoint.
int egs_server_wait(egs_server *s, int32_t *done, int32_t cap) {
    if (!s) return sfail(EGS_E_ARG, "null server");
    if (!s->slice_running) return 0;
    int k = 0;
    s->last_slice_ms = 0;
    if (s->slice_kernel) {
        SHIP(hipEventSynchronize(s->ev_slice));
        float ms = 0.f;
        SHIP(hipEventElapsedTime(&ms, s->ev_a, s->ev_slice));
        s->last_slice_ms = ms;
        for (int i : s->active) {
            Slot &sl = s->slots[i];
            if (s->states[i].converged && !sl.finished) {
                sl.finished = true;
                if (done && k < cap) done[k] = i;
                k++;
            }
        }
    }
    // large-route completions; with no slice to wait for, block until one arrives
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (k == 0 && !s->slice_kernel)
            s->cv_done.wait(lk, [&] { return !s->large_done.empty() || (s->large_queue.empty() && s->busy == 0); });
        for (int t : s->large_done) {
            s->slots[t].finished = true;
            if (done && k < cap) done[k] = t;
            k++;
        }
        s->large_done.clear();
    }
    s->slice_running = false;
    s->slice_kernel = false;
    return k;
}

// 1 when wait() would return without blocking (no slice running, the slice's kernel and
// copies are done, or a large-route game is ready to report), else 0.
int egs_server_ready(egs_server *s) {
    if (!s) return sfail(EGS_E_ARG, "null server");
    if (!s->slice_running) return 1;
    if (s->slice_kernel) {
        const hipError_t q = hipEventQuery(s->ev_slice);
        if (q == hipErrorNotReady) return 0;
        if (q != hipSuccess) return sfail(EGS_E_HIP, "hipEventQuery: %s", hipGetErrorString(q));
        return 1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    return !s->large_done.empty() || (s->large_queue.empty() && s->busy == 0) ? 1 : 0;
}

double egs_server_last_slice_ms(egs_server *s) { return s ? s->last_slice_ms : 0.0; }

int egs_server_status(egs_server *s, int ticket) {
    if (!s || ticket < 0 || ticket >= s->max_games || !s->slots[ticket].used) return sfail(EGS_E_ARG, "bad ticket");
    return s->slots[ticket].finished ? 1 : 0;
}

// Copy a finished game's values out (if values != NULL) and free its slot and memory.
int egs_server_fetch(egs_server *s, int ticket, int32_t *values, int64_t *sweeps, int64_t *edges) {
    if (!s || ticket < 0 || ticket >= s->max_games || !s->slots[ticket].used) return sfail(EGS_E_ARG, "bad ticket");
    if (s->slice_running) return sfail(EGS_E_STATE, "wait for the running slice first");
    Slot &sl = s->slots[ticket];
    if (!sl.finished) return sfail(EGS_E_STATE, "game %d has not finished", ticket);
    if (sl.large && sl.rc) {
        const int rc = sl.rc;
        const std::string msg = sl.err;
        if (sl.read_order) SHIP(hipMemset(s->d_meta + ticket, 0, sizeof(egs::ReadOrderMeta)));
        s->alloc.release(sl.at, sl.bytes);
        sl = Slot{};
        return sfail(rc, "game %d: %s", ticket, msg.c_str());
    }
    SHIP(hipSetDevice(s->device));
    if (values && sl.read_order) {
        // internal vertex i is read index ext_of_int[i]
        std::vector<int32_t> v((size_t)sl.n), ext((size_t)sl.n);
        SHIP(hipMemcpy(v.data(), s->descs[ticket].values, 4 * (size_t)sl.n, hipMemcpyDeviceToHost));
        SHIP(hipMemcpy(ext.data(), (char *)s->arena + sl.ext_at, 4 * (size_t)sl.n, hipMemcpyDeviceToHost));
        for (int32_t i = 0; i < sl.n; i++) values[ext[i]] = v[i];
    } else if (values) {
        SHIP(hipMemcpy(values, s->descs[ticket].values, 4 * (size_t)sl.n, hipMemcpyDeviceToHost));
    }
    if (sl.read_order) SHIP(hipMemset(s->d_meta + ticket, 0, sizeof(egs::ReadOrderMeta)));
    reclaim_transients(s);
    if (sweeps) *sweeps = sl.large ? sl.sweeps : s->states[ticket].sweeps;
    if (edges) *edges = sl.large ? sl.edges : s->states[ticket].edges;
    s->alloc.release(sl.at, sl.bytes);
    sl = Slot{};
    return EGS_OK;
}

}  // extern "C"
