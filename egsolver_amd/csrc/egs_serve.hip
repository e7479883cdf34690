// egs_serve.hip -- serving loop: games arrive over time, are admitted into a device
// arena by memory footprint, stay resident (graph, values, proof state) across time
// slices of the many-games kernel, and leave when their fixpoint is reached.
//
// No reference counterpart.  Each game's values are the least fixpoint the single-game
// path gives (the slices resume the same computation; see GameState in egs_games.h).
//
// Host protocol: submit() packs a game into pinned staging and copies it into its arena
// chunk on the copy stream (overlapping a running slice); launch() starts one slice over
// every resident unfinished game on the compute stream after the pending uploads;
// wait() ends the slice and reports the games that finished; fetch() copies a result
// out and frees the chunk.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "../../include/egs.h"
#include "egs_games.h"
#include "egs_pack.h"

namespace egs {
int pack_shift(int n, int wmax);
int lanes_for(double avg_deg, const egs_batch_options &o);
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
    size_t at = 0, bytes = 0;
    int n = 0;
    int64_t m = 0;
    int64_t seq = 0;      // admission order
};

constexpr int kSliceKedges = 2048;   // default edge budget per workgroup per slice, in thousands

size_t footprint(int n, int64_t m, int wmax) {
    const int shift = egs::pack_shift(n, wmax);
    return al256(4 * ((size_t)n + 1)) + al256((shift ? 4 : 8) * (size_t)m) + al256(4 * (size_t)n) +
           al256(4 * (size_t)n) + al256(8 * (size_t)n);
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
    void *staging = nullptr;
    size_t staging_bytes = 0;
    int lds_limit = 0, cus = 0;
    bool slice_running = false;
    double last_slice_ms = 0;
    int64_t admitted = 0;     // admissions so far (slot sequence numbers)
    int hi_slot = -1;         // highest slot index ever used (descriptor/state copies cover 0..hi_slot)
};

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
    s->max_games = max_games;
    if (hipMalloc(&s->arena, (size_t)arena_bytes) != hipSuccess) {
        delete s;
        return sfail(EGS_E_NOMEM, "arena hipMalloc of %lld bytes failed", (long long)arena_bytes);
    }
    s->alloc.init((size_t)arena_bytes);
    const size_t small = al256(sizeof(egs::GameDesc) * max_games) + al256(sizeof(egs::GameState) * max_games) +
                         al256(4 * (size_t)max_games) + 256;
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
    *out = s;
    return EGS_OK;
}

int egs_server_destroy(egs_server *s) {
    if (!s) return EGS_OK;
    (void)hipSetDevice(s->device);
    if (s->compute) (void)hipStreamSynchronize(s->compute);
    if (s->copy) (void)hipStreamSynchronize(s->copy);
    if (s->arena) (void)hipFree(s->arena);
    if (s->d_small) (void)hipFree(s->d_small);
    if (s->staging) (void)hipHostFree(s->staging);
    if (s->descs) (void)hipHostFree(s->descs);
    if (s->states) (void)hipHostFree(s->states);
    if (s->order) (void)hipHostFree(s->order);
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

// Admit a game: returns its ticket, or EGS_E_CAPACITY when the arena or the slot table is
// full right now (the caller retries after fetching finished games).
int egs_server_submit(egs_server *s, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
                      const int32_t *succ, const int32_t *w) {
    if (!s || n < 1 || n0 < 0 || n0 > n || m < 1 || mg < 0 || !ptr || !succ || !w)
        return sfail(EGS_E_ARG, "egs_server_submit: bad game");
    int32_t wm = 0, dm = 0;
    int64_t bad = -1;
    switch (egs_pk_check(n, m, ptr, succ, w, &wm, &dm, &bad)) {
    case EGS_PK_BAD_OFFSETS: return sfail(EGS_E_ARG, "egs_server_submit: CSR offsets must run 0..m");
    case EGS_PK_NO_EDGE: return sfail(EGS_E_ARG, "egs_server_submit: vertex %lld has no out-edge", (long long)bad);
    case EGS_PK_BAD_SUCC:
        return sfail(EGS_E_ARG, "egs_server_submit: successor out of range at edge %lld", (long long)bad);
    default: break;
    }
    int ticket = -1;
    for (int i = 0; i < s->max_games; i++)
        if (!s->slots[i].used) { ticket = i; break; }
    if (ticket < 0) return sfail(EGS_E_CAPACITY, "no free game slot");
    const size_t bytes = footprint(n, m, wm);
    size_t at = 0;
    if (!s->alloc.alloc(bytes, at)) return sfail(EGS_E_CAPACITY, "arena full (%zu bytes wanted)", bytes);
    SHIP(hipSetDevice(s->device));
    const int shift = egs::pack_shift(n, wm);
    const size_t ptr_b = 4 * ((size_t)n + 1), edge_b = (shift ? 4 : 8) * (size_t)m;
    if (s->staging_bytes < ptr_b + edge_b) {
        SHIP(hipStreamSynchronize(s->copy));
        if (s->staging) (void)hipHostFree(s->staging);
        s->staging_bytes = (ptr_b + edge_b) * 2;
        if (hipHostMalloc(&s->staging, s->staging_bytes, hipHostMallocDefault) != hipSuccess) {
            s->staging = nullptr;
            s->staging_bytes = 0;
            s->alloc.release(at, bytes);
            return sfail(EGS_E_NOMEM, "pinned staging allocation failed");
        }
    }
    // the previous upload must be done before the staging buffer is overwritten
    SHIP(hipStreamSynchronize(s->copy));
    char *stg = (char *)s->staging;
    memcpy(stg, ptr, ptr_b);
    if (shift)
        egs_pk_pack32(m, succ, w, shift, (uint32_t *)(stg + ptr_b));
    else
        egs_pk_pack64(m, succ, w, (int32_t *)(stg + ptr_b));
    char *base = (char *)s->arena + at;
    egs::GameDesc d{};
    d.ptr = (const int *)base;
    d.edges = base + al256(ptr_b);
    d.values = (int *)(base + al256(ptr_b) + al256(edge_b));
    d.sigma = (int *)((char *)d.values + al256(4 * (size_t)n));
    d.h = (long long *)((char *)d.sigma + al256(4 * (size_t)n));
    d.state = s->d_states + ticket;
    d.m = m;
    d.n = n;
    d.n0 = n0;
    d.mg = mg;
    d.wmax = wm;
    d.shift = shift;
    d.group = 1;
    SHIP(hipMemcpyAsync((void *)d.ptr, stg, ptr_b, hipMemcpyHostToDevice, s->copy));
    SHIP(hipMemcpyAsync((void *)d.edges, stg + ptr_b, edge_b, hipMemcpyHostToDevice, s->copy));
    SHIP(hipMemsetAsync(s->d_states + ticket, 0, sizeof(egs::GameState), s->copy));
    SHIP(hipEventRecord(s->ev_upload, s->copy));
    Slot &sl = s->slots[ticket];
    sl.used = true;
    sl.finished = false;
    sl.in_slice = false;
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
    for (int i = 0; i < s->max_games; i++) {
        Slot &sl = s->slots[i];
        sl.in_slice = sl.used && !sl.finished;
        if (sl.in_slice) {
            s->descs[i].group = egs::lanes_for((double)sl.m / sl.n, bo);
            s->active.push_back(i);
        }
    }
    if (s->active.empty()) return 0;
    // oldest admission first: every workgroup's edge budget bounds the slice, so the games
    // at the front of the order progress every slice and none waits behind a large one
    std::stable_sort(s->active.begin(), s->active.end(),
                     [&](int a, int c) { return s->slots[a].seq < s->slots[c].seq; });
    const int threads = o->threads > 0 ? o->threads : 1024;
    if (threads != 256 && threads != 512 && threads != 1024) return sfail(EGS_E_ARG, "threads must be 256, 512 or 1024");
    int lds_ints = o->lds_ints > 0 ? o->lds_ints : 1 << 30;   // default: all the LDS a workgroup may hold
    lds_ints = std::min(lds_ints, (s->lds_limit - 2048) / 4);
    const int grid = std::max(1, s->cus * (o->blocks_per_cu > 0 ? o->blocks_per_cu : 1));
    const int64_t max_sweeps = o->slice_sweeps > 0 ? o->slice_sweeps : INT64_MAX;
    const int64_t wg_budget = 1000 * (int64_t)(o->slice_kedges > 0 ? o->slice_kedges : kSliceKedges);
    egs::CertArgs ca{o->certify_every == 0 ? 16 : (o->certify_every < 0 ? 0 : o->certify_every),
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
    SHIP(hipEventRecord(s->ev_a, s->compute));
    hipLaunchKernelGGL(egs::k_games, dim3(grid), dim3(threads), sizeof(int) * (size_t)lds_ints, s->compute,
                       (const egs::GameDesc *)s->d_descs, (const int *)s->d_order, (int)s->active.size(),
                       s->d_counter, lds_ints, wg_budget, max_sweeps, ca);
    SHIP(hipGetLastError());
    SHIP(hipMemcpyAsync(s->states, s->d_states, sizeof(egs::GameState) * (size_t)(s->hi_slot + 1),
                        hipMemcpyDeviceToHost, s->compute));
    SHIP(hipEventRecord(s->ev_slice, s->compute));
    s->slice_running = true;
    return (int)s->active.size();
}

// Finish the running slice; writes the tickets that reached their fixpoint.
int egs_server_wait(egs_server *s, int32_t *done, int32_t cap) {
    if (!s) return sfail(EGS_E_ARG, "null server");
    if (!s->slice_running) return 0;
    SHIP(hipEventSynchronize(s->ev_slice));
    float ms = 0.f;
    SHIP(hipEventElapsedTime(&ms, s->ev_a, s->ev_slice));
    s->last_slice_ms = ms;
    s->slice_running = false;
    int k = 0;
    for (int i : s->active) {
        Slot &sl = s->slots[i];
        if (s->states[i].converged && !sl.finished) {
            sl.finished = true;
            if (done && k < cap) done[k] = i;
            k++;
        }
    }
    return k;
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
    SHIP(hipSetDevice(s->device));
    if (values) SHIP(hipMemcpy(values, s->descs[ticket].values, 4 * (size_t)sl.n, hipMemcpyDeviceToHost));
    if (sweeps) *sweeps = s->states[ticket].sweeps;
    if (edges) *edges = s->states[ticket].edges;
    s->alloc.release(sl.at, sl.bytes);
    sl = Slot{};
    return EGS_OK;
}

}  // extern "C"
