// egs_batch.hip -- many independent energy games solved in one persistent launch.
//
// No reference counterpart: the reference solves one game per process.  Each game's
// values are the same least fixpoint the single-game path and the reference's
// src/cpu_solver.c give.
//
// Games are staged on the host, then copied into one device block: per game its CSR
// offsets, its edges packed to 32 bits as (w << shift) | succ when the vertex count and
// weights allow (else int2), its values and proof scratch.  One launch of k_games
// (egs_games.h) solves them all; games are ordered by estimated work, largest first.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/egs.h"
#include "egs_games.h"
#include "egs_internal.h"
#include "egs_pack.h"

namespace egs {

// Bits for successor ids of an n-vertex game, or 0 when weights up to wmax do not fit the
// remaining signed bits (the game then keeps 64-bit edges).
int pack_shift(int n, int wmax) {
    int shift = 1;
    while ((1 << shift) < n) shift++;
    if (shift > 30) return 0;
    const int64_t lim = 1LL << (31 - shift);          // signed range of 32 - shift bits
    return ((int64_t)wmax < lim) ? shift : 0;
}

int lanes_for(double avg_deg, const egs_batch_options &o) {
    const int x16 = o.lanes_per_edge_x16 > 0 ? o.lanes_per_edge_x16 : 2;
    const int gmax = o.group_max > 0 ? o.group_max : 64;
    const int gmin = o.group_min > 0 ? o.group_min : 1;
    int G = 1;
    while (G < gmax && G < avg_deg * x16 / 16.0) G <<= 1;
    return std::max(gmin, std::min(G, gmax));
}

}  // namespace egs

namespace {
thread_local char b_err[512];
int bfail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b_err, sizeof(b_err), fmt, ap);
    va_end(ap);
    return code;
}
#define BHIP(expr)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bfail(EGS_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

struct Staged {
    int n, n0, mg, wmax;
    int64_t m;
    size_t ptr_at, succ_at;    // offsets into the host vectors
};
}  // namespace

struct egs_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    std::vector<Staged> games;
    std::vector<int32_t> h_ptr, h_succ, h_w;
    std::vector<egs::GameDesc> descs;
    std::vector<egs::GameState> states;
    void *d_block = nullptr;
    egs::GameDesc *d_descs = nullptr;
    egs::GameState *d_states = nullptr;
    int *d_order = nullptr, *d_counter = nullptr;
    bool uploaded = false;
    int lds_limit = 0, cus = 0;
    int packed_games = 0;
    int *d_values = nullptr;     // every game's values, contiguous in game order
    int64_t total_n = 0;
    // read-back staging: two pinned chunks, filled by the device while the host copies
    // the other one out
    int32_t *h_stage[2] = {nullptr, nullptr};
    hipEvent_t ev_stage[2] = {nullptr, nullptr};
    int32_t *h_values = nullptr;   // pinned copy of every value (egs_batch_read_all_pinned)
    int64_t h_values_n = 0;
};

constexpr size_t kStageInts = (size_t)2 << 20;   // 8 MB per read-back chunk

static void batch_free_device(egs_batch *b) {
    if (b->d_block) (void)hipFree(b->d_block);
    b->d_block = nullptr;
    b->d_values = nullptr;
    b->total_n = 0;
    b->uploaded = false;
}

extern "C" {

const char *egs_batch_last_error(void) { return b_err; }

int egs_batch_create(int device, egs_batch **out) {
    if (!out) return bfail(EGS_E_ARG, "null output");
    *out = nullptr;
    hipDeviceProp_t p;
    BHIP(hipGetDeviceProperties(&p, device));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return bfail(EGS_E_HIP, "device %d is %s; libegs.so is built for gfx950 only", device, p.gcnArchName);
    BHIP(hipSetDevice(device));
    egs_batch *b = new egs_batch();
    b->device = device;
    b->lds_limit = (int)p.sharedMemPerBlock;
    b->cus = p.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev_a) != hipSuccess || hipEventCreate(&b->ev_b) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_stage[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_stage[1], hipEventDisableTiming) != hipSuccess) {
        delete b;
        return bfail(EGS_E_HIP, "stream/event creation failed");
    }
    *out = b;
    return EGS_OK;
}

int egs_batch_destroy(egs_batch *b) {
    if (!b) return EGS_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free_device(b);
    if (b->ev_a) (void)hipEventDestroy(b->ev_a);
    if (b->ev_b) (void)hipEventDestroy(b->ev_b);
    for (int k = 0; k < 2; k++) {
        if (b->ev_stage[k]) (void)hipEventDestroy(b->ev_stage[k]);
        if (b->h_stage[k]) (void)hipHostFree(b->h_stage[k]);
    }
    if (b->h_values) (void)hipHostFree(b->h_values);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return EGS_OK;
}

int egs_batch_clear(egs_batch *b) {
    if (!b) return bfail(EGS_E_ARG, "null batch");
    b->games.clear();
    b->h_ptr.clear();
    b->h_succ.clear();
    b->h_w.clear();
    b->descs.clear();
    b->states.clear();
    b->uploaded = false;
    return EGS_OK;
}

int egs_batch_add(egs_batch *b, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
                  const int32_t *succ, const int32_t *w) {
    if (!b || n < 1 || n0 < 0 || n0 > n || m < 1 || mg < 0 || !ptr || !succ || !w)
        return bfail(EGS_E_ARG, "egs_batch_add: bad game");
    int32_t wm = 0, dm = 0;
    int64_t bad = -1;
    switch (egs_pk_check(n, m, ptr, succ, w, &wm, &dm, &bad)) {
    case EGS_PK_BAD_OFFSETS: return bfail(EGS_E_ARG, "egs_batch_add: CSR offsets must run 0..m");
    case EGS_PK_NO_EDGE: return bfail(EGS_E_ARG, "egs_batch_add: vertex %lld has no out-edge", (long long)bad);
    case EGS_PK_BAD_SUCC: return bfail(EGS_E_ARG, "egs_batch_add: successor out of range at edge %lld", (long long)bad);
    default: break;
    }
    Staged g{n, n0, mg, wm, m, b->h_ptr.size(), b->h_succ.size()};
    b->games.push_back(g);
    b->h_ptr.insert(b->h_ptr.end(), ptr, ptr + n + 1);
    b->h_succ.insert(b->h_succ.end(), succ, succ + m);
    b->h_w.insert(b->h_w.end(), w, w + m);
    b->uploaded = false;
    return (int)b->games.size() - 1;
}

int egs_batch_size(egs_batch *b) { return b ? (int)b->games.size() : 0; }

int egs_batch_upload(egs_batch *b) {
    if (!b || b->games.empty()) return bfail(EGS_E_STATE, "egs_batch_upload: no games");
    BHIP(hipSetDevice(b->device));
    batch_free_device(b);
    const size_t ng = b->games.size();
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // all values first (contiguous, game order: one copy reads them back), then per game:
    // ptr, edges, rise stamps, h
    std::vector<size_t> at(ng);
    std::vector<int> shift(ng);
    int64_t total_n = 0;
    for (size_t i = 0; i < ng; i++) total_n += b->games[i].n;
    const size_t o_values = al(sizeof(egs::GameDesc) * ng) + al(sizeof(egs::GameState) * ng) + al(4 * ng) + 256;
    size_t off = o_values + al(4 * (size_t)total_n);
    for (size_t i = 0; i < ng; i++) {
        const Staged &g = b->games[i];
        shift[i] = egs::pack_shift(g.n, g.wmax);
        at[i] = off;
        off += al(4 * ((size_t)g.n + 1)) + al((shift[i] ? 4 : 8) * (size_t)g.m) + al(4 * (size_t)g.n) +
               al(4 * (size_t)g.n);
    }
    if (hipMalloc(&b->d_block, off) != hipSuccess) return bfail(EGS_E_NOMEM, "hipMalloc of %zu bytes failed", off);
    char *base = (char *)b->d_block;
    b->d_descs = (egs::GameDesc *)base;
    b->d_states = (egs::GameState *)(base + al(sizeof(egs::GameDesc) * ng));
    b->d_order = (int *)(base + al(sizeof(egs::GameDesc) * ng) + al(sizeof(egs::GameState) * ng));
    b->d_counter = (int *)((char *)b->d_order + al(4 * ng));
    b->d_values = (int *)(base + o_values);
    b->total_n = total_n;
    b->descs.assign(ng, egs::GameDesc{});
    b->states.assign(ng, egs::GameState{});
    b->packed_games = 0;
    std::vector<uint32_t> packed;
    std::vector<int2> wide;
    int64_t vat = 0;
    for (size_t i = 0; i < ng; i++) {
        const Staged &g = b->games[i];
        char *p = base + at[i];
        egs::GameDesc &d = b->descs[i];
        d.values = b->d_values + vat;
        vat += g.n;
        d.ptr = (const int *)p;
        p += al(4 * ((size_t)g.n + 1));
        d.edges = p;
        p += al((shift[i] ? 4 : 8) * (size_t)g.m);
        d.rise = (int *)p;
        p += al(4 * (size_t)g.n);
        d.h = (int *)p;
        d.state = b->d_states + i;
        d.m = g.m;
        d.n = g.n;
        d.n0 = g.n0;
        d.mg = g.mg;
        d.wmax = g.wmax;
        d.shift = shift[i];
        d.group = 1;
        BHIP(hipMemcpy((void *)d.ptr, b->h_ptr.data() + g.ptr_at, 4 * ((size_t)g.n + 1), hipMemcpyHostToDevice));
        const int32_t *su = b->h_succ.data() + g.succ_at, *wv = b->h_w.data() + g.succ_at;
        if (shift[i]) {
            packed.resize(g.m);
            egs_pk_pack32(g.m, su, wv, shift[i], packed.data());
            BHIP(hipMemcpy((void *)d.edges, packed.data(), 4 * (size_t)g.m, hipMemcpyHostToDevice));
            b->packed_games++;
        } else {
            wide.resize(g.m);
            egs_pk_pack64(g.m, su, wv, (int32_t *)wide.data());
            BHIP(hipMemcpy((void *)d.edges, wide.data(), 8 * (size_t)g.m, hipMemcpyHostToDevice));
        }
    }
    b->uploaded = true;
    return EGS_OK;
}

int egs_batch_solve(egs_batch *b, const egs_batch_options *o, egs_batch_stats *out) {
    if (!b || !b->uploaded) return bfail(EGS_E_STATE, "egs_batch_solve: upload first");
    egs_batch_options def{};
    if (!o) o = &def;
    BHIP(hipSetDevice(b->device));
    const int threads = o->threads > 0 ? o->threads : 1024;
    if (threads != 256 && threads != 512 && threads != 1024)
        return bfail(EGS_E_ARG, "threads must be 256, 512 or 1024");
    int lds_ints = o->lds_ints > 0 ? o->lds_ints : 1 << 30;   // default: all the LDS a workgroup may hold
    lds_ints = std::min(lds_ints, (b->lds_limit - 2048) / 4);
    const int per_cu = o->blocks_per_cu > 0 ? o->blocks_per_cu : 1;
    const int grid = std::max(1, b->cus * per_cu);
    const int ng = (int)b->games.size();
    std::vector<int> order(ng);
    std::vector<double> cost(ng);
    for (int i = 0; i < ng; i++) {
        egs::GameDesc &d = b->descs[i];
        d.group = egs::lanes_for((double)d.m / d.n, *o);
        cost[i] = (double)d.n * (double)d.m;
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    BHIP(hipMemcpyAsync(b->d_descs, b->descs.data(), sizeof(egs::GameDesc) * ng, hipMemcpyHostToDevice, b->stream));
    BHIP(hipMemsetAsync(b->d_states, 0, sizeof(egs::GameState) * ng, b->stream));
    BHIP(hipMemcpyAsync(b->d_order, order.data(), 4 * (size_t)ng, hipMemcpyHostToDevice, b->stream));
    BHIP(hipMemsetAsync(b->d_counter, 0, 4, b->stream));
    const int64_t max_sweeps = o->max_sweeps > 0 ? o->max_sweeps : INT64_MAX;
    egs::CertArgs ca{o->certify_every == 0 ? egs::kGameCertifyEvery : (o->certify_every < 0 ? 0 : o->certify_every),
                     o->cert_max_passes > 0 ? o->cert_max_passes : 1024, egs::kGameFirstCap,
                     egs::kGameCandidates};
    BHIP(hipEventRecord(b->ev_a, b->stream));
    hipLaunchKernelGGL(egs::k_games, dim3(grid), dim3(threads), sizeof(int) * (size_t)lds_ints, b->stream,
                       (const egs::GameDesc *)b->d_descs, (const int *)b->d_order, ng, b->d_counter, lds_ints,
                       (int64_t)INT64_MAX, max_sweeps, ca);
    BHIP(hipGetLastError());
    BHIP(hipEventRecord(b->ev_b, b->stream));
    BHIP(hipEventSynchronize(b->ev_b));
    float ms = 0.f;
    BHIP(hipEventElapsedTime(&ms, b->ev_a, b->ev_b));
    BHIP(hipMemcpy(b->states.data(), b->d_states, sizeof(egs::GameState) * ng, hipMemcpyDeviceToHost));
    if (out) {
        memset(out, 0, sizeof(*out));
        out->games = ng;
        out->ms = ms;
        out->packed_edges = b->packed_games == ng ? 1 : 0;
        out->grid = grid;
        out->threads = threads;
        out->lds_ints = lds_ints;
        for (const auto &s : b->states) {
            out->sweeps += s.sweeps;
            out->edges_relaxed += s.edges;
            out->lifted += s.lifted;
            out->converged += s.converged;
            out->certificates += s.certificates;
            out->cert_passes += s.cert_passes;
        }
    }
    return EGS_OK;
}

int egs_batch_game_stats(egs_batch *b, int game, int64_t *sweeps, int64_t *edges, int32_t *converged) {
    if (!b || game < 0 || game >= (int)b->states.size()) return bfail(EGS_E_ARG, "no stats for game %d", game);
    const auto &s = b->states[game];
    if (sweeps) *sweeps = s.sweeps;
    if (edges) *edges = s.edges;
    if (converged) *converged = s.converged;
    return EGS_OK;
}

int egs_batch_read_values(egs_batch *b, int game, int32_t *out) {
    if (!b || !b->uploaded || game < 0 || game >= (int)b->games.size()) return bfail(EGS_E_ARG, "bad game %d", game);
    BHIP(hipSetDevice(b->device));
    BHIP(hipMemcpy(out, b->descs[game].values, 4 * (size_t)b->games[game].n, hipMemcpyDeviceToHost));
    return EGS_OK;
}

int egs_batch_read_all(egs_batch *b, int32_t *out) {
    if (!b || !b->uploaded) return bfail(EGS_E_STATE, "nothing uploaded");
    BHIP(hipSetDevice(b->device));
    for (int k = 0; k < 2; k++)
        if (!b->h_stage[k] && hipHostMalloc((void **)&b->h_stage[k], 4 * kStageInts, hipHostMallocDefault) != hipSuccess)
            return bfail(EGS_E_NOMEM, "hipHostMalloc of the read-back staging failed");
    const int64_t chunks = (b->total_n + (int64_t)kStageInts - 1) / (int64_t)kStageInts;
    auto len = [&](int64_t c) { return (size_t)std::min<int64_t>((int64_t)kStageInts, b->total_n - c * (int64_t)kStageInts); };
    for (int64_t c = 0; c <= chunks; c++) {
        if (c < chunks) {   // the buffer's previous chunk (c - 2) was copied out in step c - 1
            BHIP(hipMemcpyAsync(b->h_stage[c & 1], b->d_values + c * (int64_t)kStageInts, 4 * len(c),
                                hipMemcpyDeviceToHost, b->stream));
            BHIP(hipEventRecord(b->ev_stage[c & 1], b->stream));
        }
        if (c > 0) {
            BHIP(hipEventSynchronize(b->ev_stage[(c - 1) & 1]));
            egs_pk_copy(out + (c - 1) * (int64_t)kStageInts, b->h_stage[(c - 1) & 1], 4 * len(c - 1));
        }
    }
    return EGS_OK;
}

int egs_batch_read_all_pinned(egs_batch *b, const int32_t **values) {
    if (!b || !b->uploaded || !values) return bfail(EGS_E_STATE, "nothing uploaded");
    BHIP(hipSetDevice(b->device));
    if (b->h_values_n < b->total_n) {
        if (b->h_values) (void)hipHostFree(b->h_values);
        b->h_values = nullptr;
        b->h_values_n = 0;
        if (hipHostMalloc((void **)&b->h_values, 4 * (size_t)b->total_n, hipHostMallocDefault) != hipSuccess)
            return bfail(EGS_E_NOMEM, "hipHostMalloc of %lld values failed", (long long)b->total_n);
        b->h_values_n = b->total_n;
    }
    BHIP(hipMemcpyAsync(b->h_values, b->d_values, 4 * (size_t)b->total_n, hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    *values = b->h_values;
    return EGS_OK;
}

}  // extern "C"
