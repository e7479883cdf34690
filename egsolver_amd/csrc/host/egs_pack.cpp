// egs_pack.cpp -- OpenMP-parallel host checks and packing of a game's CSR (egs_pack.h).
//
// The fast paths are single vectorisable passes with OR/MAX reductions; the index of an
// offending vertex or edge is searched for only after a pass has seen one.
#include "../egs_pack.h"

#include <omp.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace {
constexpr int64_t kParallelMin = 1 << 15;   // below this many items one thread is faster

// Team size of the passes: memory-bound, they saturate at a few threads, and a larger team
// oversubscribes the cores that a serving process also needs for its solver threads and
// the HIP runtime (16 threads on a 16-core share made admissions several times slower).
inline int team() { return std::min(8, omp_get_max_threads()); }

// Degree check over the offsets: max out-degree, or the first vertex without an edge.
int check_offsets(int32_t n, int64_t m, const int32_t *ptr, int32_t *maxdeg, int64_t *bad) {
    if (ptr[0] != 0 || ptr[n] != m) return EGS_PK_BAD_OFFSETS;
    int32_t dm = 0, dmin = INT32_MAX;
#pragma omp parallel for num_threads(team()) if (n > kParallelMin) schedule(static) reduction(max : dm) reduction(min : dmin)
    for (int32_t v = 0; v < n; v++) {
        const int32_t d = ptr[v + 1] - ptr[v];
        dm = std::max(dm, d);
        dmin = std::min(dmin, d);
    }
    if (dmin < 1) {
        for (int32_t v = 0; v < n; v++)
            if (ptr[v + 1] - ptr[v] < 1) {
                *bad = v;
                break;
            }
        return EGS_PK_NO_EDGE;
    }
    *maxdeg = dm;
    return EGS_PK_OK;
}

int64_t first_bad_succ(int32_t n, int64_t m, const int32_t *succ) {
    for (int64_t e = 0; e < m; e++)
        if ((uint32_t)succ[e] >= (uint32_t)n) return e;
    return -1;
}

inline uint32_t absw(int32_t w) { return w < 0 ? 0u - (uint32_t)w : (uint32_t)w; }
}  // namespace

extern "C" int egs_pk_check(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                            int32_t *wmax, int32_t *maxdeg, int64_t *bad) {
    *bad = -1;
    int rc = check_offsets(n, m, ptr, maxdeg, bad);
    if (rc) return rc;
    uint32_t wm = 0, sbad = 0;
#pragma omp parallel for num_threads(team()) if (m > kParallelMin) schedule(static) reduction(max : wm) reduction(| : sbad)
    for (int64_t e = 0; e < m; e++) {
        sbad |= (uint32_t)((uint32_t)succ[e] >= (uint32_t)n);
        wm = std::max(wm, absw(w[e]));
    }
    if (sbad) {
        *bad = first_bad_succ(n, m, succ);
        return EGS_PK_BAD_SUCC;
    }
    *wmax = (int32_t)std::min<uint32_t>(wm, INT32_MAX);
    return EGS_PK_OK;
}

extern "C" int egs_pk_check_pack32(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ,
                                   const int32_t *w, int shift, uint32_t *out, int32_t *wmax, int32_t *maxdeg,
                                   int64_t *bad) {
    *bad = -1;
    int rc = check_offsets(n, m, ptr, maxdeg, bad);
    if (rc) return rc;
    uint32_t wm = 0, sbad = 0;
#pragma omp parallel for num_threads(team()) if (m > kParallelMin) schedule(static) reduction(max : wm) reduction(| : sbad)
    for (int64_t e = 0; e < m; e++) {
        const int32_t s = succ[e], x = w[e];
        sbad |= (uint32_t)((uint32_t)s >= (uint32_t)n);
        wm = std::max(wm, absw(x));
        out[e] = ((uint32_t)x << shift) | (uint32_t)s;
    }
    if (sbad) {
        *bad = first_bad_succ(n, m, succ);
        return EGS_PK_BAD_SUCC;
    }
    *wmax = (int32_t)std::min<uint32_t>(wm, INT32_MAX);
    return EGS_PK_OK;
}

extern "C" int egs_pk_check_pack64(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ,
                                   const int32_t *w, int32_t *out, int32_t *wmax, int32_t *maxdeg, int64_t *bad) {
    *bad = -1;
    int rc = check_offsets(n, m, ptr, maxdeg, bad);
    if (rc) return rc;
    uint32_t wm = 0, sbad = 0;
#pragma omp parallel for num_threads(team()) if (m > kParallelMin) schedule(static) reduction(max : wm) reduction(| : sbad)
    for (int64_t e = 0; e < m; e++) {
        const int32_t s = succ[e], x = w[e];
        sbad |= (uint32_t)((uint32_t)s >= (uint32_t)n);
        wm = std::max(wm, absw(x));
        out[2 * e] = s;
        out[2 * e + 1] = x;
    }
    if (sbad) {
        *bad = first_bad_succ(n, m, succ);
        return EGS_PK_BAD_SUCC;
    }
    *wmax = (int32_t)std::min<uint32_t>(wm, INT32_MAX);
    return EGS_PK_OK;
}

extern "C" int egs_pk_check_copy(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                                 int32_t *out_ptr, int32_t *out_succ, int32_t *out_w, int32_t *wmax,
                                 int32_t *maxdeg, int64_t *bad) {
    *bad = -1;
    int rc = check_offsets(n, m, ptr, maxdeg, bad);
    if (rc) return rc;
    egs_pk_copy(out_ptr, ptr, 4 * ((size_t)n + 1));
    uint32_t wm = 0, sbad = 0;
#pragma omp parallel for num_threads(team()) if (m > kParallelMin) schedule(static) reduction(max : wm) reduction(| : sbad)
    for (int64_t e = 0; e < m; e++) {
        const int32_t s = succ[e], x = w[e];
        sbad |= (uint32_t)((uint32_t)s >= (uint32_t)n);
        wm = std::max(wm, absw(x));
        out_succ[e] = s;
        out_w[e] = x;
    }
    if (sbad) {
        *bad = first_bad_succ(n, m, succ);
        return EGS_PK_BAD_SUCC;
    }
    *wmax = (int32_t)std::min<uint32_t>(wm, INT32_MAX);
    return EGS_PK_OK;
}

extern "C" void egs_pk_pack32(int64_t m, const int32_t *succ, const int32_t *w, int shift, uint32_t *out) {
#pragma omp parallel for num_threads(team()) if (m > kParallelMin) schedule(static)
    for (int64_t e = 0; e < m; e++) out[e] = ((uint32_t)w[e] << shift) | (uint32_t)succ[e];
}

extern "C" void egs_pk_pack64(int64_t m, const int32_t *succ, const int32_t *w, int32_t *out) {
#pragma omp parallel for num_threads(team()) if (m > kParallelMin) schedule(static)
    for (int64_t e = 0; e < m; e++) {
        out[2 * e] = succ[e];
        out[2 * e + 1] = w[e];
    }
}

extern "C" void egs_pk_copy(void *dst, const void *src, size_t bytes) {
    constexpr size_t kChunk = 64 << 10;
    const int64_t chunks = (int64_t)((bytes + kChunk - 1) / kChunk);
#pragma omp parallel for num_threads(team()) if (bytes > 4 * kChunk) schedule(static)
    for (int64_t c = 0; c < chunks; c++) {
        const size_t at = (size_t)c * kChunk;
        memcpy((char *)dst + at, (const char *)src + at, std::min(kChunk, bytes - at));
    }
}
