// egs_pack.cpp -- OpenMP-parallel host checks and packing of a game's CSR (egs_pack.h).
#include "../egs_pack.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace {
constexpr int64_t kParallelMin = 1 << 15;   // below this many edges one thread is faster
}

extern "C" int egs_pk_check(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                            int32_t *wmax, int32_t *maxdeg, int64_t *bad) {
    *bad = -1;
    if (ptr[0] != 0 || ptr[n] != m) return EGS_PK_BAD_OFFSETS;
    int64_t first_v = INT64_MAX, first_e = INT64_MAX;
    int32_t wm = 0, dm = 0;
#pragma omp parallel for if (m > kParallelMin) schedule(static) reduction(min : first_v) reduction(max : dm)
    for (int32_t v = 0; v < n; v++) {
        const int32_t d = ptr[v + 1] - ptr[v];
        if (d < 1 && v < first_v) first_v = v;
        dm = std::max(dm, d);
    }
    if (first_v != INT64_MAX) {
        *bad = first_v;
        return EGS_PK_NO_EDGE;
    }
#pragma omp parallel for if (m > kParallelMin) schedule(static) reduction(min : first_e) reduction(max : wm)
    for (int64_t e = 0; e < m; e++) {
        if ((succ[e] < 0 || succ[e] >= n) && e < first_e) first_e = e;
        const int64_t a = std::llabs((long long)w[e]);
        wm = std::max(wm, (int32_t)std::min<int64_t>(a, INT32_MAX));
    }
    if (first_e != INT64_MAX) {
        *bad = first_e;
        return EGS_PK_BAD_SUCC;
    }
    *wmax = wm;
    *maxdeg = dm;
    return EGS_PK_OK;
}

extern "C" void egs_pk_pack32(int64_t m, const int32_t *succ, const int32_t *w, int shift, uint32_t *out) {
#pragma omp parallel for if (m > kParallelMin) schedule(static)
    for (int64_t e = 0; e < m; e++) out[e] = ((uint32_t)w[e] << shift) | (uint32_t)succ[e];
}

extern "C" void egs_pk_pack64(int64_t m, const int32_t *succ, const int32_t *w, int32_t *out) {
#pragma omp parallel for if (m > kParallelMin) schedule(static)
    for (int64_t e = 0; e < m; e++) {
        out[2 * e] = succ[e];
        out[2 * e + 1] = w[e];
    }
}

extern "C" void egs_pk_copy(void *dst, const void *src, size_t bytes) {
    constexpr size_t kChunk = 1 << 20;
    const int64_t chunks = (int64_t)((bytes + kChunk - 1) / kChunk);
#pragma omp parallel for if (bytes > 4 * kChunk) schedule(static)
    for (int64_t c = 0; c < chunks; c++) {
        const size_t at = (size_t)c * kChunk;
        memcpy((char *)dst + at, (const char *)src + at, std::min(kChunk, bytes - at));
    }
}
