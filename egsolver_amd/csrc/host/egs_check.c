/*
 * egs_check.c -- host-side verification of a solver result (no GPU code).
 *
 * Two O(m) checks that share nothing with the device kernels:
 *   egs_check_fixpoint         lift(f) == f, the lifting of src/cpu_solver.c:300-305 with
 *                              OMINUS of src/common.h:270-272.  A fixpoint lies above the
 *                              least fixpoint, so the least fixpoint's TOP set is inside
 *                              {f == TOP}.
 *   egs_check_top_certificate  potentials that prove {f == TOP} is won by player 1, so the
 *                              TOP set is inside the least fixpoint's TOP set.
 * Together they pin the TOP set of f to the least fixpoint's exactly.
 */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>

#include "egs_host.h"

#define TOP INT_MAX

static inline int ominus(int a, int w, int mg) {
    if (a < TOP && (int64_t)a - w <= mg) {
        const int64_t d = (int64_t)a - w;
        return d > 0 ? (int)d : 0;
    }
    return TOP;
}

static int lift_of(int v, int n0, int mg, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                   const int32_t *f) {
    int best = v < n0 ? TOP : 0;
    for (int32_t e = ptr[v]; e < ptr[v + 1]; e++) {
        const int x = ominus(f[succ[e]], w[e], mg);
        if (v < n0) { if (x < best) best = x; }
        else        { if (x > best) best = x; }
    }
    return best;
}

int64_t egs_check_fixpoint(int32_t n, int32_t n0, int32_t mg, const int32_t *ptr, const int32_t *succ,
                           const int32_t *w, const int32_t *values, int64_t *first_bad) {
    int64_t bad = 0, first = INT64_MAX;
#pragma omp parallel for schedule(static, 4096) reduction(+ : bad) reduction(min : first)
    for (int32_t v = 0; v < n; v++) {
        if (ptr[v + 1] <= ptr[v] || lift_of(v, n0, mg, ptr, succ, w, values) != values[v]) {
            bad++;
            if (v < first) first = v;
        }
    }
    if (first_bad) *first_bad = bad ? first : -1;
    return bad;
}

/* What is wrong with vertex v's part of the certificate (EGS_CERT_OK when nothing). */
static int cert_vertex(int32_t v, int32_t n, int32_t n0, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                       const int32_t *values, const int64_t *h) {
    const int top = values[v] == TOP;
    if (top != (h[v] >= 0)) return EGS_CERT_MEMBERSHIP;
    if (!top) return EGS_CERT_OK;
    const __int128 scale = (__int128)n + 1;
    int found = 0;
    for (int32_t e = ptr[v]; e < ptr[v + 1]; e++) {
        const int32_t u = succ[e];
        const int in = values[u] == TOP && h[u] >= 0;
        const int tight = in && (__int128)h[v] >= scale * w[e] + 1 + h[u];
        if (v < n0) {
            if (!in) return EGS_CERT_LEAVES;
            if (!tight) return EGS_CERT_POTENTIAL;
        } else if (tight) {
            found = 1;
            break;
        }
    }
    if (v >= n0 && !found) return EGS_CERT_NO_EDGE;
    if (v < n0 && ptr[v + 1] <= ptr[v]) return EGS_CERT_NO_EDGE;
    return EGS_CERT_OK;
}

int egs_check_top_certificate(int32_t n, int32_t n0, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                              const int32_t *values, const int64_t *h, int64_t *bad_vertex) {
    int64_t first = INT64_MAX;
#pragma omp parallel for schedule(static, 4096) reduction(min : first)
    for (int32_t v = 0; v < n; v++)
        if (cert_vertex(v, n, n0, ptr, succ, w, values, h) != EGS_CERT_OK && v < first) first = v;
    if (first == INT64_MAX) {
        if (bad_vertex) *bad_vertex = -1;
        return EGS_CERT_OK;
    }
    if (bad_vertex) *bad_vertex = first;
    return cert_vertex((int32_t)first, n, n0, ptr, succ, w, values, h);
}
