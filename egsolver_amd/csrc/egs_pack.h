// egs_pack.h -- host-side CSR checks and edge packing shared by the solver, batch and
// server entry points (egs_pack.cpp, OpenMP-parallel above a few thousand edges).
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { EGS_PK_OK = 0, EGS_PK_BAD_OFFSETS = -1, EGS_PK_NO_EDGE = -2, EGS_PK_BAD_SUCC = -3 };

// Checks that ptr runs from 0 to m, that every vertex has an out-edge and that every
// successor lies in [0, n).  Fills *wmax = max |w| and *maxdeg = max out-degree.  On a
// failure returns one of the negative codes above with *bad = the first offending vertex
// (EGS_PK_NO_EDGE) or edge (EGS_PK_BAD_SUCC) index.
int egs_pk_check(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ, const int32_t *w, int32_t *wmax,
                 int32_t *maxdeg, int64_t *bad);

// egs_pk_check fused with the packing or copying of the edges, so the edge arrays are read
// once: out = ((uint32_t)w << shift) | succ (the caller repacks as int2 when the returned
// wmax does not fit 32 - shift signed bits), out = int2 {succ, w}, or plain copies of the
// three arrays.  Outputs are written even when the check fails.
int egs_pk_check_pack32(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ, const int32_t *w, int shift,
                        uint32_t *out, int32_t *wmax, int32_t *maxdeg, int64_t *bad);
int egs_pk_check_pack64(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                        int32_t *out, int32_t *wmax, int32_t *maxdeg, int64_t *bad);
int egs_pk_check_copy(int32_t n, int64_t m, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                      int32_t *out_ptr, int32_t *out_succ, int32_t *out_w, int32_t *wmax, int32_t *maxdeg,
                      int64_t *bad);

// out[e] = ((uint32_t)w[e] << shift) | succ[e]   (w must fit in 32 - shift signed bits)
void egs_pk_pack32(int64_t m, const int32_t *succ, const int32_t *w, int shift, uint32_t *out);

// out[2e] = succ[e], out[2e+1] = w[e]   (the int2 edge layout)
void egs_pk_pack64(int64_t m, const int32_t *succ, const int32_t *w, int32_t *out);

// memcpy split over threads (staging copies of large games)
void egs_pk_copy(void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
