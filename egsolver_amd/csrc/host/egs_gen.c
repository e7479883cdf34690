/*
 * egs_gen.c -- seeded synthetic energy games.
 *
 * The reference ships only two generators (clique and ladder families,
 * script/multi/mkclique.c:42-65 and mkladder.c:42-51) and no random-game generator, so
 * the random families are this project's own and frozen here:
 *
 *   PRNG     xoshiro256** seeded through SplitMix64.  Game g of a family with seed s
 *            draws from the stream seeded with splitmix64(s * 0x9E3779B97F4A7C15 + g).
 *   bounded  Lemire's multiply-shift with rejection: unbiased integers in [0, k).
 *   vertex   for v = 0..n-1 in external-id order: owner (1 bit), out-degree, then for
 *            each edge a successor (uniform over V\{v}, redrawn on a repeat) and a
 *            weight uniform in [-W, W].
 *
 * Families (degree rule per vertex):
 *   EGS_DEG_FIXED     d = deg_lo
 *   EGS_DEG_UNIFORM   d uniform in [deg_lo, deg_hi]
 *   EGS_DEG_POWERLAW  P(d) proportional to d^-alpha on [deg_lo, deg_hi] (inverse CDF)
 *   EGS_DEG_COMPLETE  every other vertex is a successor, ascending id order
 * Degrees are clipped to n-1.  Arrays are written in read order (external ids), the
 * order a .mpg file lists them; egs_internalise() applies the owner partition.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "egs_host.h"

static inline uint64_t sm64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s[4]; } rng_t;

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    uint64_t out = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return out;
}

static void rng_seed(rng_t *r, uint64_t seed, uint64_t stream) {
    uint64_t x = seed * 0x9E3779B97F4A7C15ULL + stream;
    uint64_t y = sm64(&x);
    for (int i = 0; i < 4; i++) r->s[i] = sm64(&y);
}

/* Unbiased integer in [0, k), k >= 1 (Lemire 2019). */
static inline uint64_t rng_below(rng_t *r, uint64_t k) {
    uint64_t x = rng_next(r) >> 32;           /* 32 random bits: k < 2^32 here */
    uint64_t m = x * k;
    uint32_t l = (uint32_t)m;
    if (l < (uint32_t)k) {
        uint32_t t = (uint32_t)(-(uint32_t)k) % (uint32_t)k;
        while (l < t) {
            x = rng_next(r) >> 32;
            m = x * k;
            l = (uint32_t)m;
        }
    }
    return m >> 32;
}

static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

uint64_t egs_rng_test_stream(uint64_t seed, uint64_t stream, int count, uint64_t *out) {
    rng_t r;
    rng_seed(&r, seed, stream);
    for (int i = 0; i < count; i++) out[i] = rng_next(&r);
    return (uint64_t)count;
}

/* Draw the out-degree of one vertex. */
static int draw_degree(rng_t *r, const egs_gen_params *p, const double *cdf, int n) {
    int d;
    switch (p->deg_kind) {
    case EGS_DEG_UNIFORM:
        d = p->deg_lo + (int)rng_below(r, (uint64_t)(p->deg_hi - p->deg_lo + 1));
        break;
    case EGS_DEG_POWERLAW: {
        double u = rng_unit(r);
        int lo = 0, hi = p->deg_hi - p->deg_lo;   /* smallest k with cdf[k] > u */
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (cdf[mid] > u) hi = mid; else lo = mid + 1;
        }
        d = p->deg_lo + lo;
        break;
    }
    case EGS_DEG_COMPLETE:
        d = n - 1;
        break;
    default:
        d = p->deg_lo;
    }
    if (d > n - 1) d = n - 1;
    if (d < 1) d = 1;
    return d;
}

static double *powerlaw_cdf(const egs_gen_params *p) {
    int k = p->deg_hi - p->deg_lo + 1;
    double *cdf = (double *)malloc(sizeof(double) * (size_t)k);
    double acc = 0.0;
    for (int i = 0; i < k; i++) { acc += pow((double)(p->deg_lo + i), -p->alpha); cdf[i] = acc; }
    for (int i = 0; i < k; i++) cdf[i] /= acc;
    cdf[k - 1] = 2.0;   /* every u in [0,1) maps inside the table */
    return cdf;
}

int64_t egs_gen_count_edges(const egs_gen_params *p, uint64_t stream) {
    /* Degrees are drawn interleaved with the edges, so the count replays the stream. */
    egs_gen_read_order out = {0};
    return egs_gen_game(p, stream, &out, 1);
}

/* Open-addressing set for distinct-successor draws of high-degree vertices. */
typedef struct { int *slot; int cap; } iset;

static int iset_add(iset *s, int key) {          /* 1 if inserted, 0 if present */
    uint32_t h = (uint32_t)key * 2654435761u;
    int mask = s->cap - 1;
    for (int i = (int)(h & (uint32_t)mask);; i = (i + 1) & mask) {
        if (s->slot[i] == -1) { s->slot[i] = key; return 1; }
        if (s->slot[i] == key) return 0;
    }
}

int64_t egs_gen_game(const egs_gen_params *p, uint64_t stream, egs_gen_read_order *out,
                     int count_only) {
    const int n = p->n;
    rng_t r;
    rng_seed(&r, p->seed, stream);
    double *cdf = (p->deg_kind == EGS_DEG_POWERLAW) ? powerlaw_cdf(p) : NULL;
    const uint64_t wspan = (uint64_t)(2 * (int64_t)p->wmax + 1);
    int64_t e = 0;
    int *pick = NULL;
    iset set = {NULL, 0};
    int pick_cap = 0;
    if (!count_only) out->ptr[0] = 0;
    for (int v = 0; v < n; v++) {
        int owner = (int)(rng_next(&r) >> 63);
        int d = draw_degree(&r, p, cdf, n);
        if (!count_only) out->owner[v] = (uint8_t)owner;
        if (p->deg_kind == EGS_DEG_COMPLETE) {
            for (int u = 0; u < n; u++) {
                if (u == v) continue;
                int wv = (int)((int64_t)rng_below(&r, wspan) - p->wmax);
                if (!count_only) { out->succ[e] = u; out->w[e] = wv; }
                e++;
            }
        } else {
            if (d > pick_cap) {
                pick_cap = d;
                pick = (int *)realloc(pick, sizeof(int) * (size_t)pick_cap);
            }
            int use_set = d > 32;
            if (use_set) {
                int cap = 64;
                while (cap < 2 * d) cap <<= 1;
                if (cap > set.cap) { free(set.slot); set.slot = (int *)malloc(sizeof(int) * (size_t)cap); set.cap = cap; }
                memset(set.slot, 0xff, sizeof(int) * (size_t)set.cap);
            }
            for (int k = 0; k < d; k++) {
                int u;
                for (;;) {
                    u = (int)rng_below(&r, (uint64_t)(n - 1));
                    if (u >= v) u++;
                    int fresh = 1;
                    if (use_set) fresh = iset_add(&set, u);
                    else for (int j = 0; j < k; j++) if (pick[j] == u) { fresh = 0; break; }
                    if (fresh) break;
                }
                pick[k] = u;
                int wv = (int)((int64_t)rng_below(&r, wspan) - p->wmax);
                if (!count_only) { out->succ[e] = u; out->w[e] = wv; }
                e++;
            }
        }
        if (!count_only) out->ptr[v + 1] = (int32_t)e;
    }
    free(pick);
    free(set.slot);
    free(cdf);
    return e;
}

/* Owner partition and internal CSR (the reference's remap_instance/postparsing with its
 * default ordering, src/egsolver.c:146-214,253-281): owner-0 vertices keep their read
 * order and come first, owner-1 vertices follow in read order; successor ids are renamed
 * through the same map.  Here read index == external id.  Returns n0. */
int egs_internalise(int n, const uint8_t *owner, const int32_t *rptr, const int32_t *rsucc,
                    const int32_t *rw, int32_t *ptr, int32_t *succ, int32_t *w,
                    int32_t *ext_of_int) {
    int32_t *rev = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int k = 0;
    for (int i = 0; i < n; i++) if (owner[i] == 0) { ext_of_int[k] = i; rev[i] = k; k++; }
    int n0 = k;
    for (int i = 0; i < n; i++) if (owner[i] != 0) { ext_of_int[k] = i; rev[i] = k; k++; }
    int64_t e = 0;
    for (int i = 0; i < n; i++) {
        ptr[i] = (int32_t)e;
        e += rptr[ext_of_int[i] + 1] - rptr[ext_of_int[i]];
    }
    ptr[n] = (int32_t)e;
#pragma omp parallel for schedule(dynamic, 4096) if (e > (1 << 16))
    for (int i = 0; i < n; i++) {
        const int r = ext_of_int[i];
        int64_t at = ptr[i];
        for (int j = rptr[r]; j < rptr[r + 1]; j++, at++) { succ[at] = rev[rsucc[j]]; w[at] = rw[j]; }
    }
    free(rev);
    return n0;
}

/* MG = sum over vertices of max(0, max_e -w_e) (src/egsolver.c:283-304).  Returns -1
 * when the sum would pass INT32_MAX (the reference aborts on that input). */
int egs_compute_mg(int n, const int32_t *ptr, const int32_t *w, int32_t *mg_out) {
    /* every term is >= 0, so the running sum passes INT32_MAX iff the total does */
    int64_t mg = 0;
#pragma omp parallel for schedule(static) reduction(+ : mg) if (n > (1 << 16))
    for (int v = 0; v < n; v++) {
        int64_t worst = 0;
        for (int e = ptr[v]; e < ptr[v + 1]; e++) if (-(int64_t)w[e] > worst) worst = -(int64_t)w[e];
        mg += worst;
    }
    if (mg > INT32_MAX) return -1;
    *mg_out = (int32_t)mg;
    return 0;
}

/* Write a game in the .mpg arena format read by src/the_parser.y:128-160, one vertex per
 * line in read order: "id MAX|MIN s:w,...;".  gz != 0 writes gzip. */
int egs_write_mpg(const char *path, int n, const uint8_t *owner, const int32_t *rptr,
                  const int32_t *rsucc, const int32_t *rw, int gz) {
    gzFile zf = NULL;
    FILE *f = NULL;
    if (gz) { zf = gzopen(path, "wb1"); if (!zf) return -1; }
    else { f = fopen(path, "w"); if (!f) return -1; }
    size_t cap = 1 << 20, len = 0;
    char *buf = (char *)malloc(cap);
#define FLUSHBUF() do { if (gz) gzwrite(zf, buf, (unsigned)len); else fwrite(buf, 1, len, f); len = 0; } while (0)
    len += (size_t)snprintf(buf + len, cap - len, "%d;\n", n - 1);
    for (int v = 0; v < n; v++) {
        if (cap - len < 64) FLUSHBUF();
        len += (size_t)snprintf(buf + len, cap - len, "%d %s ", v, owner[v] ? "MIN" : "MAX");
        for (int j = rptr[v]; j < rptr[v + 1]; j++) {
            if (cap - len < 64) FLUSHBUF();
            len += (size_t)snprintf(buf + len, cap - len, "%d:%d%c", rsucc[j], rw[j],
                                    j + 1 < rptr[v + 1] ? ',' : ';');
        }
        buf[len++] = '\n';
    }
    FLUSHBUF();
#undef FLUSHBUF
    free(buf);
    if (gz) gzclose(zf); else fclose(f);
    return 0;
}

static int log_uniform_int(rng_t *r, double lo, double hi) {
    double u = rng_unit(r);
    double x = exp(log(lo) + u * (log(hi) - log(lo)));
    int v = (int)floor(x + 0.5);
    if (v < (int)lo) v = (int)lo;
    if (v > (int)hi) v = (int)hi;
    return v;
}

void egs_gen_family_params(int family, uint64_t seed, int index, egs_gen_params *out) {
    rng_t r;
    rng_seed(&r, seed, 0xFFFFFFFF00000000ULL | (uint64_t)(uint32_t)index);
    double nhi = (family == EGS_FAMILY_SERVE) ? 500000.0 : 50000.0;
    int n = log_uniform_int(&r, 1000.0, nhi);
    int wmax = log_uniform_int(&r, 10.0, 1000.0);
    int clique = (index % 16) == 15;
    int k = 200 + (int)rng_below(&r, 301);
    memset(out, 0, sizeof(*out));
    out->seed = seed;
    out->wmax = wmax;
    if (clique) {
        out->n = k;
        out->deg_kind = EGS_DEG_COMPLETE;
        out->deg_lo = out->deg_hi = k - 1;
    } else {
        out->n = n;
        out->deg_kind = EGS_DEG_UNIFORM;
        out->deg_lo = 2;
        out->deg_hi = 16;
    }
}

long egs_max_loop(long m, long n, long mg, long maxloop) {
    long res = maxloop;
    if (res < 0L) {
        if (mg > (0x7fffffffffffffffL - 1) / (n * m)) res = 0x7fffffffffffffffL - 1;
        else res = n * m * mg + 1;
    }
    return res;
}
