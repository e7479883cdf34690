/*
 * egs_csr.c -- binary cache of an internalised game (no counterpart in the reference).
 *
 * Reading the .mpg text and partitioning by owner (src/the_parser.y, src/egsolver.c:107-315)
 * costs far more than a solve on the device; a game saved once in its internal numbering
 * loads again at file-read speed.  Layout, little-endian, int32 arrays:
 *   64-byte header: magic "EGSCSR1", version, header size, n, n0, m, MG, max |w|, checksum
 *   ptr[n+1]  succ[m]  w[m]  ext_of_int[n]
 * The checksum covers the header fields n, n0, m, MG and max |w| and the four arrays as one
 * sequence of 32-bit words: each block of kBlockWords words is hashed on its own (so blocks
 * hash in parallel) and the block hashes are folded in order, then the header fields.  A
 * load checks the header against the file size, the checksum, the CSR itself (offsets from
 * 0 to m, non-decreasing, an edge per vertex, successors in range, external ids in range)
 * and recomputes MG and max |w| from the weights, so a damaged or foreign file is refused
 * (a wrong n0 or MG would otherwise change every value the solver returns).
 */
#include <limits.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "egs_host.h"

#define EGS_CSR_MAGIC "EGSCSR1"
#define EGS_CSR_VERSION 2u

typedef struct {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int32_t n, n0;
    int64_t m;
    int32_t mg, max_abs_w;
    uint64_t checksum;
    uint8_t reserved[16];
} csr_header;

_Static_assert(sizeof(csr_header) == 64, "the header is 64 bytes");

static const int64_t kBlockWords = 1 << 18;

static void cerr(char *buf, int len, const char *fmt, ...) {
    if (!buf || len <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, (size_t)len, fmt, ap);
    va_end(ap);
}

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* Hash of the word sequence a[0..na) b[0..nb) c[0..nc) d[0..nd). */
static uint64_t words_hash(const int32_t *a, int64_t na, const int32_t *b, int64_t nb, const int32_t *c, int64_t nc,
                           const int32_t *d, int64_t nd) {
    const int64_t total = na + nb + nc + nd;
    const int64_t blocks = (total + kBlockWords - 1) / kBlockWords;
    uint64_t *bh = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(blocks > 0 ? blocks : 1));
    if (!bh) return 0;
#pragma omp parallel for schedule(static) if (blocks > 4)
    for (int64_t k = 0; k < blocks; k++) {
        uint64_t h = 0x9e3779b97f4a7c15ull ^ (uint64_t)k;
        const int64_t lo = k * kBlockWords, hi = lo + kBlockWords < total ? lo + kBlockWords : total;
        for (int64_t i = lo; i < hi; i++) {
            const int32_t x = i < na ? a[i] : i < na + nb ? b[i - na] : i < na + nb + nc ? c[i - na - nb] : d[i - na - nb - nc];
            h = (h ^ (uint32_t)x) * 0x100000001b3ull;
        }
        bh[k] = mix64(h);
    }
    uint64_t H = mix64((uint64_t)total);
    for (int64_t k = 0; k < blocks; k++) H = mix64(H ^ bh[k]);
    free(bh);
    return H;
}

/* The checksum: the arrays' hash folded with the header fields it depends on. */
static uint64_t file_hash(const csr_header *h, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                          const int32_t *ext_of_int) {
    uint64_t H = words_hash(ptr, (int64_t)h->n + 1, succ, h->m, w, h->m, ext_of_int, h->n);
    H = mix64(H ^ (uint64_t)(uint32_t)h->n);
    H = mix64(H ^ (uint64_t)(uint32_t)h->n0);
    H = mix64(H ^ (uint64_t)h->m);
    H = mix64(H ^ (uint64_t)(uint32_t)h->mg);
    return mix64(H ^ (uint64_t)(uint32_t)h->max_abs_w);
}

static int32_t max_abs(const int32_t *w, int64_t m) {
    int32_t mx = 0;
#pragma omp parallel for schedule(static) reduction(max : mx) if (m > (1 << 16))
    for (int64_t e = 0; e < m; e++) {
        const int32_t a = w[e] < 0 ? (w[e] == INT32_MIN ? INT32_MAX : -w[e]) : w[e];
        if (a > mx) mx = a;
    }
    return mx;
}

int egs_csr_save(const char *path, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
                 const int32_t *succ, const int32_t *w, const int32_t *ext_of_int, char *errbuf, int errlen) {
    if (!path || n < 1 || n0 < 0 || n0 > n || m < 0 || m > INT32_MAX || !ptr || !succ || !w || !ext_of_int) {
        cerr(errbuf, errlen, "egs_csr_save: bad arguments");
        return -1;
    }
    csr_header h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, EGS_CSR_MAGIC, 8);
    h.version = EGS_CSR_VERSION;
    h.header_bytes = sizeof(h);
    h.n = n;
    h.n0 = n0;
    h.m = m;
    h.mg = mg;
    h.max_abs_w = max_abs(w, m);
    h.checksum = file_hash(&h, ptr, succ, w, ext_of_int);
    FILE *f = fopen(path, "wb");
    if (!f) { cerr(errbuf, errlen, "cannot create %s", path); return -1; }
    int ok = fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(ptr, 4, (size_t)n + 1, f) == (size_t)n + 1 &&
             fwrite(succ, 4, (size_t)m, f) == (size_t)m && fwrite(w, 4, (size_t)m, f) == (size_t)m &&
             fwrite(ext_of_int, 4, (size_t)n, f) == (size_t)n;
    ok = (fclose(f) == 0) && ok;
    if (!ok) { cerr(errbuf, errlen, "write error on %s", path); return -1; }
    return 0;
}

void egs_csr_free(egs_csr *g) {
    if (!g) return;
    free(g->ptr);
    free(g->succ);
    free(g->w);
    free(g->ext_of_int);
    memset(g, 0, sizeof(*g));
}

int egs_csr_load(const char *path, egs_csr *out, char *errbuf, int errlen) {
    memset(out, 0, sizeof(*out));
    FILE *f = fopen(path, "rb");
    if (!f) { cerr(errbuf, errlen, "cannot open %s", path); return -1; }
    csr_header h;
    if (fread(&h, sizeof(h), 1, f) != 1 || memcmp(h.magic, EGS_CSR_MAGIC, 8) != 0) {
        fclose(f);
        cerr(errbuf, errlen, "%s is not a game cache (bad magic)", path);
        return -1;
    }
    if (h.version != EGS_CSR_VERSION || h.header_bytes != sizeof(h)) {
        fclose(f);
        cerr(errbuf, errlen, "%s: unsupported cache version %u", path, h.version);
        return -1;
    }
    if (h.n < 1 || h.n0 < 0 || h.n0 > h.n || h.m < 0 || h.m > INT32_MAX) {
        fclose(f);
        cerr(errbuf, errlen, "%s: bad sizes in the header", path);
        return -1;
    }
    const int64_t n = h.n, m = h.m;
    const int64_t want = (int64_t)sizeof(h) + 4 * ((n + 1) + 2 * m + n);
    if (fseek(f, 0, SEEK_END) != 0 || ftell(f) != want || fseek(f, (long)sizeof(h), SEEK_SET) != 0) {
        fclose(f);
        cerr(errbuf, errlen, "%s: file size does not match its header", path);
        return -1;
    }
    out->ptr = (int32_t *)malloc(4 * (size_t)(n + 1));
    out->succ = (int32_t *)malloc(4 * (size_t)(m > 0 ? m : 1));
    out->w = (int32_t *)malloc(4 * (size_t)(m > 0 ? m : 1));
    out->ext_of_int = (int32_t *)malloc(4 * (size_t)n);
    if (!out->ptr || !out->succ || !out->w || !out->ext_of_int) {
        fclose(f);
        egs_csr_free(out);
        cerr(errbuf, errlen, "out of memory loading %s", path);
        return -1;
    }
    int ok = fread(out->ptr, 4, (size_t)n + 1, f) == (size_t)n + 1 && fread(out->succ, 4, (size_t)m, f) == (size_t)m &&
             fread(out->w, 4, (size_t)m, f) == (size_t)m && fread(out->ext_of_int, 4, (size_t)n, f) == (size_t)n;
    fclose(f);
    if (!ok) { egs_csr_free(out); cerr(errbuf, errlen, "read error on %s", path); return -1; }
    if (file_hash(&h, out->ptr, out->succ, out->w, out->ext_of_int) != h.checksum) {
        egs_csr_free(out);
        cerr(errbuf, errlen, "%s: checksum mismatch (damaged file)", path);
        return -1;
    }
    int bad = out->ptr[0] != 0 || out->ptr[n] != m;
#pragma omp parallel for schedule(static) reduction(| : bad) if (n > (1 << 16))
    for (int64_t v = 0; v < n; v++) bad |= out->ptr[v + 1] <= out->ptr[v] || (uint32_t)out->ext_of_int[v] >= (uint32_t)n;
#pragma omp parallel for schedule(static) reduction(| : bad) if (m > (1 << 16))
    for (int64_t e = 0; e < m; e++) bad |= (uint32_t)out->succ[e] >= (uint32_t)n;
    if (bad) {
        egs_csr_free(out);
        cerr(errbuf, errlen, "%s: the stored graph is not a valid game", path);
        return -1;
    }
    int32_t mg = 0;
    if (egs_compute_mg(h.n, out->ptr, out->w, &mg) != 0 || mg != h.mg || max_abs(out->w, m) != h.max_abs_w) {
        egs_csr_free(out);
        cerr(errbuf, errlen, "%s: MG or max |w| in the header does not match the weights", path);
        return -1;
    }
    out->n = h.n;
    out->n0 = h.n0;
    out->m = h.m;
    out->mg = h.mg;
    out->max_abs_w = h.max_abs_w;
    return 0;
}
