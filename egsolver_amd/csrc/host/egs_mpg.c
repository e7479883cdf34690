/*
 * egs_mpg.c -- hand-written reader for the .mpg arena format.
 *
 * Accepts the language of the reference's flex/bison front end, which cannot be
 * regenerated here (no flex or bison in the toolchain):
 *   tokens   src/the_lex.l:33-52  blanks, newlines, "arena", "MAX", "MIN", ';', ':', ',',
 *            '-', comments from '#', '%' or "//" to end of line, FLOAT, INT, quoted
 *            STRING; any other character is skipped
 *   grammar  src/the_parser.y:121-160
 *            header : [arena] INT [INT] ';'
 *            line   : INT (MAX|MIN) edge (',' edge)* [STRING] ';'
 *            edge   : INT ':' ['-'] INT
 *   actions  src/the_parser.y:186-337: MAX -> owner 0, MIN -> owner 1; every id 0..max
 *            exactly once; successors <= max; weight rewrite modes; max |w|.
 * Input may be gzip-compressed (zlib reads plain text transparently, like the
 * reference's gzread front end, src/egsolver.c:47-59).
 */
#include <ctype.h>
#include <limits.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <zlib.h>

#include "egs_host.h"

enum { T_EOF, T_ARENA, T_MAX, T_MIN, T_SEMI, T_COL, T_COMMA, T_MINUS, T_INT, T_FLOAT, T_STRING, T_BAD };

typedef struct {
    const char *p, *end;
    int line;
    int tok;
    long long ival;
    const char *sbeg;   /* STRING token body (without quotes) */
    int slen;
    char *err;
    int errlen;
    int failed;
} lexer;

static void lerr(lexer *L, const char *fmt, ...) {
    if (L->failed) return;
    L->failed = 1;
    va_list ap;
    va_start(ap, fmt);
    if (L->err && L->errlen > 0) vsnprintf(L->err, (size_t)L->errlen, fmt, ap);
    va_end(ap);
}

static int starts(const lexer *L, const char *kw) {
    size_t k = strlen(kw);
    return (size_t)(L->end - L->p) >= k && memcmp(L->p, kw, k) == 0;
}

static void next(lexer *L) {
    for (;;) {
        if (L->p >= L->end) { L->tok = T_EOF; return; }
        char c = *L->p;
        if (c == ' ' || c == '\t') { L->p++; continue; }
        if (c == '\n') { L->line++; L->p++; continue; }
        if (c == '#' || c == '%' || (c == '/' && L->p + 1 < L->end && L->p[1] == '/')) {
            while (L->p < L->end && *L->p != '\n') L->p++;
            continue;
        }
        if (isdigit((unsigned char)c)) {
            const char *q = L->p;
            while (q < L->end && isdigit((unsigned char)*q)) q++;
            if (q + 1 < L->end && *q == '.' && isdigit((unsigned char)q[1])) {
                q++;
                while (q < L->end && isdigit((unsigned char)*q)) q++;
                L->p = q;
                L->tok = T_FLOAT;
                return;
            }
            long long v = 0;
            for (const char *r = L->p; r < q; r++) {
                v = v * 10 + (*r - '0');
                if (v > INT_MAX) { lerr(L, "CONVERSION ERROR out-of-range value in input (line %d)", L->line); L->tok = T_BAD; return; }
            }
            L->ival = v;
            L->p = q;
            L->tok = T_INT;
            return;
        }
        if (c == '"') {
            const char *q = L->p + 1;
            while (q < L->end && *q != '"') {
                if (*q == '\\' && q + 1 < L->end) q += 2;
                else { if (*q == '\n') { /* counted when re-scanned only if unmatched */ } q++; }
            }
            if (q < L->end) {
                for (const char *r = L->p; r < q; r++) if (*r == '\n') L->line++;
                L->sbeg = L->p + 1;
                L->slen = (int)(q - L->p - 1);
                L->p = q + 1;
                L->tok = T_STRING;
                return;
            }
            L->p++;            /* unmatched quote: skipped like any other character */
            continue;
        }
        if (starts(L, "arena")) { L->p += 5; L->tok = T_ARENA; return; }
        if (starts(L, "MAX")) { L->p += 3; L->tok = T_MAX; return; }
        if (starts(L, "MIN")) { L->p += 3; L->tok = T_MIN; return; }
        L->p++;
        switch (c) {
        case ';': L->tok = T_SEMI; return;
        case ':': L->tok = T_COL; return;
        case ',': L->tok = T_COMMA; return;
        case '-': L->tok = T_MINUS; return;
        default: continue;   /* catch-all rule: skip */
        }
    }
}

static const char *tokname(int t) {
    static const char *names[] = {"end of input", "arena", "MAX", "MIN", "';'", "':'", "','", "'-'",
                                  "integer", "float", "string", "bad token"};
    return names[t];
}

static int expect(lexer *L, int t) {
    if (L->failed) return 0;
    if (L->tok != t) {
        lerr(L, "Error reading input:  syntax error line %d (expected %s, found %s)", L->line, tokname(t), tokname(L->tok));
        return 0;
    }
    return 1;
}

void egs_mpg_free(egs_mpg *g) {
    if (!g) return;
    if (g->names) { for (int i = 0; i < g->n; i++) free(g->names[i]); }
    free(g->names); free(g->ext_of_read); free(g->read_of_ext); free(g->owner);
    free(g->rptr); free(g->rsucc); free(g->rw);
    memset(g, 0, sizeof(*g));
}

int egs_mpg_read_mem(const char *text, int64_t len, int weight_mode, int rnd_lo, int rnd_hi,
                     egs_mpg *out, char *errbuf, int errlen) {
    memset(out, 0, sizeof(*out));
    lexer L = {text, text + len, 1, T_EOF, 0, NULL, 0, errbuf, errlen, 0};
    next(&L);
    if (L.tok == T_ARENA) next(&L);
    if (!expect(&L, T_INT)) return -1;
    long long maxid = L.ival;
    next(&L);
    if (L.tok == T_INT) next(&L);
    if (!expect(&L, T_SEMI)) return -1;
    next(&L);
    const int n = (int)maxid + 1;
    if (maxid >= INT_MAX) { lerr(&L, "Error in input. max node id too large"); return -1; }
    out->n = n;
    out->ext_of_read = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    out->read_of_ext = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    out->owner = (uint8_t *)malloc((size_t)n);
    out->rptr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    out->names = (char **)calloc((size_t)n, sizeof(char *));
    size_t cap = 2 * (size_t)(n + 1), m = 0;
    out->rsucc = (int32_t *)malloc(sizeof(int32_t) * cap);
    out->rw = (int32_t *)malloc(sizeof(int32_t) * cap);
    if (!out->ext_of_read || !out->read_of_ext || !out->owner || !out->rptr || !out->names || !out->rsucc || !out->rw) {
        lerr(&L, "ERROR in re/malloc (reader)");
        egs_mpg_free(out);
        return -1;
    }
    for (int i = 0; i <= n; i++) out->read_of_ext[i] = -1;
    int count = 0, maxabs = 0, mind = INT_MAX, maxd = 0;
    out->rptr[0] = 0;
    while (!L.failed && L.tok != T_EOF) {
        if (!expect(&L, T_INT)) break;
        long long ext = L.ival;
        next(&L);
        int owner;
        if (L.tok == T_MAX) owner = 0;
        else if (L.tok == T_MIN) owner = 1;
        else { lerr(&L, "Error reading input:  syntax error line %d (expected MAX or MIN)", L.line); break; }
        if (ext > maxid) { lerr(&L, "Error in input. Unexpected node identifier: %lld greater that %lld (input line %d)", ext, maxid, L.line); break; }
        if (out->read_of_ext[ext] != -1) { lerr(&L, "Error in input. Repeated node identifier:  %lld (input line %d)", ext, L.line); break; }
        if (count >= n) { lerr(&L, "Error in input. more vertex lines than %d", n); break; }
        out->read_of_ext[ext] = count;
        out->ext_of_read[count] = (int32_t)ext;
        out->owner[count] = (uint8_t)owner;
        next(&L);
        size_t first = m;
        for (;;) {
            if (!expect(&L, T_INT)) break;
            long long s = L.ival;
            next(&L);
            if (!expect(&L, T_COL)) break;
            next(&L);
            long long w;
            if (L.tok == T_MINUS) {
                next(&L);
                if (!expect(&L, T_INT)) break;
                w = -L.ival;
            } else {
                if (!expect(&L, T_INT)) break;
                w = L.ival;
            }
            next(&L);
            if (s > maxid) { lerr(&L, "Error in input. Unexpected successor identifier: %lld greater that %lld (input line %d)", s, maxid, L.line); break; }
            if (m + 1 >= cap) {
                cap *= 2;
                int32_t *a = (int32_t *)realloc(out->rsucc, sizeof(int32_t) * cap);
                int32_t *b = (int32_t *)realloc(out->rw, sizeof(int32_t) * cap);
                if (!a || !b) { lerr(&L, "ERROR in re/malloc (reader edges)"); if (a) out->rsucc = a; if (b) out->rw = b; break; }
                out->rsucc = a; out->rw = b;
            }
            int wv;
            switch (weight_mode) {
            case EGS_W_UNIT: wv = 1; break;
            case EGS_W_RND: wv = rnd_lo + (rand() % (1 + rnd_hi - rnd_lo)); break;
            case EGS_W_EXP: wv = 0; break;
            default: wv = (int)w;
            }
            out->rsucc[m] = (int32_t)s;
            out->rw[m] = wv;
            int a = wv < 0 ? -wv : wv;
            if (a > maxabs) maxabs = a;
            m++;
            if (L.tok == T_COMMA) { next(&L); continue; }
            break;
        }
        if (L.failed) break;
        if (L.tok == T_STRING) {
            char *nm = (char *)malloc((size_t)L.slen + 1);
            memcpy(nm, L.sbeg, (size_t)L.slen);
            nm[L.slen] = '\0';
            out->names[count] = nm;
            next(&L);
        }
        if (!expect(&L, T_SEMI)) break;
        next(&L);
        int d = (int)(m - first);
        if (d > maxd) maxd = d;
        if (d < mind) mind = d;
        if (d < 4) out->low_outdeg[d]++;
        count++;
        out->rptr[count] = (int32_t)m;
    }
    if (!L.failed && count != n) lerr(&L, "Error in input. Expecting %d nodes, read %d nodes (input line %d)", n, count, L.line);
    if (L.failed) { egs_mpg_free(out); return -1; }
    out->m = (int64_t)m;
    out->max_abs_w = maxabs;
    out->min_outdeg = mind;
    out->max_outdeg = maxd;
    return 0;
}

int egs_mpg_read(const char *path, int weight_mode, int rnd_lo, int rnd_hi, egs_mpg *out,
                 char *errbuf, int errlen) {
    gzFile f = path ? gzopen(path, "r") : gzdopen(dup(fileno(stdin)), "rb");
    if (!f) {
        if (errbuf) snprintf(errbuf, (size_t)errlen, "ERROR in opening file %s", path ? path : "stdin");
        return -1;
    }
    size_t cap = 1 << 20, len = 0;
    char *buf = (char *)malloc(cap);
    for (;;) {
        if (cap - len < (1 << 16)) {
            cap *= 2;
            char *nb = (char *)realloc(buf, cap);
            if (!nb) { free(buf); gzclose(f); if (errbuf) snprintf(errbuf, (size_t)errlen, "ERROR in re/malloc (input)"); return -1; }
            buf = nb;
        }
        int got = gzread(f, buf + len, (unsigned)(cap - len));
        if (got < 0) { free(buf); gzclose(f); if (errbuf) snprintf(errbuf, (size_t)errlen, "Error reading input (in gzread)"); return -1; }
        if (got == 0) break;
        len += (size_t)got;
    }
    gzclose(f);
    int rc = egs_mpg_read_mem(buf, (int64_t)len, weight_mode, rnd_lo, rnd_hi, out, errbuf, errlen);
    free(buf);
    return rc;
}

int egs_mpg_internalise(const egs_mpg *g, int32_t *mapping, int32_t *ptr, int32_t *succ, int32_t *w) {
    const int n = g->n;
    int32_t *rev = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int k = 0;
    for (int i = 0; i < n; i++) if (g->owner[i] == 0) { mapping[k] = i; rev[i] = k; k++; }
    const int n0 = k;
    for (int i = 0; i < n; i++) if (g->owner[i] != 0) { mapping[k] = i; rev[i] = k; k++; }
    int64_t e = 0;
    for (int i = 0; i < n; i++) {
        ptr[i] = (int32_t)e;
        const int r = mapping[i];
        for (int j = g->rptr[r]; j < g->rptr[r + 1]; j++) {
            succ[e] = rev[g->read_of_ext[g->rsucc[j]]];
            w[e] = g->rw[j];
            e++;
        }
    }
    ptr[n] = (int32_t)e;
    free(rev);
    return n0;
}
