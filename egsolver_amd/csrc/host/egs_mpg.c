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
#include <omp.h>
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
    const char *ts;     /* start of the current token */
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
        if (L->p >= L->end) { L->ts = L->end; L->tok = T_EOF; return; }
        char c = *L->p;
        if (c == ' ' || c == '\t') { L->p++; continue; }
        if (c == '\n') { L->line++; L->p++; continue; }
        if (c == '#' || c == '%' || (c == '/' && L->p + 1 < L->end && L->p[1] == '/')) {
            while (L->p < L->end && *L->p != '\n') L->p++;
            continue;
        }
        L->ts = L->p;
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
    lexer L = {text, text + len, 1, T_EOF, 0, NULL, 0, errbuf, errlen, 0, text};
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
    out->chunks = 1;
    return 0;
}

/* ---- parallel reader ------------------------------------------------------------------
 * The text after the header is cut at line starts into one chunk per thread, and each
 * chunk parses whole statements on its own.  A chunk k > 0 starts lexing at the first line
 * start s_k at or after its share of the text; chunk k-1 parses until the first statement
 * whose first token lies at or after s_k.  The cut is consistent when that token is also
 * the first token chunk k finds from s_k: s_k then lies between two statements (a
 * statement or string spanning the line start would make the two disagree).  Anything
 * else -- an inconsistent cut, any input the sequential reader would reject, a repeated
 * id across chunks, or random weights (their rand() draws follow read order) -- falls
 * back to the sequential reader, so results and messages are always its own. */

typedef struct {
    int64_t nv, mv, capv, cape;
    int32_t *ext, *deg, *succ, *w;
    uint8_t *own;
    char **names;
    int64_t first_tok, stop_tok;   /* offsets: first token from the chunk start, first statement at/after the limit */
    int maxabs, mind, maxd, low[4];
    int bad;
} pchunk;

static void pchunk_free(pchunk *c) {
    if (c->names) for (int64_t i = 0; i < c->nv; i++) free(c->names[i]);
    free(c->ext); free(c->deg); free(c->succ); free(c->w); free(c->own); free(c->names);
    memset(c, 0, sizeof(*c));
}

static int pchunk_grow_v(pchunk *c) {
    int64_t cap = c->capv ? 2 * c->capv : 1024;
    int32_t *a = (int32_t *)realloc(c->ext, sizeof(int32_t) * (size_t)cap);
    if (a) c->ext = a;
    int32_t *b = (int32_t *)realloc(c->deg, sizeof(int32_t) * (size_t)cap);
    if (b) c->deg = b;
    uint8_t *o = (uint8_t *)realloc(c->own, (size_t)cap);
    if (o) c->own = o;
    char **nm = (char **)realloc(c->names, sizeof(char *) * (size_t)cap);
    if (nm) c->names = nm;
    if (!a || !b || !o || !nm) return -1;
    c->capv = cap;
    return 0;
}

static int pchunk_grow_e(pchunk *c) {
    int64_t cap = c->cape ? 2 * c->cape : 8192;
    int32_t *a = (int32_t *)realloc(c->succ, sizeof(int32_t) * (size_t)cap);
    if (a) c->succ = a;
    int32_t *b = (int32_t *)realloc(c->w, sizeof(int32_t) * (size_t)cap);
    if (b) c->w = b;
    if (!a || !b) return -1;
    c->cape = cap;
    return 0;
}

/* Statements of one chunk: from offset beg until the first statement token at or after
 * lim.  Mirrors the statement grammar and checks of egs_mpg_read_mem; c->bad on anything
 * it would report. */
static void parse_chunk(const char *text, int64_t len, int64_t beg, int64_t lim, long long maxid, int weight_mode,
                        pchunk *c) {
    lexer L = {text + beg, text + len, 1, T_EOF, 0, NULL, 0, NULL, 0, 0, text + beg};
    c->mind = INT_MAX;
    next(&L);
    c->first_tok = L.ts - text;
    c->stop_tok = len;
    while (!L.failed && L.tok != T_EOF) {
        if (L.ts - text >= lim) { c->stop_tok = L.ts - text; return; }
        if (L.tok != T_INT) { c->bad = 1; return; }
        const long long ext = L.ival;
        next(&L);
        int owner;
        if (L.tok == T_MAX) owner = 0;
        else if (L.tok == T_MIN) owner = 1;
        else { c->bad = 1; return; }
        if (ext > maxid) { c->bad = 1; return; }
        if (c->nv == c->capv && pchunk_grow_v(c)) { c->bad = 1; return; }
        next(&L);
        const int64_t first = c->mv;
        for (;;) {
            if (L.tok != T_INT) { c->bad = 1; return; }
            const long long sc = L.ival;
            next(&L);
            if (L.tok != T_COL) { c->bad = 1; return; }
            next(&L);
            long long wt;
            if (L.tok == T_MINUS) {
                next(&L);
                if (L.tok != T_INT) { c->bad = 1; return; }
                wt = -L.ival;
            } else {
                if (L.tok != T_INT) { c->bad = 1; return; }
                wt = L.ival;
            }
            next(&L);
            if (sc > maxid) { c->bad = 1; return; }
            if (c->mv == c->cape && pchunk_grow_e(c)) { c->bad = 1; return; }
            const int wv = weight_mode == EGS_W_UNIT ? 1 : weight_mode == EGS_W_EXP ? 0 : (int)wt;
            c->succ[c->mv] = (int32_t)sc;
            c->w[c->mv] = wv;
            const int a = wv < 0 ? -wv : wv;
            if (a > c->maxabs) c->maxabs = a;
            c->mv++;
            if (L.tok == T_COMMA) { next(&L); continue; }
            break;
        }
        if (L.failed) { c->bad = 1; return; }
        char *nm = NULL;
        if (L.tok == T_STRING) {
            nm = (char *)malloc((size_t)L.slen + 1);
            if (!nm) { c->bad = 1; return; }
            memcpy(nm, L.sbeg, (size_t)L.slen);
            nm[L.slen] = '\0';
            next(&L);
        }
        if (L.tok != T_SEMI) { free(nm); c->bad = 1; return; }
        next(&L);
        const int d = (int)(c->mv - first);
        c->ext[c->nv] = (int32_t)ext;
        c->own[c->nv] = (uint8_t)owner;
        c->deg[c->nv] = d;
        c->names[c->nv] = nm;
        c->nv++;
        if (d > c->maxd) c->maxd = d;
        if (d < c->mind) c->mind = d;
        if (d < 4) c->low[d]++;
    }
    if (L.failed) c->bad = 1;
}

int egs_mpg_read_mem_par(const char *text, int64_t len, int weight_mode, int rnd_lo, int rnd_hi, egs_mpg *out,
                         char *errbuf, int errlen, int threads) {
    if (threads <= 0) threads = omp_get_max_threads();
    if (threads <= 1 || weight_mode == EGS_W_RND || len < (1 << 20))
        return egs_mpg_read_mem(text, len, weight_mode, rnd_lo, rnd_hi, out, errbuf, errlen);
    /* header, as the sequential reader parses it */
    lexer H = {text, text + len, 1, T_EOF, 0, NULL, 0, NULL, 0, 0, text};
    next(&H);
    if (H.tok == T_ARENA) next(&H);
    if (H.tok != T_INT || H.ival >= INT_MAX) return egs_mpg_read_mem(text, len, weight_mode, rnd_lo, rnd_hi, out, errbuf, errlen);
    const long long maxid = H.ival;
    next(&H);
    if (H.tok == T_INT) next(&H);
    if (H.tok != T_SEMI) return egs_mpg_read_mem(text, len, weight_mode, rnd_lo, rnd_hi, out, errbuf, errlen);
    const int64_t body = H.p - text;
    const int n = (int)maxid + 1;
    /* chunk starts: line starts at or after equal shares of the body */
    const int T = threads;
    int64_t *start = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T + 1));
    pchunk *ch = (pchunk *)calloc((size_t)T, sizeof(pchunk));
    if (!start || !ch) { free(start); free(ch); return egs_mpg_read_mem(text, len, weight_mode, rnd_lo, rnd_hi, out, errbuf, errlen); }
    start[0] = body;
    start[T] = len;
    for (int k = 1; k < T; k++) {
        int64_t x = body + (len - body) * k / T;
        if (x < start[k - 1]) x = start[k - 1];
        const char *nl = (const char *)memchr(text + x, '\n', (size_t)(len - x));
        start[k] = nl ? (int64_t)(nl - text) + 1 : len;
    }
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int k = 0; k < T; k++) parse_chunk(text, len, start[k], start[k + 1], maxid, weight_mode, &ch[k]);
    int ok = 1;
    int64_t nv = 0, mv = 0;
    for (int k = 0; k < T; k++) {
        if (ch[k].bad) ok = 0;
        if (k > 0 && ch[k].first_tok != ch[k - 1].stop_tok) ok = 0;
        nv += ch[k].nv;
        mv += ch[k].mv;
    }
    if (ok && (nv != n || mv > INT32_MAX)) ok = 0;
    int rc = -1;
    if (ok) {
        memset(out, 0, sizeof(*out));
        out->n = n;
        out->m = mv;
        out->ext_of_read = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
        out->read_of_ext = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
        out->owner = (uint8_t *)malloc((size_t)n);
        out->rptr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
        out->names = (char **)calloc((size_t)n, sizeof(char *));
        out->rsucc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(mv > 0 ? mv : 1));
        out->rw = (int32_t *)malloc(sizeof(int32_t) * (size_t)(mv > 0 ? mv : 1));
        if (!out->ext_of_read || !out->read_of_ext || !out->owner || !out->rptr || !out->names || !out->rsucc || !out->rw) {
            egs_mpg_free(out);
            ok = 0;
        }
    }
    if (ok) {
        int64_t *vb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T + 1)), *eb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T + 1));
        if (!vb || !eb) { free(vb); free(eb); egs_mpg_free(out); ok = 0; }
        else {
            vb[0] = eb[0] = 0;
            for (int k = 0; k < T; k++) { vb[k + 1] = vb[k] + ch[k].nv; eb[k + 1] = eb[k] + ch[k].mv; }
#pragma omp parallel for num_threads(T) schedule(static)
            for (int i = 0; i <= n; i++) out->read_of_ext[i] = -1;
            int dup = 0;
#pragma omp parallel for num_threads(T) schedule(static, 1) reduction(| : dup)
            for (int k = 0; k < T; k++) {
                const pchunk *c = &ch[k];
                int64_t e = eb[k];
                for (int64_t i = 0; i < c->nv; i++) {
                    const int64_t r = vb[k] + i;
                    if (!__sync_bool_compare_and_swap(&out->read_of_ext[c->ext[i]], -1, (int32_t)r)) dup = 1;
                    out->ext_of_read[r] = c->ext[i];
                    out->owner[r] = c->own[i];
                    out->names[r] = c->names[i];
                    out->rptr[r] = (int32_t)e;
                    e += c->deg[i];
                }
                if (c->mv) {
                    memcpy(out->rsucc + eb[k], c->succ, sizeof(int32_t) * (size_t)c->mv);
                    memcpy(out->rw + eb[k], c->w, sizeof(int32_t) * (size_t)c->mv);
                }
            }
            out->rptr[n] = (int32_t)mv;
            for (int k = 0; k < T; k++) ch[k].nv = 0;     /* names now belong to out */
            int maxabs = 0, mind = INT_MAX, maxd = 0;
            for (int k = 0; k < T; k++) {
                if (ch[k].maxabs > maxabs) maxabs = ch[k].maxabs;
                if (ch[k].mind < mind) mind = ch[k].mind;
                if (ch[k].maxd > maxd) maxd = ch[k].maxd;
                for (int d = 0; d < 4; d++) out->low_outdeg[d] += ch[k].low[d];
            }
            out->max_abs_w = maxabs;
            out->min_outdeg = mind;
            out->max_outdeg = maxd;
            out->chunks = T;
            free(vb);
            free(eb);
            if (dup) { egs_mpg_free(out); ok = 0; }
            else rc = 0;
        }
    }
    for (int k = 0; k < T; k++) pchunk_free(&ch[k]);
    free(ch);
    free(start);
    if (!ok) return egs_mpg_read_mem(text, len, weight_mode, rnd_lo, rnd_hi, out, errbuf, errlen);
    return rc;
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
    int rc = egs_mpg_read_mem_par(buf, (int64_t)len, weight_mode, rnd_lo, rnd_hi, out, errbuf, errlen, 0);
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
        e += g->rptr[mapping[i] + 1] - g->rptr[mapping[i]];
    }
    ptr[n] = (int32_t)e;
    /* edge copy: each internal vertex's list is independent */
#pragma omp parallel for schedule(dynamic, 4096) if (e > (1 << 16))
    for (int i = 0; i < n; i++) {
        const int r = mapping[i];
        int64_t at = ptr[i];
        for (int j = g->rptr[r]; j < g->rptr[r + 1]; j++, at++) {
            succ[at] = rev[g->read_of_ext[g->rsucc[j]]];
            w[at] = g->rw[j];
        }
    }
    free(rev);
    return n0;
}
