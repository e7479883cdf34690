/*
 * egs_host.h -- host-side C library of the MI355X energy-game solver (libegs_host.so).
 *
 * Input side of the solver path: the .mpg reader, the owner partition into the internal
 * numbering, the MG threshold, and this project's seeded game generator.  Plain C, no
 * GPU calls.  The reader replaces the reference's flex/bison front end
 * (src/the_lex.l:33-52, src/the_parser.y:63-102,128-295) and remap/postparsing
 * (src/egsolver.c:107-315).
 */
#ifndef EGS_HOST_H
#define EGS_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- parsed game in read order ------------------------------------------------------ */

/* Weight rewriting modes applied while reading (src/the_parser.y:324-327,
 * options --unit-weights / --rnd-weights / --exp-weights / --no-weights). */
enum { EGS_W_KEEP = 0, EGS_W_UNIT = 1, EGS_W_EXP = 2, EGS_W_RND = 3 };

typedef struct egs_mpg {
    int32_t n;            /* vertices: max id + 1 */
    int64_t m;            /* edges */
    int32_t *ext_of_read; /* external id of the i-th vertex read */
    int32_t *read_of_ext; /* inverse, -1 where unused */
    uint8_t *owner;       /* owner of the i-th vertex read: 0 = MAX token, 1 = MIN token */
    int32_t *rptr;        /* n+1 offsets into rsucc/rw (read order) */
    int32_t *rsucc;       /* successor EXTERNAL ids */
    int32_t *rw;          /* weights after the rewrite mode */
    int32_t max_abs_w;    /* max |w| (src/the_parser.y:330) */
    int32_t min_outdeg, max_outdeg;
    int32_t low_outdeg[4];
    char **names;         /* optional "string" per vertex read, or NULL entries */
    int32_t chunks;       /* chunks the text was parsed in (1: the sequential reader) */
} egs_mpg;

/* Parse a .mpg file (plain or gzip, path NULL = stdin).  Weight mode and the
 * rnd interval follow the reference; EGS_W_RND draws rand() per edge in read order
 * after the caller's srand().  Returns 0, or -1 with a message in errbuf (the reference
 * prints the same facts and exits). */
int egs_mpg_read(const char *path, int weight_mode, int rnd_lo, int rnd_hi, egs_mpg *out,
                 char *errbuf, int errlen);
int egs_mpg_read_mem(const char *text, int64_t len, int weight_mode, int rnd_lo, int rnd_hi,
                     egs_mpg *out, char *errbuf, int errlen);
/* The same result as egs_mpg_read_mem (values, order, names, messages), parsed by up to
 * `threads` threads (0 = OpenMP's default) over line-aligned chunks of the text.  Inputs
 * under 1 MB, random weights and anything the chunks cannot parse on their own (a cut
 * inside a statement, an error) go through egs_mpg_read_mem.  egs_mpg_read uses it. */
int egs_mpg_read_mem_par(const char *text, int64_t len, int weight_mode, int rnd_lo, int rnd_hi,
                         egs_mpg *out, char *errbuf, int errlen, int threads);
void egs_mpg_free(egs_mpg *g);

/* Internal numbering of a parsed game: owner-0 vertices first, both blocks in read order
 * (src/egsolver.c:146-214 with the default comparator, src/sorting_criteria.h:91-102),
 * then the internal CSR (src/egsolver.c:253-281).  mapping[i] = read index of internal
 * vertex i.  Returns n0 (the reference's counter_nodi0). */
int egs_mpg_internalise(const egs_mpg *g, int32_t *mapping, int32_t *ptr, int32_t *succ,
                        int32_t *w);

/* Same partition for arrays whose read order equals external ids (generator output). */
int egs_internalise(int n, const uint8_t *owner, const int32_t *rptr, const int32_t *rsucc,
                    const int32_t *rw, int32_t *ptr, int32_t *succ, int32_t *w,
                    int32_t *ext_of_int);

/* MG = sum_v max(0, max_e -w_e) (src/egsolver.c:283-304).  -1 on int32 overflow. */
int egs_compute_mg(int n, const int32_t *ptr, const int32_t *w, int32_t *mg_out);

/* Loop bound of src/egsolver.c:64-90: maxloop < 0 means n*m*MG+1, saturating at
 * LONG_MAX-1. */
long egs_max_loop(long m, long n, long mg, long maxloop);

/* ---- generator ---------------------------------------------------------------------- */

enum { EGS_DEG_FIXED = 0, EGS_DEG_UNIFORM = 1, EGS_DEG_POWERLAW = 2, EGS_DEG_COMPLETE = 3 };

typedef struct egs_gen_params {
    uint64_t seed;
    int32_t n;
    int32_t deg_kind;     /* EGS_DEG_* */
    int32_t deg_lo, deg_hi;
    double alpha;         /* power-law exponent */
    int32_t wmax;         /* weights uniform in [-wmax, wmax] */
    int32_t reserved;
} egs_gen_params;

typedef struct egs_gen_read_order {
    uint8_t *owner;       /* n */
    int32_t *ptr;         /* n+1 */
    int32_t *succ;        /* m */
    int32_t *w;           /* m */
} egs_gen_read_order;

/* Generate game `stream` of the family: arrays in external-id order.  With count_only
 * the arrays are untouched and only the edge count is returned.  Returns m. */
int64_t egs_gen_game(const egs_gen_params *p, uint64_t stream, egs_gen_read_order *out,
                     int count_only);
int64_t egs_gen_count_edges(const egs_gen_params *p, uint64_t stream);

/* Instance parameters of the multi-game families: index i of family `family`
 * (EGS_FAMILY_BATCH: n log-uniform in [1e3, 5e4]; EGS_FAMILY_SERVE: [1e3, 5e5]);
 * out-degree uniform in [2,16]; W log-uniform in [10,1000]; every 16th instance
 * (i % 16 == 15) a complete digraph on 200..500 vertices. */
enum { EGS_FAMILY_BATCH = 4, EGS_FAMILY_SERVE = 5 };
void egs_gen_family_params(int family, uint64_t seed, int index, egs_gen_params *out);

/* Raw PRNG output, for the generator's own known-answer test. */
uint64_t egs_rng_test_stream(uint64_t seed, uint64_t stream, int count, uint64_t *out);

/* ---- binary game cache (no reference counterpart) ------------------------------------ */

/* A game in its internal numbering, as egs_csr_load returns it (arrays owned by it). */
typedef struct egs_csr {
    int32_t n, n0;
    int64_t m;
    int32_t mg, max_abs_w;
    int32_t *ptr;         /* n+1 */
    int32_t *succ;        /* m */
    int32_t *w;           /* m */
    int32_t *ext_of_int;  /* n: external id of each internal vertex */
} egs_csr;

/* Save / load a game in the internal numbering with a checksum; a load refuses a file
 * whose header, size, checksum or graph does not hold up.  0 or -1 with a message. */
int egs_csr_save(const char *path, int32_t n, int32_t n0, int64_t m, int32_t mg, const int32_t *ptr,
                 const int32_t *succ, const int32_t *w, const int32_t *ext_of_int, char *errbuf, int errlen);
int egs_csr_load(const char *path, egs_csr *out, char *errbuf, int errlen);
void egs_csr_free(egs_csr *g);

/* ---- result verification (no reference counterpart) ---------------------------------- */

/* Vertices with lift(values)(v) != values(v), for the lifting of src/cpu_solver.c:300-305
 * (owner 0 the min, owner 1 the max over out-edges of OMINUS, src/common.h:270-272), plus
 * vertices without an out-edge.  *first_bad (may be NULL) gets the first such vertex or -1.
 * Zero means values is a fixpoint, hence at or above the least fixpoint. */
int64_t egs_check_fixpoint(int32_t n, int32_t n0, int32_t mg, const int32_t *ptr, const int32_t *succ,
                           const int32_t *w, const int32_t *values, int64_t *first_bad);

/* Checks a certificate (egs_solver_top_certificate in egs.h) that T = {v : values[v] ==
 * TOP} is won by player 1: h[v] >= 0 exactly on T; with w' = (n+1) w + 1, every owner-0
 * vertex of T has all successors in T and h(v) >= w'(v,u) + h(u) on each, and every owner-1
 * vertex of T has a successor u in T with h(v) >= w'(v,u) + h(u) (exact 128-bit arithmetic).
 * Returns EGS_CERT_OK or the first failure's code, with its vertex in *bad_vertex (-1 when
 * none; may be NULL). */
enum {
    EGS_CERT_OK = 0,
    EGS_CERT_MEMBERSHIP = 1, /* h >= 0 does not match value == TOP */
    EGS_CERT_LEAVES = 2,     /* an owner-0 vertex of T has a successor outside T */
    EGS_CERT_POTENTIAL = 3,  /* an owner-0 edge inside T breaks the inequality */
    EGS_CERT_NO_EDGE = 4     /* an owner-1 vertex of T has no edge into T that meets it */
};
int egs_check_top_certificate(int32_t n, int32_t n0, const int32_t *ptr, const int32_t *succ, const int32_t *w,
                              const int32_t *values, const int64_t *h, int64_t *bad_vertex);

/* Write a game in .mpg format (read order, external ids 0..n-1); gz != 0 for gzip. */
int egs_write_mpg(const char *path, int n, const uint8_t *owner, const int32_t *rptr,
                  const int32_t *rsucc, const int32_t *rw, int gz);

#ifdef __cplusplus
}
#endif
#endif /* EGS_HOST_H */
