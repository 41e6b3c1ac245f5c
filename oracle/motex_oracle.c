/*
 * motex_oracle.c -- CPU restatement of MoTeX's all-pairs motif-extraction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke entry and
 * bench.py's cpu_baseline leg load it to compute expected results (or, for the baseline,
 * to time the reference algorithm on host cores).  The product library (motex_amd) never
 * links, loads or calls it, and must fail loudly rather than fall back to it.
 *
 * What is restated here, with the reference file:line each part follows
 * (paths relative to the MoTeX repository root):
 *
 *   orc_hd_pair       per-pair Hamming operator: motifs_extraction_hd, functions.cc:526-559.
 *                     The FLASM step it calls (libflasm::flasm_hd, libFLASM.zip
 *                     libFLASM/include/libflasm.h:74) is binary-only; it is restated as the
 *                     MaxShift bit-vector recurrence of functions.cc:289-302 with reports
 *                     restricted to full text windows (pos_t >= k-1, pos_x >= k-1).
 *   orc_ed_pair       per-pair edit-distance operator: motifs_extraction_ed,
 *                     functions.cc:734-768, over the MaxShift edit recurrence of
 *                     functions.cc:437-447 and bitminmax of functions.cc:923-967
 *                     (libflasm::flasm_ed, libflasm.h:71, is binary-only).
 *   orc_struct_pair   structured_motifs_extraction_{hd,ed}, functions.cc:564-705 / 773-915.
 *   orc_all_pairs     the all-pairs driver of motex.cc:494-583 (OpenMP over x as in
 *                     motex.cc:509-511).
 *   orc_recurse       cartesian product of gap lengths, functions.cc:710-729 and
 *                     motex.cc:439-486.
 *   orc_write_*       writers: functions.cc:1010-1142 (header, motifs), 1147-1260
 *                     (structured), 1265-1461 (SMILE), 1466-1608 (structured SMILE).
 *
 * Pinning status (what checks this restatement; see tests/test_oracle.py):
 *   - orc_hd_pair: pinned.  Brute-force full-window Hamming on random small inputs
 *     (a different algorithm), the black-box probe facts recorded for libflasm
 *     (TTTTTACGTC / CGTCGGGGGG, k=5, e=1: ACGTC u=1 v=1, TTTTT v=2, GGGGG v=3), e=0
 *     equals exact k-mer counting, and the transpose symmetry H(x,y) = H(y,x)^T.
 *   - orc_ed_pair: PARTIALLY PINNED.  At e=0 every hit is an exact occurrence (the
 *     greedy bitminmax choice can also miss some, which is a property of the recurrence,
 *     not of this restatement); the recorded probe facts (an interior exact hit yields
 *     three end positions at e=1; a length-(k-1) text prefix at the text start is an
 *     occurrence) hold; bitminmax's
 *     case analysis equals "largest word among those of least popcount" exhaustively
 *     on small words.  Whether libflasm's binary equals the MaxShift recurrence for
 *     every e >= 1 input is parity unpinned: the binary may not be run here and the
 *     reference ships no golden outputs.  Row 0 is initialised to min(j, k) ones, the
 *     evident intent of functions.cc:438 (whose unparenthesised min() macro would give
 *     j+1 ones for j < k).
 *   - orc_struct_pair: pinned against a direct restatement of the merge definition on
 *     brute-force per-box hit sets (tests), for both distances.
 *   - writers: pinned by format facts (IEEE-double quorum filter edge cases,
 *     first-occurrence dedup, SMILE header padding widths) in tests.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_DNA  "ACGTN"
#define ORC_PROT "ARNDCQEGHILKMFPSTWYVOUBZJX*"
#define ORC_USR  "0123456789"

typedef uint64_t word_t;

static inline unsigned pc64(word_t w) { return (unsigned)__builtin_popcountll(w); }

/* ones(r) for 1 <= r <= 64 */
static inline word_t ones(unsigned r) { return r >= 64 ? ~(word_t)0 : (((word_t)1 << r) - 1); }

/* bitminmax, functions.cc:923-967: among a, b, c return one with the fewest set bits;
 * ties are broken towards the numerically largest.  Restated case by case. */
word_t orc_bitminmax(word_t a, word_t b, word_t c)
{
    unsigned x = pc64(a), y = pc64(b), z = pc64(c);
    unsigned lo = x < y ? x : y;
    if (z < lo) return c;
    if (z == lo && x != y) {
        word_t other = x < y ? a : b;
        return c > other ? c : other;
    }
    if (z == x && x == y) {
        word_t mbc = b > c ? b : c;
        return a > mbc ? a : mbc;
    }
    if (z > lo && x != y) return x < y ? a : b;
    return a > b ? a : b;
}

/* ------------------------------------------------------------------------------------ */
/* Per-pair hit enumeration.  A "hit" is the tuple (pos_t, pos_x, err) FLASM returns.    */
/* ------------------------------------------------------------------------------------ */

typedef void (*hit_fn)(void *ctx, uint32_t pos_t, uint32_t pos_x, uint32_t err);

/* MaxShift Hamming, functions.cc:289-302: D[i][j] = ((D[i-1][j-1] & y) << 1) | delta.
 * Reports are restricted to complete windows on both sides (libflasm semantics). */
/* Factors longer than one 64-bit word (libflasm's multi-word "Limit", libflasm.h:44-54):
 * the same full-window Hamming hits, from a sliding mismatch count along each diagonal
 * d = pos_t - pos_x. */
static int hd_hits_long(const unsigned char *x, uint32_t m, const unsigned char *t, uint32_t n,
                        uint32_t ell, uint32_t err, hit_fn fn, void *ctx)
{
    for (int64_t d = -(int64_t)m + 1; d < (int64_t)n; ++d) {
        uint32_t cnt = 0;
        int64_t j = d < 0 ? -d : 0; /* pos_x of the diagonal's first cell */
        for (int64_t i = j + d; i < (int64_t)n && j < (int64_t)m; ++i, ++j) {
            cnt += t[i] != x[j];
            if (i >= (int64_t)ell && j >= (int64_t)ell) cnt -= t[i - ell] != x[j - ell];
            if (j + 1 >= (int64_t)ell && i + 1 >= (int64_t)ell && cnt <= err)
                fn(ctx, (uint32_t)i, (uint32_t)j, cnt);
        }
    }
    return 0;
}

static int hd_hits(const unsigned char *x, uint32_t m, const unsigned char *t, uint32_t n,
                   uint32_t ell, uint32_t err, hit_fn fn, void *ctx)
{
    if (ell == 0) return -1;
    if (ell > 64) return hd_hits_long(x, m, t, n, ell, err, fn, ctx);
    word_t ymask = ones(ell - 1);
    word_t *D = (word_t *)malloc((size_t)(m + 1) * sizeof(word_t));
    if (!D) return -2;
    D[0] = 0;
    for (uint32_t j = 1; j <= m; ++j) D[j] = ones(j < ell ? j : ell);
    for (uint32_t i = 1; i <= n; ++i) {
        unsigned char tc = t[i - 1];
        /* descending j so D[j-1] still holds row i-1 */
        for (uint32_t j = m; j >= 1; --j) {
            word_t d = ((D[j - 1] & ymask) << 1) | (word_t)(tc != x[j - 1]);
            D[j] = d;
            if (j >= ell && i >= ell) {
                unsigned p = pc64(d);
                if (p <= err) fn(ctx, i - 1, j - 1, p);
            }
        }
        D[0] = 0;
    }
    free(D);
    return 0;
}

/* Edit-distance words: 128 bits.  A selected word is narrower than 2k+1 bits, so 128 bits
 * carry factors of up to k = 64 exactly.  For k <= 32 every word fits the reference's 64-bit
 * WORD1 and the results are those of its 64-bit arithmetic; for 32 < k <= 64 the 64-bit
 * restatement would drop the high bits of the unmasked "up" candidate, and the full word is
 * kept, as libflasm's multi-word bit arrays (libflasm.h:44-54) keep a factor's errors. */
typedef unsigned __int128 wide_t;

static inline unsigned pcw(wide_t w) { return pc64((word_t)w) + pc64((word_t)(w >> 64)); }
static inline wide_t onesw(unsigned r) { return r >= 128 ? ~(wide_t)0 : (((wide_t)1 << r) - 1); }

/* bitminmax over 128-bit words: the same case analysis as orc_bitminmax. */
static wide_t bitminmax_w(wide_t a, wide_t b, wide_t c)
{
    unsigned x = pcw(a), y = pcw(b), z = pcw(c);
    unsigned lo = x < y ? x : y;
    if (z < lo) return c;
    if (z == lo && x != y) {
        wide_t other = x < y ? a : b;
        return c > other ? c : other;
    }
    if (z == x && x == y) {
        wide_t mbc = b > c ? b : c;
        return a > mbc ? a : mbc;
    }
    if (z > lo && x != y) return x < y ? a : b;
    return a > b ? a : b;
}

/* for tests: bitminmax_w on words given as (lo, hi) halves, result's halves in out[2] */
void orc_bitminmax_w(uint64_t alo, uint64_t ahi, uint64_t blo, uint64_t bhi, uint64_t clo, uint64_t chi,
                     uint64_t *out)
{
    wide_t r = bitminmax_w(((wide_t)ahi << 64) | alo, ((wide_t)bhi << 64) | blo, ((wide_t)chi << 64) | clo);
    out[0] = (uint64_t)r;
    out[1] = (uint64_t)(r >> 64);
}

/* MaxShift edit distance, functions.cc:437-447 with bitminmax (functions.cc:923-967).
 * up = D[i-1][j], left = D[i][j-1], diag = D[i-1][j-1]. */
static int ed_hits(const unsigned char *x, uint32_t m, const unsigned char *t, uint32_t n,
                   uint32_t ell, uint32_t err, hit_fn fn, void *ctx)
{
    if (ell == 0 || ell > 64) return -1;
    wide_t ymask = onesw(ell - 1);
    wide_t *prev = (wide_t *)malloc((size_t)(m + 1) * sizeof(wide_t));
    wide_t *cur = (wide_t *)malloc((size_t)(m + 1) * sizeof(wide_t));
    if (!prev || !cur) { free(prev); free(cur); return -2; }
    prev[0] = 0;
    for (uint32_t j = 1; j <= m; ++j) prev[j] = onesw(j < ell ? j : ell);
    for (uint32_t i = 1; i <= n; ++i) {
        unsigned char tc = t[i - 1];
        cur[0] = 0;
        for (uint32_t j = 1; j <= m; ++j) {
            wide_t dl = (wide_t)(tc != x[j - 1]);
            wide_t d;
            if (j <= ell)
                d = bitminmax_w((prev[j] << 1) | 1, (cur[j - 1] << 1) | 1, (prev[j - 1] << 1) | dl);
            else
                d = bitminmax_w(((cur[j - 1] & ymask) << 1) | 1, (prev[j] << 1) | 1,
                                ((prev[j - 1] & ymask) << 1) | dl);
            if (ell < 64 && (d >> (2 * ell))) abort();  /* wider than 2k bits: the width bound failed */
            cur[j] = d;
            if (j >= ell) {
                unsigned p = pcw(d);
                if (p <= err) fn(ctx, i - 1, j - 1, p);
            }
        }
        wide_t *tmp = prev; prev = cur; cur = tmp;
    }
    free(prev); free(cur);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Per-pair operators: fold hits into caller-owned u, v (functions.cc:542-555).           */
/* ------------------------------------------------------------------------------------ */

typedef struct { uint32_t *u, *v; uint8_t *occ; } fold_ctx;

static void fold_hit(void *c, uint32_t pos_t, uint32_t pos_x, uint32_t err)
{
    (void)pos_t; (void)err;
    fold_ctx *f = (fold_ctx *)c;
    if (!f->occ[pos_x]) { f->u[pos_x] += 1; f->occ[pos_x] = 1; }
    f->v[pos_x] += 1;
}

/* distance: 0 Hamming, 1 edit.  Accumulates (+=) into u[0..m), v[0..m).  Returns 0 on
 * success, negative on failure (k = 0, edit distance with k > 64, or allocation). */
int orc_pair(int distance, const char *x, uint32_t m, const char *y, uint32_t n,
             uint32_t k, uint32_t e, uint32_t *u, uint32_t *v)
{
    fold_ctx f = { u, v, (uint8_t *)calloc(m ? m : 1, 1) };
    if (!f.occ) return -2;
    int rc = distance == 0
        ? hd_hits((const unsigned char *)x, m, (const unsigned char *)y, n, k, e, fold_hit, &f)
        : ed_hits((const unsigned char *)x, m, (const unsigned char *)y, n, k, e, fold_hit, &f);
    free(f.occ);
    return rc;
}

int orc_hd_pair(const char *x, uint32_t m, const char *y, uint32_t n, uint32_t k, uint32_t e,
                uint32_t *u, uint32_t *v)
{ return orc_pair(0, x, m, y, n, k, e, u, v); }

int orc_ed_pair(const char *x, uint32_t m, const char *y, uint32_t n, uint32_t k, uint32_t e,
                uint32_t *u, uint32_t *v)
{ return orc_pair(1, x, m, y, n, k, e, u, v); }

/* Raw tuple dump for tests: writes up to cap tuples (pos_t, pos_x, err) into out[3*i..],
 * returns the total number of tuples (which may exceed cap). */
typedef struct { uint32_t *out; uint64_t cap, cnt; } dump_ctx;
static void dump_hit(void *c, uint32_t pos_t, uint32_t pos_x, uint32_t err)
{
    dump_ctx *d = (dump_ctx *)c;
    if (d->cnt < d->cap) {
        d->out[3 * d->cnt] = pos_t; d->out[3 * d->cnt + 1] = pos_x; d->out[3 * d->cnt + 2] = err;
    }
    d->cnt++;
}

int64_t orc_hits(int distance, const char *x, uint32_t m, const char *y, uint32_t n,
                 uint32_t k, uint32_t e, uint32_t *out, uint64_t cap)
{
    dump_ctx d = { out, cap, 0 };
    int rc = distance == 0
        ? hd_hits((const unsigned char *)x, m, (const unsigned char *)y, n, k, e, dump_hit, &d)
        : ed_hits((const unsigned char *)x, m, (const unsigned char *)y, n, k, e, dump_hit, &d);
    return rc < 0 ? rc : (int64_t)d.cnt;
}

/* ------------------------------------------------------------------------------------ */
/* Structured motifs (functions.cc:564-705 Hamming, 773-915 edit).                        */
/* ------------------------------------------------------------------------------------ */

/* A box matrix entry B[b][i][j] (i in 0..n, j in 0..m) holds the FLASM error of the
 * tuple (pos_t = i-1, pos_x = j-1) and m where there is none (functions.cc:597-598,
 * 641).  Only "B <= err_b" is ever read, so a flag matrix is kept. */
typedef struct { uint8_t *ok; uint32_t m1; } box_ctx;
static void box_hit(void *c, uint32_t pos_t, uint32_t pos_x, uint32_t err)
{
    (void)err;
    box_ctx *b = (box_ctx *)c;
    b->ok[(size_t)(pos_t + 1) * b->m1 + (pos_x + 1)] = 1;
}

/* S is row-major nb_structs x nb_gaps (the order orc_recurse produces). u and v hold
 * nb_structs * m entries, index s*m + pos_x. */
int orc_struct_pair(int distance, const char *x, uint32_t m, const char *y, uint32_t n,
                    uint32_t k, uint32_t e, uint32_t nb_gaps, const uint32_t *blens,
                    const uint32_t *berrs, uint32_t nb_structs, const uint32_t *S,
                    uint32_t *u, uint32_t *v)
{
    uint32_t nb_boxes = nb_gaps + 1;
    size_t cells = (size_t)(n + 1) * (m + 1);
    uint8_t *ok = (uint8_t *)malloc(cells * nb_boxes);
    uint8_t *occ = (uint8_t *)calloc((size_t)nb_structs * (m + 1), 1);
    if (!ok || !occ) { free(ok); free(occ); return -2; }
    for (uint32_t b = 0; b < nb_boxes; ++b) {
        uint32_t ell = b ? blens[b - 1] : k;
        uint32_t err = b ? berrs[b - 1] : e;
        uint8_t *okb = ok + cells * b;
        /* entries without a tuple hold m; "m <= err" makes them pass */
        memset(okb, m <= err ? 1 : 0, cells);
        box_ctx bc = { okb, m + 1 };
        int rc = distance == 0
            ? hd_hits((const unsigned char *)x, m, (const unsigned char *)y, n, ell, err, box_hit, &bc)
            : ed_hits((const unsigned char *)x, m, (const unsigned char *)y, n, ell, err, box_hit, &bc);
        if (rc < 0) { free(ok); free(occ); return rc; }
    }
    /* merge, functions.cc:647-690 */
    for (uint32_t i = 0; i <= n; ++i) {
        for (uint32_t j = 0; j <= m; ++j) {
            if (!ok[(size_t)i * (m + 1) + j]) continue;
            for (uint32_t kk = 0; kk < nb_structs; ++kk) {
                for (uint32_t l = 0; l < nb_structs; ++l) {
                    uint32_t dj = j, di = i, r = 1;
                    for (uint32_t g = 0; g < nb_gaps; ++g) {
                        dj += S[(size_t)kk * nb_gaps + g] + blens[g];
                        di += S[(size_t)l * nb_gaps + g] + blens[g];
                        if (di < n + 1 && dj < m + 1) {
                            if (ok[cells * (g + 1) + (size_t)di * (m + 1) + dj]) {
                                r++;
                                if (r == nb_boxes) {
                                    size_t o = (size_t)kk * m + j - 1;
                                    if (!occ[(size_t)kk * (m + 1) + j - 1]) {
                                        u[o] += 1;
                                        occ[(size_t)kk * (m + 1) + j - 1] = 1;
                                    }
                                    v[o] += 1;
                                }
                            } else {
                                break;
                            }
                        }
                    }
                }
            }
        }
    }
    free(ok); free(occ);
    return 0;
}

/* Cartesian product of gap ranges in the order of functions.cc:710-729: the last gap
 * varies fastest.  Returns nb_structs; writes S (row-major) when S != NULL. */
uint32_t orc_recurse(uint32_t nb_gaps, const uint32_t *gmin, const uint32_t *gmax, uint32_t *S)
{
    uint32_t total = 1;
    for (uint32_t g = 0; g < nb_gaps; ++g) total *= gmax[g] - gmin[g] + 1;
    if (!S) return total;
    for (uint32_t s = 0; s < total; ++s) {
        uint32_t rem = s;
        for (int g = (int)nb_gaps - 1; g >= 0; --g) {
            uint32_t span = gmax[g] - gmin[g] + 1;
            S[(size_t)s * nb_gaps + g] = gmin[g] + rem % span;
            rem /= span;
        }
    }
    return total;
}

/* ------------------------------------------------------------------------------------ */
/* All-pairs driver, motex.cc:494-583.                                                     */
/* ------------------------------------------------------------------------------------ */

/* seqs: concatenated bytes, off[i]..off[i]+len[i].  u/v: flat, entry i starts at
 * uv_off[i] and holds nb_structs*len[i] words (caller-zeroed, accumulated).  threads <= 0
 * means the OpenMP default.  Returns the number of skipped (short) pairs, or negative
 * on failure. */
int64_t orc_all_pairs_range(int distance, const char *seqs, const uint64_t *off,
                            const uint32_t *len, uint32_t num, uint32_t x0, uint32_t x1,
                            uint32_t k, uint32_t e, uint32_t nb_gaps, const uint32_t *gmin,
                            const uint32_t *gmax, const uint32_t *blens, const uint32_t *berrs,
                            const uint64_t *uv_off, uint32_t *u, uint32_t *v, int threads);

int64_t orc_all_pairs(int distance, const char *seqs, const uint64_t *off, const uint32_t *len,
                      uint32_t num, uint32_t k, uint32_t e, uint32_t nb_gaps,
                      const uint32_t *gmin, const uint32_t *gmax, const uint32_t *blens,
                      const uint32_t *berrs, const uint64_t *uv_off, uint32_t *u, uint32_t *v,
                      int threads)
{
    return orc_all_pairs_range(distance, seqs, off, len, num, 0, num, k, e, nb_gaps, gmin, gmax,
                               blens, berrs, uv_off, u, v, threads);
}

/* x sequences [x0, x1) against every y: a bounded sample of the all-pairs loop. */
int64_t orc_all_pairs_range(int distance, const char *seqs, const uint64_t *off,
                            const uint32_t *len, uint32_t num, uint32_t x0, uint32_t x1,
                            uint32_t k, uint32_t e, uint32_t nb_gaps, const uint32_t *gmin,
                            const uint32_t *gmax, const uint32_t *blens, const uint32_t *berrs,
                            const uint64_t *uv_off, uint32_t *u, uint32_t *v, int threads)
{
    uint32_t total_mot = k;
    for (uint32_t g = 0; g < nb_gaps; ++g) total_mot += gmin[g] + blens[g];
    uint32_t nb_structs = orc_recurse(nb_gaps, gmin, gmax, NULL);
    uint32_t *S = NULL;
    if (nb_gaps) {
        S = (uint32_t *)malloc((size_t)nb_structs * nb_gaps * sizeof(uint32_t));
        if (!S) return -2;
        orc_recurse(nb_gaps, gmin, gmax, S);
    }
    int64_t skipped = 0;
    int failed = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : skipped)
#endif
    for (int64_t i = x0; i < (int64_t)x1; ++i) {
        uint32_t m = len[i];
        for (uint32_t j = 0; j < num; ++j) {
            uint32_t n = len[j];
            if (total_mot > n || total_mot > m) { skipped++; continue; }
            int rc;
            if (nb_gaps == 0)
                rc = orc_pair(distance, seqs + off[i], m, seqs + off[j], n, k, e,
                              u + uv_off[i], v + uv_off[i]);
            else
                rc = orc_struct_pair(distance, seqs + off[i], m, seqs + off[j], n, k, e,
                                     nb_gaps, blens, berrs, nb_structs, S,
                                     u + uv_off[i], v + uv_off[i]);
            if (rc < 0) failed = 1;
        }
    }
    free(S);
    return failed ? -1 : skipped;
}

/* ------------------------------------------------------------------------------------ */
/* Writers.                                                                                */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    const char *alphabet;   /* "DNA" | "PROT" | "USR" */
    const char *input_filename, *output_filename, *boxes_filename;
    uint32_t q, Q, k, d, e, n, N;
    uint32_t nb_gaps;
    const uint32_t *gmin, *gmax, *blens, *berrs;
    uint32_t total_length;  /* sum of sequence lengths */
} orc_sw;

/* Header, functions.cc:1010-1068.  The Rundate line and the timing line vary from run
 * to run; callers compare outputs with those two lines masked. */
static void write_header(FILE *f, const orc_sw *sw, uint32_t num_seqs, double secs, int P)
{
    fprintf(f, "####################################\n");
    fprintf(f, "# Program: MoTeX\n");
    fprintf(f, "# Rundate: (masked)\n");
    fprintf(f, "# Input file: %s\n", sw->input_filename);
    fprintf(f, "# Output file: %s\n", sw->output_filename);
    fprintf(f, "# For %d input sequences\n", (int)num_seqs);
    fprintf(f, "#     d = %d\n", (int)sw->d);
    fprintf(f, "#     q = %d\n", (int)sw->q);
    if (sw->Q != 100) fprintf(f, "#     Q = %d\n", (int)sw->Q);
    if (sw->n != 1) fprintf(f, "#     n = %d\n", (int)sw->n);
    if (sw->N != 10000) fprintf(f, "#     N = %d\n", (int)sw->N);
    if (sw->nb_gaps == 0) {
        fprintf(f, "#     l = %d\n", (int)sw->k);
        fprintf(f, "#     e = %d\n", (int)sw->e);
    } else {
        fprintf(f, "# Number of boxes: %d\n", (int)(sw->nb_gaps + 1));
        fprintf(f, "#     l0 = %d\n", (int)sw->k);
        fprintf(f, "#     e0 = %d\n", (int)sw->e);
        for (uint32_t g = 0; g < sw->nb_gaps; ++g) {
            fprintf(f, "#     g%d = [%d,%d]\n", (int)g + 1, (int)sw->gmin[g], (int)sw->gmax[g]);
            fprintf(f, "#     l%d = %d\n", (int)g + 1, (int)sw->blens[g]);
            fprintf(f, "#     e%d = %d\n", (int)g + 1, (int)sw->berrs[g]);
        }
        fprintf(f, "# Boxes input file: %s\n", sw->boxes_filename);
    }
    fprintf(f, "# Run on %d proc(s) in %lf secs\n", P, secs);
    fprintf(f, "####################################\n\n");
}

/* The quorum/occurrence filter in double arithmetic, functions.cc:1107-1108. */
static int passes(const orc_sw *sw, uint32_t num_seqs, uint32_t uu, uint32_t vv)
{
    double r = ((double)uu / num_seqs) * 100;
    return r >= sw->q && vv >= sw->n && r <= sw->Q && vv <= sw->N;
}

/* A tiny open-addressing string set for first-occurrence dedup (the reference uses a
 * libdatrie trie, functions.cc:1092-1098, 1114-1116; only membership is observable). */
typedef struct { char **keys; size_t cap, cnt; } strset;
static uint64_t fnv(const char *s) { uint64_t h = 1469598103934665603ull; while (*s) { h ^= (unsigned char)*s++; h *= 1099511628211ull; } return h; }
static int strset_add(strset *st, const char *s)  /* 1 if newly added */
{
    if ((st->cnt + 1) * 2 > st->cap) {
        size_t nc = st->cap ? st->cap * 2 : 1024;
        char **nk = (char **)calloc(nc, sizeof(char *));
        for (size_t i = 0; i < st->cap; ++i)
            if (st->keys[i]) {
                size_t h = fnv(st->keys[i]) & (nc - 1);
                while (nk[h]) h = (h + 1) & (nc - 1);
                nk[h] = st->keys[i];
            }
        free(st->keys); st->keys = nk; st->cap = nc;
    }
    size_t h = fnv(s) & (st->cap - 1);
    while (st->keys[h]) {
        if (!strcmp(st->keys[h], s)) return 0;
        h = (h + 1) & (st->cap - 1);
    }
    st->keys[h] = strdup(s); st->cnt++;
    return 1;
}
static void strset_free(strset *st) { for (size_t i = 0; i < st->cap; ++i) free(st->keys[i]); free(st->keys); }

static const char *alpha_str(const char *name)
{
    if (!strcmp(name, "DNA")) return ORC_DNA;
    if (!strcmp(name, "PROT")) return ORC_PROT;
    return ORC_USR;
}

/* Build the printed motif for entry (i, j, s): plain k-mer, or boxes joined by '_'
 * (functions.cc:1191-1231). */
static void motif_text(const orc_sw *sw, const char *seq, uint32_t j, const uint32_t *Srow, char *buf)
{
    uint32_t jj = j - sw->k + 1, idx = 0;
    memcpy(buf, seq + jj, sw->k); idx = sw->k;
    if (sw->nb_gaps) {
        uint32_t offset = 0, ell = sw->k;
        buf[idx++] = '_';
        for (uint32_t b = 1; b <= sw->nb_gaps; ++b) {
            offset += Srow[b - 1] + ell;
            ell = sw->blens[b - 1];
            memcpy(buf + idx, seq + jj + offset, ell); idx += ell;
            buf[idx++] = '_';
        }
        idx--;
    }
    buf[idx] = 0;
}

/* Writes the .motex body (functions.cc:1073-1142 / 1147-1260) and, when smile != NULL,
 * the SMILE body (functions.cc:1265-1461 / 1466-1608).  u, v flat with uv_off as in
 * orc_all_pairs.  Returns the number of motifs written. */
int64_t orc_write(const orc_sw *sw, const char *seqs, const uint64_t *off, const uint32_t *len,
                  uint32_t num, const uint64_t *uv_off, const uint32_t *u, const uint32_t *v,
                  double secs, int P, const char *out_path, const char *smile_path)
{
    FILE *f = fopen(out_path, "w");
    if (!f) return -1;
    FILE *sf = NULL;
    if (smile_path) { sf = fopen(smile_path, "w"); if (!sf) { fclose(f); return -1; } }
    write_header(f, sw, num, secs, P);

    uint32_t nb_structs = orc_recurse(sw->nb_gaps, sw->gmin, sw->gmax, NULL);
    uint32_t *S = NULL;
    if (sw->nb_gaps) {
        S = (uint32_t *)malloc((size_t)nb_structs * sw->nb_gaps * sizeof(uint32_t));
        orc_recurse(sw->nb_gaps, sw->gmin, sw->gmax, S);
    }
    uint32_t tot_sm = sw->k;
    for (uint32_t g = 0; g < sw->nb_gaps; ++g) tot_sm += sw->blens[g];
    const char *astr = alpha_str(sw->alphabet);

    /* SMILE header, functions.cc:1265-1365 */
    if (sf) {
        char line1[1024];
        int need = (int)ceil((double)((double)sw->q * num) / (double)100);
        if (!sw->nb_gaps) {
            snprintf(line1, sizeof line1, "%%%%%% 1 %d/%d %d %d %d %d alphabet %s$", need, (int)num,
                     (int)sw->total_length, (int)sw->k, (int)sw->k, (int)sw->e, astr);
        } else {
            uint32_t mn = sw->k, mx = sw->k, te = sw->e;
            for (uint32_t g = 0; g < sw->nb_gaps; ++g) { mn += sw->blens[g]; mx += sw->blens[g]; te += sw->berrs[g]; }
            int p = snprintf(line1, sizeof line1, "%%%%%% %d %d/%d %d %d %d %d %d %d %d ",
                             (int)(sw->nb_gaps + 1), need, (int)num, (int)sw->total_length, (int)mn,
                             (int)mx, (int)te, (int)sw->k, (int)sw->k, (int)sw->e);
            for (uint32_t g = 0; g < sw->nb_gaps; ++g)
                p += snprintf(line1 + p, sizeof line1 - p, "%d %d %d %d %d ", (int)sw->gmin[g],
                              (int)sw->gmax[g], (int)sw->blens[g], (int)sw->blens[g], (int)sw->berrs[g]);
            snprintf(line1 + p, sizeof line1 - p, "alphabet %s$", astr);
        }
        fprintf(sf, "%s", line1);
        if (!strcmp(sw->alphabet, "DNA")) {
            for (size_t i = strlen(line1); i < 79; ++i) fputc(' ', sf);
            fputc('\n', sf);
            for (int r = 0; r < 2; ++r) fprintf(sf, "%79s\n", "");
        }
        if (!strcmp(sw->alphabet, "PROT")) {
            for (size_t i = strlen(line1); i < 159; ++i) fputc(' ', sf);
            fputc('\n', sf);
            fprintf(sf, "%79s\n", "");
        }
        for (int r = 0; r < 79; ++r) fputc('=', sf);
        fputc('\n', sf);
    }

    strset st = { 0, 0, 0 }, sst = { 0, 0, 0 };
    int64_t valid = 0, svalid = 0;
    char *buf = (char *)malloc(tot_sm + sw->nb_gaps + 2);
    char *id = (char *)malloc(tot_sm + sw->nb_gaps + 2);
    size_t alen = strlen(astr);
    for (uint32_t i = 0; i < num; ++i) {
        uint32_t m = len[i];
        const char *seq = seqs + off[i];
        for (uint32_t j = sw->k - 1; j < m; ++j) {
            for (uint32_t s = 0; s < nb_structs; ++s) {
                size_t o = uv_off[i] + (size_t)s * m + j;
                if (!passes(sw, num, u[o], v[o])) continue;
                motif_text(sw, seq, j, S ? S + (size_t)s * sw->nb_gaps : NULL, buf);
                if (strset_add(&st, buf)) {
                    fprintf(f, "%s %d %d %lf %d\n", buf, (int)u[o], (int)num, (double)u[o] / num, (int)v[o]);
                    valid++;
                }
                if (sf && strset_add(&sst, buf)) {
                    size_t L = strlen(buf);
                    for (size_t c = 0; c < L; ++c) {
                        size_t a = 0;
                        while (a < alen && buf[c] != astr[a]) a++;
                        id[c] = (sw->nb_gaps && a >= alen) ? '-' : (char)('0' + a);
                    }
                    id[L] = 0;
                    fprintf(sf, "%s %s %d\t%d\n", buf, id, (int)u[o], (int)v[o]);
                    svalid++;
                }
            }
        }
    }
    if (sw->nb_gaps)
        fprintf(f, "\n#A total of %d valid structured motifs were extracted.\n\n", (int)valid);
    else
        fprintf(f, "\n#A total of %d valid motifs were extracted.\n\n", (int)valid);
    if (sf) {
        fprintf(sf, "Nb models: %d\n", (int)svalid);
        fprintf(sf, "User time: (masked)\n");
        fclose(sf);
    }
    fclose(f);
    strset_free(&st); strset_free(&sst);
    free(buf); free(id); free(S);
    return valid;
}
