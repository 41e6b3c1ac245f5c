/*
 * motex.h -- C ABI of the MI355X (gfx950) all-pairs motif-extraction library.
 *
 * This is the drop-in boundary for MoTeX's extraction step.  It replaces two C++
 * interfaces of the reference (paths relative to the MoTeX repository root):
 *
 *   1. the all-pairs driver loop of motex.cc:494-583, together with the per-pair
 *      operators it calls, motexdefs.h:107-110 / functions.cc:526-559, 564-705, 734-768,
 *      773-915 (motifs_extraction_{hd,ed}, structured_motifs_extraction_{hd,ed});
 *   2. the FLASM library those operators call, libFLASM.zip
 *      libFLASM/include/libflasm.h:71-74 (flasm_ed, flasm_hd), whose result tuples are
 *      folded on the device and never materialised.
 *
 * Conventions (all entry points):
 *   - return 0 on success and a negative MOTEX_E* code on failure; the message for the
 *     calling thread is available from motex_last_error();
 *   - sequences are the FASTA reader's upper-case bytes; character equality on raw bytes
 *     decides a match, as libflasm's delta(a,b) = (a != b) does (libflasm.h:42);
 *   - u and v are caller-owned and ACCUMULATED (+=), exactly like the reference's
 *     g_occur / g_all_occur rows (motex.cc:517-529, functions.cc:548-553): u[i] and v[i]
 *     hold nb_structs * lens[i] words, entry s*lens[i] + pos_x;
 *   - results are deterministic: the device folds with integer atomics only;
 *   - a motex_ctx is used by one host thread at a time; the serving calls
 *     (motex_submit / motex_poll / motex_wait / motex_job_release) are thread-safe.
 */
#ifndef MOTEX_H
#define MOTEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOTEX_OK 0
#define MOTEX_EINVAL (-1)    /* invalid argument (the reference's validation, motex.cc:98-201) */
#define MOTEX_ENOMEM (-2)    /* host or device allocation failed, or over the HBM budget */
#define MOTEX_EDEVICE (-3)   /* HIP runtime error */
#define MOTEX_EUNSUPPORTED (-4) /* parameters outside what the kernels implement (Hamming k > 1024,
                                  structured boxes > 64, edit k > 64, ...) */
#define MOTEX_EBUSY (-5)     /* serving: job not finished yet */
#define MOTEX_ENOTFOUND (-6) /* serving: unknown job id */

#define MOTEX_ALPHA_DNA 0    /* "ACGTN"  (motexdefs.h:23) */
#define MOTEX_ALPHA_PROT 1   /* "ARNDCQEGHILKMFPSTWYVOUBZJX*" (motexdefs.h:24) */
#define MOTEX_ALPHA_USR 2    /* "0123456789" (motexdefs.h:25) */

typedef struct motex_ctx motex_ctx;
typedef struct motex_batch motex_batch;
typedef uint64_t motex_job_id;

/* Extraction parameters: the subset of struct TSwitch (motexdefs.h:32-61) the
 * extraction step reads.  The structure arrays follow the boxes file (motex.cc:303-357):
 * gap g lies between box g and box g+1 and has length in [gap_min[g], gap_max[g]];
 * box g+1 has length box_len[g] and at most box_err[g] errors.  Box 0 is (k, e).
 * The nb_structs gap combinations are enumerated in the order of functions.cc:710-729
 * (last gap fastest). */
typedef struct motex_params {
    uint32_t k;             /* -k, motif (box 0) length */
    uint32_t e;             /* -e, errors allowed in box 0 */
    uint32_t distance;      /* -d, 0 = Hamming, 1 = edit */
    uint32_t alphabet;      /* MOTEX_ALPHA_*, used for validation only */
    uint32_t nb_gaps;       /* 0 for single motifs (no -s) */
    const uint32_t *gap_min;
    const uint32_t *gap_max;
    const uint32_t *box_len;
    const uint32_t *box_err;
    uint32_t long_mode;     /* -L: scheduling hint, never changes results */
} motex_params;

/* Per-run statistics of a batch (filled by motex_batch_stats). */
typedef struct motex_stats {
    double cells;           /* algorithmic cells: sum over computed pairs of (m-k+1)*n
                               (per distinct box DP for structured runs) */
    double kernel_ms;       /* device time of the last run (HIP events on the run stream) */
    double main_kernel_ms;  /* device time of the dominant kernel of the last run */
    uint64_t work_items;    /* items in the device work queue */
    uint64_t pairs;         /* ordered pairs computed (after the short-sequence skip) */
    uint64_t skipped_pairs; /* pairs skipped as too short (motex.cc:536-540) */
    uint32_t tile_blocks;   /* autotuned tile shape used by the batched kernel */
    uint32_t tile_rows;
    uint32_t waves_per_cu;
    uint32_t kernel_variant;  /* edit distance: 0/1/4 one text per lane (32/64/128-bit words),
                                 2 text-parallel one wave per item, 3 text-parallel pipelined;
                                 structured: 0 one pair per item, 1 range-batched (Hamming boxes) */
    uint32_t graph_replay;  /* 1 if the last run replayed the batch's HIP graph */
    uint32_t symmetric;     /* 1 if each unordered pair is computed once and credited to
                               both orders (Hamming; cells still count every ordered pair) */
    double swept_cells;     /* cells the kernel actually sweeps (Hamming: wave rows x
                               diagonals of their windows, warm-up and padding included) */
    uint64_t handed_pairs;  /* range-batched structured runs: pairs whose range overflowed its
                               hit queue and went to the one-pair-per-item kernel */
} motex_stats;

/* A serving job.  The caller keeps seqs/lens/params and the output arrays alive until
 * the job completes; u and v are accumulated into on completion. */
typedef struct motex_job {
    const char *const *seqs;
    const uint32_t *lens;
    uint32_t num_seqs;
    motex_params params;
    uint32_t *const *u;
    uint32_t *const *v;
} motex_job;

/* --- context ---------------------------------------------------------------------- */

/* Opens a context on HIP device `device`.  hbm_budget_bytes bounds the device memory the
 * context may hold for batches and serving jobs (0 = 90% of free memory). */
int motex_ctx_create(int device, size_t hbm_budget_bytes, motex_ctx **out);
void motex_ctx_destroy(motex_ctx *ctx);

/* Thread-local description of the last failure on this thread. */
const char *motex_last_error(void);

/* --- extraction: replaces motex.cc:494-583 + the per-pair operators --------------- */

/* Batched persistent form: every ordered pair (i, j), j == i included, in one
 * persistent kernel over a length-sorted device work queue.  Pairs whose sequences are
 * shorter than the total motif length are skipped, as motex.cc:536-540 does. */
int motex_extract_all_pairs(motex_ctx *ctx, const char *const *seqs, const uint32_t *lens,
                            uint32_t num_seqs, const motex_params *params,
                            uint32_t *const *u, uint32_t *const *v);

/* Plain form: one pair, one launch -- the per-pair operator of motexdefs.h:107-110
 * (x = p, m; y = t, n).  Accumulates into u, v (nb_structs*m words each). */
int motex_extract_pair(motex_ctx *ctx, const char *x, uint32_t m, const char *y, uint32_t n,
                       const motex_params *params, uint32_t *u, uint32_t *v);

/* --- extraction without device allocation (no counterpart in the reference) -------- */

/* Bytes of the per-device scratch array the no-allocation entry point works in (64 MiB). */
#define MOTEX_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* motex_extract_all_pairs for Hamming single motifs (distance 0, nb_gaps 0) that calls no
 * hipMalloc/hipFree and no allocating device library routine: every device buffer of the
 * run comes from one static array of MOTEX_NOALLOC_SCRATCH_BYTES per device, which each GPU
 * holds from the context's creation until the process ends.  For callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, accumulation and results as motex_extract_all_pairs; MOTEX_ENOMEM when the
 * input needs more scratch than the array holds (about 4 M nt in 1000-nt sequences),
 * MOTEX_EUNSUPPORTED for edit distance or structured motifs.  The array keeps the last call's
 * contents; calls on one device share it, so the library runs them one at a time within a
 * process, and calls from several processes on one GPU race. */
int motex_extract_all_pairs_noalloc(motex_ctx *ctx, const char *const *seqs, const uint32_t *lens,
                                    uint32_t num_seqs, const motex_params *params,
                                    uint32_t *const *u, uint32_t *const *v);

/* Cross form, for re-searching motifs in a set of sequences (-I, motex.cc:883-1137): every
 * x sequence (a motif string read back from an earlier run's unmatched-motif file) against
 * every text y, with the per-pair operator of the loop at motex.cc:940-990
 * (motifs_extraction_{hd,ed}, x = fseqs[i], y = seqs[j]).  u[i], v[i] hold xlens[i] words,
 * accumulated; a text shorter than k is skipped as motex.cc:948-952 does.  Single motifs only
 * (the reference refuses -I with -s, motex.cc:193-199). */
int motex_extract_cross(motex_ctx *ctx, const char *const *xs, const uint32_t *xlens, uint32_t num_x,
                        const char *const *ys, const uint32_t *ylens, uint32_t num_y,
                        const motex_params *params, uint32_t *const *u, uint32_t *const *v);

/* --- device-resident batches (the timed form: inputs already in HBM) --------------- */

/* Uploads and packs the sequences, builds the work queue and picks tile shapes. */
int motex_batch_create(motex_ctx *ctx, const char *const *seqs, const uint32_t *lens,
                       uint32_t num_seqs, const motex_params *params, motex_batch **out);
/* Zeroes the device accumulators and runs the batched persistent kernel over all pairs;
 * blocks until done.  Results stay on the device. */
int motex_batch_run(motex_batch *b);
/* Plain one-launch-per-pair form over x sequences [first_x, first_x + num_x) against
 * every y; accumulates on the device (call motex_batch_clear first to start from 0). */
int motex_batch_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x);
int motex_batch_clear(motex_batch *b);
/* Copies device u/v back, accumulating into the caller's rows. */
int motex_batch_download(motex_batch *b, uint32_t *const *u, uint32_t *const *v);
int motex_batch_stats(const motex_batch *b, motex_stats *out);
void motex_batch_destroy(motex_batch *b);

/* --- tile-shape autotuner (no counterpart in the reference) ------------------------ */

/* Tile shape picked by motex_autotune. */
typedef struct motex_tile {
    uint32_t blocks;        /* Hamming: 2048-diagonal blocks per wave */
    uint32_t rows;          /* Hamming: x rows per work item */
    uint32_t waves_per_cu;  /* persistent grid = CUs x waves_per_cu */
    uint32_t candidates;    /* shapes timed */
    double kernel_ms;       /* batched-kernel time of the winner on the sample */
} motex_tile;

/* Times the candidate tile shapes of the batched kernel for these parameters on the given
 * sample of sequences and keeps the fastest in the context: later batches and jobs with
 * the same (distance/structured, alphabet size, k, e) use it.  Results never depend on the
 * tile shape, only the speed does. */
int motex_autotune(motex_ctx *ctx, const char *const *seqs, const uint32_t *lens,
                   uint32_t num_seqs, const motex_params *params, motex_tile *best);

/* --- serving: continuous batching of extraction jobs ------------------------------- */

/* Queues a job; it is admitted to the device when its HBM footprint fits the budget and
 * runs concurrently with jobs already resident.  Returns at once. */
int motex_submit(motex_ctx *ctx, const motex_job *job, motex_job_id *id);
/* *done = 1 once the job's u, v have been accumulated into the caller's arrays. */
int motex_poll(motex_ctx *ctx, motex_job_id id, int *done);
/* Blocks until the job is done. */
int motex_wait(motex_ctx *ctx, motex_job_id id);
/* Forgets a finished job (its id becomes invalid). */
int motex_job_release(motex_ctx *ctx, motex_job_id id);
/* Per-job timing of a finished job: submit->done wall time and device time, in ms. */
int motex_job_times(motex_ctx *ctx, motex_job_id id, double *latency_ms, double *device_ms,
                    double *cells);

#ifdef __cplusplus
}
#endif
#endif /* MOTEX_H */
