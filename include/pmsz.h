/*
 * pmsz.h -- C-ABI of libpmsz.so, the MI355X (gfx950) extrema + Morse-Smale
 * segmentation correction engine.  Plain C types only.
 *
 * The reference exposes no library interface: its correction entry point is the
 * internal template c_loops<T>(...) (pMSz.cu:1872-1878, 35 positional arguments
 * plus globals) driven by main (pMSz.cu:2596-3188).  Each function below names
 * the reference code it replaces.
 *
 * Conventions
 *   - dtype 0 = float32, 1 = float64; arrays are C order with x fastest
 *     (index = x + nx*(y + ny*z)), as the reference's raw reader (pMSz.cu:289-309).
 *   - "d_" pointers are device pointers (HIP), "h_" pointers host pointers.
 *   - Return value 0 = success, negative = error (pmsz_strerror); the reference
 *     prints and calls MPI_Abort instead (pMSz.cu:1899-1910).
 *   - The library owns its scratch state (codes, delta counters, marks, work
 *     lists), allocated on first use per device and reused; the caller owns f / g.
 *   - Calls are serialised per context by its lock and run on the context's
 *     own stream (the functions without a context argument use the library's
 *     default context); they return when the result is complete.
 */
#ifndef PMSZ_H
#define PMSZ_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { PMSZ_F32 = 0, PMSZ_F64 = 1 } pmsz_dtype;
/* lazy ghost sync (pMSz.cu) or per-iteration sync (sync_pMSz.cu) */
/* PMSZ_NAIVE: the naive-MSz baseline (naive_MSz.cu:1887-2219): extrema-only
 * sweeps, then path re-tracing against the labels, per outer iteration; one
 * domain only (sub = 1,1,1), and not in batches or the autotuner. */
typedef enum { PMSZ_LAZY = 0, PMSZ_SYNC = 1, PMSZ_NAIVE = 2 } pmsz_sync_mode;
/* clamp of the edit primitive: the reference's g = f - eb (pMSz.cu:636) in the
 * value type, or -- when that value misses |f - g| <= eb exactly -- the smallest
 * value of the type that meets it */
typedef enum { PMSZ_CLAMP_REFERENCE = 0, PMSZ_CLAMP_BOUND_SAFE = 1 } pmsz_clamp_mode;
/* ghost cells exchanged between subdomains: the reference's faces + 4 edges
 * (pMSz.cu:394-488) or all 14 stencil directions */
typedef enum { PMSZ_GHOST_REFERENCE = 0, PMSZ_GHOST_FULL = 1 } pmsz_ghost_policy;

enum {
  PMSZ_OK = 0,
  PMSZ_EINVAL = -1,        /* bad parameter, NaN / sentinel value in a field */
  PMSZ_ENOMEM = -2,        /* device allocation failed */
  PMSZ_NOT_CONVERGED = -3, /* max_sweeps reached; stats are filled */
  PMSZ_EDEVICE = -4,       /* HIP runtime error / no device */
  PMSZ_ETOOBIG = -5,       /* more voxels than the index width supports */
  PMSZ_EBUSY = -6,         /* batch / server capacity exhausted */
};

typedef struct {
  int dtype;             /* pmsz_dtype */
  int64_t nx, ny, nz;    /* field extent (argv[1] "path,W,H,D", pMSz.cu:2648-2661) */
  double abs_eb;         /* absolute error bound (argv[2], pMSz.cu:2628) */
  int q;                 /* decrements before clamping; 5 in the reference (pMSz.cu:2630) */
  int sub[3];            /* subdomain grid (argv[4..6] PX PY PZ, pMSz.cu:2631-2633) */
  int mode;              /* pmsz_sync_mode */
  int max_sweeps;        /* cap on sweeps / outer iterations; the reference has none */
  int clamp_mode;        /* pmsz_clamp_mode */
  int ghost_policy;      /* pmsz_ghost_policy */
} pmsz_params;

typedef struct {
  int64_t edit_sweeps;     /* detect+edit sweeps of the reference schedule (summed over subdomains) */
  int64_t check_sweeps;    /* checkFinal sweeps of the reference schedule */
  int64_t outer_iters;     /* outer iterations = ghost syncs */
  int64_t ite;             /* the reference's "iteration is:" counter (pMSz.cu:2073) */
  int64_t touched_voxels;  /* voxels edited at least once (final delta counter > 0) */
  int64_t clamped_voxels;  /* voxels clamped (final delta counter == q) */
  int64_t total_targets;   /* edits applied, summed over sweeps */
  uint32_t final_cnt[3];   /* counters of the last check: false max, false min, direction */
  int status;              /* PMSZ_OK or PMSZ_NOT_CONVERGED */
  int64_t dense_passes;    /* full-volume kernel passes actually run */
  int64_t sparse_voxels;   /* voxels re-examined by work-list sweeps */
  double t_total_s;        /* device time of the correction (events on the library stream) */
  double t_dense_s;        /* device time of detection 1, a full-volume detection sweep */
  double t_label_s;        /* device time of the original labeling pass */
  double bytes_algorithmic;/* algorithmic HBM bytes of the passes actually run */
  int64_t trace_n;         /* sweeps recorded below (the first min(64, sweeps)) */
  uint32_t trace_cnt[64][3];   /* per detection: false max, false min, direction counters */
  uint32_t trace_targets[64];  /* per detection: distinct targets */
} pmsz_stats;

/* Parameter defaults: q = 5, sub = 1,1,1, lazy, max_sweeps = 100000,
 * bound-safe clamp, full ghost policy. */
void pmsz_default_params(pmsz_params *p);

/* Correct g in place so that the extrema and both steepest-gradient labelings of
 * g equal those of f, every edit staying within abs_eb.  Replaces c_loops
 * (pMSz.cu:1872-2074 lazy / sync_pMSz.cu:2035-2204 sync) together with the
 * buffers main sets up for it (pMSz.cu:2833-2900).  d_dc (nullable, N bytes)
 * receives the final delta counters (delta_counter, pMSz.cu:173). */
int pmsz_correct(const pmsz_params *p, const void *d_f, void *d_g, uint8_t *d_dc,
                 pmsz_stats *stats);

/* With sub = (PX, PY, PZ) != (1, 1, 1), pmsz_correct splits the field into that
 * grid of blocks (pMSz.cu:2690-2711), one padded subdomain block each, and runs
 * the reference's per-rank loops with the ghost synchronisation between them:
 * the lazy outer loop (pMSz.cu:1943-2060) or the per-iteration one
 * (sync_pMSz.cu:2102-2200), the initial ghost values of exchange_ghost_layers
 * (pMSz.cu:320-608) and the delta-counter exchange (pMSz.cu:909-1272).  All
 * blocks live in this GPU's memory. */

/* Batch: n independent fields of one dtype, mode and max_sweeps (each with its
 * own extent, bound, q and clamp mode), corrected together; field k's result and
 * stats[k] equal those of pmsz_correct on that field alone.  The reference runs
 * one field per process (main, pMSz.cu:2596-3188); this is its many-field form.
 * d_dc may be NULL, as may any of its entries.  sub must be 1,1,1. */
int pmsz_correct_batch(int n, const pmsz_params *params, const void *const *d_f,
                       void *const *d_g, uint8_t *const *d_dc, pmsz_stats *stats);

/* Contexts.  The functions above run in the library's default context; a
 * context of one's own has its own HIP stream, scratch state and lock, so runs
 * in different contexts overlap on the device (the serving loop keeps one per
 * lane).  A context belongs to the device current at its creation.  The
 * reference has one correction per MPI rank and no such object; these entry
 * points are the concurrent form of pmsz_correct / pmsz_correct_batch.  At most
 * 8 contexts exist at once, the default one included (PMSZ_EBUSY beyond):
 * their runs' device-wide phases are sized to fit on the device together. */
typedef struct pmsz_ctx pmsz_ctx;
int pmsz_ctx_create(pmsz_ctx **out);
int pmsz_ctx_destroy(pmsz_ctx *ctx);
int pmsz_ctx_correct(pmsz_ctx *ctx, const pmsz_params *p, const void *d_f, void *d_g,
                     uint8_t *d_dc, pmsz_stats *stats);
int pmsz_ctx_correct_batch(pmsz_ctx *ctx, int n, const pmsz_params *params,
                           const void *const *d_f, void *const *d_g, uint8_t *const *d_dc,
                           pmsz_stats *stats);
/* Library-owned device bytes the context holds now (its buffers only grow). */
int64_t pmsz_ctx_bytes(pmsz_ctx *ctx);
/* Frees the context's per-field buffers (they are re-grown on the next call). */
int pmsz_ctx_trim(pmsz_ctx *ctx);
/* Library-owned device bytes a batch run of these n fields needs (codes, delta
 * counters, marks, work lists); the caller's f and g are not included.  Used to
 * admit work by HBM budget. */
int64_t pmsz_workspace_bytes(int n, const pmsz_params *params);

/* The correction server: continuous batching of correction jobs over lanes,
 * admitted by an HBM budget.  The reference corrects one field per process run
 * (main, pMSz.cu:2596-3188); this is its served form.  A server keeps `lanes`
 * contexts (pmsz_ctx) and one host thread per lane.  A free lane takes the oldest
 * queued job it can admit -- and, when that job has fewer than alone_voxels
 * voxels, the further queued small jobs of the same dtype, mode and max_sweeps,
 * up to max_batch -- while the library-owned device bytes of all lanes (each
 * running lane's workspace, pmsz_workspace_bytes, or what its context holds if
 * more, idle lanes included) stay within hbm_budget, and corrects them as one
 * batched run (pmsz_ctx_correct_batch); idle lanes free their buffers when they
 * block an admission.  Jobs submitted while others run are admitted as lanes and
 * memory free up, and runs on different lanes overlap on the device.  Each
 * job's result equals pmsz_correct on that job alone.  The server's contexts
 * count against the 8 contexts of the process. */
typedef struct pmsz_server pmsz_server;
typedef struct {
  int lanes;            /* contexts, one host thread each: 1..7 (4) */
  int max_batch;        /* most jobs in one batched run (32) */
  int64_t hbm_budget;   /* library-owned device bytes of all lanes together (64 GiB) */
  int64_t alone_voxels; /* a job of this many voxels or more runs by itself (2^20) */
} pmsz_server_config;
typedef struct {
  int status;           /* the job's return code: PMSZ_OK, PMSZ_NOT_CONVERGED or an error */
  int batch_size;       /* jobs of the run that corrected it */
  int lane;             /* the lane that ran it */
  double t_submit, t_start, t_done; /* CLOCK_MONOTONIC seconds */
} pmsz_job_info;
void pmsz_default_server_config(pmsz_server_config *c);
/* The lanes' contexts belong to the device current at this call. */
int pmsz_server_create(const pmsz_server_config *c, pmsz_server **out);
/* Corrects every queued job, then frees the server (unpolled jobs are dropped). */
int pmsz_server_destroy(pmsz_server *s);
/* Queues the correction of d_g (device, in place) against d_f with parameters p
 * (sub = 1,1,1; not PMSZ_NAIVE); d_dc (nullable) receives the delta counters.
 * The inputs are read once `stream` (a hipStream_t of the caller; NULL: the null
 * stream) has reached this call.  *job identifies the job.  PMSZ_EBUSY: the
 * job's workspace alone exceeds the budget, or the server is closing. */
int pmsz_submit(pmsz_server *s, const pmsz_params *p, const void *d_f, void *d_g,
                uint8_t *d_dc, void *stream, uint64_t *job);
/* Waits up to timeout_ms (0: no wait, negative: no limit) for the job.
 * PMSZ_OK: it is done, its results are in d_g / d_dc, *stats and *info
 * (nullable) are filled, and the job id is released; PMSZ_EBUSY: not done yet;
 * PMSZ_EINVAL: unknown job. */
int pmsz_poll(pmsz_server *s, uint64_t job, int timeout_ms, pmsz_stats *stats,
              pmsz_job_info *info);
/* Runs made, jobs corrected and jobs queued so far (each nullable). */
int pmsz_server_stats(pmsz_server *s, int64_t *runs, int64_t *jobs_done, int64_t *queued);

/* Tuning of a context: parameters of the schedule and of the full-volume
 * kernel's tiles.  They change the speed of a run, never its result.  The
 * reference has no counterpart (its schedule is fixed: a full sweep per
 * iteration, c_loops pMSz.cu:1943-2060). */
typedef struct {
  int dense_div;     /* detection j+1 over every voxel when 15 x targets >= voxels / dense_div (32) */
  int tail_targets;  /* the tail kernel takes over at or below this many targets (4096) */
  int tail_blocks;   /* workgroups (256 threads each) of the tail kernel, 1..256 (32) */
  int solo_targets;  /* the tail continues in one workgroup, with workgroup barriers, once a
                        sweep has at most this many targets (64; 0: never) */
  int zchunk;        /* planes per column unit of the full-volume kernel: 16, 32 or 64 (64) */
  int graphs;        /* 1: multi-domain runs (batches, subdomains) replay each chain of 8 sweeps
                        as a captured HIP graph, kept by the context per job layout (0) */
  int stream_div;    /* one field: the targets of a full-volume detection with at least
                        voxels / stream_div violations are edited by one pass over the target
                        byte map instead of a target list (32; 0: never) */
} pmsz_tuning;
void pmsz_default_tuning(pmsz_tuning *t);
/* ctx NULL: the default context. */
int pmsz_ctx_set_tuning(pmsz_ctx *ctx, const pmsz_tuning *t);
int pmsz_ctx_get_tuning(pmsz_ctx *ctx, pmsz_tuning *t);
/* Autotuner: corrects copies of the given fields (n == 1: one field, any sub;
 * n > 1: a batch, as pmsz_ctx_correct_batch) under candidate tunings -- a
 * coordinate descent over zchunk, dense_div, tail_targets, tail_blocks,
 * solo_targets, (for one field) stream_div and (for several domains) graphs, each
 * candidate timed as the best of `reps` runs -- and keeps the fastest in the
 * context.  d_g0 are not modified; the library allocates the copies for the
 * duration of the call.  *best / *best_seconds (nullable) receive the result. */
int pmsz_autotune(pmsz_ctx *ctx, int n, const pmsz_params *params, const void *const *d_f,
                  const void *const *d_g0, int reps, pmsz_tuning *best, double *best_seconds);

/* The same with host arrays: copies in, corrects, copies the result back
 * (the H2D / D2H copies of main, pMSz.cu:2870-2900, 2971). */
int pmsz_correct_host(const pmsz_params *p, const void *h_f, void *h_g, uint8_t *h_dc,
                      pmsz_stats *stats);

/* Original classification codes of a field (iscriticle data_type=1,
 * pMSz.cu:649-784): per voxel, low nibble = rank of the steepest-ascent
 * candidate, high nibble = rank of the steepest-descent candidate, where the 15
 * candidates (the voxel and its 14 stencil neighbours, pMSz.cu:186-193) are
 * ranked 0..14 by linear index and rank 7 is the voxel itself (an extremum). */
int pmsz_codes(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *d_field,
               uint8_t *d_codes);

/* Path-compressed ascending / descending Morse-Smale labels: the linear index of
 * the maximum / minimum each voxel's steepest path ends at.  Replaces
 * PathCompression (pMSz.cu:255-271) over iscriticle's pointers
 * (preserveMSS=1, pMSz.cu:739-760).  N must be < 2^31. */
int pmsz_labels(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *d_field,
                int32_t *d_asc, int32_t *d_desc);

/* Per-voxel delta symbols and the edited-voxel count (extract_edits_kernel,
 * pMSz.cu:2090-2138): symbol q where g is the clamp value, round((g0-g)/delta)
 * where g < g0, else 0. */
int pmsz_edit_symbols(const pmsz_params *p, const void *d_f, const void *d_g0,
                      const void *d_g, uint8_t *d_symbols, int64_t *edited);

/* The edit side channel of one corrected field, per rank block: the blocks of
 * sub = {PX, PY, PZ} (pMSz.cu:2690-2706) in the reference's rank order, which
 * MPI_Cart_create without reordering numbers row-major over {PX, PY, PZ}
 * (pMSz.cu:2673-2682): the z coordinate varies fastest. */
typedef struct {
  int64_t voxels;        /* core voxels of the block */
  int64_t packed_offset; /* byte offset of its packed symbol stream in d_packed */
  int64_t packed_bytes;  /* ceil(voxels * nbits / 8) */
  int64_t value_offset;  /* element offset of its clamp values in d_values */
  int64_t values;        /* clamp values kept */
} pmsz_edit_block;

typedef struct {
  int32_t nbits;          /* bits per symbol: bits_for_max(q) (pMSz.cu:2352-2356) */
  int32_t blocks;         /* PX * PY * PZ */
  int64_t edited;         /* voxels with g < g0 (h_count, pMSz.cu:2108) */
  int64_t values;         /* clamp values kept, all blocks */
  int64_t packed_capacity; /* bytes d_packed must hold (pmsz_edit_packed_bytes) */
} pmsz_edit_info;

/* Bytes the packed symbol streams of pmsz_encode_edits take at d_packed: each
 * block's stream starts 16-byte aligned. */
int pmsz_edit_packed_bytes(const pmsz_params *p, int64_t *bytes);

/* Replaces run_extract_edits (pMSz.cu:2141-2226) with what gather_cost_sparse
 * then does on the host before zstd (pMSz.cu:2469-2483): for each rank block,
 * in core order (x fastest), the delta symbols of extract_edits_kernel
 * (pMSz.cu:2090-2138) packed LSB first at nbits bits (pack_nbits,
 * pMSz.cu:2328-2349), and the clamp values of its symbol-q voxels without the
 * values equal to zero (compact_edits_dense_to_sparse, pMSz.cu:227-250).
 * d_packed: pmsz_edit_packed_bytes bytes, 16-byte aligned; d_values: room for
 * nx*ny*nz values of the dtype; blocks: PX*PY*PZ entries (host, nullable); info
 * nullable.  A symbol wider than nbits bits (g0 is not the field the correction
 * started from) gives PMSZ_EINVAL, as pack_nbits throws out_of_range.  A voxel
 * with g > g0 gets symbol 0; the reference leaves its byte unwritten. */
int pmsz_encode_edits(const pmsz_params *p, const void *d_f, const void *d_g0, const void *d_g,
                      uint8_t *d_packed, void *d_values, pmsz_edit_block *blocks,
                      pmsz_edit_info *info);

/* pmsz_encode_edits without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same inputs,
 * outputs and results as pmsz_encode_edits; the encoder's scratch (block table,
 * look-back words, per-tile counts, totals) is a fixed 4 MiB device array of the
 * library, one per device, held until the process ends.  Limits: at most 256
 * rank blocks, and tiles of 4096 (float) / 2048 (double) voxels whose 12 bytes
 * each fit that array -- about 1.4 G float or 0.7 G double voxels; larger
 * requests return PMSZ_ETOOBIG before touching the device.  Runs on `stream` (a
 * hipStream_t; NULL: the null stream) and returns when the results are there.
 * Calls on one device share the array: calls that overlap on one device race. */
int pmsz_encode_edits_noalloc(const pmsz_params *p, const void *d_f, const void *d_g0,
                              const void *d_g, uint8_t *d_packed, void *d_values,
                              pmsz_edit_block *blocks, pmsz_edit_info *info, void *stream);

/* The GPU compressor front end: an SZ-style Lorenzo codec with dual
 * quantisation (cuSZ, J. Tian et al., PACT 2020), standing in for SZ3 in
 * compressLocalData_coreOnly (pMSz.cu:1536-1624; SZ3 is absent).  Every voxel
 * is quantised on its own, k = round(f * (1 / (2 abs_eb))) in double (0 when
 * the rounded product's magnitude is >= 2^30, or f is NaN), and decompressed to
 * g0 = k * 2 abs_eb in the value type, or to f itself (verbatim) when that value
 * misses |f - g0| <= abs_eb exactly.  The Lorenzo residuals of k over the 7 lower neighbours, core-only
 * per rank block of sub (rank order as for pmsz_encode_edits), are histogrammed
 * per block.  hist (host, nullable): blocks x 1026 counts, residual d in slot
 * d + 512 for |d| < 512, then the escapes and the verbatim voxels.  The
 * estimated compressed size sums over blocks the entropy of the residual
 * symbols (the escapes as one symbol), 32 bits per escape and the value of
 * every verbatim voxel. */
typedef struct {
  int64_t compressed_bytes; /* estimated size of the compressed stream */
  int64_t escapes;          /* residuals with |d| >= 512 */
  int64_t verbatim;         /* voxels stored verbatim */
} pmsz_codec_info;
int pmsz_codec_dq(const pmsz_params *p, const void *d_f, void *d_g0, uint64_t *hist,
                  pmsz_codec_info *info);

/* The compressed stream of the same codec: per rank block, a canonical Huffman
 * code (at most 24 bits) over the Lorenzo residual symbols -- d + 512 for
 * |d| < 512, else an escape symbol with d in an escape list -- built from that
 * block's histogram, plus the verbatim voxels' indices and values.  The stream
 * replaces SZ_compress's byte output (pMSz.cu:1568-1572), whose size gives the
 * reference's CR (pMSz.cu:3135-3144); its layout is given in
 * csrc/pmsz_dqstream.h.  Encoding and decoding run on the device.
 * pmsz_codec_dq_stream_bound: a capacity that holds the stream of any field of
 * these parameters.  pmsz_codec_dq_compress: quantises d_f as pmsz_codec_dq
 * (g0 into d_g0) and writes the stream to d_stream (device, 16-byte aligned,
 * `capacity` bytes); *stream_bytes receives its size (PMSZ_ETOOBIG, with the
 * size needed, when the capacity is short); info as pmsz_codec_dq (nullable).
 * pmsz_codec_dq_stream_params: dtype, extent, sub and abs_eb of a stream (the
 * other fields their defaults).  pmsz_codec_dq_decompress: g0 from a stream
 * (d_g0: nx*ny*nz values of the stream's dtype), bit-exactly the g0 that
 * compression produced; PMSZ_EINVAL for a stream that fails its checks. */
int pmsz_codec_dq_stream_bound(const pmsz_params *p, int64_t *bytes);
int pmsz_codec_dq_compress(const pmsz_params *p, const void *d_f, void *d_g0, uint8_t *d_stream,
                           int64_t capacity, int64_t *stream_bytes, pmsz_codec_info *info);
int pmsz_codec_dq_stream_params(const uint8_t *d_stream, int64_t bytes, pmsz_params *p);
int pmsz_codec_dq_decompress(const uint8_t *d_stream, int64_t bytes, void *d_g0);

/* Average device times of the original labeling pass (avg_seconds[0]) and of
 * detection 1, a full-volume detection sweep on g (avg_seconds[1]), over
 * `iters` launches each, timed with HIP events on the library stream.
 * Measurement aid for the roofline report; it leaves no result. */
int pmsz_time_first_pass(const pmsz_params *p, const void *d_f, const void *d_g, int iters,
                         double *avg_seconds);

/* Test hook (no counterpart in the reference): the step budgets the edits of one
 * field store in the high nibble of a delta counter.  For each i < n (device
 * arrays; dtype PMSZ_F32 or PMSZ_F64 values f[i], g[i], counters d[i] < 16),
 * d_out[i] = the largest B <= q - 1 such that every edit made while the counter
 * is below B is provably the plain step g - abs_eb/q within the bound (clamp_mode
 * as in pmsz_params), or 0.  Needs 2 <= q <= 15. */
int pmsz_test_step_budget(int dtype, int64_t n, const void *d_f, const void *d_g,
                          const uint8_t *d_d, double abs_eb, int q, int clamp_mode,
                          uint8_t *d_out);

const char *pmsz_strerror(int code);
/* Library version string. */
const char *pmsz_version(void);

#ifdef __cplusplus
}
#endif
#endif
