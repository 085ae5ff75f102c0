/*
 * pmsz_oracle.h -- CPU restatement of pMSz's extrema + Morse-Smale segmentation
 * correction.  TEST INFRASTRUCTURE ONLY: the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark load this library as the checker.  The product
 * (libpmsz.so) never links, loads or calls it.
 *
 * Every function restates the reference algorithm under *snapshot semantics*:
 * within one sweep every classification reads the start-of-sweep field and every
 * target receives exactly one edit, evaluated on its start-of-sweep (f, g, dc).
 * The reference kernel (pMSz.cu:786-898) reads neighbours while other threads of
 * the same launch lower them, so its own output is schedule dependent; the
 * snapshot order is the deterministic restatement both this oracle and the HIP
 * kernels implement.  Its per-sweep edit dedup mirrors the atomicCAS on edits[]
 * at pMSz.cu:795.
 *
 * Layout: every subdomain is held as the reference holds one rank's block, a
 * padded (lx+2)(ly+2)(lz+2) array whose 1-voxel halo holds ghost values or the
 * missing-neighbour sentinel 1e30 (pMSz.cu:215, 2709-2711).  One subdomain
 * (sub = 1,1,1) is the single-rank case of the reference.
 */
#ifndef PMSZ_ORACLE_H
#define PMSZ_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_F32 = 0, ORC_F64 = 1 };
enum { ORC_LAZY = 0, ORC_SYNC = 1, ORC_NAIVE = 2 };
enum { ORC_CLAMP_REFERENCE = 0, ORC_CLAMP_BOUND_SAFE = 1 };
enum { ORC_GHOST_REFERENCE = 0, ORC_GHOST_FULL = 1 };

enum {
  ORC_OK = 0,
  ORC_EINVAL = -1,
  ORC_ENOMEM = -2,
  ORC_NOT_CONVERGED = -3,
};

typedef struct {
  int dtype;            /* ORC_F32 / ORC_F64 */
  int64_t nx, ny, nz;   /* x fastest */
  double abs_eb;        /* absolute error bound (pMSz.cu:2628) */
  int q;                /* decrements before clamping, 5 in the reference (pMSz.cu:2630) */
  int sub[3];           /* subdomain grid PX PY PZ (pMSz.cu:2631-2633) */
  int mode;             /* ORC_LAZY (pMSz.cu), ORC_SYNC (sync_pMSz.cu) or ORC_NAIVE
                           (naive_MSz.cu, one domain) */
  int max_sweeps;       /* cap on edit sweeps; the reference has none */
  int clamp_mode;       /* ORC_CLAMP_* */
  int ghost_policy;     /* ORC_GHOST_* */
} orc_params;

typedef struct {
  int64_t edit_sweeps;      /* dense detect+edit sweeps (summed over subdomains' max) */
  int64_t check_sweeps;     /* checkFinal sweeps */
  int64_t outer_iters;      /* outer iterations (ghost syncs) */
  int64_t ite;              /* the reference's printed "iteration is:" counter (max over ranks) */
  int64_t edited_voxels;    /* h_count of extract_edits_kernel (pMSz.cu:2116-2119) */
  int64_t clamped_voxels;   /* voxels whose final dc == q */
  int64_t total_targets;    /* sum over sweeps of distinct targets edited */
  uint32_t final_cnt[3];    /* counters of the last check sweep: max, min, dir */
  int status;
} orc_stats;

/* Per-sweep trace (optional): the counters and the number of distinct targets of
 * every edit sweep, in execution order (subdomain-major within an outer iteration). */
typedef struct {
  int64_t cap;              /* capacity of the arrays below */
  int64_t n;                /* entries written */
  uint32_t *cnt;            /* 3 per entry */
  int64_t *targets;         /* 1 per entry */
} orc_trace;

/* Original classification codes of a whole field: for every voxel the stencil
 * direction (0..13) of its steepest-ascent neighbour in the low nibble and of its
 * steepest-descent neighbour in the high nibble, 14 meaning the voxel itself
 * (a maximum / minimum).  Restates iscriticle data_type=1, pMSz.cu:649-784. */
int orc_codes(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *field,
              uint8_t *codes);

/* Path-compressed ascending / descending Morse-Smale labels (global linear index of
 * the extremum each voxel flows to).  Restates PathCompression, pMSz.cu:255-271,
 * over the pointers of iscriticle(preserveMSS=1), pMSz.cu:739-760. */
int orc_labels(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *field,
               int64_t *asc, int64_t *desc);

/* One snapshot sweep on a single domain: classify g, count, and (edit=1) apply
 * one edit to every target.  counts[3] = max, min, dir; *ntargets = distinct
 * targets; target_mask (nullable, N bytes) marks them.  dc is updated in place. */
int orc_sweep(const orc_params *p, const void *f, void *g, uint8_t *dc, int edit,
              uint32_t counts[3], int64_t *ntargets, uint8_t *target_mask);

/* The full correction: g holds the decompressed field on entry and the corrected
 * field on exit; dc (nullable) receives the final per-voxel delta counters.
 * Restates c_loops, pMSz.cu:1872-2074 (lazy) and sync_pMSz.cu:2102-2200 (sync). */
int orc_correct(const orc_params *p, const void *f, void *g, uint8_t *dc,
                orc_stats *st, orc_trace *trace);

/* Per-voxel delta symbols and the edited-voxel count of a correction
 * (extract_edits_kernel, pMSz.cu:2090-2138). */
int orc_edit_symbols(const orc_params *p, const void *f, const void *g0,
                     const void *g, uint8_t *symbols, int64_t *edited);

/* One rank block of the edit side channel: the same fields, layout and rank
 * order as pmsz_edit_block / pmsz_encode_edits of the product. */
typedef struct {
  int64_t voxels, packed_offset, packed_bytes, value_offset, values;
} orc_edit_block;

/* bits_for_max (pMSz.cu:2352-2356). */
unsigned orc_bits_for_max(unsigned q);

/* The edit side channel per rank block, in rank order: packed symbol streams
 * (each starting 16-byte aligned in packed) and the nonzero clamp values
 * (pMSz.cu:227-250, 2090-2138, 2328-2356).  packed must hold the padded total,
 * vals nx*ny*nz values.  ORC_EINVAL when a symbol exceeds bits_for_max(q) bits. */
int orc_encode_edits(const orc_params *p, const void *f, const void *g0, const void *g,
                     uint8_t *packed, void *vals, orc_edit_block *blocks, int64_t *edited);

/* The dual-quantisation Lorenzo codec (the GPU front end's stand-in for SZ3,
 * pMSz.cu:1536-1624): g0 and, per rank block, ORC_DQ_BINS residual bins
 * (residual d in bin d + ORC_DQ_RADIUS), the escapes and the verbatim voxels. */
#define ORC_DQ_RADIUS 512
#define ORC_DQ_BINS (2 * ORC_DQ_RADIUS)
int orc_codec_dq(const orc_params *p, const void *f, void *g0, uint64_t *hist);

/* The compressed stream of that codec (the layout of the product's
 * csrc/pmsz_dqstream.h): orc_dq_compress writes it to stream (host, capacity
 * bytes; *bytes its size, ORC_EINVAL when short) and g0 as orc_codec_dq;
 * orc_dq_decompress gives g0 back from it. */
int orc_dq_compress(const orc_params *p, const void *f, void *g0, uint8_t *stream,
                    int64_t capacity, int64_t *bytes);
int orc_dq_decompress(const uint8_t *stream, int64_t bytes, void *g0);

/* Balanced block split of one dimension (pMSz.cu:2690-2706). */
void orc_split(int64_t n, int parts, int coord, int64_t *local, int64_t *offset);

/* The exact bound predicate |a - b| <= eb over the reals (a, b, eb doubles). */
int orc_within(double a, double b, double eb);

/* The clamp value of the edit primitive for original value fv in float32
 * (orc_clamp_f32; fv and eb rounded to float for the reference value) or
 * float64 (orc_clamp_f64), in clamp mode ORC_CLAMP_REFERENCE / BOUND_SAFE. */
double orc_clamp_f32(double fv, double eb, int clamp_mode);
double orc_clamp_f64(double fv, double eb, int clamp_mode);

/* Test hooks of the subdomain path.  orc_blocks: the padded blocks of every
 * subdomain after set-up (core values and the initial ghost values of the ghost
 * policy, exchange_ghost_layers, pMSz.cu:320-608), f into fo and g into go,
 * blocks back to back in the order cx fastest, then cy, then cz, each
 * (lz+2)(ly+2)(lx+2) values.  orc_sync_blocks: one delta-counter ghost sync
 * (pMSz.cu:909-1272) on padded blocks in that layout: gb and dcb are updated in
 * place; send[k] is block k's send_flag. */
int orc_blocks(const orc_params *p, const void *f, const void *g, void *fo, void *go);
int orc_sync_blocks(const orc_params *p, void *fb, void *gb, uint8_t *dcb, const int *send);

#ifdef __cplusplus
}
#endif
#endif
