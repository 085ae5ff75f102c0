/*
 * pmsz_oracle.c -- CPU restatement of pMSz's correction.  TEST INFRASTRUCTURE
 * ONLY: loaded by the tests, the smoke check and the benchmark's CPU-baseline
 * leg as the checker; the product never links or calls it.
 *
 * Reference lines restated (file:line in the pMSz sources):
 *   stencil + order ........ pMSz.cu:186-193, 610-620
 *   original labeling ...... pMSz.cu:649-784
 *   detection / targets .... pMSz.cu:786-898
 *   edit primitive ......... pMSz.cu:623-645
 *   loops .................. pMSz.cu:1872-2074 (lazy), sync_pMSz.cu:2102-2200 (sync)
 *   ghost values / sync .... pMSz.cu:320-608, 909-1272
 *   path compression ....... pMSz.cu:255-271
 *   edit symbols ........... pMSz.cu:2090-2138
 *   edit side channel ...... pMSz.cu:223-250, 2328-2356, 2469-2483, 2673-2682
 *   naive-MSz loop ......... naive_MSz.cu:644-850, 1767-1864, 1887-2219
 *   GPU codec stand-in ..... pMSz.cu:1536-1624 (SZ3 absent: cuSZ-style dual-quant Lorenzo)
 *   block split ............ pMSz.cu:2690-2711
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "pmsz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The 14 stencil offsets (dx, dy, dz) in the reference's order (pMSz.cu:186-193):
 * the 6 faces, then the Freudenthal diagonals split along (-1, 1, 1).  Pairs
 * (2k, 2k+1) are opposite directions. */
static const int ORC_DIRS[14][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},   {0, 0, -1},  {-1, 1, 0},
    {1, -1, 0}, {-1, 0, 1}, {1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {-1, 1, 1},  {1, -1, -1}};

/* Missing-neighbour sentinel (INVALID_VAL, pMSz.cu:215), compared in the value
 * type of the field. */
#define ORC_INVALID 1e30

void orc_split(int64_t n, int parts, int coord, int64_t *local, int64_t *offset) {
  int64_t base = n / parts, rem = n % parts;
  *local = base + (coord < rem ? 1 : 0);
  *offset = (int64_t)coord * base + (coord < rem ? coord : rem);
}

/* |a - b| <= eb over the reals.  s = fl(a - b) with its exact rounding error e
 * (Knuth's TwoSum), so a - b = s + e exactly; then decide by cases, using that
 * |s| - eb is exact when eb/2 <= |s| <= 2 eb (Sterbenz). */
int orc_within(double a, double b, double eb) {
  double s = a - b;
  double bb = s - a;
  double e = (a - (s - bb)) + (-b - bb);
  if (s < 0) { s = -s; e = -e; }
  if (s < eb) {
    /* s <= pred(eb), and |e| <= ulp(s)/2 < eb - s */
    return 1;
  }
  if (s > 2 * eb) return 0;
  double r = s - eb; /* exact: eb <= s <= 2 eb */
  return r <= -e;
}

static int orc_ghost_dir_exchanged(int policy, int d) {
  if (policy == ORC_GHOST_FULL) return 1;
  return d < 10; /* faces and the four edges of pMSz.cu:394-488 */
}

static int orc_dir_of(int64_t PX, int64_t PY, int64_t i, int64_t j) {
  if (i == j) return 14;
  for (int d = 0; d < 14; d++) {
    int64_t off = ORC_DIRS[d][0] + PX * (ORC_DIRS[d][1] + PY * ORC_DIRS[d][2]);
    if (i + off == j) return d;
  }
  return 15; /* unreachable for a valid pointer */
}

static int64_t orc_neighbor_global(int64_t nx, int64_t ny, int64_t i, int d) {
  if (d >= 14) return i;
  return i + ORC_DIRS[d][0] + nx * (ORC_DIRS[d][1] + ny * ORC_DIRS[d][2]);
}

/* Root of every pointer chain (PathCompression, pMSz.cu:255-271).  The reference
 * chases each chain independently; walking a chain and then rewriting it gives
 * the same roots. */
static void orc_compress_pointers(int64_t *lab, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    int64_t r = i;
    while (lab[r] != r) r = lab[r];
    int64_t c = i;
    while (lab[c] != r) {
      int64_t nx = lab[c];
      lab[c] = r;
      c = nx;
    }
  }
}

/* Steps of a naive-MSz path walk before it is abandoned (a walk that follows
 * f's steepest path never revisits a voxel; the cap only bounds a malformed one). */
#define ORC_NAIVE_WALK_CAP (1 << 20)

/* f's and g's pointers as or_types / de_direction hold them (naive_MSz.cu:740-760,
 * 830-842): slot 0 of f is -1 at a maximum, -2 at a minimum, else the descent;
 * slot 1 is the descent at a maximum, else the ascent. */
static int64_t orc_naive_f0(const orc_params *p, const uint8_t *cf, int64_t c) {
  if ((cf[c] & 15) == 14) return -1;
  if ((cf[c] >> 4) == 14) return -2;
  return orc_neighbor_global(p->nx, p->ny, c, cf[c] >> 4);
}
static int64_t orc_naive_f1(const orc_params *p, const uint8_t *cf, int64_t c) {
  if ((cf[c] & 15) == 14) return orc_neighbor_global(p->nx, p->ny, c, cf[c] >> 4);
  return orc_neighbor_global(p->nx, p->ny, c, cf[c] & 15);
}

static float next_up_f32(float c) { return nextafterf(c, INFINITY); }
static double next_up_f64(double c) { return nextafter(c, INFINITY); }
static float next_down_f32(float c) { return nextafterf(c, -INFINITY); }
static double next_down_f64(double c) { return nextafter(c, -INFINITY); }
static float round_half_away_f32(float r) { return roundf(r); }
static double round_half_away_f64(double r) { return round(r); }

/* ---- the compressed stream of the dual-quantisation codec (untyped parts) ---- */
#define ORC_DQS_SYMBOLS (ORC_DQ_BINS + 1)
#define ORC_DQS_ESCAPE ORC_DQ_BINS
#define ORC_DQS_CHUNK 2048
#define ORC_DQS_MAXLEN 24
#define ORC_DQS_MAGIC 0x31514450u

/* Huffman code lengths for counts[0..ns) (0: unused), at most ORC_DQS_MAXLEN:
 * leaves in (count, symbol) order, two queues with the leaf taken on equal
 * weights, depths as lengths; longer than the limit: every count halved
 * (rounded up) and again.  One used symbol: length 1. */
static void orc_dqs_lengths(const uint64_t *counts, int ns, uint8_t *len) {
  uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)ns);
  int *sym = (int *)malloc(sizeof(int) * (size_t)ns);
  uint64_t *w = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(2 * ns));
  int *par = (int *)malloc(sizeof(int) * (size_t)(2 * ns));
  int *dep = (int *)malloc(sizeof(int) * (size_t)(2 * ns));
  memcpy(cnt, counts, sizeof(uint64_t) * (size_t)ns);
  for (;;) {
    memset(len, 0, (size_t)ns);
    int m = 0;
    for (int s = 0; s < ns; s++)
      if (cnt[s]) sym[m++] = s;
    if (m == 0) break;
    if (m == 1) { len[sym[0]] = 1; break; }
    /* stable insertion sort by count (symbols start in increasing order) */
    for (int a = 1; a < m; a++) {
      int v = sym[a], b = a - 1;
      while (b >= 0 && cnt[sym[b]] > cnt[v]) { sym[b + 1] = sym[b]; b--; }
      sym[b + 1] = v;
    }
    for (int k = 0; k < m; k++) w[k] = cnt[sym[k]];
    int i = 0, j = m;
    for (int k = m; k < 2 * m - 1; k++) {
      int pick[2];
      for (int t = 0; t < 2; t++) {
        if (i < m && (j >= k || w[i] <= w[j])) pick[t] = i++;
        else pick[t] = j++;
      }
      w[k] = w[pick[0]] + w[pick[1]];
      par[pick[0]] = par[pick[1]] = k;
    }
    dep[2 * m - 2] = 0;
    int maxd = 0;
    for (int k = 2 * m - 3; k >= 0; k--) {
      dep[k] = dep[par[k]] + 1;
      if (k < m && dep[k] > maxd) maxd = dep[k];
    }
    if (maxd <= ORC_DQS_MAXLEN) {
      for (int k = 0; k < m; k++) len[sym[k]] = (uint8_t)dep[k];
      break;
    }
    for (int s = 0; s < ns; s++)
      if (cnt[s]) cnt[s] = (cnt[s] + 1) / 2;
  }
  free(cnt); free(sym); free(w); free(par); free(dep);
}

/* Canonical codes (DEFLATE's assignment) of the lengths. */
static void orc_dqs_codes(const uint8_t *len, int ns, uint32_t *code) {
  uint32_t bl[ORC_DQS_MAXLEN + 2] = {0}, next[ORC_DQS_MAXLEN + 2] = {0};
  for (int s = 0; s < ns; s++) if (len[s]) bl[len[s]]++;
  uint32_t c = 0;
  for (int b = 1; b <= ORC_DQS_MAXLEN; b++) { c = (c + bl[b - 1]) << 1; next[b] = c; }
  for (int s = 0; s < ns; s++) code[s] = len[s] ? next[len[s]]++ : 0;
}

/* LSB-first bit writer into 32-bit words; a code goes out first bit first. */
typedef struct { uint32_t *w; uint64_t pos; } orc_bitw;
static void orc_put_code(orc_bitw *bw, uint32_t code, int len) {
  for (int k = len - 1; k >= 0; k--) {
    if ((code >> k) & 1u) bw->w[bw->pos >> 5] |= 1u << (bw->pos & 31);
    bw->pos++;
  }
}

static uint64_t orc_up8(uint64_t x) { return (x + 7) & ~(uint64_t)7; }

#define T float
#define SFX(x) x##_f32
#include "oracle_impl.inc"
#undef T
#undef SFX

#define T double
#define SFX(x) x##_f64
#include "oracle_impl.inc"
#undef T
#undef SFX

static int orc_check(const orc_params *p) {
  if (!p) return ORC_EINVAL;
  if (p->dtype != ORC_F32 && p->dtype != ORC_F64) return ORC_EINVAL;
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return ORC_EINVAL;
  if (p->nx * p->ny * p->nz < 2) return ORC_EINVAL;
  if (!(p->abs_eb > 0) || isinf(p->abs_eb)) return ORC_EINVAL;
  if (p->q < 1 || p->q > 255) return ORC_EINVAL;
  for (int a = 0; a < 3; a++)
    if (p->sub[a] < 1) return ORC_EINVAL;
  if (p->sub[0] > p->nx || p->sub[1] > p->ny || p->sub[2] > p->nz) return ORC_EINVAL;
  if (p->max_sweeps < 1) return ORC_EINVAL;
  return ORC_OK;
}

int orc_codes(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *field,
              uint8_t *codes) {
  if (nx < 1 || ny < 1 || nz < 1) return ORC_EINVAL;
  if (dtype == ORC_F32) return codes_f32(nx, ny, nz, (const float *)field, codes);
  if (dtype == ORC_F64) return codes_f64(nx, ny, nz, (const double *)field, codes);
  return ORC_EINVAL;
}

int orc_labels(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *field,
               int64_t *asc, int64_t *desc) {
  if (nx < 1 || ny < 1 || nz < 1) return ORC_EINVAL;
  if (dtype == ORC_F32) return labels_f32(nx, ny, nz, (const float *)field, asc, desc);
  if (dtype == ORC_F64) return labels_f64(nx, ny, nz, (const double *)field, asc, desc);
  return ORC_EINVAL;
}

int orc_sweep(const orc_params *p, const void *f, void *g, uint8_t *dc, int edit,
              uint32_t counts[3], int64_t *ntargets, uint8_t *target_mask) {
  int rc = orc_check(p);
  if (rc) return rc;
  if (p->dtype == ORC_F32)
    return sweep_single_f32(p, (const float *)f, (float *)g, dc, edit, counts, ntargets,
                            target_mask);
  return sweep_single_f64(p, (const double *)f, (double *)g, dc, edit, counts, ntargets,
                          target_mask);
}

int orc_correct(const orc_params *p, const void *f, void *g, uint8_t *dc, orc_stats *st,
                orc_trace *trace) {
  int rc = orc_check(p);
  if (rc) return rc;
  orc_stats tmp;
  if (!st) st = &tmp;
  if (trace) trace->n = 0;
  if (p->mode == ORC_NAIVE) {
    if (p->sub[0] * p->sub[1] * p->sub[2] != 1) return ORC_EINVAL;
    if (p->dtype == ORC_F32)
      return correct_naive_f32(p, (const float *)f, (float *)g, dc, st, trace);
    return correct_naive_f64(p, (const double *)f, (double *)g, dc, st, trace);
  }
  if (p->dtype == ORC_F32) {
    rc = correct_f32(p, (const float *)f, (float *)g, dc, st, trace);
  } else {
    rc = correct_f64(p, (const double *)f, (double *)g, dc, st, trace);
  }
  return rc;
}

int orc_edit_symbols(const orc_params *p, const void *f, const void *g0, const void *g,
                     uint8_t *symbols, int64_t *edited) {
  int rc = orc_check(p);
  if (rc) return rc;
  if (p->dtype == ORC_F32)
    edit_symbols_f32(p, (const float *)f, (const float *)g0, (const float *)g, symbols,
                     edited);
  else
    edit_symbols_f64(p, (const double *)f, (const double *)g0, (const double *)g, symbols,
                     edited);
  return ORC_OK;
}

unsigned orc_bits_for_max(unsigned q) {
  /* ceil(log2(q + 1)), 1 for q = 0 (pMSz.cu:2352-2356), counted in integers */
  unsigned b = 0;
  while (b < 32 && (1ull << b) <= q) b++;
  return b ? b : 1;
}

int orc_encode_edits(const orc_params *p, const void *f, const void *g0, const void *g,
                     uint8_t *packed, void *vals, orc_edit_block *blocks, int64_t *edited) {
  int rc = orc_check(p);
  if (rc) return rc;
  const unsigned nbits = orc_bits_for_max((unsigned)p->q);
  const int px = p->sub[0], py = p->sub[1], pz = p->sub[2];
  const size_t es = p->dtype == ORC_F32 ? 4 : 8;
  int64_t poff = 0, voff = 0, ne = 0;
  int bad = 0;
  /* MPI_Cart_create(..., reorder = 0) over dims {PX, PY, PZ} (pMSz.cu:2673-2682)
   * numbers the ranks row-major: rank = (cx * PY + cy) * PZ + cz. */
  for (int r = 0; r < px * py * pz; r++) {
    int cx = r / (py * pz), cy = (r / pz) % py, cz = r % pz;
    int64_t l[3], o[3], nv = 0;
    orc_split(p->nx, px, cx, &l[0], &o[0]);
    orc_split(p->ny, py, cy, &l[1], &o[1]);
    orc_split(p->nz, pz, cz, &l[2], &o[2]);
    char *vp = (char *)vals + voff * es;
    if (p->dtype == ORC_F32)
      bad |= encode_block_f32(p, (const float *)f, (const float *)g0, (const float *)g, l, o,
                              nbits, packed + poff, (float *)vp, &nv, &ne);
    else
      bad |= encode_block_f64(p, (const double *)f, (const double *)g0, (const double *)g, l, o,
                              nbits, packed + poff, (double *)vp, &nv, &ne);
    int64_t vox = l[0] * l[1] * l[2], pb = (vox * nbits + 7) / 8;
    if (blocks) {
      blocks[r].voxels = vox;
      blocks[r].packed_offset = poff;
      blocks[r].packed_bytes = pb;
      blocks[r].value_offset = voff;
      blocks[r].values = nv;
    }
    poff += (pb + 15) & ~(int64_t)15;
    voff += nv;
  }
  if (edited) *edited = ne;
  return bad ? ORC_EINVAL : ORC_OK;
}

int orc_codec_dq(const orc_params *p, const void *f, void *g0, uint64_t *hist) {
  int rc = orc_check(p);
  if (rc) return rc;
  if (p->dtype == ORC_F32) return codec_dq_f32(p, (const float *)f, (float *)g0, hist);
  return codec_dq_f64(p, (const double *)f, (double *)g0, hist);
}

int orc_blocks(const orc_params *p, const void *f, const void *g, void *fo, void *go) {
  int rc = orc_check(p);
  if (rc) return rc;
  if (p->dtype == ORC_F32)
    return blocks_hook_f32(p, (const float *)f, (const float *)g, (float *)fo, (float *)go);
  return blocks_hook_f64(p, (const double *)f, (const double *)g, (double *)fo, (double *)go);
}

int orc_sync_blocks(const orc_params *p, void *fb, void *gb, uint8_t *dcb, const int *send) {
  int rc = orc_check(p);
  if (rc) return rc;
  if (p->dtype == ORC_F32) return sync_hook_f32(p, (float *)fb, (float *)gb, dcb, send);
  return sync_hook_f64(p, (double *)fb, (double *)gb, dcb, send);
}

int orc_dq_compress(const orc_params *p, const void *f, void *g0, uint8_t *stream,
                    int64_t capacity, int64_t *bytes) {
  int rc = orc_check(p);
  if (rc) return rc;
  if (p->dtype == ORC_F32)
    return dq_compress_f32(p, (const float *)f, (float *)g0, stream, capacity, bytes);
  return dq_compress_f64(p, (const double *)f, (double *)g0, stream, capacity, bytes);
}

int orc_dq_decompress(const uint8_t *stream, int64_t bytes, void *g0) {
  if (!stream || bytes < 64) return ORC_EINVAL;
  uint32_t dtype;
  memcpy(&dtype, stream + 4, 4);
  if (dtype == ORC_F32) return dq_decompress_f32(stream, bytes, (float *)g0);
  if (dtype == ORC_F64) return dq_decompress_f64(stream, bytes, (double *)g0);
  return ORC_EINVAL;
}
