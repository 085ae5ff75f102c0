// fieldio.cpp -- host-side front end of the correction path: synthetic Perlin
// fields, the error-bounded codecs that produce the decompressed field g0, and
// raw little-endian field I/O.  These stand where the reference's data reader
// (read_subblock, pMSz.cu:274-313) and compressor adapters
// (compressLocalData_coreOnly / compressLocalDataZFP_coreOnly, pMSz.cu:1436-1624)
// stand; the correction itself lives in libpmsz.so.
//
// * Perlin: improved gradient noise (K. Perlin, "Improving Noise", SIGGRAPH 2002)
//   summed over octaves (fractional Brownian motion), evaluated in double at voxel
//   centres and rounded once to the field's value type.  The permutation is a
//   Fisher-Yates shuffle of 0..255 driven by splitmix64(seed), duplicated to 512
//   entries, the structure of generate_perlin.py:105-108.
// * SZ-like codec: SZ3 is not available here, so the "sz3" compressor id maps to
//   an SZ-style Lorenzo predictor with linear quantisation into 2*eb bins; every
//   reconstructed value is checked with the exact predicate |f - g0| <= eb and
//   stored verbatim if it fails, so g0 always honours the bound.
// * ZFP: fixed-accuracy mode of the system libzfp (0.5.5 in this image), loaded
//   at run time; absent library => PMSZ_IO_EUNAVAILABLE, never a fake result.
// * Edit files: zstd frames of the packed symbols and clamp values, written and
//   read as gather_cost_sparse / validateResult do (pMSz.cu:2429-2590), with the
//   system libzstd loaded at run time the same way.
#include "../../include/pmsz_io.h"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

namespace {

uint64_t splitmix64(uint64_t &state) {
  uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Perm {
  int p[512];
  explicit Perm(uint64_t seed) {
    int base[256];
    for (int i = 0; i < 256; i++) base[i] = i;
    uint64_t st = seed;
    for (int i = 255; i > 0; i--) {
      int j = (int)(splitmix64(st) % (uint64_t)(i + 1));
      int t = base[i];
      base[i] = base[j];
      base[j] = t;
    }
    for (int i = 0; i < 512; i++) p[i] = base[i & 255];
  }
};

inline double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
inline double lerp(double t, double a, double b) { return a + t * (b - a); }
inline double grad(int h, double x, double y, double z) {
  h &= 15;
  double u = h < 8 ? x : y;
  double v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
  return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}

double noise3(const Perm &P, double x, double y, double z) {
  double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  int X = (int)((int64_t)fx & 255), Y = (int)((int64_t)fy & 255), Z = (int)((int64_t)fz & 255);
  x -= fx;
  y -= fy;
  z -= fz;
  double u = fade(x), v = fade(y), w = fade(z);
  const int *p = P.p;
  int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
  int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
  return lerp(w,
              lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                   lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
              lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                   lerp(u, grad(p[AB + 1], x, y - 1, z - 1),
                        grad(p[BB + 1], x - 1, y - 1, z - 1))));
}

// |a - b| <= eb over the reals (TwoSum error term + Sterbenz), the same predicate
// the correction kernels use in bound-safe mode.
bool within_exact(double a, double b, double eb) {
  double s = a - b;
  double bb = s - a;
  double e = (a - (s - bb)) + (-b - bb);
  if (s < 0) {
    s = -s;
    e = -e;
  }
  if (s < eb) return true;
  if (s > 2 * eb) return false;
  double r = s - eb;
  return r <= -e;
}

template <typename T>
void perlin_fill(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, double freq, int octaves,
                 double persistence, T *out) {
  Perm P(seed);
  double norm = 0, amp = 1;
  for (int o = 0; o < octaves; o++) {
    norm += amp;
    amp *= persistence;
  }
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < nz; z++)
    for (int64_t y = 0; y < ny; y++)
      for (int64_t x = 0; x < nx; x++) {
        double s = 0, a = 1, fr = freq;
        for (int o = 0; o < octaves; o++) {
          s += a * noise3(P, fr * (x + 0.5) / nx, fr * (y + 0.5) / ny, fr * (z + 0.5) / nz);
          a *= persistence;
          fr *= 2.0;
        }
        out[x + nx * (y + ny * z)] = (T)(s / norm);
      }
}

// SZ-style Lorenzo + linear quantisation over reconstructed values.
template <typename T>
void szlike(int64_t nx, int64_t ny, int64_t nz, const T *f, double eb, T *g,
            int64_t *unpred, double *est_bytes) {
  const int64_t R = 32768;  // quantisation radius (SZ's default bin count 65536)
  const double bin = 2.0 * eb;
  std::vector<uint32_t> hist(2 * R + 1, 0);
  int64_t nun = 0;
  auto at = [&](int64_t x, int64_t y, int64_t z) -> double {
    if (x < 0 || y < 0 || z < 0) return 0.0;
    return (double)g[x + nx * (y + ny * z)];
  };
  for (int64_t z = 0; z < nz; z++)
    for (int64_t y = 0; y < ny; y++)
      for (int64_t x = 0; x < nx; x++) {
        double pr = at(x - 1, y, z) + at(x, y - 1, z) + at(x, y, z - 1) - at(x - 1, y - 1, z) -
                    at(x - 1, y, z - 1) - at(x, y - 1, z - 1) + at(x - 1, y - 1, z - 1);
        int64_t i = x + nx * (y + ny * z);
        double fv = (double)f[i];
        double k = std::floor((fv - pr) / bin + 0.5);
        bool ok = false;
        if (std::fabs(k) < (double)R) {
          T r = (T)(pr + bin * k);
          if (within_exact(fv, (double)r, eb)) {
            g[i] = r;
            hist[(size_t)((int64_t)k + R)]++;
            ok = true;
          }
        }
        if (!ok) {
          g[i] = f[i];
          hist[0]++;
          nun++;
        }
      }
  double n = (double)(nx * ny * nz), bits = 0;
  for (uint32_t c : hist)
    if (c) bits -= c * std::log2(c / n);
  *unpred = nun;
  *est_bytes = bits / 8.0 + (double)nun * sizeof(T);
}

// ---- ZFP via the system library, resolved at run time ----------------------
struct ZfpApi {
  void *h = nullptr;
  void *(*field_3d)(void *, int, unsigned, unsigned, unsigned) = nullptr;
  void (*field_free)(void *) = nullptr;
  void *(*stream_open)(void *) = nullptr;
  void (*stream_close)(void *) = nullptr;
  double (*set_accuracy)(void *, double) = nullptr;
  size_t (*max_size)(const void *, const void *) = nullptr;
  void (*set_bit_stream)(void *, void *) = nullptr;
  void (*rewind)(void *) = nullptr;
  size_t (*compress)(void *, const void *) = nullptr;
  size_t (*decompress)(void *, void *) = nullptr;
  void *(*bs_open)(void *, size_t) = nullptr;
  void (*bs_close)(void *) = nullptr;
  bool ok = false;
  ZfpApi() {
    const char *names[] = {"libzfp.so.0", "libzfp.so", "/opt/conda/lib/libzfp.so.0"};
    for (const char *n : names) {
      h = dlopen(n, RTLD_NOW | RTLD_LOCAL);
      if (h) break;
    }
    if (!h) return;
#define PMSZ_SYM(dst, name) *(void **)(&dst) = dlsym(h, name)
    PMSZ_SYM(field_3d, "zfp_field_3d");
    PMSZ_SYM(field_free, "zfp_field_free");
    PMSZ_SYM(stream_open, "zfp_stream_open");
    PMSZ_SYM(stream_close, "zfp_stream_close");
    PMSZ_SYM(set_accuracy, "zfp_stream_set_accuracy");
    PMSZ_SYM(max_size, "zfp_stream_maximum_size");
    PMSZ_SYM(set_bit_stream, "zfp_stream_set_bit_stream");
    PMSZ_SYM(rewind, "zfp_stream_rewind");
    PMSZ_SYM(compress, "zfp_compress");
    PMSZ_SYM(decompress, "zfp_decompress");
    PMSZ_SYM(bs_open, "stream_open");
    PMSZ_SYM(bs_close, "stream_close");
#undef PMSZ_SYM
    ok = field_3d && field_free && stream_open && stream_close && set_accuracy && max_size &&
         set_bit_stream && rewind && compress && decompress && bs_open && bs_close;
  }
};

ZfpApi &zfp_api() {
  static ZfpApi api;
  return api;
}

// The zstd entry points the edit files need (zstd.h declares them with these
// C signatures); the image ships the library without its header.
struct ZstdApi {
  void *h = nullptr;
  size_t (*compress)(void *, size_t, const void *, size_t, int) = nullptr;
  size_t (*bound)(size_t) = nullptr;
  size_t (*decompress)(void *, size_t, const void *, size_t) = nullptr;
  unsigned (*is_error)(size_t) = nullptr;
  unsigned (*version)(void) = nullptr;
  bool ok = false;
  ZstdApi() {
    const char *names[] = {"libzstd.so.1", "libzstd.so"};
    for (const char *n : names) {
      h = dlopen(n, RTLD_NOW | RTLD_LOCAL);
      if (h) break;
    }
    if (!h) return;
#define PMSZ_SYM(dst, name) *(void **)(&dst) = dlsym(h, name)
    PMSZ_SYM(compress, "ZSTD_compress");
    PMSZ_SYM(bound, "ZSTD_compressBound");
    PMSZ_SYM(decompress, "ZSTD_decompress");
    PMSZ_SYM(is_error, "ZSTD_isError");
    PMSZ_SYM(version, "ZSTD_versionNumber");
#undef PMSZ_SYM
    ok = compress && bound && decompress && is_error && version;
  }
};

ZstdApi &zstd_api() {
  static ZstdApi api;
  return api;
}

// One zstd frame of n bytes appended to fp; nothing for n = 0.
// One zstd frame of n bytes into out (empty for n = 0).
int zstd_frame(const void *src, size_t n, int level, std::vector<uint8_t> &out) {
  out.clear();
  if (n == 0) return PMSZ_IO_OK;
  ZstdApi &z = zstd_api();
  out.resize(z.bound(n));
  size_t c = z.compress(out.data(), out.size(), src, n, level);
  if (z.is_error(c)) return PMSZ_IO_EFAIL;
  out.resize(c);
  return PMSZ_IO_OK;
}

bool read_file(const std::string &path, std::vector<uint8_t> &buf) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) return false;
  buf.clear();
  uint8_t tmp[1 << 16];
  size_t k;
  while ((k = std::fread(tmp, 1, sizeof(tmp), fp)) > 0) buf.insert(buf.end(), tmp, tmp + k);
  std::fclose(fp);
  return true;
}

unsigned bits_for_max(unsigned q) {
  unsigned b = 0;
  while (b < 32 && (1ull << b) <= q) b++;
  return b ? b : 1;
}

}  // namespace

extern "C" {

int pmsz_io_perlin(int dtype, int64_t nx, int64_t ny, int64_t nz, uint64_t seed,
                   double base_freq, int octaves, double persistence, void *out) {
  if (nx < 1 || ny < 1 || nz < 1 || octaves < 1 || !out) return PMSZ_IO_EINVAL;
  if (dtype == 0)
    perlin_fill(nx, ny, nz, seed, base_freq, octaves, persistence, (float *)out);
  else if (dtype == 1)
    perlin_fill(nx, ny, nz, seed, base_freq, octaves, persistence, (double *)out);
  else
    return PMSZ_IO_EINVAL;
  return PMSZ_IO_OK;
}

int pmsz_io_range(int dtype, int64_t n, const void *data, double *vmin, double *vmax) {
  if (n < 1 || !data) return PMSZ_IO_EINVAL;
  double lo = INFINITY, hi = -INFINITY;
  if (dtype == 0) {
    const float *a = (const float *)data;
    for (int64_t i = 0; i < n; i++) {
      lo = a[i] < lo ? a[i] : lo;
      hi = a[i] > hi ? a[i] : hi;
    }
  } else if (dtype == 1) {
    const double *a = (const double *)data;
    for (int64_t i = 0; i < n; i++) {
      lo = a[i] < lo ? a[i] : lo;
      hi = a[i] > hi ? a[i] : hi;
    }
  } else {
    return PMSZ_IO_EINVAL;
  }
  *vmin = lo;
  *vmax = hi;
  return PMSZ_IO_OK;
}

int pmsz_io_szlike(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *f, double abs_eb,
                   void *g0, int64_t *unpredictable, double *est_bytes) {
  if (nx < 1 || ny < 1 || nz < 1 || !f || !g0 || !(abs_eb > 0)) return PMSZ_IO_EINVAL;
  int64_t un = 0;
  double eb = 0;
  if (dtype == 0)
    szlike(nx, ny, nz, (const float *)f, abs_eb, (float *)g0, &un, &eb);
  else if (dtype == 1)
    szlike(nx, ny, nz, (const double *)f, abs_eb, (double *)g0, &un, &eb);
  else
    return PMSZ_IO_EINVAL;
  if (unpredictable) *unpredictable = un;
  if (est_bytes) *est_bytes = eb;
  return PMSZ_IO_OK;
}

int pmsz_io_zfp_available(void) { return zfp_api().ok ? 1 : 0; }

int pmsz_io_zfp(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *f, double tolerance,
                void *g0, int64_t *compressed_bytes) {
  if (nx < 1 || ny < 1 || nz < 1 || !f || !g0 || !(tolerance > 0)) return PMSZ_IO_EINVAL;
  if (dtype != 0 && dtype != 1) return PMSZ_IO_EINVAL;
  if (nx > 0xffffffffLL || ny > 0xffffffffLL || nz > 0xffffffffLL) return PMSZ_IO_EINVAL;
  ZfpApi &z = zfp_api();
  if (!z.ok) return PMSZ_IO_EUNAVAILABLE;
  const int ztype = dtype == 0 ? 3 : 4;  // zfp_type_float / zfp_type_double
  void *field = z.field_3d(const_cast<void *>(f), ztype, (unsigned)nx, (unsigned)ny, (unsigned)nz);
  void *zs = z.stream_open(nullptr);
  z.set_accuracy(zs, tolerance);
  size_t cap = z.max_size(zs, field);
  void *buf = std::malloc(cap);
  if (!buf) {
    z.stream_close(zs);
    z.field_free(field);
    return PMSZ_IO_ENOMEM;
  }
  void *bs = z.bs_open(buf, cap);
  z.set_bit_stream(zs, bs);
  z.rewind(zs);
  size_t nbytes = z.compress(zs, field);
  int rc = PMSZ_IO_OK;
  if (nbytes == 0) rc = PMSZ_IO_EFAIL;
  if (rc == PMSZ_IO_OK) {
    void *ofield = z.field_3d(g0, ztype, (unsigned)nx, (unsigned)ny, (unsigned)nz);
    z.rewind(zs);
    if (z.decompress(zs, ofield) == 0) rc = PMSZ_IO_EFAIL;
    z.field_free(ofield);
  }
  z.bs_close(bs);
  z.stream_close(zs);
  z.field_free(field);
  std::free(buf);
  if (compressed_bytes) *compressed_bytes = (int64_t)nbytes;
  return rc;
}

int pmsz_io_max_abs_err(int dtype, int64_t n, const void *a, const void *b, double bound,
                        double *max_err, int64_t *violations) {
  if (n < 0 || !a || !b) return PMSZ_IO_EINVAL;
  double m = 0;
  int64_t v = 0;
  for (int64_t i = 0; i < n; i++) {
    double x = dtype == 0 ? (double)((const float *)a)[i] : ((const double *)a)[i];
    double y = dtype == 0 ? (double)((const float *)b)[i] : ((const double *)b)[i];
    double d = std::fabs(x - y);
    if (d > m) m = d;
    if (!within_exact(x, y, bound)) v++;
  }
  if (max_err) *max_err = m;
  if (violations) *violations = v;
  return PMSZ_IO_OK;
}

int pmsz_io_zstd_version(void) {
  ZstdApi &z = zstd_api();
  return z.ok ? (int)z.version() : PMSZ_IO_EUNAVAILABLE;
}

int pmsz_io_write_edits(const char *prefix, int dtype, int nblocks, const pmsz_edit_block *blocks,
                        const uint8_t *packed, const void *values, int level_deltas,
                        int level_edits, int64_t *deltas_bytes, int64_t *edits_bytes) {
  if (!prefix || nblocks < 1 || !blocks || !packed || (dtype != 0 && dtype != 1))
    return PMSZ_IO_EINVAL;
  if (!zstd_api().ok) return PMSZ_IO_EUNAVAILABLE;
  const size_t es = dtype == 0 ? 4 : 8;
  for (int b = 0; b < nblocks; b++)
    if (blocks[b].packed_bytes < 0 || blocks[b].values < 0 || (blocks[b].values && !values))
      return PMSZ_IO_EINVAL;
  const std::string pre(prefix);
  FILE *fd = std::fopen((pre + "_deltas.bin").c_str(), "wb");
  if (!fd) return PMSZ_IO_EFAIL;
  FILE *fe = std::fopen((pre + "_edits.bin").c_str(), "wb");
  if (!fe) {
    std::fclose(fd);
    return PMSZ_IO_EFAIL;
  }
  // the blocks' frames are independent: compress them on up to 16 threads (the
  // system libzstd has no threads of its own), then write them in rank order
  std::vector<std::vector<uint8_t>> fdel(nblocks), fval(nblocks);
  std::vector<int> brc(nblocks, PMSZ_IO_OK);
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int b; (b = next.fetch_add(1)) < nblocks;) {
      brc[b] = zstd_frame(packed + blocks[b].packed_offset, (size_t)blocks[b].packed_bytes,
                          level_deltas, fdel[b]);
      if (brc[b] == PMSZ_IO_OK)
        brc[b] = zstd_frame((const uint8_t *)values + blocks[b].value_offset * es,
                            (size_t)blocks[b].values * es, level_edits, fval[b]);
    }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const int nt = std::min(nblocks, (int)std::min(16u, hw ? hw : 1u));
  std::vector<std::thread> pool;
  for (int k = 1; k < nt; k++) pool.emplace_back(work);
  work();
  for (std::thread &th : pool) th.join();
  int64_t nd = 0, ne = 0;
  int rc = PMSZ_IO_OK;
  for (int b = 0; b < nblocks && rc == PMSZ_IO_OK; b++) {
    rc = brc[b];
    if (rc == PMSZ_IO_OK && std::fwrite(fdel[b].data(), 1, fdel[b].size(), fd) != fdel[b].size())
      rc = PMSZ_IO_EFAIL;
    if (rc == PMSZ_IO_OK && std::fwrite(fval[b].data(), 1, fval[b].size(), fe) != fval[b].size())
      rc = PMSZ_IO_EFAIL;
    nd += (int64_t)fdel[b].size();
    ne += (int64_t)fval[b].size();
  }
  if (std::fclose(fd) != 0) rc = PMSZ_IO_EFAIL;
  if (std::fclose(fe) != 0) rc = PMSZ_IO_EFAIL;
  if (rc) return rc;
  if (deltas_bytes) *deltas_bytes = nd;
  if (edits_bytes) *edits_bytes = ne;
  return PMSZ_IO_OK;
}

int pmsz_io_read_edits(const char *prefix, int dtype, int q, int nblocks,
                       const int64_t *block_voxels, uint8_t *symbols, void *values,
                       int64_t values_cap, int64_t *nvalues) {
  if (!prefix || nblocks < 1 || !block_voxels || !symbols || q < 0 || values_cap < 0 ||
      (values_cap && !values) || (dtype != 0 && dtype != 1))
    return PMSZ_IO_EINVAL;
  ZstdApi &z = zstd_api();
  if (!z.ok) return PMSZ_IO_EUNAVAILABLE;
  const unsigned nbits = bits_for_max((unsigned)q);
  if (nbits > 16) return PMSZ_IO_EINVAL;
  const std::string pre(prefix);
  std::vector<uint8_t> comp;
  // deltas: the concatenated frames decode to the blocks' packed streams in order
  if (!read_file(pre + "_deltas.bin", comp)) return PMSZ_IO_EFAIL;
  size_t need = 0;
  for (int b = 0; b < nblocks; b++) {
    if (block_voxels[b] < 0) return PMSZ_IO_EINVAL;
    need += ((size_t)block_voxels[b] * nbits + 7) / 8;
  }
  std::vector<uint8_t> pk(std::max<size_t>(need, 1));
  size_t got = need ? z.decompress(pk.data(), need, comp.data(), comp.size()) : 0;
  if (z.is_error(got) || got != need) return PMSZ_IO_EFAIL;
  const uint64_t mask = (1ull << nbits) - 1ull;
  size_t in = 0, o = 0;
  for (int b = 0; b < nblocks; b++) {
    uint64_t buf = 0;
    unsigned cnt = 0;
    for (int64_t i = 0; i < block_voxels[b]; i++) {
      while (cnt < nbits) {
        buf |= (uint64_t)pk[in++] << cnt;
        cnt += 8;
      }
      symbols[o++] = (uint8_t)(buf & mask);
      buf >>= nbits;
      cnt -= nbits;
    }
  }
  // edits: the concatenated frames of clamp values
  if (!read_file(pre + "_edits.bin", comp)) return PMSZ_IO_EFAIL;
  const size_t es = dtype == 0 ? 4 : 8;
  size_t vb = 0;
  if (!comp.empty()) {
    if (!values_cap) return PMSZ_IO_EFAIL;
    vb = z.decompress(values, (size_t)values_cap * es, comp.data(), comp.size());
    if (z.is_error(vb) || vb % es) return PMSZ_IO_EFAIL;
  }
  if (nvalues) *nvalues = (int64_t)(vb / es);
  return PMSZ_IO_OK;
}

int pmsz_io_read_raw(const char *path, int dtype, int64_t n, void *out) {
  if (!path || !out || n < 0) return PMSZ_IO_EINVAL;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return PMSZ_IO_EFAIL;
  size_t s = dtype == 0 ? 4 : 8;
  size_t got = std::fread(out, s, (size_t)n, fp);
  std::fclose(fp);
  return got == (size_t)n ? PMSZ_IO_OK : PMSZ_IO_EFAIL;
}

int pmsz_io_write_raw(const char *path, int dtype, int64_t n, const void *data) {
  if (!path || !data || n < 0) return PMSZ_IO_EINVAL;
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return PMSZ_IO_EFAIL;
  size_t s = dtype == 0 ? 4 : 8;
  size_t put = std::fwrite(data, s, (size_t)n, fp);
  std::fclose(fp);
  return put == (size_t)n ? PMSZ_IO_OK : PMSZ_IO_EFAIL;
}

}  // extern "C"
