/*
 * pmsz_io.h -- C-ABI of libpmsz_io.so, the host-side front end that feeds the
 * correction: synthetic Perlin fields, the error-bounded codecs that produce the
 * decompressed field, the zstd files of the edit side channel, and raw field I/O.  None of this is on the correction's
 * hot path; it replaces the reference's reader and compressor adapters so the
 * CLI and the benchmarks can build their inputs.
 *
 * dtype: 0 = float32, 1 = float64.  Arrays are C order, x fastest.
 */
#ifndef PMSZ_IO_H
#define PMSZ_IO_H
#include <stdint.h>

#include "pmsz.h" /* pmsz_edit_block */

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PMSZ_IO_OK = 0,
  PMSZ_IO_EINVAL = -1,
  PMSZ_IO_ENOMEM = -2,
  PMSZ_IO_EFAIL = -5,
  PMSZ_IO_EUNAVAILABLE = -6,
};

/* Fractal improved-Perlin field: sum over octaves o < octaves of
 * persistence^o * noise(base_freq * 2^o * ((x+0.5)/nx, (y+0.5)/ny, (z+0.5)/nz)),
 * divided by the sum of the amplitudes; computed in double, rounded once.
 * Stands in for generate_perlin.py:17-135 with the parameters of the workload. */
int pmsz_io_perlin(int dtype, int64_t nx, int64_t ny, int64_t nz, uint64_t seed,
                   double base_freq, int octaves, double persistence, void *out);

/* min / max of n values. */
int pmsz_io_range(int dtype, int64_t n, const void *data, double *vmin, double *vmax);

/* SZ-style Lorenzo + linear-quantisation codec with absolute bound abs_eb:
 * writes the decompressed field g0 (|f - g0| <= abs_eb exactly, every voxel).
 * Replaces SZ_compress / SZ_decompress of compressLocalData_coreOnly
 * (pMSz.cu:1536-1624), whose library is not available. */
int pmsz_io_szlike(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *f,
                   double abs_eb, void *g0, int64_t *unpredictable, double *est_bytes);

/* ZFP fixed-accuracy round trip (compressLocalDataZFP_coreOnly,
 * pMSz.cu:1436-1534).  PMSZ_IO_EUNAVAILABLE when libzfp cannot be loaded. */
int pmsz_io_zfp_available(void);
int pmsz_io_zfp(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *f,
                double tolerance, void *g0, int64_t *compressed_bytes);

/* Largest |a - b| and the number of voxels where |a - b| <= bound fails exactly. */
int pmsz_io_max_abs_err(int dtype, int64_t n, const void *a, const void *b, double bound,
                        double *max_err, int64_t *violations);

/* zstd, from the system libzstd loaded at run time (the reference links zstd:
 * compress_zstd / zstd_compress, pMSz.cu:2261-2288).  Returns
 * ZSTD_versionNumber(), or PMSZ_IO_EUNAVAILABLE when the library is absent. */
int pmsz_io_zstd_version(void);

/* The edit files of gather_cost_sparse (pMSz.cu:2469-2523): for each rank
 * block in rank order, one zstd frame of its packed symbol stream at level
 * level_deltas (5 in the reference, compress_symbols_0_N) appended to
 * <prefix>_deltas.bin, and one frame of its clamp values at level level_edits
 * (3, compress_zstd) appended to <prefix>_edits.bin; a block without clamp
 * values adds nothing there (compress_zstd returns no bytes for an empty
 * vector).  packed, values and blocks are pmsz_encode_edits' outputs, copied to
 * the host.  The files are truncated first (the reference opens them without
 * MPI_MODE_TRUNC).  deltas_bytes / edits_bytes (nullable) receive the file
 * sizes; their sum is the reference's storage overhead (pMSz.cu:2495). */
int pmsz_io_write_edits(const char *prefix, int dtype, int nblocks, const pmsz_edit_block *blocks,
                        const uint8_t *packed, const void *values, int level_deltas,
                        int level_edits, int64_t *deltas_bytes, int64_t *edits_bytes);

/* Reads the edit files back (validateResult, pMSz.cu:2545-2590): decompresses
 * every frame of <prefix>_deltas.bin and unpacks each block's stream
 * (unpack_nbits, pMSz.cu:2393-2408) into symbols (sum of block_voxels bytes:
 * the blocks' symbols in rank order, each block x fastest), and decompresses
 * <prefix>_edits.bin into values (room for values_cap values), setting
 * *nvalues.  PMSZ_IO_EFAIL on a missing file, a zstd error or a size mismatch. */
int pmsz_io_read_edits(const char *prefix, int dtype, int q, int nblocks,
                       const int64_t *block_voxels, uint8_t *symbols, void *values,
                       int64_t values_cap, int64_t *nvalues);

/* Raw little-endian arrays (the reference reads raw f64, pMSz.cu:274-313). */
int pmsz_io_read_raw(const char *path, int dtype, int64_t n, void *out);
int pmsz_io_write_raw(const char *path, int dtype, int64_t n, const void *data);

#ifdef __cplusplus
}
#endif
#endif
