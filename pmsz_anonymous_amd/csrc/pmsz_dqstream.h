// pmsz_dqstream.h -- the compressed stream of the dual-quantisation Lorenzo
// codec (pmsz_codec_dq): canonical Huffman codes over the Lorenzo residual
// symbols of each rank block, encoded and decoded on the device.
//
// Symbols.  The residual d of quantum k (k - Lorenzo prediction over the 7
// lower neighbours of the block's core, zero outside it) is symbol d + 512 when
// |d| < 512, else the escape symbol 1024 with d itself (mod 2^32) in the escape
// list.  Voxels stored verbatim keep their residual symbol; their values follow
// in the verbatim list, with their index in the block.
//
// Stream, little endian (all offsets from the stream's first byte):
//   DqsHeader                      64 B
//   DqsBlock x nblocks             64 B each, rank order
//   code lengths                   1025 B per block, padded to 8 B
//   chunk word offsets             uint64 x (chunks + 1): word offset of each chunk
//                                  in the bitstream, then the total
//   chunk escape offsets           uint64 x chunks: index of each chunk's first
//                                  escape in the escape list
//   bitstream                      uint32 words; chunk c holds the codes of
//                                  symbols [c * 2048, ...) of its block in core
//                                  order (x fastest), packed LSB first, each code
//                                  bit-reversed (its first bit lowest), padded
//                                  to a whole word
//   escapes                        uint32 x escapes (all blocks, in order)
//   verbatim indices               uint32 x verbatim (index in the block's core
//                                  order), padded to 8 B
//   verbatim values                T x verbatim
// The codes of a block are the canonical code (DEFLATE's assignment) of the
// Huffman code lengths of its symbol counts, at most 24 bits.
//
// Decoding: Huffman decode per chunk (one thread each, an LDS table for codes of
// up to 11 bits, canonical decode beyond), then the inverse Lorenzo transform --
// the residual d is the mixed difference of k along x, y and z, so k is the
// prefix sum of d along x, then y, then z (in wrapping 32-bit arithmetic, exact
// since k fits 31 bits) -- and g0 = k * 2 eb in the value type, the verbatim
// values last.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pmsz_codec.h"

namespace pmsz {

constexpr int kDqsSymbols = kDqBins + 1;  // residual bins and the escape symbol
constexpr int kDqsEscape = kDqBins;
constexpr int kDqsChunk = 2048;           // symbols per chunk
constexpr int kDqsPer = 8;                // symbols per thread of the encoder
constexpr int kDqsThreads = kDqsChunk / kDqsPer;
constexpr int kDqsMaxLen = 24;            // longest code
constexpr int kDqsFastBits = 11;          // decoder table: codes of up to 11 bits
constexpr int kDqsChunkWords = (kDqsChunk * kDqsMaxLen + 31) / 32;
constexpr uint32_t kDqsMagic = 0x31514450u;  // "PDQ1"

struct DqsHeader {
  uint32_t magic;
  uint32_t dtype;
  int64_t nx, ny, nz;
  int32_t sub[3];
  uint32_t nblocks;
  double abs_eb;
  uint32_t chunk;
  uint32_t reserved;
};
static_assert(sizeof(DqsHeader) == 64, "stream header layout");

struct DqsBlock {
  uint64_t voxels;
  uint64_t escapes, verbatim;
  uint64_t esc_offset, verb_offset;  // entries of the earlier blocks
  uint32_t first_chunk, nchunks;
  uint64_t reserved[2];
};
static_assert(sizeof(DqsBlock) == 64, "stream block record layout");

// Device view of one rank block for the stream kernels.
struct DqsGeo {
  int32_t x0, y0, z0;   // first voxel
  int32_t lx, ly, lz;   // extent
  uint32_t first_chunk;  // first chunk of the block
  uint32_t nchunks;
  uint64_t voxels;
  uint64_t esc_offset, verb_offset;
  uint32_t wg0;  // first decoder workgroup of the block
  uint32_t pad;
};

// Residual symbol of local voxel (x, y, z) of block B (kq: 2k + verbatim flag).
__device__ __forceinline__ int64_t dqs_residual(const int32_t *__restrict__ kq, int64_t i,
                                                int32_t x, int32_t y, int32_t z, int64_t nx,
                                                int64_t sxy) {
  const bool hx = x > 0, hy = y > 0, hz = z > 0;
  auto K = [&](bool ok, int64_t j) -> int64_t { return ok ? (int64_t)(kq[j] >> 1) : 0; };
  const int64_t pred = K(hx, i - 1) + K(hy, i - nx) + K(hz, i - sxy) - K(hx && hy, i - 1 - nx) -
                       K(hx && hz, i - 1 - sxy) - K(hy && hz, i - nx - sxy) +
                       K(hx && hy && hz, i - 1 - nx - sxy);
  return (int64_t)(kq[i] >> 1) - pred;
}

__device__ __forceinline__ uint32_t dqs_symbol(int64_t d) {
  return (d > -kDqRadius && d < kDqRadius) ? (uint32_t)(d + kDqRadius) : (uint32_t)kDqsEscape;
}

// The block of chunk c (chunks of a block are consecutive).
__device__ __forceinline__ uint32_t dqs_block_of(const DqsGeo *__restrict__ geo, uint32_t nb,
                                                 uint32_t c) {
  uint32_t lo = 0, hi = nb - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (geo[mid].first_chunk <= c) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Block-wide exclusive scan of one value per thread (kDqsThreads threads);
// returns the thread's prefix and sets *total.
template <typename V>
__device__ __forceinline__ V dqs_block_scan(V v, V *s_tmp, V *total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  V x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const V y = __shfl_up(x, o, 64);
    if ((int)lane >= o) x += y;
  }
  if (lane == 63) s_tmp[wave] = x;
  __syncthreads();
  V before = 0, all = 0;
  for (uint32_t w = 0; w < (uint32_t)(kDqsThreads / 64); w++) {
    if (w < wave) before += s_tmp[w];
    all += s_tmp[w];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// Steps the local coordinates (x, y, z) of a block of extent (lx, ly) one voxel
// on in core order.
__device__ __forceinline__ void dqs_step(int32_t &x, int32_t &y, int32_t &z, int32_t lx,
                                         int32_t ly) {
  if (++x == lx) {
    x = 0;
    if (++y == ly) {
      y = 0;
      z++;
    }
  }
}

// Per chunk: code bits, escapes and verbatim voxels.  codes: per block 1025
// entries (reversed code << 5 | length).
__global__ __launch_bounds__(kDqsThreads) void k_dqs_count(
    const int32_t *__restrict__ kq, const DqsGeo *__restrict__ geo, uint32_t nb, int64_t nx,
    int64_t sxy, const uint32_t *__restrict__ codes, uint32_t *__restrict__ cbits,
    uint32_t *__restrict__ cesc, uint32_t *__restrict__ cverb) {
  __shared__ uint32_t s_len[kDqsSymbols];
  __shared__ uint32_t s_red[3][kDqsThreads / 64];
  const uint32_t c = blockIdx.x;
  const uint32_t b = dqs_block_of(geo, nb, c);
  const DqsGeo G = geo[b];
  for (int s = threadIdx.x; s < kDqsSymbols; s += blockDim.x)
    s_len[s] = codes[(size_t)b * kDqsSymbols + s] & 31u;
  __syncthreads();
  const uint64_t li0 = (uint64_t)(c - G.first_chunk) * kDqsChunk + threadIdx.x * kDqsPer;
  uint32_t bits = 0, esc = 0, verb = 0;
  if (li0 < G.voxels) {
    const uint64_t plane = (uint64_t)G.lx * G.ly;
    int32_t z = (int32_t)(li0 / plane);
    const uint64_t r = li0 - (uint64_t)z * plane;
    int32_t y = (int32_t)(r / (uint32_t)G.lx), x = (int32_t)(r - (uint64_t)y * G.lx);
    for (int u = 0; u < kDqsPer && li0 + u < G.voxels; u++) {
      const int64_t i = (int64_t)(G.z0 + z) * sxy + (int64_t)(G.y0 + y) * nx + G.x0 + x;
      const uint32_t sym = dqs_symbol(dqs_residual(kq, i, x, y, z, nx, sxy));
      bits += s_len[sym];
      esc += sym == (uint32_t)kDqsEscape;
      verb += (uint32_t)(kq[i] & 1);
      dqs_step(x, y, z, G.lx, G.ly);
    }
  }
  // block sums
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t v[3] = {bits, esc, verb};
#pragma unroll
  for (int q = 0; q < 3; q++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
    if (lane == 0) s_red[q][wave] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t[3] = {0, 0, 0};
    for (int w = 0; w < kDqsThreads / 64; w++)
      for (int q = 0; q < 3; q++) t[q] += s_red[q][w];
    cbits[c] = t[0];
    cesc[c] = t[1];
    cverb[c] = t[2];
  }
}

// Exclusive scans over the chunks (one workgroup): word offsets of the chunks
// (cwoff, chunks + 1 entries, the last one the total), escape and verbatim
// offsets (ceoff, cvoff; totals into tot[0], tot[1]).
__global__ __launch_bounds__(1024) void k_dqs_scan(const uint32_t *__restrict__ cbits,
                                                   const uint32_t *__restrict__ cesc,
                                                   const uint32_t *__restrict__ cverb,
                                                   uint32_t chunks, uint64_t *__restrict__ cwoff,
                                                   uint64_t *__restrict__ ceoff,
                                                   uint64_t *__restrict__ cvoff,
                                                   uint64_t *__restrict__ tot) {
  __shared__ uint64_t s_part[3][1024];
  const uint32_t t = threadIdx.x, nt = blockDim.x;
  const uint32_t per = (chunks + nt - 1) / nt;
  const uint32_t a = t * per, e = a + per < chunks ? a + per : chunks;
  uint64_t sw = 0, se = 0, sv = 0;
  for (uint32_t c = a; c < e; c++) {
    sw += (cbits[c] + 31u) / 32u;
    se += cesc[c];
    sv += cverb[c];
  }
  s_part[0][t] = sw;
  s_part[1][t] = se;
  s_part[2][t] = sv;
  __syncthreads();
  if (t == 0) {
    uint64_t r[3] = {0, 0, 0};
    for (uint32_t k = 0; k < nt; k++)
      for (int q = 0; q < 3; q++) {
        const uint64_t v = s_part[q][k];
        s_part[q][k] = r[q];
        r[q] += v;
      }
    cwoff[chunks] = r[0];
    tot[0] = r[1];
    tot[1] = r[2];
  }
  __syncthreads();
  uint64_t ow = s_part[0][t], oe = s_part[1][t], ov = s_part[2][t];
  for (uint32_t c = a; c < e; c++) {
    cwoff[c] = ow;
    ceoff[c] = oe;
    cvoff[c] = ov;
    ow += (cbits[c] + 31u) / 32u;
    oe += cesc[c];
    ov += cverb[c];
  }
}

// The encoder: per chunk, every thread's codes into an LDS bit buffer at its
// prefix, the chunk's words to the bitstream, and the escapes and verbatim
// entries at their offsets.
template <typename T>
__global__ __launch_bounds__(kDqsThreads) void k_dqs_encode(
    const int32_t *__restrict__ kq, const T *__restrict__ g0, const DqsGeo *__restrict__ geo,
    uint32_t nb, int64_t nx, int64_t sxy, const uint32_t *__restrict__ codes,
    const uint64_t *__restrict__ cwoff, const uint64_t *__restrict__ ceoff,
    const uint64_t *__restrict__ cvoff, uint32_t *__restrict__ words,
    uint32_t *__restrict__ escs, uint32_t *__restrict__ vidx, T *__restrict__ vval) {
  __shared__ uint32_t s_code[kDqsSymbols];
  __shared__ uint32_t s_buf[kDqsChunkWords];
  __shared__ uint32_t s_tmp[kDqsThreads / 64];
  const uint32_t c = blockIdx.x;
  const uint32_t b = dqs_block_of(geo, nb, c);
  const DqsGeo G = geo[b];
  for (int s = threadIdx.x; s < kDqsSymbols; s += blockDim.x)
    s_code[s] = codes[(size_t)b * kDqsSymbols + s];
  for (int w = threadIdx.x; w < kDqsChunkWords; w += blockDim.x) s_buf[w] = 0u;
  __syncthreads();
  const uint64_t cfirst = (uint64_t)(c - G.first_chunk) * kDqsChunk;
  const uint64_t li0 = cfirst + threadIdx.x * kDqsPer;
  uint32_t sym[kDqsPer];
  int64_t dv[kDqsPer];
  int64_t gi[kDqsPer];
  uint32_t n = 0, bits = 0, esc = 0, verb = 0;
  if (li0 < G.voxels) {
    const uint64_t plane = (uint64_t)G.lx * G.ly;
    int32_t z = (int32_t)(li0 / plane);
    const uint64_t r = li0 - (uint64_t)z * plane;
    int32_t y = (int32_t)(r / (uint32_t)G.lx), x = (int32_t)(r - (uint64_t)y * G.lx);
    for (int u = 0; u < kDqsPer && li0 + u < G.voxels; u++) {
      const int64_t i = (int64_t)(G.z0 + z) * sxy + (int64_t)(G.y0 + y) * nx + G.x0 + x;
      gi[u] = i;
      dv[u] = dqs_residual(kq, i, x, y, z, nx, sxy);
      sym[u] = dqs_symbol(dv[u]);
      bits += s_code[sym[u]] & 31u;
      esc += sym[u] == (uint32_t)kDqsEscape;
      verb += (uint32_t)(kq[i] & 1);
      n++;
      dqs_step(x, y, z, G.lx, G.ly);
    }
  }
  uint32_t tot;
  const uint32_t boff = dqs_block_scan<uint32_t>(bits, s_tmp, &tot);
  const uint32_t eoff = dqs_block_scan<uint32_t>(esc, s_tmp, &tot);
  const uint32_t voff = dqs_block_scan<uint32_t>(verb, s_tmp, &tot);
  uint32_t pos = boff;
  uint64_t ep = ceoff[c] + eoff, vp = cvoff[c] + voff;
  for (uint32_t u = 0; u < n; u++) {
    const uint32_t e = s_code[sym[u]];
    const uint32_t len = e & 31u, rev = e >> 5;
    const uint32_t w = pos >> 5, sh = pos & 31u;
    atomicOr(&s_buf[w], rev << sh);
    if (sh + len > 32u) atomicOr(&s_buf[w + 1], rev >> (32u - sh));
    pos += len;
    if (sym[u] == (uint32_t)kDqsEscape) escs[ep++] = (uint32_t)(uint64_t)dv[u];
    if (kq[gi[u]] & 1) {
      vidx[vp] = (uint32_t)(li0 + u);
      vval[vp] = g0[gi[u]];
      vp++;
    }
  }
  __syncthreads();
  const uint64_t w0 = cwoff[c];
  const uint32_t nw = (uint32_t)(cwoff[c + 1] - w0);
  for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) words[w0 + w] = s_buf[w];
}

// Decoder tables of one block: fast[2^11] = symbol | length << 16 for codes of
// up to 11 bits (0 where a longer code starts); canonical decode for longer ones
// from first[len], count[len], index[len] and the symbols sorted by (length,
// symbol).
struct DqsDecTable {
  uint32_t fast[1 << kDqsFastBits];
  uint32_t first[kDqsMaxLen + 1];
  uint32_t count[kDqsMaxLen + 1];
  uint32_t index[kDqsMaxLen + 1];
  uint16_t sorted[kDqsSymbols + 1];
};

// The decoder: one thread per chunk, the chunks of one block per workgroup
// (geo[b].wg0 up to the next block's); residuals d (mod 2^32) into kd at the
// voxel's index in the field.  cwoff / ceoff: the chunks' word and escape
// offsets (checked by the host: non-decreasing, within the stream); an escape
// beyond nesc or a bit pattern that is no code sets *err.
constexpr int kDqsDecThreads = 64;
__global__ __launch_bounds__(kDqsDecThreads) void k_dqs_decode(
    const uint32_t *__restrict__ words, const uint64_t *__restrict__ cwoff,
    const uint64_t *__restrict__ ceoff, const uint32_t *__restrict__ escs, uint64_t nesc,
    const DqsGeo *__restrict__ geo, uint32_t nb, int64_t nx, int64_t sxy,
    const DqsDecTable *__restrict__ tables, uint32_t *__restrict__ kd, uint32_t *__restrict__ err) {
  __shared__ DqsDecTable s_t;
  uint32_t lo = 0, hi = nb - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (geo[mid].wg0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t b = lo;
  const DqsGeo G = geo[b];
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&tables[b]);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&s_t);
    for (uint32_t k = threadIdx.x; k < sizeof(DqsDecTable) / 4; k += blockDim.x) dst[k] = src[k];
  }
  __syncthreads();
  const uint32_t cl = (blockIdx.x - G.wg0) * kDqsDecThreads + threadIdx.x;  // chunk in block
  if (cl >= G.nchunks) return;
  const uint32_t c = G.first_chunk + cl;
  const uint64_t li0 = (uint64_t)cl * kDqsChunk;
  const uint64_t nsym = G.voxels - li0 < (uint64_t)kDqsChunk ? G.voxels - li0 : kDqsChunk;
  uint64_t wp = cwoff[c];
  const uint64_t wend = cwoff[c + 1];
  uint64_t buf = 0;
  uint32_t have = 0;
  const uint64_t plane = (uint64_t)G.lx * G.ly;
  int32_t z = (int32_t)(li0 / plane);
  const uint64_t r = li0 - (uint64_t)z * plane;
  int32_t y = (int32_t)(r / (uint32_t)G.lx), x = (int32_t)(r - (uint64_t)y * G.lx);
  uint64_t eidx = ceoff[c];
  uint32_t bad = 0;
  for (uint64_t u = 0; u < nsym; u++) {
    while (have <= 32u) {  // at least 33 bits in the buffer
      const uint64_t w = wp < wend ? (uint64_t)words[wp] : 0ull;
      wp++;
      buf |= w << have;
      have += 32u;
    }
    const uint32_t e = s_t.fast[(uint32_t)buf & ((1u << kDqsFastBits) - 1u)];
    uint32_t sym = 0, len = 0;
    if (e) {
      sym = e & 0xffffu;
      len = e >> 16;
    } else {  // canonical decode, the code's first bit lowest
      uint32_t code = 0;
      for (uint32_t l = 1; l <= (uint32_t)kDqsMaxLen; l++) {
        code = (code << 1) | (uint32_t)((buf >> (l - 1)) & 1u);
        if (code - s_t.first[l] < s_t.count[l]) {
          sym = s_t.sorted[s_t.index[l] + code - s_t.first[l]];
          len = l;
          break;
        }
      }
      if (!len) {  // no code: a corrupt stream
        bad = 1;
        len = 1;
      }
    }
    buf >>= len;
    have -= len;
    uint32_t d;
    if (sym == (uint32_t)kDqsEscape) {
      if (eidx < nesc) {
        d = escs[eidx];
      } else {
        d = 0;
        bad = 1;
      }
      eidx++;
    } else {
      d = (uint32_t)((int32_t)sym - kDqRadius);
    }
    kd[(int64_t)(G.z0 + z) * sxy + (int64_t)(G.y0 + y) * nx + G.x0 + x] = d;
    dqs_step(x, y, z, G.lx, G.ly);
  }
  if (bad) atomicOr(err, 1u);
}

// Inverse Lorenzo, along x: one wave per row of a block, 64 voxels per step
// (wrapping 32-bit sums).
__global__ __launch_bounds__(256) void k_dqs_prefix_x(uint32_t *__restrict__ kd,
                                                      const DqsGeo *__restrict__ geo,
                                                      uint32_t b, int64_t nx, int64_t sxy) {
  const DqsGeo G = geo[b];
  const uint64_t row = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= (uint64_t)G.ly * G.lz) return;
  const uint32_t lane = threadIdx.x & 63u;
  const int32_t y = (int32_t)(row % (uint32_t)G.ly), z = (int32_t)(row / (uint32_t)G.ly);
  uint32_t *p = kd + (int64_t)(G.z0 + z) * sxy + (int64_t)(G.y0 + y) * nx + G.x0;
  uint32_t carry = 0;
  for (int32_t x0 = 0; x0 < G.lx; x0 += 64) {
    const int32_t x = x0 + (int32_t)lane;
    uint32_t v = x < G.lx ? p[x] : 0u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(v, o, 64);
      if ((int)lane >= o) v += t;
    }
    v += carry;
    if (x < G.lx) p[x] = v;
    carry = __shfl(v, 63, 64);
  }
}

// Inverse Lorenzo along y (stride nx) or z (stride sxy): one thread per line.
__global__ __launch_bounds__(256) void k_dqs_prefix_yz(uint32_t *__restrict__ kd,
                                                       const DqsGeo *__restrict__ geo,
                                                       uint32_t b, int64_t nx, int64_t sxy,
                                                       int axis) {
  const DqsGeo G = geo[b];
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t lines = axis == 1 ? (uint64_t)G.lx * G.lz : (uint64_t)G.lx * G.ly;
  if (t >= lines) return;
  const int32_t x = (int32_t)(t % (uint32_t)G.lx);
  const int32_t o = (int32_t)(t / (uint32_t)G.lx);  // z for axis 1, y for axis 2
  const int64_t stride = axis == 1 ? nx : sxy;
  const int32_t len = axis == 1 ? G.ly : G.lz;
  int64_t i = axis == 1 ? (int64_t)(G.z0 + o) * sxy + (int64_t)G.y0 * nx + G.x0 + x
                        : (int64_t)G.z0 * sxy + (int64_t)(G.y0 + o) * nx + G.x0 + x;
  uint32_t s = 0;
  for (int32_t k = 0; k < len; k++, i += stride) {
    s += kd[i];
    kd[i] = s;
  }
}

// g0 = k * 2 eb in the value type, over every voxel of the field.
template <typename T>
__global__ __launch_bounds__(256) void k_dqs_values(const uint32_t *__restrict__ kd,
                                                    T *__restrict__ g0, int64_t n, double q2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    g0[i] = (T)((double)(int32_t)kd[i] * q2);
}

// The verbatim voxels: their stored values, at their index in their block.
template <typename T>
__global__ __launch_bounds__(256) void k_dqs_verbatim(const uint32_t *__restrict__ vidx,
                                                      const T *__restrict__ vval,
                                                      const DqsGeo *__restrict__ geo,
                                                      uint32_t nb, uint64_t total, int64_t nx,
                                                      int64_t sxy, T *__restrict__ g0,
                                                      uint32_t *__restrict__ err) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nb - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (geo[mid].verb_offset <= e) lo = mid;
      else hi = mid - 1;
    }
    // blocks without verbatim voxels share their offset with the next one: the
    // last block whose range holds e
    const DqsGeo &G = geo[lo];
    const uint64_t li = vidx[e];
    if (li >= G.voxels) {  // a corrupt stream
      atomicOr(err, 2u);
      continue;
    }
    const uint64_t plane = (uint64_t)G.lx * G.ly;
    const int32_t z = (int32_t)(li / plane);
    const uint64_t r = li - (uint64_t)z * plane;
    const int32_t y = (int32_t)(r / (uint32_t)G.lx), x = (int32_t)(r - (uint64_t)y * G.lx);
    g0[(int64_t)(G.z0 + z) * sxy + (int64_t)(G.y0 + y) * nx + G.x0 + x] = vval[e];
  }
}

}  // namespace pmsz
