// pmsz_codec.h -- the GPU compressor front end: an SZ-style Lorenzo codec with
// dual quantisation (cuSZ: J. Tian et al., PACT 2020), standing in for SZ3
// (compressLocalData_coreOnly, pMSz.cu:1536-1624), which is absent here.
//
// Dual quantisation removes the serial dependence of SZ's Lorenzo predictor on
// reconstructed values: every voxel is first quantised on its own,
//   k = round(f * (1 / (2 eb)))  (in double, the reciprocal rounded once;
//                                 0 when the product's magnitude is >= 2^30),
//   g0 = (T)(k * 2 eb)      (f itself, "verbatim", when that misses the bound),
// and the Lorenzo residual d = k - pred(k) over the 7 lower neighbours is then
// lossless integer work.  Only the residuals' histogram is needed here: it gives
// the entropy-coded size of the compressed stream.  Prediction is core-only per
// rank block (neighbours outside the block count as 0), as the reference
// compresses each rank's core separately.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pmsz_common.h"

namespace pmsz {

constexpr int kDqRadius = 512;
constexpr int kDqBins = 2 * kDqRadius;            // residual d in bin d + kDqRadius
constexpr int kDqSlots = kDqBins + 2;             // bins, escapes, verbatim voxels
constexpr int kDqRows = 64;                       // rows of a rank block per workgroup
constexpr double kDqLimit = 1073741824.0;         // 2^30

// One rank block of the codec, in rank order.
struct DqBlock {
  int32_t x0, y0, z0;     // first voxel
  int32_t lx, ly, lz;     // extent
  uint32_t chunk0;        // first workgroup of the block
  uint32_t pad;
};

// Pass 1: the quantum of every voxel, stored as 2k + verbatim, and g0.  A
// thread takes 16 bytes of f (vector loads and stores when f, g0 and the quanta
// are 16-byte aligned).
template <typename T>
__device__ __forceinline__ int32_t dq_quantum(T fv, T &gv, double eb, double inv_q2) {
  const double q2 = 2.0 * eb;
  // the rounded quantum must stay below 2^30 in magnitude, so that 2k + 1 fits
  // an int32; NaN fails the test (k = 0) and is then stored verbatim
  const double r = round((double)fv * inv_q2);
  const int32_t k = fabs(r) < kDqLimit ? (int32_t)r : 0;
  const T v = (T)((double)k * q2);
  const bool keep = within_exact((double)fv, (double)v, eb);
  gv = keep ? v : fv;
  return 2 * k + (keep ? 0 : 1);
}

template <typename T>
__global__ __launch_bounds__(256) void k_dq_quant(const T *__restrict__ f, T *__restrict__ g0,
                                                  int32_t *__restrict__ kq, int64_t n, double eb,
                                                  double inv_q2, bool aligned) {
  constexpr int kV = 16 / sizeof(T);
  using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  using K = typename std::conditional<sizeof(T) == 4, int4, int2>::type;
  const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kV;
  if (i0 >= n) return;
  if (aligned && i0 + kV <= n) {
    const V a = *reinterpret_cast<const V *>(f + i0);
    V b;
    K k;
    const T *pa = reinterpret_cast<const T *>(&a);
    T *pb = reinterpret_cast<T *>(&b);
    int32_t *pk = reinterpret_cast<int32_t *>(&k);
#pragma unroll
    for (int u = 0; u < kV; u++) pk[u] = dq_quantum<T>(pa[u], pb[u], eb, inv_q2);
    *reinterpret_cast<V *>(g0 + i0) = b;
    *reinterpret_cast<K *>(kq + i0) = k;
  } else {
    for (int64_t i = i0; i < i0 + kV && i < n; i++) {
      T gv;
      kq[i] = dq_quantum<T>(f[i], gv, eb, inv_q2);
      g0[i] = gv;
    }
  }
}

// Pass 2: Lorenzo residuals of one rank block's rows into an LDS histogram, then
// into the block's global histogram (one atomic per nonzero slot).
__global__ __launch_bounds__(256) void k_dq_hist(const int32_t *__restrict__ kq,
                                                 const DqBlock *__restrict__ blocks,
                                                 uint32_t nblocks, int64_t nx, int64_t sxy,
                                                 unsigned long long *__restrict__ hist) {
  __shared__ uint32_t s_h[kDqSlots];
  for (int b = threadIdx.x; b < kDqSlots; b += blockDim.x) s_h[b] = 0;
  uint32_t lo = 0, hi = nblocks - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (blocks[mid].chunk0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const DqBlock B = blocks[lo];
  __syncthreads();
  const int64_t rows = (int64_t)B.ly * B.lz;
  const int64_t r0 = (int64_t)(blockIdx.x - B.chunk0) * kDqRows;
  const int64_t r1 = r0 + kDqRows < rows ? r0 + kDqRows : rows;
  const uint32_t lane = threadIdx.x & 63u;
  // (y, z) of the first row, then stepped: no division per row
  int32_t y = (int32_t)(r0 % B.ly), z = (int32_t)(r0 / B.ly);
  for (int64_t row = r0; row < r1; row++) {
    const int64_t base = (int64_t)(B.z0 + z) * sxy + (int64_t)(B.y0 + y) * nx + B.x0;
    const bool hy = y > 0, hz = z > 0;
    // whole waves step through the row together, so the ballot below is uniform
    for (int32_t x0 = 0; x0 < B.lx; x0 += (int32_t)blockDim.x) {
      const int32_t x = x0 + (int32_t)threadIdx.x;
      const bool in = x < B.lx;
      int64_t d = 0;
      bool verb = false;
      if (in) {
        const int64_t i = base + x;
        const bool hx = x > 0;
        auto K = [&](bool ok, int64_t j) -> int64_t { return ok ? (int64_t)(kq[j] >> 1) : 0; };
        const int32_t c = kq[i];
        verb = c & 1;
        const int64_t pred = K(hx, i - 1) + K(hy, i - nx) + K(hz, i - sxy) -
                             K(hx && hy, i - 1 - nx) - K(hx && hz, i - 1 - sxy) -
                             K(hy && hz, i - nx - sxy) + K(hx && hy && hz, i - 1 - nx - sxy);
        d = (int64_t)(c >> 1) - pred;
      }
      // the zero residual is most of a smooth field: one LDS atomic per wave
      const uint64_t zm = __ballot(in && d == 0);
      if (lane == 0 && zm) atomicAdd(&s_h[kDqRadius], (uint32_t)__popcll(zm));
      if (in && d != 0) {
        const int slot = (d > -kDqRadius && d < kDqRadius) ? (int)d + kDqRadius : kDqBins;
        atomicAdd(&s_h[slot], 1u);
      }
      if (in && verb) atomicAdd(&s_h[kDqBins + 1], 1u);
    }
    if (++y == B.ly) {
      y = 0;
      z++;
    }
  }
  __syncthreads();
  unsigned long long *h = hist + (size_t)lo * kDqSlots;
  for (int b = threadIdx.x; b < kDqSlots; b += blockDim.x)
    if (s_h[b]) atomicAdd(&h[b], (unsigned long long)s_h[b]);
}

// Pass 2 with 4 consecutive voxels of a row per thread, when every block's x
// offset and extent and the row pitch are multiples of 4: the quanta of the
// voxel, its x-1 neighbour and the three lower rows come as 4 vector loads and 4
// scalar loads per 4 voxels instead of 32 scalar loads.
__global__ __launch_bounds__(256) void k_dq_hist4(const int32_t *__restrict__ kq,
                                                  const DqBlock *__restrict__ blocks,
                                                  uint32_t nblocks, int64_t nx, int64_t sxy,
                                                  unsigned long long *__restrict__ hist) {
  __shared__ uint32_t s_h[kDqSlots];
  for (int b = threadIdx.x; b < kDqSlots; b += blockDim.x) s_h[b] = 0;
  uint32_t lo = 0, hi = nblocks - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (blocks[mid].chunk0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const DqBlock B = blocks[lo];
  __syncthreads();
  const int64_t rows = (int64_t)B.ly * B.lz;
  const int64_t r0 = (int64_t)(blockIdx.x - B.chunk0) * kDqRows;
  const int64_t r1 = r0 + kDqRows < rows ? r0 + kDqRows : rows;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t quads = (uint32_t)((r1 - r0) * B.lx / 4);  // groups of 4 in the chunk
  const uint32_t qrow = (uint32_t)B.lx / 4;
  // whole waves step together (the bound is uniform), so the ballots are uniform
  for (uint32_t q0 = 0; q0 < quads; q0 += blockDim.x) {
    const uint32_t q = q0 + threadIdx.x;
    const bool in = q < quads;
    int64_t d[4] = {0, 0, 0, 0};
    uint32_t verb = 0;
    if (in) {
      const uint32_t rr = q / qrow;
      const int32_t x = (int32_t)(q - rr * qrow) * 4;
      const uint32_t row = (uint32_t)r0 + rr;  // rows of a block < 2^32
      const int32_t y = (int32_t)(row % (uint32_t)B.ly), z = (int32_t)(row / (uint32_t)B.ly);
      const int64_t i = (int64_t)(B.z0 + z) * sxy + (int64_t)(B.y0 + y) * nx + B.x0 + x;
      const bool hx = x > 0, hy = y > 0, hz = z > 0;
      // rows: the voxel's (0), y-1 (1), z-1 (2), y-1 and z-1 (3); lane 0 of each is x-1
      int32_t v[4][5];
      const int64_t off[4] = {0, -nx, -sxy, -nx - sxy};
      const bool ok[4] = {true, hy, hz, hy && hz};
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int4 w = make_int4(0, 0, 0, 0);
        int32_t left = 0;
        if (ok[r]) {
          w = *reinterpret_cast<const int4 *>(kq + i + off[r]);
          if (hx) left = kq[i + off[r] - 1];
        }
        v[r][0] = left;
        v[r][1] = w.x;
        v[r][2] = w.y;
        v[r][3] = w.z;
        v[r][4] = w.w;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        auto K = [&](int r, int c) -> int64_t { return (int64_t)(v[r][c] >> 1); };
        const int c = u + 1;  // the voxel's column; c - 1 is its x-1 neighbour (0 at x = 0)
        const int64_t pred = K(0, c - 1) + K(1, c) + K(2, c) - K(1, c - 1) - K(2, c - 1) -
                             K(3, c) + K(3, c - 1);
        d[u] = K(0, c) - pred;
        verb |= (uint32_t)(v[0][c] & 1) << u;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      // the zero residual is most of a smooth field: one LDS atomic per wave
      const uint64_t zm = __ballot(in && d[u] == 0);
      if (lane == 0 && zm) atomicAdd(&s_h[kDqRadius], (uint32_t)__popcll(zm));
      if (in && d[u] != 0) {
        const int slot = (d[u] > -kDqRadius && d[u] < kDqRadius) ? (int)d[u] + kDqRadius : kDqBins;
        atomicAdd(&s_h[slot], 1u);
      }
      if (in && ((verb >> u) & 1u)) atomicAdd(&s_h[kDqBins + 1], 1u);
    }
  }
  __syncthreads();
  unsigned long long *h = hist + (size_t)lo * kDqSlots;
  for (int b = threadIdx.x; b < kDqSlots; b += blockDim.x)
    if (s_h[b]) atomicAdd(&h[b], (unsigned long long)s_h[b]);
}

}  // namespace pmsz
