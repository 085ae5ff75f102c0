// pmsz_naive.h -- kernels of the naive-MSz correction (naive_MSz.cu:1887-2219) on
// one unpadded domain: extrema-only sweeps and path re-tracing.
//
// Codes are those of the full-volume labeling kernel: per voxel, the rank of the
// steepest-ascent candidate (low nibble) and of the steepest-descent candidate
// (high nibble), rank 7 being the voxel itself.  f's codes stand for or_types
// (naive_MSz.cu:740-760), g's for de_direction_as / de_direction_ds
// (naive_MSz.cu:830-842).  A sweep marks its targets in a byte map on the
// start-of-sweep g; k_naive_apply then edits each marked voxel once (the edits[]
// atomicCAS dedup of the reference).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pmsz_common.h"
#include "pmsz_device.h"

namespace pmsz {

// counters of a naive step
enum { kNvMax = 0, kNvMin = 1, kNvDir = 2, kNvTargets = 3, kNvTouched = 4, kNvClamped = 5 };
// steps of a path walk before it is abandoned: a walk along f's steepest path
// never revisits a voxel, so only a malformed one reaches it
constexpr uint32_t kNaiveWalkCap = 1u << 20;

// The kernels below stride over the voxels with a grid of a few blocks per CU
// and add their counts once per block: one device atomic per wave on the same
// few counters serialises (about 7 ns each; 27 ms for one pass over 512^3).
__device__ __forceinline__ void naive_flush(unsigned *cnt, int k, unsigned v) {
  __shared__ unsigned s_c[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (__lane_id() == 0) s_c[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = s_c[0] + s_c[1] + s_c[2] + s_c[3];
    if (t) atomicAdd(&cnt[k], t);
  }
  __syncthreads();
}

// or_types of f at voxel c: slot 0 is -1 at a maximum, -2 at a minimum, else
// the steepest descent; slot 1 is the descent at a maximum, else the ascent.
__device__ __forceinline__ int64_t naive_f0(int64_t c, uint32_t code, const Grid3 &d) {
  if ((code & 15u) == (uint32_t)kSelf) return -1;
  if ((code >> 4) == (uint32_t)kSelf) return -2;
  return c + rank_off((int)(code >> 4), d);
}
__device__ __forceinline__ int64_t naive_f1(int64_t c, uint32_t code, const Grid3 &d) {
  if ((code & 15u) == (uint32_t)kSelf) return c + rank_off((int)(code >> 4), d);
  return c + rank_off((int)(code & 15u), d);
}

// False extrema of g against f (iscriticle without the direction fix,
// naive_MSz.cu:780-826).  A false or lost maximum lowers the voxel (g has a
// maximum f lacks) or g's steepest ascent (f's maximum is lost); a false or lost
// minimum lowers f's steepest descent or the voxel.  mark == nullptr: count only.
__global__ __launch_bounds__(256) void k_naive_extrema(const uint8_t *__restrict__ cf,
                                                       const uint8_t *__restrict__ cg, Grid3 d,
                                                       uint8_t *__restrict__ mark,
                                                       unsigned *__restrict__ cnt) {
  unsigned nmax = 0, nmin = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t a = cf[i], b = cg[i];
    const bool fmax = (a & 15u) == (uint32_t)kSelf;
    const bool fmin = !fmax && (a >> 4) == (uint32_t)kSelf;
    const bool gmax = (b & 15u) == (uint32_t)kSelf, gmin = (b >> 4) == (uint32_t)kSelf;
    const bool fm = gmax != fmax, fn = gmin != fmin;
    nmax += fm;
    nmin += fn;
    if (mark) {
      if (fm) mark[fmax ? i + rank_off((int)(b & 15u), d) : i] = 1;
      if (fn) mark[fmin ? i : i + rank_off((int)(a >> 4), d)] = 1;
    }
  }
  naive_flush(cnt, kNvMax, nmax);
  naive_flush(cnt, kNvMin, nmin);
}

// extract_false_path (naive_MSz.cu:1767-1864).  From a voxel whose descending
// label of g differs from f's, walk f's steepest descent while g's descent
// agrees; at the first voxel where they differ, lower f's descent target there
// (the true index).  Likewise for the ascending label along the ascent, lowering
// g's ascent target (the false index).  A walk that ends on one of f's minima
// would index edits[-2] in the reference (naive_MSz.cu:1806-1811); it edits
// nothing.  mark == nullptr: count the voxels with a differing label
// (extract_false_path with fix = 1).
__global__ __launch_bounds__(256) void k_naive_paths(
    const uint8_t *__restrict__ cf, const uint8_t *__restrict__ cg,
    const int32_t *__restrict__ fa, const int32_t *__restrict__ fd,
    const int32_t *__restrict__ ga, const int32_t *__restrict__ gd, Grid3 d,
    uint8_t *__restrict__ mark, unsigned *__restrict__ cnt) {
  unsigned ndir = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const bool ds = fd[i] != gd[i];
    const bool as = fa[i] != ga[i];
    if (!mark) {
      ndir += ds || as;
      continue;
    }
    ndir += (unsigned)ds + (unsigned)as;
    if (ds) {
      int64_t cur = i;
      uint32_t steps = 0;
      for (;;) {
        const int64_t nf = naive_f0(cur, cf[cur], d);
        const int64_t ng = cur + rank_off((int)(cg[cur] >> 4), d);
        if (nf != ng || nf == cur) break;
        cur = nf;
        if (++steps > kNaiveWalkCap) {
          cur = -1;
          break;
        }
      }
      if (cur >= 0) {
        const uint32_t a = cf[cur];
        const int64_t fi = cur + rank_off((int)(cg[cur] >> 4), d);
        const int64_t f0 = naive_f0(cur, a, d);
        const int64_t ti = f0 == -1 ? naive_f1(cur, a, d) : f0;
        if (ti >= 0 && fi != ti) mark[ti] = 1;
      }
    }
    if (as) {
      int64_t cur = i;
      uint32_t steps = 0;
      for (;;) {
        const int64_t nf = naive_f1(cur, cf[cur], d);
        const int64_t ng = cur + rank_off((int)(cg[cur] & 15u), d);
        if (nf != ng || nf == cur) break;
        cur = nf;
        if (++steps > kNaiveWalkCap) {
          cur = -1;
          break;
        }
      }
      if (cur >= 0) {
        const int64_t fi = cur + rank_off((int)(cg[cur] & 15u), d);
        const int64_t ti = naive_f1(cur, cf[cur], d);
        if (fi != ti) mark[fi] = 1;
      }
    }
  }
  naive_flush(cnt, kNvDir, ndir);
}

// One edit per marked voxel on its start-of-sweep state (applyDeltaBuffer_local,
// naive_MSz.cu:618-643), clearing the mark.
template <typename T>
__global__ __launch_bounds__(256) void k_naive_apply(const T *__restrict__ f, T *__restrict__ g,
                                                     uint8_t *__restrict__ dc,
                                                     uint8_t *__restrict__ mark, int64_t n,
                                                     EditParams ep, unsigned *__restrict__ cnt) {
  unsigned nt = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!mark[i]) continue;
    T gv = g[i];
    uint8_t c = dc[i];
    edit_value<T>(f[i], gv, c, ep);
    g[i] = gv;
    dc[i] = c;
    mark[i] = 0;
    nt++;
  }
  naive_flush(cnt, kNvTargets, nt);
}

// Voxels edited at least once and voxels clamped, from the delta counters.
__global__ __launch_bounds__(256) void k_naive_dc_stats(const uint8_t *__restrict__ dc, int64_t n,
                                                        int q, unsigned *__restrict__ cnt) {
  unsigned touched = 0, clamped = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = dc[i];
    touched += c != 0;
    clamped += c == (uint32_t)q;
  }
  naive_flush(cnt, kNvTouched, touched);
  naive_flush(cnt, kNvClamped, clamped);
}

}  // namespace pmsz
