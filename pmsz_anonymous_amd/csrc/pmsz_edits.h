// pmsz_edits.h -- the edit side channel on the device: per-voxel delta symbols
// packed at nbits bits and the clamp values compacted in voxel order, for each
// rank block of the field, in one pass over g0 and g.
//
// Reference: extract_edits_kernel (pMSz.cu:2090-2138) writes a symbol byte and
// a clamp value per core voxel; compact_edits_dense_to_sparse (pMSz.cu:227-250)
// keeps the nonzero clamp values in order (thrust::copy_if); pack_nbits
// (pMSz.cu:2328-2349) packs the symbols LSB first at bits_for_max(q) bits on the
// host.  Here one kernel does all three: a tile of 2048 voxels is 256 threads x 8
// symbols, and 8 symbols of nbits bits are exactly nbits bytes, so each thread
// owns whole bytes of the packed stream.  The clamp values of a tile are placed
// by a single-pass scan across tiles (decoupled look-back, Merrill & Garland
// 2016): tiles take their index from a ticket in launch order, publish their
// count, then add up their predecessors' counts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pmsz_common.h"

namespace pmsz {

constexpr int kEncThreads = 256;
constexpr int kEncPer = 8;  // symbols per group: nbits bytes of the stream
#ifndef PMSZ_ENC_GROUPS
#define PMSZ_ENC_GROUPS 2
#endif
#ifndef PMSZ_ENC_TICKET
#define PMSZ_ENC_TICKET 0
#endif
// groups per thread; group j of thread t is group j * 256 + t of the tile
constexpr int kEncGroups = PMSZ_ENC_GROUPS;
constexpr int kEncTile = kEncThreads * kEncPer * kEncGroups;
static_assert(kEncGroups >= 1 && kEncGroups <= 4, "16-bit count fields of one u64");
// look-back word: status in bits 63..62, count in bits 61..0
constexpr unsigned long long kLookAgg = 1ull << 62;   // the tile's own count
constexpr unsigned long long kLookIncl = 2ull << 62;  // count of tiles 0..k
constexpr unsigned long long kLookVal = (1ull << 62) - 1;
// polls of a predecessor's word before the tile gives up and flags an error; the
// predecessors are resident and never wait on this tile (see k_edit_encode):
// the bound only guards against a broken launch
constexpr uint32_t kLookSpin = 1u << 20;
constexpr uint32_t kEncErrSymbol = 1;  // a symbol exceeds nbits bits
constexpr uint32_t kEncErrLook = 2;    // look-back bound reached

// One rank block, in rank order.
struct EncBlock {
  uint32_t lx, ly;   // extents along x and y
  uint32_t n;        // voxels
  uint32_t tile0;    // first tile of the block
  int64_t base;      // global index of the block's first voxel
  int64_t pk_off;    // byte offset of its packed stream (16-byte aligned)
  int64_t pk_bytes;  // ceil(n * nbits / 8)
};

__device__ inline unsigned long long enc_look_load(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void enc_look_store(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// clamp_value with its search out of line: the search runs only where f - eb
// in T breaks the bound, and inlined at each of a thread's voxels it would
// multiply the kernel's code many times over.
template <typename T>
__device__ __noinline__ T clamp_search_call(T fv, EditParams ep) {
  return clamp_search(fv, ep);
}
template <typename T>
__device__ inline T clamp_value_lean(T fv, const EditParams &ep) {
  T c = fv - (T)ep.eb;
  if (ep.clamp_mode == 1 && c < fv && !within_exact((double)fv, (double)c, ep.eb))
    c = clamp_search_call(fv, ep);
  return c;
}

// Tile order.  With the ticket, a tile's index is the order in which its
// workgroup started, so every predecessor is resident or done.  Without it the
// index is blockIdx.x: workgroups are dispatched in increasing index order
// (within each XCD), so the lowest unfinished tile is always resident and has
// all its predecessors done -- the scan cannot stall either way.
//
// kFlat: one block covering the whole field and f, g0, g 16-byte aligned, so a
// group's 8 voxels are contiguous and aligned.
template <typename T, bool kFlat>
__global__ __launch_bounds__(kEncThreads) void k_edit_encode(
    const T *__restrict__ f, const T *__restrict__ g0, const T *__restrict__ g,
    const EncBlock *__restrict__ blocks, uint32_t nblocks, int64_t nx, int64_t sxy, int nbits,
    EditParams ep, uint8_t *__restrict__ packed, T *__restrict__ vals,
    unsigned long long *__restrict__ look, uint32_t *__restrict__ tile_edited,
    uint32_t *__restrict__ ticket, uint32_t *__restrict__ err) {
  __shared__ uint8_t s_pk[kEncTile];
  __shared__ unsigned long long s_wc[kEncThreads / 64];
  __shared__ uint32_t s_we[kEncThreads / 64];
  __shared__ uint32_t s_tile;
  __shared__ unsigned long long s_prefix;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  uint32_t tile;
  if (PMSZ_ENC_TICKET) {
    if (t == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    tile = s_tile;
  } else {
    tile = blockIdx.x;
  }
  uint32_t lo = 0, hi = nblocks - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (blocks[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const EncBlock B = blocks[lo];
  const uint32_t tl = tile - B.tile0;

  const T delta = (T)ep.delta_d;
  const uint32_t lim = (1u << nbits) - 1u;
  bool bad = false;
  uint32_t nedit = 0;
  uint32_t cmask[kEncGroups];
  T cv[kEncGroups][kEncPer];
  unsigned long long counts = 0;  // clamp values of group j in bits 16j..16j+15
#pragma unroll
  for (int j = 0; j < kEncGroups; j++) {
    const uint32_t grp = (uint32_t)j * kEncThreads + t;
    const uint32_t l0 = tl * (uint32_t)kEncTile + grp * (uint32_t)kEncPer;
    T va[kEncPer], vb[kEncPer], vf[kEncPer];
    int64_t idx[kEncPer];
    if (kFlat && l0 + kEncPer <= B.n) {
      constexpr int kV = 16 / sizeof(T);
      using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
#pragma unroll
      for (int h = 0; h < kEncPer / kV; h++) {
        const V a = *reinterpret_cast<const V *>(g + l0 + h * kV);
        const V b = *reinterpret_cast<const V *>(g0 + l0 + h * kV);
        // f where g < g0 only, but those voxels touch nearly every line of f:
        // loading it with g and g0 keeps one memory latency per group
        const V c = *reinterpret_cast<const V *>(f + l0 + h * kV);
        const T *pa = reinterpret_cast<const T *>(&a), *pb = reinterpret_cast<const T *>(&b);
        const T *pc = reinterpret_cast<const T *>(&c);
#pragma unroll
        for (int k = 0; k < kV; k++) {
          va[h * kV + k] = pa[k];
          vb[h * kV + k] = pb[k];
          vf[h * kV + k] = pc[k];
        }
      }
#pragma unroll
      for (int k = 0; k < kEncPer; k++) idx[k] = (int64_t)l0 + k;
    } else {
      const uint32_t sl = B.lx * B.ly;
#pragma unroll
      for (int k = 0; k < kEncPer; k++) {
        const uint32_t l = l0 + k;
        va[k] = vb[k] = vf[k] = (T)0;
        idx[k] = -1;
        if (l < B.n) {
          int64_t i;
          if (kFlat) {
            i = (int64_t)l;
          } else {
            const uint32_t z = l / sl, r = l - z * sl, y = r / B.lx, x = r - y * B.lx;
            i = B.base + (int64_t)z * sxy + (int64_t)y * nx + x;
          }
          idx[k] = i;
          va[k] = g[i];
          vb[k] = g0[i];
          vf[k] = f[i];
        }
      }
    }
    unsigned long long bits = 0;
    cmask[j] = 0;
#pragma unroll
    for (int k = 0; k < kEncPer; k++) {
      cv[j][k] = (T)0;
      uint32_t s = 0;
      const T a = va[k], b = vb[k];
      if (idx[k] >= 0 && a != b && a < b) {
        nedit++;
        const T c = clamp_value_lean<T>(vf[k], ep);
        if (a != c) {
          s = (uint8_t)(int)std::round((b - a) / delta);
        } else {
          s = (uint32_t)ep.q;
          // copy_if keeps the nonzero values only (is_nonzero, pMSz.cu:223-225)
          if (c != (T)0) {
            cv[j][k] = c;
            cmask[j] |= 1u << k;
          }
        }
      }
      bad |= s > lim;
      bits |= (unsigned long long)(s & lim) << (k * nbits);
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k < nbits) s_pk[grp * nbits + k] = (uint8_t)(bits >> (8 * k));
    counts |= (unsigned long long)__popc(cmask[j]) << (16 * j);
  }
  if (bad) atomicOr(err, kEncErrSymbol);

  // tile-wide exclusive scan of the clamp counts, all groups at once in 16-bit
  // fields (a field's total is at most 2048), and the edited count
  unsigned long long inc = counts;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long u = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += u;
  }
  uint32_t ne = nedit;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ne += __shfl_xor(ne, o, 64);
  if (lane == 63) s_wc[wave] = inc;
  if (lane == 0) s_we[wave] = ne;
  __syncthreads();
  unsigned long long wbase = 0, tot = 0;
  uint32_t edited = 0;
#pragma unroll
  for (int w = 0; w < kEncThreads / 64; w++) {
    if ((uint32_t)w < wave) wbase += s_wc[w];
    tot += s_wc[w];
    edited += s_we[w];
  }
  uint32_t agg = 0;
#pragma unroll
  for (int j = 0; j < kEncGroups; j++) agg += (uint32_t)(tot >> (16 * j)) & 0xffffu;

  if (wave == 0) {
    unsigned long long prefix = 0;
    if (tile == 0) {
      if (lane == 0) enc_look_store(&look[0], kLookIncl | agg);
    } else {
      if (lane == 0) enc_look_store(&look[tile], kLookAgg | agg);
      int64_t top = (int64_t)tile - 1;
      for (;;) {
        const int64_t j = top - (int64_t)lane;
        // before tile 0 the count is 0 and final
        unsigned long long w = j >= 0 ? enc_look_load(&look[j]) : kLookIncl;
        uint32_t spins = 0;
        for (;;) {
          const bool wait = (w >> 62) == 0;
          if (!__any(wait)) break;
          if (++spins > kLookSpin) {
            if (lane == 0) atomicOr(err, kEncErrLook);
            if (wait) w = kLookIncl;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          if (wait) w = enc_look_load(&look[j]);
        }
        const unsigned long long incl = __ballot((w >> 62) == 2);
        const uint32_t first = incl ? (uint32_t)__ffsll((long long)incl) - 1u : 64u;
        prefix += wave_sum_u64(lane <= first ? (w & kLookVal) : 0ull);
        if (incl) break;
        top -= 64;
      }
      if (lane == 0) enc_look_store(&look[tile], kLookIncl | (prefix + agg));
    }
    if (lane == 0) {
      s_prefix = prefix;
      tile_edited[tile] = edited;
    }
  }
  __syncthreads();

  // packed bytes of the tile: whole words, then the tail of the stream
  const int64_t tb = (int64_t)tl * (kEncTile / 8) * nbits;
  const uint32_t nb = (uint32_t)std::min<int64_t>(B.pk_bytes - tb, (int64_t)(kEncTile / 8) * nbits);
  uint8_t *dst = packed + B.pk_off + tb;
  for (uint32_t w = t; w < nb / 4; w += kEncThreads)
    reinterpret_cast<uint32_t *>(dst)[w] = reinterpret_cast<const uint32_t *>(s_pk)[w];
  for (uint32_t b = (nb & ~3u) + t; b < nb; b += kEncThreads) dst[b] = s_pk[b];

  // clamp values, in voxel order: group j of every thread after group j - 1
  const unsigned long long excl = wbase + inc - counts;
  uint32_t before = 0;
#pragma unroll
  for (int j = 0; j < kEncGroups; j++) {
    if (cmask[j]) {
      T *out = vals + s_prefix + before + ((uint32_t)(excl >> (16 * j)) & 0xffffu);
#pragma unroll
      for (int k = 0; k < kEncPer; k++)
        if (cmask[j] & (1u << k)) *out++ = cv[j][k];
    }
    before += (uint32_t)(tot >> (16 * j)) & 0xffffu;
  }
}

// Totals after k_edit_encode, out[] zeroed beforehand: out[0] = edited voxels,
// out[1 + b] = clamp values of block b, out[1 + nblocks] != 0 when some tile did
// not finish its scan.
__global__ __launch_bounds__(256) void k_edit_totals(const EncBlock *__restrict__ blocks,
                                                     uint32_t nblocks, uint32_t ntiles,
                                                     const unsigned long long *__restrict__ look,
                                                     const uint32_t *__restrict__ tile_edited,
                                                     unsigned long long *__restrict__ out) {
  const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  unsigned long long e = 0;
  bool open = false;
  for (uint32_t k = gt; k < ntiles; k += stride) {
    e += tile_edited[k];
    open |= (look[k] >> 62) != 2;
  }
  e = wave_sum_u64(e);
  const bool any_open = __any(open);
  if ((threadIdx.x & 63u) == 0) {
    if (e) atomicAdd(&out[0], e);
    if (any_open) atomicOr(&out[1 + nblocks], 1ull);
  }
  for (uint32_t b = gt; b < nblocks; b += stride) {
    const uint32_t last = (b + 1 < nblocks ? blocks[b + 1].tile0 : ntiles) - 1;
    const unsigned long long hi = look[last] & kLookVal;
    const unsigned long long lo = b ? (look[blocks[b].tile0 - 1] & kLookVal) : 0ull;
    out[1 + b] = hi - lo;
  }
}

}  // namespace pmsz
