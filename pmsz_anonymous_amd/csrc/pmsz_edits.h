// pmsz_edits.h -- the edit side channel on the device: per-voxel delta symbols
// packed at nbits bits and the clamp values compacted in voxel order, for each
// rank block of the field, in one pass over g0 and g.
//
// Reference: extract_edits_kernel (pMSz.cu:2090-2138) writes a symbol byte and
// a clamp value per core voxel; compact_edits_dense_to_sparse (pMSz.cu:227-250)
// keeps the nonzero clamp values in order (thrust::copy_if); pack_nbits
// (pMSz.cu:2328-2349) packs the symbols LSB first at bits_for_max(q) bits on the
// host.  Here one kernel does all three: a tile is 256 threads x 8-symbol
// groups, and 8 symbols of nbits bits are exactly nbits bytes, so each thread
// owns whole bytes of the packed stream.  The clamp values of a tile are placed
// by a single-pass scan across tiles (decoupled look-back, Merrill & Garland
// 2016): each tile publishes its count, then adds up its predecessors' counts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pmsz_common.h"

namespace pmsz {

constexpr int kEncThreads = 256;
constexpr int kEncPer = 8;  // symbols per group: nbits bytes of the stream
// groups per thread (group j of thread t is group j * 256 + t of the tile): two
// for float, one for double, so a thread's g, g0 and f stay in registers
#ifndef PMSZ_ENC_GROUPS_F32
#define PMSZ_ENC_GROUPS_F32 2
#endif
template <typename T>
constexpr int enc_groups() {
  return sizeof(T) == 4 ? PMSZ_ENC_GROUPS_F32 : 1;
}
template <typename T>
constexpr int enc_tile() {
  return kEncThreads * kEncPer * enc_groups<T>();
}
// lowered voxels (g < g0) go through a per-wave queue in rounds of 4 slots of
// every lane, at most 256 entries
constexpr int kEncRound = 4;
constexpr int kEncQueue = 64 * kEncRound;
// look-back word: status in bits 63..62, count in bits 61..0
constexpr unsigned long long kLookAgg = 1ull << 62;   // the tile's own count
constexpr unsigned long long kLookIncl = 2ull << 62;  // count of tiles 0..k
constexpr unsigned long long kLookVal = (1ull << 62) - 1;
// polls of a predecessor's word before the tile gives up and flags an error; the
// predecessors are resident or done and never wait on this tile (see
// k_edit_encode): the bound only guards against a broken launch
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

// Tile order: a tile's index is blockIdx.x.  Workgroups are dispatched in
// increasing index order (within each XCD), so the lowest unfinished tile is
// resident and has all its predecessors done: the scan cannot stall.
//
// Per group of 8 voxels a thread holds g, g0 and f.  The voxels with g < g0
// (about one in seven after a correction) need the clamp test and a division;
// rather than run that path for every slot of every lane, each wave queues its
// lowered voxels in LDS, works the queue with all 64 lanes, and hands the
// results back by queue position.
//
// kFlat: one block covering the whole field and f, g0, g 16-byte aligned, so a
// group's 8 voxels are contiguous and aligned.  Otherwise a group is read with
// vector loads where it lies in one row and is aligned (aligned: f, g0 and g are
// 16-byte aligned), else voxel by voxel.
template <typename T, bool kFlat>
__global__ __launch_bounds__(kEncThreads) void k_edit_encode(
    const T *__restrict__ f, const T *__restrict__ g0, const T *__restrict__ g,
    const EncBlock *__restrict__ blocks, uint32_t nblocks, int64_t nx, int64_t sxy, int nbits,
    bool aligned, EditParams ep, uint8_t *__restrict__ packed, T *__restrict__ vals,
    unsigned long long *__restrict__ look, uint32_t *__restrict__ tile_edited,
    uint32_t *__restrict__ err) {
  constexpr int G = enc_groups<T>();
  constexpr int kTile = enc_tile<T>();
  constexpr int S = G * kEncPer;  // slots of a thread
  static_assert(G <= 4, "16-bit count fields of one u64");
  __shared__ uint8_t s_pk[kTile];
  __shared__ T q_a[kEncThreads / 64][kEncQueue], q_d[kEncThreads / 64][kEncQueue],
      q_f[kEncThreads / 64][kEncQueue];
  __shared__ uint16_t q_r[kEncThreads / 64][kEncQueue];
  __shared__ unsigned long long s_wc[kEncThreads / 64];
  __shared__ uint32_t s_we[kEncThreads / 64];
  __shared__ unsigned long long s_prefix;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t tile = blockIdx.x;
  uint32_t lo = 0, hi = nblocks - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (blocks[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const EncBlock B = blocks[lo];
  const uint32_t tl = tile - B.tile0;

  // g, g0, f of the thread's slots; slot j * 8 + k is voxel k of group j
  T va[S], vb[S], vf[S];
  uint32_t low = 0;  // slots with g < g0
#pragma unroll
  for (int j = 0; j < G; j++) {
    const uint32_t grp = (uint32_t)j * kEncThreads + t;
    const uint32_t l0 = tl * (uint32_t)kTile + grp * (uint32_t)kEncPer;
    // the group's global index where its 8 voxels are contiguous and 16-byte
    // aligned in f, g0 and g; -1 otherwise
    int64_t vbase = -1;
    uint32_t x = 0, y = 0;
    int64_t row = 0;
    if (kFlat) {
      if (l0 + kEncPer <= B.n) vbase = l0;
    } else if (l0 < B.n) {
      const uint32_t sl = B.lx * B.ly, z = l0 / sl, r = l0 - z * sl;
      y = r / B.lx;
      x = r - y * B.lx;
      row = B.base + (int64_t)z * sxy + (int64_t)y * nx;
      if (aligned && x + kEncPer <= B.lx && ((row + x) & (16 / (int)sizeof(T) - 1)) == 0)
        vbase = row + x;
    }
    if (vbase >= 0) {
      constexpr int kV = 16 / sizeof(T);
      using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
#pragma unroll
      for (int h = 0; h < kEncPer / kV; h++) {
        const V a = *reinterpret_cast<const V *>(g + vbase + h * kV);
        const V b = *reinterpret_cast<const V *>(g0 + vbase + h * kV);
        // f is needed where g < g0 only, but those voxels touch nearly every
        // line of f: loading it with g and g0 keeps one memory latency
        const V c = *reinterpret_cast<const V *>(f + vbase + h * kV);
        const T *pa = reinterpret_cast<const T *>(&a), *pb = reinterpret_cast<const T *>(&b);
        const T *pc = reinterpret_cast<const T *>(&c);
#pragma unroll
        for (int k = 0; k < kV; k++) {
          va[j * kEncPer + h * kV + k] = pa[k];
          vb[j * kEncPer + h * kV + k] = pb[k];
          vf[j * kEncPer + h * kV + k] = pc[k];
        }
      }
    } else {
      // from (x, y, z) of the group's first voxel, step along x with wrap
#pragma unroll
      for (int k = 0; k < kEncPer; k++) {
        const uint32_t l = l0 + k;
        T a = (T)0, b = (T)0, c = (T)0;
        if (l < B.n) {
          const int64_t i = kFlat ? (int64_t)l : row + x;
          a = g[i];
          b = g0[i];
          c = f[i];
        }
        va[j * kEncPer + k] = a;
        vb[j * kEncPer + k] = b;
        vf[j * kEncPer + k] = c;
        if (++x == B.lx) {
          x = 0;
          row += nx;
          if (++y == B.ly) {
            y = 0;
            row += sxy - (int64_t)B.ly * nx;
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < S; k++)
    if (va[k] != vb[k] && va[k] < vb[k]) low |= 1u << k;  // 0 outside the block: equal

  // symbol and clamp flag of each lowered slot, through the wave's queue
  uint32_t sym[S];
  uint32_t cflag = 0;
  const T delta = (T)ep.delta_d;
#pragma unroll
  for (int r0 = 0; r0 < S; r0 += kEncRound) {
    const uint32_t mine = (low >> r0) & ((1u << kEncRound) - 1u);
    const uint32_t cnt = __popc(mine);
    uint32_t inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = __shfl_up(inc, o, 64);
      if (lane >= (uint32_t)o) inc += u;
    }
    const uint32_t nq = __shfl(inc, 63, 64);
    uint32_t pos = inc - cnt;
#pragma unroll
    for (int k = 0; k < kEncRound; k++)
      if (mine & (1u << k)) {
        q_a[wave][pos] = va[r0 + k];
        q_d[wave][pos] = vb[r0 + k] - va[r0 + k];
        q_f[wave][pos] = vf[r0 + k];
        pos++;
      }
    __syncthreads();
    for (uint32_t e = lane; e < nq; e += 64) {
      const T a = q_a[wave][e];
      const T c = clamp_value_lean<T>(q_f[wave][e], ep);
      uint32_t r;
      if (a != c) r = (uint8_t)(int)std::round(q_d[wave][e] / delta);
      else r = 0x100u | (uint32_t)ep.q;
      q_r[wave][e] = (uint16_t)r;
    }
    __syncthreads();
    pos = inc - cnt;
#pragma unroll
    for (int k = 0; k < kEncRound; k++) {
      uint32_t r = 0;
      if (mine & (1u << k)) r = q_r[wave][pos++];
      sym[r0 + k] = r & 0xffu;
      // copy_if keeps the nonzero clamp values only (is_nonzero, pMSz.cu:223-225);
      // a clamped voxel's value is its g
      if ((r & 0x100u) && va[r0 + k] != (T)0) cflag |= 1u << (r0 + k);
    }
  }

  const uint32_t lim = (1u << nbits) - 1u;
  bool bad = false;
  unsigned long long counts = 0;  // kept clamp values of group j in bits 16j..16j+15
#pragma unroll
  for (int j = 0; j < G; j++) {
    const uint32_t grp = (uint32_t)j * kEncThreads + t;
    unsigned long long bits = 0;
#pragma unroll
    for (int k = 0; k < kEncPer; k++) {
      const uint32_t sv = sym[j * kEncPer + k];
      bad |= sv > lim;
      bits |= (unsigned long long)(sv & lim) << (k * nbits);
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k < nbits) s_pk[grp * nbits + k] = (uint8_t)(bits >> (8 * k));
    counts |= (unsigned long long)__popc((cflag >> (j * kEncPer)) & 0xffu) << (16 * j);
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
  uint32_t ne = __popc(low);
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
  for (int j = 0; j < G; j++) agg += (uint32_t)(tot >> (16 * j)) & 0xffffu;

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
  const int64_t tb = (int64_t)tl * (kTile / 8) * nbits;
  const uint32_t nb = (uint32_t)std::min<int64_t>(B.pk_bytes - tb, (int64_t)(kTile / 8) * nbits);
  uint8_t *dst = packed + B.pk_off + tb;
  for (uint32_t w = t; w < nb / 4; w += kEncThreads)
    reinterpret_cast<uint32_t *>(dst)[w] = reinterpret_cast<const uint32_t *>(s_pk)[w];
  for (uint32_t b = (nb & ~3u) + t; b < nb; b += kEncThreads) dst[b] = s_pk[b];

  // clamp values, in voxel order: group j of every thread after group j - 1
  const unsigned long long excl = wbase + inc - counts;
  uint32_t before = 0;
#pragma unroll
  for (int j = 0; j < G; j++) {
    const uint32_t cm = (cflag >> (j * kEncPer)) & 0xffu;
    if (cm) {
      T *out = vals + s_prefix + before + ((uint32_t)(excl >> (16 * j)) & 0xffffu);
#pragma unroll
      for (int k = 0; k < kEncPer; k++)
        if (cm & (1u << k)) *out++ = va[j * kEncPer + k];
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
