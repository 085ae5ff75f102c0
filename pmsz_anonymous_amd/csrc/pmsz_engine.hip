// pmsz_engine.hip -- gfx950 kernels and the C-ABI (include/pmsz.h) of the
// single-domain correction path.
//
// Data in HBM (N = nx*ny*nz voxels, s = sizeof(T)):
//   f, g        T[N]        caller-owned original / working field, x fastest
//   code        u8[N]       original steepest-ascent rank (low nibble) and
//                           steepest-descent rank (high nibble); rank 7 = the voxel
//   dc          u8[N]       delta counters (delta_counter, pMSz.cu:173)
//   tmark       u32[N/32]   targets of the running sweep: the per-sweep edit dedup
//                           (edits[] + atomicCAS, pMSz.cu:795) as one bit per voxel
//   wmark       u32[N/32]   voxels queued for the next sweep
//   tlist/wlist u32[N]      the same two sets as compact work lists
//   ring        Counters[]  ring of per-sweep counter blocks
//   run         RunState    phase hand-over between kernels, run totals
//
// Snapshot semantics: a sweep detects on the start-of-sweep g, writing only
// marks, lists and counters; a separate phase then applies one edit per target
// (applyDeltaBuffer_local, pMSz.cu:623-645).
//
// Schedule.  Sweep 1 is one full pass fused with the original labeling.  After
// that, the apply of sweep j decides on the device how sweep j+1 detects:
//   - over every voxel (tile kernel) while the targets are dense,
//   - over the closed stencil neighbourhoods of sweep j's targets otherwise: a
//     voxel's classification depends only on g over its closed neighbourhood and
//     the stencil is symmetric, so no other voxel can change -- the counters and
//     targets equal those of a full sweep;
//   - in a single-workgroup tail kernel once the targets are few, where a sweep
//     costs two workgroup barriers instead of two kernel launches.
// The host enqueues the whole chain without waiting and synchronises once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/pmsz.h"
#include "pmsz_dense.h"
#include "pmsz_device.h"

namespace pmsz {

constexpr int kTX = 64;  // dense tile: 64 x 8 voxels per plane, 16 planes
constexpr int kTY = 8;
constexpr int kTZ = 16;
constexpr int kThreads = 256;  // each thread classifies 2 rows of the tile
constexpr int kPX = kTX + 2, kPY = kTY + 2, kPZ = kTZ + 2;  // tile + halo
constexpr int kPlane = kPX * kPY;                             // cells of one staged plane
constexpr int kEPT = (kPlane + kThreads - 1) / kThreads;      // staged cells per thread
constexpr int kPlaneP = kEPT * kThreads;                       // padded slot size
constexpr int kSlots = 4;                                     // LDS plane ring
constexpr int kRows = kPY * kPZ;                              // bitmap rows of 66 cells
constexpr int kFlushPT = (kRows * 4 + kThreads - 1) / kThreads;
constexpr int kApplyThreads = 256;
constexpr int kTailThreads = 1024;
constexpr uint32_t kTailTargets = 2048;  // the tail kernel takes over at or below this

struct Sched {
  int mode;  // PMSZ_LAZY / PMSZ_SYNC
  uint32_t max_sweeps;
  uint32_t dense_min_work;  // 15 * targets at or above this -> next detection dense
};

// ---------------------------------------------------------------------------
// Tile kernel: with kLabel the original labeling of f (codes), with kDetect a
// detection sweep on g (sweep 1 when both are set; a later dense sweep reading
// the codes when only kDetect is set).
//
// Persistent blocks walk 64 x 8 x 16 tiles.  Blocks are dealt to the 8 XCDs
// round-robin by workgroup id, so block b takes tiles from the b % 8 -th
// contiguous eighth of the tile space: z- and y-neighbouring tiles, whose halos
// overlap, then share one XCD's L2.  A tile walks its planes through a ring of
// four LDS slots: while plane z is classified, plane z+2 is already in LDS and
// plane z+3 is in flight to registers, so one barrier per plane suffices and the
// loads overlap the compute.  Targets are collected in an LDS bitmap over the
// tile and its halo and flushed once per tile as row masks (one global atomicOr
// per 32 cells, one list reservation per block).
template <typename T, bool kLabel, bool kDetect>
__global__ __launch_bounds__(kThreads) void k_tiles(
    const T *__restrict__ f, const T *__restrict__ g, uint8_t *__restrict__ code,
    uint8_t *__restrict__ dc, uint32_t *__restrict__ tmark, uint32_t *__restrict__ tlist,
    Counters *__restrict__ ring, uint32_t jslot_prev, uint32_t jslot, RunState *run, Grid3 d) {
  if (!kLabel) {
    if (ring[jslot_prev].mode != kModeDense) return;  // uniform
  }
  constexpr int kFields = (kLabel ? 1 : 0) + (kDetect ? 1 : 0);
  constexpr int kF = 0, kG = kLabel ? 1 : 0;
  constexpr int kRowsPT = kTY / (kThreads / kTX);  // tile rows per thread
  constexpr bool kKeys = sizeof(T) == 4;  // float32: keys, float64: rank scan
  __shared__ double sv[kFields][kSlots][kPlaneP];
  __shared__ uint32_t sbits[kDetect ? kRows : 1][3];
  const int tid = threadIdx.x;
  const int tx = tid % kTX, ty0 = tid / kTX;
  // all voxel indices fit in 32 bits (checked on the host)
  const uint32_t nx = (uint32_t)d.nx, ny = (uint32_t)d.ny, nz = (uint32_t)d.nz;
  const uint32_t sxy = (uint32_t)d.sxy;
  const uint32_t ntx = (nx + kTX - 1) / kTX, nty = (ny + kTY - 1) / kTY, ntz = (nz + kTZ - 1) / kTZ;
  const uint32_t ntiles = ntx * nty * ntz;
  const uint32_t per_xcd_blocks = gridDim.x / 8;
  const uint32_t chunk = (ntiles + 7) / 8;
  const uint32_t t_begin = (blockIdx.x % 8) * chunk;
  const uint32_t t_end = t_begin + chunk < ntiles ? t_begin + chunk : ntiles;
  Counters *out = kDetect ? &ring[jslot] : nullptr;
  uint32_t lc[3] = {0, 0, 0};
  uint32_t lerr = 0;

  for (uint32_t tile = t_begin + blockIdx.x / 8; tile < t_end; tile += per_xcd_blocks) {
    const uint32_t tzi = tile / (ntx * nty), rem = tile - tzi * ntx * nty;
    const uint32_t tyi = rem / ntx, txi = rem - tyi * ntx;
    const uint32_t bx = txi * kTX, by = tyi * kTY, bz = tzi * kTZ;
    const uint32_t x = bx + tx;

    if (kDetect)
      for (int e = tid; e < kRows * 3; e += kThreads) (&sbits[0][0])[e] = 0;

    // staged cells of this thread: in-plane offset and in-plane validity, fixed
    // for the whole tile; loads are unconditional (clamped to cell 0) so the
    // compiler can issue them back to back
    uint32_t soff[kEPT];
    bool sok[kEPT];
#pragma unroll
    for (int q = 0; q < kEPT; q++) {
      const int e = tid + q * kThreads;
      const int lx = e % kPX, ly = e / kPX;
      const int gx = (int)bx + lx - 1, gy = (int)by + ly - 1;
      sok[q] = e < kPlane && gx >= 0 && gx < (int)nx && gy >= 0 && gy < (int)ny;
      soff[q] = sok[q] ? (uint32_t)gx + nx * (uint32_t)gy : 0u;
    }
    // out-of-domain cells hold whatever cell 0 holds: no candidate reads them
    // (interior voxels have every neighbour in the domain, boundary voxels use the
    // validity mask), so the staged values need no select
    T pf[kEPT], pg[kEPT];
    uint32_t pok = 0;  // bit q: staged cell q lies in the domain
    auto prefetch = [&](int pz) {
      const bool zok = pz >= 0 && pz < (int)nz;
      const uint32_t zoff = zok ? sxy * (uint32_t)pz : 0u;
      pok = 0;
#pragma unroll
      for (int q = 0; q < kEPT; q++) {
        const bool ok = zok && sok[q];
        pok |= (ok ? 1u : 0u) << q;
        const uint32_t gi = ok ? zoff + soff[q] : 0u;
        if (kLabel) pf[q] = f[gi];
        if (kDetect) pg[q] = g[gi];
      }
    };
    auto store = [&](int slot) {
#pragma unroll
      for (int q = 0; q < kEPT; q++) {
        const int e = tid + q * kThreads;  // slots are padded to kEPT * kThreads cells
        if (kLabel) {
          const T v = pf[q];
          if (((pok >> q) & 1u) && (!(v == v) || v == sentinel<T>())) lerr |= 1u;
          sv[kF][slot][e] = stage_value(v);
        }
        if (kDetect) {
          const T v = pg[q];
          if (kLabel && ((pok >> q) & 1u) && (!(v == v) || v == sentinel<T>())) lerr |= 1u;
          sv[kG][slot][e] = stage_value(v);
        }
      }
    };
    // codes of this thread's voxels, one plane ahead (dense sweeps after the first)
    uint8_t cpre[kRowsPT];
    auto prefetch_codes = [&](int pz) {
#pragma unroll
      for (int h = 0; h < kRowsPT; h++) {
        const uint32_t y = by + ty0 + h * (kThreads / kTX);
        const bool ok = x < nx && y < ny && pz < (int)nz;
        const uint8_t c = code[ok ? x + nx * y + sxy * (uint32_t)pz : 0u];
        cpre[h] = c;
      }
    };
    prefetch((int)bz - 1);
    store(0);
    prefetch((int)bz);
    store(1);
    prefetch((int)bz + 1);
    store(2);
    prefetch((int)bz + 2);
    if (!kLabel) prefetch_codes((int)bz);
    lds_barrier();
    const int kz = (int)((nz - bz) < (uint32_t)kTZ ? (nz - bz) : (uint32_t)kTZ);
    for (int k = 0; k < kz; k++) {
      const uint32_t z = bz + k;
      store((k + 3) % kSlots);  // plane z+2
      prefetch((int)z + 3);
      uint8_t ccur[kRowsPT];
      if (!kLabel) {
#pragma unroll
        for (int h = 0; h < kRowsPT; h++) ccur[h] = cpre[h];
        prefetch_codes((int)z + 1);
      }
      const int s0 = k % kSlots, s1 = (k + 1) % kSlots, s2 = (k + 2) % kSlots;
#pragma unroll
      for (int h = 0; h < kRowsPT; h++) {
        const int ty = ty0 + h * (kThreads / kTX);
        const uint32_t y = by + ty;
        if (x < nx && y < ny) {
          const uint32_t i = x + nx * y + sxy * z;
          const bool interior =
              x > 0 && x + 1 < nx && y > 0 && y + 1 < ny && z > 0 && z + 1 < nz;
          const uint32_t vm = interior ? 0x7fffu : valid_mask(x, y, z, d);
          int oa, od;
          auto cell = [&](int r) {
            const int sl = rank_dz(r) < 0 ? s0 : (rank_dz(r) == 0 ? s1 : s2);
            return sl * kPlaneP + (ty + 1 + rank_dy(r)) * kPX + tx + 1 + rank_dx(r);
          };
          auto classify = [&](int field, int &a, int &dsc) {
            double st[kCand];
#pragma unroll
            for (int r = 0; r < kCand; r++) st[r] = (&sv[field][0][0])[cell(r)];
            if (kKeys) {
              classify_keys(st, interior, vm, a, dsc);
            } else if (interior) {
              scan_full(st, a, dsc);
            } else {
              scan_masked(st, vm, a, dsc);
            }
          };
          if (kLabel) {
            classify(kF, oa, od);
            code[i] = (uint8_t)(oa | (od << 4));
            if (kDetect) dc[i] = 0;
          } else {
            oa = ccur[h] & 15;
            od = ccur[h] >> 4;
          }
          if (kDetect) {
            int da, dd;
            classify(kG, da, dd);
            int tr[4];
            const int nt = detect_targets(oa, od, da, dd, tr, lc);
            for (int q = 0; q < nt; q++) {
              const int r = tr[q];
              const int lx = tx + 1 + rank_dx(r), ly = ty + 1 + rank_dy(r),
                        lz = k + 1 + rank_dz(r);
              atomicOr(&sbits[lz * kPY + ly][lx >> 5], 1u << (lx & 31));
            }
          }
        }
      }
      lds_barrier();
    }
    if (kDetect) {
      // flush: (row, word) items, kRows rows x 4 global words
      uint32_t claimed[kFlushPT];
      uint32_t wbase[kFlushPT];
      uint32_t n = 0;
#pragma unroll
      for (int q = 0; q < kFlushPT; q++) {
        claimed[q] = 0;
        wbase[q] = 0;
        const int item = tid + q * kThreads;
        if (item >= kRows * 4) continue;
        const int row = item >> 2, kw = item & 3;
        const int ly = row % kPY, lz = row / kPY;
        const int gy = (int)by + ly - 1, gz = (int)bz + lz - 1;
        if (gy < 0 || gy >= (int)ny || gz < 0 || gz >= (int)nz) continue;
        // linear index of cell lx = 0 (may be -1 for the very first row)
        const int64_t i0 = (int64_t)bx - 1 + (int64_t)nx * gy + (int64_t)sxy * gz;
        const int64_t w0 = i0 >> 5;  // floor
        const int sh = (int)(i0 - (w0 << 5));
        const uint32_t m = bits32_at(sbits[row][0], sbits[row][1], sbits[row][2], 32 * kw - sh);
        if (m) {
          wbase[q] = (uint32_t)(w0 + kw);
          claimed[q] = m & ~atomicOr(&tmark[wbase[q]], m);
          n += __popc(claimed[q]);
        }
      }
      uint32_t pos = block_reserve<kThreads>(n, &out->ntargets);
#pragma unroll
      for (int q = 0; q < kFlushPT; q++) {
        uint32_t cm = claimed[q];
        while (cm) {
          const int b = __ffs(cm) - 1;
          cm &= cm - 1;
          tlist[pos++] = wbase[q] * 32u + (uint32_t)b;
        }
      }
    }
  }
  if (kLabel && lerr) atomicOr(&run->err, 1u);
  if (!kDetect) return;
  block_add3<kThreads>(lc, out->cnt);
  if (tid == 0 && blockIdx.x == 0 && !kLabel) atomicAdd(&run->dense_passes, 1ull);
}

// ---------------------------------------------------------------------------
// Shared pieces of the apply phase.

// Decision of the apply of sweep j (identical in every thread): the mode of
// detection j+1, or kModeTail, or kModeDone with *done = 1 (converged) / 2 (cap).
__device__ __forceinline__ uint32_t decide(const Counters &cj, uint32_t j, const Sched &s,
                                           bool tail_allowed, uint32_t *done) {
  bool zero = cj.cnt[0] == 0 && cj.cnt[1] == 0 && cj.cnt[2] == 0;
  *done = 0;
  if (zero) {
    *done = 1;
    return kModeDone;
  }
  if (s.mode == PMSZ_SYNC && j > s.max_sweeps) {
    *done = 2;
    return kModeDone;
  }
  if (tail_allowed && cj.ntargets <= kTailTargets) return kModeTail;
  if (s.mode == PMSZ_LAZY && j >= s.max_sweeps) {
    *done = 2;  // the edits of sweep j still land (they belong to sweep j)
    return kModeDone;
  }
  return (uint64_t)cj.ntargets * kCand >= s.dense_min_work ? kModeDense : kModeSparse;
}

// One edit (applyDeltaBuffer_local) on target t, counting first edits / clamps.
template <typename T>
__device__ __forceinline__ void apply_one(T *g, const T *f, uint8_t *dc, uint32_t *tmark,
                                          uint32_t t, const EditParams &ep, uint32_t &touched,
                                          uint32_t &clamped) {
  T gv = g[t];
  uint8_t c0 = dc[t], c = c0;
  edit_value<T>(f[t], gv, c, ep);
  g[t] = gv;
  dc[t] = c;
  touched += c0 == 0;
  clamped += (c == ep.q && c0 != ep.q);
  // every set bit of tmark is a target of this sweep and is applied in this phase,
  // so the word can be cleared with a plain store
  tmark[t >> 5] = 0u;
}

// Claims the closed stencil neighbourhood of t in wmark; returns the new voxels.
__device__ __forceinline__ int queue_neighbourhood(uint32_t t, uint32_t *wmark, const Grid3 &d,
                                                   uint32_t (&u)[kCand]) {
  int64_t x, y, z;
  decompose(t, d, x, y, z);
  const uint32_t vm = valid_mask(x, y, z, d);
  // issue all claims before looking at any result, so the atomics overlap
  uint32_t old[kCand];
#pragma unroll
  for (int r = 0; r < kCand; r++) {
    const uint32_t v = (uint32_t)((int64_t)t + rank_off(r, d));
    old[r] = ((vm >> r) & 1u) ? atomicOr(&wmark[v >> 5], 1u << (v & 31u)) : 0xffffffffu;
  }
  int n = 0;
#pragma unroll
  for (int r = 0; r < kCand; r++) {
    const uint32_t v = (uint32_t)((int64_t)t + rank_off(r, d));
    if (!((old[r] >> (v & 31u)) & 1u)) u[n++] = v;
  }
  return n;
}

// Detection at one listed voxel (global loads); returns the claimed targets.
template <typename T>
__device__ __forceinline__ int detect_one(const T *g, const uint8_t *code, uint32_t *tmark,
                                          uint32_t v, const Grid3 &d, uint32_t (&lc)[3],
                                          uint32_t (&out)[4]) {
  int64_t x, y, z;
  decompose(v, d, x, y, z);
  uint32_t vm = valid_mask(x, y, z, d);
  T cg[kCand];
#pragma unroll
  for (int r = 0; r < kCand; r++)
    cg[r] = ((vm >> r) & 1u) ? g[(int64_t)v + rank_off(r, d)] : (T)0;
  int da, dd;
  if (vm == 0x7fffu) scan_full(cg, da, dd);
  else scan_masked(cg, vm, da, dd);
  uint8_t oc = code[v];
  int tr[4];
  int nt = detect_targets(oc & 15, oc >> 4, da, dd, tr, lc);
  uint32_t tv[4], old[4];
  for (int q = 0; q < nt; q++) {
    tv[q] = (uint32_t)((int64_t)v + rank_off(tr[q], d));
    old[q] = atomicOr(&tmark[tv[q] >> 5], 1u << (tv[q] & 31u));
  }
  int n = 0;
  for (int q = 0; q < nt; q++)
    if (!((old[q] >> (tv[q] & 31u)) & 1u)) out[n++] = tv[q];
  return n;
}

__device__ __forceinline__ void zero_counters(Counters *c) {
  uint32_t *w = reinterpret_cast<uint32_t *>(c);
  for (int k = 0; k < (int)(sizeof(Counters) / 4); k++) w[k] = 0;
}

// ---------------------------------------------------------------------------
// Apply of sweep j (grid-stride over its targets), queueing the next work list
// when detection j+1 is sparse.
template <typename T>
__global__ __launch_bounds__(kApplyThreads) void k_apply(
    T *__restrict__ g, const T *__restrict__ f, uint8_t *__restrict__ dc,
    uint32_t *__restrict__ tmark, const uint32_t *__restrict__ tlist,
    uint32_t *__restrict__ wmark, uint32_t *__restrict__ wlist, Counters *__restrict__ ring,
    RunState *run, uint32_t j, Grid3 d, EditParams ep, Sched s) {
  Counters &prev = ring[(j - 1) % kCtrRing];
  Counters &cj = ring[j % kCtrRing];
  const uint32_t pm = prev.mode;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (pm == kModeTail || pm == kModeDone || pm == kModeUnset) {
    if (lead) cj.mode = pm == kModeUnset ? kModeDone : pm;
    return;
  }
  uint32_t done;
  const uint32_t mode = decide(cj, j, s, true, &done);
  if (lead) {
    cj.mode = mode;
    if (mode == kModeTail) run->tail_j = j;
    if (done) {
      run->done = done;
      run->last_j = j;
    }
  }
  if (mode == kModeTail) return;
  if (mode == kModeDone && (done == 1 || s.mode == PMSZ_SYNC)) return;
  if (lead) zero_counters(&ring[(j + 1) % kCtrRing]);
  const uint32_t nt = cj.ntargets;
  const bool queue = mode == kModeSparse;
  uint32_t touched = 0, clamped = 0;
  const uint32_t stride = gridDim.x * kApplyThreads;
  for (uint32_t base = blockIdx.x * kApplyThreads; base < nt; base += stride) {
    const uint32_t i = base + threadIdx.x;
    uint32_t u[kCand];
    int n = 0;
    if (i < nt) {
      const uint32_t t = tlist[i];
      apply_one<T>(g, f, dc, tmark, t, ep, touched, clamped);
      if (queue) n = queue_neighbourhood(t, wmark, d, u);
    }
    if (queue) {
      uint32_t pos = block_reserve<kApplyThreads>((uint32_t)n, &cj.nwork);
      for (int q = 0; q < n; q++) wlist[pos + q] = u[q];
    }
  }
  uint32_t ts = wave_sum(touched), cs = wave_sum(clamped);
  if (lane_id() == 0) {
    if (ts) atomicAdd(&run->touched, (unsigned long long)ts);
    if (cs) atomicAdd(&run->clamped, (unsigned long long)cs);
  }
  if (lead) atomicAdd(&run->total_targets, (unsigned long long)nt);
}

// Sparse detection j+1 over the work list queued by apply j (grid-stride).
template <typename T>
__global__ __launch_bounds__(kApplyThreads) void k_detect_list(
    const T *__restrict__ g, const uint8_t *__restrict__ code, uint32_t *__restrict__ tmark,
    uint32_t *__restrict__ tlist, uint32_t *__restrict__ wmark,
    const uint32_t *__restrict__ wlist, Counters *__restrict__ ring, RunState *run, uint32_t j,
    Grid3 d) {
  const Counters &cj = ring[j % kCtrRing];
  if (cj.mode != kModeSparse) return;
  Counters &cn = ring[(j + 1) % kCtrRing];
  const uint32_t nw = cj.nwork;
  uint32_t lc[3] = {0, 0, 0};
  const uint32_t stride = gridDim.x * kApplyThreads;
  for (uint32_t base = blockIdx.x * kApplyThreads; base < nw; base += stride) {
    const uint32_t i = base + threadIdx.x;
    uint32_t out[4];
    int n = 0;
    if (i < nw) {
      const uint32_t v = wlist[i];
      wmark[v >> 5] = 0u;  // every set bit is a listed voxel of this sweep
      n = detect_one<T>(g, code, tmark, v, d, lc, out);
    }
    uint32_t pos = block_reserve<kApplyThreads>((uint32_t)n, &cn.ntargets);
    for (int q = 0; q < n; q++) tlist[pos + q] = out[q];
  }
  block_add3<kApplyThreads>(lc, cn.cnt);
  if (blockIdx.x == 0 && threadIdx.x == 0)
    atomicAdd(&run->sparse_voxels, (unsigned long long)nw);
}

// Single-workgroup tail: runs apply / detect sweeps to the end once the target
// sets are small, with workgroup barriers between the phases.
template <typename T>
__global__ __launch_bounds__(kTailThreads) void k_tail(
    T *__restrict__ g, const T *__restrict__ f, uint8_t *__restrict__ dc,
    const uint8_t *__restrict__ code, uint32_t *__restrict__ tmark, uint32_t *__restrict__ tlist,
    uint32_t *__restrict__ wmark, uint32_t *__restrict__ wlist, Counters *__restrict__ ring,
    RunState *run, Grid3 d, EditParams ep, Sched s) {
  __shared__ uint32_t s_ctl[2];
  uint32_t j = run->tail_j;
  if (j == 0 || run->done) return;
  const bool lead = threadIdx.x == 0;
  unsigned long long tt = 0, sv = 0;
  uint32_t touched = 0, clamped = 0;
  while (true) {
    Counters &cj = ring[j % kCtrRing];
    if (lead) {
      uint32_t done;
      const uint32_t mode = decide(cj, j, s, false, &done);
      s_ctl[0] = mode;
      s_ctl[1] = done;
      cj.mode = mode == kModeDone ? kModeDone : kModeTail;
      if (done) {
        run->done = done;
        run->last_j = j;
      }
    }
    __syncthreads();
    const uint32_t mode = s_ctl[0], done = s_ctl[1];
    if (done == 1 || (done == 2 && s.mode == PMSZ_SYNC)) break;
    if (lead) zero_counters(&ring[(j + 1) % kCtrRing]);
    // apply sweep j, queueing the next work list unless this was the last sweep
    const uint32_t nt = cj.ntargets;
    const bool queue = mode != kModeDone;
    for (uint32_t base = 0; base < nt; base += kTailThreads) {
      const uint32_t i = base + threadIdx.x;
      uint32_t u[kCand];
      int n = 0;
      if (i < nt) {
        const uint32_t t = tlist[i];
        apply_one<T>(g, f, dc, tmark, t, ep, touched, clamped);
        if (queue) n = queue_neighbourhood(t, wmark, d, u);
      }
      if (queue) {
        uint32_t pos = block_reserve<kTailThreads>((uint32_t)n, &cj.nwork);
        for (int q = 0; q < n; q++) wlist[pos + q] = u[q];
      }
    }
    tt += nt;
    __syncthreads();
    if (!queue) break;  // lazy sweep cap: sweep j's edits landed, no detection j+1
    // detection j+1 over the work list
    Counters &cn = ring[(j + 1) % kCtrRing];
    const uint32_t nw = cj.nwork;
    uint32_t lc[3] = {0, 0, 0};
    for (uint32_t base = 0; base < nw; base += kTailThreads) {
      const uint32_t i = base + threadIdx.x;
      uint32_t out[4];
      int n = 0;
      if (i < nw) {
        const uint32_t v = wlist[i];
        atomicAnd(&wmark[v >> 5], ~(1u << (v & 31u)));
        n = detect_one<T>(g, code, tmark, v, d, lc, out);
      }
      uint32_t pos = block_reserve<kTailThreads>((uint32_t)n, &cn.ntargets);
      for (int q = 0; q < n; q++) tlist[pos + q] = out[q];
    }
    block_add3<kTailThreads>(lc, cn.cnt);
    sv += nw;
    j++;
  }
  uint32_t ts = wave_sum(touched), cs = wave_sum(clamped);
  if (lane_id() == 0) {
    if (ts) atomicAdd(&run->touched, (unsigned long long)ts);
    if (cs) atomicAdd(&run->clamped, (unsigned long long)cs);
  }
  if (lead) {
    atomicAdd(&run->total_targets, tt);
    atomicAdd(&run->sparse_voxels, sv);
  }
}

// ---------------------------------------------------------------------------
// Labels: pointers from the codes, then each voxel follows its chain to the
// root, reading entries other threads have already shortened.
__global__ void k_init_pointers(const uint8_t *__restrict__ code, int32_t *__restrict__ asc,
                                int32_t *__restrict__ desc, Grid3 d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  uint8_t c = code[i];
  asc[i] = (int32_t)(i + rank_off(c & 15, d));
  desc[i] = (int32_t)(i + rank_off(c >> 4, d));
}

__global__ void k_compress(int32_t *lab, Grid3 d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  int32_t cur = lab[i];
  // chains strictly increase in the (value, index) order, so they end at a root
  // within N steps; the cap only guards against a corrupted pointer field
  for (int64_t step = 0; step < d.n; step++) {
    int32_t nx = lab[cur];
    if (nx == cur) break;
    cur = nx;
  }
  lab[i] = cur;
}

// extract_edits_kernel (pMSz.cu:2090-2138).
template <typename T>
__global__ void k_edit_symbols(const T *__restrict__ f, const T *__restrict__ g0,
                               const T *__restrict__ g, uint8_t *__restrict__ sym,
                               unsigned long long *__restrict__ count, int64_t n,
                               EditParams ep) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool edited = false;
  if (i < n) {
    T a = g[i], b = g0[i];
    uint8_t s = 0;
    if (a != b && a < b) {
      edited = true;
      if (a != clamp_value<T>(f[i], ep)) {
        T delta = (T)ep.delta_d;
        s = (uint8_t)(int)std::round((b - a) / delta);
      } else {
        s = (uint8_t)ep.q;
      }
    }
    sym[i] = s;
  }
  uint64_t m = __ballot(edited);
  if (m && __lane_id() == 0) atomicAdd(count, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------------------
// Host side.

struct Ctx {
  int device = -1;
  int cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  int64_t cap = 0;  // voxels
  uint8_t *code = nullptr, *dc = nullptr;
  uint32_t *tmark = nullptr, *wmark = nullptr, *tlist = nullptr, *wlist = nullptr;
  Counters *ring = nullptr;
  RunState *run = nullptr;
  Counters *h_ring = nullptr;  // pinned mirror of the ring
  RunState *h_run = nullptr;   // pinned mirror of the run state
};

std::mutex g_mu;
Ctx g_ctx;

hipError_t &last_hip_error() {
  static hipError_t e = hipSuccess;
  return e;
}

#define PMSZ_HIP(call)       \
  do {                       \
    hipError_t e_ = (call);  \
    if (e_ != hipSuccess) {  \
      last_hip_error() = e_; \
      return PMSZ_EDEVICE;   \
    }                        \
  } while (0)

void release(Ctx &c) {
  (void)hipFree(c.code);
  (void)hipFree(c.dc);
  (void)hipFree(c.tmark);
  (void)hipFree(c.wmark);
  (void)hipFree(c.tlist);
  (void)hipFree(c.wlist);
  c.code = c.dc = nullptr;
  c.tmark = c.wmark = c.tlist = c.wlist = nullptr;
  c.cap = 0;
}

int ensure(Ctx &c, int64_t n) {
  int dev = 0;
  PMSZ_HIP(hipGetDevice(&dev));
  if (c.device != dev) {
    // a new device: the state of the previous one is left to process exit
    c = Ctx();
    c.device = dev;
    PMSZ_HIP(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, dev));
    PMSZ_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    PMSZ_HIP(hipEventCreate(&c.ev0));
    PMSZ_HIP(hipEventCreate(&c.ev1));
    PMSZ_HIP(hipEventCreate(&c.ev2));
    PMSZ_HIP(hipEventCreate(&c.ev3));
    PMSZ_HIP(hipMalloc(&c.ring, sizeof(Counters) * kCtrRing));
    PMSZ_HIP(hipMalloc(&c.run, sizeof(RunState)));
    PMSZ_HIP(hipHostMalloc(&c.h_ring, sizeof(Counters) * kCtrRing, hipHostMallocDefault));
    PMSZ_HIP(hipHostMalloc(&c.h_run, sizeof(RunState), hipHostMallocDefault));
  }
  if (n <= c.cap) return PMSZ_OK;
  release(c);
  int64_t words = (n + 31) / 32 + 1;
  if (hipMalloc(&c.code, (size_t)n) != hipSuccess || hipMalloc(&c.dc, (size_t)n) != hipSuccess ||
      hipMalloc(&c.tmark, sizeof(uint32_t) * words) != hipSuccess ||
      hipMalloc(&c.wmark, sizeof(uint32_t) * words) != hipSuccess ||
      hipMalloc(&c.tlist, sizeof(uint32_t) * (size_t)n) != hipSuccess ||
      hipMalloc(&c.wlist, sizeof(uint32_t) * (size_t)n) != hipSuccess) {
    release(c);
    return PMSZ_ENOMEM;
  }
  c.cap = n;
  return PMSZ_OK;
}

int check_params(const pmsz_params *p) {
  if (!p) return PMSZ_EINVAL;
  if (p->dtype != PMSZ_F32 && p->dtype != PMSZ_F64) return PMSZ_EINVAL;
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return PMSZ_EINVAL;
  int64_t n = p->nx * p->ny * p->nz;
  if (n < 2) return PMSZ_EINVAL;
  if (!(p->abs_eb > 0) || std::isinf(p->abs_eb)) return PMSZ_EINVAL;
  if (p->q < 1 || p->q > 255) return PMSZ_EINVAL;
  if (p->max_sweeps < 1) return PMSZ_EINVAL;
  if (p->mode != PMSZ_LAZY && p->mode != PMSZ_SYNC) return PMSZ_EINVAL;
  if (p->clamp_mode != 0 && p->clamp_mode != 1) return PMSZ_EINVAL;
  if (p->ghost_policy != 0 && p->ghost_policy != 1) return PMSZ_EINVAL;
  for (int a = 0; a < 3; a++)
    if (p->sub[a] < 1) return PMSZ_EINVAL;
  if (n >= (int64_t)0xffffffffLL) return PMSZ_ETOOBIG;
  return PMSZ_OK;
}

Grid3 grid_of(int64_t nx, int64_t ny, int64_t nz) {
  Grid3 d;
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.sxy = nx * ny;
  d.n = nx * ny * nz;
  d.dnx.init((uint32_t)std::min<int64_t>(nx, 0xffffffffLL));
  d.dsxy.init((uint32_t)std::min<int64_t>(d.sxy, 0xffffffffLL));
  return d;
}

int64_t tile_count(const Grid3 &d) {
  return ((d.nx + kTX - 1) / kTX) * ((d.ny + kTY - 1) / kTY) * ((d.nz + kTZ - 1) / kTZ);
}

// Persistent grid of the tile kernel: a multiple of 8 (one share per XCD), at most
// the resident capacity and no more than the tiles need.
unsigned tile_blocks(const Ctx &c, const Grid3 &d) {
  int64_t per_xcd = (tile_count(d) + 7) / 8;
  int64_t cap = std::max<int64_t>(1, (int64_t)c.cus * 4 / 8);
  return (unsigned)(8 * std::max<int64_t>(1, std::min(per_xcd, cap)));
}

unsigned compact_blocks(const Ctx &c, int64_t words) {
  const int64_t per = (int64_t)kCompactThreads * kCompactPT;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((words + per - 1) / per, c.cus * 4));
}

// Persistent grid of the column kernel: one wave per work unit up to about the
// resident capacity, a multiple of 8 blocks (one share per XCD).
unsigned col_blocks(const Ctx &c, const Grid3 &d) {
  const int64_t units = ((d.nx + kColOwned - 1) / kColOwned) * ((d.ny + kColRows - 1) / kColRows) *
                        ((d.nz + kColZ - 1) / kColZ);
  const int64_t blocks = (units + kColWaves - 1) / kColWaves;
  const int64_t per_xcd = (blocks + 7) / 8;
  const int64_t cap = std::max<int64_t>(1, (int64_t)c.cus * 8 / 8);
  return (unsigned)(8 * std::max<int64_t>(1, std::min(per_xcd, cap)));
}

EditParams edit_params(const pmsz_params *p) {
  EditParams ep;
  ep.eb = p->abs_eb;
  ep.delta_d = p->abs_eb / p->q;
  ep.q = p->q;
  ep.clamp_mode = p->clamp_mode;
  return ep;
}

constexpr uint32_t kChain = 8;  // apply/detect steps enqueued before each host check

template <typename T>
int correct_single(Ctx &c, const pmsz_params *p, const T *f, T *g, uint8_t *dc_out,
                   pmsz_stats *st) {
  const Grid3 d = grid_of(p->nx, p->ny, p->nz);
  int rc = ensure(c, d.n);
  if (rc) return rc;
  const EditParams ep = edit_params(p);
  Sched s;
  s.mode = p->mode;
  s.max_sweeps = (uint32_t)p->max_sweeps;
  s.dense_min_work = (uint32_t)std::min<int64_t>(std::max<int64_t>(d.n / 16, 1), 0xffffffffLL);
  const int64_t words = (d.n + 31) / 32 + 1;
  const unsigned grid_list = (unsigned)std::max(1, c.cus * 4);
  std::memset(st, 0, sizeof(*st));

  PMSZ_HIP(hipEventRecord(c.ev0, c.stream));
  PMSZ_HIP(hipMemsetAsync(c.tmark, 0, sizeof(uint32_t) * words, c.stream));
  PMSZ_HIP(hipMemsetAsync(c.wmark, 0, sizeof(uint32_t) * words, c.stream));
  PMSZ_HIP(hipMemsetAsync(c.ring, 0, sizeof(Counters) * kCtrRing, c.stream));
  PMSZ_HIP(hipMemsetAsync(c.run, 0, sizeof(RunState), c.stream));
  // apply(1) reads slot 0's mode: mark detection 1 as a running sweep
  static const uint32_t kRunning = kModeDense;
  PMSZ_HIP(hipMemcpyAsync(&c.ring[0].mode, &kRunning, sizeof(uint32_t), hipMemcpyHostToDevice,
                          c.stream));
  PMSZ_HIP(hipEventRecord(c.ev2, c.stream));
  k_columns<T, true, true><<<col_blocks(c, d), kColThreads, 0, c.stream>>>(
      f, g, c.code, c.dc, c.tmark, c.tlist, c.ring, 0, 1, c.run, d);
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipEventRecord(c.ev3, c.stream));
  k_compact<<<compact_blocks(c, words), kCompactThreads, 0, c.stream>>>(c.tmark, c.tlist, c.ring,
                                                                        0, 1, true, (uint32_t)words);

  uint32_t j0 = 1;
  while (true) {
    for (uint32_t j = j0; j < j0 + kChain; j++) {
      k_apply<T><<<grid_list, kApplyThreads, 0, c.stream>>>(g, f, c.dc, c.tmark, c.tlist,
                                                            c.wmark, c.wlist, c.ring, c.run, j,
                                                            d, ep, s);
      k_columns<T, false, true><<<col_blocks(c, d), kColThreads, 0, c.stream>>>(
          nullptr, g, c.code, nullptr, c.tmark, c.tlist, c.ring, j % kCtrRing,
          (j + 1) % kCtrRing, c.run, d);
      k_compact<<<compact_blocks(c, words), kCompactThreads, 0, c.stream>>>(
          c.tmark, c.tlist, c.ring, j % kCtrRing, (j + 1) % kCtrRing, false, (uint32_t)words);
      k_detect_list<T><<<grid_list, kApplyThreads, 0, c.stream>>>(
          g, c.code, c.tmark, c.tlist, c.wmark, c.wlist, c.ring, c.run, j, d);
    }
    k_tail<T><<<1, kTailThreads, 0, c.stream>>>(g, f, c.dc, c.code, c.tmark, c.tlist, c.wmark,
                                                c.wlist, c.ring, c.run, d, ep, s);
    PMSZ_HIP(hipGetLastError());
    PMSZ_HIP(hipMemcpyAsync(c.h_run, c.run, sizeof(RunState), hipMemcpyDeviceToHost, c.stream));
    PMSZ_HIP(hipStreamSynchronize(c.stream));
    if (c.h_run->err) return PMSZ_EINVAL;
    if (c.h_run->done) break;
    // the chain ended before the tail took over: continue it where it stopped
    j0 += kChain;
  }
  PMSZ_HIP(hipEventRecord(c.ev1, c.stream));
  PMSZ_HIP(hipMemcpyAsync(c.h_ring, c.ring, sizeof(Counters) * kCtrRing, hipMemcpyDeviceToHost,
                          c.stream));
  if (dc_out)
    PMSZ_HIP(hipMemcpyAsync(dc_out, c.dc, (size_t)d.n, hipMemcpyDeviceToDevice, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  float ms = 0, dms = 0;
  PMSZ_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  PMSZ_HIP(hipEventElapsedTime(&dms, c.ev2, c.ev3));
  st->t_total_s = ms * 1e-3;
  st->t_dense_s = dms * 1e-3;

  const RunState &r = *c.h_run;
  const uint32_t m = r.last_j;
  const Counters &cm = c.h_ring[m % kCtrRing];
  const bool conv = r.done == 1;
  const int status = conv ? PMSZ_OK : PMSZ_NOT_CONVERGED;
  st->status = status;
  st->final_cnt[0] = cm.cnt[0];
  st->final_cnt[1] = cm.cnt[1];
  st->final_cnt[2] = cm.cnt[2];
  if (m <= (uint32_t)kCtrRing) {
    st->trace_n = std::min<int64_t>(m, 64);
    for (int64_t k = 0; k < st->trace_n; k++) {
      const Counters &ck = c.h_ring[(k + 1) % kCtrRing];
      for (int q = 0; q < 3; q++) st->trace_cnt[k][q] = ck.cnt[q];
      st->trace_targets[k] = ck.ntargets;
    }
  }
  st->total_targets = (int64_t)r.total_targets;
  st->touched_voxels = (int64_t)r.touched;
  st->clamped_voxels = (int64_t)r.clamped;
  st->sparse_voxels = (int64_t)r.sparse_voxels;
  const int64_t extra_dense = (int64_t)r.dense_passes;
  st->dense_passes = 1 + extra_dense;
  if (p->mode == PMSZ_LAZY) {
    st->edit_sweeps = m;
    st->check_sweeps = conv ? 1 : 0;
    st->outer_iters = conv ? 1 : 0;
    st->ite = m + 1;
  } else {
    int64_t applies = conv ? std::max<int64_t>((int64_t)m - 1, 1) : (int64_t)m - 1;
    st->edit_sweeps = applies;
    st->check_sweeps = applies;
    st->outer_iters = applies;
    st->ite = 2 * applies;
  }
  // first pass: f, g read, code, dc written; later dense sweeps: g, code read;
  // edits: f, g, dc read and g, dc written; listed voxels: g and code read
  st->bytes_algorithmic = (double)d.n * (2.0 * sizeof(T) + 2.0) +
                          (double)d.n * (sizeof(T) + 1.0) * (double)extra_dense +
                          (double)st->total_targets * (3.0 * sizeof(T) + 2.0) +
                          (double)st->sparse_voxels * (sizeof(T) + 1.0);
  return status;
}

// Timing of the fused first pass alone (events around each launch), for the
// roofline report.
template <typename T>
int time_first_pass(Ctx &c, const pmsz_params *p, const T *f, const T *g, int iters,
                    double *avg_s) {
  const Grid3 d = grid_of(p->nx, p->ny, p->nz);
  int rc = ensure(c, d.n);
  if (rc) return rc;
  const int64_t words = (d.n + 31) / 32 + 1;
  double tot = 0;
  for (int it = 0; it < iters; it++) {
    PMSZ_HIP(hipMemsetAsync(c.tmark, 0, sizeof(uint32_t) * words, c.stream));
    PMSZ_HIP(hipMemsetAsync(c.ring, 0, sizeof(Counters) * 2, c.stream));
    PMSZ_HIP(hipEventRecord(c.ev0, c.stream));
    k_columns<T, true, true><<<col_blocks(c, d), kColThreads, 0, c.stream>>>(
        f, g, c.code, c.dc, c.tmark, c.tlist, c.ring, 0, 1, c.run, d);
    PMSZ_HIP(hipGetLastError());
    k_compact<<<compact_blocks(c, words), kCompactThreads, 0, c.stream>>>(
        c.tmark, c.tlist, c.ring, 0, 1, true, (uint32_t)words);
    PMSZ_HIP(hipEventRecord(c.ev1, c.stream));
    PMSZ_HIP(hipStreamSynchronize(c.stream));
    float ms = 0;
    PMSZ_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    tot += ms * 1e-3;
  }
  *avg_s = tot / std::max(1, iters);
  return PMSZ_OK;
}

}  // namespace pmsz

using namespace pmsz;

extern "C" {

void pmsz_default_params(pmsz_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->q = 5;
  p->sub[0] = p->sub[1] = p->sub[2] = 1;
  p->mode = PMSZ_LAZY;
  p->max_sweeps = 100000;
  p->clamp_mode = PMSZ_CLAMP_BOUND_SAFE;
  p->ghost_policy = PMSZ_GHOST_FULL;
}

int pmsz_correct(const pmsz_params *p, const void *d_f, void *d_g, uint8_t *d_dc,
                 pmsz_stats *stats) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!d_f || !d_g) return PMSZ_EINVAL;
  if (p->sub[0] * p->sub[1] * p->sub[2] != 1) return PMSZ_EINVAL;  // subdomains: separate path
  pmsz_stats tmp;
  if (!stats) stats = &tmp;
  std::lock_guard<std::mutex> lk(g_mu);
  if (p->dtype == PMSZ_F32)
    return correct_single<float>(g_ctx, p, (const float *)d_f, (float *)d_g, d_dc, stats);
  return correct_single<double>(g_ctx, p, (const double *)d_f, (double *)d_g, d_dc, stats);
}

int pmsz_time_first_pass(const pmsz_params *p, const void *d_f, const void *d_g, int iters,
                         double *avg_seconds) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!d_f || !d_g || !avg_seconds || iters < 1) return PMSZ_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  if (p->dtype == PMSZ_F32)
    return time_first_pass<float>(g_ctx, p, (const float *)d_f, (const float *)d_g, iters,
                                  avg_seconds);
  return time_first_pass<double>(g_ctx, p, (const double *)d_f, (const double *)d_g, iters,
                                 avg_seconds);
}

int pmsz_correct_host(const pmsz_params *p, const void *h_f, void *h_g, uint8_t *h_dc,
                      pmsz_stats *stats) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!h_f || !h_g) return PMSZ_EINVAL;
  size_t n = (size_t)(p->nx * p->ny * p->nz);
  size_t bytes = n * (p->dtype == PMSZ_F32 ? 4 : 8);
  void *df = nullptr, *dg = nullptr;
  uint8_t *ddc = nullptr;
  if (hipMalloc(&df, bytes) != hipSuccess || hipMalloc(&dg, bytes) != hipSuccess ||
      (h_dc && hipMalloc(&ddc, n) != hipSuccess)) {
    (void)hipFree(df);
    (void)hipFree(dg);
    return PMSZ_ENOMEM;
  }
  rc = PMSZ_OK;
  if (hipMemcpy(df, h_f, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dg, h_g, bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = PMSZ_EDEVICE;
  if (rc == PMSZ_OK) rc = pmsz_correct(p, df, dg, ddc, stats);
  if (rc == PMSZ_OK || rc == PMSZ_NOT_CONVERGED) {
    if (hipMemcpy(h_g, dg, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = PMSZ_EDEVICE;
    if (h_dc && hipMemcpy(h_dc, ddc, n, hipMemcpyDeviceToHost) != hipSuccess) rc = PMSZ_EDEVICE;
  }
  (void)hipFree(df);
  (void)hipFree(dg);
  (void)hipFree(ddc);
  return rc;
}

int pmsz_codes(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *d_field,
               uint8_t *d_codes) {
  pmsz_params p;
  pmsz_default_params(&p);
  p.dtype = dtype;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.abs_eb = 1.0;
  int rc = check_params(&p);
  if (rc) return rc;
  if (!d_field || !d_codes) return PMSZ_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  Ctx &c = g_ctx;
  const Grid3 d = grid_of(nx, ny, nz);
  rc = ensure(c, 1);
  if (rc) return rc;
  PMSZ_HIP(hipMemsetAsync(c.run, 0, sizeof(RunState), c.stream));
  if (dtype == PMSZ_F32)
    k_columns<float, true, false><<<col_blocks(c, d), kColThreads, 0, c.stream>>>(
        (const float *)d_field, nullptr, d_codes, nullptr, nullptr, nullptr, c.ring, 0, 0, c.run,
        d);
  else
    k_columns<double, true, false><<<col_blocks(c, d), kColThreads, 0, c.stream>>>(
        (const double *)d_field, nullptr, d_codes, nullptr, nullptr, nullptr, c.ring, 0, 0,
        c.run, d);
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipMemcpyAsync(c.h_run, c.run, sizeof(RunState), hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  return c.h_run->err ? PMSZ_EINVAL : PMSZ_OK;
}

int pmsz_labels(int dtype, int64_t nx, int64_t ny, int64_t nz, const void *d_field,
                int32_t *d_asc, int32_t *d_desc) {
  if (nx < 1 || ny < 1 || nz < 1) return PMSZ_EINVAL;
  int64_t n = nx * ny * nz;
  if (n >= ((int64_t)1 << 31)) return PMSZ_ETOOBIG;
  if (!d_asc || !d_desc) return PMSZ_EINVAL;
  uint8_t *codes = nullptr;
  if (hipMalloc(&codes, (size_t)n) != hipSuccess) return PMSZ_ENOMEM;
  int rc = pmsz_codes(dtype, nx, ny, nz, d_field, codes);
  if (rc) {
    (void)hipFree(codes);
    return rc;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  Ctx &c = g_ctx;
  const Grid3 d = grid_of(nx, ny, nz);
  unsigned blocks = (unsigned)((n + 255) / 256);
  k_init_pointers<<<blocks, 256, 0, c.stream>>>(codes, d_asc, d_desc, d);
  k_compress<<<blocks, 256, 0, c.stream>>>(d_asc, d);
  k_compress<<<blocks, 256, 0, c.stream>>>(d_desc, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)hipFree(codes);
  if (e != hipSuccess) {
    last_hip_error() = e;
    return PMSZ_EDEVICE;
  }
  return PMSZ_OK;
}

int pmsz_edit_symbols(const pmsz_params *p, const void *d_f, const void *d_g0, const void *d_g,
                      uint8_t *d_symbols, int64_t *edited) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!d_f || !d_g0 || !d_g || !d_symbols) return PMSZ_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  Ctx &c = g_ctx;
  rc = ensure(c, 1);
  if (rc) return rc;
  int64_t n = p->nx * p->ny * p->nz;
  const EditParams ep = edit_params(p);
  unsigned long long *dcount = &c.run->total_targets;
  PMSZ_HIP(hipMemsetAsync(dcount, 0, sizeof(unsigned long long), c.stream));
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (p->dtype == PMSZ_F32)
    k_edit_symbols<float><<<blocks, 256, 0, c.stream>>>((const float *)d_f, (const float *)d_g0,
                                                        (const float *)d_g, d_symbols, dcount,
                                                        n, ep);
  else
    k_edit_symbols<double><<<blocks, 256, 0, c.stream>>>(
        (const double *)d_f, (const double *)d_g0, (const double *)d_g, d_symbols, dcount, n,
        ep);
  PMSZ_HIP(hipGetLastError());
  unsigned long long h = 0;
  PMSZ_HIP(hipMemcpyAsync(&h, dcount, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  if (edited) *edited = (int64_t)h;
  return PMSZ_OK;
}

const char *pmsz_strerror(int code) {
  switch (code) {
    case PMSZ_OK: return "ok";
    case PMSZ_EINVAL: return "invalid argument (parameters, NaN or the 1e30 sentinel in a field)";
    case PMSZ_ENOMEM: return "device allocation failed";
    case PMSZ_NOT_CONVERGED: return "not converged within max_sweeps";
    case PMSZ_EDEVICE: return "HIP runtime error";
    case PMSZ_ETOOBIG: return "field too large for 32-bit voxel indices";
    case PMSZ_EBUSY: return "capacity exhausted";
    default: return "unknown error";
  }
}

const char *pmsz_version(void) { return "pmsz-mi355x 0.2"; }

}  // extern "C"
