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
//   ctr         Counters[]  one block of counters per sweep
//
// Snapshot semantics: a sweep detects on the start-of-sweep g, writing only
// marks, lists and counters; a separate launch then applies one edit per target
// (applyDeltaBuffer_local, pMSz.cu:623-645).  The classification of a voxel
// depends on g over its closed stencil neighbourhood only, so after the first
// sweep only the neighbourhoods of the previous sweep's targets can change
// (the stencil is symmetric).  Sweep 1 is a full pass fused with the original
// labeling; every later sweep examines exactly that work list and produces the
// same counters and targets as a full sweep would.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/pmsz.h"
#include "pmsz_common.h"

namespace pmsz {

struct Grid3 {
  int64_t nx, ny, nz;
  int64_t sxy;  // nx*ny
  int64_t n;    // nx*ny*nz
};

// Counters of one sweep.  cnt: false max / false min / direction (count_f_max,
// count_f_min, count_f_dir, pMSz.cu:174); ntargets: entries of tlist; nwork:
// entries of wlist queued by this sweep's apply; err bit 0: NaN or the sentinel
// value in an input field; touched / clamped: voxels whose delta counter left 0 /
// reached q in this sweep's apply.
struct Counters {
  uint32_t cnt[3];
  uint32_t ntargets;
  uint32_t nwork;
  uint32_t err;
  uint32_t touched;
  uint32_t clamped;
};

constexpr int kTX = 64;
constexpr int kTY = 4;
constexpr int kTZ = 8;
constexpr int kThreads = kTX * kTY;
constexpr int kCtrCap = 1024;

__device__ __forceinline__ int64_t rank_off(int r, const Grid3 &d) {
  return (int64_t)rank_dx(r) + d.nx * (int64_t)rank_dy(r) + d.sxy * (int64_t)rank_dz(r);
}

// Validity of the 15 candidates of (x, y, z): bit r set iff candidate r lies in
// the domain (out-of-domain neighbours are the reference's INVALID halo).
__device__ __forceinline__ uint32_t valid_mask(int64_t x, int64_t y, int64_t z, const Grid3 &d) {
  uint32_t m = 0;
#pragma unroll
  for (int r = 0; r < kCand; r++) {
    int64_t a = x + rank_dx(r), b = y + rank_dy(r), c = z + rank_dz(r);
    bool ok = a >= 0 && a < d.nx && b >= 0 && b < d.ny && c >= 0 && c < d.nz;
    m |= (ok ? 1u : 0u) << r;
  }
  return m;
}

// Steepest ascent / descent candidate over all 15 candidates.  Visiting ranks in
// increasing order, ">=" keeps the highest rank among equal maxima and "<" the
// lowest rank among equal minima: the (value, index) order of pMSz.cu:610-620.
template <typename T>
__device__ __forceinline__ void scan_full(const T (&c)[kCand], int &asc, int &desc) {
  T bv = c[0], sv = c[0];
  int br = 0, sr = 0;
#pragma unroll
  for (int r = 1; r < kCand; r++) {
    bool ge = c[r] >= bv;
    bv = ge ? c[r] : bv;
    br = ge ? r : br;
    bool lt = c[r] < sv;
    sv = lt ? c[r] : sv;
    sr = lt ? r : sr;
  }
  asc = br;
  desc = sr;
}

// The same over the valid candidates only (domain boundary).
template <typename T>
__device__ __forceinline__ void scan_masked(const T (&c)[kCand], uint32_t valid, int &asc,
                                            int &desc) {
  T bv = c[kSelf], sv = c[kSelf];
  int br = kSelf, sr = kSelf;
#pragma unroll
  for (int r = 0; r < kCand; r++) {
    if (r == kSelf || !((valid >> r) & 1u)) continue;
    bool gt = c[r] > bv || (c[r] == bv && r > br);
    bv = gt ? c[r] : bv;
    br = gt ? r : br;
    bool lt = c[r] < sv || (c[r] == sv && r < sr);
    sv = lt ? c[r] : sv;
    sr = lt ? r : sr;
  }
  asc = br;
  desc = sr;
}

__device__ __forceinline__ uint32_t lane_prefix(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Wave-aggregated append: every lane of the wave must call it.
__device__ __forceinline__ void wave_append(bool has, uint32_t val, uint32_t *counter,
                                            uint32_t *list) {
  uint64_t m = __ballot(has);
  if (m == 0) return;
  uint32_t base = 0;
  int leader = __ffsll((long long)m) - 1;
  if ((int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (has) list[base + lane_prefix(m)] = val;
}

// Set the mark bit of v; true if it was clear (the caller then owns v).
__device__ __forceinline__ bool claim(uint32_t *marks, uint32_t v) {
  uint32_t bit = 1u << (v & 31u);
  uint32_t old = atomicOr(&marks[v >> 5], bit);
  return (old & bit) == 0;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Detection at one voxel from its original codes (oa, od) and the current
// steepest candidates (da, dd): iscriticle data_type=0, pMSz.cu:786-898.
// Returns up to 4 target ranks in t[], counters in c[].
__device__ __forceinline__ int detect_targets(int oa, int od, int da, int dd, int (&t)[4],
                                              uint32_t (&c)[3]) {
  bool omax = oa == kSelf;
  bool omin = od == kSelf && !omax;  // a voxel that is both counts as a maximum (pMSz.cu:750)
  bool dmax = da == kSelf, dmin = dd == kSelf;
  bool fm = dmax != omax, fn = dmin != omin;
  int nt = 0;
  if (fm) {
    c[0]++;
    t[nt++] = omax ? da : kSelf;  // lower the new top neighbour, or the false maximum
  }
  if (fn) {
    c[1]++;
    t[nt++] = omin ? kSelf : od;  // lower the lost minimum, or the true descent neighbour
  }
  if (fm || fn) {
    c[2]++;
  } else {
    if (da != oa) {
      c[2]++;
      t[nt++] = da;  // the wrong steepest-ascent neighbour is too high
    }
    if (dd != od) {
      c[2]++;
      t[nt++] = od;  // the true steepest-descent neighbour is not low enough
    }
  }
  return nt;
}

// ---------------------------------------------------------------------------
// Full pass: original labeling of f (codes) and, with kDetect, the first
// detection sweep on g.  One block = a 64 x 4 column of voxels walking 8
// planes in z; the three planes a voxel's stencil touches are staged in LDS
// with a one-voxel halo.
template <typename T, bool kDetect>
__global__ __launch_bounds__(kThreads) void k_dense_pass(
    const T *__restrict__ f, const T *__restrict__ g, uint8_t *__restrict__ code,
    uint8_t *__restrict__ dc, uint32_t *__restrict__ tmark, uint32_t *__restrict__ tlist,
    Counters *__restrict__ ctr, Grid3 d) {
  constexpr int PX = kTX + 2, PY = kTY + 2;
  __shared__ T sf[3][PY][PX];
  __shared__ T sg[kDetect ? 3 : 1][PY][PX];
  const int tid = threadIdx.x;
  const int tx = tid % kTX, ty = tid / kTX;
  const int64_t bx = (int64_t)blockIdx.x * kTX, by = (int64_t)blockIdx.y * kTY,
                bz = (int64_t)blockIdx.z * kTZ;
  const int64_t x = bx + tx, y = by + ty;

  auto load_plane = [&](int64_t pz, int slot) {
    for (int e = tid; e < PX * PY; e += kThreads) {
      int lx = e % PX, ly = e / PX;
      int64_t gx = bx + lx - 1, gy = by + ly - 1;
      T vf = 0, vg = 0;
      if (pz >= 0 && pz < d.nz && gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny) {
        int64_t gi = gx + d.nx * gy + d.sxy * pz;
        vf = f[gi];
        if (kDetect) vg = g[gi];
      }
      sf[slot][ly][lx] = vf;
      if (kDetect) sg[slot][ly][lx] = vg;
    }
  };

  load_plane(bz - 1, 0);
  load_plane(bz, 1);
  uint32_t lc[3] = {0, 0, 0};
  uint32_t lerr = 0;
  for (int k = 0; k < kTZ; k++) {
    const int64_t z = bz + k;
    if (z >= d.nz) break;  // uniform over the block
    load_plane(z + 1, (k + 2) % 3);
    __syncthreads();
    const int slot[3] = {k % 3, (k + 1) % 3, (k + 2) % 3};
    const bool active = x < d.nx && y < d.ny;
    int tr[4];
    int nt = 0;
    int64_t i = 0;
    if (active) {
      i = x + d.nx * y + d.sxy * z;
      T cf[kCand];
#pragma unroll
      for (int r = 0; r < kCand; r++)
        cf[r] = sf[slot[rank_dz(r) + 1]][ty + 1 + rank_dy(r)][tx + 1 + rank_dx(r)];
      const bool interior =
          x > 0 && x < d.nx - 1 && y > 0 && y < d.ny - 1 && z > 0 && z < d.nz - 1;
      const uint32_t vm = interior ? 0x7fffu : valid_mask(x, y, z, d);
      int oa, od;
      if (interior) scan_full(cf, oa, od);
      else scan_masked(cf, vm, oa, od);
      code[i] = (uint8_t)(oa | (od << 4));
      T fv = cf[kSelf];
      if (!(fv == fv) || fv == sentinel<T>()) lerr |= 1u;
      if (kDetect) {
        dc[i] = 0;
        T cg[kCand];
#pragma unroll
        for (int r = 0; r < kCand; r++)
          cg[r] = sg[slot[rank_dz(r) + 1]][ty + 1 + rank_dy(r)][tx + 1 + rank_dx(r)];
        T gv = cg[kSelf];
        if (!(gv == gv) || gv == sentinel<T>()) lerr |= 1u;
        int da, dd;
        if (interior) scan_full(cg, da, dd);
        else scan_masked(cg, vm, da, dd);
        nt = detect_targets(oa, od, da, dd, tr, lc);
      }
    }
    if (kDetect) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        bool has = j < nt;
        uint32_t tv = 0;
        if (has) {
          tv = (uint32_t)(i + rank_off(tr[j], d));
          has = claim(tmark, tv);
        }
        wave_append(has, tv, &ctr->ntargets, tlist);
      }
    }
    __syncthreads();
  }
  if (kDetect) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      uint32_t s = wave_sum(lc[c]);
      if (__lane_id() == 0 && s) atomicAdd(&ctr->cnt[c], s);
    }
  }
  uint64_t em = __ballot(lerr != 0);
  if (em && __lane_id() == 0) atomicOr(&ctr->err, 1u);
}

// Apply one edit to every target of the sweep, then queue the closed stencil
// neighbourhood of each target for the next sweep.
template <typename T>
__global__ __launch_bounds__(256) void k_apply(T *__restrict__ g, const T *__restrict__ f,
                                               uint8_t *__restrict__ dc,
                                               uint32_t *__restrict__ tmark,
                                               const uint32_t *__restrict__ tlist,
                                               uint32_t ntargets, uint32_t *__restrict__ wmark,
                                               uint32_t *__restrict__ wlist,
                                               Counters *__restrict__ ctr, Grid3 d,
                                               EditParams ep) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < ntargets;
  uint32_t t = 0;
  int64_t x = 0, y = 0, z = 0;
  bool first = false, clamped = false;
  if (act) {
    t = tlist[i];
    T gv = g[t];
    uint8_t c0 = dc[t], c = c0;
    edit_value<T>(f[t], gv, c, ep);
    g[t] = gv;
    dc[t] = c;
    first = c0 == 0;
    clamped = c == ep.q && c0 != ep.q;
    atomicAnd(&tmark[t >> 5], ~(1u << (t & 31u)));
    x = (int64_t)t % d.nx;
    y = ((int64_t)t / d.nx) % d.ny;
    z = (int64_t)t / d.sxy;
  }
  uint64_t mf = __ballot(first), mc = __ballot(clamped);
  if (__lane_id() == 0) {
    if (mf) atomicAdd(&ctr->touched, (uint32_t)__popcll(mf));
    if (mc) atomicAdd(&ctr->clamped, (uint32_t)__popcll(mc));
  }
#pragma unroll
  for (int r = 0; r < kCand; r++) {
    int64_t a = x + rank_dx(r), b = y + rank_dy(r), c = z + rank_dz(r);
    bool has = act && a >= 0 && a < d.nx && b >= 0 && b < d.ny && c >= 0 && c < d.nz;
    uint32_t u = 0;
    if (has) {
      u = (uint32_t)((int64_t)t + rank_off(r, d));
      has = claim(wmark, u);
    }
    wave_append(has, u, &ctr->nwork, wlist);
  }
}

// A detection sweep over the work list queued by the previous apply.
template <typename T>
__global__ __launch_bounds__(256) void k_detect_sparse(
    const T *__restrict__ g, const uint8_t *__restrict__ code, uint32_t *__restrict__ tmark,
    uint32_t *__restrict__ tlist, uint32_t *__restrict__ wmark,
    const uint32_t *__restrict__ wlist, const Counters *__restrict__ prev,
    Counters *__restrict__ ctr, Grid3 d) {
  const uint32_t nwork = prev->nwork;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < nwork;
  uint32_t lc[3] = {0, 0, 0};
  int tr[4];
  int nt = 0;
  uint32_t v = 0;
  if (act) {
    v = wlist[i];
    atomicAnd(&wmark[v >> 5], ~(1u << (v & 31u)));
    int64_t x = (int64_t)v % d.nx, y = ((int64_t)v / d.nx) % d.ny, z = (int64_t)v / d.sxy;
    uint32_t vm = valid_mask(x, y, z, d);
    T cg[kCand];
#pragma unroll
    for (int r = 0; r < kCand; r++)
      cg[r] = ((vm >> r) & 1u) ? g[(int64_t)v + rank_off(r, d)] : (T)0;
    int da, dd;
    if (vm == 0x7fffu) scan_full(cg, da, dd);
    else scan_masked(cg, vm, da, dd);
    uint8_t oc = code[v];
    nt = detect_targets(oc & 15, oc >> 4, da, dd, tr, lc);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    bool has = j < nt;
    uint32_t tv = 0;
    if (has) {
      tv = (uint32_t)((int64_t)v + rank_off(tr[j], d));
      has = claim(tmark, tv);
    }
    wave_append(has, tv, &ctr->ntargets, tlist);
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    uint32_t s = wave_sum(lc[c]);
    if (__lane_id() == 0 && s) atomicAdd(&ctr->cnt[c], s);
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
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  int64_t cap = 0;  // voxels
  uint8_t *code = nullptr, *dc = nullptr;
  uint32_t *tmark = nullptr, *wmark = nullptr, *tlist = nullptr, *wlist = nullptr;
  Counters *ctr = nullptr;
  Counters *h_ctr = nullptr;  // pinned mirror of two consecutive Counters blocks
};

std::mutex g_mu;
Ctx g_ctx;

#define PMSZ_HIP(call)                  \
  do {                                  \
    hipError_t e_ = (call);             \
    if (e_ != hipSuccess) {             \
      last_hip_error() = e_;            \
      return PMSZ_EDEVICE;              \
    }                                   \
  } while (0)

hipError_t &last_hip_error() {
  static hipError_t e = hipSuccess;
  return e;
}

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
    PMSZ_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    PMSZ_HIP(hipEventCreate(&c.ev0));
    PMSZ_HIP(hipEventCreate(&c.ev1));
    PMSZ_HIP(hipEventCreate(&c.ev2));
    PMSZ_HIP(hipMalloc(&c.ctr, sizeof(Counters) * kCtrCap));
    PMSZ_HIP(hipHostMalloc(&c.h_ctr, 2 * sizeof(Counters), hipHostMallocDefault));
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
  return d;
}

dim3 dense_grid(const Grid3 &d) {
  return dim3((unsigned)((d.nx + kTX - 1) / kTX), (unsigned)((d.ny + kTY - 1) / kTY),
              (unsigned)((d.nz + kTZ - 1) / kTZ));
}

// Copies counter blocks k (and k+1 when two = true) into the pinned mirror.
int read_counters(Ctx &c, int k, bool two = false) {
  PMSZ_HIP(hipMemcpyAsync(c.h_ctr, c.ctr + k, (two ? 2 : 1) * sizeof(Counters),
                          hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  return PMSZ_OK;
}

template <typename T>
int correct_single(Ctx &c, const pmsz_params *p, const T *f, T *g, uint8_t *dc_out,
                   pmsz_stats *st) {
  const Grid3 d = grid_of(p->nx, p->ny, p->nz);
  int rc = ensure(c, d.n);
  if (rc) return rc;
  EditParams ep;
  ep.eb = p->abs_eb;
  ep.delta_d = p->abs_eb / p->q;
  ep.q = p->q;
  ep.clamp_mode = p->clamp_mode;
  const int64_t words = (d.n + 31) / 32 + 1;
  std::memset(st, 0, sizeof(*st));
  PMSZ_HIP(hipEventRecord(c.ev0, c.stream));
  PMSZ_HIP(hipMemsetAsync(c.tmark, 0, sizeof(uint32_t) * words, c.stream));
  PMSZ_HIP(hipMemsetAsync(c.wmark, 0, sizeof(uint32_t) * words, c.stream));
  PMSZ_HIP(hipMemsetAsync(c.ctr, 0, sizeof(Counters) * kCtrCap, c.stream));
  PMSZ_HIP(hipEventRecord(c.ev2, c.stream));
  k_dense_pass<T, true><<<dense_grid(d), kThreads, 0, c.stream>>>(f, g, c.code, c.dc, c.tmark,
                                                                    c.tlist, c.ctr, d);
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipEventRecord(c.ev1, c.stream));
  st->dense_passes = 1;
  double bytes = (double)d.n * (2.0 * sizeof(T) + 2.0);
  rc = read_counters(c, 0);
  if (rc) return rc;
  if (c.h_ctr->err) return PMSZ_EINVAL;
  {
    float dms = 0;
    PMSZ_HIP(hipEventElapsedTime(&dms, c.ev2, c.ev1));
    st->t_dense_s = dms * 1e-3;
  }
  int k = 0;        // counters slot of the running sweep
  int64_t j = 1;    // detections done
  int64_t applies = 0;
  int status = PMSZ_OK;
  uint32_t cnt[3] = {c.h_ctr->cnt[0], c.h_ctr->cnt[1], c.h_ctr->cnt[2]};
  uint32_t nt = c.h_ctr->ntargets;
  bool pending = false;  // an apply whose counters were not read back yet
  while (true) {
    bool zero = cnt[0] == 0 && cnt[1] == 0 && cnt[2] == 0;
    if (p->mode == PMSZ_LAZY) {
      if (zero) break;
      if (j >= p->max_sweeps) {
        // the sweep's edits still land (they belong to sweep j)
        status = PMSZ_NOT_CONVERGED;
      }
    } else {
      if (applies >= p->max_sweeps) {
        status = PMSZ_NOT_CONVERGED;
        break;
      }
    }
    // apply sweep j's targets and queue the next work list
    if (nt > 0) {
      unsigned blocks = (nt + 255) / 256;
      k_apply<T><<<blocks, 256, 0, c.stream>>>(g, f, c.dc, c.tmark, c.tlist, nt, c.wmark,
                                               c.wlist, c.ctr + k, d, ep);
      PMSZ_HIP(hipGetLastError());
      bytes += (double)nt * (3.0 * sizeof(T) + 2.0);
      st->total_targets += nt;
    }
    applies++;
    pending = nt > 0;
    if (status != PMSZ_OK) break;
    if (p->mode == PMSZ_SYNC && zero) break;
    // next detection over the work list
    int kn = k + 1;
    if (kn >= kCtrCap) {
      PMSZ_HIP(hipMemcpyAsync(c.ctr, c.ctr + k, sizeof(Counters), hipMemcpyDeviceToDevice,
                              c.stream));
      PMSZ_HIP(hipMemsetAsync(c.ctr + 1, 0, sizeof(Counters) * (kCtrCap - 1), c.stream));
      k = 0;
      kn = 1;
    }
    int64_t upper = std::min<int64_t>((int64_t)nt * kCand, d.n);
    if (upper > 0) {
      unsigned blocks = (unsigned)((upper + 255) / 256);
      k_detect_sparse<T><<<blocks, 256, 0, c.stream>>>(g, c.code, c.tmark, c.tlist, c.wmark,
                                                       c.wlist, c.ctr + k, c.ctr + kn, d);
      PMSZ_HIP(hipGetLastError());
    }
    // slot k holds the apply's work-list size, slot kn the new sweep's counters
    rc = read_counters(c, k, true);
    if (rc) return rc;
    pending = false;
    st->sparse_voxels += c.h_ctr[0].nwork;
    st->touched_voxels += c.h_ctr[0].touched;
    st->clamped_voxels += c.h_ctr[0].clamped;
    bytes += (double)c.h_ctr[0].nwork * (sizeof(T) + 1.0);
    k = kn;
    j++;
    cnt[0] = c.h_ctr[1].cnt[0];
    cnt[1] = c.h_ctr[1].cnt[1];
    cnt[2] = c.h_ctr[1].cnt[2];
    nt = c.h_ctr[1].ntargets;
    if (p->mode == PMSZ_SYNC && cnt[0] == 0 && cnt[1] == 0 && cnt[2] == 0) break;
  }
  if (pending) {
    rc = read_counters(c, k);
    if (rc) return rc;
    st->touched_voxels += c.h_ctr[0].touched;
    st->clamped_voxels += c.h_ctr[0].clamped;
  }
  PMSZ_HIP(hipEventRecord(c.ev1, c.stream));
  if (dc_out)
    PMSZ_HIP(hipMemcpyAsync(dc_out, c.dc, (size_t)d.n, hipMemcpyDeviceToDevice, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  float ms = 0;
  PMSZ_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  st->t_total_s = ms * 1e-3;
  st->status = status;
  st->final_cnt[0] = cnt[0];
  st->final_cnt[1] = cnt[1];
  st->final_cnt[2] = cnt[2];
  if (p->mode == PMSZ_LAZY) {
    st->edit_sweeps = j;
    st->check_sweeps = status == PMSZ_OK ? 1 : 0;
    st->outer_iters = status == PMSZ_OK ? 1 : 0;
    st->ite = j + 1;
  } else {
    st->edit_sweeps = applies;
    st->check_sweeps = applies;
    st->outer_iters = applies;
    st->ite = 2 * applies;
  }
  st->bytes_algorithmic = bytes;
  return status;
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
  PMSZ_HIP(hipMemsetAsync(c.ctr, 0, sizeof(Counters), c.stream));
  if (dtype == PMSZ_F32)
    k_dense_pass<float, false><<<dense_grid(d), kThreads, 0, c.stream>>>(
        (const float *)d_field, nullptr, d_codes, nullptr, nullptr, nullptr, c.ctr, d);
  else
    k_dense_pass<double, false><<<dense_grid(d), kThreads, 0, c.stream>>>(
        (const double *)d_field, nullptr, d_codes, nullptr, nullptr, nullptr, c.ctr, d);
  PMSZ_HIP(hipGetLastError());
  rc = read_counters(c, 0);
  if (rc) return rc;
  return c.h_ctr->err ? PMSZ_EINVAL : PMSZ_OK;
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
  EditParams ep;
  ep.eb = p->abs_eb;
  ep.delta_d = p->abs_eb / p->q;
  ep.q = p->q;
  ep.clamp_mode = p->clamp_mode;
  unsigned long long *dcount = (unsigned long long *)c.ctr;
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

const char *pmsz_version(void) { return "pmsz-mi355x 0.1"; }

}  // extern "C"
