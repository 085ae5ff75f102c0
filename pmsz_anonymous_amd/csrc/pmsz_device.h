// pmsz_device.h -- device-side building blocks shared by the correction kernels:
// grid geometry, per-sweep counters, the candidate scans, the detection rules and
// block-aggregated work-list appends.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pmsz_common.h"

namespace pmsz {

// Unsigned 32-bit division by an invariant divisor through one multiply-high
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, the round-up variant valid for every 32-bit n).
struct FastDiv {
  uint32_t m;
  int sh1, sh2;
  __host__ void init(uint32_t d) {
    int l = 0;
    while (((uint64_t)1 << l) < d) l++;
    m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    sh1 = l < 1 ? l : 1;
    sh2 = l - 1 > 0 ? l - 1 : 0;
  }
  __device__ __forceinline__ uint32_t div(uint32_t n) const {
    const uint32_t t = __umulhi(m, n);
    return (t + ((n - t) >> sh1)) >> sh2;
  }
};

struct Grid3 {
  int64_t nx, ny, nz;
  int64_t sxy;  // nx*ny
  int64_t n;    // nx*ny*nz
  FastDiv dnx, dsxy;  // division by nx and by nx*ny (32-bit voxel indices)
};

// (x, y, z) of a 32-bit voxel index.
__device__ __forceinline__ void decompose(uint32_t v, const Grid3 &d, int64_t &x, int64_t &y,
                                          int64_t &z) {
  const uint32_t zz = d.dsxy.div(v);
  const uint32_t r = v - zz * (uint32_t)d.sxy;
  const uint32_t yy = d.dnx.div(r);
  x = (int64_t)(r - yy * (uint32_t)d.nx);
  y = (int64_t)yy;
  z = (int64_t)zz;
}

// One domain of a run.  A single correction has one; a batch has one per field;
// a subdomain run has one per padded block (core + 1-cell ghost halo, the
// reference's padded subdomain, pMSz.cu:2709-2711).  Voxel ids used by the
// bitmaps and lists are (domain << shift) | local index.
struct Dom {
  const void *f;      // original field of the stored block
  void *g;            // working field of the stored block
  uint8_t *code;      // original codes (library-owned)
  uint8_t *dc;        // delta counters (library-owned)
  uint8_t *snap;      // ghost-sync snapshot of dc (padded blocks only)
  Grid3 d;            // extent of the stored block
  int32_t lo[3], hi[3];  // box of classified voxels [lo, hi): the core
  int32_t off[3];     // global coordinate of local cell lo (subdomains)
  int32_t cc[3];      // coordinates in the subdomain grid
  uint32_t unit_base; // first column unit of this domain
  uint32_t nux, nuy, nuz;  // column units of the box (waves per z chunk, 1, z chunks)
  uint32_t nhx, nyb;       // columns across x and y blocks of a z chunk
  uint32_t half;           // 1: half-wave columns (30 x), 0: wave columns (62 x)
  uint32_t zc;             // planes per column unit (16, 32 or 64)
  uint32_t shell_n;   // cells of the 2-layer shell (padded blocks)
  uint32_t padded;    // 1: cells holding the sentinel are missing neighbours
  uint32_t gbase;     // domain << shift
  uint32_t cbase;     // offset of the block's first cell in the run's target byte map
  uint32_t wpre;      // bitmap words of the domains before this one (gbase is 32-aligned)
  EditParams ep;      // bound, step, q and clamp mode of this domain's edits
};

// Per-domain state of a multi-domain run.
enum : uint32_t {
  kDomActive = 1,  // takes part in the next detection
  kDomEmit = 2,    // its targets of that detection are edited (else it only counts)
  kDomSend = 4,    // an edit touched its 2-layer boundary since the last ghost sync
};
enum : uint32_t { kDomRunning = 0, kDomConverged = 1, kDomCapped = 2 };

struct DomState {
  uint32_t cnt[2][3];   // counters of the detections of parity 0 / 1
  uint32_t flags;
  uint32_t sweeps;      // edit sweeps counted for this domain
  uint32_t status;
  uint32_t last_cnt[3]; // counters of its last detection
  uint32_t last_j;      // index of its last detection
  uint32_t pad_;
  unsigned long long targets;  // edits applied
  unsigned long long touched;  // voxels with dc > 0 at the end (core)
  unsigned long long clamped;  // voxels with dc == q at the end (core)
};

// Next-sweep decision written by the apply of sweep j into its counter block.
enum : uint32_t {
  kModeUnset = 0,
  kModeDense = 1,   // detection j+1 runs over every voxel (tile kernel)
  kModeSparse = 2,  // detection j+1 runs over the work list
  kModeTail = 3,    // the single-workgroup tail kernel takes over at sweep j
  kModeDone = 4,    // finished (converged or capped)
};

// Device-scope atomics on one address serialise (about 11 ns each on MI355X), so
// a counter that every block or wave of a launch adds to is kept in kStripes
// copies 16 bytes apart: a block adds to copy blockIdx.x % kStripes, readers sum.
constexpr int kStripes = 16;

// Counters of detection j and of the apply that follows it (one block per sweep,
// in a ring).  cs: false max / false min / direction (count_f_max, count_f_min,
// count_f_dir, pMSz.cu:174), by stripe; ntargets: entries of tlist; nwork:
// entries of wlist queued by apply j; mode: the decision of apply j.
struct Counters {
  uint32_t cs[kStripes][4];
  uint32_t ntargets;
  uint32_t nwork;
  uint32_t mode;
  uint32_t err;
  uint32_t nactive;  // multi-domain: domains taking part in detection j+1
  uint32_t ncapped;  // multi-domain: domains stopped by the sweep cap
  uint32_t pad[10];
};
static_assert(sizeof(Counters) == 320, "Counters layout");

// Slot 3 of each stripe: the targets of a full-volume detection that the
// streamed apply edited (no list counted them); not a violation counter.
constexpr int kCntStreamed = 3;
__device__ __forceinline__ void cnt_add(Counters &c, int q, uint32_t v) {
  if (v) atomicAdd(&c.cs[blockIdx.x % kStripes][q], v);
}
// Counter q of a detection (relaxed device-scope loads: valid after a grid
// barrier as well as after a kernel boundary).
__device__ __forceinline__ uint32_t cnt_get(const Counters &c, int q) {
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kStripes; k++)
    s += __hip_atomic_load(&c.cs[k][q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return s;
}
// Nonzero iff any counter of the detection is (plain loads: after a kernel
// boundary; uniform addresses, so scalar loads).
__device__ __forceinline__ uint32_t cnt_any(const Counters &c) {
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kStripes; k++) s |= c.cs[k][0] | c.cs[k][1] | c.cs[k][2];
  return s;
}
// Violations of a detection (plain loads: after a kernel boundary).
__device__ __forceinline__ uint32_t cnt_sum(const Counters &c) {
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kStripes; k++) s += c.cs[k][0] + c.cs[k][1] + c.cs[k][2];
  return s;
}
__device__ __forceinline__ uint32_t cnt_any_agent(const Counters &c) {
  return cnt_get(c, 0) | cnt_get(c, 1) | cnt_get(c, 2);
}
__host__ inline uint32_t cnt_host(const Counters &c, int q) {
  uint32_t s = 0;
  for (int k = 0; k < kStripes; k++) s += c.cs[k][q];
  return s;
}

// Schedule parameters of a run.
struct Sched {
  int mode;  // PMSZ_LAZY / PMSZ_SYNC
  uint32_t max_sweeps;
  uint32_t dense_min_work;  // 15 * targets at or above this -> next detection dense
  uint32_t tail_ok;         // the cooperative tail may take over
  uint32_t tail_targets;    // the tail takes over at or below this many targets
  uint32_t solo_targets;    // the tail continues in one workgroup at or below this many
  uint32_t dense_ok;        // 0: detections run over work lists only
  uint32_t stream_min;      // one domain: the targets of a full-volume detection with at
                            // least this many violations are edited by one pass over the
                            // target byte map instead of a list
  uint32_t append;          // work-list sweeps list claimed bits (else bitmaps are compacted)
  uint32_t tag0;            // detection jd marks its targets with detect_tag(jd + tag0)
};

constexpr uint32_t kErrInput = 1u, kErrBarrier = 2u;

// Run-wide state, written only by kernels that end a phase.
enum : int { kAccTargets = 0, kAccTouched = 1, kAccClamped = 2, kAccSparse = 3 };
struct RunState {
  uint32_t tail_j;     // sweep whose apply the tail kernel starts with (0: none)
  uint32_t done;       // 0 running, 1 converged, 2 sweep cap reached
  uint32_t last_j;     // detections performed when done
  uint32_t err;        // bit 0 (kErrInput): NaN or sentinel value in an input;
                       // bit 1 (kErrBarrier): a grid barrier of the tail timed out
  unsigned long long dense_passes;
  uint32_t bar_epoch;  // last epoch of the tail's grid barrier in this run
  uint32_t last_mode;  // the mode the latest apply decided
  unsigned long long pad[4];
  // by stripe: edits applied, voxels edited at least once, voxels clamped,
  // voxels re-examined by work-list sweeps
  unsigned long long acc[kStripes][4];
};

__device__ __forceinline__ void acc_add(RunState *r, int k, unsigned long long v) {
  if (v) atomicAdd(&r->acc[blockIdx.x % kStripes][k], v);
}
__host__ inline unsigned long long acc_host(const RunState &r, int k) {
  unsigned long long s = 0;
  for (int q = 0; q < kStripes; q++) s += r.acc[q][k];
  return s;
}

constexpr int kCtrRing = 256;  // counter blocks kept; consecutive sweeps use distinct slots

__device__ __forceinline__ int64_t rank_off(int r, const Grid3 &d) {
  return (int64_t)rank_dx(r) + d.nx * (int64_t)rank_dy(r) + d.sxy * (int64_t)rank_dz(r);
}

// Bit r set iff candidate r of (x, y, z) lies in the domain; out-of-domain
// neighbours are the reference's INVALID halo (pMSz.cu:709).
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

__device__ __forceinline__ uint32_t valid_mask(uint32_t x, uint32_t y, uint32_t z,
                                               const Grid3 &d) {
  return valid_mask((int64_t)x, (int64_t)y, (int64_t)z, d);
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

// Detection at one voxel from its original codes (oa, od) and the current
// steepest candidates (da, dd): iscriticle data_type=0, pMSz.cu:786-898.
// Returns the number of target ranks written to t[], counters in c[].
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

// The same rules without branches: at most one target on the ascent side (t1:
// the new top neighbour, the false maximum, or the wrong steepest-ascent
// neighbour) and one on the descent side (t2: the lost minimum, the true descent
// neighbour, or the true steepest-descent neighbour).  v1 / v2 tell whether each
// exists.  Counters as detect_targets.
__device__ __forceinline__ void detect_pair(int oa, int od, int da, int dd, uint32_t (&c)[3],
                                            int &t1, bool &v1, int &t2, bool &v2) {
  const bool omax = oa == kSelf;
  const bool omin = od == kSelf && !omax;
  const bool fm = (da == kSelf) != omax, fn = (dd == kSelf) != omin;
  const bool crit = fm || fn;
  const bool dira = !crit && da != oa, dird = !crit && dd != od;
  c[0] += fm ? 1u : 0u;
  c[1] += fn ? 1u : 0u;
  c[2] += crit ? 1u : (dira ? 1u : 0u) + (dird ? 1u : 0u);
  t1 = (fm && !omax) ? kSelf : da;
  v1 = fm || dira;
  t2 = (fn && omin) ? kSelf : od;
  v2 = fn || dird;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Workgroup barrier ordering LDS only: global stores issued before it need not
// complete, so loads in flight (prefetches) are not drained at the barrier.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Set the mark bit of v; true if it was clear (the caller then owns v).
__device__ __forceinline__ bool claim(uint32_t *marks, uint32_t v) {
  uint32_t bit = 1u << (v & 31u);
  uint32_t old = atomicOr(&marks[v >> 5], bit);
  return (old & bit) == 0;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block-wide exclusive scan of per-thread counts with one global atomicAdd on
// `counter` for the block's total; returns this thread's first list position.
// All threads of the block must call it (it holds two barriers).
template <int kBlock>
__device__ __forceinline__ uint32_t block_reserve(uint32_t n, uint32_t *counter) {
  constexpr int kWaves = kBlock / 64;
  __shared__ uint32_t s_off[kWaves];
  __shared__ uint32_t s_base;
  const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
  uint32_t incl = wave_incl_scan(n);
  if (lane == 63) s_off[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < kWaves; w++) {
      uint32_t v = s_off[w];
      s_off[w] = run;
      run += v;
    }
    s_base = run ? atomicAdd(counter, run) : 0u;
  }
  __syncthreads();
  uint32_t pos = s_base + s_off[wave] + incl - n;
  __syncthreads();
  return pos;
}

// A block's list entries staged in LDS across the iterations of a grid-stride
// loop and copied to the global list in large pieces, so the block makes one
// reservation (a same-address atomic) per kCap entries instead of one per
// iteration.  Entries keep no order.
template <int kBlock, int kCap>
struct StageLds {
  uint32_t buf[kCap];
  uint32_t off[kBlock / 64];
  uint32_t fill, total, base;
};

template <int kBlock, int kCap>
__device__ __forceinline__ void stage_init(StageLds<kBlock, kCap> &st) {
  if (threadIdx.x == 0) st.fill = 0;
  __syncthreads();
}

// Copies the staged entries to list[atomicAdd(counter, fill) ...]; every thread
// of the block calls it.
template <int kBlock, int kCap>
__device__ __forceinline__ void stage_flush(StageLds<kBlock, kCap> &st, uint32_t *list,
                                            uint32_t *counter) {
  __syncthreads();  // the staged writes are complete
  const uint32_t fill = st.fill;
  if (threadIdx.x == 0) st.base = fill ? atomicAdd(counter, fill) : 0u;
  __syncthreads();
  const uint32_t base = st.base;
  for (uint32_t k = threadIdx.x; k < fill; k += kBlock) list[base + k] = st.buf[k];
  __syncthreads();  // every thread has read fill, base and the entries
  if (threadIdx.x == 0) st.fill = 0;
  __syncthreads();
}

// Room for n entries of this thread (n * kBlock <= kCap); returns the first
// buffer position.  Flushes first when the block's entries would not fit.
// Every thread of the block calls it; the caller then writes st.buf[pos ...].
template <int kBlock, int kCap>
__device__ __forceinline__ uint32_t stage_reserve(StageLds<kBlock, kCap> &st, uint32_t n,
                                                  uint32_t *list, uint32_t *counter) {
  constexpr int kWaves = kBlock / 64;
  const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
  const uint32_t incl = wave_incl_scan(n);
  if (lane == 63) st.off[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < kWaves; w++) {
      const uint32_t v = st.off[w];
      st.off[w] = run;
      run += v;
    }
    st.total = run;
  }
  __syncthreads();
  const uint32_t total = st.total;
  if (st.fill + total > (uint32_t)kCap) stage_flush(st, list, counter);  // uniform
  const uint32_t pos = st.fill + st.off[wave] + incl - n;
  __syncthreads();  // every thread has read fill and off
  if (threadIdx.x == 0) st.fill += total;
  return pos;
}

// Block-wide sums of three detection counters added to dst with one atomic each.
template <int kBlock>
__device__ __forceinline__ void block_add3(const uint32_t (&c)[3], Counters &dst) {
  constexpr int kWaves = kBlock / 64;
  __shared__ uint32_t s_c[3][kWaves];
  const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    uint32_t s = wave_sum(c[k]);
    if (lane == 0) s_c[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    uint32_t s = 0;
    for (int w = 0; w < kWaves; w++) s += s_c[threadIdx.x][w];
    cnt_add(dst, (int)threadIdx.x, s);
  }
  __syncthreads();
}

// Bit r set iff candidate r of (x, y, z) lies in the box [lo, hi).
__device__ __forceinline__ uint32_t box_mask(int64_t x, int64_t y, int64_t z, const int32_t *lo,
                                             const int32_t *hi) {
  uint32_t m = 0;
#pragma unroll
  for (int r = 0; r < kCand; r++) {
    int64_t a = x + rank_dx(r), b = y + rank_dy(r), c = z + rank_dz(r);
    bool ok = a >= lo[0] && a < hi[0] && b >= lo[1] && b < hi[1] && c >= lo[2] && c < hi[2];
    m |= (ok ? 1u : 0u) << r;
  }
  return m;
}

// Domain of a voxel id and its local index.
template <bool kMulti>
__device__ __forceinline__ const Dom &dom_of(uint32_t gid, const Dom &d0, const Dom *doms,
                                             uint32_t shift, uint32_t &k, uint32_t &local) {
  if (!kMulti) {
    k = 0;
    local = gid;
    return d0;
  }
  k = gid >> shift;
  local = gid & ((1u << shift) - 1u);
  return doms[k];
}

// Adds three per-lane counters of lanes with `has` to ds[dom].cnt[par], one atomic
// per distinct domain in the wave.  Every lane of the wave must call it.
__device__ __forceinline__ void dom_add3(bool has, uint32_t dom, const uint32_t (&c)[3],
                                         DomState *ds, uint32_t par) {
  const uint32_t lane = lane_id();
  uint64_t pending = __ballot(has && (c[0] | c[1] | c[2]) != 0);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t dl = (uint32_t)__shfl((int)dom, leader);
    const bool mine = ((pending >> lane) & 1ull) && dom == dl;
    const uint64_t m = __ballot(mine);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const uint32_t s = wave_sum(mine ? c[q] : 0u);
      if (lane == (uint32_t)leader && s) atomicAdd(&ds[dl].cnt[par][q], s);
    }
    pending &= ~m;
  }
}

// Adds a per-lane count to ds[dom].targets, one atomic per distinct domain.
__device__ __forceinline__ void dom_add_targets(bool has, uint32_t dom, DomState *ds) {
  const uint32_t lane = lane_id();
  uint64_t pending = __ballot(has);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t dl = (uint32_t)__shfl((int)dom, leader);
    const uint64_t m = __ballot(((pending >> lane) & 1ull) && dom == dl);
    if (lane == (uint32_t)leader) atomicAdd(&ds[dl].targets, (unsigned long long)__popcll(m));
    pending &= ~m;
  }
}

// Cells of the 2-layer shell of a P0 x P1 x P2 block (every cell with a
// coordinate in {0, 1, P-2, P-1}); the whole block when an extent is below 4.
__host__ __device__ inline uint32_t shell_count(uint32_t P0, uint32_t P1, uint32_t P2) {
  if (P0 < 4 || P1 < 4 || P2 < 4) return P0 * P1 * P2;
  return 4 * P0 * P1 + 4 * P0 * (P2 - 4) + 4 * (P1 - 4) * (P2 - 4);
}
__host__ __device__ inline void shell_cell(uint32_t i, uint32_t P0, uint32_t P1, uint32_t P2,
                                           uint32_t &x, uint32_t &y, uint32_t &z) {
  if (P0 < 4 || P1 < 4 || P2 < 4) {
    x = i % P0;
    y = (i / P0) % P1;
    z = i / (P0 * P1);
    return;
  }
  const uint32_t a = 4 * P0 * P1;
  if (i < a) {  // planes z in {0, 1, P2-2, P2-1}
    const uint32_t zi = i / (P0 * P1), r = i - zi * P0 * P1;
    z = zi < 2 ? zi : P2 - 4 + zi;
    y = r / P0;
    x = r - y * P0;
    return;
  }
  i -= a;
  const uint32_t b = 4 * P0 * (P2 - 4);
  if (i < b) {  // rows y in {0, 1, P1-2, P1-1} of planes 2..P2-3
    const uint32_t yi = i / (P0 * (P2 - 4)), r = i - yi * P0 * (P2 - 4);
    y = yi < 2 ? yi : P1 - 4 + yi;
    z = 2 + r / P0;
    x = r - (z - 2) * P0;
    return;
  }
  i -= b;  // columns x in {0, 1, P0-2, P0-1} of rows 2..P1-3, planes 2..P2-3
  const uint32_t c = (P1 - 4) * (P2 - 4);
  const uint32_t xi = i / c, r = i - xi * c;
  x = xi < 2 ? xi : P0 - 4 + xi;
  z = 2 + r / (P1 - 4);
  y = 2 + r - (z - 2) * (P1 - 4);
}

// Bits [p, p+32) of the 96-bit little-endian value (w0, w1, w2); bits outside
// [0, 96) read as zero.  p in [-31, 95].
__device__ __forceinline__ uint32_t bits32_at(uint32_t w0, uint32_t w1, uint32_t w2, int p) {
  if (p < 0) return w0 << (-p);
  if (p >= 96) return 0u;
  int q = p >> 5, r = p & 31;
  uint32_t lo = q == 0 ? w0 : (q == 1 ? w1 : w2);
  uint32_t hi = q == 0 ? w1 : (q == 1 ? w2 : 0u);
  return r ? ((lo >> r) | (hi << (32 - r))) : lo;
}

}  // namespace pmsz
