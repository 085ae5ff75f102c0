// pmsz_engine.hip -- gfx950 kernels and the C-ABI (include/pmsz.h) of the
// correction path: one field, a batch of independent fields, or one field split
// into padded subdomains with ghost synchronisation.
//
// Data in HBM, per domain (one field of a batch, or one padded subdomain block):
//   f, g        T[n]       original / working field, x fastest (the caller's for
//                          a field; library-owned padded blocks for subdomains)
//   code        u8[n]      original steepest-ascent rank (low nibble) and
//                          steepest-descent rank (high nibble); rank 7 = the voxel
//   dc          u8[n]      delta counters (delta_counter, pMSz.cu:173)
//   snap        u8[n]      subdomains: the dc of the 2-layer shell at a ghost sync
// Per run:
//   tbyte       u8[]       target byte map of full-volume detections: detection
//                          jd marks each target cell with its tag (1..255)
//   tmark       u32[]      targets of a work-list sweep, one bit per voxel id: the
//                          per-sweep edit dedup (edits[] + atomicCAS, pMSz.cu:795)
//   wmark       u32[]      voxels queued for the next work-list sweep
//   tlist/wlist u32[]      the same sets as lists of voxel ids
//   ring        Counters[] per-sweep counters (striped) and schedule decisions
//   run         RunState   phase hand-over between kernels, run totals (striped)
//   doms, ds    Dom[], DomState[]  per-domain description and state
// A voxel id is (domain << shift) | local index; a single field is domain 0 with
// id = index.
//
// Snapshot semantics: a sweep detects on the start-of-sweep g, writing only
// marks, lists and counters; a separate phase then applies one edit per target
// (applyDeltaBuffer_local, pMSz.cu:623-645).
//
// Schedule.  The original labeling and detection 1 are full passes
// (pmsz_dense.h).  After that, the apply of sweep j decides on the device how
// detection j+1 runs:
//   - over every voxel (column kernel, targets to tbyte, then compacted) while
//     the targets are dense;
//   - over the closed stencil neighbourhoods of sweep j's targets otherwise: a
//     voxel's classification depends only on g over its closed neighbourhood and
//     the stencil is symmetric, so no other voxel can change -- the counters and
//     targets equal those of a full sweep;
//   - in the tail kernel once the targets are few: one cooperative launch runs
//     the remaining sweeps with grid barriers instead of kernel launches.
// The host enqueues a chain of sweeps without waiting and synchronises once per
// chain.
//
// Several domains run in lockstep in the same launches.  Each keeps its own
// counters and flags: it takes part in detections while its last detection
// found violations, exactly as its own loop would (the domains are independent
// between ghost syncs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/pmsz.h"
#include "pmsz_dense.h"
#include "pmsz_device.h"
#include "pmsz_edits.h"
#include "pmsz_naive.h"
#include "pmsz_codec.h"
#include "pmsz_dqstream.h"

namespace pmsz {


constexpr int kApplyThreads = 256;
constexpr int kTailThreads = 256;
// the tail takes over once a sweep has at most this many targets: heavier work-list
// sweeps run as separate launches, which bring more threads to their atomics
// most blocks of the tail (its grid barriers cost about 45 ns per block on MI355X;
// the tuning picks the grid, 32 blocks of 256 threads by default)
constexpr int kTailBlocks = 256;
constexpr size_t kBarBytes = sizeof(uint32_t) * 16 * (kTailBlocks + 1);
constexpr int kSyncThreads = 256;
constexpr uint32_t kMaxDomains = 1024;  // domain bookkeeping: one thread each
static_assert(kMaxDomains <= (uint32_t)kCompactDoms, "compaction domain table");
constexpr int kMaxContexts = 8;         // live contexts, the default one included

// Stencil directions of the reference in its order d = 0..13 (directions_host,
// pMSz.cu:186-193): component a of direction d.  The opposite of d is d ^ 1.
__host__ __device__ constexpr int dir_comp(int d, int a) {
  return a == 0 ? (d == 0 || d == 7 || d == 9 || d == 13 ? 1
                   : d == 1 || d == 6 || d == 8 || d == 12 ? -1 : 0)
       : a == 1 ? (d == 2 || d == 6 || d == 10 || d == 12 ? 1
                   : d == 3 || d == 7 || d == 11 || d == 13 ? -1 : 0)
                : (d == 4 || d == 8 || d == 10 || d == 12 ? 1
                   : d == 5 || d == 9 || d == 11 || d == 13 ? -1 : 0);
}

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Decisions of the apply of sweep j (identical in every thread): the mode of
// detection j+1, or kModeTail, or kModeDone with *done = 1 (converged) / 2 (cap).

// One domain: from the counters of detection j.  kAtomic: relaxed device-scope
// loads, for counters written by other blocks of the same launch (the tail);
// plain loads suffice after a kernel boundary.
template <bool kAtomic>
__device__ __forceinline__ uint32_t decide(const Counters &cj, uint32_t j, const Sched &s,
                                           bool tail_allowed, uint32_t *done) {
  const bool zero = (kAtomic ? cnt_any_agent(cj) : cnt_any(cj)) == 0;
  const uint32_t nt = kAtomic ? ld_agent(&cj.ntargets) : cj.ntargets;
  *done = 0;
  if (zero) {
    *done = 1;
    return kModeDone;
  }
  if (s.mode == PMSZ_SYNC && j > s.max_sweeps) {
    *done = 2;
    return kModeDone;
  }
  const bool dense = s.dense_ok && (uint64_t)nt * kCand >= s.dense_min_work;
  if (tail_allowed && !dense && nt <= s.tail_targets) return kModeTail;
  if (s.mode == PMSZ_LAZY && j >= s.max_sweeps) {
    *done = 2;  // the edits of sweep j still land (they belong to sweep j)
    return kModeDone;
  }
  return dense ? kModeDense : kModeSparse;
}

// One domain, the targets of detection j streamed from the byte map (their
// number is counted while they are edited): detection j+1 runs over every voxel.
__device__ __forceinline__ uint32_t decide_stream(const Counters &cj, uint32_t j, const Sched &s,
                                                  uint32_t *done) {
  *done = 0;
  if (cnt_any(cj) == 0) {
    *done = 1;
    return kModeDone;
  }
  if ((s.mode == PMSZ_SYNC && j > s.max_sweeps) || (s.mode == PMSZ_LAZY && j >= s.max_sweeps)) {
    *done = 2;  // lazy: the edits of sweep j still land
    return kModeDone;
  }
  return kModeDense;
}

// Several domains: from the domain bookkeeping after detection j (the caps are
// per domain, applied there).  The targets of detection j are always applied:
// a domain that must not edit them emitted none.
__device__ __forceinline__ uint32_t decide_multi(const Counters &cj, const Sched &s,
                                                 bool tail_allowed, uint32_t *done) {
  *done = 0;
  if (ld_agent(&cj.nactive) == 0) {
    *done = ld_agent(&cj.ncapped) ? 2 : 1;
    return kModeDone;
  }
  const uint32_t nt = ld_agent(&cj.ntargets);
  const bool dense = s.dense_ok && (uint64_t)nt * kCand >= s.dense_min_work;
  if (tail_allowed && !dense && nt <= s.tail_targets) return kModeTail;
  return dense ? kModeDense : kModeSparse;
}

// ---------------------------------------------------------------------------
// Domain bookkeeping after a detection (pMSz.cu:1943-1989 per subdomain / field,
// sync_pMSz.cu:2102-2146 in sync mode).
enum : uint32_t {
  kOpDetect = 0,    // after a detection of the running loop
  kOpPreCheck = 1,  // before the check detection that follows a ghost sync
  kOpCheck = 2,     // after that check detection
};

__device__ __forceinline__ void dom_update(DomState &st, uint32_t jd, uint32_t op,
                                           const Sched &s, uint32_t emit_ok, uint32_t &active,
                                           uint32_t &capped) {
  const uint32_t par = jd & 1u;
  uint32_t fl = ld_agent(&st.flags);
  if (op == kOpPreCheck) {
    // every domain is checked; one whose loop may not run another sweep only counts
    const bool emit = emit_ok && st.sweeps < s.max_sweeps;
    fl = kDomActive | (emit ? kDomEmit : 0u);
    for (int q = 0; q < 3; q++) st.cnt[par][q] = 0;
  } else if (fl & kDomActive) {
    uint32_t c[3];
    for (int q = 0; q < 3; q++) c[q] = ld_agent(&st.cnt[par][q]);
    const bool nonzero = (c[0] | c[1] | c[2]) != 0;
    const bool emit = (fl & kDomEmit) != 0;
    // an edit sweep of this domain: every emitting detection of its loop; after a
    // ghost sync, the check doubles as the first sweep of the next inner loop
    if (emit && (op == kOpDetect || nonzero)) st.sweeps++;
    for (int q = 0; q < 3; q++) st.last_cnt[q] = c[q];
    st.last_j = jd;
    uint32_t nf = fl & kDomSend;
    if (!nonzero) {
      st.status = kDomConverged;
    } else if (!emit) {
      st.status = kDomCapped;
    } else if (s.mode == PMSZ_LAZY && st.sweeps >= s.max_sweeps) {
      st.status = kDomCapped;  // its edits of this sweep land; it sweeps no more
    } else {
      st.status = kDomRunning;
      // sync mode: the detection after the last permitted edit sweep only counts
      const bool next_emit = !(s.mode == PMSZ_SYNC && st.sweeps >= s.max_sweeps);
      nf |= kDomActive | (next_emit ? kDomEmit : 0u);
    }
    fl = nf;
  }
  __hip_atomic_store(&st.flags, fl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (op != kOpPreCheck)
    for (int q = 0; q < 3; q++) st.cnt[par ^ 1u][q] = 0;
  active = (fl & kDomActive) ? 1u : 0u;
  capped = st.status == kDomCapped ? 1u : 0u;
}

// Runs after detection jd when the apply of sweep jprev let it run (gate), or
// unconditionally (gate = 0).
__global__ __launch_bounds__(kMaxDomains) void k_domains(DomState *__restrict__ ds, uint32_t ndom,
                                                         Counters *__restrict__ ring,
                                                         uint32_t jprev, uint32_t gate,
                                                         uint32_t jd, uint32_t op, Sched s,
                                                         uint32_t emit_ok) {
  if (gate) {
    const uint32_t m = ring[jprev % kCtrRing].mode;
    if (m != kModeDense && m != kModeSparse) return;
  }
  __shared__ uint32_t s_act, s_cap;
  if (threadIdx.x == 0) {
    s_act = 0;
    s_cap = 0;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < ndom; k += blockDim.x) {
    uint32_t a, c;
    dom_update(ds[k], jd, op, s, emit_ok, a, c);
    if (a) atomicAdd(&s_act, 1u);
    if (c) atomicAdd(&s_cap, 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0 && op != kOpPreCheck) {
    ring[jd % kCtrRing].nactive = s_act;
    ring[jd % kCtrRing].ncapped = s_cap;
  }
}

// ---------------------------------------------------------------------------
// Shared pieces of the apply and work-list phases.

// Step budgets, for one field with q <= 15 (EditParams::budgets; the ghost syncs
// of subdomain runs and the counter scans of batches read plain counters).  A
// delta counter's high nibble then holds a budget B: the low nibble d counts the
// edits, and while d < B the next edit is the plain step g - delta, which was
// proved to honour the bound when B was set, so it needs no f.  Every edit that
// takes the exact path (it reads f) sets the cell's budget; counters start at 0
// (no budget).  Counters are handed out with the nibble cleared.
__host__ __device__ inline bool budgets_on(const EditParams &ep) { return ep.budgets != 0; }

// The budget of a cell with counter d < q, value gv and original fv: the largest
// B <= q - 1 such that every edit made while the counter is below B is the plain
// step and honours the bound in both clamp modes.  With h_0 = gv and
// h_j = fl(h_{j-1} - delta), each h_j lies within j * (delta + M*u) below gv (u:
// half an ulp relative to M >= |h_j|, plus the smallest subnormal), and no higher
// than the first step.  So the steps land in [f - eb', f + eb'] with eb' =
// eb (1 - 2^-20), which lies below T's rounded bound and its predecessor (the
// reference mode's test) and below eb (the bound-safe test).  A margin covers the
// double arithmetic here.  Returns 0 when nothing can be proved.
template <typename T>
__device__ __forceinline__ uint32_t step_budget(T fv, T gv, uint32_t d, const EditParams &ep) {
  const uint32_t qm1 = (uint32_t)ep.q - 1u;
  if (d >= qm1) return 0;  // the next edit clamps
  const double dl = (double)(T)ep.delta_d;
  const double f = (double)fv, g = (double)gv;
  const double u = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
  const double tiny = sizeof(T) == 4 ? 0x1p-149 : 0x1p-1074;
  const double qd = (double)ep.q * dl;
  const double err = (fabs(g) + qd) * u + tiny;                   // rounding of one step
  const double marg = (fabs(g) + fabs(f) + ep.eb + qd) * 0x1p-48;  // of this arithmetic
  const double ebp = ep.eb * (1.0 - 0x1p-20);
  // every step lowers g, and the first one lands at or below f + eb'
  if (!(dl > 4.0 * err) || !(g - f - dl + err + marg <= ebp)) return 0;
  const double room = g - f + ebp - marg;  // how far g may go down
  if (!(room > 0.0)) return 0;
  // v_rcp_f64 is accurate to about 2^-23 (2^29 ulp); the quotient gives up 2^-20.
  // A subnormal step gets no budget (the reciprocal may flush it).
  if (!(dl + err >= 0x1p-1022)) return 0;
  const double k = floor(room * __builtin_amdgcn_rcp(dl + err) * (1.0 - 0x1p-20));
  const uint32_t b = k >= (double)qm1 ? qm1 : (uint32_t)k;
  return d + b < qm1 ? d + b : qm1;
}

// The same bound for float fields in float arithmetic (half the registers of the
// double version): every float operation here rounds by at most 2^-24 of the
// magnitudes involved, which the margin of 2^-20 of them covers, and the quotient
// gives up 2^-16 (v_rcp_f32 is exact to 1 ulp).  eb' is formed in double and
// rounded to float, which keeps it below eb (1 - 2^-20).
template <>
__device__ __forceinline__ uint32_t step_budget<float>(float fv, float gv, uint32_t d,
                                                       const EditParams &ep) {
  const uint32_t qm1 = (uint32_t)ep.q - 1u;
  if (d >= qm1) return 0;
  const float dl = (float)ep.delta_d;  // the step the edits take
  const float ebp = (float)(ep.eb * (1.0 - 0x1p-19));  // at most eb (1 - 2^-20)
  const float qd = (float)ep.q * dl;
  const float err = 2.0f * ((fabsf(gv) + qd) * 0x1p-24f) + 0x1p-149f;
  const float marg = (fabsf(gv) + fabsf(fv) + ebp + qd) * 0x1p-20f;
  if (!(dl > 4.0f * err) || !(gv - fv - dl + err + marg <= ebp)) return 0;
  const float room = gv - fv + ebp - marg;
  if (!(room > 0.0f)) return 0;
  if (!(dl + err >= 0x1p-126f)) return 0;  // subnormal: the reciprocal may flush it
  const float k = floorf(room * __builtin_amdgcn_rcpf(dl + err) * (1.0f - 0x1p-16f));
  const uint32_t b = k >= (float)qm1 ? qm1 : (uint32_t)k;
  return d + b < qm1 ? d + b : qm1;
}

// The budgets of n cells (the test hook pmsz_test_step_budget).
template <typename T>
__global__ void k_test_step_budget(int64_t n, const T *f, const T *g, const uint8_t *d,
                                   EditParams ep, uint8_t *out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (uint8_t)step_budget<T>(f[i], g[i], d[i], ep);
}

// Whether any of the cells picked by the 4-bit mask m, whose counter bytes are
// the bytes of cw, is not within its step budget (takes the exact edit).
__device__ __forceinline__ bool exact_any(uint32_t cw, uint32_t m) {
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t b = (cw >> (8 * k)) & 0xffu;
    any |= ((m >> k) & 1u) && (b & 15u) >= (b >> 4);
  }
  return any;
}

// One edit (applyDeltaBuffer_local) of a cell with counter byte c0 and value gv:
// within its budget the plain step, otherwise the exact edit with fv() (the only
// read of f) and, with kSet, the cell's new budget (without, the old one stays; it
// is spent).  Budgets are set where the cells are edited again later: the edits
// of full-volume detections' targets.  Returns the new counter byte; counts first
// edits and clamps.
template <typename T, bool kSet, typename FV>
__device__ __forceinline__ uint8_t edit_cell(T &gv, uint8_t c0, FV fv, const EditParams &ep,
                                             uint32_t &touched, uint32_t &clamped) {
  const bool bon = budgets_on(ep);
  const uint32_t d = bon ? c0 & 15u : c0;
  uint32_t bud = bon ? (uint32_t)(c0 >> 4) : 0u;
  uint8_t c = (uint8_t)d;
  if (d < bud) {
    gv = gv - (T)ep.delta_d;
    c = (uint8_t)(d + 1u);
  } else {
    const T f = fv();
    edit_value<T>(f, gv, c, ep);
    if (kSet && bon) bud = step_budget<T>(f, gv, c, ep);
  }
  touched += d == 0;
  clamped += (c == ep.q && d != (uint32_t)ep.q);
  return bon ? (uint8_t)(c | (bud << 4)) : c;
}

// One edit on local cell t, where edits are latency-bound (work lists, the tail,
// batches), in two parts: edit_load issues the cell's loads -- f together with g
// and the counter, so that an edit outside its budget waits for no second load
// (one within its budget ignores it) -- and edit_store edits and stores.  A
// wave's vector memory operations complete in issue order, so work that does not
// depend on the edit (the claims of the target's neighbourhood) is issued between
// the two: it then waits neither for the cell's loads to be used nor behind the
// edit's stores.
template <typename T>
struct CellLoad {
  T gv, fv;
  uint8_t c0;
};
template <typename T>
__device__ __forceinline__ CellLoad<T> edit_load(const T *g, const T *f, const uint8_t *dc,
                                                 uint32_t t) {
  CellLoad<T> e;
  e.gv = g[t];
  e.c0 = dc[t];
  e.fv = f[t];
  return e;
}
template <typename T>
__device__ __forceinline__ void edit_store(CellLoad<T> e, T *g, uint8_t *dc, uint32_t t,
                                           const EditParams &ep, uint32_t &touched,
                                           uint32_t &clamped) {
  asm volatile("" ::"v"(e.fv));  // f arrives with g, not in the exact path
  const uint8_t c = edit_cell<T, false>(e.gv, e.c0, [&] { return e.fv; }, ep, touched, clamped);
  g[t] = e.gv;
  dc[t] = c;
}
template <typename T>
__device__ __forceinline__ void apply_one(T *g, const T *f, uint8_t *dc, uint32_t t,
                                          const EditParams &ep, uint32_t &touched,
                                          uint32_t &clamped) {
  edit_store<T>(edit_load<T>(g, f, dc, t), g, dc, t, ep, touched, clamped);
}

// Claims the classified voxels of the closed stencil neighbourhood of local cell
// t = (x, y, z) in wmark: queue_issue issues every claim (old: the atomics'
// results), queue_collect returns the ids whose claim was new.
__device__ __forceinline__ void queue_issue(uint32_t t, int64_t x, int64_t y, int64_t z,
                                            const Dom &D, uint32_t *wmark,
                                            uint32_t (&old)[kCand]) {
  const uint32_t vm = box_mask(x, y, z, D.lo, D.hi);
#pragma unroll
  for (int r = 0; r < kCand; r++) {
    const uint32_t v = D.gbase + (uint32_t)((int64_t)t + rank_off(r, D.d));
    old[r] = ((vm >> r) & 1u) ? atomicOr(&wmark[v >> 5], 1u << (v & 31u)) : 0xffffffffu;
  }
}
__device__ __forceinline__ int queue_collect(uint32_t t, const Dom &D,
                                             const uint32_t (&old)[kCand],
                                             uint32_t (&u)[kCand]) {
  int n = 0;
#pragma unroll
  for (int r = 0; r < kCand; r++) {
    const uint32_t v = D.gbase + (uint32_t)((int64_t)t + rank_off(r, D.d));
    if (!((old[r] >> (v & 31u)) & 1u)) u[n++] = v;
  }
  return n;
}

// Marks the classified voxels of the closed stencil neighbourhood of local cell
// t = (x, y, z) of domain D in wmark, without looking at the old bits: the
// stencil's 7 rows of consecutive x take one atomic OR each (two where a row
// crosses a word; ids of a domain are consecutive from the 32-aligned D.gbase).
__device__ __forceinline__ void mark_neighbourhood(uint32_t t, int64_t x, int64_t y, int64_t z,
                                                   const Dom &D, uint32_t *wmark) {
  const uint32_t vm = box_mask(x, y, z, D.lo, D.hi);
  constexpr int kRowFirst[7] = {0, 2, 4, 6, 9, 11, 13};  // first rank of each row
  constexpr int kRowLen[7] = {2, 2, 2, 3, 2, 2, 2};
#pragma unroll
  for (int k = 0; k < 7; k++) {
    const uint32_t rb = (vm >> kRowFirst[k]) & ((1u << kRowLen[k]) - 1u);
    if (!rb) continue;
    const uint32_t v0 = D.gbase + (uint32_t)((int64_t)t + rank_off(kRowFirst[k], D.d));
    const uint64_t m = (uint64_t)rb << (v0 & 31u);
    atomicOr(&wmark[v0 >> 5], (uint32_t)m);
    if (m >> 32) atomicOr(&wmark[(v0 >> 5) + 1], (uint32_t)(m >> 32));
  }
}

// The same row-wise marks with returning atomics.  Claims: per stencil row, the
// bits of words w and w + 1 that this call set (claims of other threads
// excluded).
struct Claims {
  uint32_t w[7], lo[7], hi[7];
  __device__ __forceinline__ uint32_t count() const {
    uint32_t n = 0;
#pragma unroll
    for (int k = 0; k < 7; k++) n += __popc(lo[k]) + __popc(hi[k]);
    return n;
  }
  // the claimed ids, in row order, to list[pos ...]
  __device__ __forceinline__ void emit(uint32_t *list, uint32_t pos) const {
#pragma unroll
    for (int k = 0; k < 7; k++) {
      for (uint32_t b = lo[k]; b; b &= b - 1) list[pos++] = w[k] * 32u + (uint32_t)(__ffs(b) - 1);
      for (uint32_t b = hi[k]; b; b &= b - 1)
        list[pos++] = (w[k] + 1) * 32u + (uint32_t)(__ffs(b) - 1);
    }
  }
};
// claim_issue issues every row's claims (c.lo / c.hi: the bits asked for; ol /
// oh: the atomics' results); claim_finish keeps the bits this call set.
__device__ __forceinline__ void claim_issue(uint32_t t, int64_t x, int64_t y, int64_t z,
                                            const Dom &D, uint32_t *wmark, Claims &c,
                                            uint32_t (&ol)[7], uint32_t (&oh)[7]) {
  const uint32_t vm = box_mask(x, y, z, D.lo, D.hi);
  constexpr int kRowFirst[7] = {0, 2, 4, 6, 9, 11, 13};
  constexpr int kRowLen[7] = {2, 2, 2, 3, 2, 2, 2};
  // every claim is issued before any result is looked at, so the atomics overlap
#pragma unroll
  for (int k = 0; k < 7; k++) {
    const uint32_t rb = (vm >> kRowFirst[k]) & ((1u << kRowLen[k]) - 1u);
    const uint32_t v0 = D.gbase + (uint32_t)((int64_t)t + rank_off(kRowFirst[k], D.d));
    const uint64_t m = (uint64_t)rb << (v0 & 31u);
    c.w[k] = v0 >> 5;
    c.lo[k] = (uint32_t)m;
    c.hi[k] = (uint32_t)(m >> 32);
    ol[k] = c.lo[k] ? atomicOr(&wmark[c.w[k]], c.lo[k]) : 0u;
    oh[k] = c.hi[k] ? atomicOr(&wmark[c.w[k] + 1], c.hi[k]) : 0u;
  }
}
__device__ __forceinline__ void claim_finish(Claims &c, const uint32_t (&ol)[7],
                                             const uint32_t (&oh)[7]) {
#pragma unroll
  for (int k = 0; k < 7; k++) {
    c.lo[k] &= ~ol[k];
    c.hi[k] &= ~oh[k];
  }
}

// Detection at one listed voxel (global loads); its targets are marked in tmark
// without looking at the old bits (the list is compacted afterwards).  In a
// padded block a neighbour holding the sentinel is missing (pMSz.cu:709); a
// domain that must not edit counts only.
template <typename T>
__device__ __forceinline__ void detect_mark(const Dom &D, uint32_t *tmark, uint32_t local,
                                            bool padded, bool emit, uint32_t (&lc)[3]) {
  const T *g = (const T *)D.g;
  int64_t x, y, z;
  decompose(local, D.d, x, y, z);
  uint32_t vm = valid_mask(x, y, z, D.d);
  T cg[kCand];
#pragma unroll
  for (int r = 0; r < kCand; r++)
    cg[r] = ((vm >> r) & 1u) ? g[(int64_t)local + rank_off(r, D.d)] : (T)0;
  if (padded) {
#pragma unroll
    for (int r = 0; r < kCand; r++)
      if (cg[r] == sentinel<T>()) vm &= ~(1u << r);
  }
  int da, dd;
  if (vm == 0x7fffu) scan_full(cg, da, dd);
  else scan_masked(cg, vm, da, dd);
  const uint8_t oc = D.code[local];
  int tr[4];
  const int nt = detect_targets(oc & 15, oc >> 4, da, dd, tr, lc);
  if (!emit) return;
  for (int q = 0; q < nt; q++) {
    const uint32_t v = D.gbase + (uint32_t)((int64_t)local + rank_off(tr[q], D.d));
    atomicOr(&tmark[v >> 5], 1u << (v & 31u));
  }
}

// List of the set bits of a bitmap (ids), each nonzero word cleared as it is
// read; runs when gate->mode == need.  Only the words that hold ids are read:
// the domains' word ranges, back to back (dense index u; domain k covers
// [wpre_k, wpre_k + ceil(n_k / 32)) and starts at word gbase_k / 32).  A thread
// takes kBitsPT words per step, interleaved across the block so loads coalesce;
// a block makes one list reservation per step.
constexpr int kBitsThreads = 256;
constexpr int kBitsPT = 16;
__global__ __launch_bounds__(kBitsThreads) void k_compact_bits(
    uint32_t *__restrict__ bits, const Dom *__restrict__ doms, uint32_t ndom, uint32_t wdense,
    uint32_t *__restrict__ list, uint32_t *__restrict__ count, const Counters *__restrict__ gate,
    uint32_t need) {
  if (gate->mode != need) return;  // uniform
  __shared__ uint32_t s_wp[kCompactDoms], s_wb[kCompactDoms];
  if (ndom > 1) {
    for (uint32_t k = threadIdx.x; k < ndom; k += kBitsThreads) {
      s_wp[k] = doms[k].wpre;
      s_wb[k] = doms[k].gbase >> 5;
    }
    __syncthreads();
  }
  auto word_of = [&](uint32_t u) -> uint32_t {  // dense index -> bitmap word
    if (ndom <= 1) return u;
    uint32_t lo = 0, hi = ndom;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (s_wp[mid] <= u) lo = mid;
      else hi = mid;
    }
    return s_wb[lo] + (u - s_wp[lo]);
  };
  const uint32_t per_step = kBitsThreads * kBitsPT;
  for (uint32_t u0 = blockIdx.x * per_step; u0 < wdense; u0 += gridDim.x * per_step) {
    uint32_t w[kBitsPT];
    uint32_t n = 0;
#pragma unroll
    for (int v = 0; v < kBitsPT; v++) {
      const uint32_t u = u0 + (uint32_t)v * kBitsThreads + threadIdx.x;
      w[v] = u < wdense ? bits[word_of(u)] : 0u;
      n += __popc(w[v]);
    }
    uint32_t pos = block_reserve<kBitsThreads>(n, count);
#pragma unroll
    for (int v = 0; v < kBitsPT; v++) {
      uint32_t m = w[v];
      if (!m) continue;
      const uint32_t wi = word_of(u0 + (uint32_t)v * kBitsThreads + threadIdx.x);
      bits[wi] = 0u;
      while (m) {
        const uint32_t b = (uint32_t)(__ffs(m) - 1);
        m &= m - 1;
        list[pos++] = wi * 32u + b;
      }
    }
  }
}

// Detection at one listed voxel (global loads); returns the claimed targets.  In a
// padded block a neighbour holding the sentinel is missing (pMSz.cu:709).
template <typename T>
__device__ __forceinline__ int detect_one(const Dom &D, uint32_t *tmark, uint32_t local,
                                          bool padded, bool emit, uint32_t (&lc)[3],
                                          uint32_t (&out)[4]) {
  const T *g = (const T *)D.g;
  int64_t x, y, z;
  decompose(local, D.d, x, y, z);
  uint32_t vm = valid_mask(x, y, z, D.d);
  T cg[kCand];
#pragma unroll
  for (int r = 0; r < kCand; r++)
    cg[r] = ((vm >> r) & 1u) ? g[(int64_t)local + rank_off(r, D.d)] : (T)0;
  if (padded) {
#pragma unroll
    for (int r = 0; r < kCand; r++)
      if (cg[r] == sentinel<T>()) vm &= ~(1u << r);
  }
  int da, dd;
  if (vm == 0x7fffu) scan_full(cg, da, dd);
  else scan_masked(cg, vm, da, dd);
  const uint8_t oc = D.code[local];
  int tr[4];
  const int nt = detect_targets(oc & 15, oc >> 4, da, dd, tr, lc);
  if (!emit) return 0;
  uint32_t tv[4], old[4];
  for (int q = 0; q < nt; q++) {
    tv[q] = D.gbase + (uint32_t)((int64_t)local + rank_off(tr[q], D.d));
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

// Apply of one target id (shared by the apply and tail kernels): the edit, the
// mark clear, the ghost-sync flag and the claims of the neighbourhood -- queued
// into u (queue: returns their count), marked without results (mark), or as row
// claims into *cl.  One field: the claims go out between the cell's loads and the
// edit's stores (edit_load).  Several domains: after the stores, because holding
// the claims' results across the edit costs the batch kernels more occupancy
// than the overlap gains (measured).
template <typename T, bool kMulti>
__device__ __forceinline__ int apply_target(uint32_t t, const Dom &d0, const Dom *doms,
                                            uint32_t shift, DomState *ds, uint32_t *tmark,
                                            uint32_t *wmark, bool queue, bool mark, Claims *cl,
                                            uint32_t &dk, uint32_t &touched, uint32_t &clamped,
                                            uint32_t (&u)[kCand]) {
  uint32_t local;
  // the descriptor by value: its fields are loaded once, before the edit's stores
  // (a reference into the domain table would be reloaded after them, and a load
  // issued after a store waits for it)
  const Dom D = dom_of<kMulti>(t, d0, doms, shift, dk, local);
  const CellLoad<T> e = edit_load<T>((const T *)D.g, (const T *)D.f, D.dc, local);
  int64_t x, y, z;
  uint32_t old[kCand], ol[7], oh[7];
  auto claims = [&] {
    decompose(local, D.d, x, y, z);
    if (mark) mark_neighbourhood(local, x, y, z, D, wmark);
    if (queue) queue_issue(local, x, y, z, D, wmark, old);
    if (cl) claim_issue(local, x, y, z, D, wmark, *cl, ol, oh);
  };
  if (!kMulti) claims();
  edit_store<T>(e, (T *)D.g, D.dc, local, D.ep, touched, clamped);
  // every set bit of tmark is a target of this sweep and is applied in this phase,
  // so the word can be cleared with a plain store (tmark is null when the targets
  // come from a full-volume detection, which marks none)
  if (tmark) tmark[t >> 5] = 0u;
  if (kMulti) claims();
  if (kMulti && D.padded) {
    // send_flag: an edit within two layers of the padded boundary (pMSz.cu:642)
    if (x <= 1 || x >= D.d.nx - 2 || y <= 1 || y >= D.d.ny - 2 || z <= 1 || z >= D.d.nz - 2)
      if (!(ld_agent(&ds[dk].flags) & kDomSend)) atomicOr(&ds[dk].flags, kDomSend);
  }
  if (cl) claim_finish(*cl, ol, oh);
  return queue ? queue_collect(local, D, old, u) : 0;
}

// ---------------------------------------------------------------------------
// Edits of the targets of a full-volume detection, streamed from the target
// byte map (one domain).
// * A lane takes 4 consecutive cells (a lane unit: one tag word, one delta-
//   counter word, 4 values of f and g), four neighbouring lanes a 16-cell group,
//   so every load of a wave is contiguous.
// * A thread takes 8 units per step; their tag words are loaded a step ahead.
// * Their data come in rounds of two units.  A group holding any tagged cell is
//   loaded and stored whole by its four lanes, so its lines are written whole;
//   other groups cost their tag word only, and a round in which no group of the
//   wave has a target costs a few scalar instructions.
// * Within a round the tagged cells are queued in LDS (values, counters and the
//   queue of cell slots) and edited one per lane, so the edit's instructions run
//   once per 64 targets rather than once per cell slot of every lane.
// * In the run's first sweep (zero_dc) every delta counter is zero, cleared by
//   the labeling, so the counters are stored but not loaded.
// Counts the targets (for the statistics) into cj's kCntStreamed slot.
constexpr int kStreamRC = 8;  // cells per lane and round (two lane units)
template <typename T>
struct StreamLds {
  T f[kApplyThreads / 64][64 * kStreamRC];  // [wave][lane * kStreamRC + slot]
  T g[kApplyThreads / 64][64 * kStreamRC];
  uint8_t c[kApplyThreads / 64][64 * kStreamRC];
  uint16_t q[kApplyThreads / 64][64 * kStreamRC];  // queued cell slots
  uint32_t red[3][kApplyThreads / 64];
};

__device__ __forceinline__ uint32_t quad_or(uint32_t v) {
  // OR over the 4 lanes of a quad (DPP quad_perm [1,0,3,2], then [2,3,0,1])
  v |= (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false);
  v |= (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false);
  return v;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t b) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

template <typename T>
__device__ __forceinline__ void apply_stream(const Dom &d0, const uint8_t *__restrict__ tbyte,
                                             uint32_t ngroups, uint32_t tag, bool live,
                                             bool zero_dc, Counters &cj, RunState *run,
                                             StreamLds<T> &sl) {
  constexpr int kVec = 4 * (int)sizeof(T) / 16;  // 16-byte vectors per 4 cells of T
  constexpr int kTU = 8;                          // lane units per thread and step
  // clang vector types (first-class values, unlike uint4), so a round's data
  // stay in registers
  using v4u = uint32_t __attribute__((ext_vector_type(4)));
  T *g = (T *)d0.g;
  const T *f = (const T *)d0.f;
  uint8_t *dc = d0.dc;
  const uint64_t n = (uint64_t)d0.d.nx * (uint64_t)d0.d.ny * (uint64_t)d0.d.nz;
  const uint32_t nunits = ngroups * 4u;               // tag words of the map
  const uint32_t nfull = (uint32_t)(n / 16u) * 4u;    // units of groups wholly in the field
  const uint32_t rep = tag * 0x01010101u;
  const uint32_t *t32 = reinterpret_cast<const uint32_t *>(tbyte);
  uint32_t *dc32 = reinterpret_cast<uint32_t *>(dc);
  const v4u *f16 = reinterpret_cast<const v4u *>(f);
  v4u *g16 = reinterpret_cast<v4u *>(g);
  const uint32_t lane = lane_id(), w = threadIdx.x / 64;
  T *sf = sl.f[w];
  T *sg = sl.g[w];
  uint8_t *sc = sl.c[w];
  uint16_t *sq = sl.q[w];
  v4u *sf16 = reinterpret_cast<v4u *>(sf + lane * kStreamRC);
  v4u *sg16 = reinterpret_cast<v4u *>(sg + lane * kStreamRC);
  uint2 *sc8 = reinterpret_cast<uint2 *>(sc + lane * kStreamRC);
  uint32_t touched = 0, clamped = 0, count = 0;
  const uint32_t per_block = (uint32_t)kTU * kApplyThreads;
  const uint32_t stride = gridDim.x * per_block;
  uint32_t base = blockIdx.x * per_block;  // uniform in the block
  uint32_t tw[kTU];
#pragma unroll
  for (int q = 0; q < kTU; q++) {
    const uint32_t u = base + (uint32_t)q * kApplyThreads + threadIdx.x;
    tw[q] = u < nunits ? t32[u] : 0u;
  }
  for (; base < nunits; base += stride) {
    // the step's tagged cells, 4 bits per unit, and the same ORed over each quad
    uint32_t mp = 0;
#pragma unroll
    for (int q = 0; q < kTU; q++) mp |= byte_bits(tag_bytes(tw[q], rep)) << (4 * q);
#pragma unroll
    for (int q = 0; q < kTU; q++) {  // the next step's tag words
      const uint32_t u = base + stride + (uint32_t)q * kApplyThreads + threadIdx.x;
      tw[q] = u < nunits ? t32[u] : 0u;
    }
    const uint32_t gp = quad_or(mp);
    for (int r = 0; r < kTU; r += 2) {
      const uint32_t gr = (gp >> (4 * r)) & 0xffu;  // this round's two units, by group
      if (__ballot(gr != 0) == 0) continue;
      const uint32_t u0 = base + (uint32_t)r * kApplyThreads + threadIdx.x;
      const uint32_t u1 = u0 + kApplyThreads;
      const bool full0 = u0 < nfull, full1 = u1 < nfull;
      const bool ld0 = (gr & 15u) && full0, ld1 = (gr >> 4) && full1;
      uint32_t mr = (mp >> (4 * r)) & 0xffu;  // this lane's tagged cells
      // the field's last, partial group: cell by cell (tags lie on cells only)
      const uint32_t mpart = (full0 ? 0u : mr & 15u) | (full1 ? 0u : mr & 0xf0u);
      mr &= ~mpart;
      count += __popc(mr | mpart);
      for (uint32_t mm = mpart; mm; mm &= mm - 1) {
        const uint32_t k = (uint32_t)(__ffs(mm) - 1);
        apply_one<T>(g, f, dc, (k < 4 ? u0 : u1) * 4u + (k & 3u), d0.ep, touched, clamped);
      }
      v4u fa[kVec], fb[kVec], ga[kVec], gb[kVec];
      uint32_t ca = 0, cb = 0;
#pragma unroll
      for (int k = 0; k < kVec; k++) {
        fa[k] = fb[k] = ga[k] = gb[k] = v4u{0u, 0u, 0u, 0u};
        if (ld0) ga[k] = g16[(size_t)u0 * kVec + k];
        if (ld1) gb[k] = g16[(size_t)u1 * kVec + k];
      }
      // (the run's first sweep: the labeling has just cleared every counter)
      if (ld0 && !zero_dc) ca = dc32[u0];
      if (ld1 && !zero_dc) cb = dc32[u1];
      // f only for this lane's own tagged cells that take the exact edit (before
      // the first edits of a run no cell has a budget, so the tags decide alone)
      const bool fna = (mr & 15u) && (!live || exact_any(ca, mr & 15u));
      const bool fnb = (mr >> 4) && (!live || exact_any(cb, mr >> 4));
#pragma unroll
      for (int k = 0; k < kVec; k++) {
        if (fna) fa[k] = f16[(size_t)u0 * kVec + k];
        if (fnb) fb[k] = f16[(size_t)u1 * kVec + k];
      }
      // queue the tagged cells: slot k of lane L is entry L * 8 + k
      if (mr) {
#pragma unroll
        for (int k = 0; k < kVec; k++) {
          sf16[k] = fa[k];
          sf16[kVec + k] = fb[k];
          sg16[k] = ga[k];
          sg16[kVec + k] = gb[k];
        }
        *sc8 = make_uint2(ca, cb);
      }
      uint32_t nq = 0;
#pragma unroll
      for (int k = 0; k < kStreamRC; k++) {
        const bool on = (mr >> k) & 1u;
        const uint64_t bk = __ballot(on);
        if (on) sq[nq + lanes_below(bk)] = (uint16_t)(lane * kStreamRC + (uint32_t)k);
        nq += (uint32_t)__popcll(bk);
      }
      wave_lds_sync();
      for (uint32_t e = lane; e < nq; e += 64) {
        const uint32_t idx = sq[e];
        T gv = sg[idx];
        sc[idx] =
            edit_cell<T, true>(gv, sc[idx], [&] { return sf[idx]; }, d0.ep, touched, clamped);
        sg[idx] = gv;
      }
      wave_lds_sync();
      if (mr) {
#pragma unroll
        for (int k = 0; k < kVec; k++) {
          ga[k] = sg16[k];
          gb[k] = sg16[kVec + k];
        }
        const uint2 cc = *sc8;
        ca = cc.x;
        cb = cc.y;
      }
      if (ld0) {
#pragma unroll
        for (int k = 0; k < kVec; k++) g16[(size_t)u0 * kVec + k] = ga[k];
        dc32[u0] = ca;
      }
      if (ld1) {
#pragma unroll
        for (int k = 0; k < kVec; k++) g16[(size_t)u1 * kVec + k] = gb[k];
        dc32[u1] = cb;
      }
      wave_lds_sync();  // the slots and the queue are rewritten by the next round
    }
  }
  // one set of atomics per block: the target count goes to the counters' fourth
  // stripe slot (same-address atomics serialise; one per wave cost ~80 us)
  const uint32_t ts = wave_sum(touched), cs = wave_sum(clamped), ns = wave_sum(count);
  if (lane == 0) {
    sl.red[0][w] = ts;
    sl.red[1][w] = cs;
    sl.red[2][w] = ns;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t a = 0, b = 0, c = 0;
#pragma unroll
    for (int k = 0; k < kApplyThreads / 64; k++) {
      a += sl.red[0][k];
      b += sl.red[1][k];
      c += sl.red[2][k];
    }
    acc_add(run, kAccTouched, a);
    acc_add(run, kAccClamped, b);
    acc_add(run, kAccTargets, c);
    cnt_add(cj, kCntStreamed, c);
  }
}

// LDS of the apply kernel: the list staging or (one domain) the streamed
// apply's queues.
template <typename T>
union ApplyLds {
  StageLds<kApplyThreads, kApplyThreads * kCand> st;
  StreamLds<T> sl;
};
struct ApplyLdsMulti {
  StageLds<kApplyThreads, kApplyThreads * kCand> st;
};

// ---------------------------------------------------------------------------
// Apply of sweep j (grid-stride over its targets), queueing the next work list
// when detection j+1 is sparse.
template <typename T, bool kMulti>
__global__ __launch_bounds__(kApplyThreads) void k_apply(
    Dom d0, const Dom *__restrict__ doms, uint32_t shift, DomState *__restrict__ ds,
    uint32_t *__restrict__ tmark, const uint32_t *__restrict__ tlist,
    uint32_t *__restrict__ wmark, uint32_t *__restrict__ wlist, Counters *__restrict__ ring,
    RunState *run, uint32_t j, Sched s, const uint8_t *__restrict__ tbyte, uint32_t ngroups) {
  Counters &prev = ring[(j - 1) % kCtrRing];
  Counters &cj = ring[j % kCtrRing];
  const uint32_t pm = prev.mode;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (pm == kModeTail || pm == kModeDone || pm == kModeUnset) {
    if (lead) cj.mode = pm == kModeUnset ? kModeDone : pm;
    return;
  }
  // one domain, many targets of a full-volume detection: no list was made; the
  // byte map is streamed (the same decision as the compaction's)
  const bool stream = !kMulti && pm == kModeDense && s.stream_min && cnt_sum(cj) >= s.stream_min;
  uint32_t done;
  const uint32_t mode = kMulti   ? decide_multi(cj, s, s.tail_ok != 0, &done)
                        : stream ? decide_stream(cj, j, s, &done)
                                 : decide<false>(cj, j, s, true, &done);
  if (lead) {
    cj.mode = mode;
    run->last_mode = mode;
    if (mode == kModeTail) run->tail_j = j;
    if (done) {
      run->done = done;
      run->last_j = j;
    }
  }
  if (mode == kModeTail) return;
  if (!kMulti && mode == kModeDone && (done == 1 || s.mode == PMSZ_SYNC)) return;
  __shared__ std::conditional_t<kMulti, ApplyLdsMulti, ApplyLds<T>> lds;
  if constexpr (!kMulti) {
    if (stream) {
      if (lead && mode != kModeDone) zero_counters(&ring[(j + 1) % kCtrRing]);
      // budgets exist once a sweep has edited (j > 1)
      apply_stream<T>(d0, tbyte, ngroups, detect_tag(j + s.tag0), budgets_on(d0.ep) && j > 1,
                      j == 1, cj, run, lds.sl);
      return;
    }
  }
  const uint32_t nt = cj.ntargets;
  if (kMulti && mode == kModeDone && nt == 0) return;
  if (lead && mode != kModeDone) zero_counters(&ring[(j + 1) % kCtrRing]);
  const bool queue = mode == kModeSparse;
  uint32_t touched = 0, clamped = 0;
  const uint32_t stride = gridDim.x * kApplyThreads;
  // claimed lists (s.append): the targets of a work-list detection are claimed in
  // tmark; every set bit is a target of this sweep, so their words are cleared with
  // plain stores
  const bool clear = (!kMulti || s.append) && pm == kModeSparse;
  auto &st = lds.st;
  if (!kMulti && queue) {
    // one domain, work-list detection next: the edit, then the neighbourhood's
    // claims in wmark, listed at once (staged per block)
    stage_init(st);
    for (uint32_t base = blockIdx.x * kApplyThreads; base < nt; base += stride) {
      const uint32_t i = base + threadIdx.x;
      Claims cl;
      uint32_t n = 0;
      if (i < nt) {
        const uint32_t t = tlist[i];
        // the claims go out between the cell's loads and the edit's stores
        const CellLoad<T> e = edit_load<T>((const T *)d0.g, (const T *)d0.f, d0.dc, t);
        int64_t x, y, z;
        decompose(t, d0.d, x, y, z);
        uint32_t ol[7], oh[7];
        claim_issue(t, x, y, z, d0, wmark, cl, ol, oh);
        edit_store<T>(e, (T *)d0.g, d0.dc, t, d0.ep, touched, clamped);
        if (clear) tmark[t >> 5] = 0u;
        claim_finish(cl, ol, oh);
        n = cl.count();
      }
      const uint32_t pos = stage_reserve(st, n, wlist, &cj.nwork);
      if (n) cl.emit(st.buf, pos);
    }
    stage_flush(st, wlist, &cj.nwork);
  } else if (kMulti && queue && s.append) {
    // several domains without ghost syncs, work-list detection next: the edit,
    // then the neighbourhood's claims, listed at once (staged per block)
    stage_init(st);
    for (uint32_t base = blockIdx.x * kApplyThreads; base < nt; base += stride) {
      const uint32_t i = base + threadIdx.x;
      uint32_t u[kCand];
      uint32_t dk = 0, n = 0;
      Claims cl;
      const bool has = i < nt;
      if (has) {
        apply_target<T, kMulti>(tlist[i], d0, doms, shift, ds, clear ? tmark : nullptr, wmark,
                                false, false, &cl, dk, touched, clamped, u);
        n = cl.count();
      }
      dom_add_targets(has, dk, ds);
      const uint32_t pos = stage_reserve(st, n, wlist, &cj.nwork);
      if (n) cl.emit(st.buf, pos);
    }
    stage_flush(st, wlist, &cj.nwork);
  } else if (kMulti) {
    // several domains: the edit (and the ghost-sync flag), the neighbourhood's
    // bits in wmark when detection j+1 runs over a work list
    for (uint32_t base = blockIdx.x * kApplyThreads; base < nt; base += stride) {
      const uint32_t i = base + threadIdx.x;
      uint32_t u[kCand];
      uint32_t dk = 0;
      const bool has = i < nt;
      if (has)
        apply_target<T, kMulti>(tlist[i], d0, doms, shift, ds, clear ? tmark : nullptr, wmark,
                                false, queue, nullptr, dk, touched, clamped, u);
      dom_add_targets(has, dk, ds);
    }
  } else {
    // edits only (the targets of a full-volume detection, no work list): four
    // targets per thread, all their loads issued before any edit is stored
    constexpr int kU = 4;
    T *g = (T *)d0.g;
    const T *f = (const T *)d0.f;
    uint8_t *dc = d0.dc;
    for (uint32_t base = blockIdx.x * kApplyThreads * kU; base < nt; base += stride * kU) {
      uint32_t t[kU];
      T fv[kU], gv[kU];
      uint8_t cv[kU];
#pragma unroll
      for (int q = 0; q < kU; q++) {
        const uint32_t i = base + (uint32_t)q * kApplyThreads + threadIdx.x;
        t[q] = i < nt ? tlist[i] : 0xffffffffu;
        if (clear && t[q] != 0xffffffffu) tmark[t[q] >> 5] = 0u;
      }
#pragma unroll
      for (int q = 0; q < kU; q++) {
        if (t[q] != 0xffffffffu) {
          gv[q] = g[t[q]];
          cv[q] = dc[t[q]];
        }
      }
      // f only where the edit is not within a step budget
      const bool bon = budgets_on(d0.ep);
#pragma unroll
      for (int q = 0; q < kU; q++) {
        fv[q] = (T)0;
        if (t[q] != 0xffffffffu && !(bon && (cv[q] & 15u) < (uint32_t)(cv[q] >> 4)))
          fv[q] = f[t[q]];
      }
#pragma unroll
      for (int q = 0; q < kU; q++) {
        if (t[q] != 0xffffffffu) {
          const uint8_t c =
              edit_cell<T, true>(gv[q], cv[q], [&] { return fv[q]; }, d0.ep, touched, clamped);
          g[t[q]] = gv[q];
          dc[t[q]] = c;
        }
      }
    }
  }
  if (!kMulti) {
    uint32_t ts = wave_sum(touched), cs = wave_sum(clamped);
    if (lane_id() == 0) {
      acc_add(run, kAccTouched, ts);
      acc_add(run, kAccClamped, cs);
    }
  }
  if (lead) acc_add(run, kAccTargets, nt);
}

// Work-list detection of one listed voxel id (shared by k_detect_list / k_tail).
template <typename T, bool kMulti>
__device__ __forceinline__ int detect_listed(uint32_t v, const Dom &d0, const Dom *doms,
                                             uint32_t shift, const DomState *ds,
                                             uint32_t *tmark, uint32_t &dk, bool &has,
                                             uint32_t (&vc)[3], uint32_t (&out)[4]) {
  uint32_t local;
  const Dom &D = dom_of<kMulti>(v, d0, doms, shift, dk, local);
  uint32_t fl = kDomActive | kDomEmit;
  if (kMulti) fl = ds[dk].flags;
  has = (fl & kDomActive) != 0;
  if (!has) return 0;
  return detect_one<T>(D, tmark, local, kMulti && D.padded, (fl & kDomEmit) != 0, vc, out);
}

// Sparse detection j+1 over the work list queued by apply j (grid-stride).
template <typename T, bool kMulti>
__global__ __launch_bounds__(kApplyThreads) void k_detect_list(
    Dom d0, const Dom *__restrict__ doms, uint32_t shift, DomState *__restrict__ ds,
    uint32_t *__restrict__ tmark, uint32_t *__restrict__ tlist, uint32_t *__restrict__ wmark,
    const uint32_t *__restrict__ wlist, Counters *__restrict__ ring, RunState *run, uint32_t j,
    uint32_t append) {
  const Counters &cj = ring[j % kCtrRing];
  if (cj.mode != kModeSparse) return;
  Counters &cn = ring[(j + 1) % kCtrRing];
  const uint32_t par = (j + 1) & 1u;
  const uint32_t nw = cj.nwork;
  uint32_t lc[3] = {0, 0, 0};
  const uint32_t stride = gridDim.x * kApplyThreads;
  __shared__ StageLds<kApplyThreads, kApplyThreads * 4> st;
  if (!kMulti || append) {  // one field always lists its claims
    // targets claimed in tmark and listed at once; every set bit of wmark is a
    // listed voxel of this sweep, so its words are cleared with plain stores
    stage_init(st);
    for (uint32_t base = blockIdx.x * kApplyThreads; base < nw; base += stride) {
      const uint32_t i = base + threadIdx.x;
      uint32_t out[4];
      uint32_t vc[3] = {0, 0, 0};
      uint32_t dk = 0;
      bool has = false;
      int n = 0;
      if (i < nw) {
        const uint32_t v = wlist[i];
        n = detect_listed<T, kMulti>(v, d0, doms, shift, ds, tmark, dk, has, vc, out);
        wmark[v >> 5] = 0u;  // after the detection's loads, which would wait for it
      }
      if (kMulti) dom_add3(has, dk, vc, ds, par);
#pragma unroll
      for (int q = 0; q < 3; q++) lc[q] += vc[q];
      const uint32_t pos = stage_reserve(st, (uint32_t)n, tlist, &cn.ntargets);
      for (int q = 0; q < n; q++) st.buf[pos + q] = out[q];
    }
    stage_flush(st, tlist, &cn.ntargets);
  } else {
    for (uint32_t base = blockIdx.x * kApplyThreads; base < nw; base += stride) {
      const uint32_t i = base + threadIdx.x;
      uint32_t vc[3] = {0, 0, 0};
      uint32_t dk = 0;
      bool has = false;
      if (i < nw) {
        uint32_t local;
        const Dom &D = dom_of<kMulti>(wlist[i], d0, doms, shift, dk, local);
        const uint32_t fl = ds[dk].flags;
        has = (fl & kDomActive) != 0;
        if (has) detect_mark<T>(D, tmark, local, D.padded != 0, (fl & kDomEmit) != 0, vc);
      }
      dom_add3(has, dk, vc, ds, par);
#pragma unroll
      for (int q = 0; q < 3; q++) lc[q] += vc[q];
    }
  }
  block_add3<kApplyThreads>(lc, cn);
  if (blockIdx.x == 0 && threadIdx.x == 0)
    acc_add(run, kAccSparse, nw);
}

// Grid barrier of the tail (all its blocks are resident: cooperative launch).
// Block b publishes epoch in its own flag (release); block 0 gathers every flag
// (acquire) and publishes the epoch in the generation word; the others wait for
// it.  Epochs only grow within a run (the state is cleared before each run), so
// no flag needs resetting.  bar[16 * b]: flag of block b; bar[16 * kTailBlocks]:
// the generation.  Every wait is bounded: after kBarrierTicks of the 100 MHz
// wall clock (2 s; a barrier of the tail takes microseconds) the block sets
// kErrBarrier in run->err and the barrier returns false, so a grid whose blocks
// are not all resident ends with an error instead of hanging.
constexpr long long kBarrierTicks = 200000000ll;
__device__ __forceinline__ bool bar_wait(const uint32_t *w, uint32_t epoch) {
  const long long t0 = wall_clock64();
  while (__hip_atomic_load(w, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
    if (wall_clock64() - t0 > kBarrierTicks) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}
__device__ __forceinline__ bool tail_barrier(uint32_t *bar, uint32_t epoch, RunState *run) {
  __shared__ uint32_t s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    s_ok = 1u;
    __hip_atomic_store(&bar[16 * blockIdx.x], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    bool ok = true;
    for (uint32_t b = threadIdx.x; b < gridDim.x && ok; b += blockDim.x)
      ok = bar_wait(&bar[16 * b], epoch);
    if (!ok) s_ok = 0u;
    __syncthreads();
    if (threadIdx.x == 0 && s_ok)
      __hip_atomic_store(&bar[16 * kTailBlocks], epoch, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
  } else if (threadIdx.x == 0) {
    if (!bar_wait(&bar[16 * kTailBlocks], epoch)) s_ok = 0u;
  }
  __syncthreads();
  const bool ok = s_ok != 0u;
  if (!ok && threadIdx.x == 0) atomicOr(&run->err, kErrBarrier);
  __syncthreads();  // s_ok is rewritten by the next barrier
  return ok;
}

// The tail: once a sweep's targets are few and detection j+1 would run over a
// work list, one cooperative kernel runs the remaining sweeps -- apply, grid barrier, work-list detection,
// grid barrier -- instead of several launches per sweep.  Every block decides
// alike from counters read after a barrier.  A grid of several blocks is
// launched only through hipLaunchCooperativeKernel, which guarantees that all
// blocks are resident; where the runtime refuses that, the host launches one
// block, which runs in solo mode (workgroup barriers only) from the start.
template <typename T, bool kMulti>
__global__ __launch_bounds__(kTailThreads) void k_tail(
    Dom d0, const Dom *__restrict__ doms, uint32_t shift, uint32_t ndom,
    DomState *__restrict__ ds, uint32_t *__restrict__ tmark, uint32_t *__restrict__ tlist,
    uint32_t *__restrict__ wmark, uint32_t *__restrict__ wlist, Counters *__restrict__ ring,
    RunState *run, Sched s, uint32_t *bar) {
  uint32_t epoch = ld_agent(&run->bar_epoch);
  __shared__ uint32_t s_act, s_cap;
  uint32_t j = ld_agent(&run->tail_j);
  if (j == 0 || ld_agent(&run->done)) return;  // uniform: written by earlier kernels
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  // once a sweep's targets are few, block 0 carries on alone: its phases are then
  // separated by workgroup barriers instead of grid barriers
  bool solo = gridDim.x == 1;
  uint32_t nthreads = gridDim.x * kTailThreads;
  uint32_t first = blockIdx.x * kTailThreads;
  // false: a grid barrier timed out (kErrBarrier is set); the block leaves
  auto barrier = [&]() -> bool {
    if (solo) {
      __syncthreads();
      return true;
    }
    return tail_barrier(bar, ++epoch, run);
  };
  unsigned long long tt = 0, sv = 0;
  uint32_t touched = 0, clamped = 0;
  __shared__ uint32_t s_mode, s_done, s_nt;
  while (true) {
    Counters &cj = ring[j % kCtrRing];
    // one thread per block reads the counters (after the grid barrier) and
    // decides; every block decides alike
    if (threadIdx.x == 0) {
      uint32_t dn;
      s_mode = kMulti ? decide_multi(cj, s, false, &dn) : decide<true>(cj, j, s, false, &dn);
      s_done = dn;
      s_nt = ld_agent(&cj.ntargets);
    }
    __syncthreads();
    const uint32_t mode = s_mode, done = s_done, nt = s_nt;
    __syncthreads();  // s_* are rewritten in the next sweep
    if (lead) {
      cj.mode = mode == kModeDone ? kModeDone : kModeTail;
      if (done) {
        run->done = done;
        run->last_j = j;
      }
    }
    if (!kMulti && (done == 1 || (done == 2 && s.mode == PMSZ_SYNC))) break;
    if (kMulti && mode == kModeDone && nt == 0) break;
    if (!solo && nt <= s.solo_targets) {  // uniform: every block read the same counters
      if (blockIdx.x != 0) break;
      solo = true;
      nthreads = kTailThreads;
      first = 0;
    }
    if (lead && mode != kModeDone) zero_counters(&ring[(j + 1) % kCtrRing]);
    // apply sweep j, queueing the next work list unless this was the last sweep
    const bool queue = mode != kModeDone;
    for (uint32_t base = first; base < nt; base += nthreads) {
      const uint32_t i = base + threadIdx.x;
      uint32_t u[kCand];
      int n = 0;
      uint32_t dk = 0;
      const bool has = i < nt;
      if (has)
        n = apply_target<T, kMulti>(tlist[i], d0, doms, shift, ds, tmark, wmark, queue, false,
                                    nullptr, dk, touched, clamped, u);
      if (kMulti) dom_add_targets(has, dk, ds);
      if (queue) {
        uint32_t pos = block_reserve<kTailThreads>((uint32_t)n, &cj.nwork);
        for (int q = 0; q < n; q++) wlist[pos + q] = u[q];
      }
    }
    if (lead) tt += nt;
    if (!barrier()) return;
    if (!queue) break;  // sweep cap: sweep j's edits landed, no detection j+1
    // detection j+1 over the work list
    Counters &cn = ring[(j + 1) % kCtrRing];
    const uint32_t par = (j + 1) & 1u;
    const uint32_t nw = ld_agent(&cj.nwork);
    uint32_t lc[3] = {0, 0, 0};
    for (uint32_t base = first; base < nw; base += nthreads) {
      const uint32_t i = base + threadIdx.x;
      uint32_t out[4];
      uint32_t vc[3] = {0, 0, 0};
      uint32_t dk = 0;
      bool has = false;
      int n = 0;
      if (i < nw) {
        const uint32_t v = wlist[i];
        n = detect_listed<T, kMulti>(v, d0, doms, shift, ds, tmark, dk, has, vc, out);
        wmark[v >> 5] = 0u;  // every set bit is a listed voxel of this sweep
      }
      if (kMulti) dom_add3(has, dk, vc, ds, par);
#pragma unroll
      for (int q = 0; q < 3; q++) lc[q] += vc[q];
      uint32_t pos = block_reserve<kTailThreads>((uint32_t)n, &cn.ntargets);
      for (int q = 0; q < n; q++) tlist[pos + q] = out[q];
    }
    block_add3<kTailThreads>(lc, cn);
    if (lead) sv += nw;
    if (!barrier()) return;
    if (kMulti) {
      if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
          s_act = 0;
          s_cap = 0;
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < ndom; k += kTailThreads) {
          uint32_t a, c;
          dom_update(ds[k], j + 1, kOpDetect, s, 1u, a, c);
          if (a) atomicAdd(&s_act, 1u);
          if (c) atomicAdd(&s_cap, 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          cn.nactive = s_act;
          cn.ncapped = s_cap;
        }
      }
      if (!barrier()) return;
    }
    j++;
  }
  if (!kMulti) {
    uint32_t ts = wave_sum(touched), cs = wave_sum(clamped);
    if (lane_id() == 0) {
      acc_add(run, kAccTouched, ts);
      acc_add(run, kAccClamped, cs);
    }
  }
  if (lead) {
    acc_add(run, kAccTargets, tt);
    acc_add(run, kAccSparse, sv);
    run->bar_epoch = epoch;  // the next tail launch of this run continues from here
  }
}

// Hand-over to the check detection after a ghost sync: detection m+1 runs over
// the work list that the sync queues into slot m.
__global__ void k_reset_phase(Counters *ring, RunState *run, uint32_t m) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Counters &cm = ring[m % kCtrRing];
  cm.mode = kModeSparse;
  cm.nwork = 0;
  zero_counters(&ring[(m + 1) % kCtrRing]);
  run->done = 0;
  run->tail_j = 0;
}

// Start of a run: clears both mark bitmaps (`words` 32-bit words each), the
// counter ring, the run state and the tail's barrier, marks detection 1 as a
// running dense sweep (slot 0's mode, read by the first apply), and clears the
// first tgroups 16-byte groups of the target byte map -- one launch instead of a
// memset per buffer.  Every pointer is a hipMalloc base (16-byte aligned).
__global__ void k_init_run(uint32_t *__restrict__ tmark, uint32_t *__restrict__ wmark,
                           uint64_t words, Counters *__restrict__ ring, RunState *__restrict__ run,
                           uint32_t *__restrict__ bar, uint4 *__restrict__ tb, uint64_t tgroups) {
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nt = (uint64_t)gridDim.x * blockDim.x;
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  const uint64_t w4 = words / 4;
  for (uint64_t i = t0; i < w4; i += nt) {
    reinterpret_cast<uint4 *>(tmark)[i] = z;
    reinterpret_cast<uint4 *>(wmark)[i] = z;
  }
  for (uint64_t i = 4 * w4 + t0; i < words; i += nt) tmark[i] = wmark[i] = 0u;
  for (uint64_t i = t0; i < tgroups; i += nt) tb[i] = z;
  constexpr uint64_t kRingWords = sizeof(Counters) * kCtrRing / 4;
  constexpr uint64_t kModeWord = offsetof(Counters, mode) / 4;
  static_assert(sizeof(Counters) % 4 == 0 && sizeof(RunState) % 4 == 0, "word-sized state");
  for (uint64_t i = t0; i < kRingWords; i += nt)
    reinterpret_cast<uint32_t *>(ring)[i] = i == kModeWord ? (uint32_t)kModeDense : 0u;
  for (uint64_t i = t0; i < sizeof(RunState) / 4; i += nt)
    reinterpret_cast<uint32_t *>(run)[i] = 0u;
  for (uint64_t i = t0; i < kBarBytes / 4; i += nt) bar[i] = 0u;
}

// The caller's copy of n delta counters: the low nibble of each (the counters
// without their step budgets).  src is a hipMalloc base; dst may have any alignment.
__global__ void k_dc_out(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src, uint64_t n) {
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nt = (uint64_t)gridDim.x * blockDim.x;
  if (((uintptr_t)dst & 15u) == 0) {
    const uint64_t n16 = n / 16;
    for (uint64_t i = t0; i < n16; i += nt) {
      uint4 v = reinterpret_cast<const uint4 *>(src)[i];
      v.x &= 0x0f0f0f0fu;
      v.y &= 0x0f0f0f0fu;
      v.z &= 0x0f0f0f0fu;
      v.w &= 0x0f0f0f0fu;
      reinterpret_cast<uint4 *>(dst)[i] = v;
    }
    for (uint64_t i = 16 * n16 + t0; i < n; i += nt) dst[i] = src[i] & 15u;
  } else {
    for (uint64_t i = t0; i < n; i += nt) dst[i] = src[i] & 15u;
  }
}

// Zero fill of n16 16-byte groups (the target byte map when the tags start over;
// a kernel, not a memset node, so that a captured chain replays it reliably).
__global__ void k_fill16(uint4 *p, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------------------------
// Subdomains: padded blocks, the delta-counter ghost sync, the gather.

// Index of the stencil direction with components (ax, ay, az), -1 if none.
__device__ __forceinline__ int dir_index(int ax, int ay, int az) {
  int r = -1;
#pragma unroll
  for (int d = 0; d < 14; d++)
    if (dir_comp(d, 0) == ax && dir_comp(d, 1) == ay && dir_comp(d, 2) == az) r = d;
  return r;
}

__device__ __forceinline__ bool sub_exists(const int32_t *cc, int d, int sx, int sy, int sz) {
  const int a = cc[0] + dir_comp(d, 0), b = cc[1] + dir_comp(d, 1), c = cc[2] + dir_comp(d, 2);
  return a >= 0 && a < sx && b >= 0 && b < sy && c >= 0 && c < sz;
}
__device__ __forceinline__ int sub_of(const int32_t *cc, int d, int sx, int sy) {
  return (cc[0] + dir_comp(d, 0)) + sx * ((cc[1] + dir_comp(d, 1)) + sy * (cc[2] + dir_comp(d, 2)));
}

// The padded blocks of every subdomain (blockIdx.y): the core of the global
// field (read_subblock, pMSz.cu:274-313) and the initial ghost values
// (exchange_ghost_layers, pMSz.cu:320-608).  A halo cell in an exchanged
// stencil direction whose neighbour subdomain exists takes the neighbour's core
// value; every other halo cell holds the sentinel.  The reference exchanges the
// faces and the four edges d = 6..9 (pMSz.cu:394-488); the full policy all 14.
template <typename T>
__global__ __launch_bounds__(kSyncThreads) void k_scatter(const Dom *__restrict__ doms,
                                                         const T *__restrict__ F,
                                                         const T *__restrict__ G, int64_t gnx,
                                                         int64_t gny, int sx, int sy, int sz,
                                                         int policy, RunState *run) {
  const Dom &D = doms[blockIdx.y];
  T *fp = (T *)const_cast<void *>(D.f);
  T *gp = (T *)D.g;
  uint32_t err = 0;
  const uint32_t n = (uint32_t)D.d.n;
  for (uint32_t i = blockIdx.x * kSyncThreads + threadIdx.x; i < n;
       i += gridDim.x * kSyncThreads) {
    int64_t x, y, z;
    decompose(i, D.d, x, y, z);
    const int ax = x == 0 ? -1 : (x == D.d.nx - 1 ? 1 : 0);
    const int ay = y == 0 ? -1 : (y == D.d.ny - 1 ? 1 : 0);
    const int az = z == 0 ? -1 : (z == D.d.nz - 1 ? 1 : 0);
    const bool core = ax == 0 && ay == 0 && az == 0;
    bool take = core;
    if (!core) {
      const int d = dir_index(ax, ay, az);
      take = d >= 0 && (policy == PMSZ_GHOST_FULL || d < 10) && sub_exists(D.cc, d, sx, sy, sz);
    }
    T fv = sentinel<T>(), gv = sentinel<T>();
    if (take) {
      const int64_t gx = x - 1 + D.off[0], gy = y - 1 + D.off[1], gz = z - 1 + D.off[2];
      const int64_t gi = gx + gnx * (gy + gny * gz);
      fv = F[gi];
      gv = G[gi];
      if (core && (!(fv == fv) || fv == sentinel<T>() || !(gv == gv) || gv == sentinel<T>()))
        err = 1;
    }
    fp[i] = fv;
    gp[i] = gv;
    D.dc[i] = 0;
  }
  if (__ballot(err != 0) && lane_id() == 0) atomicOr(&run->err, 1u);
}

// Snapshot of the delta counters of every shell cell (the pack of
// exchange_ghost_layers_cuda_aware, pMSz.cu:930-1001): a subdomain with no
// boundary edit since the last sync sends nothing; cells outside the global
// field or holding the sentinel send 0.  Runs before any unpack.
template <typename T>
__global__ __launch_bounds__(kSyncThreads) void k_pack(const Dom *__restrict__ doms,
                                                      const DomState *__restrict__ ds,
                                                      int64_t gnx, int64_t gny, int64_t gnz,
                                                      const Counters *__restrict__ ring,
                                                      uint32_t j, uint32_t gate) {
  if (gate) {
    const uint32_t m = ring[j % kCtrRing].mode;
    if (m != kModeDense && m != kModeSparse) return;
  }
  const uint32_t k = blockIdx.y;
  const Dom &D = doms[k];
  const bool send = (ds[k].flags & kDomSend) != 0;
  const T *g = (const T *)D.g;
  const uint32_t P0 = (uint32_t)D.d.nx, P1 = (uint32_t)D.d.ny, P2 = (uint32_t)D.d.nz;
  for (uint32_t i = blockIdx.x * kSyncThreads + threadIdx.x; i < D.shell_n;
       i += gridDim.x * kSyncThreads) {
    uint32_t x, y, z;
    shell_cell(i, P0, P1, P2, x, y, z);
    const uint32_t li = x + P0 * (y + P1 * z);
    uint8_t v = 0;
    if (send) {
      const int64_t gx = (int64_t)x - 1 + D.off[0], gy = (int64_t)y - 1 + D.off[1],
                    gz = (int64_t)z - 1 + D.off[2];
      const bool in = gx >= 0 && gx < gnx && gy >= 0 && gy < gny && gz >= 0 && gz < gnz;
      if (in && g[li] != sentinel<T>()) v = D.dc[li];
    }
    D.snap[li] = v;
  }
}

// Receive side of the delta-counter sync (unpack, pMSz.cu:1015-1103): each shell
// cell of subdomain blockIdx.y merges, in direction order d = 0..ndir-1, the
// snapshot value the neighbour in direction d sent for the same global cell --
// the cell's layer (0 or 1) in the receive region toward d is the same along
// every nonzero component of d.  A larger value r replaces the counter and
// lowers g by delta * (r - cur), or clamps when r reaches q.  Cells whose value
// changed queue their classified neighbourhood for the next detection.
template <typename T>
__global__ __launch_bounds__(kSyncThreads) void k_unpack(const Dom *__restrict__ doms,
                                                        DomState *__restrict__ ds, int sx,
                                                        int sy, int sz, int ndir,
                                                        uint32_t *__restrict__ wmark,
                                                        uint32_t *__restrict__ wlist,
                                                        Counters *__restrict__ ring, uint32_t j,
                                                        uint32_t gate) {
  Counters &cj = ring[j % kCtrRing];
  const uint32_t mode = cj.mode;
  if (gate && mode != kModeDense && mode != kModeSparse) return;
  const bool queue = mode == kModeSparse;
  const uint32_t k = blockIdx.y;
  const Dom &D = doms[k];
  T *g = (T *)D.g;
  const T *f = (const T *)D.f;
  const T delta = (T)D.ep.delta_d;
  const int L0 = (int)D.d.nx - 2, L1 = (int)D.d.ny - 2, L2 = (int)D.d.nz - 2;
  const uint32_t P0 = (uint32_t)D.d.nx, P1 = (uint32_t)D.d.ny, P2 = (uint32_t)D.d.nz;
  for (uint32_t base = blockIdx.x * kSyncThreads; base < D.shell_n;
       base += gridDim.x * kSyncThreads) {
    const uint32_t i = base + threadIdx.x;
    if (i < D.shell_n) {
      uint32_t x, y, z;
      shell_cell(i, P0, P1, P2, x, y, z);
      const uint32_t li = x + P0 * (y + P1 * z);
      T gv = g[li];
      if (gv != sentinel<T>()) {
        uint8_t cur = D.dc[li];
        bool changed = false;
        const int c[3] = {(int)x, (int)y, (int)z};
        const int L[3] = {L0, L1, L2};
#pragma unroll
        for (int d = 0; d < 14; d++) {
          if (d >= ndir) continue;
          bool in = true;
          int layer = -1;
#pragma unroll
          for (int a = 0; a < 3; a++) {
            const int dd = dir_comp(d, a);
            if (dd == 0) {
              in = in && c[a] >= 1 && c[a] <= L[a];
            } else {
              const int l = dd < 0 ? c[a] : L[a] + 1 - c[a];
              in = in && (l == 0 || l == 1) && (layer < 0 || layer == l);
              layer = l;
            }
          }
          if (!in || !sub_exists(D.cc, d, sx, sy, sz)) continue;
          const Dom &B = doms[sub_of(D.cc, d, sx, sy)];
          const int64_t bx = (int64_t)x + D.off[0] - B.off[0];
          const int64_t by = (int64_t)y + D.off[1] - B.off[1];
          const int64_t bz = (int64_t)z + D.off[2] - B.off[2];
          const uint8_t r = B.snap[bx + B.d.nx * (by + B.d.ny * bz)];
          if (r == 0 || cur >= r) continue;
          if (r < D.ep.q) {
            const T step = delta * (T)(r - cur);
            gv = gv - step;
          } else {
            gv = clamp_value<T>(f[li], D.ep);
          }
          cur = r;
          changed = true;
        }
        if (changed) {
          g[li] = gv;
          D.dc[li] = cur;
          // the next work list is compacted from wmark
          if (queue) mark_neighbourhood(li, x, y, z, D, wmark);
        }
      }
    }
  }
  // send_flag = 0 after the exchange (pMSz.cu:2000-2007); the pack has read it
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAnd(&ds[k].flags, ~kDomSend);
}

// Per-domain end-of-run statistics over the classified box (touched: dc > 0,
// clamped: dc == q) and, with G non-null, the gather of the cores of padded
// blocks back into the global field and counters (the D2H copy, pMSz.cu:2971).
template <typename T>
__global__ __launch_bounds__(kSyncThreads) void k_finish(const Dom *__restrict__ doms,
                                                        DomState *__restrict__ ds,
                                                        T *__restrict__ G,
                                                        uint8_t *__restrict__ DC, int64_t gnx,
                                                        int64_t gny) {
  const uint32_t k = blockIdx.y;
  const Dom &D = doms[k];
  const uint32_t bx = (uint32_t)(D.hi[0] - D.lo[0]), by = (uint32_t)(D.hi[1] - D.lo[1]),
                 bz = (uint32_t)(D.hi[2] - D.lo[2]);
  const uint32_t rows = by * bz;
  const uint8_t q = (uint8_t)D.ep.q;
  uint32_t touched = 0, clamped = 0;
  if (!G && !D.padded) {
    // a field of a batch: its delta counters are one run of bytes, counted 16 at a
    // time (an unaligned head and the tail byte by byte)
    const uint8_t *dcp = D.dc;
    const uint64_t n = (uint64_t)D.d.nx * (uint64_t)D.d.ny * (uint64_t)D.d.nz;
    const uint64_t head = std::min<uint64_t>(n, (16u - ((uintptr_t)dcp & 15u)) & 15u);
    const uint64_t nv = (n - head) / 16u, tail0 = head + nv * 16u;
    const uint64_t tid = (uint64_t)blockIdx.x * kSyncThreads + threadIdx.x;
    const uint64_t nthr = (uint64_t)gridDim.x * kSyncThreads;
    const uint32_t qrep = (uint32_t)q * 0x01010101u;
    const uint4 *v16 = reinterpret_cast<const uint4 *>(dcp + head);
    for (uint64_t v = tid; v < nv; v += nthr) {
      const uint4 w = v16[v];
      const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int k2 = 0; k2 < 4; k2++) {
        touched += 4u - (uint32_t)__popc(byte_bits(tag_bytes(ws[k2], 0u)));
        clamped += (uint32_t)__popc(byte_bits(tag_bytes(ws[k2], qrep)));
      }
    }
    for (uint64_t i = tid; i < head + (n - tail0); i += nthr) {
      const uint8_t c = dcp[i < head ? i : tail0 + (i - head)];
      touched += c > 0;
      clamped += c == q;
    }
  } else {
    // one x row per block step (coalesced), rows dealt grid-stride
    for (uint32_t r = blockIdx.x; r < rows; r += gridDim.x) {
      const uint32_t y = r % by, z = r / by;  // uniform
      const int64_t l0 = (int64_t)D.lo[0] +
                         D.d.nx * ((int64_t)(y + D.lo[1]) + D.d.ny * (int64_t)(z + D.lo[2]));
      const int64_t g0 = G ? (int64_t)D.off[0] + gnx * ((int64_t)(y + D.off[1]) +
                                                         gny * (int64_t)(z + D.off[2]))
                           : 0;
      for (uint32_t x = threadIdx.x; x < bx; x += kSyncThreads) {
        const uint8_t c = D.dc[l0 + x];
        touched += c > 0;
        clamped += c == q;
        if (G) {
          G[g0 + x] = ((const T *)D.g)[l0 + x];
          if (DC) DC[g0 + x] = c;
        }
      }
    }
  }
  // one pair of atomics per block
  __shared__ uint32_t s_t[kSyncThreads / 64], s_c[kSyncThreads / 64];
  const uint32_t ts = wave_sum(touched), cs = wave_sum(clamped);
  if (lane_id() == 0) {
    s_t[threadIdx.x / 64] = ts;
    s_c[threadIdx.x / 64] = cs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t a = 0, c = 0;
    for (int w = 0; w < kSyncThreads / 64; w++) {
      a += s_t[w];
      c += s_c[w];
    }
    if (a) atomicAdd(&ds[k].touched, (unsigned long long)a);
    if (c) atomicAdd(&ds[k].clamped, (unsigned long long)c);
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

// Path compression by pointer jumping (the label of a voxel is the root of its
// pointer chain, PathCompression, pMSz.cu:255-271): in each round every voxel's
// pointer becomes its pointer's pointer, so chains of length L settle in about
// log2(L) rounds.  Two label fields (ascending, descending) share each round:
// threads [0, n) take lab0, [n, 2n) lab1.  A wave whose 64 voxels all point at
// roots (ballot) sets its bit in `settled` and later rounds skip it; a wave that
// still moved a pointer sets *changed.  Updates in place are safe: a read of a
// pointer being rewritten sees either value, and both are ancestors.
__global__ void k_jump(int32_t *__restrict__ lab0, int32_t *__restrict__ lab1, int64_t n,
                       uint32_t *__restrict__ settled, uint32_t *__restrict__ changed) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t w = (uint64_t)t >> 6;  // wave index (blocks are whole waves)
  if ((settled[w >> 5] >> (w & 31u)) & 1u) return;  // uniform over the wave
  bool moved = false;
  if (t < 2 * n) {
    int32_t *lab = t < n ? lab0 : lab1;
    const int64_t i = t < n ? t : t - n;
    const int32_t p = lab[i];
    const int32_t q = lab[p];
    if (q != p) {
      lab[i] = q;
      moved = true;
    }
  }
  const uint64_t mv = __ballot(moved);
  if (lane_id() == 0) {
    if (mv == 0) atomicOr(&settled[w >> 5], 1u << (w & 31u));
    else *changed = 1u;
  }
}

// extract_edits_kernel (pMSz.cu:2090-2138).  A grid of a few blocks per CU
// strides over the voxels and adds its count once per block (same-address
// device atomics serialise).
template <typename T>
__global__ __launch_bounds__(256) void k_edit_symbols(const T *__restrict__ f,
                                                      const T *__restrict__ g0,
                                                      const T *__restrict__ g,
                                                      uint8_t *__restrict__ sym,
                                                      unsigned long long *__restrict__ count,
                                                      int64_t n, EditParams ep) {
  __shared__ unsigned long long s_c[4];
  unsigned long long edited = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T a = g[i], b = g0[i];
    uint8_t s = 0;
    if (a != b && a < b) {
      edited++;
      if (a != clamp_value<T>(f[i], ep)) {
        T delta = (T)ep.delta_d;
        s = (uint8_t)(int)std::round((b - a) / delta);
      } else {
        s = (uint8_t)ep.q;
      }
    }
    sym[i] = s;
  }
  edited = wave_sum_u64(edited);
  if (__lane_id() == 0) s_c[threadIdx.x >> 6] = edited;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = s_c[0] + s_c[1] + s_c[2] + s_c[3];
    if (t) atomicAdd(count, t);
  }
}

// ---------------------------------------------------------------------------
// Host side.

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

// A device buffer that only grows.
template <typename E>
struct Buf {
  E *p = nullptr;
  size_t cap = 0;  // elements
  int ensure(size_t n) {
    if (n <= cap) return PMSZ_OK;
    (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, sizeof(E) * std::max<size_t>(n, 1)) != hipSuccess) {
      p = nullptr;
      return PMSZ_ENOMEM;
    }
    cap = n;
    return PMSZ_OK;
  }
};

pmsz_tuning default_tuning() {
  pmsz_tuning t;
  // measured on the 512^3 and 64-field workloads (autotuner), MI355X
  t.dense_div = 32;
  t.tail_targets = 4096;
  t.tail_blocks = 32;
  t.solo_targets = 64;
  t.zchunk = kColZ;
  t.graphs = 0;
  t.stream_div = 32;
  return t;
}

bool tuning_valid(const pmsz_tuning &t) {
  return t.dense_div >= 1 && t.dense_div <= 4096 && t.tail_targets >= 0 &&
         t.tail_blocks >= 1 && t.tail_blocks <= kTailBlocks && t.solo_targets >= 0 &&
         (t.zchunk == 16 || t.zchunk == 32 || t.zchunk == 64) && (t.graphs == 0 || t.graphs == 1) &&
         t.stream_div >= 0 && t.stream_div <= (1 << 24);
}

struct Ctx {
  std::mutex mu;  // one call at a time per context
  pmsz_tuning tune = default_tuning();
  // captured chains of multi-domain runs, by every value their launches take
  std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
  int device = -1;
  int cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr, ev4 = nullptr;
  Buf<uint8_t> code, dc, snap;   // per stored cell
  Buf<uint8_t> tbyte;            // target byte map of full-volume detections (16-padded)
  Buf<uint8_t> fpad, gpad;       // padded subdomain blocks (bytes)
  Buf<uint32_t> tmark, wmark;    // one bit per voxel id
  Buf<uint32_t> tlist, wlist;    // voxel ids
  Buf<Dom> doms;
  Buf<DomState> ds;
  Counters *ring = nullptr;
  RunState *run = nullptr;
  uint32_t *bar = nullptr;       // grid barrier state of the tail (kBarBytes)
  Counters *h_ring = nullptr;  // pinned mirror of the ring
  RunState *h_run = nullptr;   // pinned mirror of the run state
  DomState *h_ds = nullptr;    // pinned mirror of the domain states
  size_t h_ds_cap = 0;
  // the runtime refused a cooperative launch of the tail in this context: its
  // tails run as one workgroup from then on
  bool coop_refused = false;
  // Tags of the target byte map go on from run to run: since the map was last
  // cleared whole, it holds no tag from tag_next up to 255 (so a run can start
  // at tag_next without clearing).  0: its contents are unknown (new memory, or
  // scratch of another kernel) and the next run clears it whole.
  uint32_t tag_next = 0;
};

// Records a refused cooperative launch: the context falls back to one-workgroup
// tails, and the HIP error is reported on stderr once per process.
void note_coop_refused(Ctx &c, hipError_t e) {
  c.coop_refused = true;
  static std::atomic<int> said{0};
  if (said.exchange(1) == 0)
    fprintf(stderr, "pmsz: cooperative launch of the tail refused (%s: %s); tails run as one "
                    "workgroup in this context\n", hipGetErrorName(e), hipGetErrorString(e));
}

void clear_graphs(Ctx &c) {
  for (auto &kv : c.graphs) (void)hipGraphExecDestroy(kv.second);
  c.graphs.clear();
}

void ctx_release(Ctx &c) {
  clear_graphs(c);
  for (Buf<uint8_t> *b : {&c.code, &c.dc, &c.snap, &c.fpad, &c.gpad, &c.tbyte}) {
    (void)hipFree(b->p);
    *b = Buf<uint8_t>();
  }
  for (Buf<uint32_t> *b : {&c.tmark, &c.wmark, &c.tlist, &c.wlist}) {
    (void)hipFree(b->p);
    *b = Buf<uint32_t>();
  }
  (void)hipFree(c.doms.p);
  c.doms = Buf<Dom>();
  (void)hipFree(c.ds.p);
  c.ds = Buf<DomState>();
  (void)hipFree(c.ring);
  (void)hipFree(c.run);
  (void)hipFree(c.bar);
  if (c.h_ring) (void)hipHostFree(c.h_ring);
  if (c.h_run) (void)hipHostFree(c.h_run);
  if (c.h_ds) (void)hipHostFree(c.h_ds);
  c.ring = nullptr;
  c.run = nullptr;
  c.bar = nullptr;
  c.h_ring = nullptr;
  c.h_run = nullptr;
  c.h_ds = nullptr;
  c.h_ds_cap = 0;
  c.tag_next = 0;
  for (hipEvent_t *e : {&c.ev0, &c.ev1, &c.ev2, &c.ev3, &c.ev4}) {
    if (*e) (void)hipEventDestroy(*e);
    *e = nullptr;
  }
  if (c.stream) (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
  c.device = -1;
}

// Library-owned device bytes a context holds.
size_t ctx_bytes(const Ctx &c) {
  return c.code.cap + c.dc.cap + c.snap.cap + c.fpad.cap + c.gpad.cap + c.tbyte.cap +
         4 * (c.tmark.cap + c.wmark.cap + c.tlist.cap + c.wlist.cap) +
         sizeof(Dom) * c.doms.cap + sizeof(DomState) * c.ds.cap +
         (c.ring ? sizeof(Counters) * kCtrRing + sizeof(RunState) + kBarBytes : 0);
}

// The context of the calling thread's current device (created on first use; a
// context created on another device is rebuilt).
int ctx_device(Ctx &c) {
  int dev = 0;
  PMSZ_HIP(hipGetDevice(&dev));
  if (c.device == dev) return PMSZ_OK;
  if (c.device >= 0) ctx_release(c);
  c.device = dev;
  PMSZ_HIP(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, dev));
  PMSZ_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  PMSZ_HIP(hipEventCreate(&c.ev0));
  PMSZ_HIP(hipEventCreate(&c.ev1));
  PMSZ_HIP(hipEventCreate(&c.ev2));
  PMSZ_HIP(hipEventCreate(&c.ev3));
  PMSZ_HIP(hipEventCreate(&c.ev4));
  PMSZ_HIP(hipMalloc(&c.ring, sizeof(Counters) * kCtrRing));
  PMSZ_HIP(hipMalloc(&c.run, sizeof(RunState)));
  PMSZ_HIP(hipMalloc(&c.bar, kBarBytes));
  PMSZ_HIP(hipHostMalloc(&c.h_ring, sizeof(Counters) * kCtrRing, hipHostMallocDefault));
  PMSZ_HIP(hipHostMalloc(&c.h_run, sizeof(RunState), hipHostMallocDefault));
  return PMSZ_OK;
}

int ensure_host_ds(Ctx &c, size_t n) {
  if (n <= c.h_ds_cap) return PMSZ_OK;
  if (c.h_ds) (void)hipHostFree(c.h_ds);
  c.h_ds = nullptr;
  c.h_ds_cap = 0;
  PMSZ_HIP(hipHostMalloc(&c.h_ds, sizeof(DomState) * n, hipHostMallocDefault));
  c.h_ds_cap = n;
  return PMSZ_OK;
}

// The full-volume kernel addresses a plane with 32-bit byte offsets: a plane of
// the field, and of a padded block of it, stays below 4 GiB.
bool plane_fits(int dtype, int64_t nx, int64_t ny) {
  const int64_t s = dtype == PMSZ_F64 ? 8 : 4;
  return (nx + 2) * (ny + 2) * s < ((int64_t)1 << 32);
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
  if (p->mode != PMSZ_LAZY && p->mode != PMSZ_SYNC && p->mode != PMSZ_NAIVE) return PMSZ_EINVAL;
  if (p->clamp_mode != 0 && p->clamp_mode != 1) return PMSZ_EINVAL;
  if (p->ghost_policy != 0 && p->ghost_policy != 1) return PMSZ_EINVAL;
  for (int a = 0; a < 3; a++)
    if (p->sub[a] < 1) return PMSZ_EINVAL;
  if (p->sub[0] > p->nx || p->sub[1] > p->ny || p->sub[2] > p->nz) return PMSZ_EINVAL;
  if (n >= (int64_t)0xffffffffLL) return PMSZ_ETOOBIG;
  if (!plane_fits(p->dtype, p->nx, p->ny)) return PMSZ_ETOOBIG;
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

// Block split of one axis (pMSz.cu:2690-2706): the first n % parts blocks hold
// one more cell.
void block_split(int64_t n, int parts, int coord, int64_t &len, int64_t &off) {
  const int64_t base = n / parts, rem = n % parts;
  len = base + (coord < rem ? 1 : 0);
  off = (int64_t)coord * base + (coord < rem ? coord : rem);
}

EditParams edit_params(const pmsz_params *p) {
  EditParams ep;
  ep.eb = p->abs_eb;
  ep.delta_d = p->abs_eb / p->q;
  ep.q = p->q;
  ep.clamp_mode = p->clamp_mode;
  ep.budgets = 0;
  return ep;
}

// A domain whose classified box is the whole stored block (one field).
Dom field_dom(const pmsz_params *p, const void *f, void *g, uint8_t *code, uint8_t *dc) {
  Dom D;
  std::memset(&D, 0, sizeof(D));
  D.f = f;
  D.g = g;
  D.code = code;
  D.dc = dc;
  D.d = grid_of(p->nx, p->ny, p->nz);
  D.hi[0] = (int32_t)p->nx;
  D.hi[1] = (int32_t)p->ny;
  D.hi[2] = (int32_t)p->nz;
  D.zc = (uint32_t)kColZ;
  col_units(D, p->nx, p->ny, p->nz, kColZ);
  D.ep = edit_params(p);
  return D;
}

uint32_t dom_units(const Dom &D) { return D.nux * D.nuy * D.nuz; }

// Resident blocks per CU of a kernel (its register and LDS use), at least 1.
template <typename K>
int occupancy(K kernel, int threads) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess)
    nb = 1;
  return std::max(nb, 1);
}

// Persistent grid of the column kernel: one wave per unit up to the resident
// capacity, a multiple of 8 blocks (one share per XCD).
unsigned col_blocks(const Ctx &c, uint32_t units, int occ) {
  const int64_t blocks = ((int64_t)units + kColWaves - 1) / kColWaves;
  const int64_t per_xcd = (blocks + 7) / 8;
  const int64_t cap = std::max<int64_t>(1, (int64_t)(c.cus / 8) * occ);
  return (unsigned)(8 * std::max<int64_t>(1, std::min(per_xcd, cap)));
}

unsigned compact_blocks(const Ctx &c, int64_t groups) {
  const int64_t per = (int64_t)kCompactThreads * kCompactPT;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((groups + per - 1) / per, c.cus * 8));
}

constexpr uint32_t kChain = 8;  // apply/detect steps enqueued before each host check

// A run: its domains (device table in c.doms), the id layout and the schedule.
struct Job {
  std::vector<Dom> doms;
  uint32_t shift = 32;     // id = (domain << shift) | local (multi-domain runs)
  uint64_t wdense = 0;     // bitmap words holding ids, over all domains
  uint64_t words = 0;      // bitmap words
  uint64_t groups = 0;     // 16-byte groups of the target byte map
  uint64_t list_cap = 0;   // list entries (distinct cells)
  Sched s;
  bool subdomains = false; // padded blocks with ghost syncs
  int sub[3] = {1, 1, 1};
  int ndir = 14;
  int64_t gn[3] = {0, 0, 0};
};

// What a run leaves for the statistics.
struct Outcome {
  uint32_t last_j = 0;     // index of the last detection
  uint32_t done = 0;       // 1 converged, 2 capped
  int64_t outer_done = 0;  // ghost syncs + checks completed (subdomains)
  int64_t outer_started = 0;
  float ms_total = 0, ms_label = 0, ms_first = 0;
};

// The chain of one run: detection 1 (fused with the labeling), then apply /
// detect steps decided on the device, the tail, and -- for subdomains in lazy
// mode -- a ghost sync and a check detection whenever every subdomain's loop has
// ended.  In sync mode with subdomains the ghost sync follows every apply.
constexpr size_t kMaxGraphs = 64;  // captured chains a context keeps

// Launches the captured chain with this key, capturing and instantiating it on
// first use (enqueue(j0) records the chain's launches on the context's stream).
template <typename F>
int chain_graph(Ctx &c, const std::vector<uint64_t> &key, uint32_t j0, F &&enqueue) {
  auto it = c.graphs.find(key);
  if (it == c.graphs.end()) {
    if (c.graphs.size() >= kMaxGraphs) clear_graphs(c);
    hipGraph_t g = nullptr;
    PMSZ_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(j0);
    const hipError_t e = hipStreamEndCapture(c.stream, &g);
    if (rc) {
      if (e == hipSuccess && g) (void)hipGraphDestroy(g);
      return rc;
    }
    PMSZ_HIP(e);
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PMSZ_HIP(ei);
    it = c.graphs.emplace(key, x).first;
  }
  PMSZ_HIP(hipGraphLaunch(it->second, c.stream));
  return PMSZ_OK;
}

template <typename T, bool kMulti>
int drive(Ctx &c, Job &job, Outcome &out) {
  // the streamed apply is for one domain whose f, g and dc take 16-byte vector accesses
  if (kMulti ||
      (((uintptr_t)job.doms[0].f | (uintptr_t)job.doms[0].g | (uintptr_t)job.doms[0].dc) & 15u))
    job.s.stream_min = 0;
  // work lists from claims, except where ghost syncs mark the bitmaps too
  job.s.append = job.subdomains ? 0u : 1u;
  const uint32_t ndom = (uint32_t)job.doms.size();
  const Dom &d0 = job.doms[0];
  const Dom *doms = kMulti ? c.doms.p : nullptr;
  DomState *ds = kMulti ? c.ds.p : nullptr;
  uint32_t nunits = 0;
  for (const Dom &D : job.doms) nunits += dom_units(D);
  static const int occ_label = occupancy(k_columns<T, true, false, kMulti>, kColThreads);
  static const int occ_dense = occupancy(k_columns<T, false, true, kMulti>, kColThreads);
  static const int occ_dense1 = occupancy(k_columns<T, false, true, kMulti, true>, kColThreads);
  const unsigned g_label = col_blocks(c, nunits, occ_label);
  const unsigned g_dense = col_blocks(c, nunits, occ_dense);
  const unsigned g_dense1 = col_blocks(c, nunits, occ_dense1);
  // list kernels are latency-bound: as many resident waves as the device holds
  static const int occ_apply = occupancy(k_apply<T, kMulti>, kApplyThreads);
  static const int occ_dlist = occupancy(k_detect_list<T, kMulti>, kApplyThreads);
  const unsigned g_apply = (unsigned)std::max(1, c.cus * std::min(occ_apply, 8));
  const unsigned g_list = (unsigned)std::max(1, c.cus * std::min(occ_dlist, 8));
  // the tail: a grid of several blocks is launched cooperatively (the runtime
  // checks that every block can be resident); it takes at most
  // 1/(kMaxContexts+1) of what the device can hold of it, so the tails of all
  // contexts of this process fit together
  static const int occ_tail = occupancy(k_tail<T, kMulti>, kTailThreads);
  const unsigned g_tail = (unsigned)std::max(
      1, std::min(std::min(kTailBlocks, c.tune.tail_blocks), c.cus * occ_tail / (kMaxContexts + 1)));
  const unsigned g_compact = compact_blocks(c, (int64_t)job.groups);
  const unsigned g_bits = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>(((int64_t)job.wdense + kBitsThreads * kBitsPT - 1) /
                               (kBitsThreads * kBitsPT),
                           (int64_t)c.cus * 4));
  const bool sync_every = job.subdomains && job.s.mode == PMSZ_SYNC;
  const bool sync_lazy = job.subdomains && job.s.mode == PMSZ_LAZY;
  const int sx = job.sub[0], sy = job.sub[1], sz = job.sub[2];
  uint32_t shell_max = 0;
  for (const Dom &D : job.doms) shell_max = std::max(shell_max, D.shell_n);
  const dim3 g_shell(
      (unsigned)std::max<uint32_t>(1, std::min<uint32_t>((shell_max + kSyncThreads - 1) / kSyncThreads,
                                                         (uint32_t)c.cus * 4)),
      ndom);

  // Target tags go on from the previous run of this context: detection jd uses
  // detect_tag(jd + tag0), and the map holds none of this run's tags until they
  // start over, when the whole map is cleared (the map may be larger than this
  // run's cells: a later, larger run reads the rest).  tag0 = 0 clears it now.
  const uint64_t map_groups = c.tbyte.cap / 16;
  const uint32_t tag0 = c.tag_next ? (c.tag_next - 1u) % 255u : 0u;
  job.s.tag0 = tag0;
  c.tag_next = 0;  // unknown until the run has ended
  uint32_t max_jd = 1;  // the last detection enqueued
  const uint64_t init_work = std::max<uint64_t>(job.words / 4, tag0 == 0 ? map_groups : 0);
  const unsigned g_init =
      (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((init_work + 255) / 256, 4096));
  const unsigned g_fill =
      (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((map_groups + 255) / 256, 2048));
  PMSZ_HIP(hipEventRecord(c.ev0, c.stream));
  // clears the marks, counters, run state and barrier (and the map when tag0 = 0);
  // apply(1) reads slot 0's mode, which it sets to a running dense sweep
  k_init_run<<<g_init, 256, 0, c.stream>>>(c.tmark.p, c.wmark.p, job.words, c.ring, c.run,
                                           c.bar, (uint4 *)c.tbyte.p,
                                           tag0 == 0 ? map_groups : 0);
  // the original labeling (codes, delta counters cleared), then detection 1 as
  // a dense detection sweep (slot 0 holds mode Dense)
  PMSZ_HIP(hipEventRecord(c.ev2, c.stream));
  k_columns<T, true, false, kMulti><<<g_label, kColThreads, 0, c.stream>>>(
      d0, doms, ndom, ds, c.tbyte.p, c.ring, 0, 1, 1u, c.run, 0u);
  PMSZ_HIP(hipEventRecord(c.ev3, c.stream));
  // detection 1 also rejects NaN / sentinel values of g (padded blocks: the scatter did)
  if (job.subdomains)
    k_columns<T, false, true, kMulti><<<g_dense, kColThreads, 0, c.stream>>>(
        d0, doms, ndom, ds, c.tbyte.p, c.ring, 0, 1, 1u, c.run, detect_tag(1 + tag0));
  else
    k_columns<T, false, true, kMulti, true><<<g_dense1, kColThreads, 0, c.stream>>>(
        d0, doms, ndom, ds, c.tbyte.p, c.ring, 0, 1, 1u, c.run, detect_tag(1 + tag0));
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipEventRecord(c.ev4, c.stream));
  if (kMulti) {
    k_compact<<<g_compact, kCompactThreads, 0, c.stream>>>(c.tbyte.p, c.tlist.p, c.ring, 0, 1,
                                                           true, (uint32_t)job.groups, doms,
                                                           ndom, detect_tag(1 + tag0), 0u);
    k_domains<<<1, kMaxDomains, 0, c.stream>>>(ds, ndom, c.ring, 0, 0, 1, kOpDetect, job.s, 1);
  }

  uint32_t j0 = 1;
  out.outer_started = 1;
  // once a chain has ended in a work-list sweep, the next chain holds work-list
  // sweeps only (its applies may not choose a full-volume detection, which gives
  // the same results), so no launch of it waits for a dense phase
  bool sparse_chain = false;
  uint32_t chain = kChain;  // sweeps of the next chain
  Sched s_sparse = job.s;
  s_sparse.dense_ok = 0;
  while (true) {
    // the chain of sweeps j0 .. j0 + chain - 1, enqueued on the context's stream
    auto enqueue_chain = [&](uint32_t j0) -> int {
      for (uint32_t j = j0; j < j0 + chain; j++) {
        if (!kMulti && !sparse_chain) {
          // after a full-volume detection j: its target list
          k_compact<<<g_compact, kCompactThreads, 0, c.stream>>>(
              c.tbyte.p, c.tlist.p, c.ring, (j - 1) % kCtrRing, j % kCtrRing, false,
              (uint32_t)job.groups, doms, ndom, detect_tag(j + tag0), job.s.stream_min);
        }
        k_apply<T, kMulti><<<g_apply, kApplyThreads, 0, c.stream>>>(
            d0, doms, job.shift, ds, c.tmark.p, c.tlist.p, c.wmark.p, c.wlist.p, c.ring, c.run, j,
            sparse_chain ? s_sparse : job.s, c.tbyte.p, (uint32_t)job.groups);
        if (sync_every) {
          k_pack<T><<<g_shell, kSyncThreads, 0, c.stream>>>(doms, ds, job.gn[0], job.gn[1],
                                                            job.gn[2], c.ring, j, 1);
          k_unpack<T><<<g_shell, kSyncThreads, 0, c.stream>>>(doms, ds, sx, sy, sz, job.ndir,
                                                              c.wmark.p, c.wlist.p, c.ring, j, 1);
        }
        // the work list of detection j+1 from the bitmap that apply j (and a ghost
        // sync) marked
        if (!job.s.append)
          k_compact_bits<<<g_bits, kBitsThreads, 0, c.stream>>>(
            c.wmark.p, doms, ndom, (uint32_t)job.wdense, c.wlist.p, &c.ring[j % kCtrRing].nwork,
            &c.ring[j % kCtrRing], kModeSparse);
        // the tags start over: no mark of an earlier sweep may carry detection j+1's
        if (detect_tag(j + 1 + tag0) == 1u)
          k_fill16<<<g_fill, 256, 0, c.stream>>>((uint4 *)c.tbyte.p, map_groups);
        if (!sparse_chain)
          k_columns<T, false, true, kMulti><<<g_dense, kColThreads, 0, c.stream>>>(
              d0, doms, ndom, ds, c.tbyte.p, c.ring, j % kCtrRing, (j + 1) % kCtrRing,
              (j + 1) & 1u, c.run, detect_tag(j + 1 + tag0));
        if (kMulti && !sparse_chain)
          k_compact<<<g_compact, kCompactThreads, 0, c.stream>>>(
              c.tbyte.p, c.tlist.p, c.ring, j % kCtrRing, (j + 1) % kCtrRing, false,
              (uint32_t)job.groups, doms, ndom, detect_tag(j + 1 + tag0), job.s.stream_min);
        k_detect_list<T, kMulti><<<g_list, kApplyThreads, 0, c.stream>>>(
            d0, doms, job.shift, ds, c.tmark.p, c.tlist.p, c.wmark.p, c.wlist.p, c.ring, c.run, j,
            job.s.append);
        // the targets of a work-list detection j+1, from the bitmap
        if (!job.s.append)
          k_compact_bits<<<g_bits, kBitsThreads, 0, c.stream>>>(
            c.tmark.p, doms, ndom, (uint32_t)job.wdense, c.tlist.p,
            &c.ring[(j + 1) % kCtrRing].ntargets, &c.ring[j % kCtrRing], kModeSparse);
        if (kMulti)
          k_domains<<<1, kMaxDomains, 0, c.stream>>>(ds, ndom, c.ring, j, 1, j + 1, kOpDetect,
                                                     job.s, 1);
      }
      PMSZ_HIP(hipGetLastError());
      return PMSZ_OK;
    };
    if (kMulti && c.tune.graphs) {
      // replay of a captured chain: one graph per job layout and first sweep,
      // kept by the context while the layout recurs.  The key holds every value
      // the chain's launches take (d0 is not read by multi-domain kernels).
      const std::vector<uint64_t> key = {
          sizeof(T), j0, ndom, job.shift, job.groups, job.wdense, (uint64_t)(uintptr_t)doms,
          (uint64_t)(uintptr_t)ds, (uint64_t)(uintptr_t)c.tbyte.p, (uint64_t)(uintptr_t)c.tmark.p,
          (uint64_t)(uintptr_t)c.tlist.p, (uint64_t)(uintptr_t)c.wmark.p,
          (uint64_t)(uintptr_t)c.wlist.p, (uint64_t)(uintptr_t)c.ring, (uint64_t)(uintptr_t)c.run,
          (uint64_t)job.s.mode, job.s.max_sweeps, job.s.dense_min_work, job.s.tail_ok, tag0,
          map_groups,
          job.s.tail_targets, job.s.solo_targets, job.s.dense_ok, job.s.stream_min, job.s.append,
          sparse_chain ? 1u : 0u, chain,
          g_compact,
          g_apply, g_bits, g_fill, g_dense, g_list, g_shell.x,
          g_shell.y, sync_every ? 1u : 0u, (uint64_t)sx, (uint64_t)sy, (uint64_t)sz,
          (uint64_t)job.ndir, (uint64_t)job.gn[0], (uint64_t)job.gn[1], (uint64_t)job.gn[2]};
      int rc = chain_graph(c, key, j0, enqueue_chain);
      if (rc) return rc;
    } else {
      int rc = enqueue_chain(j0);
      if (rc) return rc;
    }
    if (!sparse_chain) max_jd = std::max(max_jd, j0 + chain);
    if (job.s.tail_ok) {
      Dom a_d0 = d0;
      const Dom *a_doms = doms;
      uint32_t a_shift = job.shift, a_ndom = ndom;
      DomState *a_ds = ds;
      uint32_t *a_tmark = c.tmark.p, *a_tlist = c.tlist.p, *a_wmark = c.wmark.p,
               *a_wlist = c.wlist.p;
      Counters *a_ring = c.ring;
      RunState *a_run = c.run;
      Sched a_s = job.s;
      uint32_t *a_bar = c.bar;
      void *args[] = {&a_d0,    &a_doms,  &a_shift, &a_ndom, &a_ds,  &a_tmark,
                      &a_tlist, &a_wmark, &a_wlist, &a_ring, &a_run, &a_s, &a_bar};
      hipError_t e = hipSuccess;
      if (g_tail > 1 && !c.coop_refused) {
        e = hipLaunchCooperativeKernel((const void *)k_tail<T, kMulti>, dim3(g_tail),
                                       dim3(kTailThreads), args, 0, c.stream);
        if (e != hipSuccess) {
          // the runtime can refuse a cooperative launch (another process may hold
          // the device's cooperative resources); the context then launches the
          // tail as one workgroup, which needs no co-residency
          note_coop_refused(c, e);
          (void)hipGetLastError();
        }
      }
      if (g_tail == 1 || c.coop_refused)
        e = hipLaunchKernel((const void *)k_tail<T, kMulti>, dim3(1), dim3(kTailThreads), args, 0,
                            c.stream);
      PMSZ_HIP(e);
    }
    PMSZ_HIP(hipGetLastError());
    PMSZ_HIP(hipMemcpyAsync(c.h_run, c.run, sizeof(RunState), hipMemcpyDeviceToHost, c.stream));
    PMSZ_HIP(hipStreamSynchronize(c.stream));
    if (c.h_run->err & kErrBarrier) return PMSZ_EDEVICE;
    if (c.h_run->err) return PMSZ_EINVAL;
    if (!c.h_run->done) {
      // the chain ended before the tail took over: continue it where it stopped
      j0 += chain;
      chain = kChain;
      sparse_chain = c.h_run->last_mode == kModeSparse;
      continue;
    }
    const uint32_t m = c.h_run->last_j;
    if (!(sync_lazy && c.h_run->done == 1)) {
      out.last_j = m;
      out.done = c.h_run->done;
      break;
    }
    // every subdomain's inner loop has ended: ghost sync, then the check sweep
    // (pMSz.cu:1991-2060), which doubles as the first sweep of the next loops
    out.outer_done++;
    const uint32_t emit_ok = out.outer_done < (int64_t)job.s.max_sweeps ? 1u : 0u;
    k_reset_phase<<<1, 64, 0, c.stream>>>(c.ring, c.run, m);
    k_pack<T><<<g_shell, kSyncThreads, 0, c.stream>>>(doms, ds, job.gn[0], job.gn[1], job.gn[2],
                                                      c.ring, m, 0);
    k_unpack<T><<<g_shell, kSyncThreads, 0, c.stream>>>(doms, ds, sx, sy, sz, job.ndir,
                                                        c.wmark.p, c.wlist.p, c.ring, m, 0);
    k_compact_bits<<<g_bits, kBitsThreads, 0, c.stream>>>(
        c.wmark.p, doms, ndom, (uint32_t)job.wdense, c.wlist.p, &c.ring[m % kCtrRing].nwork,
        &c.ring[m % kCtrRing], kModeSparse);
    k_domains<<<1, kMaxDomains, 0, c.stream>>>(ds, ndom, c.ring, 0, 0, m + 1, kOpPreCheck, job.s,
                                               emit_ok);
    k_detect_list<T, kMulti><<<g_list, kApplyThreads, 0, c.stream>>>(
        d0, doms, job.shift, ds, c.tmark.p, c.tlist.p, c.wmark.p, c.wlist.p, c.ring, c.run, m,
        job.s.append);
    k_compact_bits<<<g_bits, kBitsThreads, 0, c.stream>>>(
        c.tmark.p, doms, ndom, (uint32_t)job.wdense, c.tlist.p,
        &c.ring[(m + 1) % kCtrRing].ntargets, &c.ring[m % kCtrRing], kModeSparse);
    k_domains<<<1, kMaxDomains, 0, c.stream>>>(ds, ndom, c.ring, 0, 0, m + 1, kOpCheck, job.s, 1);
    PMSZ_HIP(hipGetLastError());
    PMSZ_HIP(hipMemcpyAsync(c.h_ring + ((m + 1) % kCtrRing), c.ring + ((m + 1) % kCtrRing),
                            sizeof(Counters), hipMemcpyDeviceToHost, c.stream));
    PMSZ_HIP(hipStreamSynchronize(c.stream));
    const Counters &cc = c.h_ring[(m + 1) % kCtrRing];
    const bool violations = (cnt_host(cc, 0) | cnt_host(cc, 1) | cnt_host(cc, 2)) != 0;
    if (!violations || !emit_ok) {
      // converged, or the cap on outer iterations stops before the next one
      out.last_j = m + 1;
      out.done = violations ? 2 : 1;
      break;
    }
    // the next outer iteration starts (every subdomain's ite counts it)
    out.outer_started++;
    if (cc.nactive == 0 && cc.ntargets == 0) {
      // every subdomain with violations is capped before it may sweep again
      out.last_j = m + 1;
      out.done = 2;
      break;
    }
    // apply the check's targets (the first sweep of the new loops); when every
    // subdomain is capped after it, that apply ends the run
    j0 = m + 1;
    sparse_chain = false;
    // few targets: the first apply hands over to the tail at once, so the chain
    // before the tail launch holds that one sweep
    chain = job.s.tail_ok && cc.ntargets <= job.s.tail_targets ? 1u : kChain;
  }
  PMSZ_HIP(hipEventRecord(c.ev1, c.stream));
  PMSZ_HIP(hipMemcpyAsync(c.h_ring, c.ring, sizeof(Counters) * kCtrRing, hipMemcpyDeviceToHost,
                          c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  PMSZ_HIP(hipEventElapsedTime(&out.ms_total, c.ev0, c.ev1));
  PMSZ_HIP(hipEventElapsedTime(&out.ms_label, c.ev2, c.ev3));
  PMSZ_HIP(hipEventElapsedTime(&out.ms_first, c.ev3, c.ev4));
  // the next run starts after the last tag this one may have written
  c.tag_next = detect_tag(max_jd + tag0) + 1u;
  return PMSZ_OK;
}

// Bitmap words and list capacity of a run, allocating what is missing.
int prepare(Ctx &c, Job &job) {
  // column units with the context's tile depth
  uint32_t ubase = 0;
  for (Dom &D : job.doms) {
    D.zc = (uint32_t)c.tune.zchunk;
    col_units(D, D.hi[0] - D.lo[0], D.hi[1] - D.lo[1], D.hi[2] - D.lo[2], D.zc);
    D.unit_base = ubase;
    ubase += D.nux * D.nuy * D.nuz;
  }
  uint64_t cells = 0, ids = 0;
  for (const Dom &D : job.doms) cells += (uint64_t)D.d.n;
  if (job.doms.size() == 1) ids = (uint64_t)job.doms[0].d.n;
  else ids = (uint64_t)job.doms.size() << job.shift;
  job.words = (ids + 31) / 32 + 1;
  // the words the domains' ids occupy, back to back (each gbase is a multiple of 32)
  job.wdense = 0;
  for (Dom &D : job.doms) {
    D.wpre = (uint32_t)job.wdense;
    job.wdense += ((uint64_t)D.d.n + 31) / 32;
  }
  job.groups = (cells + 15) / 16;
  job.list_cap = cells;
  int rc;
  const size_t map_before = c.tbyte.cap;
  if ((rc = c.tbyte.ensure(16 * job.groups))) return rc;
  if (c.tbyte.cap != map_before) c.tag_next = 0;  // new memory: contents unknown
  if ((rc = c.tmark.ensure(job.words)) ||
      (rc = c.wmark.ensure(job.words)) ||
      (rc = c.tlist.ensure(job.list_cap)) || (rc = c.wlist.ensure(job.list_cap)))
    return rc;
  if (job.doms.size() > 1 || job.subdomains) {
    if ((rc = c.doms.ensure(job.doms.size())) || (rc = c.ds.ensure(job.doms.size())) ||
        (rc = ensure_host_ds(c, job.doms.size())))
      return rc;
    PMSZ_HIP(hipMemcpyAsync(c.doms.p, job.doms.data(), sizeof(Dom) * job.doms.size(),
                            hipMemcpyHostToDevice, c.stream));
    DomState init;
    std::memset(&init, 0, sizeof(init));
    init.flags = kDomActive | kDomEmit;
    for (size_t k = 0; k < job.doms.size(); k++) c.h_ds[k] = init;
    PMSZ_HIP(hipMemcpyAsync(c.ds.p, c.h_ds, sizeof(DomState) * job.doms.size(),
                            hipMemcpyHostToDevice, c.stream));
  }
  return PMSZ_OK;
}

uint32_t ceil_log2(uint64_t v) {
  uint32_t s = 0;
  while (((uint64_t)1 << s) < v) s++;
  return s;
}

Sched make_sched(const Ctx &c, const pmsz_params *p, uint64_t work_voxels, bool tail_ok) {
  Sched s;
  s.mode = p->mode;
  s.max_sweeps = (uint32_t)p->max_sweeps;
  s.dense_min_work = (uint32_t)std::min<uint64_t>(
      std::max<uint64_t>(work_voxels / (uint64_t)c.tune.dense_div, 1), 0xffffffffull);
  s.tail_ok = tail_ok ? 1u : 0u;
  s.tail_targets = (uint32_t)c.tune.tail_targets;
  s.solo_targets = (uint32_t)c.tune.solo_targets;
  s.dense_ok = 1;
  s.append = 1;
  s.tag0 = 0;
  s.stream_min = c.tune.stream_div
                     ? (uint32_t)std::min<uint64_t>(
                           std::max<uint64_t>(work_voxels / (uint64_t)c.tune.stream_div, 1),
                           0xffffffffull)
                     : 0u;
  return s;
}

void fill_single_stats(pmsz_stats *st, int mode, uint32_t m, bool conv) {
  if (mode == PMSZ_LAZY) {
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
}

template <typename T>
int correct_single(Ctx &c, const pmsz_params *p, const T *f, T *g, uint8_t *dc_out,
                   pmsz_stats *st) {
  const int64_t n = p->nx * p->ny * p->nz;
  int rc;
  if ((rc = c.code.ensure(n)) || (rc = c.dc.ensure(n))) return rc;
  Job job;
  job.doms.push_back(field_dom(p, f, g, c.code.p, c.dc.p));
  job.doms[0].ep.budgets = p->q <= 15 ? 1 : 0;
  job.s = make_sched(c, p, (uint64_t)n, true);
  if ((rc = prepare(c, job))) return rc;
  std::memset(st, 0, sizeof(*st));
  Outcome o;
  rc = drive<T, false>(c, job, o);
  if (rc) return rc;
  if (dc_out) {
    if (budgets_on(job.doms[0].ep)) {  // the counters without their step budgets
      const unsigned gb = (unsigned)std::max<int64_t>(
          1, std::min<int64_t>((n / 16 + 255) / 256, 4096));
      k_dc_out<<<gb, 256, 0, c.stream>>>(dc_out, c.dc.p, (uint64_t)n);
      PMSZ_HIP(hipGetLastError());
    } else {
      PMSZ_HIP(hipMemcpyAsync(dc_out, c.dc.p, (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    }
  }
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  st->t_total_s = o.ms_total * 1e-3;
  st->t_dense_s = o.ms_first * 1e-3;
  st->t_label_s = o.ms_label * 1e-3;
  const RunState &r = *c.h_run;
  const uint32_t m = o.last_j;
  const Counters &cm = c.h_ring[m % kCtrRing];
  const bool conv = o.done == 1;
  const int status = conv ? PMSZ_OK : PMSZ_NOT_CONVERGED;
  st->status = status;
  for (int q = 0; q < 3; q++) st->final_cnt[q] = cnt_host(cm, q);
  if (m <= (uint32_t)kCtrRing) {
    st->trace_n = std::min<int64_t>(m, 64);
    for (int64_t k = 0; k < st->trace_n; k++) {
      const Counters &ck = c.h_ring[(k + 1) % kCtrRing];
      for (int q = 0; q < 3; q++) st->trace_cnt[k][q] = cnt_host(ck, q);
      st->trace_targets[k] = ck.ntargets + cnt_host(ck, kCntStreamed);
    }
  }
  st->total_targets = (int64_t)acc_host(r, kAccTargets);
  st->touched_voxels = (int64_t)acc_host(r, kAccTouched);
  st->clamped_voxels = (int64_t)acc_host(r, kAccClamped);
  st->sparse_voxels = (int64_t)acc_host(r, kAccSparse);
  st->dense_passes = 1 + (int64_t)r.dense_passes;  // the labeling + detection sweeps
  fill_single_stats(st, p->mode, m, conv);
  // labeling: f read, code and dc written; full-volume detections: g and code
  // read; edits: f, g, dc read and g, dc written; listed voxels: g and code read
  st->bytes_algorithmic = (double)n * (sizeof(T) + 2.0) +
                          (double)n * (sizeof(T) + 1.0) * (double)r.dense_passes +
                          (double)st->total_targets * (3.0 * sizeof(T) + 2.0) +
                          (double)st->sparse_voxels * (sizeof(T) + 1.0);
  return status;
}

// Subdomains: the global field split into sub[0] x sub[1] x sub[2] padded blocks
// (pMSz.cu:2690-2711), corrected in lockstep with the ghost syncs of the lazy
// (pMSz.cu:1872-2074) or per-iteration (sync_pMSz.cu:2102-2200) loop.
template <typename T>
int correct_subdomains(Ctx &c, const pmsz_params *p, const T *F, T *G, uint8_t *dc_out,
                       pmsz_stats *st) {
  const int sx = p->sub[0], sy = p->sub[1], sz = p->sub[2];
  const uint32_t nsub = (uint32_t)(sx * sy * sz);
  if (nsub > kMaxDomains) return PMSZ_EINVAL;
  Job job;
  job.subdomains = true;
  job.sub[0] = sx;
  job.sub[1] = sy;
  job.sub[2] = sz;
  job.ndir = p->ghost_policy == PMSZ_GHOST_FULL ? 14 : 12;
  job.gn[0] = p->nx;
  job.gn[1] = p->ny;
  job.gn[2] = p->nz;
  uint64_t cells = 0, maxcells = 0, core = 0;
  std::vector<int64_t> offs;
  for (int cz = 0; cz < sz; cz++)
    for (int cy = 0; cy < sy; cy++)
      for (int cx = 0; cx < sx; cx++) {
        int64_t l[3], o[3];
        block_split(p->nx, sx, cx, l[0], o[0]);
        block_split(p->ny, sy, cy, l[1], o[1]);
        block_split(p->nz, sz, cz, l[2], o[2]);
        Dom D;
        std::memset(&D, 0, sizeof(D));
        D.d = grid_of(l[0] + 2, l[1] + 2, l[2] + 2);
        for (int a = 0; a < 3; a++) {
          D.lo[a] = 1;
          D.hi[a] = (int32_t)l[a] + 1;
          D.off[a] = (int32_t)o[a];
        }
        D.cc[0] = cx;
        D.cc[1] = cy;
        D.cc[2] = cz;
        D.zc = (uint32_t)kColZ;
        col_units(D, l[0], l[1], l[2], kColZ);
        D.shell_n = shell_count((uint32_t)D.d.nx, (uint32_t)D.d.ny, (uint32_t)D.d.nz);
        D.padded = 1;
        D.ep = edit_params(p);
        offs.push_back((int64_t)cells);
        cells += (uint64_t)D.d.n;
        core += (uint64_t)(l[0] * l[1] * l[2]);
        maxcells = std::max<uint64_t>(maxcells, (uint64_t)D.d.n);
        job.doms.push_back(D);
      }
  job.shift = std::max<uint32_t>(5, ceil_log2(maxcells));
  if (((uint64_t)nsub << job.shift) > 0xffffffffull || cells >= 0xffffffffull)
    return PMSZ_ETOOBIG;
  int rc;
  if ((rc = c.code.ensure(cells)) || (rc = c.dc.ensure(cells)) || (rc = c.snap.ensure(cells)) ||
      (rc = c.fpad.ensure(cells * sizeof(T))) || (rc = c.gpad.ensure(cells * sizeof(T))))
    return rc;
  uint32_t ubase = 0;
  for (uint32_t k = 0; k < nsub; k++) {
    Dom &D = job.doms[k];
    D.f = (const void *)((T *)c.fpad.p + offs[k]);
    D.g = (void *)((T *)c.gpad.p + offs[k]);
    D.code = c.code.p + offs[k];
    D.dc = c.dc.p + offs[k];
    D.snap = c.snap.p + offs[k];
    D.gbase = k << job.shift;
    D.cbase = (uint32_t)offs[k];
    D.unit_base = ubase;
    ubase += dom_units(D);
  }
  job.s = make_sched(c, p, core, p->mode == PMSZ_LAZY);
  if ((rc = prepare(c, job))) return rc;
  std::memset(st, 0, sizeof(*st));
  // padded blocks with their initial ghost values; run->err flags bad inputs
  PMSZ_HIP(hipMemsetAsync(c.run, 0, sizeof(RunState), c.stream));
  uint32_t maxn = 0;
  for (const Dom &D : job.doms) maxn = std::max<uint32_t>(maxn, (uint32_t)D.d.n);
  const dim3 g_all((unsigned)std::min<uint32_t>((maxn + kSyncThreads - 1) / kSyncThreads,
                                                (uint32_t)c.cus * 8),
                   nsub);
  k_scatter<T><<<g_all, kSyncThreads, 0, c.stream>>>(c.doms.p, F, G, p->nx, p->ny, sx, sy, sz,
                                                     p->ghost_policy, c.run);
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipMemcpyAsync(c.h_run, c.run, sizeof(RunState), hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  if (c.h_run->err) return PMSZ_EINVAL;
  Outcome o;
  rc = drive<T, true>(c, job, o);
  if (rc) return rc;
  uint32_t maxrows = 0;
  for (const Dom &D : job.doms)
    maxrows = std::max<uint32_t>(maxrows, (uint32_t)((D.hi[1] - D.lo[1]) * (D.hi[2] - D.lo[2])));
  const dim3 g_fin(std::min<uint32_t>(maxrows, 512u), nsub);
  k_finish<T><<<g_fin, kSyncThreads, 0, c.stream>>>(c.doms.p, c.ds.p, G, dc_out, p->nx, p->ny);
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipMemcpyAsync(c.h_ds, c.ds.p, sizeof(DomState) * nsub, hipMemcpyDeviceToHost,
                          c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  const RunState &r = *c.h_run;
  const bool conv = o.done == 1;
  st->status = conv ? PMSZ_OK : PMSZ_NOT_CONVERGED;
  st->t_total_s = o.ms_total * 1e-3;
  st->t_dense_s = o.ms_first * 1e-3;
  st->t_label_s = o.ms_label * 1e-3;
  st->sparse_voxels = (int64_t)acc_host(r, kAccSparse);
  st->dense_passes = 1 + (int64_t)r.dense_passes;
  int64_t sweeps = 0, ite = 0;
  for (uint32_t k = 0; k < nsub; k++) {
    const DomState &s = c.h_ds[k];
    sweeps += s.sweeps;
    st->total_targets += (int64_t)s.targets;
    st->touched_voxels += (int64_t)s.touched;
    st->clamped_voxels += (int64_t)s.clamped;
    for (int q = 0; q < 3; q++)
      if (!conv) st->final_cnt[q] = std::max(st->final_cnt[q], s.last_cnt[q]);
    // "iteration is:" is printed by rank 0 (pMSz.cu:2073): the block at (0, 0, 0)
    if (k == 0) ite = o.outer_started + (int64_t)s.sweeps;
  }
  if (p->mode == PMSZ_LAZY) {
    st->edit_sweeps = sweeps;
    st->outer_iters = o.outer_done;
    st->check_sweeps = (int64_t)nsub * o.outer_done;
    st->ite = ite;
  } else {
    // one edit sweep and one check per subdomain and outer iteration
    const uint32_t m = o.last_j;
    const int64_t applies = conv ? std::max<int64_t>((int64_t)m - 1, 1) : (int64_t)m - 1;
    st->edit_sweeps = (int64_t)nsub * applies;
    st->check_sweeps = (int64_t)nsub * applies;
    st->outer_iters = applies;
    st->ite = 2 * applies;
  }
  st->bytes_algorithmic = (double)cells * (sizeof(T) + 2.0) +
                          (double)cells * (sizeof(T) + 1.0) * (double)r.dense_passes +
                          (double)st->total_targets * (3.0 * sizeof(T) + 2.0) +
                          (double)st->sparse_voxels * (sizeof(T) + 1.0);
  return st->status;
}

// Batch: independent fields of one value type corrected in lockstep.  Each field
// behaves exactly as its own pmsz_correct.
template <typename T>
int correct_batch(Ctx &c, int n, const pmsz_params *p, const void *const *d_f,
                  void *const *d_g, uint8_t *const *d_dc, pmsz_stats *stats) {
  // groups of fields whose voxel ids fit 32 bits
  int first = 0;
  while (first < n) {
    uint64_t maxn = 0, cells = 0;
    int last = first;
    while (last < n && last - first < (int)kMaxDomains) {
      const uint64_t nk = (uint64_t)(p[last].nx * p[last].ny * p[last].nz);
      const uint64_t mx = std::max(maxn, nk);
      const uint32_t sh = std::max<uint32_t>(5, ceil_log2(mx));
      if (((uint64_t)(last - first + 1) << sh) > 0xffffffffull) break;
      maxn = mx;
      cells += nk;
      last++;
    }
    if (last == first) return PMSZ_ETOOBIG;
    const int cnt = last - first;
    int rc;
    if (cnt == 1) {
      rc = correct_single<T>(c, &p[first], (const T *)d_f[first], (T *)d_g[first],
                             d_dc ? d_dc[first] : nullptr, &stats[first]);
      if (rc != PMSZ_OK && rc != PMSZ_NOT_CONVERGED) return rc;
      first = last;
      continue;
    }
    if ((rc = c.code.ensure(cells)) || (rc = c.dc.ensure(cells))) return rc;
    Job job;
    job.shift = std::max<uint32_t>(5, ceil_log2(maxn));
    uint64_t off = 0;
    uint32_t ubase = 0;
    for (int k = 0; k < cnt; k++) {
      const pmsz_params *pk = &p[first + k];
      Dom D = field_dom(pk, d_f[first + k], d_g[first + k], c.code.p + off, c.dc.p + off);
      D.gbase = (uint32_t)k << job.shift;
      D.cbase = (uint32_t)off;
      D.unit_base = ubase;
      ubase += dom_units(D);
      off += (uint64_t)D.d.n;
      job.doms.push_back(D);
    }
    job.s = make_sched(c, &p[first], cells, true);
    if ((rc = prepare(c, job))) return rc;
    Outcome o;
    rc = drive<T, true>(c, job, o);
    if (rc) return rc;
    uint32_t maxrows = 0;
    for (const Dom &D : job.doms)
      maxrows = std::max<uint32_t>(maxrows, (uint32_t)((D.hi[1] - D.lo[1]) * (D.hi[2] - D.lo[2])));
    const dim3 g_fin(std::min<uint32_t>(maxrows, 256u), (unsigned)cnt);
    k_finish<T><<<g_fin, kSyncThreads, 0, c.stream>>>(c.doms.p, c.ds.p, nullptr, nullptr, 0, 0);
    PMSZ_HIP(hipGetLastError());
    off = 0;
    for (int k = 0; k < cnt; k++) {
      const uint64_t nk = (uint64_t)job.doms[k].d.n;
      if (d_dc && d_dc[first + k])
        PMSZ_HIP(hipMemcpyAsync(d_dc[first + k], c.dc.p + off, nk, hipMemcpyDeviceToDevice,
                                c.stream));
      off += nk;
    }
    PMSZ_HIP(hipMemcpyAsync(c.h_ds, c.ds.p, sizeof(DomState) * cnt, hipMemcpyDeviceToHost,
                            c.stream));
    PMSZ_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < cnt; k++) {
      const DomState &s = c.h_ds[k];
      pmsz_stats *st = &stats[first + k];
      std::memset(st, 0, sizeof(*st));
      const bool conv = s.status == kDomConverged;
      st->status = conv ? PMSZ_OK : PMSZ_NOT_CONVERGED;
      for (int q = 0; q < 3; q++) st->final_cnt[q] = s.last_cnt[q];
      st->total_targets = (int64_t)s.targets;
      st->touched_voxels = (int64_t)s.touched;
      st->clamped_voxels = (int64_t)s.clamped;
      fill_single_stats(st, p[first + k].mode, s.last_j, conv);
      st->t_total_s = o.ms_total * 1e-3;
      st->t_dense_s = o.ms_first * 1e-3;
      st->t_label_s = o.ms_label * 1e-3;
      const uint64_t nk = (uint64_t)job.doms[k].d.n;
      st->bytes_algorithmic = (double)nk * (2.0 * sizeof(T) + 3.0) +
                              (double)st->total_targets * (3.0 * sizeof(T) + 2.0);
    }
    first = last;
  }
  return PMSZ_OK;
}

// Timing of the fused first pass alone (events around each launch), for the
// roofline report.
template <typename T>
int time_first_pass(Ctx &c, const pmsz_params *p, const T *f, const T *g, int iters,
                    double *avg_s) {
  const int64_t n = p->nx * p->ny * p->nz;
  int rc;
  if ((rc = c.code.ensure(n)) || (rc = c.dc.ensure(n))) return rc;
  Job job;
  job.doms.push_back(field_dom(p, f, (void *)g, c.code.p, c.dc.p));
  job.s = make_sched(c, p, (uint64_t)n, true);
  if ((rc = prepare(c, job))) return rc;
  static const int occ_l = occupancy(k_columns<T, true, false, false>, kColThreads);
  static const int occ_d = occupancy(k_columns<T, false, true, false, true>, kColThreads);
  const unsigned gl = col_blocks(c, dom_units(job.doms[0]), occ_l);
  const unsigned gd = col_blocks(c, dom_units(job.doms[0]), occ_d);
  double tl = 0, td = 0;
  static const uint32_t kRunning = kModeDense;
  c.tag_next = 0;  // the passes below leave tag-1 marks in the map
  for (int it = 0; it < iters; it++) {
    PMSZ_HIP(hipMemsetAsync(c.tbyte.p, 0, 16 * job.groups, c.stream));
    PMSZ_HIP(hipMemsetAsync(c.ring, 0, sizeof(Counters) * 2, c.stream));
    PMSZ_HIP(hipMemcpyAsync(&c.ring[0].mode, &kRunning, sizeof(uint32_t), hipMemcpyHostToDevice,
                            c.stream));
    PMSZ_HIP(hipEventRecord(c.ev0, c.stream));
    k_columns<T, true, false, false><<<gl, kColThreads, 0, c.stream>>>(
        job.doms[0], nullptr, 1, nullptr, c.tbyte.p, c.ring, 0, 1, 1u, c.run, 0u);
    PMSZ_HIP(hipEventRecord(c.ev1, c.stream));
    k_columns<T, false, true, false, true><<<gd, kColThreads, 0, c.stream>>>(
        job.doms[0], nullptr, 1, nullptr, c.tbyte.p, c.ring, 0, 1, 1u, c.run, 1u);
    PMSZ_HIP(hipGetLastError());
    PMSZ_HIP(hipEventRecord(c.ev2, c.stream));
    PMSZ_HIP(hipStreamSynchronize(c.stream));
    float ms1 = 0, ms2 = 0;
    PMSZ_HIP(hipEventElapsedTime(&ms1, c.ev0, c.ev1));
    PMSZ_HIP(hipEventElapsedTime(&ms2, c.ev1, c.ev2));
    tl += ms1 * 1e-3;
    td += ms2 * 1e-3;
  }
  avg_s[0] = tl / std::max(1, iters);
  avg_s[1] = td / std::max(1, iters);
  return PMSZ_OK;
}

}  // namespace pmsz

using namespace pmsz;

struct pmsz_ctx {
  Ctx c;
};

namespace pmsz {
// The codes of one unpadded field (the labeling pass of the full-volume kernel).
template <typename T>
int launch_codes(Ctx &c, const pmsz_params *p, const T *field, uint8_t *codes) {
  Dom D = field_dom(p, field, nullptr, codes, nullptr);
  static const int occ = occupancy(k_columns<T, true, false, false>, kColThreads);
  k_columns<T, true, false, false><<<col_blocks(c, dom_units(D), occ), kColThreads, 0,
                                     c.stream>>>(D, nullptr, 1, nullptr, (uint8_t *)nullptr,
                                                 c.ring, 0, 0, 0u, c.run, 0u);
  PMSZ_HIP(hipGetLastError());
  return PMSZ_OK;
}

// Labels (pointer roots) from codes: pointer jumping in batches of 4 rounds until
// a batch moves nothing.  settled holds sbytes of wave bits; changed one word.
int labels_from_codes(Ctx &c, const uint8_t *codes, int32_t *asc, int32_t *desc, const Grid3 &d,
                      uint32_t *settled, size_t sbytes, uint32_t *changed) {
  const int64_t n = d.n;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  const unsigned jblocks = (unsigned)((2 * n + 255) / 256);
  hipError_t e = hipMemsetAsync(settled, 0, sbytes, c.stream);
  if (e == hipSuccess) {
    k_init_pointers<<<blocks, 256, 0, c.stream>>>(codes, asc, desc, d);
    // rounds in batches of 4; the batch's last round tells whether anything moved
    // (a chain is shorter than n < 2^31, so 32 rounds settle any field)
    bool settled_all = false;
    for (int round = 0; round < 36 && e == hipSuccess && !settled_all; round += 4) {
      for (int k = 0; k < 3; k++)
        k_jump<<<jblocks, 256, 0, c.stream>>>(asc, desc, n, settled, changed);
      e = hipMemsetAsync(changed, 0, sizeof(uint32_t), c.stream);
      k_jump<<<jblocks, 256, 0, c.stream>>>(asc, desc, n, settled, changed);
      uint32_t h = 1;
      if (e == hipSuccess) e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(&h, changed, sizeof(h), hipMemcpyDeviceToHost, c.stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
      settled_all = e == hipSuccess && h == 0;
    }
    if (e == hipSuccess && !settled_all) e = hipErrorUnknown;  // a pointer field with a cycle
  }
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    last_hip_error() = e;
    return PMSZ_EDEVICE;
  }
  return PMSZ_OK;
}

// Bytes of the wave bits k_jump keeps for n voxels (two label fields).
size_t settled_bytes(int64_t n) {
  return sizeof(uint32_t) * (size_t)(((2 * n + 63) / 64 + 31) / 32 + 1);
}

// naive_MSz's correction (c_loops, naive_MSz.cu:1887-2219) on one field, under
// snapshot sweeps.  Per outer iteration: extrema sweeps until no false maximum
// or minimum is left (naive_MSz.cu:2000-2033); the labels of g and one path
// sweep (naive_MSz.cu:2045-2058); a check of extrema and labels
// (naive_MSz.cu:2077-2165; the reference runs it twice around the ghost
// exchange, which one domain does not need).  The host reads the counters after
// every step: the naive schedule is a baseline, not a tuned path.
template <typename T>
int correct_naive(Ctx &c, const pmsz_params *p, const T *f, T *g, uint8_t *d_dc,
                  pmsz_stats *st) {
  std::memset(st, 0, sizeof(*st));
  const int64_t n = p->nx * p->ny * p->nz;
  if (n >= ((int64_t)1 << 31)) return PMSZ_ETOOBIG;
  const Grid3 d = grid_of(p->nx, p->ny, p->nz);
  const size_t nb = ((size_t)n + 255) & ~(size_t)255;
  const size_t sbytes = settled_bytes(n);
  int rc;
  if ((rc = c.code.ensure(n)) || (rc = c.snap.ensure(n)) || (rc = c.dc.ensure(n)) ||
      (rc = c.tbyte.ensure(nb + 256 + sbytes + 256)) || (rc = c.tmark.ensure(n)) ||
      (rc = c.wmark.ensure(n)) || (rc = c.tlist.ensure(n)) || (rc = c.wlist.ensure(n)))
    return rc;
  c.tag_next = 0;  // the map's memory is this mode's scratch
  uint8_t *cf = c.code.p, *cg = c.snap.p, *dc = d_dc ? d_dc : c.dc.p;
  uint8_t *mark = c.tbyte.p;
  unsigned *cnt = (unsigned *)(c.tbyte.p + nb);
  uint32_t *settled = (uint32_t *)(c.tbyte.p + nb + 256);
  uint32_t *changed = (uint32_t *)((uint8_t *)settled + sbytes);
  int32_t *fa = (int32_t *)c.tmark.p, *fd = (int32_t *)c.wmark.p;
  int32_t *ga = (int32_t *)c.tlist.p, *gd = (int32_t *)c.wlist.p;
  const EditParams ep = edit_params(p);
  // grid-stride passes: a few blocks per CU, one counter atomic per block
  const unsigned blocks = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((n + 255) / 256, (int64_t)c.cus * 8));
  unsigned h[8];
  auto read_cnt = [&]() -> int {
    PMSZ_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    PMSZ_HIP(hipStreamSynchronize(c.stream));
    return PMSZ_OK;
  };
  auto clear_cnt = [&]() -> int {
    PMSZ_HIP(hipMemsetAsync(cnt, 0, sizeof(h), c.stream));
    return PMSZ_OK;
  };
  auto trace = [&](uint32_t a, uint32_t b, uint32_t dir, uint32_t targets) {
    if (st->trace_n < 64) {
      st->trace_cnt[st->trace_n][0] = a;
      st->trace_cnt[st->trace_n][1] = b;
      st->trace_cnt[st->trace_n][2] = dir;
      st->trace_targets[st->trace_n] = targets;
      st->trace_n++;
    }
  };

  PMSZ_HIP(hipEventRecord(c.ev0, c.stream));
  PMSZ_HIP(hipMemsetAsync(c.run, 0, sizeof(RunState), c.stream));
  PMSZ_HIP(hipMemsetAsync(dc, 0, (size_t)n, c.stream));
  PMSZ_HIP(hipMemsetAsync(mark, 0, (size_t)n, c.stream));
  if ((rc = launch_codes<T>(c, p, f, cf)) || (rc = launch_codes<T>(c, p, g, cg))) return rc;
  // NaN or sentinel values in f or g (flagged by the labeling pass)
  PMSZ_HIP(hipMemcpyAsync(c.h_run, c.run, sizeof(RunState), hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  if (c.h_run->err) return PMSZ_EINVAL;
  if ((rc = labels_from_codes(c, cf, fa, fd, d, settled, sbytes, changed))) return rc;
  st->dense_passes = 2;

  uint32_t hmax = 1, hmin = 1, hdir = 0;
  int64_t sweeps = 0;
  int status = PMSZ_OK;
  for (;;) {
    st->outer_iters++;
    st->ite += 2;  // naive_MSz.cu:1966, 1979
    while (hmax || hmin) {
      if (sweeps >= p->max_sweeps) {
        status = PMSZ_NOT_CONVERGED;
        break;
      }
      st->ite++;  // naive_MSz.cu:1982
      if ((rc = launch_codes<T>(c, p, g, cg)) || (rc = clear_cnt())) return rc;
      k_naive_extrema<<<blocks, 256, 0, c.stream>>>(cf, cg, d, mark, cnt);
      k_naive_apply<T><<<blocks, 256, 0, c.stream>>>(f, g, dc, mark, n, ep, cnt);
      PMSZ_HIP(hipGetLastError());
      if ((rc = read_cnt())) return rc;
      hmax = h[kNvMax];
      hmin = h[kNvMin];
      trace(hmax, hmin, 0, h[kNvTargets]);
      st->total_targets += h[kNvTargets];
      st->edit_sweeps++;
      st->dense_passes++;
      sweeps++;
    }
    if (status) break;
    if (sweeps >= p->max_sweeps) {
      status = PMSZ_NOT_CONVERGED;
      break;
    }
    // g is unchanged since cg was taken: the last extrema sweep found nothing,
    // or the previous check classified this g
    if ((rc = labels_from_codes(c, cg, ga, gd, d, settled, sbytes, changed))) return rc;
    if ((rc = clear_cnt())) return rc;
    k_naive_paths<<<blocks, 256, 0, c.stream>>>(cf, cg, fa, fd, ga, gd, d, mark, cnt);
    k_naive_apply<T><<<blocks, 256, 0, c.stream>>>(f, g, dc, mark, n, ep, cnt);
    PMSZ_HIP(hipGetLastError());
    if ((rc = read_cnt())) return rc;
    trace(0, 0, h[kNvDir], h[kNvTargets]);
    st->total_targets += h[kNvTargets];
    st->edit_sweeps++;
    sweeps++;
    // the check
    if ((rc = launch_codes<T>(c, p, g, cg)) || (rc = clear_cnt())) return rc;
    k_naive_extrema<<<blocks, 256, 0, c.stream>>>(cf, cg, d, nullptr, cnt);
    PMSZ_HIP(hipGetLastError());
    if ((rc = labels_from_codes(c, cg, ga, gd, d, settled, sbytes, changed))) return rc;
    k_naive_paths<<<blocks, 256, 0, c.stream>>>(cf, cg, fa, fd, ga, gd, d, nullptr, cnt);
    PMSZ_HIP(hipGetLastError());
    if ((rc = read_cnt())) return rc;
    hmax = h[kNvMax];
    hmin = h[kNvMin];
    hdir = h[kNvDir];
    st->check_sweeps += 2;
    st->dense_passes++;
    st->final_cnt[0] = hmax;
    st->final_cnt[1] = hmin;
    st->final_cnt[2] = hdir;
    if (!hmax && !hmin && !hdir) break;
  }
  if ((rc = clear_cnt())) return rc;
  k_naive_dc_stats<<<blocks, 256, 0, c.stream>>>(dc, n, p->q, cnt);
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipEventRecord(c.ev1, c.stream));
  if ((rc = read_cnt())) return rc;
  float ms = 0;
  PMSZ_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  st->t_total_s = ms * 1e-3;
  st->touched_voxels = h[kNvTouched];
  st->clamped_voxels = h[kNvClamped];
  st->status = status;
  return status;
}

// Entropy-coded size of one rank block's residual stream: every residual (the
// escapes as one more symbol) at -log2 of its frequency, plus 32 bits per escaped
// residual and the value of every verbatim voxel.
int64_t dq_block_bytes(const unsigned long long *h, size_t value_bytes) {
  double total = 0, bits = 0;
  for (int s = 0; s <= kDqBins; s++) total += (double)h[s];
  for (int s = 0; s <= kDqBins; s++)
    if (h[s]) bits += (double)h[s] * std::log2(total / (double)h[s]);
  bits += 32.0 * (double)h[kDqBins] + 8.0 * (double)value_bytes * (double)h[kDqBins + 1];
  return (int64_t)std::ceil(bits / 8.0);
}

// The rank blocks of the codec in rank order (row-major over {PX, PY, PZ}: z
// fastest), each with its first histogram workgroup.
void dq_blocks(const pmsz_params *p, std::vector<DqBlock> &bl, uint32_t &chunks) {
  const int px = p->sub[0], py = p->sub[1], pz = p->sub[2];
  const uint32_t nb = (uint32_t)(px * py * pz);
  bl.clear();
  chunks = 0;
  for (uint32_t r = 0; r < nb; r++) {
    const int cx = (int)r / (py * pz), cy = ((int)r / pz) % py, cz = (int)r % pz;
    int64_t l[3], o[3];
    block_split(p->nx, px, cx, l[0], o[0]);
    block_split(p->ny, py, cy, l[1], o[1]);
    block_split(p->nz, pz, cz, l[2], o[2]);
    DqBlock B;
    B.x0 = (int32_t)o[0];
    B.y0 = (int32_t)o[1];
    B.z0 = (int32_t)o[2];
    B.lx = (int32_t)l[0];
    B.ly = (int32_t)l[1];
    B.lz = (int32_t)l[2];
    B.chunk0 = chunks;
    B.pad = 0;
    chunks += (uint32_t)((l[1] * l[2] + kDqRows - 1) / kDqRows);
    bl.push_back(B);
  }
}

// Quantisation (g0, and the quanta 2k + verbatim in c.tmark) and the residual
// histograms of every rank block (h: blocks x kDqSlots, host).
template <typename T>
int dq_quant_hist(Ctx &c, const pmsz_params *p, const T *f, T *g0,
                  std::vector<unsigned long long> &h) {
  const int64_t n = p->nx * p->ny * p->nz;
  std::vector<DqBlock> bl;
  uint32_t chunks = 0;
  dq_blocks(p, bl, chunks);
  const uint32_t nb = (uint32_t)bl.size();
  const size_t o_hist = ((nb * sizeof(DqBlock)) + 255) & ~(size_t)255;
  const size_t hbytes = (size_t)nb * kDqSlots * sizeof(unsigned long long);
  int rc;
  if ((rc = c.tmark.ensure((size_t)n)) || (rc = c.tbyte.ensure(o_hist + hbytes))) return rc;
  c.tag_next = 0;  // the map's memory is the codec's scratch
  DqBlock *d_bl = (DqBlock *)c.tbyte.p;
  unsigned long long *d_hist = (unsigned long long *)(c.tbyte.p + o_hist);
  int32_t *kq = (int32_t *)c.tmark.p;
  PMSZ_HIP(hipMemcpyAsync(d_bl, bl.data(), nb * sizeof(DqBlock), hipMemcpyHostToDevice,
                          c.stream));
  PMSZ_HIP(hipMemsetAsync(d_hist, 0, hbytes, c.stream));
  constexpr int kV = 16 / sizeof(T);
  const bool aligned = ((uintptr_t)f & 15) == 0 && ((uintptr_t)g0 & 15) == 0;  // kq: hipMalloc
  const int64_t nthreads = (n + kV - 1) / kV;
  k_dq_quant<T><<<(unsigned)((nthreads + 255) / 256), 256, 0, c.stream>>>(
      f, g0, kq, n, p->abs_eb, 1.0 / (2.0 * p->abs_eb), aligned);
  PMSZ_HIP(hipGetLastError());
  bool quads = p->nx % 4 == 0;
  for (const DqBlock &B : bl) quads = quads && B.x0 % 4 == 0 && B.lx % 4 == 0;
  if (quads)
    k_dq_hist4<<<chunks, 256, 0, c.stream>>>(kq, d_bl, nb, p->nx, p->nx * p->ny, d_hist);
  else
    k_dq_hist<<<chunks, 256, 0, c.stream>>>(kq, d_bl, nb, p->nx, p->nx * p->ny, d_hist);
  PMSZ_HIP(hipGetLastError());
  h.assign((size_t)nb * kDqSlots, 0);
  PMSZ_HIP(hipMemcpyAsync(h.data(), d_hist, hbytes, hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  return PMSZ_OK;
}

void dq_info(const std::vector<unsigned long long> &h, size_t value_bytes,
             pmsz_codec_info *info) {
  std::memset(info, 0, sizeof(*info));
  for (size_t b = 0; b * kDqSlots < h.size(); b++) {
    const unsigned long long *hb = h.data() + b * kDqSlots;
    info->compressed_bytes += dq_block_bytes(hb, value_bytes);
    info->escapes += (int64_t)hb[kDqBins];
    info->verbatim += (int64_t)hb[kDqBins + 1];
  }
}

template <typename T>
int codec_dq(Ctx &c, const pmsz_params *p, const T *f, T *g0, uint64_t *hist_out,
             pmsz_codec_info *info) {
  std::vector<unsigned long long> h;
  int rc = dq_quant_hist<T>(c, p, f, g0, h);
  if (rc) return rc;
  if (hist_out) std::memcpy(hist_out, h.data(), h.size() * sizeof(unsigned long long));
  if (info) dq_info(h, sizeof(T), info);
  return PMSZ_OK;
}

// ---------------------------------------------------------------------------
// The compressed stream (pmsz_dqstream.h).

// Huffman code lengths for the counts of ns symbols (0: unused symbol), at most
// kDqsMaxLen bits.  Leaves in (count, symbol) order; two queues, the leaf taken
// on equal weights; a code longer than the limit halves every count (rounding
// up) and starts over.  A single used symbol gets length 1.
void dqs_lengths(const unsigned long long *counts, int ns, uint8_t *len) {
  std::vector<unsigned long long> cnt(counts, counts + ns);
  for (;;) {
    std::memset(len, 0, (size_t)ns);
    std::vector<int> sym;
    for (int s = 0; s < ns; s++)
      if (cnt[s]) sym.push_back(s);
    if (sym.empty()) return;
    if (sym.size() == 1) {
      len[sym[0]] = 1;
      return;
    }
    std::stable_sort(sym.begin(), sym.end(), [&](int a, int b) { return cnt[a] < cnt[b]; });
    const int m = (int)sym.size();
    std::vector<unsigned long long> w(2 * m - 1);
    std::vector<int> parent(2 * m - 1, -1), depth(2 * m - 1, 0);
    for (int k = 0; k < m; k++) w[k] = cnt[sym[k]];
    int i = 0, j = m;
    for (int k = m; k < 2 * m - 1; k++) {
      int pick[2];
      for (int t = 0; t < 2; t++) {
        if (i < m && (j >= k || w[i] <= w[j])) pick[t] = i++;
        else pick[t] = j++;
      }
      w[k] = w[pick[0]] + w[pick[1]];
      parent[pick[0]] = parent[pick[1]] = k;
    }
    int maxd = 0;
    for (int k = 2 * m - 3; k >= 0; k--) {
      depth[k] = depth[parent[k]] + 1;
      if (k < m) maxd = std::max(maxd, depth[k]);
    }
    if (maxd <= kDqsMaxLen) {
      for (int k = 0; k < m; k++) len[sym[k]] = (uint8_t)depth[k];
      return;
    }
    for (int s = 0; s < ns; s++)
      if (cnt[s]) cnt[s] = (cnt[s] + 1) / 2;
  }
}

// Canonical codes of the lengths (DEFLATE's assignment: codes of one length are
// consecutive in symbol order), as entries (code reversed over its length) << 5
// | length; 0 for an unused symbol.
void dqs_codes(const uint8_t *len, int ns, uint32_t *entry) {
  uint32_t bl[kDqsMaxLen + 2] = {}, next[kDqsMaxLen + 2] = {};
  for (int s = 0; s < ns; s++)
    if (len[s]) bl[len[s]]++;
  uint32_t code = 0;
  for (int b = 1; b <= kDqsMaxLen; b++) {
    code = (code + bl[b - 1]) << 1;
    next[b] = code;
  }
  for (int s = 0; s < ns; s++) {
    entry[s] = 0;
    if (!len[s]) continue;
    const uint32_t cd = next[len[s]]++;
    uint32_t rev = 0;
    for (int k = 0; k < len[s]; k++) rev |= ((cd >> k) & 1u) << (len[s] - 1 - k);
    entry[s] = (rev << 5) | len[s];
  }
}

// Decoder tables of one block's code lengths (pmsz_dqstream.h).
void dqs_dec_table(const uint8_t *len, int ns, DqsDecTable &t) {
  std::memset(&t, 0, sizeof(t));
  std::vector<uint32_t> entry((size_t)ns);
  dqs_codes(len, ns, entry.data());
  for (int s = 0; s < ns; s++)
    if (len[s]) t.count[len[s]]++;
  uint32_t code = 0, idx = 0;
  for (int b = 1; b <= kDqsMaxLen; b++) {
    code = (code + t.count[b - 1]) << 1;
    t.first[b] = code;
    t.index[b] = idx;
    idx += t.count[b];
  }
  uint32_t fill[kDqsMaxLen + 1] = {};
  for (int s = 0; s < ns; s++) {  // symbols of each length in symbol order
    if (!len[s]) continue;
    t.sorted[t.index[len[s]] + fill[len[s]]++] = (uint16_t)s;
    if (len[s] <= kDqsFastBits) {
      const uint32_t rev = entry[s] >> 5;
      for (uint32_t i = rev; i < (1u << kDqsFastBits); i += 1u << len[s])
        t.fast[i] = (uint32_t)s | ((uint32_t)len[s] << 16);
    }
  }
}

// Byte offsets of the parts of a stream.
struct DqsLayout {
  uint64_t blk, len, cwo, ceo, bits, esc, vidx, vval, end;
};
DqsLayout dqs_layout(uint32_t nb, uint64_t chunks, uint64_t words, uint64_t nesc,
                     uint64_t nverb, size_t vsize) {
  auto up8 = [](uint64_t x) { return (x + 7) & ~(uint64_t)7; };
  DqsLayout L;
  L.blk = sizeof(DqsHeader);
  L.len = L.blk + (uint64_t)nb * sizeof(DqsBlock);
  L.cwo = up8(L.len + (uint64_t)nb * kDqsSymbols);
  L.ceo = L.cwo + 8 * (chunks + 1);
  L.bits = L.ceo + 8 * chunks;
  L.esc = L.bits + 4 * words;
  L.vidx = L.esc + 4 * nesc;
  L.vval = up8(L.vidx + 4 * nverb);
  L.end = L.vval + vsize * nverb;
  return L;
}

// Rank blocks of the stream with their chunks (and decoder workgroups).
void dqs_geo(const pmsz_params *p, std::vector<DqsGeo> &geo, uint32_t &chunks, uint32_t &wgs) {
  std::vector<DqBlock> bl;
  uint32_t hc = 0;
  dq_blocks(p, bl, hc);
  geo.clear();
  chunks = 0;
  wgs = 0;
  for (const DqBlock &B : bl) {
    DqsGeo G;
    std::memset(&G, 0, sizeof(G));
    G.x0 = B.x0;
    G.y0 = B.y0;
    G.z0 = B.z0;
    G.lx = B.lx;
    G.ly = B.ly;
    G.lz = B.lz;
    G.voxels = (uint64_t)B.lx * B.ly * B.lz;
    G.first_chunk = chunks;
    G.nchunks = (uint32_t)((G.voxels + kDqsChunk - 1) / kDqsChunk);
    G.wg0 = wgs;
    chunks += G.nchunks;
    wgs += (G.nchunks + kDqsDecThreads - 1) / kDqsDecThreads;
    geo.push_back(G);
  }
}

uint64_t dqs_bound(const pmsz_params *p) {
  std::vector<DqsGeo> geo;
  uint32_t chunks, wgs;
  dqs_geo(p, geo, chunks, wgs);
  const uint64_t n = (uint64_t)(p->nx * p->ny * p->nz);
  return dqs_layout((uint32_t)geo.size(), chunks, (uint64_t)chunks * kDqsChunkWords, n, n,
                    p->dtype == PMSZ_F32 ? 4 : 8)
      .end;
}

template <typename T>
int codec_dq_compress(Ctx &c, const pmsz_params *p, const T *f, T *g0, uint8_t *stream,
                      int64_t capacity, int64_t *stream_bytes, pmsz_codec_info *info) {
  std::vector<unsigned long long> h;
  int rc = dq_quant_hist<T>(c, p, f, g0, h);
  if (rc) return rc;
  if (info) dq_info(h, sizeof(T), info);
  std::vector<DqsGeo> geo;
  uint32_t chunks, wgs;
  dqs_geo(p, geo, chunks, wgs);
  const uint32_t nb = (uint32_t)geo.size();
  // code lengths and codes of every block (the escape symbol's count is the
  // histogram's escape slot)
  std::vector<uint8_t> lens((size_t)nb * kDqsSymbols);
  std::vector<uint32_t> codes((size_t)nb * kDqsSymbols);
  uint64_t nesc = 0, nverb = 0;
  for (uint32_t b = 0; b < nb; b++) {
    const unsigned long long *hb = h.data() + (size_t)b * kDqSlots;
    dqs_lengths(hb, kDqsSymbols, &lens[(size_t)b * kDqsSymbols]);
    dqs_codes(&lens[(size_t)b * kDqsSymbols], kDqsSymbols, &codes[(size_t)b * kDqsSymbols]);
    geo[b].esc_offset = nesc;
    geo[b].verb_offset = nverb;
    nesc += hb[kDqBins];
    nverb += hb[kDqBins + 1];
  }
  // scratch (c.wlist): geo | codes | chunk bits, escapes, verbatim | verbatim
  // offsets | totals
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_codes = up(nb * sizeof(DqsGeo));
  const size_t o_cb = o_codes + up((size_t)nb * kDqsSymbols * 4);
  const size_t o_ce = o_cb + up((size_t)chunks * 4);
  const size_t o_cv = o_ce + up((size_t)chunks * 4);
  const size_t o_cvo = o_cv + up((size_t)chunks * 4);
  const size_t o_tot = o_cvo + up((size_t)chunks * 8);
  const size_t sbytes = o_tot + 256;
  if ((rc = c.wlist.ensure((sbytes + 3) / 4))) return rc;
  uint8_t *sc = (uint8_t *)c.wlist.p;
  DqsGeo *d_geo = (DqsGeo *)sc;
  uint32_t *d_codes = (uint32_t *)(sc + o_codes);
  uint32_t *d_cb = (uint32_t *)(sc + o_cb), *d_ce = (uint32_t *)(sc + o_ce),
           *d_cv = (uint32_t *)(sc + o_cv);
  uint64_t *d_cvo = (uint64_t *)(sc + o_cvo), *d_tot = (uint64_t *)(sc + o_tot);
  PMSZ_HIP(hipMemcpyAsync(d_geo, geo.data(), nb * sizeof(DqsGeo), hipMemcpyHostToDevice,
                          c.stream));
  PMSZ_HIP(hipMemcpyAsync(d_codes, codes.data(), codes.size() * 4, hipMemcpyHostToDevice,
                          c.stream));
  const DqsLayout L0 = dqs_layout(nb, chunks, 0, 0, 0, sizeof(T));
  if ((int64_t)L0.bits > capacity) {
    *stream_bytes = (int64_t)L0.bits;
    PMSZ_HIP(hipStreamSynchronize(c.stream));
    return PMSZ_ETOOBIG;
  }
  const int32_t *kq = (const int32_t *)c.tmark.p;
  const int64_t nx = p->nx, sxy = p->nx * p->ny;
  k_dqs_count<<<chunks, kDqsThreads, 0, c.stream>>>(kq, d_geo, nb, nx, sxy, d_codes, d_cb, d_ce,
                                                   d_cv);
  PMSZ_HIP(hipGetLastError());
  uint64_t *cwoff = (uint64_t *)(stream + L0.cwo), *ceoff = (uint64_t *)(stream + L0.ceo);
  k_dqs_scan<<<1, 1024, 0, c.stream>>>(d_cb, d_ce, d_cv, chunks, cwoff, ceoff, d_cvo, d_tot);
  PMSZ_HIP(hipGetLastError());
  uint64_t words = 0, tot[2] = {0, 0};
  PMSZ_HIP(hipMemcpyAsync(&words, cwoff + chunks, 8, hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipMemcpyAsync(tot, d_tot, 16, hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  if (tot[0] != nesc || tot[1] != nverb) return PMSZ_EDEVICE;  // histogram and count disagree
  const DqsLayout L = dqs_layout(nb, chunks, words, nesc, nverb, sizeof(T));
  *stream_bytes = (int64_t)L.end;
  if ((int64_t)L.end > capacity) return PMSZ_ETOOBIG;
  // header, block records and code lengths
  std::vector<uint8_t> head(L.cwo, 0);
  DqsHeader H;
  std::memset(&H, 0, sizeof(H));
  H.magic = kDqsMagic;
  H.dtype = (uint32_t)p->dtype;
  H.nx = p->nx;
  H.ny = p->ny;
  H.nz = p->nz;
  for (int a = 0; a < 3; a++) H.sub[a] = p->sub[a];
  H.nblocks = nb;
  H.abs_eb = p->abs_eb;
  H.chunk = kDqsChunk;
  std::memcpy(head.data(), &H, sizeof(H));
  for (uint32_t b = 0; b < nb; b++) {
    DqsBlock R;
    std::memset(&R, 0, sizeof(R));
    const unsigned long long *hb = h.data() + (size_t)b * kDqSlots;
    R.voxels = geo[b].voxels;
    R.escapes = hb[kDqBins];
    R.verbatim = hb[kDqBins + 1];
    R.esc_offset = geo[b].esc_offset;
    R.verb_offset = geo[b].verb_offset;
    R.first_chunk = geo[b].first_chunk;
    R.nchunks = geo[b].nchunks;
    std::memcpy(head.data() + L.blk + (size_t)b * sizeof(DqsBlock), &R, sizeof(R));
  }
  std::memcpy(head.data() + L.len, lens.data(), lens.size());
  PMSZ_HIP(hipMemcpyAsync(stream, head.data(), head.size(), hipMemcpyHostToDevice, c.stream));
  k_dqs_encode<T><<<chunks, kDqsThreads, 0, c.stream>>>(
      kq, g0, d_geo, nb, nx, sxy, d_codes, cwoff, ceoff, d_cvo, (uint32_t *)(stream + L.bits),
      (uint32_t *)(stream + L.esc), (uint32_t *)(stream + L.vidx), (T *)(stream + L.vval));
  PMSZ_HIP(hipGetLastError());
  if (L.vval > L.vidx + 4 * nverb)  // the padding word
    PMSZ_HIP(hipMemsetAsync(stream + L.vidx + 4 * nverb, 0, L.vval - (L.vidx + 4 * nverb),
                            c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  return PMSZ_OK;
}

// The header of a stream in device memory, checked against its size.
int dqs_read_head(const uint8_t *stream, int64_t bytes, DqsHeader &H, std::vector<uint8_t> &head,
                  hipStream_t st) {
  if (bytes < (int64_t)sizeof(DqsHeader)) return PMSZ_EINVAL;
  PMSZ_HIP(hipMemcpyAsync(&H, stream, sizeof(H), hipMemcpyDeviceToHost, st));
  PMSZ_HIP(hipStreamSynchronize(st));
  if (H.magic != kDqsMagic || H.chunk != (uint32_t)kDqsChunk || H.dtype > 1) return PMSZ_EINVAL;
  if (H.nblocks < 1 || H.nblocks > kMaxDomains) return PMSZ_EINVAL;
  const uint64_t need = sizeof(DqsHeader) + (uint64_t)H.nblocks * (sizeof(DqsBlock) + kDqsSymbols);
  if ((uint64_t)bytes < need) return PMSZ_EINVAL;
  head.resize(need);
  PMSZ_HIP(hipMemcpyAsync(head.data(), stream, need, hipMemcpyDeviceToHost, st));
  PMSZ_HIP(hipStreamSynchronize(st));
  return PMSZ_OK;
}

pmsz_params dqs_params(const DqsHeader &H) {
  pmsz_params q;
  pmsz_default_params(&q);
  q.dtype = (int)H.dtype;
  q.nx = H.nx;
  q.ny = H.ny;
  q.nz = H.nz;
  for (int a = 0; a < 3; a++) q.sub[a] = H.sub[a];
  q.abs_eb = H.abs_eb;
  return q;
}

template <typename T>
int codec_dq_decompress(Ctx &c, const uint8_t *stream, int64_t bytes, const DqsHeader &H,
                        const std::vector<uint8_t> &head, T *g0) {
  const pmsz_params q = dqs_params(H);
  int rc = check_params(&q);
  if (rc) return rc;
  std::vector<DqsGeo> geo;
  uint32_t chunks, wgs;
  dqs_geo(&q, geo, chunks, wgs);
  const uint32_t nb = (uint32_t)geo.size();
  if (nb != H.nblocks) return PMSZ_EINVAL;
  uint64_t nesc = 0, nverb = 0;
  std::vector<DqsDecTable> tabs(nb);
  for (uint32_t b = 0; b < nb; b++) {
    DqsBlock R;
    std::memcpy(&R, head.data() + sizeof(DqsHeader) + (size_t)b * sizeof(DqsBlock), sizeof(R));
    if (R.voxels != geo[b].voxels || R.first_chunk != geo[b].first_chunk ||
        R.nchunks != geo[b].nchunks || R.esc_offset != nesc || R.verb_offset != nverb)
      return PMSZ_EINVAL;
    geo[b].esc_offset = nesc;
    geo[b].verb_offset = nverb;
    nesc += R.escapes;
    nverb += R.verbatim;
    const uint8_t *len = head.data() + sizeof(DqsHeader) + (size_t)nb * sizeof(DqsBlock) +
                         (size_t)b * kDqsSymbols;
    for (int s = 0; s < kDqsSymbols; s++)
      if (len[s] > kDqsMaxLen) return PMSZ_EINVAL;
    dqs_dec_table(len, kDqsSymbols, tabs[b]);
  }
  // the chunk offsets, checked before any kernel reads through them
  const DqsLayout L0 = dqs_layout(nb, chunks, 0, 0, 0, sizeof(T));
  if ((uint64_t)bytes < L0.bits) return PMSZ_EINVAL;
  std::vector<uint64_t> cwo((size_t)chunks + 1), ceo((size_t)chunks);
  PMSZ_HIP(hipMemcpyAsync(cwo.data(), stream + L0.cwo, 8 * (chunks + 1), hipMemcpyDeviceToHost,
                          c.stream));
  PMSZ_HIP(hipMemcpyAsync(ceo.data(), stream + L0.ceo, 8 * (size_t)chunks, hipMemcpyDeviceToHost,
                          c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  if (cwo[0] != 0) return PMSZ_EINVAL;
  for (uint32_t k = 0; k < chunks; k++)
    if (cwo[k + 1] < cwo[k] || cwo[k + 1] - cwo[k] > (uint64_t)kDqsChunkWords ||
        ceo[k] > nesc || (k && ceo[k] < ceo[k - 1]))
      return PMSZ_EINVAL;
  const uint64_t words = cwo[chunks];
  const DqsLayout L = dqs_layout(nb, chunks, words, nesc, nverb, sizeof(T));
  if ((uint64_t)bytes < L.end) return PMSZ_EINVAL;
  // scratch: the residuals / quanta in c.tmark, tables and geo in c.wlist
  const int64_t n = q.nx * q.ny * q.nz;
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_tab = up(nb * sizeof(DqsGeo));
  const size_t o_err = o_tab + up(nb * sizeof(DqsDecTable));
  if ((rc = c.tmark.ensure((size_t)n)) || (rc = c.wlist.ensure((o_err + 256) / 4))) return rc;
  uint8_t *sc = (uint8_t *)c.wlist.p;
  DqsGeo *d_geo = (DqsGeo *)sc;
  DqsDecTable *d_tab = (DqsDecTable *)(sc + o_tab);
  uint32_t *d_err = (uint32_t *)(sc + o_err);
  uint32_t *kd = c.tmark.p;
  PMSZ_HIP(hipMemcpyAsync(d_geo, geo.data(), nb * sizeof(DqsGeo), hipMemcpyHostToDevice,
                          c.stream));
  PMSZ_HIP(hipMemcpyAsync(d_tab, tabs.data(), nb * sizeof(DqsDecTable), hipMemcpyHostToDevice,
                          c.stream));
  PMSZ_HIP(hipMemsetAsync(d_err, 0, 4, c.stream));
  const int64_t nx = q.nx, sxy = q.nx * q.ny;
  k_dqs_decode<<<wgs, kDqsDecThreads, 0, c.stream>>>(
      (const uint32_t *)(stream + L.bits), (const uint64_t *)(stream + L.cwo),
      (const uint64_t *)(stream + L.ceo), (const uint32_t *)(stream + L.esc), nesc, d_geo, nb,
      nx, sxy, d_tab, kd, d_err);
  PMSZ_HIP(hipGetLastError());
  for (uint32_t b = 0; b < nb; b++) {
    const uint64_t rows = (uint64_t)geo[b].ly * geo[b].lz;
    k_dqs_prefix_x<<<(unsigned)((rows + 3) / 4), 256, 0, c.stream>>>(kd, d_geo, b, nx, sxy);
    const uint64_t ly = (uint64_t)geo[b].lx * geo[b].lz, lz = (uint64_t)geo[b].lx * geo[b].ly;
    k_dqs_prefix_yz<<<(unsigned)((ly + 255) / 256), 256, 0, c.stream>>>(kd, d_geo, b, nx, sxy, 1);
    k_dqs_prefix_yz<<<(unsigned)((lz + 255) / 256), 256, 0, c.stream>>>(kd, d_geo, b, nx, sxy, 2);
  }
  PMSZ_HIP(hipGetLastError());
  const unsigned gv = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)c.cus * 8);
  k_dqs_values<T><<<gv, 256, 0, c.stream>>>(kd, g0, n, 2.0 * q.abs_eb);
  if (nverb)
    k_dqs_verbatim<T><<<(unsigned)std::min<uint64_t>((nverb + 255) / 256, 1024), 256, 0,
                        c.stream>>>((const uint32_t *)(stream + L.vidx),
                                    (const T *)(stream + L.vval), d_geo, nb, nverb, nx, sxy, g0,
                                    d_err);
  PMSZ_HIP(hipGetLastError());
  uint32_t err = 0;
  PMSZ_HIP(hipMemcpyAsync(&err, d_err, 4, hipMemcpyDeviceToHost, c.stream));
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  return err ? PMSZ_EINVAL : PMSZ_OK;
}

// One field (whole or split into subdomains), by value type; the caller holds
// the context's lock.
int correct_dispatch(Ctx &c, const pmsz_params *p, const void *d_f, void *d_g, uint8_t *d_dc,
                     pmsz_stats *stats) {
  const bool split = p->sub[0] * p->sub[1] * p->sub[2] != 1;
  if (p->mode == PMSZ_NAIVE) {
    if (split) return PMSZ_EINVAL;
    if (p->dtype == PMSZ_F32)
      return correct_naive<float>(c, p, (const float *)d_f, (float *)d_g, d_dc, stats);
    return correct_naive<double>(c, p, (const double *)d_f, (double *)d_g, d_dc, stats);
  }
  if (p->dtype == PMSZ_F32) {
    if (split)
      return correct_subdomains<float>(c, p, (const float *)d_f, (float *)d_g, d_dc, stats);
    return correct_single<float>(c, p, (const float *)d_f, (float *)d_g, d_dc, stats);
  }
  if (split)
    return correct_subdomains<double>(c, p, (const double *)d_f, (double *)d_g, d_dc, stats);
  return correct_single<double>(c, p, (const double *)d_f, (double *)d_g, d_dc, stats);
}

// bits_for_max (pMSz.cu:2352-2356): ceil(log2(q + 1)), the bit length of q.
int edit_nbits(int q) {
  int b = 0;
  while (b < 31 && (1 << b) <= q) b++;
  return b ? b : 1;
}

// The rank blocks of p in rank order (row-major over {PX, PY, PZ}: z fastest),
// their tiles and their 16-byte aligned packed streams, into bl (room for
// PX * PY * PZ entries).
void edit_layout_into(const pmsz_params *p, int nbits, EncBlock *bl, int64_t &packed,
                      uint32_t &ntiles) {
  const uint32_t tile = p->dtype == PMSZ_F32 ? enc_tile<float>() : enc_tile<double>();
  packed = 0;
  ntiles = 0;
  const int px = p->sub[0], py = p->sub[1], pz = p->sub[2];
  for (int r = 0; r < px * py * pz; r++) {
    const int cx = r / (py * pz), cy = (r / pz) % py, cz = r % pz;
    int64_t l[3], o[3];
    block_split(p->nx, px, cx, l[0], o[0]);
    block_split(p->ny, py, cy, l[1], o[1]);
    block_split(p->nz, pz, cz, l[2], o[2]);
    EncBlock &B = bl[r];
    B.lx = (uint32_t)l[0];
    B.ly = (uint32_t)l[1];
    B.n = (uint32_t)(l[0] * l[1] * l[2]);
    B.tile0 = ntiles;
    B.base = (o[2] * p->ny + o[1]) * p->nx + o[0];
    B.pk_off = packed;
    B.pk_bytes = ((int64_t)B.n * nbits + 7) / 8;
    packed += (B.pk_bytes + 15) & ~(int64_t)15;
    ntiles += (B.n + tile - 1) / tile;
  }
}
void edit_layout(const pmsz_params *p, int nbits, std::vector<EncBlock> &bl, int64_t &packed,
                 uint32_t &ntiles) {
  bl.resize((size_t)(p->sub[0] * p->sub[1] * p->sub[2]));
  edit_layout_into(p, nbits, bl.data(), packed, ntiles);
}

// Scratch of the encoder in one device buffer: blocks | look-back words |
// per-tile edited counts | error flags | totals (each part 256-byte aligned).
struct EncScratch {
  size_t o_look, o_edit, o_tick, o_tot, total;
};
EncScratch enc_scratch(uint32_t nb, uint32_t ntiles) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  EncScratch e;
  e.o_look = up(nb * sizeof(EncBlock));
  e.o_edit = e.o_look + up((size_t)ntiles * 8);
  e.o_tick = e.o_edit + up((size_t)ntiles * 4);
  e.o_tot = e.o_tick + 256;
  e.total = e.o_tot + up((size_t)(nb + 2) * 8);
  return e;
}

// The encoder's launches on stream st with scratch s (enc_scratch bytes, 256-byte
// aligned): the block table is copied in from bl, the rest of the scratch is
// cleared by a fill kernel, and the totals (edited, values per block, look-back
// failures: nb + 2 words) and the error flags are copied back to h_tot / h_err.
// Returns when they are there.
template <typename T>
int encode_edits_on(hipStream_t st, uint8_t *s, const pmsz_params *p, const EncBlock *bl,
                    uint32_t nb, uint32_t ntiles, int nbits, const T *f, const T *g0, const T *g,
                    uint8_t *packed, T *vals, unsigned long long *h_tot, uint32_t *h_err) {
  const EncScratch sc = enc_scratch(nb, ntiles);
  EncBlock *d_bl = (EncBlock *)s;
  unsigned long long *look = (unsigned long long *)(s + sc.o_look);
  uint32_t *tile_edited = (uint32_t *)(s + sc.o_edit);
  uint32_t *err = (uint32_t *)(s + sc.o_tick);
  unsigned long long *tot = (unsigned long long *)(s + sc.o_tot);
  PMSZ_HIP(hipMemcpyAsync(d_bl, bl, nb * sizeof(EncBlock), hipMemcpyHostToDevice, st));
  const uint64_t n16 = (sc.total - sc.o_look) / 16;
  k_fill16<<<(unsigned)std::min<uint64_t>((n16 + 255) / 256, 1024), 256, 0, st>>>(
      (uint4 *)(s + sc.o_look), n16);
  PMSZ_HIP(hipGetLastError());
  const EditParams ep = edit_params(p);
  const bool aligned = ((uintptr_t)f & 15) == 0 && ((uintptr_t)g0 & 15) == 0 &&
                       ((uintptr_t)g & 15) == 0;
  const bool flat = nb == 1 && aligned;
  if (flat)
    k_edit_encode<T, true><<<ntiles, kEncThreads, 0, st>>>(
        f, g0, g, d_bl, nb, p->nx, p->nx * p->ny, nbits, aligned, ep, packed, vals, look,
        tile_edited, err);
  else
    k_edit_encode<T, false><<<ntiles, kEncThreads, 0, st>>>(
        f, g0, g, d_bl, nb, p->nx, p->nx * p->ny, nbits, aligned, ep, packed, vals, look,
        tile_edited, err);
  PMSZ_HIP(hipGetLastError());
  k_edit_totals<<<std::min<uint32_t>((ntiles + 255) / 256, 128u), 256, 0, st>>>(
      d_bl, nb, ntiles, look, tile_edited, tot);
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipMemcpyAsync(h_tot, tot, (nb + 2) * 8, hipMemcpyDeviceToHost, st));
  PMSZ_HIP(hipMemcpyAsync(h_err, err, 4, hipMemcpyDeviceToHost, st));
  PMSZ_HIP(hipStreamSynchronize(st));
  return PMSZ_OK;
}

// The caller's outputs from the totals of encode_edits_on.
int encode_edits_report(const EncBlock *bl, uint32_t nb, int nbits, int64_t packed_bytes,
                        const unsigned long long *h, uint32_t herr, pmsz_edit_block *blocks,
                        pmsz_edit_info *info) {
  if (herr & kEncErrLook || h[nb + 1]) return PMSZ_EDEVICE;
  if (herr & kEncErrSymbol) return PMSZ_EINVAL;
  int64_t voff = 0;
  for (uint32_t b = 0; b < nb; b++) {
    if (blocks) {
      blocks[b].voxels = bl[b].n;
      blocks[b].packed_offset = bl[b].pk_off;
      blocks[b].packed_bytes = bl[b].pk_bytes;
      blocks[b].value_offset = voff;
      blocks[b].values = (int64_t)h[1 + b];
    }
    voff += (int64_t)h[1 + b];
  }
  if (info) {
    info->nbits = nbits;
    info->blocks = (int32_t)nb;
    info->edited = (int64_t)h[0];
    info->values = voff;
    info->packed_capacity = packed_bytes;
  }
  return PMSZ_OK;
}

template <typename T>
int encode_edits(Ctx &c, const pmsz_params *p, const T *f, const T *g0, const T *g,
                 uint8_t *packed, T *vals, pmsz_edit_block *blocks, pmsz_edit_info *info) {
  const int nbits = edit_nbits(p->q);
  std::vector<EncBlock> bl;
  int64_t packed_bytes;
  uint32_t ntiles;
  edit_layout(p, nbits, bl, packed_bytes, ntiles);
  const uint32_t nb = (uint32_t)bl.size();
  // the scratch lives in the context's byte map
  int rc = c.tbyte.ensure(enc_scratch(nb, ntiles).total);
  if (rc) return rc;
  c.tag_next = 0;  // the map's memory is the encoder's scratch
  std::vector<unsigned long long> h(nb + 2);
  uint32_t herr = 0;
  rc = encode_edits_on<T>(c.stream, c.tbyte.p, p, bl.data(), nb, ntiles, nbits, f, g0, g, packed,
                          vals, h.data(), &herr);
  if (rc) return rc;
  return encode_edits_report(bl.data(), nb, nbits, packed_bytes, h.data(), herr, blocks, info);
}

// ---------------------------------------------------------------------------
// The encoder without allocation (pmsz_encode_edits_noalloc).  Its scratch is
// this file-scope device array: each device that runs the library holds its own
// copy until the process ends.  Host code never passes `scratch` itself (in host
// code the name is a host-side stand-in, not a device address): the device
// address comes from hipGetSymbolAddress, once per device, with that device
// current, and is kept in g_scratch_ptr.  Every call on a device shares that
// device's array and leaves its contents behind; the parts that must start at
// zero are cleared by each call, and calls that overlap on one device race.
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;  // 4 MiB
constexpr int kNoAllocBlocks = 256;                        // most rank blocks per call
constexpr int kNoAllocDevices = 64;
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocDevices];

// Device address of the current device's scratch array (looked up on the
// device's first call).
int scratch_on_device(uint8_t **out) {
  int dev = -1;
  PMSZ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) return PMSZ_EDEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      last_hip_error() = e;
      return PMSZ_EDEVICE;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = (uint8_t *)g_scratch_ptr[dev];
  return PMSZ_OK;
}

template <typename T>
int encode_edits_noalloc(hipStream_t st, const pmsz_params *p, const T *f, const T *g0,
                         const T *g, uint8_t *packed, T *vals, pmsz_edit_block *blocks,
                         pmsz_edit_info *info) {
  const int nbits = edit_nbits(p->q);
  const uint32_t nb = (uint32_t)(p->sub[0] * p->sub[1] * p->sub[2]);
  if (nb > (uint32_t)kNoAllocBlocks) return PMSZ_ETOOBIG;
  EncBlock bl[kNoAllocBlocks];
  int64_t packed_bytes;
  uint32_t ntiles;
  edit_layout_into(p, nbits, bl, packed_bytes, ntiles);
  if (enc_scratch(nb, ntiles).total > kNoAllocScratchBytes) return PMSZ_ETOOBIG;
  uint8_t *s = nullptr;
  int rc = scratch_on_device(&s);
  if (rc) return rc;
  unsigned long long h[kNoAllocBlocks + 2];
  uint32_t herr = 0;
  rc = encode_edits_on<T>(st, s, p, bl, nb, ntiles, nbits, f, g0, g, packed, vals, h, &herr);
  if (rc) return rc;
  return encode_edits_report(bl, nb, nbits, packed_bytes, h, herr, blocks, info);
}
}  // namespace pmsz

namespace {
pmsz_ctx g_default;
Ctx &default_ctx() { return g_default.c; }
std::atomic<int> g_contexts{1};  // the default context counts
}  // namespace

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
  return pmsz_ctx_correct(&g_default, p, d_f, d_g, d_dc, stats);
}

int pmsz_ctx_correct(pmsz_ctx *ctx, const pmsz_params *p, const void *d_f, void *d_g,
                     uint8_t *d_dc, pmsz_stats *stats) {
  if (!ctx) return PMSZ_EINVAL;
  int rc = check_params(p);
  if (rc) return rc;
  if (!d_f || !d_g) return PMSZ_EINVAL;
  pmsz_stats tmp;
  if (!stats) stats = &tmp;
  Ctx &c = ctx->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if ((rc = ctx_device(c))) return rc;
  return correct_dispatch(c, p, d_f, d_g, d_dc, stats);
}

int pmsz_ctx_create(pmsz_ctx **out) {
  if (!out) return PMSZ_EINVAL;
  if (g_contexts.fetch_add(1) >= kMaxContexts) {
    g_contexts.fetch_sub(1);
    return PMSZ_EBUSY;
  }
  pmsz_ctx *ctx = new (std::nothrow) pmsz_ctx();
  if (!ctx) {
    g_contexts.fetch_sub(1);
    return PMSZ_ENOMEM;
  }
  int rc = ctx_device(ctx->c);
  if (rc) {
    ctx_release(ctx->c);
    delete ctx;
    g_contexts.fetch_sub(1);
    return rc;
  }
  *out = ctx;
  return PMSZ_OK;
}

int pmsz_ctx_destroy(pmsz_ctx *ctx) {
  if (!ctx || ctx == &g_default) return PMSZ_EINVAL;
  {
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
    ctx_release(ctx->c);
  }
  delete ctx;
  g_contexts.fetch_sub(1);
  return PMSZ_OK;
}

int pmsz_ctx_trim(pmsz_ctx *ctx) {
  if (!ctx) return PMSZ_EINVAL;
  Ctx &c = ctx->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.stream) PMSZ_HIP(hipStreamSynchronize(c.stream));
  for (Buf<uint8_t> *b : {&c.code, &c.dc, &c.snap, &c.fpad, &c.gpad, &c.tbyte}) {
    (void)hipFree(b->p);
    *b = Buf<uint8_t>();
  }
  c.tag_next = 0;
  for (Buf<uint32_t> *b : {&c.tmark, &c.wmark, &c.tlist, &c.wlist}) {
    (void)hipFree(b->p);
    *b = Buf<uint32_t>();
  }
  return PMSZ_OK;
}

int64_t pmsz_ctx_bytes(pmsz_ctx *ctx) {
  if (!ctx) return PMSZ_EINVAL;
  std::lock_guard<std::mutex> lk(ctx->c.mu);
  return (int64_t)ctx_bytes(ctx->c);
}

int64_t pmsz_workspace_bytes(int n, const pmsz_params *params) {
  if (n < 1 || !params) return PMSZ_EINVAL;
  uint64_t cells = 0, maxn = 0;
  for (int k = 0; k < n; k++) {
    if (check_params(&params[k]) || params[k].mode == PMSZ_NAIVE) return PMSZ_EINVAL;
    const uint64_t nk = (uint64_t)(params[k].nx * params[k].ny * params[k].nz);
    cells += nk;
    maxn = std::max(maxn, nk);
  }
  const uint64_t ids = n == 1 ? maxn : ((uint64_t)n << std::max<uint32_t>(5, ceil_log2(maxn)));
  // code, dc and target byte per cell, two bitmaps over the ids, two lists of ids
  return (int64_t)(3 * cells + 16 + 2 * 4 * ((ids + 31) / 32 + 1) + 2 * 4 * cells +
                   (sizeof(Dom) + sizeof(DomState)) * (uint64_t)n);
}

int pmsz_correct_batch(int n, const pmsz_params *params, const void *const *d_f,
                       void *const *d_g, uint8_t *const *d_dc, pmsz_stats *stats) {
  return pmsz_ctx_correct_batch(&g_default, n, params, d_f, d_g, d_dc, stats);
}

int pmsz_ctx_correct_batch(pmsz_ctx *ctx, int n, const pmsz_params *params,
                           const void *const *d_f, void *const *d_g, uint8_t *const *d_dc,
                           pmsz_stats *stats) {
  if (!ctx || n < 1 || !params || !d_f || !d_g || !stats) return PMSZ_EINVAL;
  for (int k = 0; k < n; k++) {
    int rc = check_params(&params[k]);
    if (!rc && params[k].mode == PMSZ_NAIVE) rc = PMSZ_EINVAL;
    if (rc) return rc;
    if (!d_f[k] || !d_g[k]) return PMSZ_EINVAL;
    if (params[k].sub[0] * params[k].sub[1] * params[k].sub[2] != 1) return PMSZ_EINVAL;
    if (params[k].dtype != params[0].dtype || params[k].mode != params[0].mode ||
        params[k].max_sweeps != params[0].max_sweeps)
      return PMSZ_EINVAL;
  }
  Ctx &c = ctx->c;
  std::lock_guard<std::mutex> lk(c.mu);
  int rc = ctx_device(c);
  if (rc) return rc;
  if (params[0].dtype == PMSZ_F32) return correct_batch<float>(c, n, params, d_f, d_g, d_dc, stats);
  return correct_batch<double>(c, n, params, d_f, d_g, d_dc, stats);
}

void pmsz_default_tuning(pmsz_tuning *t) {
  if (t) *t = default_tuning();
}

int pmsz_ctx_set_tuning(pmsz_ctx *ctx, const pmsz_tuning *t) {
  if (!t || !tuning_valid(*t)) return PMSZ_EINVAL;
  Ctx &c = (ctx ? ctx : &g_default)->c;
  std::lock_guard<std::mutex> lk(c.mu);
  c.tune = *t;
  return PMSZ_OK;
}

int pmsz_ctx_get_tuning(pmsz_ctx *ctx, pmsz_tuning *t) {
  if (!t) return PMSZ_EINVAL;
  Ctx &c = (ctx ? ctx : &g_default)->c;
  std::lock_guard<std::mutex> lk(c.mu);
  *t = c.tune;
  return PMSZ_OK;
}

int pmsz_autotune(pmsz_ctx *ctx, int n, const pmsz_params *params, const void *const *d_f,
                  const void *const *d_g0, int reps, pmsz_tuning *best, double *best_seconds) {
  if (n < 1 || !params || !d_f || !d_g0 || reps < 1) return PMSZ_EINVAL;
  for (int k = 0; k < n; k++) {
    int rc = check_params(&params[k]);
    if (!rc && params[k].mode == PMSZ_NAIVE) rc = PMSZ_EINVAL;
    if (rc) return rc;
    if (!d_f[k] || !d_g0[k]) return PMSZ_EINVAL;
    if (n > 1 && (params[k].sub[0] * params[k].sub[1] * params[k].sub[2] != 1 ||
                  params[k].dtype != params[0].dtype || params[k].mode != params[0].mode ||
                  params[k].max_sweeps != params[0].max_sweeps))
      return PMSZ_EINVAL;
  }
  Ctx &c = (ctx ? ctx : &g_default)->c;
  std::lock_guard<std::mutex> lk(c.mu);
  int rc = ctx_device(c);
  if (rc) return rc;
  // scratch copies of the decompressed fields: the inputs stay untouched
  const size_t s = params[0].dtype == PMSZ_F64 ? 8 : 4;
  std::vector<size_t> bytes(n), offs(n);
  size_t total = 0;
  for (int k = 0; k < n; k++) {
    bytes[k] = s * (size_t)(params[k].nx * params[k].ny * params[k].nz);
    offs[k] = total;
    total += (bytes[k] + 255) & ~(size_t)255;
  }
  char *scratch = nullptr;
  if (hipMalloc(&scratch, total) != hipSuccess) return PMSZ_ENOMEM;
  std::vector<void *> gs(n);
  for (int k = 0; k < n; k++) gs[k] = scratch + offs[k];
  std::vector<pmsz_stats> st(n);
  const pmsz_tuning start = c.tune;
  // one candidate: the best of `reps` timed runs (host clock around the call,
  // which returns when the result is complete)
  auto measure = [&](const pmsz_tuning &t, double &secs) -> int {
    c.tune = t;
    secs = 1e30;
    for (int r = 0; r < reps; r++) {
      for (int k = 0; k < n; k++)
        PMSZ_HIP(hipMemcpyAsync(gs[k], d_g0[k], bytes[k], hipMemcpyDeviceToDevice, c.stream));
      PMSZ_HIP(hipStreamSynchronize(c.stream));
      const auto t0 = std::chrono::steady_clock::now();
      int e;
      if (n == 1) {
        e = correct_dispatch(c, &params[0], d_f[0], gs[0], nullptr, st.data());
      } else if (params[0].dtype == PMSZ_F32) {
        e = correct_batch<float>(c, n, params, d_f, gs.data(), nullptr, st.data());
      } else {
        e = correct_batch<double>(c, n, params, d_f, gs.data(), nullptr, st.data());
      }
      const auto t1 = std::chrono::steady_clock::now();
      if (e && e != PMSZ_NOT_CONVERGED) return e;
      secs = std::min(secs, std::chrono::duration<double>(t1 - t0).count());
    }
    return PMSZ_OK;
  };
  // coordinate descent over the tile depth and the schedule thresholds, one pass
  pmsz_tuning cur = start;
  double cur_s = 0;
  rc = measure(cur, cur_s);
  const int zc[] = {64, 32, 16};
  const int dd[] = {32, 16, 64, 8, 128};
  const int tt[] = {4096, 16384, 1024, 65536};
  const int tb[] = {32, 64, 16, 128};
  const int so[] = {64, 0, 256, 1024};
  auto sweep = [&](int pmsz_tuning::*field, const int *vals, int nv) {
    for (int v = 0; v < nv && rc == PMSZ_OK; v++) {
      if (cur.*field == vals[v]) continue;
      pmsz_tuning t = cur;
      t.*field = vals[v];
      double secs = 0;
      rc = measure(t, secs);
      if (rc == PMSZ_OK && secs < cur_s) {
        cur = t;
        cur_s = secs;
      }
    }
  };
  if (rc == PMSZ_OK) sweep(&pmsz_tuning::zchunk, zc, 3);
  if (rc == PMSZ_OK) sweep(&pmsz_tuning::dense_div, dd, 5);
  if (rc == PMSZ_OK) sweep(&pmsz_tuning::tail_targets, tt, 4);
  if (rc == PMSZ_OK) sweep(&pmsz_tuning::tail_blocks, tb, 4);
  if (rc == PMSZ_OK) sweep(&pmsz_tuning::solo_targets, so, 4);
  const bool multi = n > 1 || params[0].sub[0] * params[0].sub[1] * params[0].sub[2] != 1;
  const int sd[] = {32, 16, 64, 8, 0};  // the streamed apply: one-domain runs only
  if (rc == PMSZ_OK && !multi) sweep(&pmsz_tuning::stream_div, sd, 5);
  const int gr[] = {0, 1};  // captured chains: multi-domain runs only
  if (rc == PMSZ_OK && multi) sweep(&pmsz_tuning::graphs, gr, 2);
  (void)hipFree(scratch);
  c.tune = rc == PMSZ_OK ? cur : start;
  if (rc) return rc;
  if (best) *best = cur;
  if (best_seconds) *best_seconds = cur_s;
  return PMSZ_OK;
}

int pmsz_time_first_pass(const pmsz_params *p, const void *d_f, const void *d_g, int iters,
                         double *avg_seconds) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!d_f || !d_g || !avg_seconds || iters < 1) return PMSZ_EINVAL;
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  if ((rc = ctx_device(c))) return rc;
  if (p->dtype == PMSZ_F32)
    return time_first_pass<float>(c, p, (const float *)d_f, (const float *)d_g, iters,
                                  avg_seconds);
  return time_first_pass<double>(c, p, (const double *)d_f, (const double *)d_g, iters,
                                 avg_seconds);
}

int pmsz_test_step_budget(int dtype, int64_t n, const void *d_f, const void *d_g,
                          const uint8_t *d_d, double abs_eb, int q, int clamp_mode,
                          uint8_t *d_out) {
  if (n < 0 || (n > 0 && (!d_f || !d_g || !d_d || !d_out)) || q < 2 || q > 15 ||
      !(abs_eb > 0) || (dtype != PMSZ_F32 && dtype != PMSZ_F64) ||
      (clamp_mode != PMSZ_CLAMP_REFERENCE && clamp_mode != PMSZ_CLAMP_BOUND_SAFE))
    return PMSZ_EINVAL;
  if (n == 0) return PMSZ_OK;
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int rc = ctx_device(c);
  if (rc) return rc;
  EditParams ep;
  ep.eb = abs_eb;
  ep.delta_d = abs_eb / q;
  ep.q = q;
  ep.clamp_mode = clamp_mode;
  ep.budgets = 1;
  const unsigned gb = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
  if (dtype == PMSZ_F32)
    k_test_step_budget<float><<<gb, 256, 0, c.stream>>>(n, (const float *)d_f,
                                                        (const float *)d_g, d_d, ep, d_out);
  else
    k_test_step_budget<double><<<gb, 256, 0, c.stream>>>(n, (const double *)d_f,
                                                         (const double *)d_g, d_d, ep, d_out);
  PMSZ_HIP(hipGetLastError());
  PMSZ_HIP(hipStreamSynchronize(c.stream));
  return PMSZ_OK;
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
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  if ((rc = ctx_device(c))) return rc;
  Dom D = field_dom(&p, d_field, nullptr, d_codes, nullptr);
  PMSZ_HIP(hipMemsetAsync(c.run, 0, sizeof(RunState), c.stream));
  if (dtype == PMSZ_F32) {
    static const int occ = occupancy(k_columns<float, true, false, false>, kColThreads);
    k_columns<float, true, false, false><<<col_blocks(c, dom_units(D), occ), kColThreads, 0,
                                           c.stream>>>(D, nullptr, 1, nullptr, (uint8_t *)nullptr, c.ring, 0,
                                                       0, 0u, c.run, 0u);
  } else {
    static const int occ = occupancy(k_columns<double, true, false, false>, kColThreads);
    k_columns<double, true, false, false><<<col_blocks(c, dom_units(D), occ), kColThreads, 0,
                                            c.stream>>>(D, nullptr, 1, nullptr, (uint8_t *)nullptr, c.ring,
                                                        0, 0, 0u, c.run, 0u);
  }
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
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const Grid3 d = grid_of(nx, ny, nz);
  const size_t sbytes = settled_bytes(n);
  uint32_t *settled = nullptr;
  if (hipMalloc(&settled, sbytes + sizeof(uint32_t)) != hipSuccess) {
    (void)hipFree(codes);
    return PMSZ_ENOMEM;
  }
  uint32_t *changed = settled + sbytes / sizeof(uint32_t);
  rc = labels_from_codes(c, codes, d_asc, d_desc, d, settled, sbytes, changed);
  if (rc == PMSZ_OK && hipStreamSynchronize(c.stream) != hipSuccess) rc = PMSZ_EDEVICE;
  (void)hipFree(codes);
  (void)hipFree(settled);
  return rc;
}

int pmsz_edit_symbols(const pmsz_params *p, const void *d_f, const void *d_g0, const void *d_g,
                      uint8_t *d_symbols, int64_t *edited) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!d_f || !d_g0 || !d_g || !d_symbols) return PMSZ_EINVAL;
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  if ((rc = ctx_device(c))) return rc;
  int64_t n = p->nx * p->ny * p->nz;
  const EditParams ep = edit_params(p);
  unsigned long long *dcount = &c.run->acc[0][kAccTargets];
  PMSZ_HIP(hipMemsetAsync(dcount, 0, sizeof(unsigned long long), c.stream));
  const unsigned blocks =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)c.cus * 8));
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

int pmsz_codec_dq(const pmsz_params *p, const void *d_f, void *d_g0, uint64_t *hist,
                  pmsz_codec_info *info) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!d_f || !d_g0 || (uint32_t)(p->sub[0] * p->sub[1] * p->sub[2]) > kMaxDomains)
    return PMSZ_EINVAL;
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  if ((rc = ctx_device(c))) return rc;
  if (p->dtype == PMSZ_F32)
    return codec_dq<float>(c, p, (const float *)d_f, (float *)d_g0, hist, info);
  return codec_dq<double>(c, p, (const double *)d_f, (double *)d_g0, hist, info);
}

int pmsz_codec_dq_stream_bound(const pmsz_params *p, int64_t *bytes) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!bytes || (uint32_t)(p->sub[0] * p->sub[1] * p->sub[2]) > kMaxDomains) return PMSZ_EINVAL;
  *bytes = (int64_t)dqs_bound(p);
  return PMSZ_OK;
}

int pmsz_codec_dq_compress(const pmsz_params *p, const void *d_f, void *d_g0, uint8_t *d_stream,
                           int64_t capacity, int64_t *stream_bytes, pmsz_codec_info *info) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!d_f || !d_g0 || !d_stream || !stream_bytes || ((uintptr_t)d_stream & 15) ||
      (uint32_t)(p->sub[0] * p->sub[1] * p->sub[2]) > kMaxDomains)
    return PMSZ_EINVAL;
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  if ((rc = ctx_device(c))) return rc;
  if (p->dtype == PMSZ_F32)
    return codec_dq_compress<float>(c, p, (const float *)d_f, (float *)d_g0, d_stream, capacity,
                                    stream_bytes, info);
  return codec_dq_compress<double>(c, p, (const double *)d_f, (double *)d_g0, d_stream, capacity,
                                   stream_bytes, info);
}

int pmsz_codec_dq_stream_params(const uint8_t *d_stream, int64_t bytes, pmsz_params *p) {
  if (!d_stream || !p) return PMSZ_EINVAL;
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int rc = ctx_device(c);
  if (rc) return rc;
  DqsHeader H;
  std::vector<uint8_t> head;
  if ((rc = dqs_read_head(d_stream, bytes, H, head, c.stream))) return rc;
  *p = dqs_params(H);
  return check_params(p);
}

int pmsz_codec_dq_decompress(const uint8_t *d_stream, int64_t bytes, void *d_g0) {
  if (!d_stream || !d_g0 || ((uintptr_t)d_stream & 15)) return PMSZ_EINVAL;
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int rc = ctx_device(c);
  if (rc) return rc;
  DqsHeader H;
  std::vector<uint8_t> head;
  if ((rc = dqs_read_head(d_stream, bytes, H, head, c.stream))) return rc;
  if (H.dtype == PMSZ_F32)
    return codec_dq_decompress<float>(c, d_stream, bytes, H, head, (float *)d_g0);
  return codec_dq_decompress<double>(c, d_stream, bytes, H, head, (double *)d_g0);
}

int pmsz_edit_packed_bytes(const pmsz_params *p, int64_t *bytes) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!bytes || (uint32_t)(p->sub[0] * p->sub[1] * p->sub[2]) > kMaxDomains) return PMSZ_EINVAL;
  std::vector<EncBlock> bl;
  uint32_t ntiles;
  edit_layout(p, edit_nbits(p->q), bl, *bytes, ntiles);
  return PMSZ_OK;
}

int pmsz_encode_edits(const pmsz_params *p, const void *d_f, const void *d_g0, const void *d_g,
                      uint8_t *d_packed, void *d_values, pmsz_edit_block *blocks,
                      pmsz_edit_info *info) {
  int rc = check_params(p);
  if (rc) return rc;
  const size_t es = p->dtype == PMSZ_F32 ? 4 : 8;
  if (!d_f || !d_g0 || !d_g || !d_packed || !d_values) return PMSZ_EINVAL;
  if (((uintptr_t)d_packed & 15) || ((uintptr_t)d_values % es)) return PMSZ_EINVAL;
  if ((uint32_t)(p->sub[0] * p->sub[1] * p->sub[2]) > kMaxDomains) return PMSZ_EINVAL;
  Ctx &c = default_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  if ((rc = ctx_device(c))) return rc;
  if (p->dtype == PMSZ_F32)
    return encode_edits<float>(c, p, (const float *)d_f, (const float *)d_g0,
                               (const float *)d_g, d_packed, (float *)d_values, blocks, info);
  return encode_edits<double>(c, p, (const double *)d_f, (const double *)d_g0,
                              (const double *)d_g, d_packed, (double *)d_values, blocks, info);
}

int pmsz_encode_edits_noalloc(const pmsz_params *p, const void *d_f, const void *d_g0,
                              const void *d_g, uint8_t *d_packed, void *d_values,
                              pmsz_edit_block *blocks, pmsz_edit_info *info, void *stream) {
  int rc = check_params(p);
  if (rc) return rc;
  const size_t es = p->dtype == PMSZ_F32 ? 4 : 8;
  if (!d_f || !d_g0 || !d_g || !d_packed || !d_values) return PMSZ_EINVAL;
  if (((uintptr_t)d_packed & 15) || ((uintptr_t)d_values % es)) return PMSZ_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (p->dtype == PMSZ_F32)
    return encode_edits_noalloc<float>(st, p, (const float *)d_f, (const float *)d_g0,
                                       (const float *)d_g, d_packed, (float *)d_values, blocks,
                                       info);
  return encode_edits_noalloc<double>(st, p, (const double *)d_f, (const double *)d_g0,
                                      (const double *)d_g, d_packed, (double *)d_values, blocks,
                                      info);
}

const char *pmsz_strerror(int code) {
  switch (code) {
    case PMSZ_OK: return "ok";
    case PMSZ_EINVAL: return "invalid argument (parameters, NaN or the 1e30 sentinel in a field)";
    case PMSZ_ENOMEM: return "device allocation failed";
    case PMSZ_NOT_CONVERGED: return "not converged within max_sweeps";
    case PMSZ_EDEVICE: return "HIP runtime error";
    case PMSZ_ETOOBIG: return "field too large (32-bit voxel indices, or the fixed scratch of a "
                              "no-allocation entry point)";
    case PMSZ_EBUSY: return "capacity exhausted";
    default: return "unknown error";
  }
}

const char *pmsz_version(void) { return "pmsz-mi355x 0.3"; }

}  // extern "C"
