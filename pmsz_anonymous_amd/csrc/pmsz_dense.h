// pmsz_dense.h -- the full-volume kernel of the correction: the original
// labeling of f, or one detection sweep over g, for every voxel.
//
// Layout of the work: a wave owns 62 consecutive x (lanes 1..62; lanes 0 and 63
// carry the x-1 / x+1 halo), kColRows consecutive y, and walks z -- or, where a
// box's x extent would leave much of such a column empty, two half-columns of 30
// x, one per half-wave, each with its own halo lanes (col_units).  A lane loads
// its x column of kColRows+2 rows per plane from HBM (coalesced across the
// wave), the x-1 / x+1 neighbours arrive through DPP wave shifts, and z is a
// register pipeline.  No barriers; LDS holds only a table (the detection's fast
// test, the float32 labeling's code bytes) and each wave's violation queue.
//
// The 15 candidates of voxel (x, y, z) -- itself and its 14 stencil neighbours,
// ranked by linear index (pmsz_common.h) -- split by plane: ranks 0..3 lie in
// plane z-1, 4..10 in plane z, 11..14 in plane z+1.  When plane p arrives, each
// row gets three partial extremes from it: "lo" (p as the z-1 plane of voxel
// p+1), "mid" (p as the voxel's own plane) and "hi" (p as the z+1 plane of voxel
// p-1).  Voxel p-1 is then complete: extreme(lo(p-2), mid(p-1), hi(p)).
// Every cell is loaded once per pass.
//
// Candidate order.  float32: each cell becomes an exact double key (key32
// below) whose numeric order is the (value, index) order of islarger_shared /
// isless_shared (pMSz.cu:610-620); max / min are single v_max_f64 / v_min_f64.
// float64: (value, rank) pairs combined in increasing rank order, the later one
// winning ties for the maximum and the earlier one for the minimum.
// Out-of-domain cells (the reference's INVALID halo, pMSz.cu:709) are quiet NaN:
// maxnum / minnum and the pair combines skip them.
//
// Detection compares each voxel's winners in g with its code; the voxels that
// differ go through a per-wave LDS queue to the exact rules, which count them
// and mark their targets in the target byte map.
#pragma once
#include "pmsz_device.h"

namespace pmsz {

constexpr int kColOwned = 62;   // x owned by a wave (64 lanes minus the two halo lanes)
constexpr int kHalfOwned = 30;  // x owned by a half-wave (32 lanes minus its two halo lanes)
constexpr int kColRows = 4;    // y rows per wave
constexpr int kColH = kColRows + 2;
constexpr int kColZ = 64;      // planes per work unit by default (a domain's zc: 16, 32 or 64)
constexpr int kColWaves = 4;   // waves per block
constexpr int kColThreads = 64 * kColWaves;

// Value of lane i-1 / lane i+1 (lane 0 / lane 63 receive 0: bound_ctrl).
__device__ __forceinline__ uint32_t from_left_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ uint32_t from_right_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ double from_left(double v) {
  return __hiloint2double((int)from_left_u((uint32_t)__double2hiint(v)),
                          (int)from_left_u((uint32_t)__double2loint(v)));
}
__device__ __forceinline__ double from_right(double v) {
  return __hiloint2double((int)from_right_u((uint32_t)__double2hiint(v)),
                          (int)from_right_u((uint32_t)__double2loint(v)));
}

// Running extremes over candidates visited in increasing rank order (the float64
// passes; float32 takes the keys below).
template <typename T>
struct Ext;

template <>
struct Ext<double> {  // (value, rank) pairs
  double mx, mn;
  int rmx, rmn;
  __device__ __forceinline__ void first(double v, int r) {
    mx = mn = v;
    rmx = rmn = r;
  }
  // (b, rb) comes later in rank order: it wins a tie for the maximum and loses
  // it for the minimum.  fmax / fmin skip a NaN (an invalid candidate) and keep
  // the other operand, so the maximum's rank moves to rb iff b equals the new
  // maximum, and the minimum's iff the new minimum differs from the old one (a
  // NaN old value always differs; -0 equals +0, as the reference's compares have
  // it).  Two v_max_f64 / v_min_f64 and two compare-selects per combine.
  __device__ __forceinline__ void take(double b, int rb, double bn, int rbn) {
    const double nmx = fmax(mx, b);
    rmx = b == nmx ? rb : rmx;
    mx = nmx;
    const double nmn = fmin(mn, bn);
    rmn = nmn != mn ? rbn : rmn;
    mn = nmn;
  }
  __device__ __forceinline__ void put(double v, int r) { take(v, r, v, r); }
  __device__ __forceinline__ void add(double v, int r) { put(v, r); }
  __device__ __forceinline__ void merge(const Ext &o) { take(o.mx, o.rmx, o.mn, o.rmn); }
  __device__ __forceinline__ int asc() const { return rmx; }
  __device__ __forceinline__ int desc() const { return rmn; }
};

// One plane's staged column of one field, with the x-1 / x+1 neighbours.
struct ColPlane {
  double k[kColH], l[kColH], r[kColH];
};

__device__ __forceinline__ void shift_plane(ColPlane &c) {
#pragma unroll
  for (int h = 0; h < kColH; h++) {
    c.l[h] = from_left(c.k[h]);
    c.r[h] = from_right(c.k[h]);
  }
}

// Partial extremes contributed by plane p to voxel row h (1..kColRows).
template <typename T>
__device__ __forceinline__ void partial_lo(const ColPlane &c, int h, Ext<T> &e) {
  e.first(c.k[h - 1], 0);   // (x,   y-1, z-1)
  e.add(c.r[h - 1], 1);     // (x+1, y-1, z-1)
  e.add(c.k[h], 2);         // (x,   y,   z-1)
  e.add(c.r[h], 3);         // (x+1, y,   z-1)
}
template <typename T>
__device__ __forceinline__ void partial_mid(const ColPlane &c, int h, Ext<T> &e) {
  e.first(c.k[h - 1], 4);   // (x,   y-1)
  e.add(c.r[h - 1], 5);     // (x+1, y-1)
  e.add(c.l[h], 6);         // (x-1, y)
  e.add(c.k[h], 7);         // the voxel
  e.add(c.r[h], 8);         // (x+1, y)
  e.add(c.l[h + 1], 9);     // (x-1, y+1)
  e.add(c.k[h + 1], 10);    // (x,   y+1)
}
template <typename T>
__device__ __forceinline__ void partial_hi(const ColPlane &c, int h, Ext<T> &e) {
  e.first(c.l[h], 11);      // (x-1, y,   z+1)
  e.add(c.k[h], 12);        // (x,   y,   z+1)
  e.add(c.l[h + 1], 13);    // (x-1, y+1, z+1)
  e.add(c.k[h + 1], 14);    // (x,   y+1, z+1)
}

// float32 keys with a cell-intrinsic tie-break.  Within one unit a cell at lane
// L, loaded row h and plane counter pl carries the tag t = L + 3h + 5pl; its key
// is the exact double 2^52 + 512 * ord(v) + t.  Over the 15 candidates of any
// voxel the tags increase strictly in linear-index order (the rows of a
// neighbourhood are 3 apart in tag, the planes 5 apart, and the lanes differ by
// at most 2), so the numeric order of the keys is the (value, index) order of
// islarger_shared / isless_shared (pMSz.cu:610-620) with no per-candidate rank.
// The winner's tag minus the voxel's own is one of -8, -7, -5, -4, ..., 5, 7, 8
// in rank order, which decodes to the rank with two compares (diff_rank).
__device__ __forceinline__ double key32(float v, uint32_t t) {
  const float c = v + 0.0f;  // -0 -> +0: the reference compares -0 == +0
  const uint32_t b = __float_as_uint(c);
  const uint32_t ord = b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
  return __hiloint2double((int)(0x43300000u | (ord >> 23)), (int)((ord << 9) | t));
}

// max / min of two keys as single v_max_f64 / v_min_f64.  Keys are exact
// positive numbers, the quiet NaN of a missing cell, or the 0.0 a DPP shift
// brings into an edge lane (whose results are never used); none is a signalling
// NaN, so the instructions need no canonicalised operands, and a quiet NaN
// operand yields the other one.
__device__ __forceinline__ double kmax(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double kmin(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Target list from the target byte map after a full-volume detection: every
// byte equal to the sweep's tag is one distinct target, listed as its voxel id.
// A thread takes kCompactPT 16-byte groups per step (interleaved across the
// block, so loads coalesce) and keeps one 16-bit mask per group; a block makes
// one list reservation per step (same-address atomics serialise, so steps are
// large).  Nothing is cleared: the next sweep uses another tag.
constexpr int kCompactThreads = 256;
constexpr int kCompactPT = 16;     // groups per thread per step (all loads in flight)
constexpr int kCompactBatch = 16;  // loads in flight per thread
constexpr int kCompactDoms = 1024;  // most domains of a run
__device__ __forceinline__ uint32_t tag_bytes(uint32_t w, uint32_t rep) {
  // 0x01 in each byte of w equal to the tag (rep = tag * 0x01010101), else 0
  const uint32_t x = w ^ rep;
  const uint32_t nz = (((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) & 0x80808080u;
  return (~nz >> 7) & 0x01010101u;
}
// The four 0x01 / 0x00 bytes of tag_bytes() as 4 bits.
__device__ __forceinline__ uint32_t byte_bits(uint32_t t) {
  return (t | (t >> 7) | (t >> 14) | (t >> 21)) & 15u;
}
__global__ __launch_bounds__(kCompactThreads) void k_compact(
    const uint8_t *__restrict__ tbyte, uint32_t *__restrict__ tlist, Counters *__restrict__ ring,
    uint32_t jslot_prev, uint32_t jslot, bool always, uint32_t ngroups,
    const Dom *__restrict__ doms, uint32_t ndom, uint32_t tag, uint32_t stream_min) {
  if (!always && ring[jslot_prev].mode != kModeDense) return;  // uniform
  if (stream_min && cnt_sum(ring[jslot]) >= stream_min) return;  // the apply streams the map
  // the domains' first cells and ids, for the cell -> id mapping (ndom <= 1024)
  __shared__ uint32_t s_cb[kCompactDoms], s_gb[kCompactDoms];
  if (ndom > 1) {
    for (uint32_t k = threadIdx.x; k < ndom; k += kCompactThreads) {
      s_cb[k] = doms[k].cbase;
      s_gb[k] = doms[k].gbase;
    }
    __syncthreads();
  }
  uint32_t *counter = &ring[jslot].ntargets;
  const uint4 *g16 = reinterpret_cast<const uint4 *>(tbyte);
  const uint32_t rep = tag * 0x01010101u;
  const uint32_t per_step = kCompactThreads * kCompactPT;
  for (uint32_t base = blockIdx.x * per_step; base < ngroups; base += gridDim.x * per_step) {
    uint32_t m[kCompactPT];
    uint32_t n = 0;
#pragma unroll
    for (int q0 = 0; q0 < kCompactPT; q0 += kCompactBatch) {
      uint4 v[kCompactBatch];
#pragma unroll
      for (int q = 0; q < kCompactBatch; q++) {
        const uint32_t i = base + (uint32_t)(q0 + q) * kCompactThreads + threadIdx.x;
        v[q] = i < ngroups ? g16[i] : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < kCompactBatch; q++) {
        m[q0 + q] = byte_bits(tag_bytes(v[q].x, rep)) | (byte_bits(tag_bytes(v[q].y, rep)) << 4) |
                    (byte_bits(tag_bytes(v[q].z, rep)) << 8) |
                    (byte_bits(tag_bytes(v[q].w, rep)) << 12);
        n += __popc(m[q0 + q]);
      }
    }
    uint32_t pos = block_reserve<kCompactThreads>(n, counter);
#pragma unroll
    for (int q = 0; q < kCompactPT; q++) {
      uint32_t mq = m[q];
      const uint32_t i = base + (uint32_t)q * kCompactThreads + threadIdx.x;
      while (mq) {
        const uint32_t cell = i * 16u + (uint32_t)(__ffs(mq) - 1);
        mq &= mq - 1;
        uint32_t id = cell;
        if (ndom > 1) {  // the domain whose cells start at or before `cell`
          uint32_t lo = 0, hi = ndom;
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (s_cb[mid] <= cell) lo = mid;
            else hi = mid;
          }
          id = s_gb[lo] + (cell - s_cb[lo]);
        }
        tlist[pos++] = id;
      }
    }
  }
}

// Tag of the marks of detection jd in the target byte map: 1..255 in turn (the
// map is cleared whenever the tags start over).
__host__ __device__ inline uint32_t detect_tag(uint32_t jd) { return 1u + (jd - 1u) % 255u; }

// Units of one domain's box: waves walking zc planes of columns of kColRows y.
// A wave takes one column 62 x wide, or two half-columns 30 x wide (the
// half-columns of a z chunk in (x, y) order, two to a wave) where that needs
// fewer waves: boxes whose x extent leaves much of a 62-wide column empty.
__host__ inline void col_units(Dom &D, int64_t bx, int64_t by, int64_t bz, int64_t zc) {
  const uint32_t nyb = (uint32_t)((by + kColRows - 1) / kColRows);
  const uint32_t nfull = (uint32_t)((bx + kColOwned - 1) / kColOwned);
  const uint32_t nhalf = (uint32_t)((bx + kHalfOwned - 1) / kHalfOwned);
  D.half = (nhalf * nyb + 1) / 2 < nfull * nyb ? 1u : 0u;
  D.nhx = D.half ? nhalf : nfull;
  D.nyb = nyb;
  D.nux = D.half ? (nhalf * nyb + 1) / 2 : nfull * nyb;
  D.nuy = 1;
  D.nuz = (uint32_t)((bz + zc - 1) / zc);
}


// Violations of a full-volume detection are resolved off the per-voxel path.
// Every voxel runs a fast test -- are both winners in g the original's? -- and
// the lanes that fail append (cell, info) to their wave's queue in LDS; the queue
// is resolved 64 entries at a time, one per lane, by the exact rules
// (detect_pair): counters and target bytes.  info = the code byte | the
// detection's winners above it (float32: winner tag - voxel tag + 8, 5 bits
// each; float64: ranks, 4 bits each).
constexpr int kQCap = 512;  // queue entries per wave (a power of two)

// Orders one wave's LDS writes before its later LDS reads by other lanes (and
// reads before later overwrites): a wavefront-scope fence around a wave barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Winner tag - voxel tag + 8 of the candidate of rank r (float32 keys).
__host__ __device__ constexpr uint32_t tag_diff(int r) {
  return (uint32_t)(r + (r >= 2 ? 1 : 0) + (r >= 13 ? 1 : 0));
}
__device__ __forceinline__ int diff_rank(int d) { return d - (d > 2 ? 1 : 0) - (d > 14 ? 1 : 0); }

// Fast-test table of the float32 detection, indexed by the code byte: the tag
// differences (mod 512) that the original's ascent and descent winners have, in
// the low and high 16 bits.  A voxel that is both maximum and minimum of its
// neighbourhood (it has no valid neighbour) gets an ascent entry no winner can
// match, so the exact rules always see it.
__host__ __device__ inline uint32_t expect_entry(uint32_t c) {
  const int oa = (int)(c & 15u), od = (int)(c >> 4);
  uint32_t a = (tag_diff(oa) + 504u) & 511u;
  const uint32_t d = (tag_diff(od) + 504u) & 511u;
  if (oa == kSelf && od == kSelf) a = 256u;
  return a | (d << 16);
}

// Wave masks of compares (one v_cmp into a lane mask, no select-and-compare):
// integer predicates of __builtin_amdgcn_uicmp, and the lanes whose value is NaN
// or the sentinel (unordered-or-equal, v_cmp_nlg).
constexpr int kICmpEq = 32, kICmpNe = 33;
__device__ __forceinline__ uint64_t bad_mask(float v) {
  return __builtin_amdgcn_fcmpf(v, sentinel<float>(), 9);
}
__device__ __forceinline__ uint64_t bad_mask(double v) {
  return __builtin_amdgcn_fcmp(v, sentinel<double>(), 9);
}

// Loads through a buffer resource whose base is a plane's first byte: the
// in-plane byte offset is the instruction's 32-bit VGPR offset, so a load
// costs no address arithmetic.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)0xffffffff,
                                           0x00020000);
}
__device__ __forceinline__ void plane_load(__amdgpu_buffer_rsrc_t rs, uint32_t off, float &v) {
  v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
}
__device__ __forceinline__ void plane_load(__amdgpu_buffer_rsrc_t rs, uint32_t off, double &v) {
  const auto w = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
  v = __hiloint2double((int)w[1], (int)w[0]);
}

// The full-volume kernel: with kLabel the original labeling of f (codes; the
// delta counters are cleared when the domain has them), with kDetect a
// detection sweep on g against the codes (kCheck: reject NaN / sentinel values
// of g, in the first detection).  kMulti: the units of all domains of the run,
// domain state in ds (par = parity of the detection).  Targets are marked in
// the byte map tbyte with the detection's tag.
// Waves per SIMD asked for where that was measured faster (1: no request):
// * four (at most 128 VGPRs) for the float32 multi-domain detections (135 VGPRs
//   otherwise; with a few spilled registers the 64-field batch takes 3-5 % less
//   time) and the float64 labeling (140 otherwise; no spills);
// * five (at most 102) for the float32 labeling (114 otherwise; 0.277 -> 0.261 ms
//   at 512^3, two registers spilled in the one-field kernel).
// The float64 detections spill 96 bytes at 128 VGPRs and are slower, so they keep
// three waves.
constexpr int cols_waves(bool f32, bool label, bool detect, bool multi) {
  return label ? (f32 ? 5 : 4) : (detect && multi && f32) ? 4 : 1;
}
template <typename T, bool kLabel, bool kDetect, bool kMulti, bool kCheck = false>
__global__ __launch_bounds__(kColThreads) __attribute__((amdgpu_waves_per_eu(
    cols_waves(sizeof(T) == 4, kLabel, kDetect, kMulti)))) void k_columns(
    Dom d0, const Dom *__restrict__ doms, uint32_t ndom, DomState *__restrict__ ds,
    uint8_t *__restrict__ tbyte, Counters *__restrict__ ring, uint32_t jslot_prev,
    uint32_t jslot, uint32_t par, RunState *run, uint32_t tag) {
  static_assert(kLabel != kDetect, "a pass labels f or detects on g");
  static_assert(kColThreads == 256, "one fast-test table entry per thread");
  constexpr bool kF32 = sizeof(T) == 4;
  constexpr bool kCheckV = kLabel || kCheck;  // validate the loaded values
  if (kDetect) {
    if (ring[jslot_prev].mode != kModeDense) return;  // uniform
  }
  const double kNaN = __builtin_nan("");
  __shared__ uint32_t s_exp[kDetect && kF32 ? 256 : 1];
  __shared__ uint2 s_q[kDetect ? kColWaves * kQCap : 1];
  // per wave: the cell offsets of the 15 candidates in the current unit's domain
  __shared__ int32_t s_off[kDetect ? kColWaves * 16 : 1];
  // float32 labeling: the code byte of the winners' tag differences (+ 8, 5 bits
  // each, ascent low), one table read per voxel row instead of two rank decodes
  __shared__ uint8_t s_code[kLabel && kF32 ? 1024 : 1];
  if (kDetect && kF32) {
    s_exp[threadIdx.x] = expect_entry(threadIdx.x);
    __syncthreads();
  }
  if (kLabel && kF32) {
    for (uint32_t k = threadIdx.x; k < 1024u; k += kColThreads)
      s_code[k] = (uint8_t)(diff_rank((int)min(k & 31u, 16u)) |
                            (diff_rank((int)min(k >> 5, 16u)) << 4));
    __syncthreads();
  }
  const uint32_t lane = lane_id();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  uint32_t nunits;
  if (kMulti) {
    const Dom &L = doms[ndom - 1];
    nunits = L.unit_base + L.nux * L.nuy * L.nuz;
  } else {
    nunits = d0.nux * d0.nuy * d0.nuz;
  }
  // XCD-aware split: blocks are dealt to the 8 XCDs by workgroup id; each XCD
  // takes one contiguous eighth of the units, so neighbouring units share an L2
  const uint32_t waves_per_xcd = (gridDim.x / 8) * kColWaves;
  const uint32_t chunk = (nunits + 7) / 8;
  const uint32_t u_begin = (blockIdx.x % 8) * chunk;
  const uint32_t u_end = u_begin + chunk < nunits ? u_begin + chunk : nunits;
  Counters *out = kDetect ? &ring[jslot] : nullptr;
  uint32_t lc[3] = {0, 0, 0};
  uint64_t lerr = 0;  // lanes that loaded a NaN or the sentinel (a wave mask)
  const uint32_t qbase = wave * (uint32_t)kQCap;

  for (uint32_t unit = u_begin + (blockIdx.x / 8) * kColWaves + wave; unit < u_end;
       unit += waves_per_xcd) {
    uint32_t dk = 0;
    if (kMulti) {  // last domain whose first unit is <= unit (uniform)
      uint32_t a = 0, b = ndom;
      while (b - a > 1) {
        const uint32_t m = (a + b) / 2;
        if (doms[m].unit_base <= unit) a = m;
        else b = m;
      }
      dk = a;
    }
    const Dom &D = kMulti ? doms[dk] : d0;
    bool emit = true;
    if (kMulti && kDetect) {
      const uint32_t fl = ds[dk].flags;
      if (!(fl & kDomActive)) continue;
      emit = (fl & kDomEmit) != 0;
    }
    const bool padded = kMulti && D.padded;
    // the field this pass reads, and per plane its first byte: loads and stores
    // take a uniform plane base plus a 32-bit in-plane byte offset
    const char *__restrict__ src = (const char *)(kLabel ? D.f : (const void *)D.g);
    uint8_t *__restrict__ code = D.code;
    uint8_t *__restrict__ dc = D.dc;
    const uint32_t nx = (uint32_t)D.d.nx, ny = (uint32_t)D.d.ny, nz = (uint32_t)D.d.nz;
    const uint64_t sxy = (uint64_t)D.d.sxy;
    const uint32_t u = unit - D.unit_base;
    const uint32_t uz = u / D.nux, w = u - uz * D.nux;  // (nuy = 1)
    // this lane's column: the wave's, or with half-columns the one of its half
    // (lanes 0 and 31 of a half carry its x-1 / x+1 halo; a DPP shift across
    // the halves reaches only halo lanes, whose results are not used, and every
    // voxel's candidates lie in its own half)
    const bool half = D.half != 0;  // uniform
    const uint32_t cidx = half ? 2u * w + (lane >> 5) : w, cl = half ? lane & 31u : lane;
    const uint32_t cown = half ? (uint32_t)kHalfOwned : (uint32_t)kColOwned;
    const bool cvalid = cidx < D.nhx * D.nyb;
    const uint32_t cy = cidx / D.nhx, cx = cidx - cy * D.nhx;
    const int x = D.lo[0] + (int)(cx * cown) - 1 + (int)cl;  // this lane's x
    const int y0m1 = D.lo[1] + (int)(cy * kColRows) - 1;     // y of loaded row 0
    const int zs = D.lo[2] + (int)(uz * D.zc);
    const int ze = zs + (int)D.zc < D.hi[2] ? zs + (int)D.zc : D.hi[2];
    const bool xok = cvalid && x >= 0 && x < (int)nx;
    const bool owned = cvalid && cl >= 1 && cl <= cown && x < D.hi[0];
    // in-plane offsets of the loaded rows (bytes of T) and of the voxel rows
    // (cells; 0 where the lane has no such cell), fixed for the unit
    uint32_t roff[kColH], coff[kColRows];
    bool rok[kColH], vis[kColRows];
#pragma unroll
    for (int h = 0; h < kColH; h++) {
      const int y = y0m1 + h;
      rok[h] = xok && y >= 0 && y < (int)ny;
      roff[h] = rok[h] ? ((uint32_t)x + nx * (uint32_t)y) * (uint32_t)sizeof(T) : 0u;
    }
#pragma unroll
    for (int h = 0; h < kColRows; h++) {
      const int y = y0m1 + 1 + h;
      vis[h] = owned && y < D.hi[1];
      coff[h] = vis[h] ? (uint32_t)x + nx * (uint32_t)y : 0u;
    }
    // the fast test's mask of the tag bits: 511 where the lane has the voxel row
    uint32_t vmask[kColRows];
#pragma unroll
    for (int h = 0; h < kColRows; h++) vmask[h] = vis[h] ? 511u : 0u;
    // raw values of plane pz into nv; the plane's validity is uniform over the
    // wave, so an out-of-domain plane is a branch, not a select
    auto fetch = [&](int pz, T (&nv)[kColH]) {
      if (pz < 0 || pz >= (int)nz) return;
      const auto rs = plane_rsrc(src + (uint64_t)pz * sxy * sizeof(T));
#pragma unroll
      for (int h = 0; h < kColH; h++) plane_load(rs, roff[h], nv[h]);
    };
    auto fetch_codes = [&](int pz, uint32_t (&nc)[kColRows]) {  // pz lies in the domain
      const auto rs = plane_rsrc(code + (uint64_t)pz * sxy);
#pragma unroll
      for (int h = 0; h < kColRows; h++)
        nc[h] = __builtin_amdgcn_raw_buffer_load_b8(rs, coff[h], 0, 0);
    };
    // the value's validity: a padded block's sentinel cells carry no value; in
    // a field, NaN or the sentinel is an input error
    // (pin: the plane lies in the domain, uniform).  A lane outside the domain
    // loaded the plane's first cell (offset 0), a cell of the domain, so the
    // input check needs no lane mask: whatever it flags is a NaN or sentinel
    // of the input.
    auto valid = [&](bool ok, bool pin, T v) -> bool {
      if (padded) return ok && v != sentinel<T>();
      if (kCheckV && pin) lerr |= bad_mask(v);
      return ok;
    };

    // the violation queue of this wave: entries [qh, qt), resolved per unit
    uint32_t uc[3] = {0, 0, 0};  // counters of this unit
    uint32_t qh = 0, qt = 0;
    // target bytes are stored through a buffer resource based at the domain's
    // first cell, with the target's 32-bit cell index (below 2^32 - 1) as offset;
    // the candidates' offsets come from this wave's table (written here, read
    // after the wave_lds_sync that precedes every resolve; the previous unit's
    // queue was drained before)
    const auto trs = plane_rsrc(tbyte + D.cbase);
    const uint32_t obase = wave * 16u;
    if (kDetect && lane < (uint32_t)kCand) s_off[obase + lane] = (int32_t)rank_off((int)lane, D.d);
    auto resolve = [&](uint32_t count) {  // the `count` oldest entries (count <= 64)
      if (lane < count) {
        const uint2 q = s_q[qbase + ((qh + lane) & (uint32_t)(kQCap - 1))];
        const int oa = (int)(q.y & 15u), od = (int)((q.y >> 4) & 15u);
        int da, dd;
        if constexpr (kF32) {
          // bytes 1, 2: the winners' tag differences less the original winners',
          // mod 256 (the differences themselves lie in 0..16)
          da = diff_rank((int)min(((q.y >> 8) + tag_diff(oa)) & 255u, 16u));
          dd = diff_rank((int)min(((q.y >> 16) + tag_diff(od)) & 255u, 16u));
        } else {
          da = (int)((q.y >> 8) & 15u);
          dd = (int)((q.y >> 12) & 15u);
        }
        int t1, t2;
        bool v1, v2;
        detect_pair(oa, od, da, dd, uc, t1, v1, t2, v2);
        if (emit) {  // a target is a cell of the block
          if (v1)
            __builtin_amdgcn_raw_buffer_store_b8((uint8_t)tag, trs,
                                                 q.x + (uint32_t)s_off[obase + (uint32_t)t1], 0, 0);
          if (v2)
            __builtin_amdgcn_raw_buffer_store_b8((uint8_t)tag, trs,
                                                 q.x + (uint32_t)s_off[obase + (uint32_t)t2], 0, 0);
        }
      }
      qh += count;
    };
    // append the lanes of wave mask bm; `info` is evaluated only if bm is not 0
    auto enqueue = [&](uint64_t bm, int h, int e, auto info) {
      if (bm == 0) return;  // uniform
      if (__builtin_amdgcn_inverse_ballot_w64(bm)) {
        const uint32_t pos = qt + __builtin_amdgcn_mbcnt_hi(
                                      (uint32_t)(bm >> 32),
                                      __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
        const uint32_t cell = coff[h] + (uint32_t)((uint64_t)e * sxy);
        s_q[qbase + (pos & (uint32_t)(kQCap - 1))] = make_uint2(cell, info());
      }
      qt += (uint32_t)__popcll(bm);
      if (qt - qh > (uint32_t)(kQCap - 64)) {  // room for the next row's appends
        wave_lds_sync();
        while (qt - qh >= 64u) resolve(64u);
        wave_lds_sync();
      }
    };
    // code and counter bytes go through buffer resources of plane e, with the
    // voxel row's in-plane offset (no 64-bit address arithmetic per row)
    auto label = [&](int h, int e, uint32_t cbyte) {
      if (!vis[h]) return;
      const uint64_t pb = (uint64_t)e * sxy;
      __builtin_amdgcn_raw_buffer_store_b8((uint8_t)cbyte, plane_rsrc(code + pb), coff[h], 0, 0);
      if (dc)  // the run's delta counters start at zero
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)0, plane_rsrc(dc + pb), coff[h], 0, 0);
    };

    if constexpr (kF32) {
      // per voxel row: lo (plane p's share for plane p+1) and part (planes p-1, p
      // for plane p), as max / min keys
      double lo_mx[kColRows], lo_mn[kColRows], pt_mx[kColRows], pt_mn[kColRows];
#pragma unroll
      for (int h = 0; h < kColRows; h++) lo_mx[h] = lo_mn[h] = kNaN;
      // plane step p: rows rv of plane p are in registers, cc holds the codes of
      // plane p-1.  The labeling loads plane p+2 into nv (two steps ahead: it has
      // registers to spare, and a load then has two steps of every wave to
      // arrive); a detection loads plane p+1 into nv and the codes of plane p into
      // nc (one step ahead: measured faster, the detection is issue-bound)
      constexpr int kAhead = kLabel ? 2 : 1;
      auto step = [&](int p, const T (&rv)[kColH], const uint32_t (&cc)[kColRows],
                      T (&nv)[kColH], uint32_t (&nc)[kColRows]) {
        const bool pin = p >= 0 && p < (int)nz;  // uniform
        if (p + kAhead <= ze) fetch(p + kAhead, nv);
        if (kDetect && p >= zs && p < ze) fetch_codes(p, nc);
        const uint32_t tplane = lane + 5u * (uint32_t)(p - zs + 1);
        const int e = p - 1;  // plane completed at this step
        const bool emitp = e >= zs;
        const uint32_t tself0 = lane + 5u * (uint32_t)(p - zs);
        // row extremes: A over (x, x+1) and C over (x-1, x) of rows k-2 (A only)
        // and k-1; the row's own key k1 and right neighbour r1 of row k-1
        double a2x = 0, a2n = 0, a1x = 0, a1n = 0, c1x = 0, c1n = 0, k1 = 0, r1 = 0;
#pragma unroll
        for (int k = 0; k < kColH; k++) {
          const bool ok = valid(pin && rok[k], pin, rv[k]);
          const float cv = rv[k] + 0.0f;  // -0 -> +0: the reference compares -0 == +0
          const uint32_t b = __float_as_uint(cv);
          const uint32_t ord = b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
          // key32(v, tag): high word 0x43300000 | ord >> 23 by one funnel shift; a
          // missing cell is a quiet NaN whatever its low word
          const uint32_t khi = ok ? __builtin_amdgcn_alignbit(0x219800u, ord, 23) : 0x7ff80000u;
          const double kk =
              __hiloint2double((int)khi, (int)((ord << 9) | (tplane + 3u * (uint32_t)k)));
          const double l = from_left(kk), r = from_right(kk);
          const double ax = kmax(kk, r), an = kmin(kk, r);
          const double cx = kmax(l, kk), cn = kmin(l, kk);
          if (k >= 2) {
            const int h = k - 1;  // voxel row completed by row k
            // with y the voxel row: D = C(y) + C(y+1), the plane's share for the
            // voxel in the previous plane; Q = A(y-1) + the cell right of the voxel
            const double dx = kmax(c1x, cx), dn = kmin(c1n, cn);
            const double qx = kmax(a2x, r1), qn = kmin(a2n, r1);
            if (emitp) {
              const double wmx = kmax(pt_mx[h - 1], dx);
              const double wmn = kmin(pt_mn[h - 1], dn);
              const uint32_t ts = tself0 + 3u * (uint32_t)h;
              if (kLabel) {
                // a winner's tag difference + 8 lies in 0..16 and the tags sit
                // below the key's ordinal bits: (low word + 8 - ts) mod 32 is it
                const uint32_t k8 = 8u - ts;
                const uint32_t ia = ((uint32_t)__double2loint(wmx) + k8) & 31u;
                const uint32_t id = ((uint32_t)__double2loint(wmn) + k8) & 31u;
                label(h - 1, e, s_code[ia | (id << 5)]);
              } else {
                const uint32_t c = cc[h - 1];
                const uint32_t ex = s_exp[c] + ts * 0x10001u;
                const uint32_t ra = (uint32_t)__double2loint(wmx) - (ex & 0xffffu);
                const uint32_t rd = (uint32_t)__double2loint(wmn) - (ex >> 16);
                // lanes whose winners differ: one compare into a wave mask
                const uint64_t bm =
                    __builtin_amdgcn_uicmp((ra | rd) & vmask[h - 1], 0u, kICmpNe);
                // info: the code byte, then the low bytes of ra and rd
                enqueue(bm, h - 1, e, [&] {
                  return c | (__builtin_amdgcn_perm(rd, ra, 0x0c0c0400u) << 8);
                });
              }
            }
            // plane p's partials: mid for plane p (A(y-1), B(y) = C(y) + the right
            // cell, C(y+1)), lo for plane p+1 (A(y-1) + A(y) = Q + the cell itself)
            pt_mx[h - 1] = kmax(lo_mx[h - 1], kmax(dx, qx));
            pt_mn[h - 1] = kmin(lo_mn[h - 1], kmin(dn, qn));
            lo_mx[h - 1] = kmax(qx, k1);
            lo_mn[h - 1] = kmin(qn, k1);
          }
          a2x = a1x; a2n = a1n;
          a1x = ax; a1n = an;
          c1x = cx; c1n = cn;
          k1 = kk; r1 = r;
        }
      };
      // the code bytes, one per register (packed bytes would cost an extract each)
      uint32_t ca[kColRows] = {}, cb[kColRows] = {};
      int p = zs - 1;
      if constexpr (kAhead == 2) {
        // three register sets for the planes in rotation (no copies between
        // steps): step q reads set q mod 3 and loads plane q + 2 into set
        // (q + 2) mod 3
        T va[kColH], vb[kColH], vc[kColH];
        fetch(zs - 1, va);
        fetch(zs, vb);
        for (; p + 2 <= ze; p += 3) {
          step(p, va, ca, vc, cb);
          step(p + 1, vb, ca, va, cb);
          step(p + 2, vc, ca, vb, cb);
        }
        if (p <= ze) step(p, va, ca, vc, cb);
        if (p + 1 <= ze) step(p + 1, vb, ca, va, cb);
      } else {
        // two register sets for the planes, alternating
        T va[kColH], vb[kColH];
        fetch(zs - 1, va);
        for (; p + 1 <= ze; p += 2) {
          step(p, va, ca, vb, cb);
          step(p + 1, vb, cb, va, ca);
        }
        if (p <= ze) step(p, va, ca, vb, cb);
      }
    } else {
      Ext<T> lo_e[kColRows], part[kColRows];
#pragma unroll
      for (int h = 0; h < kColRows; h++) lo_e[h].first(kNaN, 0);
      T nv[kColH];
      uint32_t ccode[kColRows], ncode[kColRows] = {};
      fetch(zs - 1, nv);
      for (int p = zs - 1; p <= ze; p++) {
        ColPlane cp;
        if (p < 0 || p >= (int)nz) {
#pragma unroll
          for (int h = 0; h < kColH; h++) cp.k[h] = cp.l[h] = cp.r[h] = kNaN;
        } else {
#pragma unroll
          for (int h = 0; h < kColH; h++) cp.k[h] = valid(rok[h], true, nv[h]) ? (double)nv[h] : kNaN;
          shift_plane(cp);
        }
        if (kDetect) {
#pragma unroll
          for (int h = 0; h < kColRows; h++) ccode[h] = ncode[h];
        }
        if (p < ze) {
          fetch(p + 1, nv);
          if (kDetect && p >= zs) fetch_codes(p, ncode);
        }
        const int e = p - 1;  // plane completed at this step
        const bool emitp = e >= zs;
#pragma unroll
        for (int h = 1; h <= kColRows; h++) {
          if (emitp) {
            Ext<T> hi;
            partial_hi<T>(cp, h, hi);
            part[h - 1].merge(hi);
            const int da = part[h - 1].asc(), dd = part[h - 1].desc();
            if (kLabel) {
              label(h - 1, e, (uint32_t)da | ((uint32_t)dd << 4));
            } else {
              const uint32_t c = ccode[h - 1];
              const uint32_t w = (uint32_t)da | ((uint32_t)dd << 4);
              const uint64_t bm =
                  (__builtin_amdgcn_uicmp(w, c, kICmpNe) |
                   __builtin_amdgcn_uicmp(c, (uint32_t)(kSelf | (kSelf << 4)), kICmpEq)) &
                  __builtin_amdgcn_ballot_w64(vis[h - 1]);
              enqueue(bm, h - 1, e, [&] { return c | (w << 8); });
            }
          }
          // partials of plane p for the voxels of planes p and p+1; the old lo is
          // consumed before the new one is formed in its place, so the loop
          // carries both without copies
          Ext<T> mid;
          partial_mid<T>(cp, h, mid);
          part[h - 1] = lo_e[h - 1];
          part[h - 1].merge(mid);
          partial_lo<T>(cp, h, lo_e[h - 1]);
        }
      }
    }
    if (kDetect) {
      wave_lds_sync();
      while (qt != qh) resolve(qt - qh < 64u ? qt - qh : 64u);
      wave_lds_sync();
      if (kMulti) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const uint32_t sm = wave_sum(uc[c]);
          if (lane == 0 && sm) atomicAdd(&ds[dk].cnt[par][c], sm);
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) lc[c] += uc[c];
    }
  }
  if (lerr && lane == 0) atomicOr(&run->err, 1u);
  if (!kDetect) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const uint32_t s = wave_sum(lc[c]);
    if (lane == 0) cnt_add(*out, c, s);
  }
  if (threadIdx.x == 0 && blockIdx.x == 0) atomicAdd(&run->dense_passes, 1ull);
}

}  // namespace pmsz
