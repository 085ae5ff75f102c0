// pmsz_dense.h -- the full-volume kernel of the correction: original labeling
// of f and/or one detection sweep over g, for every voxel.
//
// Layout of the work: a wave owns 62 consecutive x (lanes 1..62; lanes 0 and 63
// carry the x-1 / x+1 halo), kColRows consecutive y, and walks z.  No LDS and no
// barriers: a lane loads its x column of kColRows+2 rows per plane from HBM
// (coalesced across the wave), the x-1 / x+1 neighbours arrive through DPP wave
// shifts, and z is a register pipeline.
//
// The 15 candidates of voxel (x, y, z) -- itself and its 14 stencil neighbours,
// ranked by linear index (pmsz_common.h) -- split by plane: ranks 0..3 lie in
// plane z-1, 4..10 in plane z, 11..14 in plane z+1.  When plane p arrives, each
// row gets three partial extremes from it: "lo" (p as the z-1 plane of voxel
// p+1), "mid" (p as the voxel's own plane) and "hi" (p as the z+1 plane of voxel
// p-1).  Voxel p-1 is then complete: extreme(lo(p-2), mid(p-1), hi(p)).
// Every cell is loaded once per pass.
//
// Candidate order.  float32: the staged key 2^52 + 16 * ord(v) plus the rank is
// an exact integer in a double whose numeric order is the (value, index) order
// of islarger_shared / isless_shared (pMSz.cu:610-620); max / min are single
// v_max_f64 / v_min_f64 and the winner's rank is the key's low 4 bits.
// float64: (value, rank) pairs combined in increasing rank order, the later one
// winning ties for the maximum and the earlier one for the minimum.
// Out-of-domain cells (the reference's INVALID halo, pMSz.cu:709) are quiet NaN:
// maxnum / minnum and the pair combines skip them.
//
// Targets (with kDetect) go to per-lane bit masks over (dx, row) for the three
// planes a voxel's targets can lie in; a completed plane is exchanged with the
// neighbouring lanes, turned into one 64-lane mask per row by ballots and OR-ed
// into the target bitmap with one atomic per 32 cells; cells whose bit was new
// are appended to the target list with one reservation per wave.
#pragma once
#include "pmsz_device.h"

namespace pmsz {

constexpr int kColOwned = 62;  // x owned by a wave (64 lanes minus the two halo lanes)
constexpr int kColRows = 4;    // y rows per wave
constexpr int kColH = kColRows + 2;
constexpr int kColZ = 64;      // planes per work unit
constexpr int kColWaves = 4;   // waves per block
constexpr int kColThreads = 64 * kColWaves;

// Value of lane i-1 / lane i+1 (lane 0 / lane 63 receive 0).
__device__ __forceinline__ uint32_t from_left_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t from_right_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ double from_left(double v) {
  return __hiloint2double((int)from_left_u((uint32_t)__double2hiint(v)),
                          (int)from_left_u((uint32_t)__double2loint(v)));
}
__device__ __forceinline__ double from_right(double v) {
  return __hiloint2double((int)from_right_u((uint32_t)__double2hiint(v)),
                          (int)from_right_u((uint32_t)__double2loint(v)));
}

// Target bit of rank r in the 54-bit plane accumulator of a lane, less the voxel
// row h: field (plane e-1, e, e+1) * 18 + (dx + 1) * kColH + dy.
__host__ __device__ constexpr uint32_t target_base(int r) {
  return (uint32_t)((rank_dz(r) + 1) * 3 * kColH + (rank_dx(r) + 1) * kColH + rank_dy(r));
}

// Offsets of rank r packed for runtime lookup: (dx+1) | (dy+1) << 2, 4 bits per rank.
__host__ __device__ constexpr uint64_t rank_xy_table() {
  uint64_t t = 0;
  for (int r = 0; r < kCand; r++)
    t |= (uint64_t)((rank_dx(r) + 1) | ((rank_dy(r) + 1) << 2)) << (4 * r);
  return t;
}

// Running extremes over candidates visited in increasing rank order.
template <typename T>
struct Ext;

template <>
struct Ext<float> {  // staged keys
  double mx, mn;
  __device__ __forceinline__ void first(double k, int r) { mx = mn = k + (double)r; }
  __device__ __forceinline__ void add(double k, int r) {
    const double v = k + (double)r;
    mx = fmax(mx, v);
    mn = fmin(mn, v);
  }
  __device__ __forceinline__ void merge(const Ext &o) {  // o holds higher ranks
    mx = fmax(mx, o.mx);
    mn = fmin(mn, o.mn);
  }
  __device__ __forceinline__ int asc() const { return __double2loint(mx) & 15; }
  __device__ __forceinline__ int desc() const { return __double2loint(mn) & 15; }
  static __device__ __forceinline__ double stage(float v) { return stage_value(v); }
};

template <>
struct Ext<double> {  // (value, rank) pairs
  double mx, mn;
  int rmx, rmn;
  __device__ __forceinline__ void first(double v, int r) {
    mx = mn = v;
    rmx = rmn = r;
  }
  // b comes later in rank order: it wins a tie for the maximum, loses it for
  // the minimum; NaN (invalid) never wins against a valid value
  __device__ __forceinline__ void put(double v, int r) {
    const bool anan = mx != mx;
    const bool tmx = v >= mx || anan;
    mx = tmx ? v : mx;
    rmx = tmx ? r : rmx;
    const bool tmn = v < mn || anan;
    mn = tmn ? v : mn;
    rmn = tmn ? r : rmn;
  }
  __device__ __forceinline__ void add(double v, int r) { put(v, r); }
  __device__ __forceinline__ void merge(const Ext &o) {
    const bool anan = mx != mx;
    const bool tmx = o.mx >= mx || anan;
    mx = tmx ? o.mx : mx;
    rmx = tmx ? o.rmx : rmx;
    const bool tmn = o.mn < mn || anan;
    mn = tmn ? o.mn : mn;
    rmn = tmn ? o.rmn : rmn;
  }
  __device__ __forceinline__ int asc() const { return rmx; }
  __device__ __forceinline__ int desc() const { return rmn; }
  static __device__ __forceinline__ double stage(double v) { return v; }
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
// in rank order, which decodes to the rank with two compares.
__device__ __forceinline__ double key32(float v, uint32_t t) {
  const float c = v + 0.0f;  // -0 -> +0: the reference compares -0 == +0
  const uint32_t b = __float_as_uint(c);
  const uint32_t ord = b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
  return __hiloint2double((int)(0x43300000u | (ord >> 23)), (int)((ord << 9) | t));
}
__device__ __forceinline__ int key_rank(double w, uint32_t tself) {
  const int d = (int)((uint32_t)__double2loint(w) & 511u) - (int)tself + 8;
  return d - (d > 2 ? 1 : 0) - (d > 14 ? 1 : 0);
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

// Flush the completed target plane `pz` of this wave's column block: `m` holds,
// per lane, bits (dx+1)*kColH + row for targets at (x+dx, y0-1+row, pz).  After
// the exchange with the neighbouring lanes each lane holds the target bits of its
// own x and marks each target's byte in the run's target byte map with the
// sweep's tag, by a plain store: several waves may mark one cell, all with the
// same value, so no atomic is needed.  The target list is compacted from the
// byte map afterwards; bytes holding other tags are marks of earlier sweeps.
__device__ __forceinline__ void flush_targets(uint32_t m, int64_t xl, int64_t y0m1, int64_t pz,
                                              const Grid3 &d, uint8_t *tbyte, uint8_t tag) {
  const uint32_t dxm = m & ((1u << kColH) - 1);                   // dx = -1
  const uint32_t dx0 = (m >> kColH) & ((1u << kColH) - 1);        // dx =  0
  const uint32_t dxp = (m >> (2 * kColH)) & ((1u << kColH) - 1);  // dx = +1
  uint32_t col = dx0 | from_left_u(dxp) | from_right_u(dxm);
  // a set bit is a target, so its cell lies in the block
  const int64_t base = xl + d.nx * y0m1 + d.sxy * pz;
  while (col) {
    const int h = __ffs(col) - 1;
    col &= col - 1;
    tbyte[base + d.nx * h] = tag;
  }
}

// Target list from the target byte map after a full-volume detection: every
// byte equal to the sweep's tag is one distinct target, listed as its voxel id.
// Each thread takes 16-byte groups; a block reserves its list range with one
// atomic per step.  Nothing is cleared: the next sweep uses another tag.
constexpr int kCompactThreads = 256;
constexpr int kCompactPT = 8;  // 16-byte groups per thread per step
__device__ __forceinline__ uint32_t tag_bytes(uint32_t w, uint32_t rep) {
  // 0x01 in each byte of w equal to the tag (rep = tag * 0x01010101), else 0
  const uint32_t x = w ^ rep;
  const uint32_t nz = (((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) & 0x80808080u;
  return (~nz >> 7) & 0x01010101u;
}
__global__ __launch_bounds__(kCompactThreads) void k_compact(
    const uint8_t *__restrict__ tbyte, uint32_t *__restrict__ tlist, Counters *__restrict__ ring,
    uint32_t jslot_prev, uint32_t jslot, bool always, uint32_t ngroups,
    const Dom *__restrict__ doms, uint32_t ndom, uint32_t tag) {
  if (!always && ring[jslot_prev].mode != kModeDense) return;  // uniform
  uint32_t *counter = &ring[jslot].ntargets;
  const uint4 *g16 = reinterpret_cast<const uint4 *>(tbyte);
  const uint32_t rep = tag * 0x01010101u;
  const uint32_t step = gridDim.x * kCompactThreads * kCompactPT;
  for (uint32_t base = blockIdx.x * kCompactThreads * kCompactPT; base < ngroups; base += step) {
    uint32_t w[kCompactPT][4];
    uint32_t n = 0;
#pragma unroll
    for (int q = 0; q < kCompactPT; q++) {
      const uint32_t i = base + (uint32_t)q * kCompactThreads + threadIdx.x;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (i < ngroups) v = g16[i];
      w[q][0] = tag_bytes(v.x, rep);
      w[q][1] = tag_bytes(v.y, rep);
      w[q][2] = tag_bytes(v.z, rep);
      w[q][3] = tag_bytes(v.w, rep);
#pragma unroll
      for (int c = 0; c < 4; c++) n += __popc(w[q][c]);
    }
    uint32_t pos = block_reserve<kCompactThreads>(n, counter);
#pragma unroll
    for (int q = 0; q < kCompactPT; q++) {
      const uint32_t i = base + (uint32_t)q * kCompactThreads + threadIdx.x;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        uint32_t m = w[q][c];
        while (m) {
          const int b = (__ffs(m) - 1) >> 3;
          m &= m - 1;
          const uint32_t cell = i * 16u + (uint32_t)(c * 4 + b);
          uint32_t id = cell;
          if (ndom > 1) {  // the domain whose cells start at or before `cell`
            uint32_t lo = 0, hi = ndom;
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) / 2;
              if (doms[mid].cbase <= cell) lo = mid;
              else hi = mid;
            }
            id = doms[lo].gbase + (cell - doms[lo].cbase);
          }
          tlist[pos++] = id;
        }
      }
    }
  }
}

// Tag of the marks of detection jd in the target byte map: 1..255 in turn (the
// map is cleared whenever the tags start over).
__host__ __device__ inline uint32_t detect_tag(uint32_t jd) { return 1u + (jd - 1u) % 255u; }

// Units (62 x 4 x 64 column blocks) of one domain's box.
__host__ inline void col_units(int64_t bx, int64_t by, int64_t bz, uint32_t &nux, uint32_t &nuy,
                               uint32_t &nuz) {
  nux = (uint32_t)((bx + kColOwned - 1) / kColOwned);
  nuy = (uint32_t)((by + kColRows - 1) / kColRows);
  nuz = (uint32_t)((bz + kColZ - 1) / kColZ);
}

// The full-volume kernel.  With kLabel the original labeling of f (codes; the
// delta counters are cleared when the domain has them), with kDetect a
// detection sweep on g reading the codes (check_g: reject NaN / sentinel values
// of g, in the first sweep).  kMulti: the units of all domains of the run,
// domain state in ds (par = parity of the detection).  Targets go to the byte
// map tbyte.
template <typename T, bool kLabel, bool kDetect, bool kMulti>
__global__ __launch_bounds__(kColThreads) void k_columns(
    Dom d0, const Dom *__restrict__ doms, uint32_t ndom, DomState *__restrict__ ds,
    uint8_t *__restrict__ tbyte, Counters *__restrict__ ring, uint32_t jslot_prev,
    uint32_t jslot, uint32_t par, RunState *run, uint32_t check_g, uint32_t tag) {
  if (!kLabel) {
    if (ring[jslot_prev].mode != kModeDense) return;  // uniform
  }
  const double kNaN = __builtin_nan("");
  __shared__ uint8_t s_tbase[16];
  if (kDetect) {
    if (threadIdx.x < kCand) s_tbase[threadIdx.x] = (uint8_t)target_base((int)threadIdx.x);
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
  uint32_t lerr = 0;

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
    const T *__restrict__ f = (const T *)D.f;
    const T *__restrict__ g = (const T *)D.g;
    uint8_t *__restrict__ code = D.code;
    uint8_t *__restrict__ dc = D.dc;
    const uint32_t nx = (uint32_t)D.d.nx, ny = (uint32_t)D.d.ny, nz = (uint32_t)D.d.nz;
    const uint32_t sxy = (uint32_t)D.d.sxy;
    const uint32_t u = unit - D.unit_base;
    const uint32_t uz = u / (D.nux * D.nuy), rem = u - uz * D.nux * D.nuy;
    const uint32_t uy = rem / D.nux, ux = rem - uy * D.nux;
    const int x = D.lo[0] + (int)(ux * kColOwned) - 1 + (int)lane;  // this lane's x
    const int y0m1 = D.lo[1] + (int)(uy * kColRows) - 1;             // y of loaded row 0
    const int zs = D.lo[2] + (int)(uz * kColZ);
    const int ze = zs + kColZ < D.hi[2] ? zs + kColZ : D.hi[2];
    const bool xok = x >= 0 && x < (int)nx;
    const bool owned = lane >= 1 && lane <= (uint32_t)kColOwned && x < D.hi[0];
    const int yhi = D.hi[1];
    // in-plane offsets of the loaded rows, fixed for the unit
    uint32_t roff[kColH];
    bool rok[kColH];
#pragma unroll
    for (int h = 0; h < kColH; h++) {
      const int y = y0m1 + h;
      rok[h] = xok && y >= 0 && y < (int)ny;
      roff[h] = rok[h] ? (uint32_t)x + nx * (uint32_t)y : 0u;
    }
    // raw values of the next plane, in flight; the plane's validity is uniform
    // over the wave, so an out-of-domain plane is a branch, not a select
    T nf[kColH], ng[kColH];
    auto fetch = [&](int pz) {
      if (pz < 0 || pz >= (int)nz) return;
      const uint32_t zoff = sxy * (uint32_t)pz;
#pragma unroll
      for (int h = 0; h < kColH; h++) {
        const uint32_t gi = zoff + roff[h];
        if (kLabel) nf[h] = f[gi];
        if (kDetect) ng[h] = g[gi];
      }
    };
    auto stage = [&](int pz, const T (&raw)[kColH], ColPlane &c, bool check) {
      if (pz < 0 || pz >= (int)nz) {
#pragma unroll
        for (int h = 0; h < kColH; h++) c.k[h] = c.l[h] = c.r[h] = kNaN;
        return;
      }
#pragma unroll
      for (int h = 0; h < kColH; h++) {
        const T v = raw[h];
        bool ok = rok[h];
        if (padded) {
          ok = ok && v != sentinel<T>();  // a ghost cell that carries no value
        } else if (check && ok && (!(v == v) || v == sentinel<T>())) {
          lerr |= 1u;
        }
        c.k[h] = ok ? Ext<T>::stage(v) : kNaN;
      }
      shift_plane(c);
    };
    uint8_t ccode[kColRows], ncode[kColRows];  // codes of the plane to emit / of the next
    auto fetch_codes = [&](int pz) {
#pragma unroll
      for (int h = 0; h < kColRows; h++) {
        const int y = y0m1 + 1 + h;
        const bool ok = owned && y < yhi && pz >= 0 && pz < (int)nz;
        ncode[h] = code[ok ? (uint32_t)x + nx * (uint32_t)y + sxy * (uint32_t)pz : 0u];
      }
    };

    // target bits of planes e-1, e, e+1 in three 18-bit fields: (dx+1)*6 + row
    uint64_t acc = 0;
    uint32_t uc[3] = {0, 0, 0};              // counters of this unit
    const int64_t cbase = (int64_t)D.cbase;
    // detection at owned voxel row h of plane e: codes, counters, target masks
    auto visit = [&](int h, int e, int oa, int od, int da, int dd) {
      const int y = y0m1 + h;
      if (!(owned && y < yhi)) return;
      const uint32_t i = (uint32_t)x + nx * (uint32_t)y + sxy * (uint32_t)e;
      if (kLabel) {
        code[i] = (uint8_t)(oa | (od << 4));
        if (dc) dc[i] = 0;  // the run's delta counters start at zero
      }
      if (kDetect) {
        int t1, t2;
        bool v1, v2;
        detect_pair(oa, od, da, dd, uc, t1, v1, t2, v2);
        if (emit) {
          const uint64_t b1 = (uint64_t)(v1 ? 1u : 0u) << (s_tbase[t1] + (uint32_t)h);
          const uint64_t b2 = (uint64_t)(v2 ? 1u : 0u) << (s_tbase[t2] + (uint32_t)h);
          acc |= b1 | b2;
        }
      }
    };
    auto flush_plane = [&](int e) {
      // plane e-1 can receive no more targets
      flush_targets((uint32_t)acc & ((1u << (3 * kColH)) - 1u), cbase + x, y0m1, (int64_t)e - 1,
                    D.d, tbyte, (uint8_t)tag);
      acc >>= 3 * kColH;
    };
    if constexpr (sizeof(T) == 4) {
      // per voxel row: lo (plane p's share for plane p+1) and part (planes p-1, p
      // for plane p), as max / min keys
      double lf_mx[kColRows], lf_mn[kColRows], pf_mx[kColRows], pf_mn[kColRows];
      double lg_mx[kColRows], lg_mn[kColRows], pg_mx[kColRows], pg_mn[kColRows];
#pragma unroll
      for (int h = 0; h < kColRows; h++) lf_mx[h] = lf_mn[h] = lg_mx[h] = lg_mn[h] = kNaN;
      auto key_of = [&](bool pin, T v, int h, uint32_t tplane, bool check) -> double {
        bool ok = pin && rok[h];
        if (padded) {
          ok = ok && v != sentinel<T>();  // a ghost cell that carries no value
        } else if (check && ok && (!(v == v) || v == sentinel<T>())) {
          lerr |= 1u;
        }
        return ok ? key32(v, tplane + 3u * (uint32_t)h) : kNaN;
      };
      // row extremes: A over (x, x+1), B over (x-1..x+1), C over (x-1, x)
      struct Win {
        double a2x, a2n, a1x, a1n, b1x, b1n, c1x, c1n;  // rows k-2 (A) and k-1
      };
      fetch(zs - 1);
      for (int p = zs - 1; p <= ze; p++) {
        const bool pin = p >= 0 && p < (int)nz;  // uniform
        const uint32_t tplane = lane + 5u * (uint32_t)(p - zs + 1);
        T rf[kColH], rg[kColH];
#pragma unroll
        for (int h = 0; h < kColH; h++) {
          if (kLabel) rf[h] = nf[h];
          if (kDetect) rg[h] = ng[h];
        }
        if (!kLabel) {
#pragma unroll
          for (int h = 0; h < kColRows; h++) ccode[h] = ncode[h];
        }
        if (p < ze) {
          fetch(p + 1);
          if (!kLabel) fetch_codes(p);
        }
        const int e = p - 1;  // plane completed at this step
        const bool emitp = e >= zs;
        const uint32_t tself0 = lane + 5u * (uint32_t)(p - zs);
        Win wf, wg;
#pragma unroll
        for (int k = 0; k < kColH; k++) {
          double ax = 0, an = 0, bx = 0, bn = 0, cx = 0, cn = 0;
          double gax = 0, gan = 0, gbx = 0, gbn = 0, gcx = 0, gcn = 0;
          if (kLabel) {
            const double kk = key_of(pin, rf[k], k, tplane, true);
            const double l = from_left(kk), r = from_right(kk);
            ax = kmax(kk, r);
            an = kmin(kk, r);
            cx = kmax(l, kk);
            cn = kmin(l, kk);
            bx = kmax(ax, l);
            bn = kmin(an, l);
          }
          if (kDetect) {
            const double kk = key_of(pin, rg[k], k, tplane, kLabel || check_g != 0);
            const double l = from_left(kk), r = from_right(kk);
            gax = kmax(kk, r);
            gan = kmin(kk, r);
            gcx = kmax(l, kk);
            gcn = kmin(l, kk);
            gbx = kmax(gax, l);
            gbn = kmin(gan, l);
          }
          if (k >= 2) {
            const int h = k - 1;  // voxel row completed by row k
            if (emitp) {
              const uint32_t ts = tself0 + 3u * (uint32_t)h;
              int oa, od, da = 0, dd = 0;
              if (kLabel) {
                oa = key_rank(kmax(pf_mx[h - 1], kmax(wf.c1x, cx)), ts);
                od = key_rank(kmin(pf_mn[h - 1], kmin(wf.c1n, cn)), ts);
              } else {
                oa = ccode[h - 1] & 15;
                od = ccode[h - 1] >> 4;
              }
              if (kDetect) {
                da = key_rank(kmax(pg_mx[h - 1], kmax(wg.c1x, gcx)), ts);
                dd = key_rank(kmin(pg_mn[h - 1], kmin(wg.c1n, gcn)), ts);
              }
              visit(h, e, oa, od, da, dd);
            }
            // plane p's partials: mid for plane p, lo for plane p+1
            if (kLabel) {
              pf_mx[h - 1] = kmax(lf_mx[h - 1], kmax(wf.a2x, kmax(wf.b1x, cx)));
              pf_mn[h - 1] = kmin(lf_mn[h - 1], kmin(wf.a2n, kmin(wf.b1n, cn)));
              lf_mx[h - 1] = kmax(wf.a2x, wf.a1x);
              lf_mn[h - 1] = kmin(wf.a2n, wf.a1n);
            }
            if (kDetect) {
              pg_mx[h - 1] = kmax(lg_mx[h - 1], kmax(wg.a2x, kmax(wg.b1x, gcx)));
              pg_mn[h - 1] = kmin(lg_mn[h - 1], kmin(wg.a2n, kmin(wg.b1n, gcn)));
              lg_mx[h - 1] = kmax(wg.a2x, wg.a1x);
              lg_mn[h - 1] = kmin(wg.a2n, wg.a1n);
            }
          }
          if (kLabel) {
            wf.a2x = wf.a1x; wf.a2n = wf.a1n;
            wf.a1x = ax; wf.a1n = an;
            wf.b1x = bx; wf.b1n = bn;
            wf.c1x = cx; wf.c1n = cn;
          }
          if (kDetect) {
            wg.a2x = wg.a1x; wg.a2n = wg.a1n;
            wg.a1x = gax; wg.a1n = gan;
            wg.b1x = gbx; wg.b1n = gbn;
            wg.c1x = gcx; wg.c1n = gcn;
          }
        }
        if (kDetect && emitp) flush_plane(e);
      }
    } else {
      Ext<T> lo_f[kColRows], part_f[kColRows], lo_g[kColRows], part_g[kColRows];
#pragma unroll
      for (int h = 0; h < kColRows; h++) {
        lo_f[h].first(kNaN, 0);
        lo_g[h].first(kNaN, 0);
      }
      fetch(zs - 1);
      for (int p = zs - 1; p <= ze; p++) {
        ColPlane cf, cg;
        if (kLabel) stage(p, nf, cf, true);
        if (kDetect) stage(p, ng, cg, kLabel || check_g != 0);
        if (!kLabel) {
#pragma unroll
          for (int h = 0; h < kColRows; h++) ccode[h] = ncode[h];
        }
        if (p < ze) {
          fetch(p + 1);
          if (!kLabel) fetch_codes(p);
        }
        const int e = p - 1;  // plane completed at this step
        const bool emitp = e >= zs;
#pragma unroll
        for (int h = 1; h <= kColRows; h++) {
          Ext<T> hf, hg, mf, mg, lf, lg;
          if (emitp) {
            int oa = 0, od = 0, da = 0, dd = 0;
            if (kLabel) {
              partial_hi<T>(cf, h, hf);
              part_f[h - 1].merge(hf);
              oa = part_f[h - 1].asc();
              od = part_f[h - 1].desc();
            } else {
              oa = ccode[h - 1] & 15;
              od = ccode[h - 1] >> 4;
            }
            if (kDetect) {
              partial_hi<T>(cg, h, hg);
              part_g[h - 1].merge(hg);
              da = part_g[h - 1].asc();
              dd = part_g[h - 1].desc();
            }
            visit(h, e, oa, od, da, dd);
          }
          // partials of plane p for the voxels of planes p and p+1
          if (kLabel) {
            partial_mid<T>(cf, h, mf);
            partial_lo<T>(cf, h, lf);
            part_f[h - 1] = lo_f[h - 1];
            part_f[h - 1].merge(mf);
            lo_f[h - 1] = lf;
          }
          if (kDetect) {
            partial_mid<T>(cg, h, mg);
            partial_lo<T>(cg, h, lg);
            part_g[h - 1] = lo_g[h - 1];
            part_g[h - 1].merge(mg);
            lo_g[h - 1] = lg;
          }
        }
        if (kDetect && emitp) flush_plane(e);
      }
    }
    if (kDetect) {
      flush_targets((uint32_t)acc & ((1u << (3 * kColH)) - 1u), cbase + x, y0m1, (int64_t)ze - 1,
                    D.d, tbyte, (uint8_t)tag);
      flush_targets((uint32_t)(acc >> (3 * kColH)) & ((1u << (3 * kColH)) - 1u), cbase + x, y0m1,
                    (int64_t)ze, D.d, tbyte, (uint8_t)tag);
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
  const uint64_t em = __ballot(lerr != 0);
  if (em && lane == 0) atomicOr(&run->err, 1u);
  if (!kDetect) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const uint32_t s = wave_sum(lc[c]);
    if (lane == 0 && s) atomicAdd(&out->cnt[c], s);
  }
  if (threadIdx.x == 0 && blockIdx.x == 0 && !kLabel) atomicAdd(&run->dense_passes, 1ull);
}

}  // namespace pmsz
