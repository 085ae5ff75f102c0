// pmsz_common.h -- definitions shared by the HIP kernels and the host code of
// libpmsz.so: the stencil in rank order, the exact bound predicate and the edit
// primitive.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmsz {

// The 15 candidates of a voxel -- itself and its 14 stencil neighbours
// (directions_host, pMSz.cu:186-193) -- ranked by linear index, i.e. in
// lexicographic (dz, dy, dx) order.  Ties in the total order (value, index) of
// islarger_shared / isless_shared (pMSz.cu:610-620) go to the higher rank for
// the maximum and to the lower rank for the minimum.  Rank 7 is the voxel itself.
constexpr int kSelf = 7;
constexpr int kCand = 15;
__host__ __device__ constexpr int rank_dx(int r) {
  return r == 0 ? 0 : r == 1 ? 1 : r == 2 ? 0 : r == 3 ? 1 : r == 4 ? 0 : r == 5 ? 1
       : r == 6 ? -1 : r == 7 ? 0 : r == 8 ? 1 : r == 9 ? -1 : r == 10 ? 0 : r == 11 ? -1
       : r == 12 ? 0 : r == 13 ? -1 : 0;
}
__host__ __device__ constexpr int rank_dy(int r) {
  return r == 0 ? -1 : r == 1 ? -1 : r == 2 ? 0 : r == 3 ? 0 : r == 4 ? -1 : r == 5 ? -1
       : r == 6 ? 0 : r == 7 ? 0 : r == 8 ? 0 : r == 9 ? 1 : r == 10 ? 1 : r == 11 ? 0
       : r == 12 ? 0 : r == 13 ? 1 : 1;
}
__host__ __device__ constexpr int rank_dz(int r) { return r < 4 ? -1 : (r < 11 ? 0 : 1); }

// Missing-neighbour sentinel of the reference (INVALID_VAL = 1e30, pMSz.cu:215),
// compared in the field's value type.
template <typename T>
__host__ __device__ inline T sentinel() { return (T)1e30; }

// |a - b| <= eb over the reals: TwoSum gives a - b = s + e exactly; |s| - eb is
// exact (Sterbenz) whenever it has to be examined.
__host__ __device__ inline bool within_exact(double a, double b, double eb) {
  double s = a - b;
  double bb = s - a;
  double e = (a - (s - bb)) + (-b - bb);
  if (s < 0) {
    s = -s;
    e = -e;
  }
  if (s < eb) return true;
  if (s > 2 * eb) return false;
  double r = s - eb;
  return r <= -e;
}

__host__ __device__ inline float next_up(float c) {
  if (c != c || c == __builtin_huge_valf()) return c;
  if (c == 0.0f) return __builtin_bit_cast(float, 1u);
  uint32_t b = __builtin_bit_cast(uint32_t, c);
  b = c > 0 ? b + 1 : b - 1;
  return __builtin_bit_cast(float, b);
}
__host__ __device__ inline double next_up(double c) {
  if (c != c || c == __builtin_huge_val()) return c;
  if (c == 0.0) return __builtin_bit_cast(double, (uint64_t)1);
  uint64_t b = __builtin_bit_cast(uint64_t, c);
  b = c > 0 ? b + 1 : b - 1;
  return __builtin_bit_cast(double, b);
}

__host__ __device__ inline float next_down(float c) { return -next_up(-c); }
__host__ __device__ inline double next_down(double c) { return -next_up(-c); }

struct EditParams {
  double eb;      // the absolute bound
  double delta_d; // eb / q in double (delta = bound / q, pMSz.cu:2814)
  int q;
  int clamp_mode; // 0 reference, 1 bound-safe
  int budgets;    // 1: the high nibble of a delta counter holds a step budget (q <= 15)
};

// Clamp value of the edit primitive (pMSz.cu:636).  Bound-safe mode keeps that
// value when |f - c| <= eb holds exactly, and otherwise takes the smallest value
// of T that honours the bound: the real f - eb rounded up.  That value is found
// from f - eb formed in double (within one unit in the last place of T of it)
// and settled with the exact predicate, one step each way at most.
template <typename T>
__host__ __device__ inline T clamp_search(T fv, const EditParams &ep) {
  T c = (T)((double)fv - ep.eb);
  for (int k = 0; k < 8 && !within_exact((double)fv, (double)c, ep.eb); k++) c = next_up(c);
  for (int k = 0; k < 8 && within_exact((double)fv, (double)next_down(c), ep.eb); k++)
    c = next_down(c);
  return c;
}

template <typename T>
__host__ __device__ inline T clamp_value(T fv, const EditParams &ep) {
  T eps = (T)ep.eb;
  T c = fv - eps;
  if (ep.clamp_mode == 1 && c < fv && !within_exact((double)fv, (double)c, ep.eb))
    c = clamp_search(fv, ep);
  return c;
}

// applyDeltaBuffer_local (pMSz.cu:623-645) on one target's start-of-sweep state.
template <typename T>
__host__ __device__ inline void edit_value(T fv, T &gv, uint8_t &dc, const EditParams &ep) {
  T delta = (T)ep.delta_d;
  T eps = (T)ep.eb;
  T cand = gv - delta;
  bool ok;
  if (ep.clamp_mode == 1) {
    ok = within_exact((double)fv, (double)cand, ep.eb);
  } else {
    T diff = fv - cand;
    if (diff < 0) diff = -diff;
    ok = diff <= eps;
  }
  if (ok && dc < ep.q - 1) {
    gv = cand;
    dc = (uint8_t)(dc + 1);
  } else {
    dc = (uint8_t)ep.q;
    gv = clamp_value(fv, ep);
  }
}

}  // namespace pmsz
