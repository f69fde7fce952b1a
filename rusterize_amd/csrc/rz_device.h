// rz_device.h -- device helpers shared by every kernel translation unit:
// Rust scalar rules, the crossing, wave-level synchronisation and
// reductions, the strip coverage of a polygon item and the seven pixel
// folds (plain and packed).
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>

#include "rz_kernels.h"

namespace rzk {

// ------------------------------------------------------ Rust scalar rules --

// `v as usize` (saturating, NaN -> 0), then clamped to `cap`.
__device__ inline int32_t sat_usize_cap(double v, int32_t cap) {
  if (!(v > 0.0)) return 0;
  if (v >= (double)cap) return cap;
  return (int32_t)v;
}
// `v as isize`: saturating, NaN -> 0.
__device__ inline int64_t sat_isize(double v) {
  if (v != v) return 0;
  if (v <= -9223372036854775808.0) return INT64_MIN;
  if (v >= 9223372036854775808.0) return INT64_MAX;
  return (int64_t)v;
}

// NaN results as the reference's x86 build produces them (SSE): an operation
// with a NaN operand returns that NaN, sign kept (the first one when both
// are NaN); an invalid operation (0 * inf, inf - inf, inf / inf) returns the
// default NaN, which is negative.  The GPU's default NaN is positive, and a
// compiled a - b may negate b first, NaN included, so the edge set-up picks
// every NaN explicitly.
__device__ inline double x86_nan(double r, double a, double b) {
  return isnan(a) ? a : (isnan(b) ? b : (isnan(r) ? -__builtin_nan("") : r));
}

// PolyEdge::intersect_at (rust/src/geo/edges.rs:50-55): x0 + ((r+0.5) - y0) * dxdy.
// k_prep leaves no NaN x0 or dxdy in a kept edge, so a NaN here comes from an
// invalid operation, which is negative on x86: span_col sends it to column 0
// like any NaN, which is where f64::total_cmp sorts a negative NaN.
__device__ inline double crossing(const PolyEdge& e, int32_t r) {
  return e.x0 + (((double)r + 0.5) - e.y0) * e.dxdy;
}

// Orders a wave's LDS accesses around this point (wave-private buffers need
// no workgroup barrier).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ----------------------------------------------------------- pixel folds --

template <typename T> __device__ inline bool isnan_t(T) { return false; }
template <> __device__ inline bool isnan_t<float>(float v) { return v != v; }
template <> __device__ inline bool isnan_t<double>(double v) { return v != v; }

template <typename T> struct Unsigned { using type = T; };
template <> struct Unsigned<int8_t> { using type = uint8_t; };
template <> struct Unsigned<int16_t> { using type = uint16_t; };
template <> struct Unsigned<int32_t> { using type = uint32_t; };
template <> struct Unsigned<int64_t> { using type = uint64_t; };

// Integer `+=` wraps (release profile); floats add in their own precision.
template <typename T> __device__ inline T add_wrap(T a, T b) {
  using U = typename Unsigned<T>::type;
  return (T)(U)((U)a + (U)b);
}
template <> __device__ inline float add_wrap<float>(float a, float b) { return a + b; }
template <> __device__ inline double add_wrap<double>(double a, double b) { return a + b; }

// The seven transitions of rust/src/rasterization/pixel_functions.rs:55-123,
// applied k >= 1 times with the same value (all writes of one geometry
// carry one value, so only their number matters; sum and count are the
// only non-idempotent transitions).
template <typename T, int FUN>
__device__ inline void fold(T& p, T v, T bg, uint32_t k) {
  if (FUN == 0) {  // sum
    for (uint32_t i = 0; i < k; ++i)
      p = (p == bg || isnan_t(p) || isnan_t(v)) ? v : add_wrap(p, v);
  } else if (FUN == 1) {  // first
    if (p == bg || isnan_t(p)) p = v;
  } else if (FUN == 2) {  // last
    p = v;
  } else if (FUN == 3) {  // min
    if (p == bg || isnan_t(p) || p > v) p = v;
  } else if (FUN == 4) {  // max
    if (p == bg || isnan_t(p) || p < v) p = v;
  } else if (FUN == 5) {  // count
    for (uint32_t i = 0; i < k; ++i) p = (p == bg || isnan_t(p)) ? (T)1 : add_wrap(p, (T)1);
  } else {  // any
    p = (T)1;
  }
}

// One write of value v, applied where `hit`: the same transitions as fold
// with k = 1, written as selects so the compiler emits no branch per pixel.
template <typename T, int FUN>
__device__ inline T fold_select(T p, T v, T bg, bool hit) {
  T n;
  if (FUN == 0) n = (p == bg || isnan_t(p) || isnan_t(v)) ? v : add_wrap(p, v);
  else if (FUN == 1) n = (p == bg || isnan_t(p)) ? v : p;
  else if (FUN == 2) n = v;
  else if (FUN == 3) n = (p == bg || isnan_t(p) || p > v) ? v : p;
  else if (FUN == 4) n = (p == bg || isnan_t(p) || p < v) ? v : p;
  else if (FUN == 5) n = (p == bg || isnan_t(p)) ? (T)1 : add_wrap(p, (T)1);
  else n = (T)1;
  return hit ? n : p;
}

// ---- packed folds for 1- and 2-byte outputs: the same transitions on the
// 4 (or 2) pixels of a 32-bit word at once, lane-wise on the word (SWAR).
// Equality and wrapping addition are bitwise, so signedness only matters for
// min / max.
template <int S> struct Lanes;
template <> struct Lanes<1> {
  static constexpr uint32_t lo7 = 0x7F7F7F7Fu, hi = 0x80808080u, one = 0x01010101u, full = 0xFFu;
  static constexpr int sh = 7;
};
template <> struct Lanes<2> {
  static constexpr uint32_t lo7 = 0x7FFF7FFFu, hi = 0x80008000u, one = 0x00010001u, full = 0xFFFFu;
  static constexpr int sh = 15;
};

// all-ones in each lane where a and b are equal
template <int S> __device__ inline uint32_t lane_eq(uint32_t a, uint32_t b) {
  using L = Lanes<S>;
  const uint32_t x = a ^ b;
  const uint32_t z = ~(((x & L::lo7) + L::lo7) | x | L::lo7);  // lane sign bit set iff the lane is 0
  return (z >> L::sh) * L::full;
}
// lane-wise wrapping a + b
template <int S> __device__ inline uint32_t lane_add(uint32_t a, uint32_t b) {
  using L = Lanes<S>;
  return ((a & L::lo7) + (b & L::lo7)) ^ ((a ^ b) & L::hi);
}
// 4 (or 2) pixel bits -> all-ones lanes
template <int S> __device__ inline uint32_t lane_mask(uint32_t bits) {
  if (S == 1) return ((bits * 0x00204081u) & 0x01010101u) * 0xFFu;
  return ((bits & 1u) * 0xFFFFu) | (((bits >> 1) & 1u) * 0xFFFF0000u);
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
template <typename V> __device__ inline V as_vec(uint32_t w) {
  V v;
  __builtin_memcpy(&v, &w, 4);
  return v;
}
template <typename V> __device__ inline uint32_t as_word(V v) {
  uint32_t w;
  __builtin_memcpy(&w, &v, 4);
  return w;
}
// lane-wise unsigned min / max of 16-bit lanes (packed 16-bit VALU)
__device__ inline uint32_t pk_min_u16(uint32_t a, uint32_t b) {
  return as_word(__builtin_elementwise_min(as_vec<u16x2>(a), as_vec<u16x2>(b)));
}
__device__ inline uint32_t pk_max_u16(uint32_t a, uint32_t b) {
  return as_word(__builtin_elementwise_max(as_vec<u16x2>(a), as_vec<u16x2>(b)));
}
// lane-wise min (MAX = false) or max of T lanes
template <typename T, bool MAX> __device__ inline uint32_t lane_minmax(uint32_t a, uint32_t b) {
  if (sizeof(T) == 2) {
    if (T(-1) < T(0)) {  // int16
      const i16x2 x = as_vec<i16x2>(a), y = as_vec<i16x2>(b);
      return as_word(MAX ? __builtin_elementwise_max(x, y) : __builtin_elementwise_min(x, y));
    }
    return MAX ? pk_max_u16(a, b) : pk_min_u16(a, b);
  }
  // bytes: the even and the odd bytes as two 16-bit lanes each; int8 lanes
  // are biased by 0x80 so that unsigned order is signed order
  const uint32_t bias = (T(-1) < T(0)) ? 0x80808080u : 0u;
  a ^= bias;
  b ^= bias;
  const uint32_t ae = a & 0x00FF00FFu, ao = (a >> 8) & 0x00FF00FFu;
  const uint32_t be = b & 0x00FF00FFu, bo = (b >> 8) & 0x00FF00FFu;
  const uint32_t e = MAX ? pk_max_u16(ae, be) : pk_min_u16(ae, be);
  const uint32_t o = MAX ? pk_max_u16(ao, bo) : pk_min_u16(ao, bo);
  return (e | (o << 8)) ^ bias;
}

// The transitions of fold_select on every lane of word p whose mask lane is
// set (vw, bgw: the value and background repeated in every lane).
template <typename T, int FUN>
__device__ inline uint32_t fold_word(uint32_t p, uint32_t vw, uint32_t bgw, uint32_t m) {
  constexpr int S = (int)sizeof(T);
  const uint32_t e = lane_eq<S>(p, bgw);  // `== bg` (integers: never NaN)
  uint32_t n;
  if (FUN == 0) n = (e & vw) | (~e & lane_add<S>(p, vw));  // sum
  else if (FUN == 1) n = (e & vw) | (~e & p);              // first
  else if (FUN == 2) n = vw;                               // last
  else if (FUN == 3) n = (e & vw) | (~e & lane_minmax<T, false>(p, vw));  // min: p > v -> v
  else if (FUN == 4) n = (e & vw) | (~e & lane_minmax<T, true>(p, vw));   // max: p < v -> v
  else if (FUN == 5) n = (e & Lanes<S>::one) | (~e & lane_add<S>(p, Lanes<S>::one));  // count
  else n = Lanes<S>::one;                                  // any
  return (m & n) | (~m & p);
}

// value bits repeated in every lane of a word
template <typename T> __device__ inline uint32_t lane_splat(uint64_t bits) {
  return sizeof(T) == 1 ? ((uint32_t)bits & 0xFFu) * 0x01010101u : ((uint32_t)bits & 0xFFFFu) * 0x00010001u;
}

template <typename T> __device__ inline T from_bits(uint64_t b) {
  T v;
  __builtin_memcpy(&v, &b, sizeof(T));
  return v;
}


typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// XOR-combine a value across the four lanes of a quad with DPP quad
// permutations ([1,0,3,2] then [2,3,0,1]): VALU data movement, no LDS.
__device__ inline uint32_t quad_xor_all(uint32_t x) {
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
  return x;
}
__device__ inline uint64_t quad_xor_all64(uint64_t x) {
  return (uint64_t)quad_xor_all((uint32_t)x) | ((uint64_t)quad_xor_all((uint32_t)(x >> 32)) << 32);
}
__device__ inline int32_t quad_max_all(int32_t x) {
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));
  return x;
}

// One 64-byte descriptor from a read-only table (four 16-byte loads).
__device__ inline ItemDesc load_desc(const RZ_CONST ItemDesc* p) {
  const RZ_CONST u32x4* q = (const RZ_CONST u32x4*)p;
  const u32x4 w[4] = {q[0], q[1], q[2], q[3]};
  ItemDesc d;
  __builtin_memcpy(&d, w, sizeof(d));
  return d;
}

// The box (r0, r1, c0, c1) and kind | band word of a descriptor: its first
// 32 bytes (two 16-byte loads), all a box test needs.
struct BoxLayer {
  int32_t r0, r1, c0, c1;
  uint32_t kind_layer;
};
__device__ inline BoxLayer load_box(const RZ_CONST ItemDesc* p) {
  const RZ_CONST u32x4* q = (const RZ_CONST u32x4*)p;
  const u32x4 a = q[0], b = q[1];
  return BoxLayer{(int32_t)a[0], (int32_t)a[1], (int32_t)a[2], (int32_t)a[3], b[3]};
}

// First index of XCD x's contiguous share of n.
__device__ inline uint32_t xcd_unit_lo(uint32_t x, uint32_t n) {
  const uint32_t per = n / NUM_XCD, rem = n % NUM_XCD;
  return x * per + min(x, rem);
}

// Coverage of one polygon item over the 16 rows [Rw, Rw + 16) (rows from
// strip_r1 on excluded) and the columns [C0, C0 + 64 * NWORD): lane
// (rl = lane / 4, q = lane % 4) returns word q of row Rw + rl's mask (lanes
// with q >= NWORD return 0).  Lane layout: the four lanes of a row take every
// 4th active edge; their crossing toggles are XOR-combined with DPP and the
// row mask follows from the note above row_coverage_word in rz_kernels.h.
// `stage` is 64 edges of wave-private LDS.
template <int NWORD>
__device__ inline uint64_t strip_coverage(const PolyEdge* ebase, uint32_t n_e, int32_t Rw,
                                          int32_t strip_r1, int32_t C0, double ncols_f,
                                          PolyEdge* stage, int lane, int wq = -1) {
  const int rl = lane >> 2, q = lane & 3;
  if (wq < 0) wq = q;  // the window word this lane returns (window from C0, NWORD words)
  const int32_t r = Rw + rl;
  const bool row_ok = r < strip_r1;
  uint64_t tog[NWORD];
#pragma unroll
  for (int w = 0; w < NWORD; ++w) tog[w] = 0ull;
  uint32_t lp = 0;
  int32_t amax = -1;
  for (uint32_t cb = 0; cb < n_e; cb += 64) {
    // stage the edges of this chunk that reach the strip, compacted
    // (two 16-byte vector loads and stores: no struct copy through scratch)
    const uint32_t e = cb + (uint32_t)lane;
    uint4 e_lo = make_uint4(0u, 0u, 0u, 0u), e_hi = make_uint4(0u, 0u, 0u, 0u);
    bool act = false;
    if (e < n_e) {
      const uint4* src = reinterpret_cast<const uint4*>(ebase + e);
      e_lo = src[0];
      e_hi = src[1];  // dxdy, ystart (z), yend (w)
      const int32_t ys = (int32_t)e_hi.z, ye = (int32_t)e_hi.w;
      act = ys < strip_r1 && ye > Rw && ys < ye;
    }
    const uint64_t m = __ballot(act);
    if (act) {
      uint4* dst = reinterpret_cast<uint4*>(stage + __popcll(m & ((1ull << lane) - 1ull)));
      dst[0] = e_lo;
      dst[1] = e_hi;
    }
    wave_sync();
    const int n = __popcll(m);
    for (int k0 = 0; k0 < n; k0 += 4) {  // uniform trip count, predicated lanes
      const int k = k0 + q;
      const PolyEdge ee = stage[k < n ? k : n - 1];
      const bool on = k < n && row_ok && ee.ystart <= r && r < ee.yend;
      const double x = crossing(ee, r);
      toggle_crossing_if<NWORD>(on, span_col(x, ncols_f), C0, tog, lp, amax);
    }
    wave_sync();
  }
  lp = quad_xor_all(lp);
  amax = quad_max_all(amax);
#pragma unroll
  for (int w = 0; w < NWORD; ++w) tog[w] = quad_xor_all64(tog[w]);
  return wq < NWORD ? row_coverage_word<NWORD>(tog, lp, amax, C0, wq) : 0ull;
}

// Largest value of v (0..16) over the lanes of the wave, from five ballots.
__device__ inline int32_t wave_max16(int32_t v) {
  int32_t m = 0;
#pragma unroll
  for (int step = 16; step >= 1; step >>= 1)
    if (__any(v >= m + step)) m += step;
  return m;
}

}  // namespace rzk
