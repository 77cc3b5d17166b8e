// Device-side building blocks of the gfx950 tracker: complex arithmetic,
// wave64 lane-group primitives, table-driven evaluation and the group LU.
//
// A "group" is G consecutive lanes of one wave64 that cooperate on ONE path:
// lane gl < n owns equation/row gl (the reference's one-thread-per-row idea,
// GPU-HC/magmaHC/gpu-kernels/kernel_HC_Solver_generalized_3views_4pts.cu:299-300,
// re-laid for wave64: floor(64/G) paths per wave, G need not be a power of 2).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace hcx {
namespace dev {

// Complex value; aligned to its size, so that one 16-byte (FP64) or 8-byte
// (FP32) LDS or global access moves it (no split into two half accesses,
// which would touch every other LDS bank twice in a lane-contiguous gather).
template <class T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

template <class T>
__device__ __forceinline__ cx<T> cmk(T a, T b) {
  cx<T> z;
  z.re = a;
  z.im = b;
  return z;
}
template <class T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) {
  return cmk<T>(a.re + b.re, a.im + b.im);
}
template <class T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) {
  return cmk<T>(a.re - b.re, a.im - b.im);
}
template <class T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a - b*c
template <class T>
__device__ __forceinline__ cx<T> cfnms(cx<T> a, cx<T> b, cx<T> c) {
  return cmk<T>(a.re - b.re * c.re + b.im * c.im, a.im - b.re * c.im - b.im * c.re);
}
// a + b*c
template <class T>
__device__ __forceinline__ cx<T> cfma(cx<T> a, cx<T> b, cx<T> c) {
  return cmk<T>(a.re + b.re * c.re - b.im * c.im, a.im + b.re * c.im + b.im * c.re);
}
template <class T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) {
  return cmk<T>(a.re * s, a.im * s);
}
template <class T>
__device__ __forceinline__ T cabs1(cx<T> a) {
  return fabs(a.re) + fabs(a.im);
}
template <class T>
__device__ __forceinline__ T cnorm2(cx<T> a) {
  return a.re * a.re + a.im * a.im;
}
// Complex division a/b by Smith's algorithm (R. L. Smith, CACM 5(8), 1962):
// scales by the larger component of b to avoid overflow in |b|^2.
template <class T>
__device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const T r = b.im / b.re;
    const T inv = T(1) / (b.re + b.im * r);
    return cmk<T>((a.re + a.im * r) * inv, (a.im - a.re * r) * inv);
  } else {
    const T r = b.re / b.im;
    const T inv = T(1) / (b.re * r + b.im);
    return cmk<T>((a.re * r + a.im) * inv, (a.im * r - a.re) * inv);
  }
}
// Reciprocals from the hardware estimate: v_rcp_f32 (1 ulp) for FP32;
// v_rcp_f64 refined by two Newton-Raphson steps r <- r + r (1 - x r) for FP64.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}
// Complex reciprocal 1/b by Smith's scaling (as cdiv) with the fast real reciprocal.
template <class T>
__device__ __forceinline__ cx<T> crecip(cx<T> b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const T r = b.im * frcp(b.re);
    const T inv = frcp(b.re + b.im * r);
    return cmk<T>(inv, -r * inv);
  } else {
    const T r = b.re * frcp(b.im);
    const T inv = frcp(b.re * r + b.im);
    return cmk<T>(r * inv, -inv);
  }
}

template <class T, class U>
__device__ __forceinline__ cx<T> ccast(cx<U> a) {
  return cmk<T>((T)a.re, (T)a.im);
}

// ------------------------------------------------------------ lane exchange
// ds_bpermute: every lane reads `v` of lane `src` (0..63) of its own wave.
__device__ __forceinline__ int lane_read(int v, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, v); }
__device__ __forceinline__ float lane_read(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <class T>
__device__ __forceinline__ cx<T> lane_read(cx<T> v, int src) {
  return cmk<T>(lane_read(v.re, src), lane_read(v.im, src));
}

// DPP lane permutations (a VALU operand modifier, no LDS round trip) inside
// 16-lane rows: quad_perm xor 1 and xor 2, half-row mirror (lane i <-> 7 - i
// of each 8), row mirror (lane i <-> 15 - i).  Composed in this order they
// reach every lane of an aligned group of 2, 4, 8 or 16 lanes.
constexpr int kDppQuadXor1 = 0xB1;
constexpr int kDppQuadXor2 = 0x4E;
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;
template <int CTRL>
__device__ __forceinline__ int dpp_read(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_read(float v) {
  return __int_as_float(dpp_read<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_read(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_read<CTRL>((int)(b & 0xffffffffll));
  const int hi = dpp_read<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Orders LDS traffic between the phases of a wave.  Kernels run with
// 64-thread workgroups (one wave); a wave's LDS instructions execute in issue
// order, so a wave-scope fence (no memory-counter drain) is enough to keep the
// compiler from moving LDS reads above the writes they depend on.
#define HCX_WAVE_SYNC() __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront")

// Group geometry of one lane.
// A path's lanes: R replicas of G lanes each (R = 1 but in launches of few
// paths).  The replicas hold identical path state and perform identical
// operations, except that each sums every R-th term of the evaluation's
// table rows; a butterfly over the replicas then gives all of them the same
// full sums.
struct Group {
  int G;      // lanes per replica (rows / equations)
  int R;      // replicas per path
  int gpw;    // full groups (paths) per wave
  int grp;    // group index of this lane
  int s;      // replica index of this lane
  int gl;     // lane index inside the replica (row / equation)
  int gbase;  // wave lane of the replica's lane 0
  int wbase;  // wave lane of the group's lane 0 (replica 0's lane 0)
  bool in_grp;// lane belongs to a full group
};

__device__ __forceinline__ Group make_group(int G, int lane, int R = 1) {
  Group g;
  const int W = G * R;
  g.G = G;
  g.R = R;
  g.gpw = 64 / W;
  g.grp = lane / W;
  const int w = lane - g.grp * W;
  g.s = R > 1 ? w / G : 0;
  g.gl = w - g.s * G;
  g.wbase = g.grp * W;
  g.gbase = g.wbase + g.s * G;
  g.in_grp = g.grp < g.gpw;
  return g;
}

// Sum of v over the lanes of the group, returned to every lane of the group.
template <class T>
__device__ __forceinline__ T group_sum(T v, const Group& g, int lane) {
  T s = v;
  for (int off = 1; off < g.G; off <<= 1) {
    const int src = lane + off < 64 ? lane + off : 63;
    const T o = lane_read(s, src);
    if (g.gl + off < g.G) s += o;
  }
  return lane_read(s, g.gbase);
}

// ------------------------------------------------------------ device tables
// One table term: scalar * c[coef] * prod x[var_q]; var index n reads the
// constant 1 stored at x[n] (reference table convention,
// dev-eval-indxing-generalized_3views_4pts.cuh:70-120).
struct __align__(16) Term {
  int s;
  int c;
  uint32_t v03;  // variable indices 0..3, 8 bits each
  uint32_t v47;  // variable indices 4..7
};

__device__ __forceinline__ int term_var(const Term& t, int q) {
  return q < 4 ? (int)((t.v03 >> (8 * q)) & 0xffu) : (int)((t.v47 >> (8 * (q - 4))) & 0xffu);
}

// Compact 8-byte form of a term for tables held in LDS: x = coefficient slot
// (bits 0-15) | scalar (bits 16-31, two's complement); y = variable indices
// 0..3, 8 bits each (terms of degree <= 4).
__device__ __forceinline__ int cterm_c(uint2 t) { return (int)(t.x & 0xffffu); }
__device__ __forceinline__ int cterm_s(uint2 t) { return (int)t.x >> 16; }
__device__ __forceinline__ int cterm_var(uint2 t, int q) { return (int)((t.y >> (8 * q)) & 0xffu); }

// Compact 4-byte form for tables of degree <= 2 (each term has at most two
// variables): coefficient slot (bits 0-11) | scalar (bits 12-15, two's
// complement, -8..7) | variable 0 (bits 16-23) | variable 1 (bits 24-31).
__device__ __forceinline__ int cterm_c(uint32_t t) { return (int)(t & 0xfffu); }
__device__ __forceinline__ int cterm_s(uint32_t t) { return (int)(t << 16) >> 28; }
__device__ __forceinline__ int cterm_var(uint32_t t, int q) { return (int)((t >> (16 + 8 * q)) & 0xffu); }

// The compact term record of a table of degree DEG.
template <int DEG>
using CTerm = typename std::conditional<(DEG <= 2), uint32_t, uint2>::type;

}  // namespace dev
}  // namespace hcx
