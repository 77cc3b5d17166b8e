// Tracker kernel templates (shared by the per-instance translation units).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "hcx_device.hpp"
#include "hcx_flags.h"

namespace hcx {

using dev::cx;
using dev::cmk;

constexpr int kMaxSeg = 8;
constexpr int kMaxN = 32;
// terms whose table records are requested together when the tables are read
// from global memory
#ifndef HCX_TERM_CHUNK
#define HCX_TERM_CHUNK 4
#endif
constexpr int kTermChunk = HCX_TERM_CHUNK;
// terms of an LDS-table row per unrolled loop iteration: two in FP64 (the
// loads in flight then fit a budget of 168 VGPRs, three waves per SIMD),
// four in FP32
template <class T>
constexpr int kTermUnroll = sizeof(T) == 8 ? 2 : 4;

// Lanes per path of the size-specialised instance for n: n, except n = 30,
// which runs on aligned 32-lane groups (still two paths per wave) so that
// its group reductions can use DPP row permutations.
constexpr int exact_lanes(int n) { return n == 30 ? 32 : n; }
constexpr bool is_pow2(int g) { return g > 0 && (g & (g - 1)) == 0; }

// Output pointers of a segment (cold): kept in device memory and read only
// when a path finishes, so they do not occupy scalar registers in the loop.
struct SegOut {
  double2* roots;              // [P][roots_stride]
  uint32_t* flags;
  int* steps;
  int* newton;
  int* refine;
  double* tfin;
  double* dtfin;
  double* resid;
  uint8_t* conv8;
  uint8_t* inf8;
  double2* debug;
  unsigned long long* inst_done;
  int* inst_left;
  int roots_stride;
};

// Per-launch device workspace of one segment (path queue, output
// descriptor, per-instance remaining-path counters, re-track path list).
struct LaunchWork {
  unsigned int* counter;
  SegOut* out;
  int* inst_left;              // capacity inst_cap
  int* plist;                  // re-track path list (capacity plist_cap)
  unsigned int* plist_count;   // [1]
  long long inst_cap;
  long long plist_cap;
};

// Hot, wave-uniform description of one segment (kernel arguments).
struct Seg {
  int n, C, M, K;
  int hxT, htT;
  int G;
  int R;                       // replicas per path (lanes per path = G * R; FP64 loop)
  int lds_group;               // bytes of one group's LDS region
  int start_stride;
  int I;
  long long P;
  const dev::Term* hx;         // [n var][hxT][n eq]
  const dev::Term* ht;         // [htT][n eq]
  const int* hx_cnt;           // [n] real terms per Jacobian column (max over equations)
  const double2* start;        // [K][start_stride]
  const double2* chx;          // [I][C+1][M+1]
  const double2* cht;          // [I][C+1][M]
  const int* cslot;            // [C+1] LDS slot of coefficient j (the tables hold slots)
  const int* cinv;             // [C+1] coefficient of LDS slot q
  const uint32_t* ctab;        // compact hx then ht tables (LDS-table instances; 4-byte terms
                               // for DEG <= 2, else 8-byte)
  unsigned int* counter;
  const SegOut* out;           // device copy written by init_kernel
  const int* plist;            // optional: track only these paths (count in *plist_count)
  const unsigned int* plist_count;
};

struct Opts {
  int max_steps, max_corr, incr, refine_iters;
  int newton32;      // mixed: FP32 corrector before the end zone
  double dt0, t_tol, dt_min, corr_tol, inf_tol, refine_tol, real_tol;
  double real_ref_tol;  // the reference's real test on the tracked root
  double ez_enter;   // enter the end zone when |1 - t| <= end_zone + 1e-7
  double ez_land;    // before the end zone, steps land on 1 - end_zone
};

struct Launch {
  Seg seg[kMaxSeg];
  int nseg;
  int first_block[kMaxSeg + 1];
  int persistent;
  int tab_lds;                 // bytes of dynamic LDS before the group regions (compact tables)
  int dot_lds;                 // FP64 mode, M = 1: stage dc/dt in LDS (else read it from global)
  int wpb;                     // waves per block (> 1 only for one-segment launches: the
                               // waves of a block share one copy of the compact tables)
  Opts o;
};

// The options and segment descriptors in the kernel-argument (constant)
// address space: loads through these types are scalar loads.
typedef const __attribute__((address_space(4))) Opts KOpts;
typedef const __attribute__((address_space(4))) Seg KSeg;

namespace dev {

// The launch's options where the kernel argument holds them (the
// descriptor is the kernels' only argument: offset 0 of the segment).
__device__ __forceinline__ const KOpts* launch_opts() {
  return (const KOpts*)((const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr() +
                        offsetof(Launch, o));
}

// Segment si's descriptor where the kernel argument holds it.
__device__ __forceinline__ const KSeg* launch_seg(int si) {
  return (const KSeg*)((const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr() +
                       offsetof(Launch, seg) + (size_t)si * sizeof(Seg));
}

// A kernel-argument view for one iteration of a path loop: the pointer
// passes through an empty asm keyed on the iteration count, so the fields
// are re-read (scalar loads) where an iteration uses them instead of
// staying live in scalar registers across the loop, where they would be
// spilled to vector-register lanes and reloaded with VALU instructions.
template <class T>
__device__ __forceinline__ T* iteration_view(T* p, int iter) {
  asm("" : "+s"(p) : "s"(iter));
  return p;
}

// ------------------------------------------------------------ evaluation
// sum_k s_k * coef[c_k] * prod_{d < NV} x[var_kd] over the cnt terms of one
// table row (records n apart from tp, full Term records in global memory or
// in an LDS copy).  The records of a chunk of terms are all requested before
// the first is used, so their latencies overlap; remap (when not null) maps
// a record's coefficient number to its index in coef.  Terms k0, k0 + ks,
// k0 + 2 ks, ... only (the replica's share of the row).
template <int NV, class T>
__device__ __forceinline__ cx<T> term_row_sum(const Term* tp, int cnt, int n, const cx<T>* sx, const cx<T>* coef,
                                              const int* remap, int k0 = 0, int ks = 1) {
  cx<T> acc = cmk<T>(0, 0);
  for (int kb = k0; kb < cnt; kb += kTermChunk * ks) {
    Term tm[kTermChunk];
    int ci[kTermChunk];
#pragma unroll
    for (int q = 0; q < kTermChunk; ++q)
      if (kb + q * ks < cnt) tm[q] = tp[(size_t)(kb + q * ks) * n];
#pragma unroll
    for (int q = 0; q < kTermChunk; ++q)
      if (kb + q * ks < cnt) ci[q] = remap ? remap[tm[q].c] : tm[q].c;
#pragma unroll
    for (int q = 0; q < kTermChunk; ++q) {
      if (kb + q * ks < cnt) {
        cx<T> v = cscale(coef[ci[q]], (T)tm[q].s);
#pragma unroll
        for (int d = 0; d < NV; ++d) v = cmul(v, sx[term_var(tm[q], d)]);
        acc = cadd(acc, v);
      }
    }
  }
  return acc;
}

// Row gl of dH/dx (dev-eval-indxing-generalized_3views_4pts.cuh:70-86).
// LT: the tables are the compact copy in LDS at lt (hx part first); else
// full Term records, from the LDS copy at lt when tl, else global memory.
// Terms k0, k0 + ks, ... of each entry only (a replica's share).
template <int NMAX, int DEG, class T, bool LT, class SegT>
__device__ __forceinline__ void eval_jac_row(const SegT& S, int n, int gle, const cx<T>* sx, const cx<T>* sc,
                                             cx<T> (&A)[NMAX], const void* lt, bool tl = false, int k0 = 0,
                                             int ks = 1) {
#pragma unroll
  for (int j = 0; j < NMAX; ++j) {
    cx<T> acc = cmk<T>(0, 0);
    if (j < n) {
      const int cnt = S.hx_cnt[j];
      if constexpr (LT) {
        const CTerm<DEG>* tp = static_cast<const CTerm<DEG>*>(lt) + (size_t)j * S.hxT * n + gle;
#pragma unroll kTermUnroll<T>
        for (int k = k0; k < cnt; k += ks) {
          const CTerm<DEG> tm = tp[(size_t)k * n];
          cx<T> v = cscale(sc[cterm_c(tm)], (T)cterm_s(tm));
#pragma unroll
          for (int q = 0; q < DEG - 1; ++q) v = cmul(v, sx[cterm_var(tm, q)]);
          acc = cadd(acc, v);
        }
      } else {
        const size_t off = (size_t)j * S.hxT * n + gle;
        if (tl)
          acc = term_row_sum<DEG - 1, T>(static_cast<const Term*>(lt) + off, cnt, n, sx, sc, nullptr, k0, ks);
        else
          acc = term_row_sum<DEG - 1, T>(S.hx + off, cnt, n, sx, sc, nullptr, k0, ks);
      }
    }
    A[j] = acc;
  }
}

// Row gl of sum_k s * coef[c] * prod x over the dHdt table: H when coef = c(t),
// dH/dt when coef = dc/dt (…cuh:89-120).
// coef is an LDS table indexed by slot, or (remap != nullptr) a global table
// indexed by coefficient.
template <int DEG, class T, bool LT, class SegT>
__device__ __forceinline__ cx<T> eval_t_row(const SegT& S, int n, int gle, const cx<T>* sx, const cx<T>* coef,
                                            const void* lt, const int* remap = nullptr) {
  cx<T> acc = cmk<T>(0, 0);
  if constexpr (LT) {
    const CTerm<DEG>* tp = static_cast<const CTerm<DEG>*>(lt) + (size_t)n * S.hxT * n + gle;
#pragma unroll kTermUnroll<T>
    for (int k = 0; k < S.htT; ++k) {
      const CTerm<DEG> tm = tp[(size_t)k * n];
      const int c = cterm_c(tm);
      cx<T> v = cscale(coef[remap ? remap[c] : c], (T)cterm_s(tm));
#pragma unroll
      for (int q = 0; q < DEG; ++q) v = cmul(v, sx[cterm_var(tm, q)]);
      acc = cadd(acc, v);
    }
  } else {
    // records (and, with remap, the coefficient numbers) of a chunk of terms
    // are requested together, as in eval_jac_row
    const Term* tp = S.ht + gle;
    for (int k0 = 0; k0 < S.htT; k0 += kTermChunk) {
      Term tm[kTermChunk];
      int ci[kTermChunk];
#pragma unroll
      for (int q = 0; q < kTermChunk; ++q)
        if (k0 + q < S.htT) tm[q] = tp[(size_t)(k0 + q) * n];
#pragma unroll
      for (int q = 0; q < kTermChunk; ++q)
        if (k0 + q < S.htT) ci[q] = remap ? remap[tm[q].c] : tm[q].c;
#pragma unroll
      for (int q = 0; q < kTermChunk; ++q) {
        if (k0 + q < S.htT) {
          cx<T> v = cscale(coef[ci[q]], (T)tm[q].s);
#pragma unroll
          for (int d = 0; d < DEG; ++d) v = cmul(v, sx[term_var(tm[q], d)]);
          acc = cadd(acc, v);
        }
      }
    }
  }
  return acc;
}

// c_j(t) and dc_j/dt(t) of order MC in t into LDS slot cslot[j] (Horner
// from the highest term, as the general loop of stage_coeffs): the terms of
// c_j are requested before the first Horner step uses one, then those of
// dc_j/dt, so two memory latencies per coefficient instead of one per term.
template <int MC, class T, class SegT>
__device__ __forceinline__ void stage_coeffs_m(const SegT& S, const Group& g, const double2* hx, const double2* ht,
                                               double t, cx<T>* sc, cx<T>* sd, bool with_dot) {
  const int C1 = S.C + 1;
#pragma unroll 1
  for (int j = g.gl + g.s * g.G; j < C1; j += g.G * g.R) {
    const int q = S.cslot[j];
    {
      const double2* ha = hx + (size_t)j * (MC + 1);
      double2 a[MC + 1];
#pragma unroll
      for (int m = 0; m <= MC; ++m) a[m] = ha[m];
      double cr = a[MC].x, ci = a[MC].y;
#pragma unroll
      for (int m = MC - 1; m >= 0; --m) {
        const double nr = cr * t + a[m].x, ni = ci * t + a[m].y;
        cr = nr;
        ci = ni;
      }
      sc[q] = cmk<T>((T)cr, (T)ci);
    }
    if (with_dot) {
      const double2* hb = ht + (size_t)j * MC;
      double2 b[MC];
#pragma unroll
      for (int m = 0; m < MC; ++m) b[m] = hb[m];
      double dr = b[MC - 1].x, di = b[MC - 1].y;
#pragma unroll
      for (int m = MC - 2; m >= 0; --m) {
        const double nr = dr * t + b[m].x, ni = di * t + b[m].y;
        dr = nr;
        di = ni;
      }
      sd[q] = cmk<T>((T)dr, (T)di);
    }
  }
}

// c_j(t) and dc_j/dt(t) of order MC in t (MC >= 1) into LDS slot cslot[j]
// for Q of the lane's coefficients at a time: every term of the Q
// coefficients (both tables) is requested before the first Horner step, so
// one memory latency per Q coefficients (the latency instances, whose
// register budget is a whole SIMD's).  Surplus entries of the last chunk
// reload coefficient C and store nothing.
template <int MC, int Q, class T, class SegT>
__device__ __forceinline__ void stage_coeffs_q(const SegT& S, const Group& g, const double2* hx, const double2* ht,
                                               double t, cx<T>* sc, cx<T>* sd, bool with_dot) {
  const int C1 = S.C + 1, st = g.G * g.R;
#pragma unroll 1
  for (int jb = g.gl + g.s * g.G; jb < C1; jb += Q * st) {
    double2 a[Q][MC + 1], b[Q][MC];
    int q[Q];
#pragma unroll
    for (int u = 0; u < Q; ++u) {
      const int j = min(jb + u * st, C1 - 1);
      q[u] = S.cslot[j];
#pragma unroll
      for (int m = 0; m <= MC; ++m) a[u][m] = hx[(size_t)j * (MC + 1) + m];
      if (with_dot) {
#pragma unroll
        for (int m = 0; m < MC; ++m) b[u][m] = ht[(size_t)j * MC + m];
      }
    }
#pragma unroll
    for (int u = 0; u < Q; ++u) {
      if (jb + u * st < C1) {
        double cr = a[u][MC].x, ci = a[u][MC].y;
#pragma unroll
        for (int m = MC - 1; m >= 0; --m) {
          const double nr = cr * t + a[u][m].x, ni = ci * t + a[u][m].y;
          cr = nr;
          ci = ni;
        }
        sc[q[u]] = cmk<T>((T)cr, (T)ci);
        if (with_dot) {
          double dr = b[u][MC - 1].x, di = b[u][MC - 1].y;
#pragma unroll
          for (int m = MC - 2; m >= 0; --m) {
            const double nr = dr * t + b[u][m].x, ni = di * t + b[u][m].y;
            dr = nr;
            di = ni;
          }
          sd[q[u]] = cmk<T>((T)dr, (T)di);
        }
      }
    }
  }
}

// Coefficient polynomials in t of the group's instance into LDS
// (eval_parameter_homotopy, …cuh:34-67), computed in FP64 and stored as T.
// Coefficient j goes to slot cslot[j]: slots follow the order in which the
// rows of a group read them (term by term), so the lanes of a group that read
// term k of their rows together hit consecutive LDS words instead of words one
// equation apart (a bank conflict per lane pair when equations have many terms).
// The replicas of a group share its LDS region and stage every R-th
// coefficient each.  HOIST: orders M <= 4 request all terms of a coefficient
// before its Horner steps (stage_coeffs_m; the FP64 loop, whose register
// budget has room for them).  CH > 0: orders M <= 4 stage CH coefficients
// per lane at a time (stage_coeffs_q; the latency instances).
template <class T, bool HOIST = false, int CH = 0, class SegT>
__device__ __forceinline__ void stage_coeffs(const SegT& S, const Group& g, bool act, int inst, double t,
                                             unsigned char* scb, bool with_dot) {
  if (!act) return;
  cx<T>* sc = reinterpret_cast<cx<T>*>(scb);
  cx<T>* sd = sc + (S.C + 1);
  const int C1 = S.C + 1, M = S.M;
  const double2* hx = S.chx + (size_t)inst * C1 * (M + 1);
  const double2* ht = S.cht + (size_t)inst * C1 * M;
  if constexpr (CH > 0) {
    // chunks of CH coefficients up to order 2, CH / 2 above (registers)
    switch (M) {
      case 1: return stage_coeffs_q<1, CH, T>(S, g, hx, ht, t, sc, sd, with_dot);
      case 2: return stage_coeffs_q<2, CH, T>(S, g, hx, ht, t, sc, sd, with_dot);
      case 3: return stage_coeffs_q<3, (CH + 1) / 2, T>(S, g, hx, ht, t, sc, sd, with_dot);
      case 4: return stage_coeffs_q<4, (CH + 1) / 2, T>(S, g, hx, ht, t, sc, sd, with_dot);
      default: break;
    }
  }
  if (M == 1) {
    // c(t) = c0 + t c1, dc/dt = Ht[j][0]
#pragma unroll kTermUnroll<T>
    for (int j = g.gl + g.s * g.G; j < C1; j += g.G * g.R) {
      const int q = S.cslot[j];
      const double2 a = hx[2 * (size_t)j], b = hx[2 * (size_t)j + 1];
      sc[q] = cmk<T>((T)(a.x + t * b.x), (T)(a.y + t * b.y));
      if (with_dot) {
        const double2 d = ht[j];
        sd[q] = cmk<T>((T)d.x, (T)d.y);
      }
    }
    return;
  }
  if constexpr (HOIST) {
    switch (M) {
      case 2: return stage_coeffs_m<2, T>(S, g, hx, ht, t, sc, sd, with_dot);
      case 3: return stage_coeffs_m<3, T>(S, g, hx, ht, t, sc, sd, with_dot);
      case 4: return stage_coeffs_m<4, T>(S, g, hx, ht, t, sc, sd, with_dot);
      default: break;
    }
  }
  for (int j = g.gl + g.s * g.G; j < C1; j += g.G * g.R) {
    double2 a = hx[(size_t)j * (M + 1) + M];
    double cr = a.x, ci = a.y;
    for (int m = M - 1; m >= 0; --m) {
      const double2 b = hx[(size_t)j * (M + 1) + m];
      const double nr = cr * t + b.x, ni = ci * t + b.y;
      cr = nr;
      ci = ni;
    }
    const int q = S.cslot[j];
    sc[q] = cmk<T>((T)cr, (T)ci);
    if (with_dot) {
      double2 d = ht[(size_t)j * M + M - 1];
      double dr = d.x, di = d.y;
      for (int m = M - 2; m >= 0; --m) {
        const double2 b = ht[(size_t)j * M + m];
        const double nr = dr * t + b.x, ni = di * t + b.y;
        dr = nr;
        di = ni;
      }
      sd[q] = cmk<T>((T)dr, (T)di);
    }
  }
}

template <class T>
__device__ __forceinline__ void publish_x(cx<T>* sx, cx<double> x, const Group& g, int n, bool act, bool row) {
  if (act && row) sx[g.gl] = ccast<T>(x);
  if (act && g.gl == 0) sx[n] = cmk<T>(T(1), T(0));
}

template <class T>
__device__ __forceinline__ T group_max(T v, const Group& g) {
  for (int s = 1; s < g.G; s <<= 1) {
    int src = g.gl + s;
    src = (src >= g.G ? src - g.G : src) + g.gbase;
    const T o = lane_read(v, src);
    v = o > v ? o : v;
  }
  return v;
}

// Argmax step of the pivot search: keep the larger candidate, ties to the
// smaller key (position label << 6 | lane).
template <class T>
__device__ __forceinline__ void piv_take(T& v, int& key, T ov, int ok) {
  const bool take = (ov > v) | ((ov == v) & (ok < key));
  v = take ? ov : v;
  key = take ? ok : key;
}

// Pivot argmax over the group: GC = compile-time lanes per path when that is
// a power of two (aligned groups; DPP inside 16-lane rows, one lane exchange
// per further doubling), else 0 (ring of ceil(log2 G) lane exchanges).
template <int GC, class T>
__device__ __forceinline__ void group_argmax(T& v, int& key, const Group& g) {
  if constexpr (GC > 0) {
    static_assert(is_pow2(GC) && GC <= 64, "aligned power-of-two groups");
    if constexpr (GC >= 2) piv_take(v, key, dpp_read<kDppQuadXor1>(v), dpp_read<kDppQuadXor1>(key));
    if constexpr (GC >= 4) piv_take(v, key, dpp_read<kDppQuadXor2>(v), dpp_read<kDppQuadXor2>(key));
    if constexpr (GC >= 8) piv_take(v, key, dpp_read<kDppHalfMirror>(v), dpp_read<kDppHalfMirror>(key));
    if constexpr (GC >= 16) piv_take(v, key, dpp_read<kDppRowMirror>(v), dpp_read<kDppRowMirror>(key));
    const int lane = g.gbase + g.gl;
    if constexpr (GC >= 32) piv_take(v, key, lane_read(v, lane ^ 16), lane_read(key, lane ^ 16));
    if constexpr (GC >= 64) piv_take(v, key, lane_read(v, lane ^ 32), lane_read(key, lane ^ 32));
  } else {
    for (int s = 1; s < g.G; s <<= 1) {
      int src = g.gl + s;
      src = (src >= g.G ? src - g.G : src) + g.gbase;
      piv_take(v, key, lane_read(v, src), lane_read(key, src));
    }
  }
}

// Sums over the group's rows of two per-row values (the corrector's and the
// refinement's |d|^2 and |x|^2), returned to every lane of the group.  Aligned
// power-of-two groups use lane exchanges; other group sizes publish the
// pairs in the group's LDS slots `q` and every lane adds them in row order
// (one LDS round trip).
template <int NMAX, int GC>
__device__ __forceinline__ void group_sum2(double& a, double& b, const Group& g, int n, bool row, int lane,
                                           cx<double>* q) {
  if constexpr (GC > 0) {
    a = group_sum(a, g, lane);
    b = group_sum(b, g, lane);
  } else {
    if (row) q[g.gl] = cmk<double>(a, b);
    HCX_WAVE_SYNC();
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      if (j < n) {
        const cx<double> v = q[j];
        sa += v.re;
        sb += v.im;
      }
    }
    HCX_WAVE_SYNC();
    a = sa;
    b = sb;
  }
}

// ------------------------------------------------------------ group solve
// Solves A x = b for the group's n x n system, row r held by lane r (A[j] =
// A[r][j]).  Returns x_gl to lane gl.  Pivoting is the reference's partial
// pivoting (dev-cgesv-batched-small.cuh:51-95): cabs1, the first maximum in
// current row order (rows carry a position label swapped like a physical
// row interchange), multipliers by the pivot reciprocal, a zero pivot column
// skipped and reported through `singular`.  Every step eliminates the pivot
// column from the rows ABOVE the pivot as well (Gauss-Jordan), in place of the
// reference's back substitution (:97-106): on a lane group the lanes of those
// rows would only wait during the step, so the extra row updates cost no
// issue cycles, while the n-step dependent chain of the back substitution
// (a lane exchange per step) and the need to keep U alive for it disappear
// (n = 30 FP64: 157 instead of about 290 registers).  The candidate rows of
// each pivot search receive exactly the updates of the LU, so the pivot
// sequence is the LU's.  At the end the lane holding position i has
// x_i = b_i / U_ii; one LDS exchange returns x_gl to lane gl.  A zero pivot
// makes every component NaN (the reference's division by a zero U_ii gives
// NaN in that component and in the rows that depend on it; the tracker only
// reacts to whether any component is not finite).
template <int NMAX, int GC, class T>
__device__ __forceinline__ cx<T> group_lu_solve(cx<T> (&A)[NMAX], cx<T> b, const Group& g, int n, bool row,
                                                cx<T>* sp, bool& singular) {
  int label = g.gl;
  cx<T> own_rinv = cmk<T>(T(0), T(0));  // 1 / U_ii of the row this lane ends up holding
  singular = false;
  // sp shares the group's LDS slots with x: the evaluation's reads of x come first
  HCX_WAVE_SYNC();
#pragma unroll
  for (int k = 0; k < NMAX; ++k) {
    if (k < n) {
      T v = (row && label >= k) ? cabs1(A[k]) : T(-1);
      if (v != v) v = T(INFINITY);
      int blabel;
      bool piv, zero;
      if constexpr (GC > 0) {
        int key = (label << 6) | g.gl;
        group_argmax<GC>(v, key, g);
        blabel = key >> 6;
        zero = !(v > T(0));
        piv = g.gl == (key & 63);
      } else {
        // other group sizes: the candidates publish their magnitudes by
        // position label and every lane scans them in current row order
        // (one LDS round trip instead of a ring of lane exchanges)
        T* mag = reinterpret_cast<T*>(sp);
        if (row && label >= k) mag[label] = v;
        HCX_WAVE_SYNC();
        T best = mag[k];
        blabel = k;
#pragma unroll
        for (int j = k + 1; j < NMAX; ++j) {
          if (j < n) {
            const T mj = mag[j];
            if (mj > best) {
              best = mj;
              blabel = j;
            }
          }
        }
        HCX_WAVE_SYNC();
        zero = !(best > T(0));
        piv = row && label == blabel;
      }
      singular = singular || zero;
      if (g.in_grp && piv) {
#pragma unroll
        for (int j = k; j < NMAX; ++j)
          if (j < n) sp[j] = A[j];
        sp[NMAX] = b;
      }
      HCX_WAVE_SYNC();
      if (piv) label = k;
      else if (label == k) label = blabel;
      const cx<T> rinv = zero ? cmk<T>(T(1), T(0)) : crecip(sp[k]);
      if (piv) own_rinv = rinv;
      if (row && !piv && !zero) {
        const cx<T> l = cmul(A[k], rinv);
#pragma unroll
        for (int j = k + 1; j < NMAX; ++j)
          if (j < n) A[j] = cfnms(A[j], l, sp[j]);
        b = cfnms(b, l, sp[NMAX]);
      }
      HCX_WAVE_SYNC();
    }
  }
  const cx<T> xi = cmul(b, own_rinv);
  if (row) sp[label] = xi;
  HCX_WAVE_SYNC();
  cx<T> sol = row ? sp[g.gl] : cmk<T>(T(0), T(0));
  HCX_WAVE_SYNC();
  if (singular) sol = cmk<T>(T(NAN), T(NAN));
  return sol;
}

// Copies a segment's compact term tables to the start of the dynamic LDS.
// Every wave of the block takes part (one copy per block) and the block
// barrier orders every wave's table reads after the copy; blocks of several
// waves run one segment only, so each wave reaches the barrier once.
template <int DEG>
__device__ __forceinline__ const void* stage_tables(const Seg& S, int n, unsigned char* smem) {
  uint32_t* d = reinterpret_cast<uint32_t*>(smem);
  const int words = (S.hxT * n + S.htT) * n * (int)(sizeof(CTerm<DEG>) / 4);
  for (int i = (int)threadIdx.x; i < words; i += (int)blockDim.x) d[i] = S.ctab[i];
  __syncthreads();
  return d;
}

// The full Term records of the generic instances (dH/dx part, then dH/dt
// part, in their global layout) copied into LDS by the block's waves; for
// launches of few paths (latency bound: no other waves hide the global
// latency of the evaluation's record loads).  Same barrier rule as
// stage_tables.
__device__ __forceinline__ const void* stage_terms(const Seg& S, int n, unsigned char* smem) {
  uint4* d = reinterpret_cast<uint4*>(smem);
  const int nhx = S.hxT * n * n, nht = S.htT * n;
  const uint4* hx = reinterpret_cast<const uint4*>(S.hx);
  const uint4* ht = reinterpret_cast<const uint4*>(S.ht);
  for (int i = (int)threadIdx.x; i < nhx; i += (int)blockDim.x) d[i] = hx[i];
  for (int i = (int)threadIdx.x; i < nht; i += (int)blockDim.x) d[nhx + i] = ht[i];
  __syncthreads();
  return d;
}

// Per-group view of the LDS region.  x (read by the evaluation) and the
// pivot-row buffer (written by the LU that follows it) are never live at the
// same time, so they share the first NMAX + 1 slots; group_lu_solve starts
// with a wave-scope fence that keeps the evaluation's reads before its writes.
template <int NMAX>
struct Lds {
  unsigned char* base;
  __device__ cx<double>* x64() const { return reinterpret_cast<cx<double>*>(base); }
  __device__ cx<double>* piv64() const { return x64(); }
  __device__ unsigned char* coef() const { return reinterpret_cast<unsigned char*>(x64() + (NMAX + 1)); }
  template <class T>
  __device__ cx<T>* x() const { return reinterpret_cast<cx<T>*>(base); }
  template <class T>
  __device__ cx<T>* piv() const { return reinterpret_cast<cx<T>*>(piv64()); }
};

// Newton corrector at fixed t in precision T (…4pts.cu:189-223).  The
// coefficients c(t) must already be in the group's LDS.  `run` enters as the
// set of groups to correct; ok / inf return the reference's r_isSuccessful /
// r_isInfFail.
template <int NMAX, int GC, int DEG, class T, bool LT, class OT, class SegT>
__device__ __forceinline__ void newton_pass(const SegT& S, int n, const OT& o, const Group& g, const Lds<NMAX>& L,
                                            const void* lt, bool run, bool row, cx<double>& x, int& nnewt, bool& ok,
                                            bool& inf) {
  ok = false;
  inf = false;
  const int gle = row ? g.gl : 0;
  cx<T>* sx = L.template x<T>();
  const cx<T>* sc = reinterpret_cast<const cx<T>*>(L.coef());
  for (int it = 0; it < o.max_corr; ++it) {
    if (!__any(run)) break;
    publish_x<T>(sx, x, g, n, run, row);
    HCX_WAVE_SYNC();
    cx<T> A[NMAX];
    eval_jac_row<NMAX, DEG, T, LT>(S, n, gle, sx, sc, A, lt);
    const cx<T> h = eval_t_row<DEG, T, LT>(S, n, gle, sx, sc, lt);
    bool sing;
    const cx<T> d = group_lu_solve<NMAX, GC, T>(A, h, g, n, row, L.template piv<T>(), sing);
    if (run && row) x = csub(x, ccast<double>(d));
    double rs = row ? (double)cnorm2(d) : 0.0;
    double rc = row ? cnorm2(x) : 0.0;
    rs = group_sum(rs, g, g.gbase + g.gl);
    rc = group_sum(rc, g, g.gbase + g.gl);
    if (run) {
      ++nnewt;
      ok = rs < o.corr_tol * rc;
      inf = rc > o.inf_tol;
      if (ok || inf) run = false;
    }
  }
}

}  // namespace dev

using namespace dev;

// One RK4 predictor step in precision T for the groups in `act`
// (…4pts.cu:139-183): stages at (x0, t), (x0 + h k1, t + h), (x0 + h k2, t + h),
// (x0 + dt k3, t + dt); returns k1 + 2 k2 + 2 k3 + k4 (component gl).
// sd_ext: dc/dt table of the group's instance in global memory (constant in t,
// M = 1) or nullptr to stage dc/dt(t) in LDS next to c(t).
template <int NMAX, int GC, int DEG, class T, bool LT, class SegT>
__device__ __forceinline__ cx<double> predict(const SegT& S, int n, const Group& g, const Lds<NMAX>& L, const void* lt,
                                              bool act, bool row, int gle, int inst, double ts, double dt,
                                              cx<double> xl, const cx<T>* sd_ext) {
  const double h = 0.5 * dt;
  cx<double> acc = cmk<double>(0, 0), kprev = cmk<double>(0, 0);
  cx<T>* sx = L.template x<T>();
  const cx<T>* sc = reinterpret_cast<const cx<T>*>(L.coef());
  const cx<T>* sd = sd_ext ? sd_ext : sc + (S.C + 1);
#pragma unroll 1
  for (int st = 0; st < 4; ++st) {
    const double a = st == 0 ? 0.0 : (st == 3 ? dt : h);
    if (st != 2) stage_coeffs<T>(S, g, act, inst, ts + a, L.coef(), sd_ext == nullptr);
    const cx<double> xs = cmk<double>(xl.re + a * kprev.re, xl.im + a * kprev.im);
    publish_x<T>(sx, xs, g, n, act, row);
    HCX_WAVE_SYNC();
    cx<T> A[NMAX];
    eval_jac_row<NMAX, DEG, T, LT>(S, n, gle, sx, sc, A, lt);
    const cx<T> bt = eval_t_row<DEG, T, LT>(S, n, gle, sx, sd, lt, sd_ext ? S.cinv : nullptr);
    bool sing;
    const cx<T> k = group_lu_solve<NMAX, GC, T>(A, cmk<T>(-bt.re, -bt.im), g, n, row, L.template piv<T>(), sing);
    kprev = ccast<double>(k);
    const double w = (st == 1 || st == 2) ? 2.0 : 1.0;
    acc = cmk<double>(acc.re + w * kprev.re, acc.im + w * kprev.im);
  }
  return acc;
}

// Mixed-precision lane-group path loop (MODE 1; the FP64 mode runs
// run_segment_fp64), phase-synchronous: FP32 RK4
// predictor (Jacobian, dH/dt and LU of the four stages), FP64 Newton
// corrector (FP32 before the end zone when o.newton32), FP64 refinement; a
// path that ends without a refined root is re-tracked from t = 0 entirely in
// FP64 (escalation): by the same group (KESC: the generic instances and
// n = 30), or, for the size-specialised instances up to n = 16, by the FP64
// pass after the launch, which keeps the FP64 predictor out of the kernel and
// with it the registers of a second predictor.  The pass skips paths already
// escalated, so either way each failed path is re-tracked once.
template <int NMAX, int DEG, int MODE, bool EXACT>
__device__ __forceinline__ void run_segment(const Launch& LN, int si, unsigned char* smem, int lane) {
  static_assert(MODE == 1, "the FP64 mode runs run_segment_fp64");
  // local copies: the fields are wave-uniform and live in SGPRs (scalar
  // kernarg loads) instead of being re-read through a generic pointer
  Seg S = LN.seg[si];
  constexpr int GE = exact_lanes(NMAX);
  constexpr bool KESC = !(EXACT && NMAX <= 16);
  if (EXACT) {
    // size-specialised instance: n = NMAX and G are compile-time constants,
    // so every loop bound and the lane-group geometry fold away
    S.n = NMAX;
    S.G = GE;
  }
  // compile-time power-of-two lanes per path: DPP group reductions
  constexpr int GC = (EXACT && is_pow2(GE)) ? GE : 0;
  const KOpts* const opts_arg = launch_opts();
  const KSeg* const seg_arg = launch_seg(si);
  int iter = 0;
  const Group g = make_group(S.G, lane);
  const int n = S.n;
  const bool row = g.in_grp && g.gl < n;
  const int gle = row ? g.gl : 0;
  // Size-specialised instances up to n = 16, and n = 30 with quadratic
  // terms, read the segment's tables from a compact copy in LDS (at the start
  // of the dynamic LDS, LN.tab_lds bytes) instead of global memory: the term
  // loads head every evaluation's dependency chain, and LDS returns them
  // several times sooner.
  constexpr bool LT = EXACT && (NMAX <= 16 || DEG <= 2);
  const void* lt = nullptr;
  if constexpr (LT) lt = stage_tables<DEG>(S, n, smem);
  // lanes outside the full groups (64 % G of them) alias region 0: they
  // track nothing and store nothing to LDS
  const int wv = (int)threadIdx.x >> 6;  // wave of the block
  Lds<NMAX> L{smem + LN.tab_lds + ((size_t)wv * g.gpw + (g.in_grp ? g.grp : 0)) * S.lds_group};

  // per-group path state, replicated over the group's lanes
  long long path = -1;
  int inst = 0, root = 0;
  double t = 0.0, dt = 0.0;
  int cnt = 0, nsteps = 0, asteps = 0, nnewt = 0, nref = 0;
  bool ez = false, esc = false;  // this path runs in FP64
  int status = 0;  // 0 idle, 1 tracking, 2 finishing
  uint32_t fl = 0;
  bool exhausted = !g.in_grp;
  long long static_path = ((long long)((int)blockIdx.x - LN.first_block[si]) * LN.wpb + wv) * g.gpw + g.grp;
  // lane-distributed state: component gl of x and of the last accepted x
  cx<double> x = cmk<double>(0, 0), xl = x;

  auto start_path = [&](const KSeg& Sv, bool fp64) {
    t = 0.0;
    dt = opts_arg->dt0;
    cnt = 0;
    asteps = 0;
    ez = false;
    esc = fp64;
    status = 1;
    if (row) {
      const double2 v = Sv.start[(size_t)root * Sv.start_stride + g.gl];
      x = cmk<double>(v.x, v.y);
    } else {
      x = cmk<double>(0, 0);
    }
    xl = x;
  };

  for (;;) {
    const KOpts& o = *iteration_view(opts_arg, iter);
    const KSeg& Sk = *iteration_view(seg_arg, iter);
    ++iter;
    // ------------------------------------------------ refill idle groups
    const bool want = status == 0 && !exhausted;
    if (__any(want)) {
      long long p;
      if (LN.persistent) {
        unsigned int q = 0;
        if (want && g.gl == 0) {
          q = atomicAdd(Sk.counter, 1u);
          // path-list launch (FP64 re-track of selected paths)
          if (Sk.plist) q = q < *Sk.plist_count ? (unsigned int)Sk.plist[q] : 0xffffffffu;
        }
        p = (long long)(unsigned int)lane_read((int)q, g.gbase);
        if (p == 0xffffffffll) p = -1;
      } else {
        p = static_path;
      }
      if (want) {
        if (p < 0 || p >= Sk.P) {
          exhausted = true;
        } else {
          path = p;
          inst = (int)(p / Sk.K);
          root = (int)(p - (long long)inst * Sk.K);
          nsteps = 0;
          nnewt = 0;
          nref = 0;
          fl = Sk.plist ? HCX_FLAG_ESCALATED : 0u;
          start_path(Sk, false);
          if (!LN.persistent) exhausted = true;
        }
      }
    }
    if (!__any(status != 0)) break;

    // ------------------------------------------------ loop-top tests (:116)
    bool stepping = false;
    if (status == 1) {
      if (!(t < 1.0 && 1.0 - t > o.t_tol)) {
        status = 2;
        fl |= HCX_FLAG_TRACKED;
      } else if (asteps > o.max_steps) {
        status = 2;
        fl |= HCX_FLAG_BUDGET;
      } else {
        stepping = true;
      }
    }

    if (__any(stepping)) {
      // end zone and step clamp (:121-131)
      if (stepping) {
        if (!ez && fabs(1.0 - t) <= o.ez_enter) ez = true;
        if (ez) {
          if (dt > fabs(1.0 - t)) dt = fabs(1.0 - t);
        } else if (dt > fabs(o.ez_land - t)) {
          dt = fabs(o.ez_land - t);
        }
      }
      const double ts = t;
      cx<double> acc;
      if constexpr (!KESC) {
        acc = predict<NMAX, GC, DEG, float, LT>(Sk, n, g, L, lt, stepping, row, gle, inst, ts, dt, xl, nullptr);
      } else {
        const bool p32 = stepping && !esc, p64 = stepping && esc;
        cx<double> a32 = cmk<double>(0, 0), a64 = a32;
        if (__any(p32)) a32 = predict<NMAX, GC, DEG, float, LT>(Sk, n, g, L, lt, p32, row, gle, inst, ts, dt, xl, nullptr);
        if (__any(p64)) {
          HCX_WAVE_SYNC();
          const cx<double>* sd_ext =
              Sk.M == 1 ? reinterpret_cast<const cx<double>*>(Sk.cht + (size_t)inst * (Sk.C + 1)) : nullptr;
          a64 = predict<NMAX, GC, DEG, double, LT>(Sk, n, g, L, lt, p64, row, gle, inst, ts, dt, xl, sd_ext);
        }
        acc = esc ? a64 : a32;
      }
      if (stepping) {
        const double f = dt * (1.0 / 6.0);
        x = cmk<double>(xl.re + f * acc.re, xl.im + f * acc.im);
        t = ts + dt;
      }
      // Newton corrector at the new t (:189-223)
      bool ok = false, inf = false;
      {
        const bool run32 = stepping && !ez && !esc && o.newton32, run64 = stepping && !run32;
        bool ok32 = false, inf32 = false, ok64 = false, inf64 = false;
        if (__any(run32)) newton_pass<NMAX, GC, DEG, float, LT>(Sk, n, o, g, L, lt, run32, row, x, nnewt, ok32, inf32);
        if (__any(run64)) {
          HCX_WAVE_SYNC();
          stage_coeffs<double>(Sk, g, run64, inst, t, L.coef(), false);
          newton_pass<NMAX, GC, DEG, double, LT>(Sk, n, o, g, L, lt, run64, row, x, nnewt, ok64, inf64);
        }
        ok = run64 ? ok64 : ok32;
        inf = run64 ? inf64 : inf32;
      }
      // step control (:225-249)
      if (stepping) {
        ++nsteps;
        ++asteps;
        if (inf) {
          status = 2;
          fl |= HCX_FLAG_INFINITY;
        } else if (!ok) {
          dt *= 0.5;
          x = xl;
          t = ts;
          cnt = 0;
          if (dt < o.dt_min) {
            status = 2;
            fl |= HCX_FLAG_MINSTEP;
          }
        } else {
          ++cnt;
          xl = x;
          if (cnt >= o.incr) {
            cnt = 0;
            dt *= 2.0;
          }
        }
      }
      HCX_WAVE_SYNC();
    }

    // ------------------------------------------------ finish: refine + outputs
    const bool fin = status == 2;
    if (__any(fin)) {
      // the reference's real test on the tracked root, before refinement
      // (Evaluations.cpp:96-104: converged and every |Im x_i| <= 1e-4)
      {
        double imt = (fin && row) ? fabs(x.im) : 0.0;
        if (imt != imt) imt = INFINITY;
        imt = group_max(imt, g);
        if (fin && (fl & HCX_FLAG_TRACKED) && imt <= o.real_ref_tol) fl |= HCX_FLAG_REAL_REF;
      }
      const bool ref = fin && (fl & HCX_FLAG_TRACKED) && o.refine_iters > 0;
      double resid = 0.0;
      if (__any(ref)) {
        stage_coeffs<double>(Sk, g, ref, inst, 1.0, L.coef(), false);
        cx<double>* sx = L.template x<double>();
        const cx<double>* sc = reinterpret_cast<const cx<double>*>(L.coef());
        bool run = ref, conv = false, sing_any = false;
        for (int it = 0; it < o.refine_iters; ++it) {
          if (!__any(run)) break;
          publish_x<double>(sx, x, g, n, run, row);
          HCX_WAVE_SYNC();
          cx<double> A[NMAX];
          eval_jac_row<NMAX, DEG, double, LT>(Sk, n, gle, sx, sc, A, lt);
          const cx<double> hv = eval_t_row<DEG, double, LT>(Sk, n, gle, sx, sc, lt);
          bool sing;
          const cx<double> d = group_lu_solve<NMAX, GC, double>(A, hv, g, n, row, L.template piv<double>(), sing);
          if (run && row) x = csub(x, d);
          double nd = row ? cnorm2(d) : 0.0, nx = row ? cnorm2(x) : 0.0;
          nd = group_sum(nd, g, lane);
          nx = group_sum(nx, g, lane);
          if (run) {
            ++nref;
            sing_any = sing;
            if (!(nx < o.inf_tol)) {
              run = false;
            } else if (sqrt(nd) <= o.refine_tol * fmax(1.0, sqrt(nx))) {
              conv = true;
              run = false;
            }
          }
        }
        // residual max_i |H_i(x, 1)|
        publish_x<double>(sx, x, g, n, ref, row);
        HCX_WAVE_SYNC();
        const cx<double> hv = eval_t_row<DEG, double, LT>(Sk, n, gle, sx, sc, lt);
        double r = row ? sqrt(cnorm2(hv)) : 0.0;
        if (r != r) r = INFINITY;
        resid = group_max(r, g);
        double im = row ? fabs(x.im) : 0.0;
        if (im != im) im = INFINITY;
        im = group_max(im, g);
        if (ref) {
          if (conv) fl |= HCX_FLAG_REFINED;
          if (conv && im <= o.real_tol) fl |= HCX_FLAG_REAL;
          if (sing_any) fl |= HCX_FLAG_SINGULAR;
        }
        HCX_WAVE_SYNC();
      }
      if (KESC && fin && !esc && !(fl & HCX_FLAG_REFINED)) {
        // precision escalation: the mixed attempt did not reach a refined
        // root; track the same path again from t = 0 in FP64 (the counters
        // then describe the FP64 attempt, as after the re-track pass)
        fl = HCX_FLAG_ESCALATED;
        nsteps = 0;
        nnewt = 0;
        nref = 0;
        start_path(Sk, true);
      } else if (fin) {
        const SegOut* O = Sk.out;
        double2* roots = O->roots;
        const int rstride = O->roots_stride;
        if (row) roots[(size_t)path * rstride + g.gl] = make_double2(x.re, x.im);
        if (g.gl == 0) {
          if (rstride > n) roots[(size_t)path * rstride + n] = make_double2(1.0, 0.0);
          if (O->flags) O->flags[path] = fl;
          if (O->steps) O->steps[path] = nsteps;
          if (O->newton) O->newton[path] = nnewt;
          if (O->refine) O->refine[path] = nref;
          if (O->tfin) O->tfin[path] = t;
          if (O->dtfin) O->dtfin[path] = dt;
          if (O->resid) O->resid[path] = resid;
          if (O->conv8) O->conv8[path] = (fl & HCX_FLAG_TRACKED) ? 1 : 0;
          if (O->inf8) O->inf8[path] = (fl & HCX_FLAG_INFINITY) ? 1 : 0;
          if (O->debug) O->debug[path] = (fl & HCX_FLAG_TRACKED) ? make_double2(1.0, 0.0) : make_double2(t, dt);
          if (O->inst_left) {
            const int left = atomicSub(&O->inst_left[inst], 1);
            if (left == 1 && O->inst_done) O->inst_done[inst] = wall_clock64();
          } else if (O->inst_done && Sk.plist) {
            atomicMax(&O->inst_done[inst], (unsigned long long)wall_clock64());
          }
        }
        status = 0;
        path = -1;
      }
    }
  }
}

// Right-hand side row of the FP64 micro-step loop: sum_k s * coef * prod x
// over the dHdt table (…cuh:89-120), with coef = c(t) from the group's LDS
// slots (H) or, for the lanes of groups in an RK4 stage (dot), dc/dt from
// `dtab`: LDS slots next to c(t), or (GDOT) the instance's table in global
// memory indexed by coefficient number (cinv maps slots to coefficients).
template <int DEG, bool LT, bool GDOT, class SegT>
__device__ __forceinline__ cx<double> eval_rhs_sel(const SegT& S, int n, int gle, const cx<double>* sx,
                                                   const cx<double>* sc, const cx<double>* dtab, bool dot,
                                                   const void* lt, bool tl = false, int k0 = 0, int ks = 1) {
  const cx<double>* base = dot ? dtab : sc;
  cx<double> acc = cmk<double>(0, 0);
  if constexpr (LT) {
    const CTerm<DEG>* tp = static_cast<const CTerm<DEG>*>(lt) + (size_t)n * S.hxT * n + gle;
#pragma unroll kTermUnroll<double>
    for (int k = k0; k < S.htT; k += ks) {
      const CTerm<DEG> tm = tp[(size_t)k * n];
      int c = cterm_c(tm);
      if (GDOT && dot) c = S.cinv[c];
      cx<double> v = cscale(base[c], (double)cterm_s(tm));
#pragma unroll
      for (int q = 0; q < DEG; ++q) v = cmul(v, sx[cterm_var(tm, q)]);
      acc = cadd(acc, v);
    }
  } else {
    const int* remap = (GDOT && dot) ? S.cinv : nullptr;
    if (tl)
      acc = term_row_sum<DEG, double>(static_cast<const Term*>(lt) + (size_t)n * S.hxT * n + gle, S.htT, n, sx, base,
                                      remap, k0, ks);
    else
      acc = term_row_sum<DEG, double>(S.ht + gle, S.htT, n, sx, base, remap, k0, ks);
  }
  return acc;
}

// Phases of a lane group in the FP64 micro-step loop.
enum : int {
  kPhIdle = 0,
  kPhLoop,     // at the loop top of the step loop (…4pts.cu:116)
  kPhStage0,   // RK4 stages 0..3 (…4pts.cu:139-183)
  kPhStage1,
  kPhStage2,
  kPhStage3,
  kPhNewton,   // corrector iteration at the new t (…4pts.cu:189-223)
  kPhRefine,   // refinement iteration on H(x, 1)
  kPhResid,    // residual H(x, 1) of the refined root
  kPhOut,      // outputs pending
};

// FP64 lane-group path loop in uniform micro-steps.  Every iteration of the
// wave performs ONE evaluation (Jacobian rows and a right-hand side) and ONE
// group solve for every active group, whatever each group's path is doing:
// an RK4 stage, a corrector iteration, a refinement iteration or the final
// residual.  A group that leaves a phase (a corrector that converges after
// one iteration, a path that ends) takes its next micro-step in the next
// iteration, so no group waits for the slowest group of its wave to finish a
// phase.  Each path's sequence of operations and results are those of the
// step loop (…4pts.cu:115-254) and of the refinement: only the interleaving
// of the groups of a wave changes.  c(t) is staged in LDS only when the
// group's evaluation point moves (RK4 stages 0, 1 and 3; the corrector
// reuses stage 3's c(t + dt)).
// REP: the latency instance (launches of few paths, one wave per SIMD at
// most): R >= 1 replicas per path (Seg::R) and chunked coefficient staging,
// compiled apart so that the throughput loop carries none of it.
// HOIST: stage_coeffs requests all terms of a coefficient at once (orders
// M <= 4; size-specialised instances up to n = 16 outside the multi-size
// kernel: the others would lose occupancy to the registers).
template <int NMAX, int DEG, bool EXACT, bool REP, bool HOIST = (EXACT && NMAX <= 16)>
__device__ __forceinline__ void run_segment_fp64(const Launch& LN, int si, unsigned char* smem, int lane) {
  Seg S = LN.seg[si];
  constexpr int GE = exact_lanes(NMAX);
  if (EXACT) {
    S.n = NMAX;
    S.G = GE;
  }
  constexpr int GC = (EXACT && is_pow2(GE)) ? GE : 0;
  const KOpts* const opts_arg = launch_opts();
  const KSeg* const seg_arg = launch_seg(si);
  int iter = 0;
  const Group g = make_group(S.G, lane, REP ? S.R : 1);
  const int k0 = REP ? g.s : 0, ks = REP ? g.R : 1;  // the replica's terms of each sum
  const int n = S.n;
  const bool row = g.in_grp && g.gl < n;
  const int gle = row ? g.gl : 0;
  constexpr bool LT = EXACT && (NMAX <= 16 || DEG <= 2);
  const void* lt = nullptr;
  // tl: the generic latency instance's full Term records are in LDS
  // (tab_lds then holds their bytes)
  bool tl = false;
  if constexpr (LT) {
    lt = stage_tables<DEG>(S, n, smem);
  } else if constexpr (!EXACT && REP) {
    if (LN.tab_lds > 0) {
      lt = stage_terms(S, n, smem);
      tl = true;
    }
  }
  const int wv = (int)threadIdx.x >> 6;  // wave of the block
  Lds<NMAX> L{smem + LN.tab_lds + ((size_t)wv * g.gpw + (g.in_grp ? g.grp : 0)) * S.lds_group};
  cx<double>* sx = L.template x<double>();
  const cx<double>* sc = reinterpret_cast<const cx<double>*>(L.coef());
  // dc/dt of the RK4 stages: constant (M = 1) -> read from the instance's
  // table in global memory (half the group's LDS), unless the launch stages
  // it in LDS (dot_lds: launches of few paths, latency bound); M > 1 -> LDS
  const bool gdot = S.M == 1 && !LN.dot_lds;

  long long path = -1;
  int inst = 0, root = 0;
  double t = 0.0, dt = 0.0, ts = 0.0, tst = -1.0, resid = 0.0;
  int cnt = 0, nsteps = 0, asteps = 0, nnewt = 0, nref = 0, nit = 0;
  bool ez = false, conv = false, sing_any = false;
  int ph = kPhIdle;
  uint32_t fl = 0;
  bool exhausted = !g.in_grp;
  long long static_path = ((long long)((int)blockIdx.x - LN.first_block[si]) * LN.wpb + wv) * g.gpw + g.grp;
  cx<double> x = cmk<double>(0, 0), xl = x, acc = x, kp = x;

  for (;;) {
    const KOpts& o = *iteration_view(opts_arg, iter);
    const KSeg& Sk = *iteration_view(seg_arg, iter);
    ++iter;
    // ------------------------------------------------ refill idle groups
    const bool want = ph == kPhIdle && !exhausted;
    if (__any(want)) {
      long long p;
      if (LN.persistent) {
        unsigned int q = 0;
        if (want && g.gl == 0 && g.s == 0) {
          q = atomicAdd(Sk.counter, 1u);
          if (Sk.plist) q = q < *Sk.plist_count ? (unsigned int)Sk.plist[q] : 0xffffffffu;
        }
        p = (long long)(unsigned int)lane_read((int)q, g.wbase);
        if (p == 0xffffffffll) p = -1;
      } else {
        p = static_path;
      }
      if (want) {
        if (p < 0 || p >= Sk.P) {
          exhausted = true;
        } else {
          path = p;
          inst = (int)(p / Sk.K);
          root = (int)(p - (long long)inst * Sk.K);
          nsteps = asteps = nnewt = nref = cnt = 0;
          fl = Sk.plist ? HCX_FLAG_ESCALATED : 0u;
          t = 0.0;
          dt = o.dt0;
          ez = false;
          tst = -1.0;  // the LDS coefficients belong to another instance
          resid = 0.0;
          if (row) {
            const double2 v = Sk.start[(size_t)root * Sk.start_stride + g.gl];
            x = cmk<double>(v.x, v.y);
          } else {
            x = cmk<double>(0, 0);
          }
          xl = x;
          ph = kPhLoop;
          if (!LN.persistent) exhausted = true;
        }
      }
    }
    if (!__any(ph != kPhIdle)) break;

    // ------------------------------------------------ loop top (:116-131)
    bool fin = false;
    if (ph == kPhLoop) {
      if (!(t < 1.0 && 1.0 - t > o.t_tol)) {
        fl |= HCX_FLAG_TRACKED;
        fin = true;
      } else if (asteps > o.max_steps) {
        fl |= HCX_FLAG_BUDGET;
        fin = true;
      } else {
        if (!ez && fabs(1.0 - t) <= o.ez_enter) ez = true;
        if (ez) {
          if (dt > fabs(1.0 - t)) dt = fabs(1.0 - t);
        } else if (dt > fabs(o.ez_land - t)) {
          dt = fabs(o.ez_land - t);
        }
        ts = t;
        acc = cmk<double>(0, 0);
        kp = acc;
        ph = kPhStage0;
      }
    }
    if (__any(fin)) {
      // the reference's real test on the tracked root, before refinement
      // (Evaluations.cpp:96-104: converged and every |Im x_i| <= 1e-4)
      double imt = (fin && row) ? fabs(x.im) : 0.0;
      if (imt != imt) imt = INFINITY;
      imt = group_max(imt, g);
      if (fin) {
        if ((fl & HCX_FLAG_TRACKED) && imt <= o.real_ref_tol) fl |= HCX_FLAG_REAL_REF;
        if ((fl & HCX_FLAG_TRACKED) && o.refine_iters > 0) {
          ph = kPhRefine;
          nit = 0;
          conv = false;
          sing_any = false;
        } else {
          ph = kPhOut;
        }
      }
    }

    // ------------------------------------------------ one micro-step
    const int pe = ph;
    const bool stg = pe >= kPhStage0 && pe <= kPhStage3;
    const bool act = stg || pe == kPhNewton || pe == kPhRefine || pe == kPhResid;
    if (__any(act)) {
      const int k = pe - kPhStage0;
      const double a = k == 0 ? 0.0 : (k == 3 ? dt : 0.5 * dt);
      const double te = stg ? ts + a : (pe == kPhNewton ? t : 1.0);
      const cx<double> xe = stg ? cmk<double>(xl.re + a * kp.re, xl.im + a * kp.im) : x;
      const bool restage = act && te != tst;
      if (__any(restage)) stage_coeffs<double, HOIST, REP ? 4 : 0>(Sk, g, restage, inst, te, L.coef(), !gdot);
      if (restage) tst = te;
      publish_x<double>(sx, xe, g, n, act, row);
      HCX_WAVE_SYNC();
      cx<double> A[NMAX];
      eval_jac_row<NMAX, DEG, double, LT>(Sk, n, gle, sx, sc, A, lt, tl, k0, ks);
      cx<double> r;
      if (gdot) {
        const cx<double>* dg = reinterpret_cast<const cx<double>*>(Sk.cht + (size_t)inst * (Sk.C + 1));
        r = eval_rhs_sel<DEG, LT, true>(Sk, n, gle, sx, sc, dg, stg, lt, tl, k0, ks);
      } else {
        r = eval_rhs_sel<DEG, LT, false>(Sk, n, gle, sx, sc, sc + (Sk.C + 1), stg, lt, tl, k0, ks);
      }
      if constexpr (REP) {
        // each replica summed every R-th term of the row's sums: a butterfly
        // over the replicas (all sums of a level exchanged together) leaves
        // the full sums, identical in every replica (each level adds a pair
        // of partial sums, commutatively)
        for (int m = 1; m < g.R; m <<= 1) {
          const int src = (lane + ((g.s ^ m) - g.s) * g.G) & 63;
#pragma unroll
          for (int j = 0; j < NMAX; ++j) {
            if (j < n) {
              A[j].re += __shfl(A[j].re, src);
              A[j].im += __shfl(A[j].im, src);
            }
          }
          r.re += __shfl(r.re, src);
          r.im += __shfl(r.im, src);
        }
      }
      bool sing;
      const cx<double> d =
          group_lu_solve<NMAX, GC, double>(A, stg ? cmk<double>(-r.re, -r.im) : r, g, n, row, L.template piv<double>(), sing);

      // RK4 stage: k_s = solution, accumulate k1 + 2 k2 + 2 k3 + k4
      if (stg) {
        kp = d;
        const double w = (k == 1 || k == 2) ? 2.0 : 1.0;
        acc = cmk<double>(acc.re + w * kp.re, acc.im + w * kp.im);
        if (k == 3) {
          const double f = dt * (1.0 / 6.0);
          x = cmk<double>(xl.re + f * acc.re, xl.im + f * acc.im);
          t = ts + dt;
          ph = kPhNewton;
          nit = 0;
        } else {
          ph = pe + 1;
        }
      }
      // corrector and refinement iterations: x -= d and the group norms
      const bool nr = pe == kPhNewton || pe == kPhRefine;
      if (nr && row) x = csub(x, d);
      double s1 = (nr && row) ? cnorm2(d) : 0.0, s2 = (nr && row) ? cnorm2(x) : 0.0;
      if (__any(nr)) group_sum2<NMAX, GC>(s1, s2, g, n, row, lane, L.template piv<double>());
      if (pe == kPhNewton) {
        ++nnewt;
        ++nit;
        const bool ok = s1 < o.corr_tol * s2, inf = s2 > o.inf_tol;
        if (ok || inf || nit >= o.max_corr) {
          // step control (:225-249)
          ++nsteps;
          ++asteps;
          if (inf) {
            fl |= HCX_FLAG_INFINITY;
            ph = kPhOut;
          } else if (!ok) {
            dt *= 0.5;
            x = xl;
            t = ts;
            cnt = 0;
            if (dt < o.dt_min) {
              fl |= HCX_FLAG_MINSTEP;
              ph = kPhOut;
            } else {
              ph = kPhLoop;
            }
          } else {
            ++cnt;
            xl = x;
            if (cnt >= o.incr) {
              cnt = 0;
              dt *= 2.0;
            }
            ph = kPhLoop;
          }
        }
      } else if (pe == kPhRefine) {
        ++nref;
        ++nit;
        sing_any = sing;
        bool stop = false;
        if (!(s2 < o.inf_tol)) {
          stop = true;
        } else if (sqrt(s1) <= o.refine_tol * fmax(1.0, sqrt(s2))) {
          conv = true;
          stop = true;
        }
        if (stop || nit >= o.refine_iters) ph = kPhResid;
      }
      // residual max_i |H_i(x, 1)| of the refined root (r = H at x)
      const bool rz = pe == kPhResid;
      if (__any(rz)) {
        double rr = (rz && row) ? sqrt(cnorm2(r)) : 0.0;
        if (rr != rr) rr = INFINITY;
        rr = group_max(rr, g);
        double im = (rz && row) ? fabs(x.im) : 0.0;
        if (im != im) im = INFINITY;
        im = group_max(im, g);
        if (rz) {
          resid = rr;
          if (conv) fl |= HCX_FLAG_REFINED;
          if (conv && im <= o.real_tol) fl |= HCX_FLAG_REAL;
          if (sing_any) fl |= HCX_FLAG_SINGULAR;
          ph = kPhOut;
        }
      }
      HCX_WAVE_SYNC();
    }

    // ------------------------------------------------ outputs
    const bool out = ph == kPhOut;
    if (__any(out) && out) {
      const SegOut* O = Sk.out;
      double2* roots = O->roots;
      const int rstride = O->roots_stride;
      // replica 0 stores (the other replicas hold the same values)
      if (row && g.s == 0) roots[(size_t)path * rstride + g.gl] = make_double2(x.re, x.im);
      if (g.gl == 0 && g.s == 0) {
        if (rstride > n) roots[(size_t)path * rstride + n] = make_double2(1.0, 0.0);
        if (O->flags) O->flags[path] = fl;
        if (O->steps) O->steps[path] = nsteps;
        if (O->newton) O->newton[path] = nnewt;
        if (O->refine) O->refine[path] = nref;
        if (O->tfin) O->tfin[path] = t;
        if (O->dtfin) O->dtfin[path] = dt;
        if (O->resid) O->resid[path] = resid;
        if (O->conv8) O->conv8[path] = (fl & HCX_FLAG_TRACKED) ? 1 : 0;
        if (O->inf8) O->inf8[path] = (fl & HCX_FLAG_INFINITY) ? 1 : 0;
        if (O->debug) O->debug[path] = (fl & HCX_FLAG_TRACKED) ? make_double2(1.0, 0.0) : make_double2(t, dt);
        if (O->inst_left) {
          const int left = atomicSub(&O->inst_left[inst], 1);
          if (left == 1 && O->inst_done) O->inst_done[inst] = wall_clock64();
        } else if (O->inst_done && Sk.plist) {
          atomicMax(&O->inst_done[inst], (unsigned long long)wall_clock64());
        }
      }
      ph = kPhIdle;
      path = -1;
    }
  }
}

// Register budget: the size-specialised instances up to n = 6 are compiled
// for three waves per SIMD (at most 168 VGPRs), which their live state fits
// without spills.
template <int NMAX, int DEG, int MODE, bool EXACT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((EXACT && NMAX <= 6) ? 3 : 1)))
void track_kernel(const Launch LN) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = (int)threadIdx.x & 63;
  int s = 0;
  while (s + 1 < LN.nseg && (int)blockIdx.x >= LN.first_block[s + 1]) ++s;
  for (int tried = 0; tried < LN.nseg; ++tried) {
    if constexpr (MODE == 0) {
      run_segment_fp64<NMAX, DEG, EXACT, false>(LN, s, smem, lane);
    } else {
      run_segment<NMAX, DEG, MODE, EXACT>(LN, s, smem, lane);
    }
    if (!LN.persistent) break;
    s = s + 1 == LN.nseg ? 0 : s + 1;
  }
}

// Latency instance: FP64 launches of one segment and few paths (one wave
// per SIMD at most, so the register budget is a full SIMD's), with R >= 1
// replicas per path.
template <int NMAX, int DEG, bool EXACT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1)))
void track_kernel_rep(const Launch LN) {
  extern __shared__ __align__(16) unsigned char smem[];
  run_segment_fp64<NMAX, DEG, EXACT, true>(LN, 0, smem, (int)threadIdx.x & 63);
}

// One launch over segments of different sizes (uneven batches): segment s
// runs the size-specialised loop of its n, n in {3, 5, 6, 9, 12} (terms of
// degree <= DEG, compact tables in LDS, G = n), so a multi-type launch keeps
// the specialised code of each type instead of the generic instance; the
// kernel's registers and LDS are those of its largest member.
template <int N, int DEG, int MODE>
__device__ __forceinline__ void run_exact(const Launch& LN, int si, unsigned char* smem, int lane) {
  if constexpr (MODE == 0) {
    run_segment_fp64<N, DEG, true, false, false>(LN, si, smem, lane);
  } else {
    run_segment<N, DEG, MODE, true>(LN, si, smem, lane);
  }
}

template <int DEG, int MODE>
__global__ __launch_bounds__(256) void track_kernel_set(const Launch LN) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = (int)threadIdx.x & 63;
  int s = 0;
  while (s + 1 < LN.nseg && (int)blockIdx.x >= LN.first_block[s + 1]) ++s;
  for (int tried = 0; tried < LN.nseg; ++tried) {
    switch (LN.seg[s].n) {
      case 3: run_exact<3, DEG, MODE>(LN, s, smem, lane); break;
      case 5: run_exact<5, DEG, MODE>(LN, s, smem, lane); break;
      case 6: run_exact<6, DEG, MODE>(LN, s, smem, lane); break;
      case 9: run_exact<9, DEG, MODE>(LN, s, smem, lane); break;
      case 12: run_exact<12, DEG, MODE>(LN, s, smem, lane); break;
      default: break;
    }
    if (!LN.persistent) break;
    s = s + 1 == LN.nseg ? 0 : s + 1;
  }
}

using KernelFn = void (*)(const Launch);

// Kernel instances live in separate translation units (compile in parallel).
// Generic instances: any n <= NMAX and any lanes-per-path G >= n; one unit
// per (NMAX, mode, term degree) and a dispatcher per (NMAX, mode).
KernelFn kernel_g8_m0(int deg);
KernelFn kernel_g8_m1(int deg);
KernelFn kernel_g16_m0(int deg);
KernelFn kernel_g16_m1(int deg);
KernelFn kernel_g32_m0(int deg);
KernelFn kernel_g32_m1(int deg);
#define HCX_GENERIC_UNITS(G, M) \
  KernelFn kernel_g##G##_m##M##_d2(); KernelFn kernel_g##G##_m##M##_d3(); KernelFn kernel_g##G##_m##M##_d6();
HCX_GENERIC_UNITS(8, 0)
HCX_GENERIC_UNITS(8, 1)
HCX_GENERIC_UNITS(16, 0)
HCX_GENERIC_UNITS(16, 1)
HCX_GENERIC_UNITS(32, 0)
HCX_GENERIC_UNITS(32, 1)
#undef HCX_GENERIC_UNITS
// Size-specialised instances (n = G = N) for the problem sizes of the
// benchmark workloads and of the reference's small problems (n = 4, 8, 10),
// DEG 2 and 3; DEG 4 for n = 4, 5 in FP64 mode.
// Multi-size instances (n in {3, 5, 6, 9, 12}, DEG 2) for uneven batches.
KernelFn kernel_set_m0();
KernelFn kernel_set_m1();
KernelFn kernel_x3_m0(int deg);
KernelFn kernel_x3_m1(int deg);
KernelFn kernel_x4_m0(int deg);
KernelFn kernel_x4_m1(int deg);
KernelFn kernel_x5_m0(int deg);
KernelFn kernel_x5_m1(int deg);
KernelFn kernel_x6_m0(int deg);
KernelFn kernel_x6_m1(int deg);
KernelFn kernel_x8_m0(int deg);
KernelFn kernel_x8_m1(int deg);
KernelFn kernel_x9_m0(int deg);
KernelFn kernel_x9_m1(int deg);
KernelFn kernel_x10_m0(int deg);
KernelFn kernel_x10_m1(int deg);
KernelFn kernel_x12_m0(int deg);
KernelFn kernel_x12_m1(int deg);
KernelFn kernel_x16_m0(int deg);
KernelFn kernel_x16_m1(int deg);
KernelFn kernel_x30_m0(int deg);
KernelFn kernel_x30_m1(int deg);
// Latency instances (FP64, one segment, few paths): size-specialised for
// n in {3, 4, 5, 6} and {8, 9, 10, 12, 16} (DEG 2 and 3), generic for NMAX 8,
// 16 and 32 (DEG 2, 3 and 6).
KernelFn kernel_rep_xs(int n, int deg);
KernelFn kernel_rep_xl(int n, int deg);
KernelFn kernel_rep_g8(int deg);
KernelFn kernel_rep_g16(int deg);
KernelFn kernel_rep_g32(int deg);

}  // namespace hcx
