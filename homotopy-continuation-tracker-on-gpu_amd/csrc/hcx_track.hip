// gfx950 path tracker: persistent lane-group kernel, launchers and the
// tracking half of the C-ABI.
//
// Method (per path) restates the reference kernel loop
// GPU-HC/magmaHC/gpu-kernels/kernel_HC_Solver_generalized_3views_4pts.cu:115-260:
// end-zone clamp (:121-131), "loopy" RK4 predictor (:139-183), Newton
// corrector with |dx|^2 < tol |x|^2 success and |x|^2 > 1e14 infinity tests
// (:189-225), halving / doubling step control (:230-249), outputs (:256-264).
// Table semantics follow dev-eval-indxing-generalized_3views_4pts.cuh:34-120,
// the LU follows dev-cgesv-batched-small.cuh:38-107 (cabs1 partial pivoting,
// first maximum in current row order, multiply-by-reciprocal elimination,
// division in back substitution).
//
// Work decomposition (MI355X-first): a wave64 holds floor(64/G) lane groups;
// each group tracks one path at a time, lane gl < n owns equation gl.  Groups
// refill from a device work counter as soon as their path ends, so a wave
// never idles on its slowest path.  Paths of several problem types
// ("segments") share one launch; a wave works a segment until its queue is
// empty and then steals from the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "hcx_internal.hpp"
#include "hcx_kernel.hpp"

namespace hcx {

using dev::cx;
using dev::cmk;

struct OutPack {
  SegOut out[kMaxSeg];
};

// Launch set-up on the device (graph-safe, no memset): path queues,
// per-instance remaining-path counters, output descriptors, start stamp.
__global__ void init_kernel(const Launch LN, const OutPack OP, unsigned long long* batch_start) {
  const int s = blockIdx.y;
  if (s >= LN.nseg) return;
  const Seg& S = LN.seg[s];
  const SegOut& O = OP.out[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *S.counter = 0u;
    *const_cast<SegOut*>(S.out) = O;
  }
  if (O.inst_left && i < S.I) O.inst_left[i] = S.K;
  if (batch_start && s == 0 && i == 0) *batch_start = wall_clock64();
}

// Parameter-affine P2C on the device: c_j(p) = A[j][0] + sum_q A[j][1+q] p_q
// for p0 and p1 of each instance, written as the M = 1 coefficient tables
// Hx[i][j] = {c(p0), c(p1) - c(p0)}, Ht[i][j] = {c(p1) - c(p0)}, entry C = {1, 0}
// (the role of Data_Reader::Construct_Coeffs_From_Params, Data_Reader.cpp:59-89).
__global__ void p2c_affine_kernel(const double2* __restrict__ A, int C, int k, const double2* __restrict__ p0,
                                  int p0_stride, const double2* __restrict__ p1, int I, double2* __restrict__ hx,
                                  double2* __restrict__ ht) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j > C || i >= I) return;
  double2* hxi = hx + ((size_t)i * (C + 1) + j) * 2;
  double2* hti = ht + ((size_t)i * (C + 1) + j);
  if (j == C) {
    hxi[0] = make_double2(1.0, 0.0);
    hxi[1] = make_double2(0.0, 0.0);
    hti[0] = make_double2(0.0, 0.0);
    return;
  }
  const double2* a = A + (size_t)j * (1 + k);
  const double2* q0 = p0 + (size_t)i * p0_stride;
  const double2* q1 = p1 + (size_t)i * k;
  double r0 = a[0].x, i0 = a[0].y, r1 = a[0].x, i1 = a[0].y;
  for (int q = 0; q < k; ++q) {
    const double2 c = a[1 + q], u = q0[q], v = q1[q];
    r0 += c.x * u.x - c.y * u.y;
    i0 += c.x * u.y + c.y * u.x;
    r1 += c.x * v.x - c.y * v.y;
    i1 += c.x * v.y + c.y * v.x;
  }
  hxi[0] = make_double2(r0, i0);
  hxi[1] = make_double2(r1 - r0, i1 - i0);
  hti[0] = make_double2(r1 - r0, i1 - i0);
}

// Parameter-polynomial P2C on the device: for instance i and coefficient j,
// c_j(p(t)) with p(t) = p0 + t (p1 - p0) and
// c_j(p) = sum_u scalar[j][u] prod_d p[params[j][u][d]] (p[k] = 1) as its
// t^0..t^M coefficients Hx[i][j][0..M] and those of dc_j/dt, Ht[i][j][0..M-1]
// (the role of the generated p2c-<p>.h, Data_Reader.cpp:59-89, with the
// coefficient-polynomial layout of createUnivPoly4Coeffs.m; entry C = {1, 0..}).
// Each term is a product of linear factors (q_a + t d_a), expanded by
// repeated polynomial multiplication.
constexpr int kMaxOrderT = 8;

__global__ void p2c_poly_kernel(const double2* __restrict__ scal, const int* __restrict__ prm, int C, int T, int D,
                                int k, int M, const double2* __restrict__ p0, int p0_stride,
                                const double2* __restrict__ p1, int I, double2* __restrict__ hx,
                                double2* __restrict__ ht) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j > C || i >= I) return;
  double2* hxi = hx + ((size_t)i * (C + 1) + j) * (M + 1);
  double2* hti = ht + ((size_t)i * (C + 1) + j) * M;
  double cr[kMaxOrderT + 1], ci[kMaxOrderT + 1];
#pragma unroll
  for (int m = 0; m <= kMaxOrderT; ++m) cr[m] = ci[m] = 0.0;
  if (j == C) {
    cr[0] = 1.0;
  } else {
    const double2* q0 = p0 + (size_t)i * p0_stride;
    const double2* q1 = p1 + (size_t)i * k;
    for (int u = 0; u < T; ++u) {
      const double2 sc = scal[(size_t)j * T + u];
      if (sc.x == 0.0 && sc.y == 0.0) continue;
      double tr[kMaxOrderT + 1], ti[kMaxOrderT + 1];
#pragma unroll
      for (int m = 0; m <= kMaxOrderT; ++m) tr[m] = ti[m] = 0.0;
      tr[0] = sc.x;
      ti[0] = sc.y;
      int deg = 0;
      for (int d = 0; d < D; ++d) {
        const int a = prm[((size_t)j * T + u) * D + d];
        if (a >= k) continue;  // the constant-1 padding factor
        const double2 qa = q0[a], pa = q1[a];
        const double dr = pa.x - qa.x, di = pa.y - qa.y;
        // (tr, ti) *= (qa + t d): new[m] = old[m] qa + old[m-1] d
#pragma unroll
        for (int m = kMaxOrderT; m >= 0; --m) {
          if (m > deg + 1) continue;
          double nr = tr[m] * qa.x - ti[m] * qa.y, ni = tr[m] * qa.y + ti[m] * qa.x;
          if (m > 0) {
            nr += tr[m - 1] * dr - ti[m - 1] * di;
            ni += tr[m - 1] * di + ti[m - 1] * dr;
          }
          tr[m] = nr;
          ti[m] = ni;
        }
        ++deg;
      }
#pragma unroll
      for (int m = 0; m <= kMaxOrderT; ++m) {
        cr[m] += tr[m];
        ci[m] += ti[m];
      }
    }
  }
#pragma unroll
  for (int m = 0; m <= kMaxOrderT; ++m) {
    if (m <= M) hxi[m] = make_double2(cr[m], ci[m]);
    if (m < M) hti[m] = make_double2((m + 1) * cr[m + 1], (m + 1) * ci[m + 1]);
  }
}

__device__ __forceinline__ bool same_root(const double2* a, const double2* b, int n, double tol) {
  for (int v = 0; v < n; ++v)
    if (!(fabs(a[v].x - b[v].x) < tol && fabs(a[v].y - b[v].y) < tol)) return false;
  return true;
}

// Solution collection on the device, one block per instance
// (Evaluations.cpp:88-176).  Counts [tracked, infinity, refined, real,
// unique, unique_ref, real_ref]:
//  * unique: first occurrence of a refined root among the refined roots;
//  * real_ref: tracked paths carrying HCX_FLAG_REAL_REF, the reference's real
//    test on the tracked root (Evaluate_GPUHC_Sols, :88-110);
//  * unique_ref: Find_Unique_Sols (:127-176).  Its Skip set is always the
//    Duplicate set of the last examined path that had one (bstar), so the
//    sequential pass needs, per path, only "has a later path (any outcome)
//    with the same root" and a comparison with bstar.
// The pair comparisons run in parallel over the block (thread per first
// index, roots staged in LDS when they fit) into LDS bit sets; one thread
// then runs the short sequential pass.  Dynamic LDS: 4 bit sets of
// ceil(K / 32) words, then K * n roots when stage_roots is set.
__global__ void collect_kernel(int n, int K, const double2* __restrict__ roots, uint32_t* __restrict__ flags,
                               double tol, unsigned long long* __restrict__ counts, int stage_roots) {
  extern __shared__ __align__(16) unsigned int bits[];
  const int words = (K + 31) / 32;
  unsigned int* later = bits;                // a later path (any outcome) has the same root
  unsigned int* tracked = bits + words;      // tracked ("converged")
  unsigned int* dupref = bits + 2 * words;   // refined with an earlier refined duplicate
  unsigned int* uref = bits + 3 * words;     // counted by the unique_ref pass
  double2* sroots = reinterpret_cast<double2*>(bits + ((4 * words + 3) & ~3));
  const int inst = blockIdx.x;
  const size_t base = (size_t)inst * K;
  for (int w = threadIdx.x; w < 4 * words; w += blockDim.x) bits[w] = 0u;
  const double2* x = roots + base * n;
  if (stage_roots) {
    for (int q = threadIdx.x; q < K * n; q += blockDim.x) sroots[q] = x[q];
    x = sroots;
  }
  __syncthreads();
  unsigned long long c_tr = 0, c_inf = 0, c_ref = 0, c_real = 0, c_rref = 0;
  for (int i = threadIdx.x; i < K; i += blockDim.x) {
    const uint32_t fi = flags[base + i];
    if (fi & HCX_FLAG_TRACKED) {
      ++c_tr;
      atomicOr(&tracked[i >> 5], 1u << (i & 31));
      if (fi & HCX_FLAG_REAL_REF) ++c_rref;
    }
    c_inf += (fi & HCX_FLAG_INFINITY) ? 1 : 0;
    const bool ri = (fi & HCX_FLAG_REFINED) != 0;
    if (ri) {
      ++c_ref;
      c_real += (fi & HCX_FLAG_REAL) ? 1 : 0;
    }
    bool lat = false;
    for (int j = i + 1; j < K; ++j) {
      if (!same_root(x + (size_t)i * n, x + (size_t)j * n, n, tol)) continue;
      lat = true;
      if (ri && (flags[base + j] & HCX_FLAG_REFINED)) atomicOr(&dupref[j >> 5], 1u << (j & 31));
    }
    if (lat) atomicOr(&later[i >> 5], 1u << (i & 31));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long c_uref = 0;
    int bstar = -1;
    for (int bs = 0; bs < K; ++bs) {
      if (!((tracked[bs >> 5] >> (bs & 31)) & 1u)) continue;
      if (bstar >= 0 && same_root(x + (size_t)bs * n, x + (size_t)bstar * n, n, tol)) continue;
      if ((later[bs >> 5] >> (bs & 31)) & 1u) {
        bstar = bs;
      } else {
        uref[bs >> 5] |= 1u << (bs & 31);
        ++c_uref;
      }
    }
    atomicAdd(&counts[5], c_uref);
  }
  __syncthreads();
  unsigned long long c_uni = 0;
  for (int j = threadIdx.x; j < K; j += blockDim.x) {
    uint32_t f = flags[base + j] & ~(HCX_FLAG_UNIQUE | HCX_FLAG_UNIQUE_REF);
    if ((f & HCX_FLAG_REFINED) && !((dupref[j >> 5] >> (j & 31)) & 1u)) {
      f |= HCX_FLAG_UNIQUE;
      ++c_uni;
    }
    if ((uref[j >> 5] >> (j & 31)) & 1u) f |= HCX_FLAG_UNIQUE_REF;
    flags[base + j] = f;
  }
  atomicAdd(&counts[0], c_tr);
  atomicAdd(&counts[1], c_inf);
  atomicAdd(&counts[2], c_ref);
  atomicAdd(&counts[3], c_real);
  atomicAdd(&counts[4], c_uni);
  atomicAdd(&counts[6], c_rref);
}

// Lists the paths of a mixed-precision launch that the FP64 pass re-tracks
// from t = 0: every path without a refined root that the launch did not
// already re-track in FP64 (precision escalation) and, when mark_dups is set,
// every refined path whose root equals another refined root of the same
// instance (|dRe|, |dIm| < tol in every coordinate): a path that jumped onto
// another path's root shares that root.  The listed paths are appended to
// plist.  One block per instance.
__global__ void dup_mark_kernel(int n, int K, const double2* __restrict__ roots, const uint32_t* __restrict__ flags,
                                double tol, int mark_dups, int* __restrict__ plist,
                                unsigned int* __restrict__ count) {
  const int inst = blockIdx.x;
  const size_t base = (size_t)inst * K;
  for (int j = threadIdx.x; j < K; j += blockDim.x) {
    const uint32_t fj = flags[base + j];
    if (!(fj & HCX_FLAG_REFINED)) {
      if (!(fj & HCX_FLAG_ESCALATED)) {
        const unsigned int q = atomicAdd(count, 1u);
        plist[q] = (int)(base + j);
      }
      continue;
    }
    if (!mark_dups) continue;
    const double2* xj = roots + (base + j) * n;
    bool dup = false;
    for (int i = 0; i < K && !dup; ++i) {
      if (i == j || !(flags[base + i] & HCX_FLAG_REFINED)) continue;
      const double2* xi = roots + (base + i) * n;
      bool same = true;
      for (int v = 0; v < n && same; ++v) same = fabs(xi[v].x - xj[v].x) < tol && fabs(xi[v].y - xj[v].y) < tol;
      dup = same;
    }
    if (dup) {
      const unsigned int q = atomicAdd(count, 1u);
      plist[q] = (int)(base + j);
    }
  }
}

__global__ void zero_u32_kernel(unsigned int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// ------------------------------------------------------------ dispatch table
static const int kNmaxList[] = {8, 16, 32};
static const int kDegList[] = {2, 3, 6};

static KernelFn pick_kernel(int nmax, int deg, int mode) {
  switch (nmax) {
    case 8: return mode == 0 ? kernel_g8_m0(deg) : kernel_g8_m1(deg);
    case 16: return mode == 0 ? kernel_g16_m0(deg) : kernel_g16_m1(deg);
    case 32: return mode == 0 ? kernel_g32_m0(deg) : kernel_g32_m1(deg);
    default: return nullptr;
  }
}

// The replica instance of a launch planned on `pl` (nullptr: none).
static KernelFn pick_rep(int nmax, int deg, bool exact) {
  if (exact) {
    if (deg > 4) return nullptr;
    deg = deg < 2 ? 2 : deg;
    if (nmax >= 3 && nmax <= 6) return kernel_rep_xs(nmax, deg);
    if (nmax == 8 || nmax == 9 || nmax == 10 || nmax == 12 || nmax == 16) return kernel_rep_xl(nmax, deg);
    return nullptr;
  }
  switch (nmax) {
    case 8: return kernel_rep_g8(deg);
    case 16: return kernel_rep_g16(deg);
    case 32: return kernel_rep_g32(deg);
    default: return nullptr;
  }
}

static KernelFn pick_exact(int n, int deg, int mode) {
  // degree 4 (8-byte compact terms of four variables): n = 4, 5 in FP64 mode
  if (deg > 4 || (deg == 4 && mode != 0)) return nullptr;
  deg = deg < 2 ? 2 : deg;
  switch (n) {
    case 3: return mode == 0 ? kernel_x3_m0(deg) : kernel_x3_m1(deg);
    case 4: return mode == 0 ? kernel_x4_m0(deg) : kernel_x4_m1(deg);
    case 5: return mode == 0 ? kernel_x5_m0(deg) : kernel_x5_m1(deg);
    case 6: return mode == 0 ? kernel_x6_m0(deg) : kernel_x6_m1(deg);
    case 8: return mode == 0 ? kernel_x8_m0(deg) : kernel_x8_m1(deg);
    case 9: return mode == 0 ? kernel_x9_m0(deg) : kernel_x9_m1(deg);
    case 10: return mode == 0 ? kernel_x10_m0(deg) : kernel_x10_m1(deg);
    case 12: return mode == 0 ? kernel_x12_m0(deg) : kernel_x12_m1(deg);
    case 16: return mode == 0 ? kernel_x16_m0(deg) : kernel_x16_m1(deg);
    case 30: return mode == 0 ? kernel_x30_m0(deg) : kernel_x30_m1(deg);
    default: return nullptr;
  }
}

}  // namespace hcx

#include "hcx_problem.hpp"


namespace hcx {
namespace {

std::mutex g_mu;
unsigned long long* g_counts = nullptr;   // device [kNumCounts], hcx_collect's totals
int g_cu = 0;
constexpr int kNumCounts = 7;
static_assert(kSegSlots == kMaxSeg, "workspace blocks hold one counter slot per segment");

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return HCX_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

// Compute-unit count of the device (sizes persistent grids); allocates nothing.
hcx_status ensure_device_info() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_cu > 0) return HCX_OK;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  g_cu = prop.multiProcessorCount;
  return HCX_OK;
}

// Device-memory policy of the workspace pool.
struct HipDev {
  void* alloc(size_t n) {
    void* p = nullptr;
    return hipMalloc(&p, n) == hipSuccess ? p : nullptr;
  }
  void release(void* p) { (void)hipFree(p); }
  bool sync(const void* s) { return hipStreamSynchronize((hipStream_t)s) == hipSuccess; }
  bool capturing(const void* s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing((hipStream_t)s, &st) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return st == hipStreamCaptureStatusActive;
  }
};

// Per-stream launch workspaces; lives until the process ends (blocks that
// captured graphs refer to must outlive every graph).
WorkspacePool<HipDev>& pool() {
  static auto* p = new WorkspacePool<HipDev>(HipDev{}, sizeof(SegOut));
  return *p;
}

hcx_status ws_error(WsStatus w, const char* who) {
  switch (w) {
    case WsStatus::kOk: return HCX_OK;
    case WsStatus::kCaptureTooSmall:
      set_error("%s: the stream's launch workspace is too small and cannot grow during graph capture; "
                "run the same launch once on this stream before capturing it", who);
      return HCX_ERR_INVALID_ARG;
    case WsStatus::kOutOfMemory:
      set_error("%s: out of device memory for the launch workspace", who);
      return HCX_ERR_OUT_OF_MEMORY;
    default:
      set_error("%s: waiting for the stream before growing its workspace failed", who);
      return HCX_ERR_HIP;
  }
}

int round_up_idx(const int* list, int cnt, int v) {
  for (int i = 0; i < cnt; ++i)
    if (list[i] >= v) return i;
  return -1;
}

double f_term(int d) { return 6.0 * d + 4.0; }

// LAPACK Working Note 41 counts for one complex GETRF + GETRS with 1 RHS.
double f_lu(int n) { return (8.0 / 3.0) * n * n * n + 7.0 * n * n + (7.0 / 3.0) * n; }

hcx_status validate_opts(const hcx_track_opts* o) {
  if (!o) {
    set_error("options are NULL");
    return HCX_ERR_INVALID_ARG;
  }
  if (o->max_steps < 0 || o->max_corr < 1 || o->incr_steps < 1 || !(o->dt0 > 0) || !(o->end_zone >= 0) ||
      !(o->t_tol > 0) || o->refine_iters < 0 || (o->precision != HCX_PREC_FP64 && o->precision != HCX_PREC_MIXED) ||
      o->lanes_per_path < 0 || o->lanes_per_path > 64 || o->specialize < 0 || o->specialize > 1 ||
      o->waves_per_cu < -64 || o->waves_per_cu > 64 || o->replicas < 0 || o->replicas > 64 ||
      (o->replicas & (o->replicas - 1)) != 0) {
    set_error("invalid tracker options (max_steps=%d max_corr=%d incr=%d precision=%d lanes=%d)", o->max_steps,
              o->max_corr, o->incr_steps, o->precision, o->lanes_per_path);
    return HCX_ERR_INVALID_ARG;
  }
  return HCX_OK;
}

Opts to_opts(const hcx_track_opts* o) {
  Opts r;
  r.ez_enter = o->end_zone + 1e-7;
  r.ez_land = 1.0 - o->end_zone;
  r.max_steps = o->max_steps;
  r.max_corr = o->max_corr;
  r.incr = o->incr_steps;
  r.refine_iters = o->refine_iters;
  r.dt0 = o->dt0;
  r.t_tol = o->t_tol;
  r.dt_min = o->dt_min;
  r.corr_tol = o->corr_tol;
  r.inf_tol = o->inf_tol;
  r.refine_tol = o->refine_tol;
  r.real_tol = o->real_tol;
  r.real_ref_tol = o->real_ref_tol;
  r.newton32 = o->fp32_corrector ? 1 : 0;
  return r;
}

// Per-group LDS: the x / pivot-row slots (shared, FP64 capacity) and the
// coefficient area: with M = 1 (dc/dt constant, read by the FP64 stages from
// the instance's table in global memory) FP32 c + dc/dt overlaid by FP64 c,
// else (or dot_lds: FP64 stages with dc/dt in LDS) FP64 c + dc/dt.
// The region stride is padded to 16 mod 256 bytes: the groups of a wave read
// the same offsets of their regions together (pivot-row broadcasts, x and
// coefficient gathers), and this stride spreads them over distinct LDS banks
// (64 banks x 4 bytes) for 8- and 16-byte accesses.
int lds_group_bytes(int nmax, int C, int M, bool dot_lds = false) {
  const int per = (M > 1 || dot_lds) ? 32 : 16;
  int b = 16 * (nmax + 1) + (C + 1) * per;
  b = (b + 15) & ~15;
  return ((b - 16 + 255) / 256) * 256 + 16;
}

struct Plan {
  int nmax = 0, mode = 0;
  int deg = 2;       // term degree of the instance (generic: rounded up to its instance's)
  bool exact = false;
  bool tab4 = false; // the kernel reads 4-byte compact terms (DEG <= 2), else 8-byte
  int tab_lds = 0;   // LDS bytes of the compact tables (instances that hold them there)
  size_t lds = 0;
  KernelFn fn = nullptr;
};

// Largest compact table the size-specialised instances (up to n = 16; n = 30
// with quadratic terms) stage in
// LDS; a problem with larger tables runs on the generic instance.
constexpr int kMaxTabLds = 32 * 1024;
// LDS per workgroup on gfx950
constexpr int kMaxLdsBytes = 160 * 1024;

// Kernel instance for a launch: the size-specialised one when every segment
// has the same n, runs its lanes per path (G = exact_lanes(n)) and fits
// DEG <= 3 (and, for
// n <= 16, has compact tables of at most kMaxTabLds bytes); the multi-size
// instance when the segments' sizes differ but all are in {3, 5, 6, 9, 12}
// with G = n, quadratic terms and compact tables; otherwise the generic
// instance with the smallest NMAX >= max n.
hcx_status choose_kernel(int nseg, const hcx_problem* const* probs, const int* Gs, int max_deg, int mode,
                         int specialize, Plan& pl) {
  pl.mode = mode;
  pl.tab_lds = 0;
  int max_n = 1, tab = 0;
  bool same = true, tab_ok = true, in_set = true;
  // the compact term format of the kernel's degree: 4 bytes up to degree 2
  const bool t4 = max_deg <= 2;
  pl.tab4 = t4;
  for (int s = 0; s < nseg; ++s) {
    const int n = probs[s]->n;
    max_n = std::max(max_n, n);
    same = same && n == probs[0]->n && Gs[s] == exact_lanes(n);
    in_set = in_set && (n == 3 || n == 5 || n == 6 || n == 9 || n == 12) && Gs[s] == n;
    const void* ct = t4 ? (const void*)probs[s]->d_ctab4 : (const void*)probs[s]->d_ctab;
    const int cb = t4 ? probs[s]->ctab4_bytes : probs[s]->ctab_bytes;
    tab_ok = tab_ok && ct && cb <= kMaxTabLds;
    tab = std::max(tab, cb);
  }
  if (specialize && !same && in_set && tab_ok && max_deg <= 2) {
    pl.fn = mode == 0 ? kernel_set_m0() : kernel_set_m1();
    pl.nmax = max_n;
    pl.deg = 2;
    pl.exact = true;
    pl.tab_lds = tab;
    return HCX_OK;
  }
  const int n0 = probs[0]->n;
  // n = 30 holds the tables in LDS only with quadratic terms (4-byte records)
  if (specialize && same && (tab_ok || (n0 > 16 && !t4))) {
    KernelFn fn = pick_exact(n0, max_deg, mode);
    if (fn) {
      pl.fn = fn;
      pl.nmax = n0;
      pl.deg = max_deg;
      pl.exact = true;
      pl.tab_lds = (n0 <= 16 || t4) ? tab : 0;
      return HCX_OK;
    }
  }
  const int ni = round_up_idx(kNmaxList, 3, max_n);
  const int di = round_up_idx(kDegList, 3, max_deg);
  if (ni < 0 || di < 0) {
    set_error("unsupported sizes: n=%d (max %d), degree=%d (max 6)", max_n, kMaxN, max_deg);
    return HCX_ERR_UNSUPPORTED;
  }
  pl.fn = pick_kernel(kNmaxList[ni], kDegList[di], mode);
  pl.nmax = kNmaxList[ni];
  pl.deg = kDegList[di];
  pl.exact = false;
  if (!pl.fn) {
    set_error("no kernel instance for n=%d degree=%d mode=%d", max_n, max_deg, mode);
    return HCX_ERR_UNSUPPORTED;
  }
  return HCX_OK;
}

// Fills the segment-independent part of a Seg from a problem.
void seg_from_problem(Seg& S, SegOut& O, const hcx_problem* p, int G, int mode, int nmax, bool tab4) {
  std::memset(&S, 0, sizeof(S));
  std::memset(&O, 0, sizeof(O));
  O.roots_stride = p->n;
  S.n = p->n;
  S.C = p->C;
  S.M = p->M;
  S.K = p->K;
  S.hxT = p->hxT;
  S.htT = p->htT;
  S.G = G;
  S.R = 1;
  S.lds_group = lds_group_bytes(nmax, p->C, p->M);
  S.hx = p->d_hx;
  S.ht = p->d_ht;
  S.hx_cnt = p->d_hx_cnt;
  S.cslot = p->d_cslot;
  S.cinv = p->d_cinv;
  S.ctab = tab4 ? p->d_ctab4 : p->d_ctab;
  S.start = p->d_start;
  S.start_stride = p->n;
}

// Dynamic LDS bytes of a launch: the tables (when staged) and one region per
// full group of every wave of the block; the lanes of a partial group share
// their wave's region 0 and never write LDS.
size_t plan_lds(const Launch& LN, const Plan& pl) {
  size_t lds = 0;
  for (int s = 0; s < LN.nseg; ++s) {
    const Seg& S = LN.seg[s];
    const int groups = 64 / (S.G * S.R);
    lds = std::max(lds, (size_t)std::max(LN.wpb, 1) * (size_t)groups * (size_t)S.lds_group);
  }
  return lds + (size_t)pl.tab_lds;
}

hcx_status finish_plan(Launch& LN, Plan& pl) {
  if (LN.wpb < 1) LN.wpb = 1;
  const size_t lds = plan_lds(LN, pl);
  LN.tab_lds = pl.tab_lds;
  pl.lds = lds;
  if (lds > (size_t)kMaxLdsBytes) {
    set_error("launch needs %zu bytes of LDS per block (at most %d)", lds, kMaxLdsBytes);
    return HCX_ERR_UNSUPPORTED;
  }
  if (lds > 65536) {
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(pl.fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  }
  return HCX_OK;
}

// Waves per block of a persistent one-segment launch whose instance holds
// the compact tables in LDS: the waves of a block share one copy of the
// tables, so blocks of several waves can leave room for more waves per CU.
// The candidate (1, 2 or 4 waves) with the most resident waves per CU wins,
// the smaller block on ties.
hcx_status choose_waves_per_block(Launch& LN, Plan& pl, int& best_wpb, int& best_nb) {
  best_wpb = 1;
  best_nb = 0;
  int best_waves = -1;
  for (int wpb : {1, 2, 4}) {
    LN.wpb = wpb;
    Plan q = pl;
    if (wpb > 1 && plan_lds(LN, q) > (size_t)kMaxLdsBytes) continue;
    hcx_status st = finish_plan(LN, q);
    if (st != HCX_OK) return st;
    int nb = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(pl.fn), 64 * wpb,
                                                         q.lds));
    if (nb * wpb > best_waves) {
      best_waves = nb * wpb;
      best_wpb = wpb;
      best_nb = nb;
    }
  }
  LN.wpb = best_wpb;
  return finish_plan(LN, pl);
}

// Default lanes per path: those of the size-specialised instance (n, or 32
// for n = 30; several paths per wave, G need not divide 64), widened while
// the wave's groups would need more than 64 KiB of LDS (wide coefficient
// tables), so that at least two waves share a CU.
int default_lanes(int n, int C, int M) {
  const int ni = round_up_idx(kNmaxList, 3, n);
  const int lg = lds_group_bytes(ni < 0 ? n : kNmaxList[ni], C, M);
  int G = exact_lanes(n);
  while (G < 64 && (size_t)(64 / G) * (size_t)lg > 65536) ++G;
  return G;
}

// Workspace entries a launch needs: instance counters when instance
// completion times are requested, the re-track list in mixed precision.
void ws_need(int nseg, hcx_problem* const* probs, const hcx_batch_desc* batches, const hcx_result_desc* results,
             const hcx_track_opts* o, long long& inst, long long& paths) {
  inst = paths = 0;
  for (int s = 0; s < nseg; ++s) {
    if (!probs[s] || batches[s].num_instances < 0) continue;
    const long long I = batches[s].num_instances;
    if (results[s].d_inst_done) inst += I;
    if (o && o->precision == HCX_PREC_MIXED) paths += I * probs[s]->K;
  }
}

// Carves the per-segment views of a block for the given launch.
void ws_carve(const WsBlock& b, int nseg, const hcx_batch_desc* batches, const hcx_result_desc* results,
              hcx_problem* const* probs, const hcx_track_opts* o, LaunchWork* work) {
  long long io = 0, po = 0;
  for (int s = 0; s < nseg; ++s) {
    work[s] = seg_work(b, s, io, po);
    const long long I = std::max(batches[s].num_instances, 0);
    if (results[s].d_inst_done) io += I;
    if (o->precision == HCX_PREC_MIXED && probs[s]) po += I * probs[s]->K;
  }
}

}  // namespace

LaunchWork seg_work(const WsBlock& b, int s, long long inst_off, long long path_off) {
  const WsLayout L = ws_layout(sizeof(SegOut), b.inst_cap, b.paths_cap);
  char* base = static_cast<char*>(b.base);
  LaunchWork w;
  w.counter = reinterpret_cast<unsigned int*>(base) + 64 * s;
  w.plist_count = w.counter + 32;
  w.out = reinterpret_cast<SegOut*>(base + L.out_off) + s;
  w.inst_left = reinterpret_cast<int*>(base + L.inst_off) + inst_off;
  w.plist = reinterpret_cast<int*>(base + L.plist_off) + path_off;
  w.inst_cap = b.inst_cap - inst_off;
  w.plist_cap = b.paths_cap - path_off;
  return w;
}

// ------------------------------------------------------------ static scratch
// Launch workspace of hcx_track_noalloc: a file-scope device array (each GPU
// holds its own copy until the process ends).  Host code never uses the
// array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
constexpr size_t kScratchBytes =
    ws_layout(sizeof(SegOut), HCX_NOALLOC_MAX_INSTANCES, HCX_NOALLOC_MAX_PATHS).bytes;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch as a workspace block.
hcx_status scratch_block(WsBlock& b) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("hcx_track_noalloc: device %d beyond the %d devices the scratch table covers", dev, kMaxDevices);
    return HCX_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hcx_track_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return HCX_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  b.base = g_scratch_ptr[dev];
  b.inst_cap = HCX_NOALLOC_MAX_INSTANCES;
  b.paths_cap = HCX_NOALLOC_MAX_PATHS;
  b.pinned = true;
  return HCX_OK;
}
}  // namespace

// A launch on `stream` with that stream's pooled workspace.
hcx_status track_pooled(int nseg, hcx_problem* const* probs, const hcx_batch_desc* batches,
                        const hcx_result_desc* results, const hcx_track_opts* o, hipStream_t stream, int persistent,
                        unsigned long long* batch_start, const char* who) {
  if (nseg < 1 || nseg > kMaxSeg || !probs || !batches || !results || !o) {
    set_error("%s: need 1..%d segments with problems, batches, results and options", who, kMaxSeg);
    return HCX_ERR_INVALID_ARG;
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  long long inst = 0, paths = 0;
  ws_need(nseg, probs, batches, results, o, inst, paths);
  WorkspacePool<HipDev>::Lease lease;
  hcx_status st = ws_error(pool().acquire(dev, stream, inst, paths, lease), who);
  if (st != HCX_OK) return st;
  LaunchWork work[kMaxSeg];
  ws_carve(lease.block(), nseg, batches, results, probs, o, work);
  return track_impl(nseg, probs, batches, results, o, stream, persistent, batch_start, work);
}

}  // namespace hcx

using namespace hcx;

extern "C" {

int hcx_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return c;
}

hcx_status hcx_problem_create(const hcx_problem_desc* d, hcx_problem** out) {
  clear_error();
  if (!d || !out) {
    set_error("hcx_problem_create: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  const int n = d->num_vars, C = d->num_coeffs, M = d->max_order_t;
  if (n < 1 || n > kMaxN || C < 1 || M < 1 || d->hx_terms < 1 || d->ht_terms < 1 || d->hx_parts < 3 ||
      d->ht_parts < 3 || d->hx_parts > 10 || d->ht_parts > 10 || !d->h_hx_idx || !d->h_ht_idx ||
      d->num_tracks < 1 || !d->h_start_sols || d->num_params < 0 ||
      (d->num_params > 0 && (!d->h_start_params || (!d->h_coeff_affine) == (!d->h_cmap_scalar)))) {
    set_error("hcx_problem_create: invalid descriptor (n=%d C=%d M=%d hx %dx%d ht %dx%d K=%d k=%d)", n, C, M,
              d->hx_terms, d->hx_parts, d->ht_terms, d->ht_parts, d->num_tracks, d->num_params);
    return HCX_ERR_INVALID_ARG;
  }
  if (d->h_cmap_scalar) {
    if (!d->h_cmap_params || d->cmap_terms < 1 || d->cmap_deg < 1 || d->cmap_deg > M || M > kMaxOrderT) {
      set_error("hcx_problem_create: invalid coefficient map (terms %d, degree %d, M %d; need 1 <= degree <= M <= %d)",
                d->cmap_terms, d->cmap_deg, M, kMaxOrderT);
      return HCX_ERR_INVALID_ARG;
    }
    const size_t np = (size_t)C * d->cmap_terms * d->cmap_deg;
    for (size_t q = 0; q < np; ++q)
      if (d->h_cmap_params[q] < 0 || d->h_cmap_params[q] > d->num_params) {
        set_error("hcx_problem_create: coefficient map parameter index %d out of range 0..%d", d->h_cmap_params[q],
                  d->num_params);
        return HCX_ERR_INVALID_ARG;
      }
  }
  if (hcx_device_count() < 1) {
    set_error("hcx_problem_create: no HIP device available");
    return HCX_ERR_NO_DEVICE;
  }
  hcx_status st = ensure_device_info();
  if (st != HCX_OK) return st;

  const int hxv = d->hx_parts - 2, htv = d->ht_parts - 2;
  // repack + validate tables: hx [var j][term][eq i], ht [term][eq i]
  std::vector<dev::Term> hx((size_t)n * d->hx_terms * n), ht((size_t)d->ht_terms * n);
  std::vector<int> hx_cnt(n, 0);
  auto pack = [&](const int* rec, int nv, dev::Term& tm) -> bool {
    tm.s = rec[0];
    tm.c = rec[1];
    if (tm.c < 0 || tm.c > C) return false;
    uint32_t v03 = 0, v47 = 0;
    for (int q = 0; q < 8; ++q) {
      int v = q < nv ? rec[2 + q] : n;
      if (v < 0 || v > n) return false;
      if (q < 4) v03 |= (uint32_t)v << (8 * q);
      else v47 |= (uint32_t)v << (8 * (q - 4));
    }
    tm.v03 = v03;
    tm.v47 = v47;
    return true;
  };
  int max_deg = 1;
  for (int i = 0; i < n; ++i) {
    for (int tt = 0; tt < d->ht_terms; ++tt) {
      const int* rec = d->h_ht_idx + ((size_t)i * d->ht_terms + tt) * d->ht_parts;
      if (!pack(rec, htv, ht[(size_t)tt * n + i])) {
        set_error("hcx_problem_create: dHdt entry (eq %d, term %d) out of range", i, tt);
        return HCX_ERR_INVALID_ARG;
      }
      if (rec[0] != 0) {
        int dg = 0;
        for (int q = 0; q < htv; ++q) dg += rec[2 + q] != n;
        max_deg = std::max(max_deg, dg);
      }
    }
    for (int j = 0; j < n; ++j) {
      int real = 0;
      for (int tt = 0; tt < d->hx_terms; ++tt) {
        const int* rec = d->h_hx_idx + (((size_t)i * n + j) * d->hx_terms + tt) * d->hx_parts;
        if (!pack(rec, hxv, hx[((size_t)j * d->hx_terms + tt) * n + i])) {
          set_error("hcx_problem_create: dHdx entry (eq %d, var %d, term %d) out of range", i, j, tt);
          return HCX_ERR_INVALID_ARG;
        }
        if (rec[0] != 0) {
          real = tt + 1;
          int dg = 1;
          for (int q = 0; q < hxv; ++q) dg += rec[2 + q] != n;
          max_deg = std::max(max_deg, dg);
        }
      }
      hx_cnt[j] = std::max(hx_cnt[j], real);
    }
  }
  // a padded H term still multiplies ht_vars factors; the kernel reads DEG slots
  const int deg = std::max(std::max(max_deg, htv), hxv + 1);
  if (deg > 6) {
    set_error("hcx_problem_create: terms with %d variable factors are not supported (max 6)", deg);
    return HCX_ERR_UNSUPPORTED;
  }

  // LDS slots of the coefficients in the order the lanes of a group read them:
  // dH/dt terms term by term across the equations, then dH/dx terms, then the rest
  std::vector<int> cslot(C + 1, -1), cinv(C + 1, -1);
  int next_slot = 0;
  auto take = [&](dev::Term& tm) {
    if (tm.s == 0) return;
    if (cslot[tm.c] < 0) {
      cslot[tm.c] = next_slot;
      cinv[next_slot] = tm.c;
      ++next_slot;
    }
  };
  for (int tt = 0; tt < d->ht_terms; ++tt)
    for (int i = 0; i < n; ++i) take(ht[(size_t)tt * n + i]);
  for (int j = 0; j < n; ++j)
    for (int tt = 0; tt < d->hx_terms; ++tt)
      for (int i = 0; i < n; ++i) take(hx[((size_t)j * d->hx_terms + tt) * n + i]);
  for (int c = 0; c <= C; ++c)
    if (cslot[c] < 0) {
      cslot[c] = next_slot;
      cinv[next_slot] = c;
      ++next_slot;
    }
  for (auto& tm : ht) tm.c = cslot[tm.c];
  for (auto& tm : hx) tm.c = cslot[tm.c];

  // compact 8-byte terms (hx part, then ht part) for the instances that hold
  // the tables in LDS: slot and scalar in 16 bits each, <= 4 variables
  std::vector<uint2> ctab;
  {
    bool fits = C <= 0xffff && std::max(deg, 2) <= 4;
    auto pack = [&](const dev::Term& tm) {
      if (tm.s < -32768 || tm.s > 32767) fits = false;
      return make_uint2(((uint32_t)tm.c & 0xffffu) | ((uint32_t)tm.s << 16), tm.v03);
    };
    ctab.reserve(hx.size() + ht.size());
    for (const auto& tm : hx) ctab.push_back(pack(tm));
    for (const auto& tm : ht) ctab.push_back(pack(tm));
    if (!fits) ctab.clear();
  }
  // compact 4-byte terms for degree <= 2: slot in 12 bits, scalar in 4 bits
  // signed, two variables of 8 bits
  std::vector<uint32_t> ctab4;
  if (std::max(deg, 2) <= 2 && C <= 0xfff) {
    bool fits = true;
    auto pack4 = [&](const dev::Term& tm) {
      if (tm.s < -8 || tm.s > 7) fits = false;
      return ((uint32_t)tm.c & 0xfffu) | (((uint32_t)tm.s & 0xfu) << 12) | ((tm.v03 & 0xffffu) << 16);
    };
    ctab4.reserve(hx.size() + ht.size());
    for (const auto& tm : hx) ctab4.push_back(pack4(tm));
    for (const auto& tm : ht) ctab4.push_back(pack4(tm));
    if (!fits) ctab4.clear();
  }

  auto* p = new hcx_problem();
  (void)hipGetDevice(&p->device);
  p->n = n;
  p->C = C;
  p->M = M;
  p->K = d->num_tracks;
  p->k = d->num_params;
  p->hxT = d->hx_terms;
  p->htT = d->ht_terms;
  p->hx_vars = hxv;
  p->ht_vars = htv;
  p->deg = std::max(deg, 2);
  auto fail = [&](hipError_t e, const char* what) {
    set_error("hcx_problem_create: %s: %s", what, hipGetErrorString(e));
    hcx_problem_destroy(p);
    return HCX_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&p->d_hx, hx.size() * sizeof(dev::Term))) != hipSuccess) return fail(e, "hipMalloc hx");
  if ((e = hipMalloc(&p->d_ht, ht.size() * sizeof(dev::Term))) != hipSuccess) return fail(e, "hipMalloc ht");
  if ((e = hipMalloc(&p->d_hx_cnt, n * sizeof(int))) != hipSuccess) return fail(e, "hipMalloc hx_cnt");
  if ((e = hipMalloc(&p->d_cslot, (C + 1) * sizeof(int))) != hipSuccess) return fail(e, "hipMalloc cslot");
  if ((e = hipMalloc(&p->d_cinv, (C + 1) * sizeof(int))) != hipSuccess) return fail(e, "hipMalloc cinv");
  if ((e = hipMemcpy(p->d_cslot, cslot.data(), (C + 1) * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "copy cslot");
  if ((e = hipMemcpy(p->d_cinv, cinv.data(), (C + 1) * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "copy cinv");
  if (!ctab4.empty()) {
    p->ctab4_bytes = (int)((ctab4.size() * sizeof(uint32_t) + 15) / 16 * 16);
    if ((e = hipMalloc(&p->d_ctab4, p->ctab4_bytes)) != hipSuccess) return fail(e, "hipMalloc ctab4");
    if ((e = hipMemcpy(p->d_ctab4, ctab4.data(), ctab4.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return fail(e, "copy ctab4");
  }
  if (!ctab.empty()) {
    p->ctab_bytes = (int)((ctab.size() * sizeof(uint2) + 15) / 16 * 16);
    if ((e = hipMalloc(&p->d_ctab, p->ctab_bytes)) != hipSuccess) return fail(e, "hipMalloc ctab");
    if ((e = hipMemcpy(p->d_ctab, ctab.data(), ctab.size() * sizeof(uint2), hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e, "copy ctab");
  }
  if ((e = hipMalloc(&p->d_start, (size_t)p->K * n * sizeof(double2))) != hipSuccess) return fail(e, "hipMalloc start");
  if ((e = hipMalloc(&p->launcher_ws.base, ws_layout(sizeof(SegOut), 0, 0).bytes)) != hipSuccess)
    return fail(e, "hipMalloc launcher workspace");
  if ((e = hipMemcpy(p->d_hx, hx.data(), hx.size() * sizeof(dev::Term), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "copy hx");
  if ((e = hipMemcpy(p->d_ht, ht.data(), ht.size() * sizeof(dev::Term), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "copy ht");
  if ((e = hipMemcpy(p->d_hx_cnt, hx_cnt.data(), n * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "copy hx_cnt");
  if ((e = hipMemcpy(p->d_start, d->h_start_sols, (size_t)p->K * n * sizeof(double2), hipMemcpyHostToDevice)) !=
      hipSuccess)
    return fail(e, "copy start");
  if (p->k > 0 && d->h_coeff_affine) {
    if ((e = hipMalloc(&p->d_affine, (size_t)C * (1 + p->k) * sizeof(double2))) != hipSuccess)
      return fail(e, "hipMalloc affine");
    if ((e = hipMemcpy(p->d_affine, d->h_coeff_affine, (size_t)C * (1 + p->k) * sizeof(double2),
                       hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e, "copy affine");
  }
  if (p->k > 0 && d->h_cmap_scalar) {
    p->cmap_T = d->cmap_terms;
    p->cmap_D = d->cmap_deg;
    const size_t ns = (size_t)C * p->cmap_T, np = ns * p->cmap_D;
    if ((e = hipMalloc(&p->d_cmap_scalar, ns * sizeof(double2))) != hipSuccess) return fail(e, "hipMalloc cmap");
    if ((e = hipMalloc(&p->d_cmap_params, np * sizeof(int))) != hipSuccess) return fail(e, "hipMalloc cmap");
    if ((e = hipMemcpy(p->d_cmap_scalar, d->h_cmap_scalar, ns * sizeof(double2), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return fail(e, "copy cmap");
    if ((e = hipMemcpy(p->d_cmap_params, d->h_cmap_params, np * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e, "copy cmap");
  }
  if (p->k > 0) {
    if ((e = hipMalloc(&p->d_p0, (size_t)p->k * sizeof(double2))) != hipSuccess) return fail(e, "hipMalloc p0");
    if ((e = hipMemcpy(p->d_p0, d->h_start_params, (size_t)p->k * sizeof(double2), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return fail(e, "copy p0");
  }

  // algorithmic FLOPs of the reference algorithm (padding terms excluded)
  hcx_problem_info& in = p->info;
  in.num_vars = n;
  in.num_coeffs = C;
  in.max_order_t = M;
  in.num_tracks = p->K;
  in.num_params = p->k;
  in.hx_terms = p->hxT;
  in.ht_terms = p->htT;
  in.hx_vars = hxv;
  in.ht_vars = htv;
  double fh = 0, fhx = 0;
  for (int i = 0; i < n; ++i) {
    for (int tt = 0; tt < d->ht_terms; ++tt) {
      const int* rec = d->h_ht_idx + ((size_t)i * d->ht_terms + tt) * d->ht_parts;
      if (rec[0] == 0) continue;
      int dg = 0;
      for (int q = 0; q < htv; ++q) dg += rec[2 + q] != n;
      fh += f_term(dg);
    }
    for (int j = 0; j < n; ++j)
      for (int tt = 0; tt < d->hx_terms; ++tt) {
        const int* rec = d->h_hx_idx + (((size_t)i * n + j) * d->hx_terms + tt) * d->hx_parts;
        if (rec[0] == 0) continue;
        int dg = 0;
        for (int q = 0; q < hxv; ++q) dg += rec[2 + q] != n;
        fhx += f_term(dg);
      }
  }
  in.flops_coef = (double)C * (8.0 * M - 4.0);
  in.flops_hx = fhx;
  in.flops_ht = fh;
  in.flops_h = fh;
  in.flops_lu = f_lu(n);
  in.flops_stage = in.flops_coef + fhx + fh + in.flops_lu + 16.0 * n;
  in.flops_newton = fhx + fh + in.flops_lu + 12.0 * n;
  *out = p;
  return HCX_OK;
}

hcx_status hcx_problem_destroy(hcx_problem* p) {
  if (!p) return HCX_OK;
  if (p->d_hx) (void)hipFree(p->d_hx);
  if (p->d_ht) (void)hipFree(p->d_ht);
  if (p->d_hx_cnt) (void)hipFree(p->d_hx_cnt);
  if (p->d_cslot) (void)hipFree(p->d_cslot);
  if (p->d_cinv) (void)hipFree(p->d_cinv);
  if (p->d_ctab) (void)hipFree(p->d_ctab);
  if (p->d_ctab4) (void)hipFree(p->d_ctab4);
  if (p->d_start) (void)hipFree(p->d_start);
  if (p->d_affine) (void)hipFree(p->d_affine);
  if (p->d_p0) (void)hipFree(p->d_p0);
  if (p->d_cmap_scalar) (void)hipFree(p->d_cmap_scalar);
  if (p->d_cmap_params) (void)hipFree(p->d_cmap_params);
  if (p->launcher_ws.base) (void)hipFree(p->launcher_ws.base);
  delete p;
  return HCX_OK;
}

hcx_status hcx_problem_get_info(const hcx_problem* p, hcx_problem_info* info) {
  clear_error();
  if (!p || !info) {
    set_error("hcx_problem_get_info: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  *info = p->info;
  return HCX_OK;
}

hcx_status hcx_sync(void* stream) {
  clear_error();
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return HCX_OK;
}

hcx_status hcx_stream_release(void* stream) {
  clear_error();
  if (hcx_device_count() < 1) {
    set_error("hcx_stream_release: no HIP device available");
    return HCX_ERR_NO_DEVICE;
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  return ws_error(pool().release_stream(dev, stream), "hcx_stream_release");
}

// Shared implementation of hcx_track / hcx_track_multi.
}  // extern "C"

namespace hcx {
// Shared implementation of hcx_track / hcx_track_multi / graphs / the server.
// `work` (one per segment) is the launch workspace: the stream's pooled
// block, a graph's own block, a server slot's block or the static scratch.
hcx_status track_impl(int nseg, hcx_problem* const* probs, const hcx_batch_desc* batches,
                      const hcx_result_desc* results, const hcx_track_opts* o, hipStream_t stream, int persistent,
                      unsigned long long* batch_start, const LaunchWork* work) {
  hcx_status st = validate_opts(o);
  if (st != HCX_OK) return st;
  if (nseg < 1 || nseg > kMaxSeg || !probs || !batches || !results || !work) {
    set_error("track: need 1..%d segments with problems, batches, results and a workspace", kMaxSeg);
    return HCX_ERR_INVALID_ARG;
  }
  st = ensure_device_info();
  if (st != HCX_OK) return st;
  Launch LN;
  std::memset(&LN, 0, sizeof(LN));
  OutPack OP;
  std::memset(&OP, 0, sizeof(OP));
  LN.nseg = nseg;
  LN.persistent = persistent;
  LN.o = to_opts(o);
  int max_n = 1, max_deg = 2;
  for (int s = 0; s < nseg; ++s) {
    hcx_problem* p = probs[s];
    if (!p) {
      set_error("track: segment %d has no problem", s);
      return HCX_ERR_INVALID_ARG;
    }
    for (int q = 0; q < s; ++q)
      if (probs[q] == p) {
        set_error("track: problem used by two segments of one launch");
        return HCX_ERR_INVALID_ARG;
      }
    max_n = std::max(max_n, p->n);
    max_deg = std::max(max_deg, p->deg);
  }
  int ns[kMaxSeg], Gs[kMaxSeg];
  int specialize = o->specialize;
  bool all_tuned = o->lanes_per_path == 0;
  for (int s = 0; s < nseg; ++s) {
    all_tuned = all_tuned && probs[s]->tuned;
    ns[s] = probs[s]->n;
    if (o->lanes_per_path > 0) Gs[s] = o->lanes_per_path;
    else if (probs[s]->tuned) Gs[s] = probs[s]->tuned_G;
    else Gs[s] = default_lanes(probs[s]->n, probs[s]->C, probs[s]->M);
    if (Gs[s] < ns[s]) {
      set_error("track: lanes_per_path=%d is below n=%d", Gs[s], ns[s]);
      return HCX_ERR_INVALID_ARG;
    }
  }
  if (all_tuned && nseg == 1) specialize = probs[0]->tuned_spec;
  Plan pl;
  st = choose_kernel(nseg, probs, Gs, max_deg, o->precision, specialize, pl);
  if (st != HCX_OK) return st;
  const int nmax = pl.nmax;
  long long total_groups_needed = 0;
  std::vector<long long> seg_groups(nseg);
  for (int s = 0; s < nseg; ++s) {
    hcx_problem* p = probs[s];
    const hcx_batch_desc& b = batches[s];
    const hcx_result_desc& r = results[s];
    // an empty batch (no instances) needs no buffers and tracks nothing
    const bool empty = b.num_instances == 0 || p->K == 0;
    if (b.num_instances < 0 || (!empty && (!r.d_roots || !r.d_flags))) {
      set_error("track: segment %d: bad batch or result descriptor", s);
      return HCX_ERR_INVALID_ARG;
    }
    const int G = Gs[s];
    Seg& S = LN.seg[s];
    SegOut& O = OP.out[s];
    seg_from_problem(S, O, p, G, o->precision, nmax, pl.tab4);
    S.I = b.num_instances;
    S.P = (long long)b.num_instances * p->K;
    if (empty) {
      S.chx = nullptr;
      S.cht = nullptr;
    } else if (b.d_coeffs_hx) {
      if (!b.d_coeffs_ht) {
        set_error("track: segment %d: coefficient tables need both Hx and Ht", s);
        return HCX_ERR_INVALID_ARG;
      }
      S.chx = reinterpret_cast<const double2*>(b.d_coeffs_hx);
      S.cht = reinterpret_cast<const double2*>(b.d_coeffs_ht);
    } else {
      if (!b.d_target_params || !b.d_coeff_work || p->k < 1 || (!p->d_cmap_scalar && (!p->d_affine || p->M != 1))) {
        set_error("track: segment %d: needs coefficient tables or target params + a coefficient map + workspace", s);
        return HCX_ERR_INVALID_ARG;
      }
      double2* hx = reinterpret_cast<double2*>(b.d_coeff_work);
      double2* ht = hx + (size_t)b.num_instances * (p->C + 1) * (p->M + 1);
      if (b.num_instances > 0) {
        const double2* p0 = b.d_start_params ? reinterpret_cast<const double2*>(b.d_start_params) : p->d_p0;
        const int p0_stride = b.d_start_params ? p->k : 0;
        dim3 grid((p->C + 1 + 127) / 128, b.num_instances);
        if (p->d_cmap_scalar)
          hipLaunchKernelGGL(p2c_poly_kernel, grid, dim3(128), 0, stream, p->d_cmap_scalar, p->d_cmap_params, p->C,
                             p->cmap_T, p->cmap_D, p->k, p->M, p0, p0_stride,
                             reinterpret_cast<const double2*>(b.d_target_params), b.num_instances, hx, ht);
        else
          hipLaunchKernelGGL(p2c_affine_kernel, grid, dim3(128), 0, stream,
                             reinterpret_cast<const double2*>(p->d_affine), p->C, p->k, p0, p0_stride,
                             reinterpret_cast<const double2*>(b.d_target_params), b.num_instances, hx, ht);
      }
      S.chx = hx;
      S.cht = ht;
    }
    O.roots = reinterpret_cast<double2*>(r.d_roots);
    O.flags = r.d_flags;
    O.steps = r.d_steps;
    O.newton = r.d_newton;
    O.refine = r.d_refine;
    O.tfin = r.d_t_final;
    O.dtfin = r.d_dt_final;
    O.resid = r.d_residual;
    O.inst_done = reinterpret_cast<unsigned long long*>(r.d_inst_done);
    S.counter = work[s].counter;
    S.out = work[s].out;
    O.inst_left = nullptr;
    if (r.d_inst_done && b.num_instances > 0) {
      if (!work[s].inst_left || work[s].inst_cap < b.num_instances) {
        set_error("track: launch workspace instance counters too small (%lld < %d)", work[s].inst_cap,
                  b.num_instances);
        return HCX_ERR_INVALID_ARG;
      }
      O.inst_left = work[s].inst_left;
    }
    const int gpw = 64 / G;
    seg_groups[s] = (S.P + gpw - 1) / gpw;
    total_groups_needed += seg_groups[s];
  }
  // A one-segment FP64 launch of at most one wave per SIMD is latency bound:
  // no other waves hide the latency of its evaluations.  It runs the latency
  // instance of its size when one exists; dc/dt goes to LDS next to c(t),
  // the generic instances read their Term records from an LDS copy, and
  // each path gets R replicas of its lane group, which split the
  // evaluation's term sums (the most, a power of two up to 8 with R * G <=
  // 64, that keep the launch within two waves per SIMD, or half of one above
  // R = 2 for light rows; or o->replicas).  All of it only when the group
  // regions and the records fit in 64 KiB.
  if (o->precision == HCX_PREC_FP64 && nseg == 1 && (total_groups_needed <= 4LL * g_cu || o->replicas > 1)) {
    Seg& S = LN.seg[0];
    const hcx_problem* p = probs[0];
    const int lg = lds_group_bytes(nmax, p->C, p->M, true);
    const size_t terms =
        pl.exact ? 0 : ((size_t)p->n * p->n * p->hxT + (size_t)p->n * p->htT) * sizeof(dev::Term);
    auto waves = [&](int R) {
      const long long gpw = 64 / (S.G * R);
      return (S.P + gpw - 1) / gpw;
    };
    int R = 1;
    if (o->replicas > 0) {
      R = o->replicas;
      if (R * S.G > 64) {
        set_error("track: %d replicas of %d lanes exceed a wave", R, S.G);
        return HCX_ERR_INVALID_ARG;
      }
    } else {
      // at most 8 replicas; beyond 2 and half a wave per SIMD only for rows
      // whose evaluation outweighs the butterfly (complex products per row
      // >= 8 per exchanged sum).  Measured: 583 paths of 46 quadratic terms
      // per row ran slower at R = 4 than at R = 2, 668 paths of 84 cubic
      // terms faster, 4,096 paths of a light n = 6 type faster at R = 2 (820
      // waves) than at R = 1, and 2,560 paths at n = 16 faster at R = 2
      // (1,280 waves) than at R = 1
      const int dg = std::max(p->deg, 1);
      const double row_products = (double)p->n * p->hxT * (dg - 1) + (double)p->htT * dg;
      const bool heavy = row_products >= 8.0 * (p->n + 1);
      while (R < 8 && 2 * R * S.G <= 64 && waves(2 * R) <= 8LL * g_cu &&
             (R == 1 || heavy || waves(2 * R) <= 2LL * g_cu))
        R *= 2;
    }
    // the latency instance (its register budget is a whole SIMD's), also
    // for R = 1
    const KernelFn rf = pick_rep(pl.nmax, pl.deg, pl.exact);
    if (!rf) R = 1;
    const size_t regions = (size_t)(64 / (S.G * R)) * (size_t)lg;
    if (regions + (size_t)pl.tab_lds + terms <= 65536) {
      LN.dot_lds = 1;
      S.lds_group = lg;
      S.R = R;
      if (rf) pl.fn = rf;
      // the generic latency instances read their term records from LDS
      if (!pl.exact && rf) pl.tab_lds = (int)terms;
      seg_groups[0] = waves(R);
      total_groups_needed = seg_groups[0];
    }
  }
  LN.wpb = 1;
  int occ_blocks = -1;  // blocks per CU at full occupancy, when already known
  if (persistent && nseg == 1 && pl.tab_lds > 0) {
    int wpb = 1;
    st = choose_waves_per_block(LN, pl, wpb, occ_blocks);
  } else {
    st = finish_plan(LN, pl);
  }
  if (st != HCX_OK) return st;
  const int wpb = LN.wpb;

  // grid: persistent -> occupancy-limited, split over segments by work
  long long blocks;
  if (persistent) {
    int per_cu = o->waves_per_cu;  // waves per CU
    if (per_cu == 0 && all_tuned && nseg == 1) per_cu = probs[0]->tuned_wpc;
    if (per_cu <= 0) {
      // occupancy, less -waves_per_cu waves when that is negative (room for
      // the work of other solves in flight on other streams)
      int nb = occ_blocks;
      if (nb < 0)
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(pl.fn), 64 * wpb,
                                                             pl.lds));
      per_cu = std::max(nb * wpb + std::min(per_cu, 0), 1);
    }
    // blocks of wpb waves: waves per CU rounded up to whole blocks
    const long long waves_needed = std::max<long long>(total_groups_needed, 1);
    blocks = std::min<long long>((long long)((per_cu + wpb - 1) / wpb) * g_cu, (waves_needed + wpb - 1) / wpb);
    // proportional split by estimated work (paths * n^2 as a cost proxy)
    std::vector<double> w(nseg);
    double wsum = 0;
    for (int s = 0; s < nseg; ++s) {
      w[s] = (double)LN.seg[s].P * (double)(probs[s]->n * probs[s]->n + probs[s]->C);
      wsum += w[s];
    }
    long long acc = 0;
    for (int s = 0; s < nseg; ++s) {
      LN.first_block[s] = (int)acc;
      long long nb = wsum > 0 ? (long long)std::llround(blocks * w[s] / wsum) : 0;
      nb = std::min(nb, seg_groups[s]);
      acc += nb;
    }
    LN.first_block[nseg] = (int)acc;
    if (acc < blocks) blocks = std::max<long long>(acc, 1);
    if (acc == 0) blocks = 1;
  } else {
    long long acc = 0;
    for (int s = 0; s < nseg; ++s) {
      LN.first_block[s] = (int)acc;
      acc += seg_groups[s];
    }
    LN.first_block[nseg] = (int)acc;
    blocks = std::max<long long>(acc, 1);
  }
  int max_inst = 1;
  for (int s = 0; s < nseg; ++s) max_inst = std::max(max_inst, LN.seg[s].I);
  dim3 ig((max_inst + 255) / 256, nseg);
  hipLaunchKernelGGL(init_kernel, ig, dim3(256), 0, stream, LN, OP, batch_start);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(pl.fn, dim3((unsigned)blocks), dim3(64 * wpb), pl.lds, stream, LN);
  HIP_TRY(hipGetLastError());

  // Mixed precision: re-track in FP64 from t = 0 every path that ended
  // without a refined root (precision escalation) and every path whose root
  // duplicates another root of its instance (a jump onto another path), so
  // the refined roots and counts are those of the FP64 tracker.
  if (o->precision == HCX_PREC_MIXED) {
    Launch L2 = LN;
    L2.persistent = 1;
    L2.wpb = 1;
    OutPack OP2 = OP;
    for (int s = 0; s < nseg; ++s) {
      hcx_problem* p = probs[s];
      Seg& S = L2.seg[s];
      (void)p;
      if (work[s].plist_cap < S.P || !work[s].plist || !work[s].plist_count) {
        set_error("track: launch workspace path list too small (%lld < %lld)", work[s].plist_cap, S.P);
        return HCX_ERR_INVALID_ARG;
      }
      int* plist = work[s].plist;
      unsigned int* pcount = work[s].plist_count;
      hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, stream, pcount, 1);
      if (S.I > 0)
        hipLaunchKernelGGL(dup_mark_kernel, dim3(S.I), dim3(256), 0, stream, S.n, S.K,
                           reinterpret_cast<const double2*>(results[s].d_roots), results[s].d_flags, o->unique_tol,
                           o->retrack_duplicates, plist, pcount);
      S.plist = plist;
      S.plist_count = pcount;
      OP2.out[s].inst_left = nullptr;
    }
    HIP_TRY(hipGetLastError());
    Plan p2;
    st = choose_kernel(nseg, probs, Gs, max_deg, HCX_PREC_FP64, specialize, p2);
    if (st != HCX_OK) return st;
    // few paths: latency bound, so dc/dt is staged in LDS when the groups'
    // regions stay within 64 KiB
    size_t need = 0;
    for (int s = 0; s < nseg; ++s)
      need = std::max(need, (size_t)(64 / L2.seg[s].G) *
                                (size_t)lds_group_bytes(p2.nmax, probs[s]->C, probs[s]->M, true));
    L2.dot_lds = need + (size_t)p2.tab_lds <= 65536 ? 1 : 0;
    for (int s = 0; s < nseg; ++s)
      L2.seg[s].lds_group = lds_group_bytes(p2.nmax, probs[s]->C, probs[s]->M, L2.dot_lds != 0);
    st = finish_plan(L2, p2);
    if (st != HCX_OK) return st;
    // a small persistent grid: the list holds few paths; idle waves exit at once
    long long b2 = 0;
    for (int s = 0; s < nseg; ++s) {
      L2.first_block[s] = (int)b2;
      const int gpw = 64 / L2.seg[s].G;
      b2 += std::max<long long>(1, std::min<long long>((L2.seg[s].P + gpw - 1) / gpw, g_cu / nseg + 1));
    }
    L2.first_block[nseg] = (int)b2;
    hipLaunchKernelGGL(init_kernel, dim3(1, nseg), dim3(64), 0, stream, L2, OP2, (unsigned long long*)nullptr);
    hipLaunchKernelGGL(p2.fn, dim3((unsigned)b2), dim3(64), p2.lds, stream, L2);
    HIP_TRY(hipGetLastError());
  }
  return HCX_OK;
}

}  // namespace hcx

extern "C" {

hcx_status hcx_track(hcx_problem* p, const hcx_batch_desc* b, const hcx_track_opts* o, const hcx_result_desc* r,
                     void* stream) {
  clear_error();
  if (!p || !b || !r) {
    set_error("hcx_track: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  return track_pooled(1, &p, b, r, o, (hipStream_t)stream, 1, reinterpret_cast<unsigned long long*>(r->d_batch_start),
                      "hcx_track");
}

hcx_status hcx_track_multi(int num_segments, hcx_problem* const* problems, const hcx_batch_desc* batches,
                           const hcx_result_desc* results, const hcx_track_opts* o, void* stream) {
  clear_error();
  if (num_segments < 1 || !results) {
    set_error("hcx_track_multi: no segments");
    return HCX_ERR_INVALID_ARG;
  }
  return track_pooled(num_segments, problems, batches, results, o, (hipStream_t)stream, 1,
                      reinterpret_cast<unsigned long long*>(results[0].d_batch_start), "hcx_track_multi");
}

double hcx_kernel_HC_Solver(void* stream, const hcx_problem* pc, int max_steps, int max_corr, int incr_steps,
                            const hcx_complex* d_start_sols, hcx_complex* d_track, const hcx_complex* d_coeffs_hx,
                            const hcx_complex* d_coeffs_ht, uint8_t* d_converge, uint8_t* d_infinity,
                            hcx_complex* d_debug) {
  clear_error();
  hcx_problem* p = const_cast<hcx_problem*>(pc);
  if (!p || !d_start_sols || !d_track || !d_coeffs_hx || !d_coeffs_ht || !d_converge || !d_infinity) {
    set_error("hcx_kernel_HC_Solver: null argument");
    return -1.0;
  }
  hcx_track_opts o;
  hcx_track_opts_default(&o);
  o.max_steps = max_steps;
  o.max_corr = max_corr;
  o.incr_steps = incr_steps;
  o.dt_min = 0.0;
  o.refine_iters = 0;
  if (validate_opts(&o) != HCX_OK) return -1.0;
  if (ensure_device_info() != HCX_OK) return -1.0;
  hipStream_t s = (hipStream_t)stream;
  // the problem's launcher block (counters, output descriptor) serves one
  // synchronous call at a time
  std::lock_guard<std::mutex> lk(p->launcher_mu);
  const LaunchWork w = seg_work(p->launcher_ws, 0, 0, 0);
  Launch LN;
  std::memset(&LN, 0, sizeof(LN));
  OutPack OP;
  std::memset(&OP, 0, sizeof(OP));
  LN.nseg = 1;
  LN.persistent = 0;
  LN.o = to_opts(&o);
  const int G = default_lanes(p->n, p->C, p->M);
  Plan pl;
  if (choose_kernel(1, &p, &G, p->deg, HCX_PREC_FP64, 1, pl) != HCX_OK) return -1.0;
  Seg& S = LN.seg[0];
  SegOut& O = OP.out[0];
  seg_from_problem(S, O, p, G, HCX_PREC_FP64, pl.nmax, pl.tab4);
  S.start = reinterpret_cast<const double2*>(d_start_sols);
  S.start_stride = p->n + 1;
  O.roots = reinterpret_cast<double2*>(d_track);
  O.roots_stride = p->n + 1;
  S.I = 1;
  S.P = p->K;
  S.chx = reinterpret_cast<const double2*>(d_coeffs_hx);
  S.cht = reinterpret_cast<const double2*>(d_coeffs_ht);
  S.counter = w.counter;
  S.out = w.out;
  // the reference's outputs only: converged / infinity bytes and (t, dt)
  O.flags = nullptr;
  O.conv8 = d_converge;
  O.inf8 = d_infinity;
  O.debug = reinterpret_cast<double2*>(d_debug);
  const int gpw = 64 / G;
  LN.first_block[0] = 0;
  LN.first_block[1] = (p->K + gpw - 1) / gpw;
  if (finish_plan(LN, pl) != HCX_OK) return -1.0;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    set_error("hcx_kernel_HC_Solver: hipEventCreate failed");
    return -1.0;
  }
  hipLaunchKernelGGL(init_kernel, dim3(1, 1), dim3(64), 0, s, LN, OP, (unsigned long long*)nullptr);
  (void)hipEventRecord(e0, s);
  hipLaunchKernelGGL(pl.fn, dim3(LN.first_block[1]), dim3(64), pl.lds, s, LN);
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(e1, s);
  hipError_t se = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (le != hipSuccess || se != hipSuccess) {
    set_error("hcx_kernel_HC_Solver: launch failed: %s", hipGetErrorString(le != hipSuccess ? le : se));
    return -1.0;
  }
  return ms * 1e-3;
}

// hcx_track with the static scratch as its launch workspace: nothing is
// allocated (no hipMalloc, no pool, no allocating library call).
hcx_status hcx_track_noalloc(hcx_problem* p, const hcx_batch_desc* b, const hcx_track_opts* o,
                             const hcx_result_desc* r, void* stream) {
  clear_error();
  if (!p || !b || !r || !o) {
    set_error("hcx_track_noalloc: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  long long inst = 0, paths = 0;
  ws_need(1, &p, b, r, o, inst, paths);
  if (inst > HCX_NOALLOC_MAX_INSTANCES || paths > HCX_NOALLOC_MAX_PATHS) {
    set_error("hcx_track_noalloc: the batch needs %lld instance counters and %lld re-track entries; the static "
              "scratch holds %d and %d", inst, paths, HCX_NOALLOC_MAX_INSTANCES, HCX_NOALLOC_MAX_PATHS);
    return HCX_ERR_INVALID_ARG;
  }
  WsBlock blk;
  hcx_status st = scratch_block(blk);
  if (st != HCX_OK) return st;
  LaunchWork work[1];
  ws_carve(blk, 1, b, r, &p, o, work);
  return track_impl(1, &p, b, r, o, (hipStream_t)stream, 1, reinterpret_cast<unsigned long long*>(r->d_batch_start),
                    work);
}

// Reference-style per-instance launch with the generic result layout (no
// persistence, no refill): used as the plain per-problem baseline.
hcx_status hcx_track_static(hcx_problem* p, const hcx_batch_desc* b, const hcx_track_opts* o,
                            const hcx_result_desc* r, void* stream) {
  clear_error();
  if (!p || !b || !r) {
    set_error("hcx_track_static: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  return track_pooled(1, &p, b, r, o, (hipStream_t)stream, 0, reinterpret_cast<unsigned long long*>(r->d_batch_start),
                      "hcx_track_static");
}

hcx_status hcx_collect(int n, int num_instances, int num_tracks, const hcx_complex* d_roots, uint32_t* d_flags,
                       double unique_tol, hcx_counts* out, void* stream) {
  clear_error();
  const bool empty = num_instances == 0 || num_tracks == 0;
  if (n < 1 || num_instances < 0 || num_tracks < 0 || (!empty && (!d_roots || !d_flags)) || !out) {
    set_error("hcx_collect: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  // bit sets, then the instance's roots when they fit in 48 KiB of LDS
  const size_t bit_lds = 16 * (size_t)((4 * ((num_tracks + 31) / 32) + 3) / 4);
  const size_t root_lds = sizeof(double2) * (size_t)num_tracks * n;
  if (bit_lds > 65536) {
    set_error("hcx_collect: %d paths per instance exceed the collection kernel's bit sets", num_tracks);
    return HCX_ERR_UNSUPPORTED;
  }
  const int stage = bit_lds + root_lds <= 48 * 1024 ? 1 : 0;
  const size_t lds = bit_lds + (stage ? root_lds : 0);
  hcx_status st = ensure_device_info();
  if (st != HCX_OK) return st;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_counts) HIP_TRY(hipMalloc(&g_counts, kNumCounts * sizeof(unsigned long long)));
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, g_counts, kNumCounts);
  if (!empty)
    hipLaunchKernelGGL(collect_kernel, dim3(num_instances), dim3(256), lds, s, n, num_tracks,
                       reinterpret_cast<const double2*>(d_roots), d_flags, unique_tol, g_counts, stage);
  HIP_TRY(hipGetLastError());
  unsigned long long h[kNumCounts];
  HIP_TRY(hipMemcpyAsync(h, g_counts, sizeof(h), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  out->tracked = (int64_t)h[0];
  out->infinity = (int64_t)h[1];
  out->refined = (int64_t)h[2];
  out->real = (int64_t)h[3];
  out->unique = (int64_t)h[4];
  out->unique_ref = (int64_t)h[5];
  out->real_ref = (int64_t)h[6];
  return HCX_OK;
}

}  // extern "C"

// ============================================================== autotuner + graphs
extern "C" {

// Times candidate launch shapes of problem p on a sample batch and keeps the
// fastest: lanes per path G in {n, next power of two <= 64}, the
// size-specialised or the generic kernel instance, and the persistent grid
// density (occupancy, half occupancy).  Results are bitwise-independent of
// the grid density; G and the kernel variant change only the summation order
// of group reductions.
hcx_status hcx_autotune(hcx_problem* p, const hcx_batch_desc* sample, const hcx_track_opts* o,
                        hcx_tune_result* out, void* stream) {
  clear_error();
  if (!p || !sample || !o || sample->num_instances < 1) {
    set_error("hcx_autotune: need a problem, options and a non-empty sample batch");
    return HCX_ERR_INVALID_ARG;
  }
  hcx_status st = validate_opts(o);
  if (st != HCX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  const long long P = (long long)sample->num_instances * p->K;
  double2* roots = nullptr;
  uint32_t* flags = nullptr;
  HIP_TRY(hipMalloc(&roots, sizeof(double2) * P * p->n));
  HIP_TRY(hipMalloc(&flags, sizeof(uint32_t) * P));
  hcx_result_desc r;
  std::memset(&r, 0, sizeof(r));
  r.d_roots = reinterpret_cast<hcx_complex*>(roots);
  r.d_flags = flags;
  const int G0 = default_lanes(p->n, p->C, p->M);
  int pow2 = 1;
  while (pow2 < G0) pow2 <<= 1;
  struct Cand { int G, spec, wpc; };
  std::vector<Cand> cands;
  int occ = 0;
  {
    // occupancy of the default shape, to derive the half-density candidate
    Plan pl;
    int G = G0;
    if (choose_kernel(1, &p, &G, p->deg, o->precision, 1, pl) == HCX_OK) {
      Launch LN;
      std::memset(&LN, 0, sizeof(LN));
      SegOut O;
      LN.nseg = 1;
      seg_from_problem(LN.seg[0], O, p, G, o->precision, pl.nmax, pl.tab4);
      if (finish_plan(LN, pl) == HCX_OK)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(pl.fn), 64, pl.lds);
    }
  }
  for (int spec = 1; spec >= (G0 == p->n ? 0 : 1); --spec) {
    cands.push_back({G0, spec, 0});
    if (occ > 1) cands.push_back({G0, spec, std::max(1, occ / 2)});
  }
  if (pow2 != G0 && pow2 <= 64) {
    cands.push_back({pow2, 0, 0});
    if (occ > 1) cands.push_back({pow2, 0, std::max(1, occ / 2)});
  }
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  double best = 1e300;
  Cand bc = cands[0];
  const bool was_tuned = p->tuned;
  p->tuned = false;
  int timed = 0;
  for (const Cand& c : cands) {
    hcx_track_opts oc = *o;
    oc.lanes_per_path = c.G;
    oc.specialize = c.spec;
    oc.waves_per_cu = c.wpc;
    st = track_pooled(1, &p, sample, &r, &oc, s, 1, nullptr, "hcx_autotune");  // warm-up
    if (st != HCX_OK) continue;
    HIP_TRY(hipEventRecord(e0, s));
    const int reps = 2;
    for (int q = 0; q < reps && st == HCX_OK; ++q)
      st = track_pooled(1, &p, sample, &r, &oc, s, 1, nullptr, "hcx_autotune");
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    if (st != HCX_OK) continue;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    ++timed;
    if (ms / reps < best) {
      best = ms / reps;
      bc = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(roots);
  (void)hipFree(flags);
  if (timed == 0) {
    p->tuned = was_tuned;
    char msg[512];
    std::snprintf(msg, sizeof(msg), "%s", hcx_last_error());
    set_error("hcx_autotune: no candidate ran (%s)", msg);
    return HCX_ERR_UNSUPPORTED;
  }
  p->tuned = true;
  p->tuned_G = bc.G;
  p->tuned_spec = bc.spec;
  p->tuned_wpc = bc.wpc;
  if (out) {
    out->lanes_per_path = bc.G;
    out->specialize = bc.spec;
    out->waves_per_cu = bc.wpc;
    out->best_ms = best;
    out->candidates = timed;
  }
  return HCX_OK;
}

hcx_status hcx_problem_set_launch_shape(hcx_problem* p, int lanes_per_path, int specialize, int waves_per_cu) {
  clear_error();
  if (!p || lanes_per_path < 0 || lanes_per_path > 64 || (lanes_per_path > 0 && lanes_per_path < p->n) ||
      specialize < 0 || specialize > 1 || waves_per_cu < 0) {
    set_error("hcx_problem_set_launch_shape: invalid shape");
    return HCX_ERR_INVALID_ARG;
  }
  p->tuned = lanes_per_path > 0;
  p->tuned_G = lanes_per_path;
  p->tuned_spec = specialize;
  p->tuned_wpc = waves_per_cu;
  return HCX_OK;
}

}  // extern "C"

struct hcx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the graph's own launch workspace (path queues, output descriptors,
  // instance counters, re-track list): baked into the graph, freed with it
  WsBlock ws;
  hipEvent_t last = nullptr;  // completion of the latest replay
};

extern "C" {

// Captures one multi-segment persistent launch (counter set-up, device P2C,
// tracking, and in mixed precision the FP64 re-track pass) into a HIP graph.
// Counters are cleared by kernels, not memset nodes.  The arguments are baked
// into the graph: the same device buffers are reused on every
// hcx_graph_launch.  The launch workspace belongs to the graph, so eager
// launches of the same problems on any stream, of any size, never touch the
// memory a replay uses.
hcx_status hcx_graph_capture_multi(int num_segments, hcx_problem* const* problems, const hcx_batch_desc* batches,
                                   const hcx_result_desc* results, const hcx_track_opts* o, hcx_graph** out) {
  clear_error();
  if (!out || num_segments < 1 || num_segments > kMaxSeg || !problems || !batches || !results || !o) {
    set_error("hcx_graph_capture_multi: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  for (int s = 0; s < num_segments; ++s)
    if (!problems[s]) {
      set_error("hcx_graph_capture_multi: segment %d has no problem", s);
      return HCX_ERR_INVALID_ARG;
    }
  auto* g = new hcx_graph();
  long long inst = 0, paths = 0;
  ws_need(num_segments, problems, batches, results, o, inst, paths);
  const WsLayout L = ws_layout(sizeof(SegOut), inst, paths);
  hipError_t e = hipMalloc(&g->ws.base, L.bytes);
  if (e != hipSuccess) {
    delete g;
    set_error("hcx_graph_capture_multi: workspace allocation: %s", hipGetErrorString(e));
    return HCX_ERR_OUT_OF_MEMORY;
  }
  g->ws.inst_cap = inst;
  g->ws.paths_cap = paths;
  g->ws.pinned = true;
  LaunchWork work[kMaxSeg];
  ws_carve(g->ws, num_segments, batches, results, problems, o, work);
  // an uncaptured run first: sets kernel attributes outside the capture
  hipStream_t cs = nullptr;
  hcx_status st = HCX_OK;
  if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&g->last, hipEventDisableTiming)) != hipSuccess) {
    set_error("hcx_graph_capture_multi: %s", hipGetErrorString(e));
    st = HCX_ERR_HIP;
  }
  if (st == HCX_OK)
    st = track_impl(num_segments, problems, batches, results, o, cs, 1,
                    reinterpret_cast<unsigned long long*>(results[0].d_batch_start), work);
  if (st == HCX_OK && (e = hipStreamSynchronize(cs)) != hipSuccess) {
    set_error("hcx_graph_capture_multi: %s", hipGetErrorString(e));
    st = HCX_ERR_HIP;
  }
  if (st == HCX_OK) {
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      st = track_impl(num_segments, problems, batches, results, o, cs, 1,
                      reinterpret_cast<unsigned long long*>(results[0].d_batch_start), work);
      hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
      if (st == HCX_OK && e2 != hipSuccess) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(e2));
        st = HCX_ERR_HIP;
      }
    } else {
      set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
      st = HCX_ERR_HIP;
    }
  }
  if (st == HCX_OK) {
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
      st = HCX_ERR_HIP;
    }
  }
  if (cs) (void)hipStreamDestroy(cs);
  if (st != HCX_OK) {
    hcx_graph_destroy(g);
    return st;
  }
  *out = g;
  return HCX_OK;
}

hcx_status hcx_graph_launch(hcx_graph* g, void* stream) {
  clear_error();
  if (!g || !g->exec) {
    set_error("hcx_graph_launch: no graph");
    return HCX_ERR_INVALID_ARG;
  }
  HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
  HIP_TRY(hipEventRecord(g->last, (hipStream_t)stream));
  return HCX_OK;
}

// Waits for the latest replay, then frees the graph and its workspace.
hcx_status hcx_graph_destroy(hcx_graph* g) {
  if (!g) return HCX_OK;
  if (g->last) {
    (void)hipEventSynchronize(g->last);
    (void)hipEventDestroy(g->last);
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->ws.base) (void)hipFree(g->ws.base);
  delete g;
  return HCX_OK;
}

}  // extern "C"
