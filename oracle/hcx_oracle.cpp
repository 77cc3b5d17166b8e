// hcx oracle — TEST INFRASTRUCTURE, not part of the product.
//
// A plain FP64 CPU restatement of the reference path tracker
// (C-H-Chien/Homotopy-Continuation-Tracker-on-GPU), used only by tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker.
// The product (libhcx.so) never links, loads or calls this code.
//
// Functions and the reference lines they follow:
//   orc_coeffs      eval_parameter_homotopy          dev-eval-indxing-generalized_3views_4pts.cuh:34-67
//   orc_eval        eval_Jacobian_Hx / _Ht / Homotopy dev-eval-indxing-generalized_3views_4pts.cuh:70-120
//   orc_lu_solve    cgesv_batched_small_device        dev-cgesv-batched-small.cuh:38-107
//   orc_track       HC_solver_<p> step loop           kernel_HC_Solver_generalized_3views_4pts.cu:96-264
//   orc_p2c_affine  Construct_Coeffs_From_Params      Data_Reader.cpp:59-89 (affine closed form)
//   orc_collect     Evaluate_GPUHC_Sols / Find_Unique_Sols  Evaluations.cpp:88-176
//                   (F_REAL_REF is set by orc_track on the tracked root, :96-104)
// Additions required by the tracker's contract (not in the reference): final
// FP64 Newton refinement at t = 1, residual, real classification on refined
// roots, optional minimum step.
//
// Pinning (see tests/test_oracle_*.py): the LU against LAPACK zgetrf pivots and
// numpy solves; the evaluator against direct polynomial evaluation and finite
// differences; tables and start solutions against the reference's committed
// problem data (H(start_sols, t=0) = 0); the tracker against closed forms
// (linear homotopies, polynomial solution paths) and Bezout root counts.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

using cd = std::complex<double>;

extern "C" {

typedef struct {
  int n, C, M;
  int hx_terms, hx_parts, ht_terms, ht_parts;
  const int* hx_idx;  // [n eq][n var][hx_terms][hx_parts]
  const int* ht_idx;  // [n eq][ht_terms][ht_parts]
} orc_tables;

typedef struct {
  int max_steps, max_corr, incr_steps;
  double dt0, end_zone, t_tol, dt_min, corr_tol, inf_tol;
  int refine_iters;
  double refine_tol, real_tol;
  double real_ref_tol;  // the reference's real test on the tracked root (definitions.hpp:25)
} orc_opts;

}  // extern "C"

namespace {

const uint32_t F_TRACKED = 0x01, F_INF = 0x02, F_REFINED = 0x04, F_REAL = 0x08, F_UNIQUE = 0x10, F_BUDGET = 0x20,
               F_MINSTEP = 0x40, F_SINGULAR = 0x80, F_REAL_REF = 0x200, F_UNIQUE_REF = 0x400;

// c_j(t) and dc_j/dt(t) from the coefficient-polynomial-in-t tables.
void coeffs(const orc_tables& T, const cd* hx, const cd* ht, double t, std::vector<cd>& c, std::vector<cd>& cdot) {
  const int C1 = T.C + 1, M = T.M;
  c.assign(C1, cd(0, 0));
  cdot.assign(C1, cd(0, 0));
  for (int j = 0; j < C1; ++j) {
    cd v(0, 0), tp(1, 0);
    for (int m = 0; m <= M; ++m) {
      v += hx[(size_t)j * (M + 1) + m] * tp;
      tp *= t;
    }
    c[j] = v;
    cd d(0, 0);
    tp = cd(1, 0);
    for (int m = 0; m < M; ++m) {
      d += ht[(size_t)j * M + m] * tp;
      tp *= t;
    }
    cdot[j] = d;
  }
}

cd term_value(const int* rec, int nvars, const std::vector<cd>& coef, const cd* xh) {
  cd v = (double)rec[0] * coef[rec[1]];
  for (int q = 0; q < nvars; ++q) v *= xh[rec[2 + q]];
  return v;
}

// H (with c), dH/dt (with cdot), dH/dx; x has n entries, xh = [x, 1].
void eval_all(const orc_tables& T, const std::vector<cd>& c, const std::vector<cd>& cdot, const cd* x, cd* H, cd* J,
              cd* Ht) {
  const int n = T.n;
  std::vector<cd> xh(n + 1);
  for (int i = 0; i < n; ++i) xh[i] = x[i];
  xh[n] = cd(1, 0);
  const int hv = T.ht_parts - 2, xv = T.hx_parts - 2;
  for (int i = 0; i < n; ++i) {
    cd h(0, 0), hd(0, 0);
    for (int k = 0; k < T.ht_terms; ++k) {
      const int* rec = T.ht_idx + ((size_t)i * T.ht_terms + k) * T.ht_parts;
      if (H) h += term_value(rec, hv, c, xh.data());
      if (Ht) hd += term_value(rec, hv, cdot, xh.data());
    }
    if (H) H[i] = h;
    if (Ht) Ht[i] = hd;
    if (J)
      for (int j = 0; j < n; ++j) {
        cd a(0, 0);
        for (int k = 0; k < T.hx_terms; ++k) {
          const int* rec = T.hx_idx + (((size_t)i * n + j) * T.hx_terms + k) * T.hx_parts;
          a += term_value(rec, xv, c, xh.data());
        }
        J[(size_t)i * n + j] = a;
      }
  }
}

double cabs1(cd z) { return std::fabs(z.real()) + std::fabs(z.imag()); }

// LU with partial pivoting (cabs1, first maximum, physical row interchanges),
// multiply-by-reciprocal elimination; zero pivot columns are skipped.  Then a
// column-oriented back substitution.  A (row-major) and b are overwritten.
int lu_solve(int n, cd* A, cd* b, cd* x, int* piv) {
  int info = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double mx = cabs1(A[(size_t)k * n + k]);
    for (int r = k + 1; r < n; ++r) {
      const double v = cabs1(A[(size_t)r * n + k]);
      if (v > mx) {
        mx = v;
        p = r;
      }
    }
    if (piv) piv[k] = p;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
      std::swap(b[k], b[p]);
    }
    if (mx == 0.0) {
      if (!info) info = k + 1;
      continue;
    }
    const cd rinv = cd(1, 0) / A[(size_t)k * n + k];
    for (int r = k + 1; r < n; ++r) {
      const cd l = A[(size_t)r * n + k] * rinv;
      for (int j = k + 1; j < n; ++j) A[(size_t)r * n + j] -= l * A[(size_t)k * n + j];
      b[r] -= l * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    // MAGMA_C_DIV by a zero U_ii (a skipped zero pivot column) yields NaN in
    // the reference (dev-cgesv-batched-small.cuh:97-106), not the infinity
    // of C99 complex division
    const cd uii = A[(size_t)i * n + i];
    const cd xi = (uii == cd(0, 0)) ? cd(NAN, NAN) : b[i] / uii;
    x[i] = xi;
    for (int r = 0; r < i; ++r) b[r] -= A[(size_t)r * n + i] * xi;
  }
  return info;
}

struct PathOut {
  uint32_t flags = 0;
  int steps = 0, newton = 0, refine = 0;
  double t = 0, dt = 0, resid = 0;
};

// Solve J(x,t) d = rhs where rhs = H (newton) or -dH/dt (predictor).
bool solve_at(const orc_tables& T, const cd* hx, const cd* ht, const cd* x, double t, bool predictor, cd* d,
              std::vector<cd>& c, std::vector<cd>& cdot, std::vector<cd>& J, std::vector<cd>& r) {
  const int n = T.n;
  coeffs(T, hx, ht, t, c, cdot);
  if (predictor) {
    eval_all(T, c, cdot, x, nullptr, J.data(), r.data());
    for (int i = 0; i < n; ++i) r[i] = -r[i];
  } else {
    eval_all(T, c, cdot, x, r.data(), J.data(), nullptr);
  }
  return lu_solve(n, J.data(), r.data(), d, nullptr) == 0;
}

PathOut track_one(const orc_tables& T, const cd* hx, const cd* ht, const cd* start, const orc_opts& o, cd* xout) {
  const int n = T.n;
  std::vector<cd> x(start, start + n), xl = x, k1(n), k2(n), k3(n), k4(n), xs(n), d(n);
  std::vector<cd> c, cdot, J((size_t)n * n), r(n);
  PathOut out;
  double t = 0.0, dt = o.dt0;
  bool ez = false, inf = false, minstep = false;
  int cnt = 0;
  for (int step = 0; step <= o.max_steps; ++step) {
    if (!(t < 1.0 && 1.0 - t > o.t_tol)) break;
    if (!ez && std::fabs(1.0 - t) <= o.end_zone + 1e-7) ez = true;
    if (ez) {
      if (dt > std::fabs(1.0 - t)) dt = std::fabs(1.0 - t);
    } else if (dt > std::fabs(1.0 - o.end_zone - t)) {
      dt = std::fabs(1.0 - o.end_zone - t);
    }
    const double ts = t, h = 0.5 * dt;
    solve_at(T, hx, ht, xl.data(), ts, true, k1.data(), c, cdot, J, r);
    for (int i = 0; i < n; ++i) xs[i] = xl[i] + h * k1[i];
    solve_at(T, hx, ht, xs.data(), ts + h, true, k2.data(), c, cdot, J, r);
    for (int i = 0; i < n; ++i) xs[i] = xl[i] + h * k2[i];
    solve_at(T, hx, ht, xs.data(), ts + h, true, k3.data(), c, cdot, J, r);
    for (int i = 0; i < n; ++i) xs[i] = xl[i] + dt * k3[i];
    solve_at(T, hx, ht, xs.data(), ts + dt, true, k4.data(), c, cdot, J, r);
    for (int i = 0; i < n; ++i) x[i] = xl[i] + (dt / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
    t = ts + dt;
    bool ok = false;
    inf = false;
    for (int it = 0; it < o.max_corr; ++it) {
      solve_at(T, hx, ht, x.data(), t, false, d.data(), c, cdot, J, r);
      double rs = 0, rc = 0;
      for (int i = 0; i < n; ++i) {
        x[i] -= d[i];
        rs += std::norm(d[i]);
        rc += std::norm(x[i]);
      }
      ++out.newton;
      ok = rs < o.corr_tol * rc;
      inf = rc > o.inf_tol;
      if (inf || ok) break;
    }
    ++out.steps;
    if (inf) break;
    if (!ok) {
      dt *= 0.5;
      x = xl;
      t = ts;
      cnt = 0;
      if (dt < o.dt_min) {
        minstep = true;
        break;
      }
    } else {
      ++cnt;
      xl = x;
      if (cnt >= o.incr_steps) {
        cnt = 0;
        dt *= 2.0;
      }
    }
  }
  if (inf) out.flags |= F_INF;
  else if (minstep) out.flags |= F_MINSTEP;
  else if (t >= 1.0 || 1.0 - t <= o.t_tol) out.flags |= F_TRACKED;
  else out.flags |= F_BUDGET;
  out.t = t;
  out.dt = dt;
  // Evaluate_GPUHC_Sols's real test (Evaluations.cpp:96-104) on the tracked
  // root: converged and every |Im x_i| <= ZERO_IMAG_PART_TOL_FOR_SP
  if (out.flags & F_TRACKED) {
    int real_vars = 0;
    for (int i = 0; i < n; ++i) real_vars += std::fabs(x[i].imag()) <= o.real_ref_tol ? 1 : 0;
    if (real_vars == n) out.flags |= F_REAL_REF;
  }
  if ((out.flags & F_TRACKED) && o.refine_iters > 0) {
    bool conv = false, sing = false;
    for (int it = 0; it < o.refine_iters; ++it) {
      coeffs(T, hx, ht, 1.0, c, cdot);
      eval_all(T, c, cdot, x.data(), r.data(), J.data(), nullptr);
      sing = lu_solve(n, J.data(), r.data(), d.data(), nullptr) != 0;
      double nd = 0, nx = 0;
      for (int i = 0; i < n; ++i) {
        x[i] -= d[i];
        nd += std::norm(d[i]);
        nx += std::norm(x[i]);
      }
      ++out.refine;
      if (!(nx < o.inf_tol)) break;
      if (std::sqrt(nd) <= o.refine_tol * std::max(1.0, std::sqrt(nx))) {
        conv = true;
        break;
      }
    }
    coeffs(T, hx, ht, 1.0, c, cdot);
    eval_all(T, c, cdot, x.data(), r.data(), nullptr, nullptr);
    double res = 0, im = 0;
    for (int i = 0; i < n; ++i) {
      double a = std::abs(r[i]);
      if (a != a) a = INFINITY;
      res = std::max(res, a);
      double b = std::fabs(x[i].imag());
      if (b != b) b = INFINITY;
      im = std::max(im, b);
    }
    out.resid = res;
    if (conv) out.flags |= F_REFINED;
    if (conv && im <= o.real_tol) out.flags |= F_REAL;
    if (sing) out.flags |= F_SINGULAR;
  }
  for (int i = 0; i < n; ++i) xout[i] = x[i];
  return out;
}

}  // namespace

extern "C" {

int orc_abi_version(void) { return 1; }

// c_j(t), dc_j/dt(t) for j = 0..C; out arrays hold 2*(C+1) doubles each.
int orc_coeffs(const orc_tables* T, const double* hx, const double* ht, double t, double* c, double* cdot) {
  std::vector<cd> cc, dd;
  coeffs(*T, reinterpret_cast<const cd*>(hx), reinterpret_cast<const cd*>(ht), t, cc, dd);
  std::memcpy(c, cc.data(), cc.size() * sizeof(cd));
  std::memcpy(cdot, dd.data(), dd.size() * sizeof(cd));
  return 0;
}

// H [n], J [n][n] (row = equation), dH/dt [n] at (x, t); any output may be NULL.
int orc_eval(const orc_tables* T, const double* hx, const double* ht, const double* x, double t, double* H, double* J,
             double* Ht) {
  std::vector<cd> c, cdot;
  coeffs(*T, reinterpret_cast<const cd*>(hx), reinterpret_cast<const cd*>(ht), t, c, cdot);
  eval_all(*T, c, cdot, reinterpret_cast<const cd*>(x), reinterpret_cast<cd*>(H), reinterpret_cast<cd*>(J),
           reinterpret_cast<cd*>(Ht));
  return 0;
}

// Solves A x = b (A row-major n x n complex); piv[k] = row interchanged with k.
// Returns info (0, or k+1 for the first zero pivot column).
int orc_lu_solve(int n, const double* A, const double* b, double* x, int* piv) {
  std::vector<cd> a(reinterpret_cast<const cd*>(A), reinterpret_cast<const cd*>(A) + (size_t)n * n);
  std::vector<cd> bb(reinterpret_cast<const cd*>(b), reinterpret_cast<const cd*>(b) + n);
  return lu_solve(n, a.data(), bb.data(), reinterpret_cast<cd*>(x), piv);
}

// Parameter homotopy coefficient tables from the affine map A [C][1+k]:
// hx [I][C+1][2] = {c(p0), c(p1)-c(p0)}, ht [I][C+1] = {c(p1)-c(p0)}; entry C = {1,0}/{0}.
int orc_p2c_affine(int C, int k, const double* A, const double* p0, const double* p1, int I, double* hx, double* ht) {
  const cd* a = reinterpret_cast<const cd*>(A);
  const cd* q0 = reinterpret_cast<const cd*>(p0);
  const cd* q1 = reinterpret_cast<const cd*>(p1);
  cd* ox = reinterpret_cast<cd*>(hx);
  cd* ot = reinterpret_cast<cd*>(ht);
  for (int i = 0; i < I; ++i) {
    for (int j = 0; j < C; ++j) {
      cd c0 = a[(size_t)j * (1 + k)], c1 = c0;
      for (int q = 0; q < k; ++q) {
        c0 += a[(size_t)j * (1 + k) + 1 + q] * q0[q];
        c1 += a[(size_t)j * (1 + k) + 1 + q] * q1[(size_t)i * k + q];
      }
      ox[((size_t)i * (C + 1) + j) * 2] = c0;
      ox[((size_t)i * (C + 1) + j) * 2 + 1] = c1 - c0;
      ot[(size_t)i * (C + 1) + j] = c1 - c0;
    }
    ox[((size_t)i * (C + 1) + C) * 2] = cd(1, 0);
    ox[((size_t)i * (C + 1) + C) * 2 + 1] = cd(0, 0);
    ot[(size_t)i * (C + 1) + C] = cd(0, 0);
  }
  return 0;
}

// Parameter-polynomial coefficient map c_j(p) = sum_u s[j][u] prod_d p[idx[j][u][d]]
// (p[k] = 1) on the path p(t) = p0 + t (p1 - p0): the t^m coefficients of c_j
// (hx [I][C+1][M+1]) and of dc_j/dt (ht [I][C+1][M]); entry C is the constant
// 1.  The restatement of the generated P2C code (Data_Reader.cpp:59-89 with
// the createUnivPoly4Coeffs.m layout): each term's product of linear factors
// is expanded factor by factor.
int orc_p2c_poly(int C, int T, int D, int k, int M, const double* scal, const int* idx, const double* p0,
                 const double* p1, int I, double* hx, double* ht) {
  const cd* sc = reinterpret_cast<const cd*>(scal);
  const cd* q0 = reinterpret_cast<const cd*>(p0);
  const cd* q1 = reinterpret_cast<const cd*>(p1);
  cd* ox = reinterpret_cast<cd*>(hx);
  cd* ot = reinterpret_cast<cd*>(ht);
  std::vector<cd> acc(M + 1), term(M + 2);
  for (int i = 0; i < I; ++i) {
    for (int j = 0; j <= C; ++j) {
      std::fill(acc.begin(), acc.end(), cd(0, 0));
      if (j == C) {
        acc[0] = cd(1, 0);
      } else {
        for (int u = 0; u < T; ++u) {
          const cd s = sc[(size_t)j * T + u];
          if (s == cd(0, 0)) continue;
          std::fill(term.begin(), term.end(), cd(0, 0));
          term[0] = s;
          int deg = 0;
          for (int d = 0; d < D; ++d) {
            const int a = idx[((size_t)j * T + u) * D + d];
            if (a == k) continue;
            if (a < 0 || a > k || deg + 1 > M) return -1;
            const cd qa = q0[a], da = q1[(size_t)i * k + a] - q0[a];
            for (int m = deg + 1; m >= 0; --m) term[m] = term[m] * qa + (m > 0 ? term[m - 1] * da : cd(0, 0));
            ++deg;
          }
          for (int m = 0; m <= M; ++m) acc[m] += term[m];
        }
      }
      for (int m = 0; m <= M; ++m) ox[((size_t)i * (C + 1) + j) * (M + 1) + m] = acc[m];
      for (int m = 0; m < M; ++m) ot[((size_t)i * (C + 1) + j) * M + m] = (double)(m + 1) * acc[m + 1];
    }
  }
  return 0;
}

// Tracks every (instance, root) path: path = inst * K + root.  hx/ht are
// [I][(C+1)(M+1)] / [I][(C+1)M].  Outputs per path (any may be NULL except
// roots and flags).  threads <= 0: hardware concurrency.
int orc_track(const orc_tables* T, int K, const double* start, int I, const double* hx, const double* ht,
              const orc_opts* o, int threads, double* roots, uint32_t* flags, int* steps, int* newton, int* refine,
              double* tfin, double* dtfin, double* resid) {
  const long long P = (long long)I * K;
  const int n = T->n;
  const size_t sx = (size_t)(T->C + 1) * (T->M + 1), st = (size_t)(T->C + 1) * T->M;
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  threads = (int)std::min<long long>(threads, std::max<long long>(P, 1));
  auto work = [&](long long lo, long long hi) {
    for (long long p = lo; p < hi; ++p) {
      const long long inst = p / K, root = p % K;
      PathOut r = track_one(*T, reinterpret_cast<const cd*>(hx) + inst * sx, reinterpret_cast<const cd*>(ht) + inst * st,
                            reinterpret_cast<const cd*>(start) + root * n, *o, reinterpret_cast<cd*>(roots) + p * n);
      flags[p] = r.flags;
      if (steps) steps[p] = r.steps;
      if (newton) newton[p] = r.newton;
      if (refine) refine[p] = r.refine;
      if (tfin) tfin[p] = r.t;
      if (dtfin) dtfin[p] = r.dt;
      if (resid) resid[p] = r.resid;
    }
  };
  if (threads == 1) {
    work(0, P);
    return 0;
  }
  std::vector<std::thread> pool;
  const long long chunk = (P + threads - 1) / threads;
  for (int w = 0; w < threads; ++w) {
    const long long lo = w * chunk, hi = std::min(P, lo + chunk);
    if (lo >= hi) break;
    pool.emplace_back(work, lo, hi);
  }
  for (auto& th : pool) th.join();
  return 0;
}

// Counts: [tracked, infinity, refined, real, unique, unique_ref, real_ref].
//  * unique: F_UNIQUE on the first occurrence of each refined root;
//  * real_ref: tracked paths with F_REAL_REF (Evaluate_GPUHC_Sols, Evaluations.cpp:88-110);
//  * unique_ref: Find_Unique_Sols (Evaluations.cpp:127-176) restated with its
//    own sets: for each converged bs not in Skip, Duplicate = every later ds
//    (converged or not) whose root agrees in every coordinate; an empty set
//    counts bs (F_UNIQUE_REF), a non-empty set replaces Skip.  Duplicate is
//    flushed only when Skip is non-empty, as in the reference.
int orc_collect(int n, int I, int K, const double* roots, uint32_t* flags, double tol, long long* counts) {
  for (int q = 0; q < 7; ++q) counts[q] = 0;
  const cd* x = reinterpret_cast<const cd*>(roots);
  auto same = [&](size_t a, size_t b) {
    for (int v = 0; v < n; ++v) {
      const cd p = x[a * n + v], q = x[b * n + v];
      if (!(std::fabs(p.real() - q.real()) < tol && std::fabs(p.imag() - q.imag()) < tol)) return false;
    }
    return true;
  };
  for (int inst = 0; inst < I; ++inst) {
    const size_t base = (size_t)inst * K;
    for (int j = 0; j < K; ++j) {
      uint32_t& f = flags[base + j];
      f &= ~(F_UNIQUE | F_UNIQUE_REF);
      if (f & F_TRACKED) counts[0]++;
      if (f & F_INF) counts[1]++;
      if ((f & F_TRACKED) && (f & F_REAL_REF)) counts[6]++;
      if (!(f & F_REFINED)) continue;
      counts[2]++;
      if (f & F_REAL) counts[3]++;
      bool dup = false;
      for (int i = 0; i < j && !dup; ++i) dup = (flags[base + i] & F_REFINED) && same(base + i, base + j);
      if (!dup) {
        f |= F_UNIQUE;
        counts[4]++;
      }
    }
    // Find_Unique_Sols, literally: std::set Skip / Duplicate
    std::vector<int> skip, dupl;
    for (int bs = 0; bs < K; ++bs) {
      if (!(flags[base + bs] & F_TRACKED)) continue;
      if (!skip.empty()) {
        if (std::find(skip.begin(), skip.end(), bs) != skip.end()) continue;
        dupl.clear();
      }
      for (int ds = bs + 1; ds < K; ++ds)
        if (same(base + bs, base + ds) && std::find(dupl.begin(), dupl.end(), ds) == dupl.end()) dupl.push_back(ds);
      if (dupl.empty()) {
        flags[base + bs] |= F_UNIQUE_REF;
        counts[5]++;
      } else {
        skip = dupl;
      }
    }
  }
  return 0;
}

}  // extern "C"
