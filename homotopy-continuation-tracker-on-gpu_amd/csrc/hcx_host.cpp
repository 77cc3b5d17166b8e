// Host side of libhcx: errors, options, random streams, parametric polynomial
// systems (synthetic generator, reference-format table compiler, P2C) and the
// total-degree start system.
//
// Table layout follows the reference's on-disk format:
//   dHdt_indx [eq][term][part] with parts {scalar, coeff, var x (parts-2)}
//   dHdx_indx [eq][var][term][part] with the same part layout
//   var index n means "times 1"; padding terms are {0, C, n, ...}
// (auto-data-gen-tools/autogen_gpuhc.m:503-699, read by
//  GPU-HC/magmaHC/gpu-idx-evals/dev-eval-indxing-generalized_3views_4pts.cuh:70-120).
#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <unordered_set>

#include "hcx_internal.hpp"

namespace hcx {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }

uint64_t sub_seed(uint64_t seed, uint64_t tag) {
  SplitMix64 r(seed ^ (0xD1B54A32D192ED03ull * (tag + 1)));
  r.next();
  return r.next();
}

void System::compile() {
  // equation membership and degrees
  std::vector<std::vector<int>> mono_of_eq(n);
  degree_of_eq.assign(n, 0);
  max_degree = 1;
  for (int j = 0; j < C; ++j) {
    mono_of_eq[eq[j]].push_back(j);
    int d = 0;
    for (int v = 0; v < n; ++v) d += exps[(size_t)j * n + v];
    degree_of_eq[eq[j]] = std::max(degree_of_eq[eq[j]], d);
    max_degree = std::max(max_degree, d);
  }
  max_terms = 0;
  for (int i = 0; i < n; ++i) max_terms = std::max<int>(max_terms, (int)mono_of_eq[i].size());

  const int ht_vars = max_degree;                  // H term: all factors
  const int hx_vars = std::max(max_degree - 1, 1); // dH/dx term: one factor removed
  ht_parts = 2 + ht_vars;
  hx_parts = 2 + hx_vars;
  ht_terms = std::max(max_terms, 1);

  // per-entry derivative term counts
  hx_terms = 1;
  for (int i = 0; i < n; ++i)
    for (int v = 0; v < n; ++v) {
      int cnt = 0;
      for (int j : mono_of_eq[i]) cnt += exps[(size_t)j * n + v] > 0;
      hx_terms = std::max(hx_terms, cnt);
    }

  auto factors = [&](int j, int drop_var, std::vector<int>& out) {
    out.clear();
    for (int v = 0; v < n; ++v) {
      int e = exps[(size_t)j * n + v] - (v == drop_var ? 1 : 0);
      for (int r = 0; r < e; ++r) out.push_back(v);
    }
  };

  ht_idx.assign((size_t)n * ht_terms * ht_parts, 0);
  hx_idx.assign((size_t)n * n * hx_terms * hx_parts, 0);
  std::vector<int> f;
  for (int i = 0; i < n; ++i) {
    for (int t = 0; t < ht_terms; ++t) {
      int* rec = &ht_idx[((size_t)i * ht_terms + t) * ht_parts];
      if (t < (int)mono_of_eq[i].size()) {
        const int j = mono_of_eq[i][t];
        factors(j, -1, f);
        rec[0] = 1;
        rec[1] = j;
        for (int q = 0; q < ht_vars; ++q) rec[2 + q] = q < (int)f.size() ? f[q] : n;
      } else {
        rec[0] = 0;
        rec[1] = C;
        for (int q = 0; q < ht_vars; ++q) rec[2 + q] = n;
      }
    }
    for (int v = 0; v < n; ++v) {
      int t = 0;
      for (int j : mono_of_eq[i]) {
        const int e = exps[(size_t)j * n + v];
        if (e <= 0) continue;
        int* rec = &hx_idx[(((size_t)i * n + v) * hx_terms + t) * hx_parts];
        factors(j, v, f);
        rec[0] = e;
        rec[1] = j;
        for (int q = 0; q < hx_vars; ++q) rec[2 + q] = q < (int)f.size() ? f[q] : n;
        ++t;
      }
      for (; t < hx_terms; ++t) {
        int* rec = &hx_idx[(((size_t)i * n + v) * hx_terms + t) * hx_parts];
        rec[0] = 0;
        rec[1] = C;
        for (int q = 0; q < hx_vars; ++q) rec[2 + q] = n;
      }
    }
  }
}

// All monomials of total degree 1..d in n variables, graded then lexicographic
// (exponent vectors compared from variable 0 down).
static void enumerate_monomials(int n, int d, std::vector<std::vector<int>>& out) {
  out.clear();
  std::vector<int> e(n, 0);
  for (int deg = 1; deg <= d; ++deg) {
    // compositions of deg into n parts, lexicographically descending in e[0]
    std::function<void(int, int)> rec = [&](int v, int left) {
      if (v == n - 1) {
        e[v] = left;
        out.push_back(e);
        return;
      }
      for (int a = left; a >= 0; --a) {
        e[v] = a;
        rec(v + 1, left - a);
      }
    };
    rec(0, deg);
  }
}

}  // namespace hcx

using namespace hcx;

extern "C" {

const char* hcx_last_error(void) { return g_err; }
int hcx_abi_version(void) { return HCX_ABI_VERSION; }

void hcx_track_opts_default(hcx_track_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_steps = 200;
  o->max_corr = 3;
  o->incr_steps = 3;
  o->dt0 = 0.05;
  o->end_zone = 0.05;
  o->t_tol = 1e-7;
  o->dt_min = 1e-7;
  o->corr_tol = 1e-6;
  o->inf_tol = 1e14;
  o->refine_iters = 8;
  o->refine_tol = 1e-10;
  o->real_tol = 1e-8;
  o->unique_tol = 1e-4;
  o->precision = HCX_PREC_FP64;
  o->lanes_per_path = 0;
  o->waves_per_cu = 0;
  o->specialize = 1;
  o->retrack_duplicates = 1;
  o->fp32_corrector = 0;
  o->real_ref_tol = 1e-4;
  o->replicas = 0;
}

hcx_status hcx_random_cn(uint64_t seed, int64_t count, hcx_complex* out) {
  clear_error();
  if (count < 0 || (count > 0 && !out)) {
    set_error("hcx_random_cn: bad count/out");
    return HCX_ERR_INVALID_ARG;
  }
  SplitMix64 r(seed);
  for (int64_t i = 0; i < count; ++i) out[i] = to_hc(r.cn());
  return HCX_OK;
}

hcx_complex hcx_random_gamma(uint64_t seed) {
  SplitMix64 r(seed);
  const double u = r.uniform();
  const double th = 6.283185307179586476925286766559 * u;
  return hcx_complex{std::cos(th), std::sin(th)};
}

hcx_status hcx_system_create(int n, int num_params, int num_coeffs, const int* exps, const int* eq,
                             const hcx_complex* affine, hcx_system** out) {
  clear_error();
  if (!out || n < 1 || n > 64 || num_params < 0 || num_coeffs < 1 || !exps || !eq) {
    set_error("hcx_system_create: invalid arguments (n=%d k=%d C=%d)", n, num_params, num_coeffs);
    return HCX_ERR_INVALID_ARG;
  }
  auto* s = new hcx_system();
  s->n = n;
  s->k = num_params;
  s->C = num_coeffs;
  s->exps.assign(exps, exps + (size_t)num_coeffs * n);
  s->eq.assign(eq, eq + num_coeffs);
  for (int j = 0; j < num_coeffs; ++j) {
    if (eq[j] < 0 || eq[j] >= n) {
      delete s;
      set_error("hcx_system_create: monomial %d has equation %d outside [0,%d)", j, eq[j], n);
      return HCX_ERR_INVALID_ARG;
    }
    for (int v = 0; v < n; ++v)
      if (exps[(size_t)j * n + v] < 0 || exps[(size_t)j * n + v] > 16) {
        delete s;
        set_error("hcx_system_create: exponent out of range at monomial %d", j);
        return HCX_ERR_INVALID_ARG;
      }
  }
  std::vector<int> seen(n, 0);
  for (int j = 0; j < num_coeffs; ++j) seen[eq[j]] = 1;
  for (int i = 0; i < n; ++i)
    if (!seen[i]) {
      delete s;
      set_error("hcx_system_create: equation %d has no monomial", i);
      return HCX_ERR_INVALID_ARG;
    }
  s->affine.assign((size_t)num_coeffs * (1 + num_params), cd(0, 0));
  if (affine)
    for (size_t q = 0; q < s->affine.size(); ++q) s->affine[q] = to_cd(affine[q]);
  s->compile();
  *out = s;
  return HCX_OK;
}

hcx_status hcx_system_create_synthetic(const hcx_synth_spec* spec, hcx_system** out) {
  clear_error();
  if (!spec || !out || spec->num_vars < 1 || spec->num_vars > 32 || spec->num_params < 0 ||
      spec->num_quadratic < 0 || spec->num_quadratic > spec->num_vars) {
    set_error("hcx_system_create_synthetic: invalid spec");
    return HCX_ERR_INVALID_ARG;
  }
  const int n = spec->num_vars, k = spec->num_params;
  std::vector<int> exps, eqv;
  SplitMix64 pick(sub_seed(spec->seed, 1));
  std::vector<std::vector<int>> all;
  for (int i = 0; i < n; ++i) {
    const int d = i < spec->num_quadratic ? 2 : 3;
    enumerate_monomials(n, d, all);
    // mandatory support {1, x_v^d for all v}: the leading forms then have no
    // common projective zero for generic coefficients, so the root count is
    // the Bezout number prod d_i.
    std::vector<std::vector<int>> sup;
    sup.push_back(std::vector<int>(n, 0));
    for (int v = 0; v < n; ++v) {
      std::vector<int> e(n, 0);
      e[v] = d;
      sup.push_back(e);
    }
    std::vector<std::vector<int>> rest;
    for (auto& m : all) {
      int mx = 0, tot = 0;
      for (int v = 0; v < n; ++v) {
        mx = std::max(mx, m[v]);
        tot += m[v];
      }
      if (tot == d && mx == d) continue;  // pure power already present
      rest.push_back(m);
    }
    int extra;
    if (spec->monomials <= 0) extra = (int)rest.size();
    else extra = std::min<int>((int)rest.size(), std::max(0, spec->monomials - (n + 1)));
    if (extra < (int)rest.size()) {
      // seeded partial Fisher-Yates, then restore canonical order
      std::vector<int> idx(rest.size());
      for (size_t q = 0; q < idx.size(); ++q) idx[q] = (int)q;
      for (int q = 0; q < extra; ++q) {
        const int r = q + (int)pick.below(idx.size() - q);
        std::swap(idx[q], idx[r]);
      }
      std::sort(idx.begin(), idx.begin() + extra);
      for (int q = 0; q < extra; ++q) sup.push_back(rest[idx[q]]);
    } else {
      for (auto& m : rest) sup.push_back(m);
    }
    for (auto& m : sup) {
      exps.insert(exps.end(), m.begin(), m.end());
      eqv.push_back(i);
    }
  }
  const int C = (int)eqv.size();
  std::vector<hcx_complex> aff((size_t)C * (1 + k));
  SplitMix64 ra(sub_seed(spec->seed, 2));
  for (auto& z : aff) z = to_hc(ra.cn());
  return hcx_system_create(n, k, C, exps.data(), eqv.data(), aff.data(), out);
}

hcx_status hcx_system_destroy(hcx_system* s) {
  delete s;
  return HCX_OK;
}

hcx_status hcx_system_get_info(const hcx_system* s, hcx_system_info* info) {
  clear_error();
  if (!s || !info) {
    set_error("hcx_system_get_info: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  info->num_vars = s->n;
  info->num_params = s->k;
  info->num_coeffs = s->C;
  info->max_degree = s->max_degree;
  info->max_terms = s->max_terms;
  info->hx_terms = s->hx_terms;
  info->hx_parts = s->hx_parts;
  info->ht_terms = s->ht_terms;
  info->ht_parts = s->ht_parts;
  return HCX_OK;
}

hcx_status hcx_system_get_monomials(const hcx_system* s, int* exps, int* eq) {
  clear_error();
  if (!s) {
    set_error("hcx_system_get_monomials: null system");
    return HCX_ERR_INVALID_ARG;
  }
  if (exps) std::copy(s->exps.begin(), s->exps.end(), exps);
  if (eq) std::copy(s->eq.begin(), s->eq.end(), eq);
  return HCX_OK;
}

hcx_status hcx_system_get_affine(const hcx_system* s, hcx_complex* affine) {
  clear_error();
  if (!s || !affine) {
    set_error("hcx_system_get_affine: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  for (size_t q = 0; q < s->affine.size(); ++q) affine[q] = to_hc(s->affine[q]);
  return HCX_OK;
}

hcx_status hcx_system_compile_tables(const hcx_system* s, int* hx_idx, int* ht_idx) {
  clear_error();
  if (!s) {
    set_error("hcx_system_compile_tables: null system");
    return HCX_ERR_INVALID_ARG;
  }
  if (hx_idx) std::copy(s->hx_idx.begin(), s->hx_idx.end(), hx_idx);
  if (ht_idx) std::copy(s->ht_idx.begin(), s->ht_idx.end(), ht_idx);
  return HCX_OK;
}

hcx_status hcx_system_eval_coeffs(const hcx_system* s, const hcx_complex* params, hcx_complex* coeffs) {
  clear_error();
  if (!s || !coeffs || (s->k > 0 && !params)) {
    set_error("hcx_system_eval_coeffs: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  const int k = s->k;
  for (int j = 0; j < s->C; ++j) {
    const cd* a = &s->affine[(size_t)j * (1 + k)];
    cd c = a[0];
    for (int q = 0; q < k; ++q) c += a[1 + q] * to_cd(params[q]);
    coeffs[j] = to_hc(c);
  }
  return HCX_OK;
}

hcx_status hcx_system_param_homotopy_coeffs(const hcx_system* s, const hcx_complex* p0,
                                            const hcx_complex* p1, hcx_complex* hx, hcx_complex* ht) {
  clear_error();
  if (!s || !hx || !ht) {
    set_error("hcx_system_param_homotopy_coeffs: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  std::vector<hcx_complex> c0(s->C), c1(s->C);
  hcx_status st = hcx_system_eval_coeffs(s, p0, c0.data());
  if (st != HCX_OK) return st;
  st = hcx_system_eval_coeffs(s, p1, c1.data());
  if (st != HCX_OK) return st;
  for (int j = 0; j < s->C; ++j) {
    const cd a = to_cd(c0[j]), d = to_cd(c1[j]) - a;
    hx[2 * j] = to_hc(a);
    hx[2 * j + 1] = to_hc(d);
    ht[j] = to_hc(d);
  }
  hx[2 * s->C] = hcx_complex{1, 0};
  hx[2 * s->C + 1] = hcx_complex{0, 0};
  ht[s->C] = hcx_complex{0, 0};
  return HCX_OK;
}

hcx_status hcx_system_problem_desc(const hcx_system* s, hcx_problem_desc* d) {
  clear_error();
  if (!s || !d) {
    set_error("hcx_system_problem_desc: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  std::memset(d, 0, sizeof(*d));
  d->num_vars = s->n;
  d->num_coeffs = s->C;
  d->max_order_t = 1;
  d->hx_terms = s->hx_terms;
  d->hx_parts = s->hx_parts;
  d->ht_terms = s->ht_terms;
  d->ht_parts = s->ht_parts;
  d->h_hx_idx = s->hx_idx.data();
  d->h_ht_idx = s->ht_idx.data();
  d->num_params = s->k;
  d->h_coeff_affine = s->k > 0 ? reinterpret_cast<const hcx_complex*>(s->affine.data()) : nullptr;
  return HCX_OK;
}

hcx_status hcx_total_degree_homotopy(const hcx_system* s, const hcx_complex* target_coeffs,
                                     hcx_complex gamma, hcx_system** out_sys, int* out_num_coeffs,
                                     hcx_complex* hx, hcx_complex* ht) {
  clear_error();
  if (!s) {
    set_error("hcx_total_degree_homotopy: null system");
    return HCX_ERR_INVALID_ARG;
  }
  const int n = s->n;
  // merged monomial list: F's monomials, then any missing x_i^{d_i} or 1
  std::vector<int> exps = s->exps, eqv = s->eq;
  std::vector<cd> f(s->C), g(s->C, cd(0, 0));
  for (int j = 0; j < s->C; ++j) f[j] = target_coeffs ? to_cd(target_coeffs[j]) : cd(0, 0);
  for (int i = 0; i < n; ++i) {
    const int d = s->degree_of_eq[i];
    for (int which = 0; which < 2; ++which) {
      std::vector<int> e(n, 0);
      if (which == 0) e[i] = d;
      const cd gv = which == 0 ? cd(1, 0) : cd(-1, 0);
      int found = -1;
      for (int j = 0; j < (int)eqv.size(); ++j) {
        if (eqv[j] != i) continue;
        bool same = true;
        for (int v = 0; v < n && same; ++v) same = exps[(size_t)j * n + v] == e[v];
        if (same) { found = j; break; }
      }
      if (found < 0) {
        exps.insert(exps.end(), e.begin(), e.end());
        eqv.push_back(i);
        f.push_back(cd(0, 0));
        g.push_back(gv);
      } else {
        g[found] = gv;
      }
    }
  }
  const int C2 = (int)eqv.size();
  if (out_num_coeffs) *out_num_coeffs = C2;
  if (out_sys) {
    hcx_status st = hcx_system_create(n, 0, C2, exps.data(), eqv.data(), nullptr, out_sys);
    if (st != HCX_OK) return st;
    // The merged system keeps coefficient values in its affine constant term.
    for (int j = 0; j < C2; ++j) (*out_sys)->affine[j] = f[j];
  }
  const cd gm = to_cd(gamma);
  if (hx && ht) {
    for (int j = 0; j < C2; ++j) {
      const cd a = gm * g[j], dlt = f[j] - gm * g[j];
      hx[2 * j] = to_hc(a);
      hx[2 * j + 1] = to_hc(dlt);
      ht[j] = to_hc(dlt);
    }
    hx[2 * C2] = hcx_complex{1, 0};
    hx[2 * C2 + 1] = hcx_complex{0, 0};
    ht[C2] = hcx_complex{0, 0};
  }
  return HCX_OK;
}

hcx_status hcx_total_degree_start_roots(const hcx_system* s, int count, const int64_t* root_index,
                                        hcx_complex* roots) {
  clear_error();
  if (!s || count < 0 || (count > 0 && (!root_index || !roots))) {
    set_error("hcx_total_degree_start_roots: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  const int n = s->n;
  for (int r = 0; r < count; ++r) {
    int64_t idx = root_index[r];
    for (int v = 0; v < n; ++v) {
      const int d = s->degree_of_eq[v];
      const int m = (int)(idx % d);
      idx /= d;
      // exact values for the square roots of unity, cos/sin otherwise
      cd z;
      if (d == 1) z = cd(1, 0);
      else if (d == 2) z = m == 0 ? cd(1, 0) : cd(-1, 0);
      else {
        const double th = 6.283185307179586476925286766559 * (double)m / (double)d;
        z = cd(std::cos(th), std::sin(th));
        if (m == 0) z = cd(1, 0);
      }
      roots[(size_t)r * n + v] = to_hc(z);
    }
  }
  return HCX_OK;
}

hcx_status hcx_total_degree_subset(const hcx_system* s, int count, uint64_t seed, int64_t* root_index) {
  clear_error();
  if (!s || count < 0 || (count > 0 && !root_index)) {
    set_error("hcx_total_degree_subset: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  // Bezout number, saturating at 2^62
  uint64_t total = 1;
  for (int v = 0; v < s->n; ++v) {
    const uint64_t d = (uint64_t)s->degree_of_eq[v];
    if (total > (1ull << 62) / d) { total = 1ull << 62; break; }
    total *= d;
  }
  if ((uint64_t)count >= total) {
    for (uint64_t r = 0; r < total; ++r) root_index[r] = (int64_t)r;
    for (int r = (int)total; r < count; ++r) root_index[r] = -1;
    return HCX_OK;
  }
  SplitMix64 rng(sub_seed(seed, 3));
  if ((uint64_t)count * 2 >= total) {
    // dense case: partial Fisher-Yates over the full index range
    std::vector<int64_t> idx(total);
    for (uint64_t q = 0; q < total; ++q) idx[q] = (int64_t)q;
    for (int q = 0; q < count; ++q) {
      const uint64_t r = q + rng.below(total - q);
      std::swap(idx[q], idx[r]);
    }
    std::copy(idx.begin(), idx.begin() + count, root_index);
    return HCX_OK;
  }
  std::unordered_set<int64_t> seen;
  int got = 0;
  while (got < count) {
    const int64_t r = (int64_t)rng.below(total);
    if (seen.insert(r).second) root_index[got++] = r;
  }
  return HCX_OK;
}

}  // extern "C"
