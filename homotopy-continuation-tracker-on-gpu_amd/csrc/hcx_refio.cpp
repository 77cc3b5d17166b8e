// Reference on-disk problem format and host-side solution collection.
//
// Readers follow GPU-HC/magmaHC/Data_Reader.cpp:91-225 (whitespace separated
// "re im" pairs and integer tables read with >>) and the flat YAML keys parsed
// in GPU-HC/magmaHC/GPU_HC_Solver.cpp:44-68.  Values are parsed in FP64 (the
// reference parses them as float).  Collection follows
// GPU-HC/magmaHC/Evaluations.cpp:66-176.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hcx_internal.hpp"

using namespace hcx;

namespace {

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// Flat "key: value" YAML (the only form the reference settings use); '#'
// starts a comment, "%YAML" directives are skipped.
bool read_flat_yaml(const std::string& path, std::vector<std::pair<std::string, std::string>>& kv) {
  std::ifstream f(path);
  if (!f) return false;
  std::string line;
  while (std::getline(f, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    line = trim(line);
    if (line.empty() || line[0] == '%') continue;
    size_t c = line.find(':');
    if (c == std::string::npos) continue;
    kv.emplace_back(trim(line.substr(0, c)), trim(line.substr(c + 1)));
  }
  return true;
}

}  // namespace

extern "C" {

hcx_status hcx_ref_read_settings(const char* problem_dir, hcx_ref_settings* s) {
  clear_error();
  if (!problem_dir || !s) {
    set_error("hcx_ref_read_settings: null argument");
    return HCX_ERR_INVALID_ARG;
  }
  std::memset(s, 0, sizeof(*s));
  std::vector<std::pair<std::string, std::string>> kv;
  const std::string path = std::string(problem_dir) + "/gpuhc_settings.yaml";
  if (!read_flat_yaml(path, kv)) {
    set_error("hcx_ref_read_settings: cannot open %s", path.c_str());
    return HCX_ERR_IO;
  }
  auto geti = [&](const char* key, int* dst, bool required) -> bool {
    for (auto& p : kv)
      if (p.first == key) {
        *dst = std::atoi(p.second.c_str());
        return true;
      }
    if (required) set_error("hcx_ref_read_settings: key %s missing in %s", key, path.c_str());
    return !required;
  };
  for (auto& p : kv) {
    if (p.first == "problem_name") std::snprintf(s->problem_name, sizeof(s->problem_name), "%s", p.second.c_str());
    if (p.first == "GPUHC_Type") std::snprintf(s->gpuhc_type, sizeof(s->gpuhc_type), "%s", p.second.c_str());
  }
  bool ok = geti("GPUHC_Max_Steps", &s->max_steps, true) &&
            geti("GPUHC_Max_Correction_Steps", &s->max_corr, true) &&
            geti("GPUHC_Num_Of_Steps_to_Increase_Delta_t", &s->incr_steps, true) &&
            geti("Num_Of_Vars", &s->num_vars, true) && geti("Num_Of_Params", &s->num_params, true) &&
            geti("Num_Of_Tracks", &s->num_tracks, true) && geti("dHdx_Max_Terms", &s->hx_terms, true) &&
            geti("dHdx_Max_Parts", &s->hx_parts, true) && geti("dHdt_Max_Terms", &s->ht_terms, true) &&
            geti("dHdt_Max_Parts", &s->ht_parts, true);
  if (!ok) return HCX_ERR_IO;
  geti("Max_Order_Of_T", &s->max_order_t, false);
  geti("Num_Of_Coeffs_From_Params", &s->num_coeffs, false);
  return HCX_OK;
}

hcx_status hcx_ref_read_params(const char* path, hcx_complex* out, int cap, int* count) {
  clear_error();
  if (!path || !count || cap < 0 || (cap > 0 && !out)) {
    set_error("hcx_ref_read_params: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  std::ifstream f(path);
  if (!f) {
    set_error("hcx_ref_read_params: cannot open %s", path);
    return HCX_ERR_IO;
  }
  double re, im;
  int d = 0;
  while (f >> re >> im) {
    if (d >= cap) {
      set_error("hcx_ref_read_params: %s holds more than %d values", path, cap);
      return HCX_ERR_INVALID_ARG;
    }
    out[d++] = hcx_complex{re, im};
  }
  *count = d;
  return HCX_OK;
}

hcx_status hcx_ref_read_start_sols(const char* path, int n, hcx_complex* out, int cap_tracks, int* tracks) {
  clear_error();
  if (!path || !tracks || n < 1 || cap_tracks < 0 || (cap_tracks > 0 && !out)) {
    set_error("hcx_ref_read_start_sols: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  std::ifstream f(path);
  if (!f) {
    set_error("hcx_ref_read_start_sols: cannot open %s", path);
    return HCX_ERR_IO;
  }
  double re, im;
  long long d = 0;
  while (f >> re >> im) {
    if (d >= (long long)cap_tracks * n) {
      set_error("hcx_ref_read_start_sols: %s holds more than %d tracks", path, cap_tracks);
      return HCX_ERR_INVALID_ARG;
    }
    out[d++] = hcx_complex{re, im};
  }
  if (d % n != 0) {
    set_error("hcx_ref_read_start_sols: %lld values is not a multiple of n=%d", d, n);
    return HCX_ERR_IO;
  }
  *tracks = (int)(d / n);
  return HCX_OK;
}

hcx_status hcx_ref_read_indices(const char* path, int* out, int cap, int* count) {
  clear_error();
  if (!path || !count || cap < 0 || (cap > 0 && !out)) {
    set_error("hcx_ref_read_indices: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  std::ifstream f(path);
  if (!f) {
    set_error("hcx_ref_read_indices: cannot open %s", path);
    return HCX_ERR_IO;
  }
  long long v;
  int d = 0;
  while (f >> v) {
    if (d >= cap) {
      set_error("hcx_ref_read_indices: %s holds more than %d indices", path, cap);
      return HCX_ERR_INVALID_ARG;
    }
    out[d++] = (int)v;
  }
  *count = d;
  return HCX_OK;
}

hcx_status hcx_ref_read_cmap(const char* path, int* C, int* T, int* D, int* k, hcx_complex* scalar, int* params,
                             int cap_terms) {
  clear_error();
  if (!path || !C || !T || !D || !k || (scalar && !params)) {
    set_error("hcx_ref_read_cmap: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  std::ifstream f(path);
  if (!f) {
    set_error("hcx_ref_read_cmap: cannot open %s", path);
    return HCX_ERR_IO;
  }
  int c = 0, t = 0, d = 0, kk = 0;
  if (!(f >> c >> t >> d >> kk) || c < 1 || t < 1 || d < 1 || kk < 1) {
    set_error("hcx_ref_read_cmap: %s: bad header (need \"C T D k\")", path);
    return HCX_ERR_IO;
  }
  *C = c;
  *T = t;
  *D = d;
  *k = kk;
  if (!scalar) return HCX_OK;
  const long long terms = (long long)c * t;
  if (terms > cap_terms) {
    set_error("hcx_ref_read_cmap: %lld terms exceed the buffer (%d)", terms, cap_terms);
    return HCX_ERR_INVALID_ARG;
  }
  for (long long u = 0; u < terms; ++u) {
    double re, im;
    if (!(f >> re >> im)) {
      set_error("hcx_ref_read_cmap: %s: term %lld missing", path, u);
      return HCX_ERR_IO;
    }
    scalar[u] = hcx_complex{re, im};
    for (int q = 0; q < d; ++q) {
      int a;
      if (!(f >> a) || a < 0 || a > kk) {
        set_error("hcx_ref_read_cmap: %s: term %lld has a bad parameter index", path, u);
        return HCX_ERR_IO;
      }
      params[u * d + q] = a;
    }
  }
  return HCX_OK;
}

hcx_status hcx_ref_ph_to_p2c(int n, int P, int ht_terms, int ht_parts, const int* ph_ht, int hx_terms, int hx_parts,
                             const int* ph_hx, int* C, int* ht, int* hx, int* pairs, int cap_pairs) {
  clear_error();
  if (n < 1 || P < 1 || ht_terms < 1 || hx_terms < 1 || ht_parts < 4 || hx_parts < 4 || !ph_ht || !ph_hx || !C ||
      !ht || !hx || (cap_pairs > 0 && !pairs)) {
    set_error("hcx_ref_ph_to_p2c: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  std::vector<std::pair<int, int>> found;
  auto pid = [&](int a, int b) -> int {
    if (a > b) std::swap(a, b);
    if (a == P && b == P) return -1;
    for (size_t q = 0; q < found.size(); ++q)
      if (found[q].first == a && found[q].second == b) return (int)q;
    found.emplace_back(a, b);
    return (int)found.size() - 1;
  };
  // first pass assigns the numbering; the tables are written with -1 for the
  // constant and fixed up once C is known
  const int htp = ht_parts - 1, hxp = hx_parts - 1;
  for (int i = 0; i < n; ++i)
    for (int u = 0; u < ht_terms; ++u) {
      int* o = ht + ((size_t)i * ht_terms + u) * htp;
      const int s = ph_ht[((size_t)u * ht_parts + 0) * n + i];
      const int a = ph_ht[((size_t)u * ht_parts + 1) * n + i], b = ph_ht[((size_t)u * ht_parts + 2) * n + i];
      if (a < 0 || a > P || b < 0 || b > P) {
        set_error("hcx_ref_ph_to_p2c: dH/dt parameter index out of range");
        return HCX_ERR_INVALID_ARG;
      }
      o[0] = s;
      o[1] = s != 0 ? pid(a, b) : -1;
      for (int q = 3; q < ht_parts; ++q) o[q - 1] = ph_ht[((size_t)u * ht_parts + q) * n + i];
    }
  for (int j = 0; j < n; ++j)
    for (int u = 0; u < hx_terms; ++u)
      for (int i = 0; i < n; ++i) {
        int* o = hx + (((size_t)i * n + j) * hx_terms + u) * hxp;
        const size_t base = (((size_t)j * hx_terms + u) * hx_parts) * n + i;
        const int s = ph_hx[base];
        const int a = ph_hx[base + (size_t)1 * n], b = ph_hx[base + (size_t)2 * n];
        if (a < 0 || a > P || b < 0 || b > P) {
          set_error("hcx_ref_ph_to_p2c: dH/dx parameter index out of range");
          return HCX_ERR_INVALID_ARG;
        }
        o[0] = s;
        o[1] = s != 0 ? pid(a, b) : -1;
        for (int q = 3; q < hx_parts; ++q) o[q - 1] = ph_hx[base + (size_t)q * n];
      }
  const int c = (int)found.size();
  if (c > cap_pairs) {
    set_error("hcx_ref_ph_to_p2c: %d parameter products exceed the pair buffer (%d)", c, cap_pairs);
    return HCX_ERR_INVALID_ARG;
  }
  for (size_t q = 0; q < (size_t)n * ht_terms; ++q)
    if (ht[q * htp + 1] < 0) ht[q * htp + 1] = c;
  for (size_t q = 0; q < (size_t)n * n * hx_terms; ++q)
    if (hx[q * hxp + 1] < 0) hx[q * hxp + 1] = c;
  for (int q = 0; q < c; ++q) {
    pairs[2 * q] = found[q].first;
    pairs[2 * q + 1] = found[q].second;
  }
  *C = c;
  return HCX_OK;
}

hcx_status hcx_ref_write_converged(const char* path, int n, int num_instances, int num_tracks,
                                   const hcx_complex* h_roots, const uint32_t* h_flags, uint32_t mask) {
  clear_error();
  if (!path || !h_roots || !h_flags || n < 1 || num_instances < 0 || num_tracks < 0) {
    set_error("hcx_ref_write_converged: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    set_error("hcx_ref_write_converged: cannot open %s", path);
    return HCX_ERR_IO;
  }
  for (int inst = 0; inst < num_instances; ++inst) {
    std::fprintf(f, "-------------------- RANSAC Iteration %d --------------------\n\n", inst + 1);
    for (int r = 0; r < num_tracks; ++r) {
      const size_t p = (size_t)inst * num_tracks + r;
      if ((h_flags[p] & mask) != mask) continue;
      for (int v = 0; v < n; ++v)
        std::fprintf(f, "%.20g\t%.20g\n", h_roots[p * n + v].re, h_roots[p * n + v].im);
      std::fprintf(f, "\n");
    }
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  return HCX_OK;
}

// Root-level collection (the host twin of hcx_collect).  "Unique" counts the
// first occurrence of each refined root of an instance: path j is a duplicate
// when a refined path i < j of the same instance agrees in every coordinate
// (|dRe|, |dIm| < tol).  "unique_ref" and "real_ref" are the reference's own
// counts (Evaluations.cpp:88-110, 127-176): see unique_ref_pass.
namespace {

bool same_root(const hcx_complex* a, const hcx_complex* b, int n, double tol) {
  for (int v = 0; v < n; ++v)
    if (!(std::fabs(a[v].re - b[v].re) < tol && std::fabs(a[v].im - b[v].im) < tol)) return false;
  return true;
}

// Find_Unique_Sols (Evaluations.cpp:127-176) over one instance.  Path bs is
// examined when it is tracked ("converged") and not in the Skip set; its
// Duplicate set is every LATER path ds (tracked or not) with the same root;
// an empty set counts bs as unique, a non-empty one REPLACES the Skip set.
// The Skip set is therefore always the Duplicate set of the last examined
// path that had one (bstar), and "bs in Skip" is same_root(bs, bstar).
int64_t unique_ref_pass(int n, int K, const hcx_complex* x, uint32_t* f, double tol) {
  int64_t cnt = 0;
  int bstar = -1;
  for (int bs = 0; bs < K; ++bs) {
    if (!(f[bs] & HCX_FLAG_TRACKED)) continue;
    if (bstar >= 0 && same_root(x + (size_t)bs * n, x + (size_t)bstar * n, n, tol)) continue;
    bool later = false;
    for (int ds = bs + 1; ds < K && !later; ++ds) later = same_root(x + (size_t)bs * n, x + (size_t)ds * n, n, tol);
    if (later) {
      bstar = bs;
    } else {
      f[bs] |= HCX_FLAG_UNIQUE_REF;
      ++cnt;
    }
  }
  return cnt;
}

}  // namespace

hcx_status hcx_collect_host(int n, int num_instances, int num_tracks, const hcx_complex* h_roots,
                            uint32_t* h_flags, double unique_tol, hcx_counts* out) {
  clear_error();
  if (n < 1 || num_instances < 0 || num_tracks < 0 || !h_roots || !h_flags || !out) {
    set_error("hcx_collect_host: invalid arguments");
    return HCX_ERR_INVALID_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  for (int inst = 0; inst < num_instances; ++inst) {
    const size_t base = (size_t)inst * num_tracks;
    for (int j = 0; j < num_tracks; ++j) {
      uint32_t& fj = h_flags[base + j];
      fj &= ~(HCX_FLAG_UNIQUE | HCX_FLAG_UNIQUE_REF);
      if (fj & HCX_FLAG_TRACKED) out->tracked++;
      if (fj & HCX_FLAG_INFINITY) out->infinity++;
      if ((fj & HCX_FLAG_TRACKED) && (fj & HCX_FLAG_REAL_REF)) out->real_ref++;
      if (!(fj & HCX_FLAG_REFINED)) continue;
      out->refined++;
      if (fj & HCX_FLAG_REAL) out->real++;
      const hcx_complex* xj = h_roots + (base + j) * n;
      bool dup = false;
      for (int i = 0; i < j && !dup; ++i) {
        if (!(h_flags[base + i] & HCX_FLAG_REFINED)) continue;
        dup = same_root(h_roots + (base + i) * n, xj, n, unique_tol);
      }
      if (!dup) {
        fj |= HCX_FLAG_UNIQUE;
        out->unique++;
      }
    }
    out->unique_ref += unique_ref_pass(n, num_tracks, h_roots + base * n, h_flags + base, unique_tol);
  }
  return HCX_OK;
}

}  // extern "C"
