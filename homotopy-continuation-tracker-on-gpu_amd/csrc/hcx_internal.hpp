// Internal declarations shared by the host C++ and HIP translation units of libhcx.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "hcx.h"

namespace hcx {

using cd = std::complex<double>;

// Thread-local error message behind hcx_last_error().
void set_error(const char* fmt, ...);
void clear_error();

inline cd to_cd(hcx_complex z) { return cd(z.re, z.im); }
inline hcx_complex to_hc(cd z) { return hcx_complex{z.real(), z.imag()}; }

// ---------------------------------------------------------------- random
// SplitMix64 (Steele, Lea, Flood 2014) and Box-Muller (1958) normal draws.
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in (0, 1]: 53 random mantissa bits, never 0 (safe for log).
  double uniform_open0() { return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }
  // uniform in [0, 1)
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [0, bound) by rejection (no modulo bias)
  uint64_t below(uint64_t bound) {
    const uint64_t lim = UINT64_MAX - (UINT64_MAX % bound);
    uint64_t v;
    do { v = next(); } while (v >= lim);
    return v % bound;
  }
  // CN(0,1): (a + i b)/sqrt(2) with a, b ~ N(0,1) from one Box-Muller pair
  cd cn() {
    const double u1 = uniform_open0(), u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    return cd(r * std::cos(th), r * std::sin(th)) * 0.70710678118654752440084436210485;
  }
};

// Independent sub-stream seed for purpose `tag` of a user seed.
uint64_t sub_seed(uint64_t seed, uint64_t tag);

// ---------------------------------------------------------------- systems
// A square polynomial system in n unknowns with C monomial coefficients.
// Monomial j lies in equation eq[j] with exponent row exps[j*n .. j*n+n-1]
// and coefficient c_j(p) = affine[j][0] + sum_k affine[j][1+k] p_k.
struct System {
  int n = 0, k = 0, C = 0;
  std::vector<int> exps;        // [C][n]
  std::vector<int> eq;          // [C]
  std::vector<cd> affine;       // [C][1+k]
  // compiled reference-format tables
  int hx_terms = 0, hx_parts = 0, ht_terms = 0, ht_parts = 0;
  int max_degree = 0, max_terms = 0;
  std::vector<int> hx_idx;      // [n][n][hx_terms][hx_parts]
  std::vector<int> ht_idx;      // [n][ht_terms][ht_parts]
  std::vector<int> degree_of_eq;// [n] total degree of each equation
  void compile();               // builds the tables from exps/eq
};

}  // namespace hcx

struct hcx_system : public hcx::System {};
