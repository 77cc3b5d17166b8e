// C++ host lifecycle (GPU_HC_Solver member functions) over the C-ABI.
// Follows the order of GPU-HC/magmaHC/GPU_HC_Solver.cpp: construction from
// the settings (:25-70), Allocate_Arrays (:104-140), Read_Problem_Data
// (:143-170), Data_Transfer_From_Host_To_Device (:172-205) and
// Solve_by_GPU_HC (:209-384, with Evaluations.cpp:66-176 for collection).
#include "hcx_solver.hpp"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <stdexcept>

namespace hcx {

namespace {

void check(hcx_status s, const char* what) {
  if (s != HCX_OK) throw std::runtime_error(std::string(what) + ": " + hcx_last_error());
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::vector<int> read_table(const std::string& path, size_t count) {
  std::vector<int> v(count);
  int got = 0;
  check(hcx_ref_read_indices(path.c_str(), v.data(), (int)count, &got), "read table");
  if ((size_t)got != count)
    throw std::runtime_error(path + ": expected " + std::to_string(count) + " indices, read " + std::to_string(got));
  return v;
}

}  // namespace

GPU_HC_Solver::GPU_HC_Solver(const std::string& problem_dir, const SolverOptions& opts)
    : dir_(problem_dir), opt_(opts) {
  check(hcx_ref_read_settings(dir_.c_str(), &st_), "read settings");
  if (opt_.instances < 1) throw std::runtime_error("instances must be >= 1");
  inst_ = opt_.instances;
  if (opt_.cmap_path.empty()) opt_.cmap_path = dir_ + "/coeff_map.txt";
}

GPU_HC_Solver::~GPU_HC_Solver() { release_device(); }

void GPU_HC_Solver::release_device() {
  if (prob_) hcx_problem_destroy(prob_);
  prob_ = nullptr;
  for (void** p : {&d_params_, &d_work_, &d_roots_, &d_flags_})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
  stream_ = nullptr;
  allocated_ = false;
}

void GPU_HC_Solver::Allocate_Arrays() {
  const int n = st_.num_vars, K = st_.num_tracks;
  h_roots_.assign((size_t)inst_ * K * n, hcx_complex{0, 0});
  h_flags_.assign((size_t)inst_ * K, 0u);
  h_start_sols_.assign((size_t)K * n, hcx_complex{0, 0});
  h_start_params_.assign(st_.num_params, hcx_complex{0, 0});
  allocated_ = true;
}

bool GPU_HC_Solver::Read_Problem_Data() {
  if (!allocated_) Allocate_Arrays();
  const int n = st_.num_vars, K = st_.num_tracks, k = st_.num_params;
  const std::string type(st_.gpuhc_type);
  try {
    int got = 0;
    check(hcx_ref_read_start_sols((dir_ + "/start_sols.txt").c_str(), n, h_start_sols_.data(), K, &got),
          "read start solutions");
    if (got != K) throw std::runtime_error("start_sols.txt holds " + std::to_string(got) + " of the " +
                                           std::to_string(K) + " tracks");
    check(hcx_ref_read_params((dir_ + "/start_params.txt").c_str(), h_start_params_.data(), k, &got),
          "read start parameters");
    if (got != k) throw std::runtime_error("start_params.txt holds " + std::to_string(got) + " of " + std::to_string(k));
    if (!have_targets_) {
      std::vector<hcx_complex> tp(k);
      check(hcx_ref_read_params((dir_ + "/target_params.txt").c_str(), tp.data(), k, &got), "read target parameters");
      if (got != k)
        throw std::runtime_error("target_params.txt holds " + std::to_string(got) + " of " + std::to_string(k));
      h_target_params_.clear();
      for (int i = 0; i < inst_; ++i) h_target_params_.insert(h_target_params_.end(), tp.begin(), tp.end());
    }
    if (type == "PH") {
      // transposed PH tables -> P2C tables with one coefficient per parameter product
      std::vector<int> pht = read_table(dir_ + "/dHdt_indx.txt", (size_t)st_.ht_terms * st_.ht_parts * n);
      std::vector<int> phx = read_table(dir_ + "/dHdx_indx.txt", (size_t)n * st_.hx_terms * st_.hx_parts * n);
      htP_ = st_.ht_parts - 1;
      hxP_ = st_.hx_parts - 1;
      h_ht_.assign((size_t)n * st_.ht_terms * htP_, 0);
      h_hx_.assign((size_t)n * n * st_.hx_terms * hxP_, 0);
      const int cap = (k + 1) * (k + 2) / 2;
      std::vector<int> pairs(2 * (size_t)cap);
      check(hcx_ref_ph_to_p2c(n, k, st_.ht_terms, st_.ht_parts, pht.data(), st_.hx_terms, st_.hx_parts, phx.data(),
                              &C_, h_ht_.data(), h_hx_.data(), pairs.data(), cap),
            "PH to P2C");
      M_ = 2;
      cmap_T_ = 1;
      cmap_D_ = 2;
      cmap_scalar_.assign(C_, hcx_complex{1, 0});
      cmap_params_.assign(pairs.begin(), pairs.begin() + 2 * (size_t)C_);
    } else if (type == "P2C") {
      C_ = st_.num_coeffs;
      M_ = st_.max_order_t;
      htP_ = st_.ht_parts;
      hxP_ = st_.hx_parts;
      h_ht_ = read_table(dir_ + "/dHdt_indx.txt", (size_t)n * st_.ht_terms * htP_);
      h_hx_ = read_table(dir_ + "/dHdx_indx.txt", (size_t)n * n * st_.hx_terms * hxP_);
      int c = 0, t = 0, d = 0, kk = 0;
      check(hcx_ref_read_cmap(opt_.cmap_path.c_str(), &c, &t, &d, &kk, nullptr, nullptr, 0),
            "read coefficient map (write one with: python -m hcx.compiler <problem_dir> <sys_file>)");
      if (c != C_ || kk != k)
        throw std::runtime_error("coefficient map sizes (C=" + std::to_string(c) + ", k=" + std::to_string(kk) +
                                 ") disagree with the settings");
      cmap_T_ = t;
      cmap_D_ = d;
      cmap_scalar_.resize((size_t)c * t);
      cmap_params_.resize((size_t)c * t * d);
      check(hcx_ref_read_cmap(opt_.cmap_path.c_str(), &c, &t, &d, &kk, cmap_scalar_.data(), cmap_params_.data(),
                              c * t),
            "read coefficient map");
    } else {
      throw std::runtime_error("unknown GPUHC_Type " + type);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[hcx] %s: %s\n", st_.problem_name, e.what());
    return false;
  }
  return true;
}

void GPU_HC_Solver::Set_Target_Params(const std::vector<hcx_complex>& params, int instances) {
  if (instances < 1 || params.size() != (size_t)instances * st_.num_params)
    throw std::runtime_error("Set_Target_Params: need instances * num_params values");
  inst_ = instances;
  h_target_params_ = params;
  have_targets_ = true;
  if (allocated_) Allocate_Arrays();
}

void GPU_HC_Solver::Data_Transfer_From_Host_To_Device() {
  const double t0 = now_s();
  const int n = st_.num_vars, K = st_.num_tracks, k = st_.num_params;
  if (prob_) hcx_problem_destroy(prob_);
  prob_ = nullptr;
  hcx_problem_desc d{};
  d.num_vars = n;
  d.num_coeffs = C_;
  d.max_order_t = M_;
  d.hx_terms = st_.hx_terms;
  d.hx_parts = hxP_;
  d.ht_terms = st_.ht_terms;
  d.ht_parts = htP_;
  d.h_hx_idx = h_hx_.data();
  d.h_ht_idx = h_ht_.data();
  d.num_tracks = K;
  d.h_start_sols = h_start_sols_.data();
  d.num_params = k;
  d.h_start_params = h_start_params_.data();
  d.cmap_terms = cmap_T_;
  d.cmap_deg = cmap_D_;
  d.h_cmap_scalar = cmap_scalar_.data();
  d.h_cmap_params = cmap_params_.data();
  check(hcx_problem_create(&d, &prob_), "problem create");
  if (!stream_) check_hip(hipStreamCreateWithFlags((hipStream_t*)&stream_, hipStreamNonBlocking), "stream");
  for (void** p : {&d_params_, &d_work_, &d_roots_, &d_flags_})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  const size_t P = (size_t)inst_ * K;
  check_hip(hipMalloc(&d_params_, sizeof(hcx_complex) * (size_t)inst_ * k), "hipMalloc params");
  check_hip(hipMalloc(&d_work_, sizeof(hcx_complex) * (size_t)inst_ * (C_ + 1) * (2 * M_ + 1)), "hipMalloc work");
  check_hip(hipMalloc(&d_roots_, sizeof(hcx_complex) * P * n), "hipMalloc roots");
  check_hip(hipMalloc(&d_flags_, sizeof(uint32_t) * P), "hipMalloc flags");
  check_hip(hipMemcpy(d_params_, h_target_params_.data(), sizeof(hcx_complex) * (size_t)inst_ * k,
                      hipMemcpyHostToDevice),
            "copy target params");
  transfer_h2d_time = now_s() - t0;
}

void GPU_HC_Solver::Solve_by_GPU_HC() {
  if (!prob_) throw std::runtime_error("Solve_by_GPU_HC: call Data_Transfer_From_Host_To_Device first");
  const int n = st_.num_vars, K = st_.num_tracks;
  hcx_track_opts o;
  hcx_track_opts_default(&o);
  o.max_steps = st_.max_steps;
  o.max_corr = st_.max_corr;
  o.incr_steps = st_.incr_steps;
  o.precision = opt_.precision;
  hcx_batch_desc b{};
  b.num_instances = inst_;
  b.d_target_params = (const hcx_complex*)d_params_;
  b.d_coeff_work = (hcx_complex*)d_work_;
  hcx_result_desc r{};
  r.d_roots = (hcx_complex*)d_roots_;
  r.d_flags = (uint32_t*)d_flags_;
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "event");
  check_hip(hipEventCreate(&e1), "event");
  check_hip(hipEventRecord(e0, (hipStream_t)stream_), "event record");
  check(hcx_track(prob_, &b, &o, &r, stream_), "track");
  check_hip(hipEventRecord(e1, (hipStream_t)stream_), "event record");
  check(hcx_sync(stream_), "sync");
  float ms = 0;
  check_hip(hipEventElapsedTime(&ms, e0, e1), "event time");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  gpu_time = ms * 1e-3;
  const double t0 = now_s();
  const size_t P = (size_t)inst_ * K;
  h_roots_.resize(P * n);
  h_flags_.resize(P);
  check_hip(hipMemcpy(h_roots_.data(), d_roots_, sizeof(hcx_complex) * P * n, hipMemcpyDeviceToHost), "copy roots");
  check_hip(hipMemcpy(h_flags_.data(), d_flags_, sizeof(uint32_t) * P, hipMemcpyDeviceToHost), "copy flags");
  transfer_d2h_time = now_s() - t0;
  hcx_counts c{};
  check(hcx_collect_host(n, inst_, K, h_roots_.data(), h_flags_.data(), 1e-4, &c), "collect");
  Num_Of_Coverged_Sols = (unsigned)c.tracked;
  Num_Of_Inf_Sols = (unsigned)c.infinity;
  Num_Of_Real_Sols = (unsigned)c.real_ref;
  Num_Of_Unique_Sols = (unsigned)c.unique_ref;
  Num_Of_Refined_Sols = (unsigned)c.refined;
  Num_Of_Refined_Real_Sols = (unsigned)c.real;
  Num_Of_Distinct_Sols = (unsigned)c.unique;
  if (!opt_.output_dir.empty()) {
    const std::string path = opt_.output_dir + "/GPU_Converged_HC_tracks.txt";
    check(hcx_ref_write_converged(path.c_str(), n, inst_, K, h_roots_.data(), h_flags_.data(), HCX_FLAG_REFINED),
          "write converged solutions");
  }
  if (opt_.verbose) {
    std::printf("## Solving %s (%s, %s)\n", st_.problem_name, st_.gpuhc_type,
                opt_.precision == HCX_PREC_FP64 ? "FP64" : "mixed FP32/FP64");
    std::printf("- instances x tracks:           %d x %d\n", inst_, K);
    std::printf("- GPU computation time:         %.3f ms\n", gpu_time * 1e3);
    std::printf("- transfer host->device:        %.3f ms, device->host: %.3f ms\n", transfer_h2d_time * 1e3,
                transfer_d2h_time * 1e3);
    std::printf("- converged (t = 1):            %u\n", Num_Of_Coverged_Sols);
    std::printf("- refined in FP64:              %u\n", Num_Of_Refined_Sols);
    std::printf("- infinity failures:            %u\n", Num_Of_Inf_Sols);
    std::printf("- real solutions:               %u\n", Num_Of_Real_Sols);
    std::printf("- unique solutions:             %u\n", Num_Of_Unique_Sols);
    std::printf("- real refined roots:           %u\n", Num_Of_Refined_Real_Sols);
    std::printf("- distinct refined roots:       %u\n", Num_Of_Distinct_Sols);
  }
}

}  // namespace hcx
