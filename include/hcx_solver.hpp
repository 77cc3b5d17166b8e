// C++ host lifecycle over the C-ABI of hcx.h, with the member functions of the
// reference's GPU_HC_Solver (GPU-HC/magmaHC/GPU_HC_Solver.hpp:30-140,
// driven by GPU-HC/cmd/magmaHC-main.cpp:86-115):
//
//   hcx::GPU_HC_Solver s(problem_dir, opts);   // reads gpuhc_settings.yaml
//   s.Allocate_Arrays();
//   s.Read_Problem_Data();                     // tables, params, start sols, coefficient map
//   s.Data_Transfer_From_Host_To_Device();
//   s.Solve_by_GPU_HC();                       // track + collect (+ write converged roots)
//
// Differences from the reference: problem sizes are data (no per-problem
// kernel), arithmetic is FP64 or mixed FP32/FP64 with FP64 refinement, the
// coefficients of a P2C problem come from a coefficient map file written by
// hcx/compiler.py (in place of the generated p2c-<p>.h), several target
// parameter sets can be solved in one launch, and errors are thrown as
// std::runtime_error carrying hcx_last_error().
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "hcx.h"

namespace hcx {

struct SolverOptions {
  std::string cmap_path;     // P2C coefficient map (default <problem_dir>/coeff_map.txt)
  std::string output_dir;    // where GPU_Converged_HC_tracks.txt goes ("" = not written)
  int precision = HCX_PREC_MIXED;
  int instances = 1;         // target parameter sets per launch (replicas of target_params.txt
                             // unless Set_Target_Params gives them)
  bool verbose = true;       // print the summary the reference prints
};

class GPU_HC_Solver {
 public:
  GPU_HC_Solver(const std::string& problem_dir, const SolverOptions& opts = SolverOptions());
  ~GPU_HC_Solver();
  GPU_HC_Solver(const GPU_HC_Solver&) = delete;
  GPU_HC_Solver& operator=(const GPU_HC_Solver&) = delete;

  void Allocate_Arrays();
  bool Read_Problem_Data();
  // Optional: I target parameter sets [I][k] instead of target_params.txt.
  void Set_Target_Params(const std::vector<hcx_complex>& params, int instances);
  void Data_Transfer_From_Host_To_Device();
  void Solve_by_GPU_HC();

  // collection counters (Evaluations.hpp:47-50), over all instances, with
  // the reference's semantics (Evaluations.cpp:88-176): real = converged with
  // every |Im x_i| <= 1e-4 on the tracked root; unique = Find_Unique_Sols
  unsigned Num_Of_Coverged_Sols = 0;
  unsigned Num_Of_Inf_Sols = 0;
  unsigned Num_Of_Real_Sols = 0;
  unsigned Num_Of_Unique_Sols = 0;
  // root-level counts of the FP64-refined roots
  unsigned Num_Of_Refined_Sols = 0;
  unsigned Num_Of_Refined_Real_Sols = 0;    // |Im x_i| <= 1e-8 on the refined root
  unsigned Num_Of_Distinct_Sols = 0;        // first occurrence of each refined root

  // timers (seconds)
  double gpu_time = 0, transfer_h2d_time = 0, transfer_d2h_time = 0;

  const hcx_ref_settings& settings() const { return st_; }
  int num_vars() const { return st_.num_vars; }
  int num_tracks() const { return st_.num_tracks; }
  int num_instances() const { return inst_; }
  // [instances][tracks][n] refined roots and [instances][tracks] HCX_FLAG_* words
  const std::vector<hcx_complex>& solutions() const { return h_roots_; }
  const std::vector<uint32_t>& flags() const { return h_flags_; }

 private:
  std::string dir_;
  SolverOptions opt_;
  hcx_ref_settings st_{};
  int inst_ = 1;
  int C_ = 0, M_ = 1, hxP_ = 0, htP_ = 0;
  bool allocated_ = false, have_targets_ = false;
  std::vector<int> h_hx_, h_ht_;
  std::vector<hcx_complex> h_start_sols_, h_start_params_, h_target_params_;
  std::vector<hcx_complex> cmap_scalar_;
  std::vector<int> cmap_params_;
  int cmap_T_ = 0, cmap_D_ = 0;
  std::vector<hcx_complex> h_roots_;
  std::vector<uint32_t> h_flags_;
  hcx_problem* prob_ = nullptr;
  void* stream_ = nullptr;
  void* d_params_ = nullptr;
  void* d_work_ = nullptr;
  void* d_roots_ = nullptr;
  void* d_flags_ = nullptr;
  void release_device();
};

}  // namespace hcx
