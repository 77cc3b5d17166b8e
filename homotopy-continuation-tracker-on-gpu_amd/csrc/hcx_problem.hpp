// Problem object and the shared launch entry of libhcx (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

#include "hcx.h"
#include "hcx_kernel.hpp"
#include "hcx_wspool.hpp"

// Device-resident problem type: repacked tables, start roots and the
// coefficient map.  Launch workspaces are not part of it: they belong to the
// stream a launch runs on (hcx_wspool.hpp), to a captured graph, to a server
// slot or to the caller's static scratch, so one problem serves any number of
// streams, graphs and servers at once.
struct hcx_problem {
  int device = 0;
  int n = 0, C = 0, M = 1, K = 0, k = 0;
  int hxT = 0, htT = 0, hx_vars = 0, ht_vars = 0, deg = 2;
  hcx::dev::Term* d_hx = nullptr;
  hcx::dev::Term* d_ht = nullptr;
  int* d_hx_cnt = nullptr;
  int* d_cslot = nullptr;              // [C+1] LDS slot of each coefficient
  int* d_cinv = nullptr;               // [C+1] coefficient of each slot
  uint32_t* d_ctab = nullptr;          // compact hx + ht tables (8-byte terms), or null
  int ctab_bytes = 0;                  // their size rounded up to 16 bytes
  uint32_t* d_ctab4 = nullptr;         // the same with 4-byte terms (degree <= 2), or null
  int ctab4_bytes = 0;
  double2* d_start = nullptr;
  double2* d_affine = nullptr;
  double2* d_p0 = nullptr;
  double2* d_cmap_scalar = nullptr;   // parameter-polynomial coefficient map [C][cmap_T]
  int* d_cmap_params = nullptr;       // [C][cmap_T][cmap_D]
  int cmap_T = 0, cmap_D = 0;
  // Counters and output descriptor of the reference-shaped synchronous
  // launcher (hcx_kernel_HC_Solver), allocated with the problem like the
  // reference's GPU_HC_Solver::Allocate_Arrays, so that launcher allocates
  // nothing; the mutex serialises its synchronous calls on this problem.
  hcx::WsBlock launcher_ws;
  std::mutex launcher_mu;
  // autotuned launch shape (hcx_autotune); used when the options leave it open
  bool tuned = false;
  int tuned_G = 0, tuned_spec = 1, tuned_wpc = 0;
  hcx_problem_info info{};
};

namespace hcx {
// Segment s's view of a workspace block whose instance counters start at
// inst_off and whose re-track list starts at path_off.
LaunchWork seg_work(const WsBlock& b, int s, long long inst_off, long long path_off);

hcx_status track_impl(int nseg, hcx_problem* const* probs, const hcx_batch_desc* batches,
                      const hcx_result_desc* results, const hcx_track_opts* o, hipStream_t stream, int persistent,
                      unsigned long long* batch_start, const LaunchWork* work);
}  // namespace hcx
