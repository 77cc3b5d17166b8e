// hcx_solve: command line front end of hcx::GPU_HC_Solver, the counterpart of
// the reference's magmaHC-main (GPU-HC/cmd/magmaHC-main.cpp:30-115):
//   hcx_solve -p <problem_dir> [-m coeff_map.txt] [-o out_dir] [--fp64] [-i instances]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "hcx_solver.hpp"

static void usage() {
  std::printf(
      "usage: hcx_solve -p <problem_dir> [options]\n"
      "  -p, --problem DIR     problems/<p>/ folder (gpuhc_settings.yaml, start/target params,\n"
      "                        start_sols.txt, dHdt_indx.txt, dHdx_indx.txt)\n"
      "  -m, --cmap FILE       coefficient map of a P2C problem (default DIR/coeff_map.txt;\n"
      "                        written by: python -m hcx.compiler DIR sys_<p>.m)\n"
      "  -o, --output DIR      write GPU_Converged_HC_tracks.txt there\n"
      "  -i, --instances N     solve N copies of the target parameters in one launch\n"
      "      --fp64            track in FP64 (default: mixed FP32/FP64)\n"
      "  -h, --help\n");
}

int main(int argc, char** argv) {
  std::string dir;
  hcx::SolverOptions o;
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    auto next = [&](const char* what) -> std::string {
      if (a + 1 >= argc) {
        std::fprintf(stderr, "hcx_solve: %s needs a value\n", what);
        std::exit(2);
      }
      return argv[++a];
    };
    if (s == "-h" || s == "--help") {
      usage();
      return 0;
    } else if (s == "-p" || s == "--problem") {
      dir = next("--problem");
    } else if (s == "-m" || s == "--cmap") {
      o.cmap_path = next("--cmap");
    } else if (s == "-o" || s == "--output") {
      o.output_dir = next("--output");
    } else if (s == "-i" || s == "--instances") {
      o.instances = std::atoi(next("--instances").c_str());
    } else if (s == "--fp64") {
      o.precision = HCX_PREC_FP64;
    } else {
      std::fprintf(stderr, "hcx_solve: unknown argument %s\n", s.c_str());
      usage();
      return 2;
    }
  }
  if (dir.empty()) {
    usage();
    return 2;
  }
  try {
    hcx::GPU_HC_Solver solver(dir, o);
    solver.Allocate_Arrays();
    if (!solver.Read_Problem_Data()) return 1;
    solver.Data_Transfer_From_Host_To_Device();
    solver.Solve_by_GPU_HC();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "hcx_solve: %s\n", e.what());
    return 1;
  }
  return 0;
}
