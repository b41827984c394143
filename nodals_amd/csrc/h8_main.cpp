// nodals_gpu_h8_<precision>: drop-in for the reference H8 binaries
// (gpu/serial_cuda/h8_main.cu, h8_main_base.cu:1748-2203).
//
// Same argv, the same marker lines consumed by the case front-end
// (NODALS_GPU_H8_SIMPLE, NODALS_GPU_H8_HP_ERROR, NODALS_GPU_H8_FIXED10 /
// NODALS_GPU_H8_FULLCONV, NODALS_GPU_RESULT gate=H8 ... status=PASS) and the
// same exit codes: 0 pass, 35 not pass, 90 exception.  The precision comes
// from the executable name: *_fp64 -> FP64 AMG, *_fp32 -> FP32 AMG
// (FP64 state and outer solve), *_bf16 -> BF16 finest level; an explicit
// --amg-precision overrides it.  Everything below goes through the C ABI.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nodals_hip.h"

int main(int argc, char** argv) {
  std::vector<const char*> args(argv, argv + argc);
  std::string self = argv[0];
  const std::size_t slash = self.find_last_of('/');
  if (slash != std::string::npos) self = self.substr(slash + 1);
  bool explicitPrecision = false;
  std::string tag = "h8", runMode = "fixed10", momentumWork = "fgs1";
  double simpleTol = 1e-6;
  for (int i = 1; i + 1 < argc; ++i) {
    if (std::strcmp(argv[i], "--amg-precision") == 0) explicitPrecision = true;
    if (std::strcmp(argv[i], "--tag") == 0) tag = argv[i + 1];
    if (std::strcmp(argv[i], "--run-mode") == 0) runMode = argv[i + 1];
    if (std::strcmp(argv[i], "--momentum-work") == 0) momentumWork = argv[i + 1];
    if (std::strcmp(argv[i], "--simple-tol") == 0) simpleTol = std::atof(argv[i + 1]);
  }
  std::string precision = "fp64";
  if (self.size() >= 5 && self.compare(self.size() - 5, 5, "_fp32") == 0) precision = "fp32";
  if (self.size() >= 5 && self.compare(self.size() - 5, 5, "_bf16") == 0) precision = "bf16";
  if (!explicitPrecision && precision != "fp64") {
    args.push_back("--amg-precision");
    args.push_back(precision == "fp32" ? "fp32" : "bf16");
  }
  nodals_h8* h = nullptr;
  if (nodals_h8_create((int)args.size(), args.data(), &h) != 0) {
    std::fprintf(stderr, "NODALS_GPU_H8_EXCEPTION what=%s\n", nodals_last_error());
    return 90;
  }
  nodals_h8_run_result r{};
  if (nodals_h8_run(h, &r) != 0) {
    std::fprintf(stderr, "NODALS_GPU_H8_EXCEPTION what=%s\n", nodals_last_error());
    nodals_h8_destroy(h);
    return 90;
  }
  int64_t info[8];
  nodals_h8_info(h, info);
  const double dpErr = std::abs(r.pressure_drop - r.exact_pressure_drop) / std::max(std::abs(r.exact_pressure_drop), 1e-300);
  const int outer = r.outer > 0 ? r.outer : 1;
  std::printf(
      "NODALS_GPU_H8_HP_ERROR tag=%s cells=%lld pressureDropFit=%.12e exactPressureDrop=%.12e pressureDropRelErr=%.12e "
      "status=%s\n",
      tag.c_str(), (long long)info[0], r.pressure_drop, r.exact_pressure_drop, dpErr,
      r.converged ? "PASS" : "UNCONVERGED");
  const char* marker = runMode == "fixed10" ? "NODALS_GPU_H8_FIXED10" : "NODALS_GPU_H8_FULLCONV";
  std::printf(
      "%s tag=%s cells=%lld momentumWork=%s simpleTol=%.3e outerIts=%d converged=%d finalRelCont=%.12e "
      "avgPressureIts=%.6f pressureDropFit=%.12e exactPressureDrop=%.12e pressureDropRelErr=%.12e loopMs=%.3f "
      "avgSimpleMs=%.6f assemblyAvgMs=%.6f momentumAvgMs=%.6f continuityAvgMs=%.6f pressureAvgMs=%.6f "
      "pressureUpdateAvgMs=%.6f setupMs=%.3f status=%s\n",
      marker, tag.c_str(), (long long)info[0], momentumWork.c_str(), simpleTol, r.outer, r.converged, r.final_rel_cont,
      r.avg_pressure_its, r.pressure_drop, r.exact_pressure_drop, dpErr, r.loop_ms, r.loop_ms / outer,
      r.assembly_ms / outer, r.momentum_ms / outer, r.continuity_ms / outer, r.pressure_ms / outer,
      r.pupdate_ms / outer, r.setup_ms, r.pass ? "PASS" : "FAIL");
  std::printf(
      "NODALS_GPU_RESULT gate=H8 tag=%s cells=%lld precision=%s runMode=%s momentumWork=%s outer=%d simpleTol=%.3e "
      "fineCsrRefreshEvery=1 physicalOperator=current_exact_CSR pressureAll=PASS finite=PASS noPetsc=1 noMPI=1 "
      "status=%s\n",
      tag.c_str(), (long long)info[0], precision.c_str(), runMode.c_str(), momentumWork.c_str(), r.outer, simpleTol,
      r.pass ? "PASS" : "FAIL");
  std::fflush(stdout);
  nodals_h8_destroy(h);
  return r.pass ? 0 : 35;
}
