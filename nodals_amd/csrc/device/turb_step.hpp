// Device state of the k-omega SST transport step of an H8 context.
//
// One step follows the reference's step_turbulence for the masked-P1 state
// space (turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/modules/
// coupling/g2_step_initialization.inc, g2_step_convection.inc, g2_step_k.inc,
// g2_step_omega.inc): Picard convection at the current velocity, then for k
// (with the step's starting omega) and then omega (with the new k): element
// volume terms, inlet flux, the omega wall penalty, combination, row-L1
// relaxation, a fixed number of coloured Gauss-Seidel sweeps on the vertex
// rows and the floor.  Face-bubble coefficients stay zero throughout.
#pragma once

#include <string>
#include <vector>

#include "../host/turb_setup.hpp"
#include "common.hpp"

namespace nodals {

struct TurbStepResult {
  int kClips = 0, omegaClips = 0;  // floored vertex values
  int wallInvalid = 0;             // wall points without a valid fit/target
  double ms = 0;                   // device time of the step
};

// Parses the reference's SST flags that this step honours
// (turbulence/sst_gpu_v2/nodals.py:444-509 passes them):
//   --sst-intensity --sst-length-ratio --sst-alpha-k --sst-alpha-omega
//   --sst-k-fgs-sweeps --sst-omega-fgs-sweeps --sst-gs-omega
//   --sst-k-floor --sst-omega-floor --sst-deviatoric-production 0|1
//   --sst-omega-wall-mode of_auto|none --sst-wall-penalty-gamma
//   --sst-profile-fit-y1 .. --sst-profile-fit-y5 --sst-wall-sample-scale
//   --sst-of-kappa --sst-of-E --sst-of-beta1 --sst-of-re-blend
//   --sst-of-production-limit
// Unknown flags raise kErrInvalid.
TurbParams parseTurbArgs(int argc, const char* const* argv);

struct TurbState {
  TurbTopo T;
  TurbParams P;
  DBuf<int> sref, row, col, diagPos, entOff, termOff, term, rhsOff, rhsTerm, crows;
  DBuf<int> fcell, flf, scell, counters, wcell, wlf;
  DBuf<double> fgeom, wgeom, sl, sphi, sy, coef, target, mu, conv, blk, brhs, val, diag, rhs, k, w;
  // Spalding points of the omega wall fit: the k-omega setup's active
  // points, replaced by the wall path's own count when it is initialised
  // (the reference holds one count for both, SST24B_ACTIVE_POINTS)
  int omegaFitPoints = 5;
  DBuf<double> vdist;  // vertex wall distances (duct mode), empty for the straight pipe
  DBuf<float> valF;    // FP32 copy of val for --sst-sweep-precision fp32
  DBuf<double> resid;  // residual of a scalar_solve norm
  const double* vdistPtr() const { return vdist.size() ? vdist.data() : nullptr; }
  std::vector<int> colorOff;
  GraphCache sweepGraph[2];  // k, omega
  hipEvent_t ev[2] = {nullptr, nullptr};
  int steps = 0;
  ~TurbState() {
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

}  // namespace nodals
