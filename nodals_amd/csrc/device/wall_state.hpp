// Device state of the wall-function momentum path of an H8 context (the
// k-omega SST path with Nitsche/Spalding walls; see wall_kernels.hpp).
#pragma once

#include "../host/potential_start.hpp"
#include "../host/wall_setup.hpp"
#include "common.hpp"

namespace nodals {

// Parses the wall-function options (reference names, lower-case flags):
//   --gamma G                 Nitsche normal penalty (SST22D4_GAMMA_N)
//   --wall-active-points N    Spalding fit points 1..5 (SST24B_ACTIVE_POINTS)
//   --wall-zero-limit 0|1     STAGE27_SPALDING_ZERO_LIMIT
//   --convection-form skew_symmetric|advective
//   --stress-form deviatoric|vector_laplacian
//   --m3-mode fgs|sgs  --m3-sweeps N  --m3-omega W
//   --initial-velocity potential|plug  --potential-tol T
// Unknown flags raise kErrInvalid.
WallParams parseWallArgs(int argc, const char* const* argv);

struct WallState {
  WallParams P;
  WallTopo W;
  // initial velocity [nv][3] in reference free order, and the potential
  // solve's report (iterations, rel, inlet flux, outlet flux, imbalance,
  // largest wall face-mean flux, largest vertex speed; zeros for the plug)
  std::vector<double> u0, startReport;
  // topology (device numbering)
  DBuf<int> wcell, wlf, bub, vert, dof, dofOff, dofFace, cellWallOff, cellWall, blkOff, blkIdx, bubFace, wallDof;
  DBuf<unsigned char> isBub;
  DBuf<double> wn, wh, qw, ql, bw, bp, heights;
  DBuf<int> rSrcFace, rSrcPos, rTgtPos, rTgtOff, rTgtSrc, lSrcFace, lSrcPos, lTgtPos, lTgtOff, lTgtSrc;
  int nRSrc = 0, nRTgt = 0, nLSrc = 0, nLTgt = 0;
  // work
  DBuf<double> fit, blk, pen, cpl, nuq, K, D, rau3, r3, delta3, diag3, dbuf, bn, y3, shear;
  DBuf<double> uPrev, kPrev, wPrev, momT;
  DBuf<int> counters;  // fit failures, non-finite viscosities, bad rows, failed block solves
};

}  // namespace nodals
