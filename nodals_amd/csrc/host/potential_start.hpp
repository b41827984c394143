// Potential-flow start of the wall-function path: the P1 Laplace problem
//   int grad(phi) . grad(v) = int_inlet v (U_in . n),  phi = 0 on the outlet
// vertices, Neumann walls, solved by Jacobi-preconditioned CG; the cellwise
// grad(phi) projected onto the vertices with the lumped mass; and one normal
// BF3 bubble value per boundary face that sets the face's mean normal
// velocity (U_in . n on the inlet, zero on the wall, the projected value
// plus one uniform flux-balancing correction on the outlet).
//
// Reference: turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
// modules/initialization/stage27_potential_host.hpp:44-167 (selected by the
// case's [initialization] mode = potential_flow, nodals.py:147-149, 328).
#pragma once

#include <vector>

#include "h8_setup.hpp"
#include "mesh.hpp"

namespace nodals {

struct PotentialStart {
  std::vector<double> value;  // [nPoints + nFaces][3]: vertices, then face bubbles
  int iterations = 0;
  double rel = 0;             // final ||r|| / ||b|| of the Laplace solve
  double inletFlux = 0, outletFlux = 0, imbalanceRel = 0, wallFluxMax = 0, maxSpeed = 0;
};

// Patches: S.pipe.inlet and S.pipe.outlet by index, every other boundary
// patch a wall; U_in = S.pipe.inletVelocity.  Throws when the solve does not
// reach `tol` in `maxIterations` or the flux audit fails.
PotentialStart potentialStart(const TetMesh& M, const H8Setup& S, double tol = 1e-10, int maxIterations = 12000);

}  // namespace nodals
