// Host setup of the wall-function momentum closure of the k-omega SST path:
// the wall faces with their twelve-point quadrature caches (wall trace and
// normal-column Picard functional of the eight P1+BF3 basis functions), the
// free DOFs of every wall face's bubble and vertices, and the index lists
// with which the device assembles the Nitsche/Spalding wall blocks and
// applies the exact BF3 wall-flux constraint u = T x to the momentum
// operator (K = T^T M T), to right-hand sides and to the pressure Schur
// operator, all without atomics.
//
// Reference: turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
// modules/wall/tet_wall_types.inc:27-61 (quadrature caches),
// stage22d4i_constraint.inc (T, Q, T^T), stage22d4d_coupled_matrixfree.inc:
// 734-813 (M3 right/left transforms), stage22d4b_nitsche_live.inc (face
// blocks) and coupling/stage22d4i3a_schur.cuh (constrained Schur).
#pragma once

#include <cstdint>
#include <vector>

#include "h8_setup.hpp"
#include "mesh.hpp"
#include "turb_setup.hpp"

namespace nodals {

// Options of the wall-function momentum path, named after the reference's
// controls (turbulence/sst_gpu_v2/nodals.py:330-352).
struct WallParams {
  double gammaN = 1e-6;     // Nitsche normal penalty (SST22D4_GAMMA_N)
  int activePoints = 1;     // Spalding fit points, 1..5 (SST24B_ACTIVE_POINTS)
  bool zeroLimit = true;    // molecular limit of a u_tau = 0 fit (STAGE27_SPALDING_ZERO_LIMIT)
  bool skew = true;         // skew-symmetric convection (NODALS_MOMENTUM_CONVECTION_FORM)
  bool deviatoric = true;   // implicit deviatoric stress (NODALS_MOMENTUM_STRESS_FORM)
  bool m3Symmetric = false; // forward (fgs) or symmetric (sgs) block sweeps (NODALS_M3_MODE)
  int m3Sweeps = 20;        // NODALS_M3_SWEEPS
  double m3Omega = 0.7;     // NODALS_M3_OMEGA
  // Initial velocity: the potential-flow start (default, the case's
  // [initialization] mode = potential_flow) or the inlet plug
  // (STAGE27_INITIAL_VELOCITY), and the Laplace solve's tolerance.
  bool potential = true;
  double potentialTol = 1e-10;
};

struct WallTopo {
  int nf = 0;                     // wall faces (the order of TurbTopo::wall)
  std::vector<std::int32_t> cell, lf;
  std::vector<double> n;          // [nf][3] outward unit normal
  std::vector<double> area, h;    // h = 3 |cell volume| / area
  // Twelve-point degree-6 rule on every face: weight area*w_q, barycentric
  // coordinates of the point in its cell, the basis values there (wall
  // trace) and the normal-column Picard functional (4-point Gauss-Legendre
  // mean of the basis over the first 90% of the inward normal chord through
  // the cell).
  std::vector<double> qw;         // [nf][12]
  std::vector<double> ql;         // [nf][12][4]
  std::vector<double> bw, bp;     // [nf][12][8]
  std::vector<double> ybar;       // [nf][12] mean chord height of the Picard column
  // Free DOFs (reference free numbering) of each face's bubble and its three
  // vertices (in the cell's face-vertex order).
  std::vector<std::int32_t> bub;  // [nf]
  std::vector<std::int32_t> vert; // [nf][3]
  // Wall DOFs for T^T / T application without races: row list of the DOFs
  // touched (bubbles and wall vertices), each with its faces.  For a bubble
  // the one face; for a vertex every wall face containing it.
  std::vector<std::int32_t> dof, dofOff, dofFace;
  std::vector<std::uint8_t> dofIsBubble;
  // Faces of each cell that lie on the wall: cellWallOff[nc+1], cellWall.
  std::vector<std::int32_t> cellWallOff, cellWall;
  // Gather of the wall face blocks into the momentum CSR (reference entry
  // order): entry k sums blocks blkIdx[blkOff[k] .. blkOff[k+1]) = face*64
  // + 8a + b, ascending.
  std::vector<std::int32_t> blkOff, blkIdx;
  // M3 transforms in reference CSR entry indices.
  // Right (K T): source entries (face, position of (i, bubble)) for every
  // row i coupled to the bubble; targets gather sources with coefficient 1
  // (the entry itself) or 20/27 (the entries (i, v) of the face vertices).
  std::vector<std::int32_t> rSrcFace, rSrcPos;
  std::vector<std::int32_t> rTgtPos, rTgtOff, rTgtSrc;  // rTgtSrc: src*2 + (0: coef 1, 1: coef 20/27)
  // Left (T^T K): source entries (face, position of (bubble, j)); targets
  // (bubble, j) itself and (v, j) for the face vertices.
  std::vector<std::int32_t> lSrcFace, lSrcPos;
  std::vector<std::int32_t> lTgtPos, lTgtOff, lTgtSrc;
};

// Needs the setup of a wall-function duct (S.wallFunction) and the turbulence
// topology, whose wall faces and wall-normal samples it shares.
WallTopo buildWallTopo(const TetMesh& M, const H8Setup& S, const TurbTopo& T);

// Four-point Gauss-Legendre rule on [-1, 1].
extern const double kGauss4[4][2];

}  // namespace nodals
