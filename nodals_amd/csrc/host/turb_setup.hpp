// Host setup of the k-omega SST transport step on the P1+BF3 tetrahedral
// space in the masked-P1 state space (vertex coefficients active, face-bubble
// coefficients held at zero): the scalar DOF numbering, the vertex-row
// operator pattern with its greedy colouring, the atomic-free gather maps of
// the assembled rows, the inlet face plans and the wall-normal sample points
// of the omega wall closure.
//
// Reference: turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
// modules/coupling/g2_step_k.inc, g2_step_omega.inc (step order),
// modules/coupling/g2_state.inc:12-27 (inlet k/omega, q1 mask),
// modules/transport/transport_preflight.inc + transport_finalize.inc
// (assembly and row-L1 relaxation), scalar_linear_solver.inc:33-67 (fixed
// coloured sweeps), tet_inlet_flux.inc (inlet flux), puretet_of_auto_point3.cuh
// (omega wall penalty), modules/wall/tet_wall_match_cache.inc:3-67 (sample
// location by adjacency walk) and modules/wall/tet_wall_ray_cache.inc:3-205
// (ray-traced sample placement).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "h8_setup.hpp"
#include "mesh.hpp"

namespace nodals {

struct TurbParams {
  double intensity = 0.05, lengthRatio = 0.07;  // inlet k = 1.5 (I U)^2, omega from l = ratio * D
  double alphaK = 0.3, alphaOmega = 0.3;        // row-L1 under-relaxation
  int kSweeps = 20, omegaSweeps = 20;           // fixed coloured GS sweeps per equation
  double gsOmega = 0.7;                         // sweep relaxation
  double kFloor = 1e-8, omegaFloor = 1e-3;
  bool deviatoric = false;                      // dev(2S):grad(U) production
  bool wall = true;                             // omega wall condition at the third sample
  double wallGamma = 1e6;
  // Omega wall closure (STAGE41_OMEGA_MODE, sst_application.inc:88-139;
  // of_auto_point3_rank1_penalty.inc, of_auto_trace_log_nitsche.inc):
  // 0 of_auto, the blended viscous/log target, and 1 sample_log, the pure
  // log target, each as a rank-one penalty at the third sample; 2
  // trace_log, the pure log target imposed on the wall face itself by
  // symmetric Nitsche with tau = traceGamma mu / h, h = 3 volume / area.
  // The reference binary's default is sample_log.
  int omegaMode = 1;
  double traceGamma = 50.0;
  // Storage of the relaxed sweep operator's off-diagonal values: false FP64;
  // true FP32 (an extension of this build: the sweeps read half the value
  // bytes; the state, diagonal, right-hand side and accumulation stay FP64).
  bool sweepFp32 = false;
  // Wall-normal sample heights; zeros select the reference rule
  // d_i * D / 0.004 * sampleScale (turbulence/sst_gpu_v2/nodals.py:289-292).
  std::array<double, 5> heights{{0, 0, 0, 0, 0}};
  double sampleScale = 0.5;
  double ofKappa = 0.41, ofE = 9.8, ofBeta1 = 0.075, ofReBlend = 11.0, ofProdLimit = 10.0;
  // Sample placement (SST_WALL_SAMPLING_MODE): 0 legacy_endpoint, the
  // adjacency walk to each endpoint (tet_wall_match_cache.inc:3-67); 1
  // ray_trace, the face-connected traversal of the whole wall-normal segment
  // (tet_wall_ray_cache.inc:3-205) for the active slots only: the Spalding
  // fit's points plus the omega wall point (the third), the other slots
  // aliasing the first active one.  rayLocalAdapt (STAGE27_LOCAL_RAY_ADAPT,
  // one active slot) shortens a segment that leaves the fluid to 0.90 of its
  // certified length, at least 0.25 of the height.
  int sampling = 0;
  int rayMaxCrossings = 4096;
  double rayTol = 1e-9;
  bool rayLocalAdapt = false;
  int activePoints = 5;  // Spalding fit points (SST24B_ACTIVE_POINTS)
  // Per-face third-sample heights from a previous segment's fitted friction
  // velocities (NODALS_FACE_YPLUS_UTAU_CSV, rows "face,cell,lf,u_tau,y+"):
  // y3 = target y+ nu / u_tau, clamped to [minScale, maxScale] times the
  // full-scale third height (the given third height / baseScale,
  // STAGE32_SPALDING_SAMPLE_SCALE); u_tau <= 1e-8 keeps the given height.
  // Ray tracing with one active slot and local shortening.
  std::string faceUtauCsv;
  double faceTargetSp = 30, faceTargetOmega = 30, faceMinScale = 0.02, faceMaxScale = 0.35, baseScale = 0;
};

// A boundary face seen from its tetrahedron: local face lf is the face
// opposite local vertex lf; n is the outward unit normal.
struct TurbFace {
  std::int32_t cell = -1, lf = -1;
  double n[3] = {0, 0, 0};
  double area = 0;
};

// One wall-normal sample: containing cell and its P1+BF3 basis values there,
// and its wall distance (the slot's height unless the ray was shortened).
struct TurbSample {
  std::int32_t cell = -1;
  double y = 0;
  double l[4] = {0, 0, 0, 0};
  double phi[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

struct TurbTopo {
  int nS = 0, nvert = 0, nc = 0;
  int nInlet = 0, nWallQ = 0;  // inlet faces, wall quadrature points (12 per wall face)
  std::vector<std::int32_t> cellEnt;  // [nc][8] scalar DOF (mesh entity) of each basis
  std::vector<TurbFace> inlet, wall;
  std::vector<TurbSample> samples;    // [wall face][12][5]
  // Operator of the vertex rows over vertex columns (sorted), in which the
  // coloured sweeps run (bubble coefficients are zero, so bubble columns
  // drop out of the sweeps but not of the row-L1 norms).
  std::vector<std::int32_t> row, col, diagPos;
  Coloring coloring;
  // Gather maps.  Blocks: cells [0, nc), inlet faces [nc, nc + nInlet) (on
  // the face's cell), wall quadrature points [nc + nInlet, nc + nInlet +
  // nWallQ) (on the cell of their third sample; on the wall face's own
  // cell for the trace_log closure).  Row i owns entries
  // [entOff[i], entOff[i+1]): first its CSR columns in order, then its
  // bubble columns ascending; entry e sums block values term[termOff[e] ..
  // termOff[e+1]) (block*64 + 8a + b, ascending).  Its right-hand side sums
  // rhsTerm[rhsOff[i] .. rhsOff[i+1]) (block*8 + a).
  std::vector<std::int32_t> entOff, termOff, term, rhsOff, rhsTerm;
  std::vector<std::int32_t> blockCell;  // cell of every block
  double kIn = 0, omegaIn = 0;
  std::array<double, 5> heights{{0, 0, 0, 0, 0}};  // sample heights in use
  // Placement statistics: traced (or walked) samples, those whose cell is
  // not the wall cell, the largest crossing count, rays shortened and the
  // smallest kept fraction of a shortened ray's height.
  std::int64_t placed = 0, crossed = 0, shortened = 0;
  int maxCrossings = 0;
  double minKeptFraction = 1.0;
  unsigned rayMask = 0x1f;
  // Per-face third-sample heights when loaded from fitted friction
  // velocities (empty otherwise), and the load summary: faces, fallbacks
  // (u_tau <= 1e-8), clamped Spalding and omega heights, then min, mean and
  // max of the Spalding and of the omega heights.
  std::vector<double> faceY3;
  std::array<double, 10> faceYplusStats{};
};

// Builds the topology for the H8 pipe (inlet and wall patches of S.pipe).
// Throws nodals::Error if a wall sample lies outside the mesh.
TurbTopo buildTurbTopo(const TetMesh& M, const H8Setup& S, const TurbParams& P);

// Local vertices of face lf (the face opposite vertex lf) in the reference
// order (modules/foundation/hxt4a_foundation.inc:37-39).
extern const int kTetFace[4][3];
// Twelve-point degree-6 triangle rule (Dunavant), barycentric coordinates on
// the three face vertices and weights summing to one, as tabulated by the
// reference (hxt5b_boundary.cuh:42-60).
extern const double kTri12[12][4];

}  // namespace nodals
