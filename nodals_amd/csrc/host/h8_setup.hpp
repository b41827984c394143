// Host setup of the H8 laminar SIMPLE path: Dirichlet masks, free velocity
// DOF numbering, cell plans, P1+BF3 reference tensors, the momentum CSR with
// its diffusion values, greedy coloring, fixed-value right-hand sides and the
// fine P0 Schur (B diag(rAU) B^T) CSR topology.
//
// Reference: gpu/serial_cuda/g2_host.hpp, g3_host.hpp, g4_host.hpp,
// h2b_fine_csr.cuh (host part).  Integer outputs (numbering, CSR patterns,
// slots, colors) are bit-exact with the reference; floating-point setup values
// follow the reference's operation order so they are bitwise reproducible on
// the host as well (this file must be compiled with -ffp-contract=off).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "mesh.hpp"

namespace nodals {

// Pipe description (g4_host.hpp:19-49).
struct PipeInfo {
  int wall = -1, inlet = -1, outlet = -1;
  double cx = 0, cy = 0, zIn = 0, zOut = 0, R = 0, D = 0, L = 0;
  double bulk = 1, profileScale = 1, nu = 0, re = 20;
  double inletArea = 0, inletProjectedArea = 0, circleArea = 0, hpGradient = 0, hpDrop = 0;
  double inletNormal[3] = {0, 0, -1};
  double inletVelocity[3] = {0, 0, 1};
};

// Per-tet plan (g4_host.hpp:51-60 without the DG-inlet/weak-wall extras).
// ref[a] >= 0: free velocity DOF; ref[a] < 0: -(fixedSlot+1).
// rowSlot[8a+b]: position of column ref[b] inside momentum row ref[a]
// (255 when either DOF is fixed).
struct CellPlan {
  std::int32_t ref[8];
  std::uint8_t rowSlot[64];
  double det;
  double h2;
  double invJ[9];
};

struct CsrD {
  int n = 0;
  std::vector<std::int64_t> row;
  std::vector<std::int32_t> col;
  std::vector<double> val;
  std::vector<double> diag;
};

struct Coloring {
  int ncolors = 0;
  std::vector<std::int32_t> color, offset, rows;
};

// P0 Schur operator topology.  The first block is the reference layout
// (h2b_fine_csr.cuh:21-55): row/col/diagPos, free-velocity incidence
// incOff/packed ((cell<<3)|basis) and the per-row directed contribution
// slots.  The second block is the gather layout used by the HIP refresh:
// for every nonzero k the list nzTerm[nzOff[k]..nzOff[k+1]) of shared
// velocity DOFs, packed as (g<<6)|(a<<3)|b with a the basis of DOF g in the
// row cell and b its basis in the column cell.
struct FineTopo {
  int n = 0, nv = 0;
  std::vector<std::int32_t> row, col, diagPos;
  std::vector<std::int32_t> incOff;
  std::vector<std::uint32_t> packed;
  std::vector<std::int32_t> contribOff;
  std::vector<std::uint8_t> slot;
  std::vector<std::int32_t> nzOff;
  std::vector<std::uint32_t> nzTerm;
  std::uint32_t rowMax = 0;
};

struct H8Params {
  std::string wall = "patch_0_0", inlet = "patch_2_0", outlet = "patch_1_0";
  // Lid-driven cavity instead of the pipe when non-empty: every boundary
  // face is Dirichlet, the lid patch moves at (bulk, 0, 0), nu = bulk L / re
  // with L the largest box extent, and the pressure is pinned in cell 0.
  std::string lid;
  // General duct instead of the straight z-aligned pipe: a uniform inlet
  // velocity bulk along the inward inlet normal, D the hydraulic diameter of
  // the inlet (2 sqrt(A/pi)), the pressure drop as the difference of the
  // face-area-weighted pressures next to the inlet and the outlet, and a
  // geometric wall-distance field on the vertices (for the SST terms).
  bool duct = false;
  // Duct mode only: the wall patch is a wall-function wall (free velocity
  // DOFs with a Nitsche/Spalding wall closure) instead of no-slip.
  bool wallFunction = false;
  // Duct mode only: nu = bulk * ductDiameter / re when ductDiameter > 0
  // (else the inlet's hydraulic diameter), and an explicit inlet velocity.
  double ductDiameter = 0.0;
  bool inletVelocityGiven = false;
  double inletVelocity[3] = {0, 0, 0};
  // Straight pipe only: the DG weak inlet (build_g4_setup's dgInlet,
  // gpu/serial_cuda/g4_host.hpp:289-313): inlet DOFs free, the divergence
  // without the inlet face trace and the inlet flux prescribed in the fixed
  // divergence; the setup assembly starts from the plug.
  bool dgInlet = false;
  double re = 20.0, bulk = 1.0;
  double alphaU = 0.5;
  double rauScale = 1.0;
  int threads = 0;  // 0: hardware concurrency (capped at 16)
};

// A DG inlet face seen from its cell: local face lf (opposite vertex lf),
// area and outward unit normal.
struct DgInletFace {
  std::int32_t cell = -1, lf = -1;
  double area = 0;
  double n[3] = {0, 0, 0};
};

struct H8Setup {
  PipeInfo pipe;
  bool cavity = false;
  bool duct = false;
  bool wallFunction = false;
  bool dgInlet = false;
  std::vector<DgInletFace> dgFaces;   // DG inlet faces, in inlet patch order
  std::vector<double> plug;           // [nv][3] start velocity of the DG inlet: the plug on free vertices
  int pinCell = -1;  // pressure row pinned (cavity), -1 for none
  // duct mode: cells next to the inlet / outlet with their face areas, and
  // the distance of every vertex to the nearest wall face
  std::vector<std::int32_t> inletCells, outletCells;
  std::vector<double> inletAreas, outletAreas;
  std::vector<double> wallDistance;
  int nc = 0, nv = 0, nFixed = 0;
  int nPoints = 0, nFaces = 0;
  int freeVertices = 0, freeFaces = 0, fixedVertices = 0, fixedFaces = 0;
  std::vector<std::uint8_t> fixed;          // per entity (vertices, then faces)
  std::vector<std::int32_t> g2free;         // entity -> free DOF or -1
  std::vector<std::int32_t> fixedSlot;      // entity -> fixed slot or -1
  std::vector<std::array<double, 3>> fixedValue;
  std::vector<CellPlan> cells;
  std::vector<double> bcoef;                // [nc][8][3] effective B coefficients
  CsrD mom;                                 // momentum topology + diffusion at nu (alpha 1)
  std::vector<std::int32_t> momDiagPos;
  Coloring coloring;
  std::array<std::vector<double>, 3> staticRhs;
  std::vector<double> fixedDiv, volumes;
  std::vector<double> initRau, initDelta;   // rAU of the U=0 assembly
  std::vector<double> cellZ;                // centroid z per cell (pressure-drop fit)
  double zMin = 0, zMax = 0;
};

// Reference tensors (g3_host.hpp:58-100, g4_host.hpp:92-98).
struct Tensors {
  double diff[8][8][3][3];
  double cent[8][8][8][3];
};
const Tensors& referenceTensors();

H8Setup buildH8Setup(const TetMesh& M, const H8Params& P);

// B coefficient of basis a, component d (g4_host.hpp:64-73 without inlet/wall
// modifications): (det/6) * grad(lambda_i) for vertices, -(27/20) of it for
// the face bubble opposite vertex i.
double bcoefOf(const CellPlan& cp, int a, int d);

// Host central convection + row-L1 relaxation at a free-velocity state
// (g4_host.hpp:317-422); used for the setup rAU and by host parity checks.
void hostAssembleCentral(const H8Setup& S, const std::array<std::vector<double>, 3>& U,
                         std::vector<double>& aval, std::array<std::vector<double>, 3>& conv);
std::vector<double> hostFinalizeRelax(const H8Setup& S, std::vector<double>& a, double alphaU,
                                      std::vector<double>* delta);

Coloring greedyColoring(int n, const std::int64_t* row, const std::int32_t* col);
Coloring greedyColoring32(int n, const std::int32_t* row, const std::int32_t* col);

FineTopo buildFineTopo(const H8Setup& S);

// Linear least-squares fit of p over cell-centroid z, times -(zmax-zmin)
// (g4_host.hpp:424).
double pressureDropFit(const H8Setup& S, const std::vector<double>& p);
// Duct mode: area-weighted mean pressure of the inlet cells minus that of
// the outlet cells.
double pressureDropPatches(const H8Setup& S, const std::vector<double>& p);
// Distance of every mesh vertex to the nearest face of the given patch
// (exact point-triangle distance; uniform-grid search).
std::vector<double> wallDistanceField(const TetMesh& M, int patch, int threads);

int hostThreads(int requested);

}  // namespace nodals
