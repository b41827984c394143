// Tetrahedral mesh read from an OpenFOAM polyMesh directory.
//
// Same input contract as the reference reader
// (gpu/serial_cuda/foam_mesh_g2.hpp:55): triangular faces only, owner/
// neighbour labels define the cells, every cell must be a tetrahedron.  Cell
// vertices are the sorted vertex set with vertices 1 and 2 swapped when the
// Jacobian determinant is negative, and oppFace[c][i] is the face of cell c
// that does not contain vertex i.  Velocity DOF numbering and the cell plans
// depend on these conventions, so they are kept exactly.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace nodals {

struct Patch {
  std::string name;
  std::int32_t start = 0;
  std::int32_t count = 0;
};

struct TetMesh {
  std::vector<std::array<double, 3>> points;
  std::vector<std::array<std::int32_t, 3>> faces;
  std::vector<std::int32_t> owner, neighbour;
  std::vector<Patch> patches;
  std::vector<std::int32_t> facePatch;  // -1 for internal faces
  std::vector<std::array<std::int32_t, 4>> tets;
  std::vector<std::array<std::int32_t, 4>> oppFace;
  // Velocity DOF of every (cell, basis) in the numbering of the mesh file
  // (basis a < 4: vertex tets[c][a]; a >= 4: the face opposite vertex
  // tets[c][a - 4]): an HXT1 file's own tet velocity DOFs, or for a polyMesh
  // vertices then faces in this mesh's face order.  fileNvel DOFs, the first
  // fileNvert of them the vertices.  Checkpoints store fields this way.
  std::vector<std::array<std::int32_t, 8>> fileDof;
  std::int64_t fileNvel = 0, fileNvert = 0;

  int nPoints() const { return (int)points.size(); }
  int nFaces() const { return (int)faces.size(); }
  int nInternal() const { return (int)neighbour.size(); }
  int nCells() const { return (int)tets.size(); }
  int patchIndex(const std::string& name) const;
};

// Reads points/faces/owner/neighbour/boundary from `dir` and reconstructs
// the tetrahedra.  Throws nodals::Error on malformed input.
TetMesh readPolyMesh(const std::string& dir);

// HXT1 binary mesh of the SST application (tetrahedra only), as a face list
// with the H8 duct patch names: patch_0_0 wall, patch_2_0 inlet, patch_1_0
// outlet.  Throws nodals::Error on malformed input.
TetMesh readHxt1(const std::string& path);
bool isHxt1File(const std::string& path);
// readHxt1 for a file that starts with the HXT1 magic, readPolyMesh otherwise.
TetMesh readMesh(const std::string& path);

double det3(const double J[3][3]);
// Inverse by cofactors divided by the determinant, the operation order of the
// reference's inv3 (foam_mesh_g2.hpp:54) so that setup values are bitwise
// reproducible.
void inv3(const double J[3][3], double I[3][3]);

}  // namespace nodals
