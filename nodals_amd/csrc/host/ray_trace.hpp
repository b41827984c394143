// Straight-segment traversal of a tetrahedral mesh: wall-normal sample points
// located by walking the segment from its wall origin through face-connected
// cells, so a sample is accepted only when the whole segment from the wall to
// it lies in the fluid.  A segment that reaches a boundary before its sample
// is an error, even where the sample point itself lies in another part of the
// mesh.
//
// Reference: turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/modules/
// wall/portable/tet_ray_trace.hpp (TetRayTrace; its interval traversal,
// trace_intervals :212-318, is the method used here), its consumer
// modules/wall/tet_wall_ray_cache.inc:3-205 (active slots, local shortening
// with 28 bisection steps, 0.90 safety and 0.25 floor), selected by
// SST_WALL_SAMPLING_MODE=ray_trace (modules/app/sst_application.inc:1141-1168).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace nodals {

struct RayHit {
  std::int32_t cell = -1;
  double l[4] = {0, 0, 0, 0};  // barycentric coordinates in `cell`
  std::int32_t crossings = 0;  // cells left before this sample, counted from the origin
};

class TetRayTracer {
 public:
  // points [np][3], tets [nt][4]; builds the face adjacency.  Throws
  // nodals::Error on an empty, degenerate or non-manifold mesh.
  TetRayTracer(const std::vector<std::array<double, 3>>& points, const std::vector<std::array<std::int32_t, 4>>& tets);

  // Samples at strictly increasing distances along `dir` (any length) from
  // `origin`, which must lie in `start` (barycentric tolerance `tol`).
  // Throws nodals::Error when the segment leaves the fluid, a crossing
  // cannot be followed, or more than `maxCrossings` cells are crossed.
  std::vector<RayHit> trace(int start, const std::array<double, 3>& origin, const std::array<double, 3>& dir,
                            const std::vector<double>& distances, int maxCrossings = 4096, double tol = 1e-9) const;

  // Largest distance along `dir`, to within requested * 2^-28, up to which a
  // single-sample trace succeeds (bisection over [0, requested]).
  double certifiedDistance(int start, const std::array<double, 3>& origin, const std::array<double, 3>& dir,
                           double requested, int maxCrossings, double tol) const;

  void bary(int c, const std::array<double, 3>& p, double l[4]) const;
  int cells() const { return (int)geo_.size(); }
  const std::array<std::int32_t, 4>& neighbours(int c) const { return adj_[(std::size_t)c]; }

 private:
  struct Geo {
    std::array<double, 3> o;     // vertex 0
    double g[3][3];              // lambda_{1..3} = g (p - o)
    double scale;                // longest edge from vertex 0
  };
  // [lo, hi]: the ray parameters at which every barycentric coordinate of
  // cell c is >= -tol (empty when lo > hi).
  void interval(int c, const std::array<double, 3>& origin, const std::array<double, 3>& d, double tol, double& lo,
                double& hi) const;

  std::vector<Geo> geo_;
  std::vector<std::array<std::int32_t, 4>> adj_;  // neighbour across the face opposite local vertex j
};

}  // namespace nodals
