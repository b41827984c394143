#include "ray_trace.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <tuple>

#include "error.hpp"

namespace nodals {

namespace {

using P3 = std::array<double, 3>;

P3 sub(const P3& a, const P3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
P3 along(const P3& o, double t, const P3& d) { return {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]}; }
double dot(const P3& a, const P3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
P3 cross(const P3& a, const P3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double minOf(const double l[4]) { return std::min(std::min(l[0], l[1]), std::min(l[2], l[3])); }

[[noreturn]] void rayFail(const std::string& m) { fail("RAY_TRACE " + m); }

// At most this many cells are examined around one crossing point (the cells
// meeting at an edge or vertex the segment passes through).
constexpr std::size_t kFanLimit = 256;

}  // namespace

TetRayTracer::TetRayTracer(const std::vector<P3>& points, const std::vector<std::array<std::int32_t, 4>>& tets) {
  if (points.empty() || tets.empty()) rayFail("empty mesh");
  const std::size_t nt = tets.size(), np = points.size();
  geo_.resize(nt);
  adj_.assign(nt, {-1, -1, -1, -1});
  // (sorted face vertices, cell, local face) for every cell face
  std::vector<std::tuple<std::array<std::int32_t, 3>, std::int32_t, std::int32_t>> keys;
  keys.reserve(4 * nt);
  for (std::size_t c = 0; c < nt; ++c) {
    const auto& v = tets[c];
    for (int j = 0; j < 4; ++j)
      if (v[(std::size_t)j] < 0 || (std::size_t)v[(std::size_t)j] >= np) rayFail("vertex index out of range");
    for (int a = 0; a < 4; ++a)
      for (int b = a + 1; b < 4; ++b)
        if (v[(std::size_t)a] == v[(std::size_t)b]) rayFail("repeated vertex in a tetrahedron");
    Geo& G = geo_[c];
    G.o = points[(std::size_t)v[0]];
    const P3 e1 = sub(points[(std::size_t)v[1]], G.o), e2 = sub(points[(std::size_t)v[2]], G.o),
             e3 = sub(points[(std::size_t)v[3]], G.o);
    const double det = dot(e1, cross(e2, e3));
    const double s2 = std::max(std::max(dot(e1, e1), dot(e2, e2)), dot(e3, e3));
    if (!(s2 > 0) || !std::isfinite(det) || std::abs(det) <= 1e-13 * s2 * std::sqrt(s2))
      rayFail("degenerate tetrahedron " + std::to_string(c));
    G.scale = std::sqrt(s2);
    // rows of the inverse of [e1 e2 e3]: (e2 x e3, e3 x e1, e1 x e2) / det
    const P3 r1 = cross(e2, e3), r2 = cross(e3, e1), r3 = cross(e1, e2);
    for (int d = 0; d < 3; ++d) {
      G.g[0][d] = r1[(std::size_t)d] / det;
      G.g[1][d] = r2[(std::size_t)d] / det;
      G.g[2][d] = r3[(std::size_t)d] / det;
    }
    for (int j = 0; j < 4; ++j) {
      std::array<std::int32_t, 3> k{};
      int n = 0;
      for (int a = 0; a < 4; ++a)
        if (a != j) k[(std::size_t)n++] = v[(std::size_t)a];
      std::sort(k.begin(), k.end());
      keys.emplace_back(k, (std::int32_t)c, j);
    }
  }
  std::sort(keys.begin(), keys.end());
  for (std::size_t i = 0; i < keys.size();) {
    std::size_t e = i + 1;
    while (e < keys.size() && std::get<0>(keys[e]) == std::get<0>(keys[i])) ++e;
    if (e - i > 2) rayFail("non-manifold face (shared by more than two tetrahedra)");
    if (e - i == 2) {
      const int c0 = std::get<1>(keys[i]), j0 = std::get<2>(keys[i]);
      const int c1 = std::get<1>(keys[i + 1]), j1 = std::get<2>(keys[i + 1]);
      adj_[(std::size_t)c0][(std::size_t)j0] = c1;
      adj_[(std::size_t)c1][(std::size_t)j1] = c0;
    }
    i = e;
  }
}

void TetRayTracer::bary(int c, const P3& p, double l[4]) const {
  const Geo& G = geo_[(std::size_t)c];
  const P3 r = sub(p, G.o);
  l[1] = G.g[0][0] * r[0] + G.g[0][1] * r[1] + G.g[0][2] * r[2];
  l[2] = G.g[1][0] * r[0] + G.g[1][1] * r[1] + G.g[1][2] * r[2];
  l[3] = G.g[2][0] * r[0] + G.g[2][1] * r[1] + G.g[2][2] * r[2];
  l[0] = 1.0 - l[1] - l[2] - l[3];
}

void TetRayTracer::interval(int c, const P3& origin, const P3& d, double tol, double& lo, double& hi) const {
  // Barycentric coordinates are affine along the ray: l(t) = l(origin) + t s.
  const Geo& G = geo_[(std::size_t)c];
  double l0[4], s[4];
  bary(c, origin, l0);
  s[1] = G.g[0][0] * d[0] + G.g[0][1] * d[1] + G.g[0][2] * d[2];
  s[2] = G.g[1][0] * d[0] + G.g[1][1] * d[1] + G.g[1][2] * d[2];
  s[3] = G.g[2][0] * d[0] + G.g[2][1] * d[1] + G.g[2][2] * d[2];
  s[0] = -s[1] - s[2] - s[3];
  lo = -std::numeric_limits<double>::infinity();
  hi = std::numeric_limits<double>::infinity();
  for (int j = 0; j < 4; ++j) {
    const double at = (-tol - l0[j]);  // l_j(t) >= -tol  <=>  s_j t >= at
    if (s[j] > 0) lo = std::max(lo, at / s[j]);
    else if (s[j] < 0) hi = std::min(hi, at / s[j]);
    else if (l0[j] < -tol) {
      lo = std::numeric_limits<double>::infinity();
      hi = -std::numeric_limits<double>::infinity();
      return;
    }
  }
}

std::vector<RayHit> TetRayTracer::trace(int start, const P3& origin, const P3& dir, const std::vector<double>& distances,
                                        int maxCrossings, double tol) const {
  if (start < 0 || start >= cells() || maxCrossings < 1 || !(tol > 0) || tol > 1e-5)
    rayFail("invalid start cell, crossing limit or tolerance");
  const double len = std::sqrt(dot(dir, dir));
  if (!(len > 0) || !std::isfinite(len)) rayFail("invalid direction");
  const P3 d = {dir[0] / len, dir[1] / len, dir[2] / len};
  double l[4];
  bary(start, origin, l);
  if (minOf(l) < -tol) rayFail("origin outside its start cell");
  const double coordScale = std::max({std::abs(origin[0]), std::abs(origin[1]), std::abs(origin[2]), 1e-3});
  int cell = start, crossings = 0;
  double progress = 0.0, previous = 0.0;
  std::vector<RayHit> hits;
  hits.reserve(distances.size());
  std::vector<int> fan;
  fan.reserve(kFanLimit);
  for (const double target : distances) {
    if (!std::isfinite(target) || !(target > previous)) rayFail("sample distances must be positive and increasing");
    previous = target;
    const P3 end = along(origin, target, d);
    for (;;) {
      bary(cell, end, l);
      if (minOf(l) >= -tol) {
        RayHit h;
        h.cell = cell;
        for (int j = 0; j < 4; ++j) h.l[j] = l[j];
        h.crossings = crossings;
        hits.push_back(h);
        progress = target;
        break;
      }
      // The sample lies beyond this cell: leave it where its interval ends.
      double lo, t;
      interval(cell, origin, d, tol, lo, t);
      const double eps = std::max(128.0 * std::numeric_limits<double>::epsilon() * coordScale,
                                  4.0 * tol * geo_[(std::size_t)cell].scale);
      if (!std::isfinite(t) || t < progress - eps || t >= target + eps)
        rayFail("no valid exit from cell " + std::to_string(cell));
      if (++crossings > maxCrossings) rayFail("more than " + std::to_string(maxCrossings) + " cells crossed");
      const P3 x = along(origin, t, d);
      // Cells around the crossing point, reached through faces that contain
      // it; the next cell is the one whose interval reaches farthest.
      fan.assign(1, cell);
      int best = -1;
      double bestFar = t, bestMargin = -std::numeric_limits<double>::infinity();
      for (std::size_t q = 0; q < fan.size() && q < kFanLimit; ++q) {
        const int c = fan[q];
        double lx[4];
        bary(c, x, lx);
        if (minOf(lx) < -8 * tol) continue;
        if (c != cell) {
          double clo, chi;
          interval(c, origin, d, tol, clo, chi);
          const double ceps = std::max(eps, 4.0 * tol * geo_[(std::size_t)c].scale);
          if (clo <= t + ceps && chi > t + 0.01 * ceps) {
            const double reach = std::min(target, chi);
            double lw[4];
            bary(c, along(origin, t + 0.5 * std::min(target - t, chi - t), d), lw);
            const double margin = minOf(lw);
            if (margin >= -tol &&
                (reach > bestFar + 0.01 * ceps || (std::abs(reach - bestFar) <= 0.01 * ceps && margin > bestMargin))) {
              best = c;
              bestFar = reach;
              bestMargin = margin;
            }
          }
        }
        for (int j = 0; j < 4; ++j) {
          if (std::abs(lx[j]) > 8 * tol) continue;
          const int n = adj_[(std::size_t)c][(std::size_t)j];
          if (n >= 0 && fan.size() < kFanLimit && std::find(fan.begin(), fan.end(), n) == fan.end()) fan.push_back(n);
        }
      }
      if (best < 0)
        rayFail("segment left the fluid at distance " + std::to_string(t) + " before the sample at " +
                std::to_string(target) + " (cell " + std::to_string(cell) + ")");
      if (best == cell || !(bestFar > t)) rayFail("no forward progress from cell " + std::to_string(cell));
      progress = std::max(progress, t);
      cell = best;
    }
  }
  return hits;
}

double TetRayTracer::certifiedDistance(int start, const P3& origin, const P3& dir, double requested, int maxCrossings,
                                       double tol) const {
  double lo = 0.0, hi = requested;
  for (int it = 0; it < 28; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (!(mid > 0.0)) break;
    try {
      trace(start, origin, dir, std::vector<double>{mid}, maxCrossings, tol);
      lo = mid;
    } catch (const Error&) {
      hi = mid;
    }
  }
  return lo;
}

}  // namespace nodals
