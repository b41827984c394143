// Host setup of the k-omega transport step (see turb_setup.hpp).
#include "turb_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "error.hpp"
#include "ray_trace.hpp"

namespace nodals {

const int kTetFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

const double kTri12[12][4] = {
    {0.063089014491502, 0.063089014491502, 0.873821971016996, 0.050844906370207},
    {0.063089014491502, 0.873821971016996, 0.063089014491502, 0.050844906370207},
    {0.873821971016996, 0.063089014491502, 0.063089014491502, 0.050844906370207},
    {0.249286745170910, 0.249286745170910, 0.501426509658180, 0.116786275726379},
    {0.249286745170910, 0.501426509658180, 0.249286745170910, 0.116786275726379},
    {0.501426509658180, 0.249286745170910, 0.249286745170910, 0.116786275726379},
    {0.053145049844816, 0.310352451033785, 0.636502499121399, 0.082851075618374},
    {0.053145049844816, 0.636502499121399, 0.310352451033785, 0.082851075618374},
    {0.310352451033785, 0.053145049844816, 0.636502499121399, 0.082851075618374},
    {0.310352451033785, 0.636502499121399, 0.053145049844816, 0.082851075618374},
    {0.636502499121399, 0.053145049844816, 0.310352451033785, 0.082851075618374},
    {0.636502499121399, 0.310352451033785, 0.053145049844816, 0.082851075618374}};

namespace {

using V3 = std::array<double, 3>;

V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Barycentric coordinates of p in cell c (l0 = 1 - l1 - l2 - l3).
bool barycentric(const TetMesh& M, int c, const V3& p, double l[4]) {
  const auto& t = M.tets[(std::size_t)c];
  const V3 x0 = M.points[(std::size_t)t[0]];
  const V3 e1 = sub(M.points[(std::size_t)t[1]], x0), e2 = sub(M.points[(std::size_t)t[2]], x0),
           e3 = sub(M.points[(std::size_t)t[3]], x0), r = sub(p, x0);
  const double det = dot(e1, cross(e2, e3));
  if (!(std::abs(det) > 0.0) || !std::isfinite(det)) return false;
  l[1] = dot(r, cross(e2, e3)) / det;
  l[2] = dot(e1, cross(r, e3)) / det;
  l[3] = dot(e1, cross(e2, r)) / det;
  l[0] = 1.0 - l[1] - l[2] - l[3];
  return std::isfinite(l[0]) && std::isfinite(l[1]) && std::isfinite(l[2]) && std::isfinite(l[3]);
}

void basisAt(const double l[4], double phi[8]) {
  for (int i = 0; i < 4; ++i) phi[i] = l[i];
  for (int i = 0; i < 4; ++i) {
    double p = 27.0;
    for (int j = 0; j < 4; ++j)
      if (j != i) p *= l[j];
    phi[4 + i] = p;
  }
}

TurbFace facePlan(const TetMesh& M, int f) {
  TurbFace F;
  const int c = M.owner[(std::size_t)f];
  F.cell = c;
  for (int i = 0; i < 4; ++i)
    if (M.oppFace[(std::size_t)c][(std::size_t)i] == f) F.lf = i;
  require(F.lf >= 0, "boundary face not found in its owner cell", kErrInternal);
  const auto& fv = M.faces[(std::size_t)f];
  const V3 p0 = M.points[(std::size_t)fv[0]], p1 = M.points[(std::size_t)fv[1]], p2 = M.points[(std::size_t)fv[2]];
  V3 n = cross(sub(p1, p0), sub(p2, p0));
  const double len = std::sqrt(dot(n, n));
  require(len > 0.0, "degenerate boundary face");
  F.area = 0.5 * len;
  // outward: away from the cell's vertex opposite the face
  const V3 opp = M.points[(std::size_t)M.tets[(std::size_t)c][(std::size_t)F.lf]];
  const double s = dot(n, sub(p0, opp)) > 0.0 ? 1.0 : -1.0;
  for (int d = 0; d < 3; ++d) F.n[d] = s * n[d] / len;
  return F;
}

// Per-face third-sample heights from fitted friction velocities
// (sst_application.inc:1039-1097, NODALS_FACE_YPLUS_UTAU_CSV).
std::vector<double> loadFaceY3(const TurbParams& P, TurbTopo& T, double nu) {
  require(P.activePoints == 1, "per-face y+ heights need one active Spalding point (the third)");
  require(P.faceTargetSp > 0 && P.faceTargetOmega > 0 && P.faceTargetSp <= 500 && P.faceTargetOmega <= 500 &&
              P.faceMinScale > 0 && P.faceMinScale <= P.faceMaxScale && P.faceMaxScale <= 1,
          "per-face y+ target or scale range invalid");
  require(P.faceTargetOmega == P.faceTargetSp,
          "per-face y+ heights: the omega target must equal the Spalding target (one third-sample point serves both)");
  require(P.baseScale > 0, "per-face y+ heights need the Spalding sample scale of the given heights");
  const double y3 = T.heights[2], full = y3 / P.baseScale;
  const double yMin = full * P.faceMinScale, yMax = full * P.faceMaxScale;
  std::ifstream in(P.faceUtauCsv);
  require((bool)in, "cannot open the face friction-velocity file " + P.faceUtauCsv);
  std::vector<double> out;
  out.reserve(T.wall.size());
  std::size_t fallback = 0, clampSp = 0, clampOm = 0;
  double spMin = 1e300, spMax = 0, spSum = 0, omMin = 1e300, omMax = 0, omSum = 0;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    long long fi = -1, cell = -1, lf = -1;
    double ut = 0, yp = 0;
    const std::size_t n = out.size();
    require(std::sscanf(line.c_str(), "%lld,%lld,%lld,%lf,%lf", &fi, &cell, &lf, &ut, &yp) == 5 &&
                fi == (long long)n && n < T.wall.size() && cell == T.wall[n].cell && lf == T.wall[n].lf &&
                std::isfinite(ut) && ut >= 0.0,
            "face friction-velocity file does not match the wall faces at row " + std::to_string(n));
    const bool fit = ut > 1e-8;
    if (!fit) ++fallback;
    const double sp0 = fit ? P.faceTargetSp * nu / ut : y3;
    const double om0 = fit ? P.faceTargetOmega * nu / ut : y3 * (P.faceTargetOmega / P.faceTargetSp);
    const double sp = std::min(yMax, std::max(yMin, sp0)), om = std::min(yMax, std::max(yMin, om0));
    clampSp += std::abs(sp - sp0) > 1e-13;
    clampOm += std::abs(om - om0) > 1e-13;
    spMin = std::min(spMin, sp);
    spMax = std::max(spMax, sp);
    spSum += sp;
    omMin = std::min(omMin, om);
    omMax = std::max(omMax, om);
    omSum += om;
    out.push_back(sp);
  }
  require(!out.empty() && out.size() == T.wall.size(), "face friction-velocity file has the wrong number of faces");
  const double nf = (double)out.size();
  T.faceYplusStats = {nf, (double)fallback, (double)clampSp, (double)clampOm, spMin, spSum / nf, spMax,
                      omMin, omSum / nf, omMax};
  return out;
}

}  // namespace

TurbTopo buildTurbTopo(const TetMesh& M, const H8Setup& S, const TurbParams& P) {
  require(!S.cavity, "the k-omega step needs the pipe case (inlet and wall patches)");
  require(P.alphaK > 0 && P.alphaK <= 1 && P.alphaOmega > 0 && P.alphaOmega <= 1, "under-relaxation must be in (0, 1]");
  require(P.kSweeps >= 0 && P.omegaSweeps >= 0, "sweep counts must be >= 0");
  require(P.kFloor > 0 && P.omegaFloor > 0, "k and omega floors must be positive");
  TurbTopo T;
  const int np = M.nPoints(), nf = M.nFaces(), nc = M.nCells();
  T.nvert = np;
  T.nS = np + nf;
  T.nc = nc;
  T.cellEnt.resize((std::size_t)nc * 8);
  for (int c = 0; c < nc; ++c)
    for (int a = 0; a < 8; ++a)
      T.cellEnt[(std::size_t)c * 8 + a] =
          a < 4 ? M.tets[(std::size_t)c][(std::size_t)a] : np + M.oppFace[(std::size_t)c][(std::size_t)(a - 4)];

  const double D = S.pipe.D, bulk = S.pipe.bulk;
  T.kIn = 1.5 * std::pow(P.intensity * bulk, 2);
  T.omegaIn = std::sqrt(T.kIn) / (std::pow(0.09, 0.25) * P.lengthRatio * D);
  static const double kOriginal[5] = {1.4694676501540146e-4, 2.9389353003080293e-4, 4.4084029504620439e-4,
                                      5.8778706006160585e-4, 7.3473382507700737e-4};
  for (int m = 0; m < 5; ++m)
    T.heights[(std::size_t)m] = P.heights[0] > 0 ? P.heights[(std::size_t)m] : kOriginal[m] * D / 0.004 * P.sampleScale;
  for (int m = 1; m < 5; ++m)
    require(T.heights[(std::size_t)m] > T.heights[(std::size_t)m - 1] && T.heights[0] > 0,
            "wall sample heights must be positive and increasing");

  // Boundary faces of the inlet and wall patches.
  for (int f = M.nInternal(); f < nf; ++f) {
    const int p = M.facePatch[(std::size_t)f];
    if (p == S.pipe.inlet) T.inlet.push_back(facePlan(M, f));
    if (P.wall && p == S.pipe.wall) T.wall.push_back(facePlan(M, f));
  }
  T.nInlet = (int)T.inlet.size();
  T.nWallQ = (int)T.wall.size() * 12;

  // Wall-normal samples.
  require(P.sampling == 0 || P.sampling == 1, "wall sampling must be legacy_endpoint or ray_trace");
  require(P.faceUtauCsv.empty() || P.sampling == 1, "per-face y+ sample heights need ray_trace wall sampling");
  require(P.activePoints >= 1 && P.activePoints <= 5, "active wall points must be in 1..5");
  T.samples.resize((std::size_t)T.nWallQ * 5);
  auto quadPoint = [&](const TurbFace& F, int q) {
    const auto& t = M.tets[(std::size_t)F.cell];
    V3 X = {0, 0, 0};
    for (int j = 0; j < 3; ++j) {
      const V3& v = M.points[(std::size_t)t[(std::size_t)kTetFace[F.lf][j]]];
      for (int d = 0; d < 3; ++d) X[(std::size_t)d] += kTri12[q][j] * v[(std::size_t)d];
    }
    return X;
  };
  auto place = [&](TurbSample& s, int cell, const double l[4], double y) {
    s.cell = cell;
    s.y = y;
    for (int j = 0; j < 4; ++j) s.l[j] = l[j];
    basisAt(s.l, s.phi);
  };
  if (P.sampling == 0) {
    // Adjacency walk from the wall cell towards each endpoint.
    auto across = [&](int c, int j) {
      const int f = M.oppFace[(std::size_t)c][(std::size_t)j];
      if (M.owner[(std::size_t)f] != c) return (int)M.owner[(std::size_t)f];
      return f < M.nInternal() ? (int)M.neighbour[(std::size_t)f] : -1;
    };
    std::vector<int> seen((std::size_t)nc, 0);
    int stamp = 0;
    for (std::size_t fi = 0; fi < T.wall.size(); ++fi) {
      const TurbFace& F = T.wall[fi];
      for (int q = 0; q < 12; ++q) {
        const V3 X = quadPoint(F, q);
        for (int m = 0; m < 5; ++m) {
          const double y = T.heights[(std::size_t)m];
          const V3 p = {X[0] - y * F.n[0], X[1] - y * F.n[1], X[2] - y * F.n[2]};
          int c = F.cell, found = -1, steps = 0;
          double l[4] = {0, 0, 0, 0};
          ++stamp;
          while (c >= 0 && steps < 4096 && seen[(std::size_t)c] != stamp) {
            seen[(std::size_t)c] = stamp;
            ++steps;
            if (!barycentric(M, c, p, l)) break;
            int neg = 0;
            for (int j = 1; j < 4; ++j)
              if (l[j] < l[neg]) neg = j;
            if (l[neg] >= -1e-8) {
              found = c;
              break;
            }
            c = across(c, neg);
          }
          require(found >= 0, "wall sample outside the mesh (face " + std::to_string(fi) + ", point " +
                                  std::to_string(q) + ", height " + std::to_string(y) + ")");
          place(T.samples[((std::size_t)fi * 12 + q) * 5 + m], found, l, y);
          ++T.placed;
          if (found != F.cell) ++T.crossed;
          T.maxCrossings = std::max(T.maxCrossings, steps - 1);
        }
      }
    }
    T.rayMask = 0x1f;
  } else {
    if (!P.faceUtauCsv.empty()) T.faceY3 = loadFaceY3(P, T, S.pipe.nu);
    // Ray-traced active slots: the Spalding fit's points and the omega wall
    // point (slot 3); inactive slots alias the first active slot's hit and
    // keep their own height.
    static const unsigned kFitMask[5] = {0x04u, 0x11u, 0x15u, 0x1bu, 0x1fu};
    const unsigned mask = P.activePoints == 5 ? 0x1fu : (kFitMask[P.activePoints - 1] | 0x04u);
    std::vector<int> slots;
    std::vector<double> base;
    for (int m = 0; m < 5; ++m)
      if (mask & (1u << m)) {
        slots.push_back(m);
        base.push_back(T.heights[(std::size_t)m]);
      }
    const bool adapt = P.rayLocalAdapt || !T.faceY3.empty();
    require(!adapt || slots.size() == 1,
            "ray local shortening needs exactly one active wall sample (one Spalding point at the omega slot)");
    TetRayTracer tracer(M.points, M.tets);
    for (std::size_t fi = 0; fi < T.wall.size(); ++fi) {
      const TurbFace& F = T.wall[fi];
      const std::array<double, 3> inward = {-F.n[0], -F.n[1], -F.n[2]};
      for (int q = 0; q < 12; ++q) {
        const V3 X = quadPoint(F, q);
        std::vector<double> dist = base;
        if (!T.faceY3.empty()) {
          const double ratio = T.faceY3[fi] / T.heights[2];
          for (double& v : dist) v *= ratio;
        }
        std::vector<RayHit> hits;
        try {
          hits = tracer.trace(F.cell, X, inward, dist, P.rayMaxCrossings, P.rayTol);
        } catch (const Error& e) {
          if (!adapt)
            fail("wall sample ray (face " + std::to_string(fi) + ", point " + std::to_string(q) + ", slot " +
                 std::to_string(slots[0] + 1) + "): " + e.what());
          const double want = dist[0];
          const double ok = tracer.certifiedDistance(F.cell, X, inward, want, P.rayMaxCrossings, P.rayTol);
          const double chosen = std::min(want, 0.90 * ok);
          require(chosen >= 0.25 * want && chosen < want,
                  "wall sample ray cannot be shortened safely (face " + std::to_string(fi) + ", point " +
                      std::to_string(q) + ", requested " + std::to_string(want) + ", certified " +
                      std::to_string(ok) + "): " + e.what());
          dist[0] = chosen;
          hits = tracer.trace(F.cell, X, inward, dist, P.rayMaxCrossings, P.rayTol);
          ++T.shortened;
          T.minKeptFraction = std::min(T.minKeptFraction, chosen / want);
        }
        for (std::size_t k = 0; k < slots.size(); ++k) {
          const RayHit& h = hits[k];
          place(T.samples[((std::size_t)fi * 12 + q) * 5 + slots[k]], h.cell, h.l, dist[k]);
          ++T.placed;
          if (h.cell != F.cell) ++T.crossed;
          T.maxCrossings = std::max(T.maxCrossings, (int)h.crossings);
        }
        for (int m = 0; m < 5; ++m)
          if (!(mask & (1u << m))) {
            const RayHit& h = hits[0];
            place(T.samples[((std::size_t)fi * 12 + q) * 5 + m], h.cell, h.l, T.heights[(std::size_t)m]);
          }
      }
    }
    T.rayMask = mask;
  }

  // Blocks and their cells.
  const int nb = nc + T.nInlet + T.nWallQ;
  T.blockCell.resize((std::size_t)nb);
  for (int c = 0; c < nc; ++c) T.blockCell[(std::size_t)c] = c;
  for (int i = 0; i < T.nInlet; ++i) T.blockCell[(std::size_t)nc + i] = T.inlet[(std::size_t)i].cell;
  require(P.omegaMode >= 0 && P.omegaMode <= 2, "omega wall mode must be of_auto, sample_log or trace_log");
  require(P.omegaMode != 2 || (P.traceGamma >= 10.0 && P.traceGamma <= 1e5),
          "trace_log trace_gamma must be in [10,100000]");
  for (int w = 0; w < T.nWallQ; ++w)
    T.blockCell[(std::size_t)nc + T.nInlet + w] =
        P.omegaMode == 2 ? T.wall[(std::size_t)w / 12].cell : T.samples[(std::size_t)w * 5 + 2].cell;

  // Vertex rows: sorted vertex columns, then sorted bubble columns.
  std::vector<std::vector<int>> vc((std::size_t)np), bc((std::size_t)np);
  for (int c = 0; c < nc; ++c)
    for (int a = 0; a < 4; ++a) {
      const int i = T.cellEnt[(std::size_t)c * 8 + a];
      for (int b = 0; b < 8; ++b) {
        const int j = T.cellEnt[(std::size_t)c * 8 + b];
        (j < np ? vc : bc)[(std::size_t)i].push_back(j);
      }
    }
  T.row.assign((std::size_t)np + 1, 0);
  T.entOff.assign((std::size_t)np + 1, 0);
  std::vector<int> entCol;
  for (int i = 0; i < np; ++i) {
    auto& v = vc[(std::size_t)i];
    auto& b = bc[(std::size_t)i];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    std::sort(b.begin(), b.end());
    b.erase(std::unique(b.begin(), b.end()), b.end());
    require(!v.empty(), "vertex " + std::to_string(i) + " belongs to no cell");
    T.row[(std::size_t)i + 1] = T.row[(std::size_t)i] + (int)v.size();
    T.entOff[(std::size_t)i + 1] = T.entOff[(std::size_t)i] + (int)(v.size() + b.size());
    for (int j : v) T.col.push_back(j);
    entCol.insert(entCol.end(), v.begin(), v.end());
    entCol.insert(entCol.end(), b.begin(), b.end());
  }
  T.diagPos.resize((std::size_t)np);
  for (int i = 0; i < np; ++i) {
    const auto b = T.col.begin() + T.row[(std::size_t)i], e = T.col.begin() + T.row[(std::size_t)i + 1];
    T.diagPos[(std::size_t)i] = (int)(std::lower_bound(b, e, i) - T.col.begin());
  }
  auto entryOf = [&](int i, int j) {
    const int e0 = T.entOff[(std::size_t)i], nvc = T.row[(std::size_t)i + 1] - T.row[(std::size_t)i];
    const int lo = j < np ? e0 : e0 + nvc, hi = j < np ? e0 + nvc : T.entOff[(std::size_t)i + 1];
    const auto it = std::lower_bound(entCol.begin() + lo, entCol.begin() + hi, j);
    return (int)(it - entCol.begin());
  };

  // Gather terms in ascending block order.
  const std::size_t nEnt = entCol.size();
  T.termOff.assign(nEnt + 1, 0);
  T.rhsOff.assign((std::size_t)np + 1, 0);
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int> fill(pass ? T.termOff.begin() : T.termOff.end(), T.termOff.end());
    std::vector<int> rfill(pass ? T.rhsOff.begin() : T.rhsOff.end(), T.rhsOff.end());
    if (pass) {
      T.term.resize((std::size_t)T.termOff[nEnt]);
      T.rhsTerm.resize((std::size_t)T.rhsOff[(std::size_t)np]);
    }
    for (int blk = 0; blk < nb; ++blk) {
      const int c = T.blockCell[(std::size_t)blk];
      for (int a = 0; a < 4; ++a) {
        const int i = T.cellEnt[(std::size_t)c * 8 + a];
        for (int b = 0; b < 8; ++b) {
          const int e = entryOf(i, T.cellEnt[(std::size_t)c * 8 + b]);
          if (pass)
            T.term[(std::size_t)fill[(std::size_t)e]++] = blk * 64 + 8 * a + b;
          else
            ++T.termOff[(std::size_t)e + 1];
        }
        if (pass)
          T.rhsTerm[(std::size_t)rfill[(std::size_t)i]++] = blk * 8 + a;
        else
          ++T.rhsOff[(std::size_t)i + 1];
      }
    }
    if (!pass) {
      for (std::size_t e = 0; e < nEnt; ++e) T.termOff[e + 1] += T.termOff[e];
      for (int i = 0; i < np; ++i) T.rhsOff[(std::size_t)i + 1] += T.rhsOff[(std::size_t)i];
    }
  }
  T.coloring = greedyColoring32(np, T.row.data(), T.col.data());
  return T;
}

}  // namespace nodals
