// Potential-flow start of the wall-function path (see potential_start.hpp).
#include "potential_start.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <string>

#include "error.hpp"

namespace nodals {

namespace {

using V3 = std::array<double, 3>;

V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

struct BFace {
  int f, kind;  // kind 0 wall, 1 inlet, 2 outlet
  V3 n;         // outward unit normal
  double area;
};

}  // namespace

PotentialStart potentialStart(const TetMesh& M, const H8Setup& S, double tol, int maxIterations) {
  const int np = M.nPoints(), nc = M.nCells(), nfaces = M.nFaces();
  require(np > 0 && nc > 0, "potential start: empty mesh");
  require(S.pipe.inlet >= 0 && S.pipe.outlet >= 0, "potential start: the mesh has no inlet or no outlet patch");
  const V3 uin{S.pipe.inletVelocity[0], S.pipe.inletVelocity[1], S.pipe.inletVelocity[2]};

  // P1 gradients and volumes per tetrahedron
  std::vector<std::array<V3, 4>> grad((std::size_t)nc);
  std::vector<double> vol((std::size_t)nc);
  for (int c = 0; c < nc; ++c) {
    const auto& t = M.tets[(std::size_t)c];
    const V3 p0 = M.points[(std::size_t)t[0]], a = sub(M.points[(std::size_t)t[1]], p0),
             b = sub(M.points[(std::size_t)t[2]], p0), cc = sub(M.points[(std::size_t)t[3]], p0);
    const double det = dot(a, cross(b, cc));
    require(std::isfinite(det) && std::abs(det) >= 1e-300, "potential start: degenerate tetrahedron");
    vol[(std::size_t)c] = std::abs(det) / 6.0;
    const V3 g1 = cross(b, cc), g2 = cross(cc, a), g3 = cross(a, b);
    auto& G = grad[(std::size_t)c];
    for (int j = 0; j < 3; ++j) {
      G[1][(std::size_t)j] = g1[(std::size_t)j] / det;
      G[2][(std::size_t)j] = g2[(std::size_t)j] / det;
      G[3][(std::size_t)j] = g3[(std::size_t)j] / det;
      G[0][(std::size_t)j] = -G[1][(std::size_t)j] - G[2][(std::size_t)j] - G[3][(std::size_t)j];
    }
  }

  // boundary faces with their kind, outward normal and area
  std::vector<BFace> bf;
  for (int f = M.nInternal(); f < nfaces; ++f) {
    const int patch = M.facePatch[(std::size_t)f];
    if (patch < 0) continue;
    const int c = M.owner[(std::size_t)f];
    int lf = -1;
    for (int i = 0; i < 4; ++i)
      if (M.oppFace[(std::size_t)c][(std::size_t)i] == f) lf = i;
    require(lf >= 0, "potential start: boundary face not on its owner cell");
    const auto& fv = M.faces[(std::size_t)f];
    const V3 p0 = M.points[(std::size_t)fv[0]];
    V3 n = cross(sub(M.points[(std::size_t)fv[1]], p0), sub(M.points[(std::size_t)fv[2]], p0));
    const double ln = std::sqrt(dot(n, n));
    require(ln > 0.0, "potential start: degenerate boundary face");
    const V3 opp = M.points[(std::size_t)M.tets[(std::size_t)c][(std::size_t)lf]];
    const double s = dot(n, sub(p0, opp)) > 0.0 ? 1.0 : -1.0;
    for (double& x : n) x *= s / ln;
    bf.push_back({f, patch == S.pipe.inlet ? 1 : (patch == S.pipe.outlet ? 2 : 0), n, 0.5 * ln});
  }

  // Laplace rows: entries summed in element order per (row, column)
  std::vector<std::uint8_t> outlet((std::size_t)np, 0);
  std::vector<double> rhs((std::size_t)np, 0.0);
  for (const BFace& b : bf) {
    const auto& fv = M.faces[(std::size_t)b.f];
    if (b.kind == 2)
      for (int k = 0; k < 3; ++k) outlet[(std::size_t)fv[(std::size_t)k]] = 1;
    if (b.kind == 1) {
      const double q = dot(uin, b.n);
      for (int k = 0; k < 3; ++k) rhs[(std::size_t)fv[(std::size_t)k]] += q * b.area / 3.0;
    }
  }
  struct Trip {
    std::int32_t r, c;
    double v;
  };
  std::vector<Trip> tr;
  tr.reserve((std::size_t)nc * 16);
  for (int c = 0; c < nc; ++c) {
    const auto& t = M.tets[(std::size_t)c];
    const auto& G = grad[(std::size_t)c];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) tr.push_back({t[(std::size_t)a], t[(std::size_t)b], vol[(std::size_t)c] * dot(G[(std::size_t)a], G[(std::size_t)b])});
  }
  std::stable_sort(tr.begin(), tr.end(), [](const Trip& x, const Trip& y) { return x.r != y.r ? x.r < y.r : x.c < y.c; });
  std::vector<std::int32_t> row((std::size_t)np + 1, 0), col;
  std::vector<double> val, invDiag((std::size_t)np, 0.0);
  int nOutlet = 0;
  {
    std::size_t t = 0;
    for (int i = 0; i < np; ++i) {
      if (outlet[(std::size_t)i]) {
        ++nOutlet;
        while (t < tr.size() && tr[t].r == i) ++t;
        col.push_back(i);
        val.push_back(1.0);
        invDiag[(std::size_t)i] = 1.0;
        rhs[(std::size_t)i] = 0.0;
        row[(std::size_t)i + 1] = (std::int32_t)col.size();
        continue;
      }
      while (t < tr.size() && tr[t].r == i) {
        const int j = tr[t].c;
        double s = 0.0;
        for (; t < tr.size() && tr[t].r == i && tr[t].c == j; ++t) s += tr[t].v;
        if (outlet[(std::size_t)j]) continue;
        col.push_back(j);
        val.push_back(s);
        if (j == i) invDiag[(std::size_t)i] = 1.0 / s;
      }
      require(invDiag[(std::size_t)i] > 0.0 && std::isfinite(invDiag[(std::size_t)i]),
              "potential start: non-positive Laplace diagonal");
      row[(std::size_t)i + 1] = (std::int32_t)col.size();
    }
  }
  require(nOutlet > 0, "potential start: no outlet vertices");

  auto matvec = [&](const std::vector<double>& x, std::vector<double>& y) {
    for (int i = 0; i < np; ++i) {
      double s = 0.0;
      for (int k = row[(std::size_t)i]; k < row[(std::size_t)i + 1]; ++k) s += val[(std::size_t)k] * x[(std::size_t)col[(std::size_t)k]];
      y[(std::size_t)i] = s;
    }
  };
  auto dotp = [&](const std::vector<double>& x, const std::vector<double>& y) {
    long double s = 0;
    for (int i = 0; i < np; ++i) s += (long double)x[(std::size_t)i] * y[(std::size_t)i];
    return (double)s;
  };
  PotentialStart out;
  std::vector<double> phi((std::size_t)np, 0.0), r = rhs, z((std::size_t)np), p((std::size_t)np), Ap((std::size_t)np);
  const double bnorm = std::sqrt(dotp(rhs, rhs));
  require(bnorm > 0.0 && std::isfinite(bnorm), "potential start: zero inlet flux");
  for (int i = 0; i < np; ++i) z[(std::size_t)i] = invDiag[(std::size_t)i] * r[(std::size_t)i];
  p = z;
  double rho = dotp(r, z);
  bool converged = false;
  for (int it = 0; it < maxIterations; ++it) {
    matvec(p, Ap);
    const double pq = dotp(p, Ap);
    require(pq > 0.0 && std::isfinite(pq), "potential start: CG breakdown (p.Ap not positive)");
    const double alpha = rho / pq;
    for (int i = 0; i < np; ++i) {
      phi[(std::size_t)i] += alpha * p[(std::size_t)i];
      r[(std::size_t)i] -= alpha * Ap[(std::size_t)i];
    }
    out.iterations = it + 1;
    out.rel = std::sqrt(dotp(r, r)) / bnorm;
    if (out.rel < tol) {
      converged = true;
      break;
    }
    for (int i = 0; i < np; ++i) z[(std::size_t)i] = invDiag[(std::size_t)i] * r[(std::size_t)i];
    const double next = dotp(r, z);
    require(next > 0.0 && std::isfinite(next), "potential start: CG breakdown");
    for (int i = 0; i < np; ++i) p[(std::size_t)i] = z[(std::size_t)i] + (next / rho) * p[(std::size_t)i];
    rho = next;
  }
  require(converged, "potential start: Laplace solve not converged to " + std::to_string(tol) + " in " +
                         std::to_string(maxIterations) + " iterations (rel " + std::to_string(out.rel) + ")");

  // lumped-mass projection of the cellwise gradient onto the vertices
  std::vector<V3> u((std::size_t)np, V3{0, 0, 0});
  std::vector<double> mass((std::size_t)np, 0.0);
  for (int c = 0; c < nc; ++c) {
    const auto& t = M.tets[(std::size_t)c];
    V3 g{0, 0, 0};
    for (int a = 0; a < 4; ++a)
      for (int k = 0; k < 3; ++k) g[(std::size_t)k] += phi[(std::size_t)t[(std::size_t)a]] * grad[(std::size_t)c][(std::size_t)a][(std::size_t)k];
    for (int a = 0; a < 4; ++a) {
      const int v = t[(std::size_t)a];
      mass[(std::size_t)v] += vol[(std::size_t)c] / 4.0;
      for (int k = 0; k < 3; ++k) u[(std::size_t)v][(std::size_t)k] += g[(std::size_t)k] * vol[(std::size_t)c] / 4.0;
    }
  }
  out.value.assign((std::size_t)(np + nfaces) * 3, 0.0);
  for (int i = 0; i < np; ++i) {
    require(mass[(std::size_t)i] > 0.0, "potential start: vertex in no cell");
    for (int k = 0; k < 3; ++k) {
      u[(std::size_t)i][(std::size_t)k] /= mass[(std::size_t)i];
      out.value[3 * (std::size_t)i + (std::size_t)k] = u[(std::size_t)i][(std::size_t)k];
    }
    out.maxSpeed = std::max(out.maxSpeed, std::sqrt(dot(u[(std::size_t)i], u[(std::size_t)i])));
  }

  // face bubbles: the BF3 bubble's face mean is 27/60 = 0.45 and its trace
  // vanishes on the other faces, so one normal bubble value sets a face's
  // mean normal velocity
  constexpr double kBubbleMean = 0.45;
  auto faceMean = [&](int f) {
    V3 m{0, 0, 0};
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 3; ++d) m[(std::size_t)d] += u[(std::size_t)M.faces[(std::size_t)f][(std::size_t)k]][(std::size_t)d] / 3.0;
    return m;
  };
  double outletProj = 0.0, outletArea = 0.0, inletPrescribed = 0.0;
  for (const BFace& b : bf) {
    if (b.kind == 2) {
      outletProj += b.area * dot(faceMean(b.f), b.n);
      outletArea += b.area;
    } else if (b.kind == 1) {
      inletPrescribed += b.area * dot(uin, b.n);
    }
  }
  require(outletArea > 0.0, "potential start: no outlet area");
  const double outletCorr = (-inletPrescribed - outletProj) / outletArea;
  for (const BFace& b : bf) {
    const double before = dot(faceMean(b.f), b.n);
    const double goal = b.kind == 1 ? dot(uin, b.n) : (b.kind == 2 ? before + outletCorr : 0.0);
    const double corr = (goal - before) / kBubbleMean;
    for (int d = 0; d < 3; ++d) out.value[3 * ((std::size_t)np + (std::size_t)b.f) + (std::size_t)d] = corr * b.n[(std::size_t)d];
    const double after = before + kBubbleMean * corr;
    if (b.kind == 1) out.inletFlux += after * b.area;
    else if (b.kind == 2) out.outletFlux += after * b.area;
    else out.wallFluxMax = std::max(out.wallFluxMax, std::abs(after));
  }
  out.imbalanceRel = std::abs(out.inletFlux + out.outletFlux) / std::max(1e-300, std::abs(out.inletFlux));
  require(out.wallFluxMax < 1e-10 * std::max(out.maxSpeed, 1.0) && out.inletFlux < 0.0 && out.imbalanceRel < 1e-10,
          "potential start: projected wall/inlet face-mean flux audit failed");
  return out;
}

}  // namespace nodals
