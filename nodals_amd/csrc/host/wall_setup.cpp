// Host setup of the wall-function momentum closure (see wall_setup.hpp).
#include "wall_setup.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <string>

#include "error.hpp"

namespace nodals {

const double kGauss4[4][2] = {{-0.8611363115940525752, 0.3478548451374538574},
                              {-0.3399810435848562648, 0.6521451548625461426},
                              {0.3399810435848562648, 0.6521451548625461426},
                              {0.8611363115940525752, 0.3478548451374538574}};

namespace {

void basis8(const double l[4], double phi[8]) {
  for (int i = 0; i < 4; ++i) phi[i] = l[i];
  for (int i = 0; i < 4; ++i) {
    double p = 27.0;
    for (int j = 0; j < 4; ++j)
      if (j != i) p *= l[j];
    phi[4 + i] = p;
  }
}

// Solves [e1 e2 e3] x = r by Cramer's rule.
bool solveCols(const std::array<double, 3>& e1, const std::array<double, 3>& e2, const std::array<double, 3>& e3,
               const std::array<double, 3>& r, double x[3]) {
  auto det = [](const std::array<double, 3>& a, const std::array<double, 3>& b, const std::array<double, 3>& c) {
    return a[0] * (b[1] * c[2] - b[2] * c[1]) - b[0] * (a[1] * c[2] - a[2] * c[1]) + c[0] * (a[1] * b[2] - a[2] * b[1]);
  };
  const double d = det(e1, e2, e3);
  if (!(std::abs(d) > 0.0) || !std::isfinite(d)) return false;
  x[0] = det(r, e2, e3) / d;
  x[1] = det(e1, r, e3) / d;
  x[2] = det(e1, e2, r) / d;
  return std::isfinite(x[0]) && std::isfinite(x[1]) && std::isfinite(x[2]);
}

}  // namespace

WallTopo buildWallTopo(const TetMesh& M, const H8Setup& S, const TurbTopo& T) {
  require(S.wallFunction, "the wall closure needs a wall-function duct (--wall-model spalding)");
  require(!T.wall.empty(), "the wall closure needs the wall faces of the k-omega setup");
  const int np = M.nPoints(), nc = M.nCells();
  const CsrD& A = S.mom;
  WallTopo W;
  W.nf = (int)T.wall.size();
  const int nf = W.nf;
  W.cell.resize((std::size_t)nf);
  W.lf.resize((std::size_t)nf);
  W.n.resize((std::size_t)nf * 3);
  W.area.resize((std::size_t)nf);
  W.h.resize((std::size_t)nf);
  W.qw.resize((std::size_t)nf * 12);
  W.ql.resize((std::size_t)nf * 48);
  W.bw.assign((std::size_t)nf * 96, 0.0);
  W.bp.assign((std::size_t)nf * 96, 0.0);
  W.ybar.resize((std::size_t)nf * 12);
  W.bub.resize((std::size_t)nf);
  W.vert.resize((std::size_t)nf * 3);
  for (int f = 0; f < nf; ++f) {
    const TurbFace& F = T.wall[(std::size_t)f];
    const int c = F.cell;
    W.cell[(std::size_t)f] = c;
    W.lf[(std::size_t)f] = F.lf;
    for (int d = 0; d < 3; ++d) W.n[(std::size_t)f * 3 + d] = F.n[d];
    W.area[(std::size_t)f] = F.area;
    const double vol = std::abs(S.cells[(std::size_t)c].det) / 6.0;
    W.h[(std::size_t)f] = 3.0 * vol / F.area;
    const auto& t = M.tets[(std::size_t)c];
    const std::array<double, 3> x0 = M.points[(std::size_t)t[0]];
    std::array<double, 3> e[3];
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) e[j][(std::size_t)d] = M.points[(std::size_t)t[(std::size_t)j + 1]][(std::size_t)d] - x0[(std::size_t)d];
    // barycentric direction of the inward normal -n (per unit length)
    double dl[3];
    require(solveCols(e[0], e[1], e[2], {-F.n[0], -F.n[1], -F.n[2]}, dl), "wall cell normal transform failed");
    const double dr[4] = {-dl[0] - dl[1] - dl[2], dl[0], dl[1], dl[2]};
    for (int q = 0; q < 12; ++q) {
      double lw[4] = {0, 0, 0, 0};
      for (int j = 0; j < 3; ++j) lw[kTetFace[F.lf][j]] = kTri12[q][j];
      const std::size_t iq = (std::size_t)f * 12 + q;
      for (int j = 0; j < 4; ++j) W.ql[iq * 4 + j] = lw[j];
      basis8(lw, &W.bw[iq * 8]);
      double exit = 1e300;
      for (int j = 0; j < 4; ++j)
        if (dr[j] < -1e-15) exit = std::min(exit, -lw[j] / dr[j]);
      require(exit > 0.0 && std::isfinite(exit) && exit < 1e100, "wall Picard normal column has no exit");
      W.ybar[iq] = 0.45 * exit;
      for (int m = 0; m < 4; ++m) {
        const double s = 0.45 * exit * (1.0 + kGauss4[m][0]);
        double lm[4], pb[8];
        for (int j = 0; j < 4; ++j) lm[j] = lw[j] + s * dr[j];
        basis8(lm, pb);
        for (int j = 0; j < 8; ++j) W.bp[iq * 8 + j] += 0.5 * kGauss4[m][1] * pb[j];
      }
      W.qw[iq] = F.area * kTri12[q][3];
    }
    const int bubE = np + M.oppFace[(std::size_t)c][(std::size_t)F.lf];
    W.bub[(std::size_t)f] = S.g2free[(std::size_t)bubE];
    require(W.bub[(std::size_t)f] >= 0, "a wall-function wall face bubble is fixed");
    for (int j = 0; j < 3; ++j) {
      const int v = t[(std::size_t)kTetFace[F.lf][j]];
      W.vert[(std::size_t)f * 3 + j] = S.g2free[(std::size_t)v];
      require(W.vert[(std::size_t)f * 3 + j] >= 0, "a wall-function wall vertex is fixed");
    }
  }

  // Wall DOFs with their faces.
  {
    std::map<int, std::vector<int>> faces;
    std::map<int, bool> isBub;
    for (int f = 0; f < nf; ++f) {
      faces[W.bub[(std::size_t)f]].push_back(f);
      isBub[W.bub[(std::size_t)f]] = true;
      for (int j = 0; j < 3; ++j) {
        faces[W.vert[(std::size_t)f * 3 + j]].push_back(f);
        isBub.emplace(W.vert[(std::size_t)f * 3 + j], false);
      }
    }
    W.dofOff.push_back(0);
    for (auto& kv : faces) {
      require(!(isBub[kv.first] && kv.second.size() != 1), "a wall face bubble belongs to more than one wall face",
              kErrInternal);
      W.dof.push_back(kv.first);
      W.dofIsBubble.push_back(isBub[kv.first] ? 1 : 0);
      for (int f : kv.second) W.dofFace.push_back(f);
      W.dofOff.push_back((int)W.dofFace.size());
    }
  }
  // Wall faces of each cell.
  W.cellWallOff.assign((std::size_t)nc + 1, 0);
  for (int f = 0; f < nf; ++f) ++W.cellWallOff[(std::size_t)W.cell[(std::size_t)f] + 1];
  for (int c = 0; c < nc; ++c) W.cellWallOff[(std::size_t)c + 1] += W.cellWallOff[(std::size_t)c];
  W.cellWall.resize((std::size_t)nf);
  {
    std::vector<int> fill(W.cellWallOff.begin(), W.cellWallOff.end() - 1);
    for (int f = 0; f < nf; ++f) W.cellWall[(std::size_t)fill[(std::size_t)W.cell[(std::size_t)f]]++] = f;
  }

  // Wall-block gather per momentum entry (faces ascending).
  const std::size_t nnz = A.col.size();
  W.blkOff.assign(nnz + 1, 0);
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int> fill;
    if (pass) {
      for (std::size_t k = 0; k < nnz; ++k) W.blkOff[k + 1] += W.blkOff[k];
      W.blkIdx.resize((std::size_t)W.blkOff[nnz]);
      fill.assign(W.blkOff.begin(), W.blkOff.end() - 1);
    }
    for (int f = 0; f < nf; ++f) {
      const CellPlan& cp = S.cells[(std::size_t)W.cell[(std::size_t)f]];
      for (int a = 0; a < 8; ++a) {
        if (cp.ref[a] < 0) continue;
        for (int b = 0; b < 8; ++b) {
          if (cp.ref[b] < 0) continue;  // wall blocks vanish off the face's four DOFs
          const std::size_t k = (std::size_t)(A.row[(std::size_t)cp.ref[a]] + cp.rowSlot[8 * a + b]);
          if (pass) W.blkIdx[(std::size_t)fill[k]++] = f * 64 + 8 * a + b;
          else ++W.blkOff[k + 1];
        }
      }
    }
  }

  // M3 transforms.
  auto pos = [&](int i, int j) -> int {
    const auto b = A.col.begin() + A.row[(std::size_t)i], e = A.col.begin() + A.row[(std::size_t)i + 1];
    const auto it = std::lower_bound(b, e, j);
    require(it != e && *it == j, "momentum pattern misses a BF3 wall coupling", kErrInternal);
    return (int)(it - A.col.begin());
  };
  auto buildTargets = [](const std::vector<std::pair<int, int>>& pairs, std::vector<std::int32_t>& tgtPos,
                         std::vector<std::int32_t>& tgtOff, std::vector<std::int32_t>& tgtSrc) {
    // pairs: (target position, src*2 + kind); group by target, sources ascending
    std::vector<std::pair<int, int>> p = pairs;
    std::sort(p.begin(), p.end());
    tgtOff.push_back(0);
    for (std::size_t a = 0; a < p.size();) {
      std::size_t b = a;
      while (b < p.size() && p[b].first == p[a].first) {
        tgtSrc.push_back(p[b].second);
        ++b;
      }
      tgtPos.push_back(p[a].first);
      tgtOff.push_back((int)tgtSrc.size());
      a = b;
    }
  };
  std::vector<std::pair<int, int>> rp, lp;
  for (int f = 0; f < nf; ++f) {
    const int b = W.bub[(std::size_t)f];
    for (std::int64_t kk = A.row[(std::size_t)b]; kk < A.row[(std::size_t)b + 1]; ++kk) {
      const int i = A.col[(std::size_t)kk];
      // right: column b of row i
      const int src = (int)W.rSrcPos.size();
      const int kb = pos(i, b);
      W.rSrcFace.push_back(f);
      W.rSrcPos.push_back(kb);
      rp.push_back({kb, 2 * src});
      for (int a = 0; a < 3; ++a) rp.push_back({pos(i, W.vert[(std::size_t)f * 3 + a]), 2 * src + 1});
      // left: row b, column i (= j)
      const int lsrc = (int)W.lSrcPos.size();
      W.lSrcFace.push_back(f);
      W.lSrcPos.push_back((int)kk);
      lp.push_back({(int)kk, 2 * lsrc});
      for (int a = 0; a < 3; ++a) lp.push_back({pos(W.vert[(std::size_t)f * 3 + a], i), 2 * lsrc + 1});
    }
  }
  buildTargets(rp, W.rTgtPos, W.rTgtOff, W.rTgtSrc);
  buildTargets(lp, W.lTgtPos, W.lTgtOff, W.lTgtSrc);
  return W;
}

}  // namespace nodals
