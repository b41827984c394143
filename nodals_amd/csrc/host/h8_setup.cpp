// H8 setup.  See h8_setup.hpp.  Compile with -ffp-contract=off: the
// floating-point setup values follow the reference operation order so that
// the rAU snapshot, and through it the AMG hierarchy, are reproducible.
#include "h8_setup.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <limits>
#include <mutex>
#include <thread>

#include "error.hpp"

namespace nodals {

int hostThreads(int requested) {
  int n = requested;
  if (n <= 0) {
    const unsigned h = std::thread::hardware_concurrency();
    n = h ? (int)std::min(h, 16u) : 1;
  }
  return std::max(1, std::min(n, 64));
}

namespace {

// Work-stealing loop over [0,n) in chunks.  Every output index is written by
// exactly one worker, so results do not depend on the thread count.
template <class F>
void parallelFor(int n, int nth, int chunk, F&& fn) {
  if (n <= 0) return;
  nth = std::max(1, std::min(nth, (n + chunk - 1) / chunk));
  if (nth == 1) {
    fn(0, n);
    return;
  }
  std::atomic<int> next{0};
  std::exception_ptr err;
  std::mutex em;
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; ++t) {
    pool.emplace_back([&]() {
      try {
        for (;;) {
          const int b = next.fetch_add(chunk);
          if (b >= n) break;
          fn(b, std::min(n, b + chunk));
        }
      } catch (...) {
        std::lock_guard<std::mutex> g(em);
        if (!err) err = std::current_exception();
        next.store(n);
      }
    });
  }
  for (auto& th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

using V3 = std::array<double, 3>;
V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double norm(const V3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// Boundary face area vector pointing out of its owner cell (g2_host.hpp:30-41).
V3 outwardAreaVector(const TetMesh& M, int f) {
  const auto& F = M.faces[(std::size_t)f];
  const V3& a = M.points[(std::size_t)F[0]];
  const V3& b = M.points[(std::size_t)F[1]];
  const V3& c = M.points[(std::size_t)F[2]];
  const V3 u = sub(b, a), v = sub(c, a);
  V3 sf = {0.5 * (u[1] * v[2] - u[2] * v[1]), 0.5 * (u[2] * v[0] - u[0] * v[2]),
           0.5 * (u[0] * v[1] - u[1] * v[0])};
  const int cell = M.owner[(std::size_t)f];
  const V3 fc = {(a[0] + b[0] + c[0]) / 3.0, (a[1] + b[1] + c[1]) / 3.0, (a[2] + b[2] + c[2]) / 3.0};
  V3 cc = {0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    const V3& x = M.points[(std::size_t)M.tets[(std::size_t)cell][(std::size_t)i]];
    cc[0] += 0.25 * x[0];
    cc[1] += 0.25 * x[1];
    cc[2] += 0.25 * x[2];
  }
  const double dot = sf[0] * (fc[0] - cc[0]) + sf[1] * (fc[1] - cc[1]) + sf[2] * (fc[2] - cc[2]);
  if (dot < 0.0) sf = {-sf[0], -sf[1], -sf[2]};
  return sf;
}

// Poiseuille profile 2U(1 - r^2/R^2) and its exact triangle average
// (g4_host.hpp:27-33).
double idealUz(const PipeInfo& P, double x, double y) {
  const double rr = (x - P.cx) * (x - P.cx) + (y - P.cy) * (y - P.cy);
  return 2.0 * P.bulk * (1.0 - rr / (P.R * P.R));
}
double triangleAvgUz(const PipeInfo& P, const V3 X[3]) {
  double xx[3], yy[3];
  for (int i = 0; i < 3; ++i) {
    xx[i] = X[i][0] - P.cx;
    yy[i] = X[i][1] - P.cy;
  }
  const double ax = (xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2] + xx[0] * xx[1] + xx[1] * xx[2] +
                     xx[2] * xx[0]) / 6.0;
  const double ay = (yy[0] * yy[0] + yy[1] * yy[1] + yy[2] * yy[2] + yy[0] * yy[1] + yy[1] * yy[2] +
                     yy[2] * yy[0]) / 6.0;
  return 2.0 * P.bulk * (1.0 - (ax + ay) / (P.R * P.R));
}

PipeInfo makeCavity(const TetMesh& M, const H8Params& prm) {
  PipeInfo P;
  P.wall = M.patchIndex(prm.wall);
  P.inlet = M.patchIndex(prm.lid);  // the lid carries the driving velocity
  P.outlet = -1;
  P.re = prm.re;
  P.bulk = prm.bulk;
  require(P.inlet >= 0, "lid patch not found (lid=" + prm.lid + ")");
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (const auto& x : M.points)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], x[(std::size_t)d]);
      hi[d] = std::max(hi[d], x[(std::size_t)d]);
    }
  P.L = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  P.zIn = lo[2];
  P.zOut = hi[2];
  require(P.L > 0 && prm.re > 0 && prm.bulk > 0, "invalid cavity geometry");
  P.nu = prm.bulk * P.L / prm.re;
  P.inletVelocity[0] = prm.bulk;
  P.inletVelocity[1] = 0.0;
  P.inletVelocity[2] = 0.0;
  return P;
}

// General duct: hydraulic diameter of the inlet, uniform inflow along the
// inward inlet normal.
PipeInfo makeDuct(const TetMesh& M, const H8Params& prm) {
  PipeInfo P;
  P.wall = M.patchIndex(prm.wall);
  P.inlet = M.patchIndex(prm.inlet);
  P.outlet = M.patchIndex(prm.outlet);
  P.re = prm.re;
  P.bulk = prm.bulk;
  require(P.wall >= 0 && P.inlet >= 0 && P.outlet >= 0,
          "duct patch not found (wall=" + prm.wall + " inlet=" + prm.inlet + " outlet=" + prm.outlet + ")");
  require(prm.re > 0 && prm.bulk > 0, "invalid duct flow parameters");
  const Patch& pin = M.patches[(std::size_t)P.inlet];
  V3 sfSum = {0, 0, 0}, cen = {0, 0, 0};
  for (int f = pin.start; f < pin.start + pin.count; ++f) {
    const auto& F = M.faces[(std::size_t)f];
    const V3 X[3] = {M.points[(std::size_t)F[0]], M.points[(std::size_t)F[1]], M.points[(std::size_t)F[2]]};
    const double a = 0.5 * norm(cross(sub(X[1], X[0]), sub(X[2], X[0])));
    P.inletArea += a;
    for (int d = 0; d < 3; ++d) cen[(std::size_t)d] += a * (X[0][(std::size_t)d] + X[1][(std::size_t)d] + X[2][(std::size_t)d]) / 3.0;
    const V3 sf = outwardAreaVector(M, f);
    for (int d = 0; d < 3; ++d) sfSum[(std::size_t)d] += sf[(std::size_t)d];
  }
  P.inletProjectedArea = norm(sfSum);
  require(P.inletArea > 0.0 && P.inletProjectedArea > 0.0, "degenerate duct inlet");
  for (int d = 0; d < 3; ++d) {
    P.inletNormal[d] = sfSum[(std::size_t)d] / P.inletProjectedArea;
    P.inletVelocity[d] = -prm.bulk * P.inletNormal[d];
  }
  if (prm.inletVelocityGiven)
    for (int d = 0; d < 3; ++d) P.inletVelocity[d] = prm.inletVelocity[d];
  P.cx = cen[0] / P.inletArea;
  P.cy = cen[1] / P.inletArea;
  P.D = prm.ductDiameter > 0.0 ? prm.ductDiameter : 2.0 * std::sqrt(P.inletArea / 3.14159265358979323846);
  P.R = 0.5 * P.D;
  P.nu = prm.bulk * P.D / prm.re;
  P.circleArea = P.inletArea;
  P.profileScale = 1.0;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (const auto& x : M.points)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], x[(std::size_t)d]);
      hi[d] = std::max(hi[d], x[(std::size_t)d]);
    }
  P.zIn = lo[2];
  P.zOut = hi[2];
  P.L = hi[2] - lo[2];
  P.hpGradient = 0.0;
  P.hpDrop = 0.0;  // no closed form
  return P;
}

PipeInfo makePipe(const TetMesh& M, const H8Params& prm) {
  PipeInfo P;
  P.wall = M.patchIndex(prm.wall);
  P.inlet = M.patchIndex(prm.inlet);
  P.outlet = M.patchIndex(prm.outlet);
  P.re = prm.re;
  P.bulk = prm.bulk;
  require(P.wall >= 0 && P.inlet >= 0 && P.outlet >= 0,
          "pipe patch not found (wall=" + prm.wall + " inlet=" + prm.inlet + " outlet=" + prm.outlet + ")");
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (const auto& x : M.points)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], x[(std::size_t)d]);
      hi[d] = std::max(hi[d], x[(std::size_t)d]);
    }
  P.cx = .5 * (lo[0] + hi[0]);
  P.cy = .5 * (lo[1] + hi[1]);
  P.zIn = lo[2];
  P.zOut = hi[2];
  P.L = hi[2] - lo[2];
  P.R = .5 * std::max(hi[0] - lo[0], hi[1] - lo[1]);
  P.D = 2 * P.R;
  require(P.R > 0 && P.L > 0 && prm.re > 0 && prm.bulk > 0, "invalid pipe geometry");
  P.nu = prm.bulk * P.D / prm.re;
  P.circleArea = 3.14159265358979323846 * P.R * P.R;
  const Patch& pin = M.patches[(std::size_t)P.inlet];
  double raw = 0;
  V3 sfSum = {0, 0, 0};
  for (int f = pin.start; f < pin.start + pin.count; ++f) {
    const auto& F = M.faces[(std::size_t)f];
    const V3 X[3] = {M.points[(std::size_t)F[0]], M.points[(std::size_t)F[1]], M.points[(std::size_t)F[2]]};
    const double a = 0.5 * norm(cross(sub(X[1], X[0]), sub(X[2], X[0])));
    P.inletArea += a;
    raw += a * triangleAvgUz(P, X);
    const V3 sf = outwardAreaVector(M, f);
    sfSum[0] += sf[0];
    sfSum[1] += sf[1];
    sfSum[2] += sf[2];
  }
  P.inletProjectedArea = norm(sfSum);
  require(P.inletProjectedArea > 0.0, "degenerate inlet projected area");
  for (int d = 0; d < 3; ++d) {
    P.inletNormal[d] = sfSum[(std::size_t)d] / P.inletProjectedArea;
    P.inletVelocity[d] = -prm.bulk * P.inletNormal[d];
  }
  require(raw > 0, "non-positive inlet flux");
  P.profileScale = prm.bulk * P.inletArea / raw;
  P.hpGradient = 32.0 * P.nu * prm.bulk / (P.D * P.D);
  P.hpDrop = P.hpGradient * P.L;
  return P;
}

// Collapsed (Duffy) 5x5x5 Gauss-Jacobi rule on the reference tetrahedron
// (g3_host.hpp:58-71).  Nodes and weights are the published rule.
struct QPoint {
  double lam[4];
  double w;
};
std::vector<QPoint> duffy125() {
  const double rn[5] = {0.034578939918215090, 0.17348032077169567, 0.38988638706551931,
                        0.63433347263088680, 0.85105421294701644};
  const double rw[5] = {0.081764784285771011, 0.12619896189991137, 0.089200161221590066,
                        0.032055600722961895, 0.0041138252030990035};
  const double sn[5] = {0.039809857051468722, 0.19801341787360821, 0.43797481024738616,
                        0.69546427335363614, 0.90146491420117358};
  const double sw[5] = {0.096781590226651476, 0.16717463809436969, 0.14638698708466985,
                        0.073908870072616678, 0.015747914521692299};
  const double tn[5] = {0.046910077030668018, 0.23076534494715845, 0.50000000000000000,
                        0.76923465505284150, 0.95308992296933193};
  const double tw[5] = {0.11846344252809449, 0.23931433524968326, 0.28444444444444450,
                        0.23931433524968326, 0.11846344252809449};
  std::vector<QPoint> q;
  q.reserve(125);
  for (int ir = 0; ir < 5; ++ir)
    for (int is = 0; is < 5; ++is)
      for (int it = 0; it < 5; ++it) {
        const double r = rn[ir], s = sn[is], t = tn[it];
        const double omr = 1 - r, oms = 1 - s;
        QPoint p;
        p.lam[0] = omr * oms * (1 - t);
        p.lam[1] = r;
        p.lam[2] = omr * s;
        p.lam[3] = omr * oms * t;
        p.w = rw[ir] * sw[is] * tw[it];
        q.push_back(p);
      }
  return q;
}

// Values and reference gradients of the 8 P1+BF3 functions: lambda_i and the
// face bubbles 27*lambda_j*lambda_k*lambda_l (face opposite vertex i).
void basisAt(const double lam[4], double val[8], double gr[8][3]) {
  static const double gl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 4; ++i) {
    val[i] = lam[i];
    for (int d = 0; d < 3; ++d) gr[i][d] = gl[i][d];
  }
  for (int i = 0; i < 4; ++i) {
    int js[3], k = 0;
    for (int j = 0; j < 4; ++j)
      if (j != i) js[k++] = j;
    val[4 + i] = 27 * lam[js[0]] * lam[js[1]] * lam[js[2]];
    for (int d = 0; d < 3; ++d) gr[4 + i][d] = 0.0;
    for (int a = 0; a < 3; ++a) {
      const int j = js[a], o1 = js[(a + 1) % 3], o2 = js[(a + 2) % 3];
      for (int d = 0; d < 3; ++d) gr[4 + i][d] += 27.0 * lam[o1] * lam[o2] * gl[j][d];
    }
  }
}

Tensors* makeTensors() {
  auto* T = new Tensors();
  std::fill(&T->diff[0][0][0][0], &T->diff[0][0][0][0] + 8 * 8 * 9, 0.0);
  std::fill(&T->cent[0][0][0][0], &T->cent[0][0][0][0] + 8 * 8 * 8 * 3, 0.0);
  for (const QPoint& q : duffy125()) {
    double val[8], gr[8][3];
    basisAt(q.lam, val, gr);
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) T->diff[a][b][j][k] += gr[a][j] * gr[b][k] * q.w;
    for (int a = 0; a < 8; ++a)
      for (int m = 0; m < 8; ++m)
        for (int b = 0; b < 8; ++b)
          for (int j = 0; j < 3; ++j) T->cent[a][m][b][j] += val[a] * val[m] * gr[b][j] * q.w;
  }
  return T;
}

void cellGeometry(const TetMesh& M, int c, double& det, double I[3][3], double* h2) {
  const auto& t = M.tets[(std::size_t)c];
  const V3 X[4] = {M.points[(std::size_t)t[0]], M.points[(std::size_t)t[1]], M.points[(std::size_t)t[2]],
                   M.points[(std::size_t)t[3]]};
  const double J[3][3] = {{X[1][0] - X[0][0], X[2][0] - X[0][0], X[3][0] - X[0][0]},
                          {X[1][1] - X[0][1], X[2][1] - X[0][1], X[3][1] - X[0][1]},
                          {X[1][2] - X[0][2], X[2][2] - X[0][2], X[3][2] - X[0][2]}};
  det = det3(J);
  require(det > 0.0, "non-positive tetrahedron orientation at cell " + std::to_string(c));
  inv3(J, I);
  if (h2) {
    double h = 0.0;
    for (int a = 0; a < 4; ++a)
      for (int b = a + 1; b < 4; ++b) {
        const double dx = X[a][0] - X[b][0], dy = X[a][1] - X[b][1], dz = X[a][2] - X[b][2];
        h = std::max(h, dx * dx + dy * dy + dz * dz);
      }
    *h2 = h;
  }
}

// Cell-local diffusion block nu*det*sum_{j,k,d} T[a][b][j][k] I[j][d] I[k][d]
// (g3_host.hpp:161-167).
void diffusionBlock(const Tensors& T, double nu, double det, const double I[3][3], double K[8][8]) {
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int d = 0; d < 3; ++d) s += T.diff[a][b][j][k] * I[j][d] * I[k][d];
      K[a][b] = nu * det * s;
    }
}

}  // namespace

const Tensors& referenceTensors() {
  static const Tensors* T = makeTensors();
  return *T;
}

double bcoefOf(const CellPlan& cp, int a, int d) {
  const double g1 = cp.invJ[d], g2 = cp.invJ[3 + d], g3 = cp.invJ[6 + d];
  const int i = (a < 4) ? a : a - 4;
  const double gl = i == 0 ? -(g1 + g2 + g3) : (i == 1 ? g1 : (i == 2 ? g2 : g3));
  const double base = (cp.det / 6.0) * gl;
  return (a < 4) ? base : -(27.0 / 20.0) * base;
}

Coloring greedyColoring(int n, const std::int64_t* row, const std::int32_t* col) {
  // Natural-order greedy: each row takes the smallest color unused by its
  // lower-indexed neighbours (g3_host.hpp:614-652).
  Coloring C;
  C.color.assign((std::size_t)n, -1);
  std::vector<int> mark(64, -1);
  int maxc = -1;
  for (int i = 0; i < n; ++i) {
    for (std::int64_t k = row[i]; k < row[i + 1]; ++k) {
      const int j = col[k];
      if (j >= i) continue;
      const int c = C.color[(std::size_t)j];
      if ((std::size_t)c >= mark.size()) mark.resize((std::size_t)c + 32, -1);
      mark[(std::size_t)c] = i;
    }
    int c = 0;
    while (c < (int)mark.size() && mark[(std::size_t)c] == i) ++c;
    if (c == (int)mark.size()) mark.resize(mark.size() + 32, -1);
    C.color[(std::size_t)i] = c;
    maxc = std::max(maxc, c);
  }
  C.ncolors = maxc + 1;
  C.offset.assign((std::size_t)C.ncolors + 1, 0);
  for (int c : C.color) ++C.offset[(std::size_t)c + 1];
  for (int c = 0; c < C.ncolors; ++c) C.offset[(std::size_t)c + 1] += C.offset[(std::size_t)c];
  C.rows.resize((std::size_t)n);
  std::vector<std::int32_t> next(C.offset.begin(), C.offset.end() - 1);
  for (int i = 0; i < n; ++i) C.rows[(std::size_t)next[(std::size_t)C.color[(std::size_t)i]]++] = i;
  for (int i = 0; i < n; ++i)
    for (std::int64_t k = row[i]; k < row[i + 1]; ++k) {
      const int j = col[k];
      if (j != i && C.color[(std::size_t)i] == C.color[(std::size_t)j]) fail("graph coloring conflict", kErrInternal);
    }
  return C;
}

Coloring greedyColoring32(int n, const std::int32_t* row, const std::int32_t* col) {
  std::vector<std::int64_t> r64((std::size_t)n + 1);
  for (int i = 0; i <= n; ++i) r64[(std::size_t)i] = row[i];
  return greedyColoring(n, r64.data(), col);
}

H8Setup buildH8Setup(const TetMesh& M, const H8Params& prm) {
  require(prm.alphaU > 0.0 && prm.alphaU <= 1.0, "alpha-u must be in (0,1]");
  const int nth = hostThreads(prm.threads);
  H8Setup S;
  S.cavity = !prm.lid.empty();
  S.duct = !S.cavity && prm.duct;
  S.pipe = S.cavity ? makeCavity(M, prm) : S.duct ? makeDuct(M, prm) : makePipe(M, prm);
  S.pinCell = S.cavity ? 0 : -1;
  const int np = M.nPoints(), nf = M.nFaces(), ni = M.nInternal(), nc = M.nCells();
  S.nc = nc;
  S.nPoints = np;
  S.nFaces = nf;

  // Dirichlet mask: every non-outlet boundary face and its vertices
  // (g3_host.hpp:29-55).
  S.fixed.assign((std::size_t)np + nf, 0);
  S.wallFunction = S.duct && prm.wallFunction;
  require(!prm.dgInlet || (!S.cavity && !S.duct), "the DG inlet is a straight-pipe option (not for --lid or --duct)");
  S.dgInlet = prm.dgInlet;
  for (int f = ni; f < nf; ++f) {
    if (M.facePatch[(std::size_t)f] == S.pipe.outlet) continue;
    // DG inlet: the inlet's entities stay free (the wall keeps the vertices
    // it shares with the inlet)
    if (S.dgInlet && M.facePatch[(std::size_t)f] == S.pipe.inlet) continue;
    // Wall-function walls are not Dirichlet: their vertex and face DOFs stay
    // free and carry the weak Nitsche/Spalding wall terms.
    if (S.wallFunction && M.facePatch[(std::size_t)f] == S.pipe.wall) continue;
    S.fixed[(std::size_t)np + f] = 1;
    for (auto v : M.faces[(std::size_t)f]) S.fixed[(std::size_t)v] = 1;
  }
  if (S.wallFunction)  // the wall keeps the vertices it shares with the inlet
    for (int f = ni; f < nf; ++f)
      if (M.facePatch[(std::size_t)f] == S.pipe.wall)
        for (auto v : M.faces[(std::size_t)f]) S.fixed[(std::size_t)v] = 0;
  S.g2free.assign(S.fixed.size(), -1);
  std::int32_t next = 0;
  for (int v = 0; v < np; ++v) {
    if (S.fixed[(std::size_t)v]) ++S.fixedVertices;
    else { S.g2free[(std::size_t)v] = next++; ++S.freeVertices; }
  }
  for (int f = 0; f < nf; ++f) {
    if (S.fixed[(std::size_t)np + f]) ++S.fixedFaces;
    else { S.g2free[(std::size_t)np + f] = next++; ++S.freeFaces; }
  }
  S.nv = next;
  require(S.nv > 0, "no free velocity DOFs");

  // Fixed slots and Dirichlet values: Poiseuille profile on inlet vertices
  // off the wall, face bubbles carry the (20/9) correction that makes the
  // P1+BF3 trace exact on average (g4_host.hpp:297-305).
  S.fixedSlot.assign(S.fixed.size(), -1);
  int ns = 0;
  for (std::size_t e = 0; e < S.fixed.size(); ++e)
    if (S.fixed[e]) S.fixedSlot[e] = ns++;
  S.nFixed = ns;
  S.fixedValue.assign((std::size_t)ns, {0.0, 0.0, 0.0});
  std::vector<std::uint8_t> onWall((std::size_t)np, 0), onInlet((std::size_t)np, 0);
  for (int f = ni; f < nf; ++f) {
    const int p = M.facePatch[(std::size_t)f];
    if (p == S.pipe.wall)
      for (auto v : M.faces[(std::size_t)f]) onWall[(std::size_t)v] = 1;
    if (p == S.pipe.inlet)
      for (auto v : M.faces[(std::size_t)f]) onInlet[(std::size_t)v] = 1;
  }
  if (S.cavity) {
    // Every lid vertex (its edges included) moves with the lid; each
    // boundary face bubble takes (20/9)(exact face average - vertex mean)
    // so that the P1+BF3 trace has the exact average on every face.
    for (int v = 0; v < np; ++v)
      if (onInlet[(std::size_t)v]) S.fixedValue[(std::size_t)S.fixedSlot[(std::size_t)v]][0] = S.pipe.inletVelocity[0];
    for (int f = ni; f < nf; ++f) {
      const auto& F = M.faces[(std::size_t)f];
      const double exact = M.facePatch[(std::size_t)f] == S.pipe.inlet ? S.pipe.inletVelocity[0] : 0.0;
      double vm = 0;
      for (int j = 0; j < 3; ++j) vm += S.fixedValue[(std::size_t)S.fixedSlot[(std::size_t)F[(std::size_t)j]]][0] / 3.0;
      S.fixedValue[(std::size_t)S.fixedSlot[(std::size_t)np + f]][0] = (20.0 / 9.0) * (exact - vm);
    }
  }
  if (S.duct) {
    // uniform inflow: every inlet vertex off the wall carries the inlet
    // velocity; each inlet face bubble makes the face average exact
    for (int v = 0; v < np; ++v)
      if (onInlet[(std::size_t)v] && !onWall[(std::size_t)v])
        for (int d = 0; d < 3; ++d) S.fixedValue[(std::size_t)S.fixedSlot[(std::size_t)v]][(std::size_t)d] = S.pipe.inletVelocity[d];
    const Patch& pin = M.patches[(std::size_t)S.pipe.inlet];
    for (int f = pin.start; f < pin.start + pin.count; ++f) {
      const auto& F = M.faces[(std::size_t)f];
      for (int d = 0; d < 3; ++d) {
        double vm = 0;
        for (int j = 0; j < 3; ++j) {
          // a free (wall-function wall) vertex is taken at the plug value
          const int sl = S.fixedSlot[(std::size_t)F[(std::size_t)j]];
          vm += (sl >= 0 ? S.fixedValue[(std::size_t)sl][(std::size_t)d] : S.pipe.inletVelocity[d]) / 3.0;
        }
        S.fixedValue[(std::size_t)S.fixedSlot[(std::size_t)np + f]][(std::size_t)d] =
            (20.0 / 9.0) * (S.pipe.inletVelocity[d] - vm);
      }
    }
  }
  for (int v = 0; v < np && !S.cavity && !S.duct && !S.dgInlet; ++v)
    if (onInlet[(std::size_t)v] && !onWall[(std::size_t)v]) {
      const int s = S.fixedSlot[(std::size_t)v];
      S.fixedValue[(std::size_t)s][2] =
          S.pipe.profileScale * idealUz(S.pipe, M.points[(std::size_t)v][0], M.points[(std::size_t)v][1]);
    }
  if (!S.cavity && !S.duct && !S.dgInlet) {
    const Patch& pin = M.patches[(std::size_t)S.pipe.inlet];
    for (int f = pin.start; f < pin.start + pin.count; ++f) {
      const auto& F = M.faces[(std::size_t)f];
      const V3 X[3] = {M.points[(std::size_t)F[0]], M.points[(std::size_t)F[1]], M.points[(std::size_t)F[2]]};
      const double exact = S.pipe.profileScale * triangleAvgUz(S.pipe, X);
      double vm = 0;
      for (int j = 0; j < 3; ++j) vm += S.fixedValue[(std::size_t)S.fixedSlot[(std::size_t)F[(std::size_t)j]]][2] / 3.0;
      S.fixedValue[(std::size_t)S.fixedSlot[(std::size_t)np + f]][2] = (20.0 / 9.0) * (exact - vm);
    }
  }

  // Cell DOF references and geometry.
  S.cells.resize((std::size_t)nc);
  auto entity = [&](int c, int a) {
    return a < 4 ? M.tets[(std::size_t)c][(std::size_t)a] : np + M.oppFace[(std::size_t)c][(std::size_t)(a - 4)];
  };
  parallelFor(nc, nth, 1024, [&](int cb, int ce) {
    for (int c = cb; c < ce; ++c) {
      CellPlan& cp = S.cells[(std::size_t)c];
      for (int a = 0; a < 8; ++a) {
        const int e = entity(c, a);
        const int g = S.g2free[(std::size_t)e];
        cp.ref[a] = g >= 0 ? g : -(S.fixedSlot[(std::size_t)e] + 1);
      }
      double I[3][3];
      cellGeometry(M, c, cp.det, I, &cp.h2);
      require(cp.h2 > 0.0, "zero tetrahedron diameter");
      for (int j = 0; j < 3; ++j)
        for (int d = 0; d < 3; ++d) cp.invJ[3 * j + d] = I[j][d];
      std::fill(cp.rowSlot, cp.rowSlot + 64, (std::uint8_t)255);
    }
  });

  // Free-row incidence (row -> cell, local basis) in ascending cell order.
  const int n = S.nv;
  std::vector<std::int64_t> iptr((std::size_t)n + 1, 0);
  for (int c = 0; c < nc; ++c)
    for (int a = 0; a < 8; ++a) {
      const int g = S.cells[(std::size_t)c].ref[a];
      if (g >= 0) ++iptr[(std::size_t)g + 1];
    }
  for (int r = 0; r < n; ++r) iptr[(std::size_t)r + 1] += iptr[(std::size_t)r];
  std::vector<std::int32_t> incCell((std::size_t)iptr[(std::size_t)n]);
  std::vector<std::uint8_t> incLocal((std::size_t)iptr[(std::size_t)n]);
  {
    std::vector<std::int64_t> cur(iptr.begin(), iptr.end() - 1);
    for (int c = 0; c < nc; ++c)
      for (int a = 0; a < 8; ++a) {
        const int g = S.cells[(std::size_t)c].ref[a];
        if (g < 0) continue;
        const std::int64_t z = cur[(std::size_t)g]++;
        incCell[(std::size_t)z] = c;
        incLocal[(std::size_t)z] = (std::uint8_t)a;
      }
  }

  // Momentum CSR topology: sorted unique columns of the 8x8 coupling graph.
  std::vector<std::vector<std::int32_t>> rows((std::size_t)n);
  parallelFor(n, nth, 256, [&](int rb, int re) {
    for (int r = rb; r < re; ++r) {
      auto& rr = rows[(std::size_t)r];
      for (std::int64_t z = iptr[(std::size_t)r]; z < iptr[(std::size_t)r + 1]; ++z) {
        const CellPlan& cp = S.cells[(std::size_t)incCell[(std::size_t)z]];
        for (int b = 0; b < 8; ++b)
          if (cp.ref[b] >= 0) rr.push_back(cp.ref[b]);
      }
      std::sort(rr.begin(), rr.end());
      rr.erase(std::unique(rr.begin(), rr.end()), rr.end());
      require(std::binary_search(rr.begin(), rr.end(), r), "momentum CSR row missing diagonal");
    }
  });
  CsrD& A = S.mom;
  A.n = n;
  A.row.assign((std::size_t)n + 1, 0);
  for (int r = 0; r < n; ++r) A.row[(std::size_t)r + 1] = A.row[(std::size_t)r] + (std::int64_t)rows[(std::size_t)r].size();
  A.col.resize((std::size_t)A.row[(std::size_t)n]);
  for (int r = 0; r < n; ++r) std::copy(rows[(std::size_t)r].begin(), rows[(std::size_t)r].end(), A.col.begin() + A.row[(std::size_t)r]);
  std::vector<std::vector<std::int32_t>>().swap(rows);

  // Row slots of every cell-local pair.
  parallelFor(nc, nth, 1024, [&](int cb, int ce) {
    for (int c = cb; c < ce; ++c) {
      CellPlan& cp = S.cells[(std::size_t)c];
      for (int a = 0; a < 8; ++a) {
        const int r = cp.ref[a];
        if (r < 0) continue;
        const auto first = A.col.begin() + A.row[(std::size_t)r];
        const auto last = A.col.begin() + A.row[(std::size_t)r + 1];
        for (int b = 0; b < 8; ++b) {
          if (cp.ref[b] < 0) continue;
          const auto it = std::lower_bound(first, last, cp.ref[b]);
          require(it != last && *it == cp.ref[b], "momentum row slot missing");
          const auto slot = it - first;
          require(slot < 255, "momentum row wider than 255 entries");
          cp.rowSlot[8 * a + b] = (std::uint8_t)slot;
        }
      }
    }
  });

  // Diffusion values at nu, accumulated per row in ascending cell order
  // (g3_host.hpp:150-176, alpha_u = 1).
  const Tensors& T = referenceTensors();
  const double nu = S.pipe.nu;
  A.val.assign(A.col.size(), 0.0);
  parallelFor(n, nth, 64, [&](int rb, int re) {
    for (int r = rb; r < re; ++r) {
      const std::int64_t base = A.row[(std::size_t)r];
      for (std::int64_t z = iptr[(std::size_t)r]; z < iptr[(std::size_t)r + 1]; ++z) {
        const CellPlan& cp = S.cells[(std::size_t)incCell[(std::size_t)z]];
        const int a = incLocal[(std::size_t)z];
        double I[3][3];
        for (int j = 0; j < 3; ++j)
          for (int d = 0; d < 3; ++d) I[j][d] = cp.invJ[3 * j + d];
        for (int b = 0; b < 8; ++b) {
          if (cp.ref[b] < 0) continue;
          double s = 0.0;
          for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
              for (int d = 0; d < 3; ++d) s += T.diff[a][b][j][k] * I[j][d] * I[k][d];
          A.val[(std::size_t)(base + cp.rowSlot[8 * a + b])] += nu * cp.det * s;
        }
      }
    }
  });
  A.diag.assign((std::size_t)n, 0.0);
  S.momDiagPos.assign((std::size_t)n, -1);
  for (int i = 0; i < n; ++i) {
    const auto first = A.col.begin() + A.row[(std::size_t)i];
    const auto last = A.col.begin() + A.row[(std::size_t)i + 1];
    const auto it = std::lower_bound(first, last, i);
    const std::size_t k = (std::size_t)(it - A.col.begin());
    require(A.val[k] > 0.0 && std::isfinite(A.val[k]), "non-positive diffusion diagonal");
    require(k <= (std::size_t)std::numeric_limits<std::int32_t>::max(), "momentum nnz exceeds int32 diagonal index");
    A.diag[(std::size_t)i] = A.val[k];
    S.momDiagPos[(std::size_t)i] = (std::int32_t)k;
  }

  S.coloring = greedyColoring(n, A.row.data(), A.col.data());

  // Dirichlet diffusion right-hand side (g4_host.hpp:309-310).
  for (auto& r : S.staticRhs) r.assign((std::size_t)n, 0.0);
  for (const CellPlan& cp : S.cells) {
    double I[3][3];
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) I[j][d] = cp.invJ[3 * j + d];
    double metric[3][3] = {{0}};
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) metric[j][k] += I[j][d] * I[k][d];
    for (int a = 0; a < 8; ++a) {
      const int r = cp.ref[a];
      if (r < 0) continue;
      for (int b = 0; b < 8; ++b) {
        if (cp.ref[b] >= 0) continue;
        double v = 0;
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) v += T.diff[a][b][j][k] * metric[j][k];
        const double K = nu * cp.det * v;
        const auto& fv = S.fixedValue[(std::size_t)(-cp.ref[b] - 1)];
        for (int d = 0; d < 3; ++d) S.staticRhs[(std::size_t)d][(std::size_t)r] -= K * fv[(std::size_t)d];
      }
    }
  }

  // B coefficient table, fixed divergence and volumes (g4_host.hpp:311-313).
  S.bcoef.resize((std::size_t)nc * 24);
  S.fixedDiv.assign((std::size_t)nc, 0.0);
  S.volumes.assign((std::size_t)nc, 0.0);
  S.cellZ.assign((std::size_t)nc, 0.0);
  // DG inlet: each cell's inlet face (one at most) and its outward area
  // vector (cell_inlet_face_g2, g2_host.hpp:45-50)
  std::vector<int> inletOpp;
  std::vector<V3> inletSf;
  if (S.dgInlet) {
    inletOpp.assign((std::size_t)nc, -1);
    inletSf.assign((std::size_t)nc, V3{0, 0, 0});
    const Patch& pin = M.patches[(std::size_t)S.pipe.inlet];
    for (int f = pin.start; f < pin.start + pin.count; ++f) {
      const int c = M.owner[(std::size_t)f];
      int lf = -1;
      for (int i = 0; i < 4; ++i)
        if (M.oppFace[(std::size_t)c][(std::size_t)i] == f) lf = i;
      require(lf >= 0, "inlet face " + std::to_string(f) + " is not a face of its owner cell", kErrInternal);
      require(inletOpp[(std::size_t)c] < 0, "tet has multiple DG inlet faces");
      inletOpp[(std::size_t)c] = lf;
      inletSf[(std::size_t)c] = outwardAreaVector(M, f);
      DgInletFace q;
      q.cell = c;
      q.lf = lf;
      const V3& sf = inletSf[(std::size_t)c];
      q.area = std::sqrt(sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]);
      require(q.area > 0.0, "degenerate DG inlet face");
      for (int d = 0; d < 3; ++d) q.n[d] = sf[(std::size_t)d] / q.area;
      S.dgFaces.push_back(q);
    }
  }
  // Effective B: the DG inlet removes the inlet face's trace, -sf/3 on the
  // face's vertices and -(9/20) sf on its bubble (g4_effective_bcoeff_host,
  // g4_host.hpp:77-86, same operation order).
  auto effB = [&](int c, const CellPlan& cp, int a, int d) {
    double v = bcoefOf(cp, a, d);
    if (S.dgInlet && inletOpp[(std::size_t)c] >= 0) {
      const int o = inletOpp[(std::size_t)c];
      if (a < 4 && a != o) v -= inletSf[(std::size_t)c][(std::size_t)d] / 3.0;
      else if (a == 4 + o) v -= (9.0 / 20.0) * inletSf[(std::size_t)c][(std::size_t)d];
    }
    return v;
  };
  for (int c = 0; c < nc; ++c) {
    const CellPlan& cp = S.cells[(std::size_t)c];
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) S.bcoef[(std::size_t)c * 24 + (std::size_t)(3 * a + d)] = effB(c, cp, a, d);
    S.volumes[(std::size_t)c] = cp.det / 6.0;
    double q = 0;
    if (S.dgInlet && inletOpp[(std::size_t)c] >= 0)  // the prescribed inflow flux (g4_host.hpp:313)
      q += S.pipe.inletVelocity[0] * inletSf[(std::size_t)c][0] + S.pipe.inletVelocity[1] * inletSf[(std::size_t)c][1] +
           S.pipe.inletVelocity[2] * inletSf[(std::size_t)c][2];
    for (int a = 0; a < 8; ++a) {
      if (cp.ref[a] >= 0) continue;
      const auto& fv = S.fixedValue[(std::size_t)(-cp.ref[a] - 1)];
      for (int d = 0; d < 3; ++d) q += effB(c, cp, a, d) * fv[(std::size_t)d];
    }
    S.fixedDiv[(std::size_t)c] = q;
    double z = 0;
    for (int i = 0; i < 4; ++i) z += M.points[(std::size_t)M.tets[(std::size_t)c][(std::size_t)i]][2] * .25;
    S.cellZ[(std::size_t)c] = z;
  }
  // Wall-function walls: B_N = B - int_wall (n . u), the divergence without
  // the wall's normal flux (s22d3_patch_normal_wall_divergence,
  // turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/modules/wall/
  // stage22d3_wall_b_host.inc:4-20, applied to the production B at
  // sst_application.inc:802-808).  On a wall face the three vertex traces
  // integrate to area/3 and the face's own bubble to 27/60 area (the
  // reference's 12-point rule is exact for both); every other basis has zero
  // trace there.  Wall DOFs are free, so the fixed divergence is unchanged.
  if (S.wallFunction) {
    const Patch& wp = M.patches[(std::size_t)S.pipe.wall];
    for (int f = wp.start; f < wp.start + wp.count; ++f) {
      const int c = M.owner[(std::size_t)f];
      int lf = -1;
      for (int i = 0; i < 4; ++i)
        if (M.oppFace[(std::size_t)c][(std::size_t)i] == f) lf = i;
      if (lf < 0) fail("wall face " + std::to_string(f) + " is not a face of its owner cell");
      const V3 sf = outwardAreaVector(M, f);
      for (int a = 0; a < 8; ++a) {
        const double w = a < 4 ? (a == lf ? 0.0 : 1.0 / 3.0) : (a == 4 + lf ? 27.0 / 60.0 : 0.0);
        if (w == 0.0) continue;
        for (int d = 0; d < 3; ++d) S.bcoef[(std::size_t)c * 24 + (std::size_t)(3 * a + d)] -= w * sf[(std::size_t)d];
      }
    }
  }
  S.zMin = 1e300;
  S.zMax = -1e300;
  for (const auto& x : M.points) {
    S.zMin = std::min(S.zMin, x[2]);
    S.zMax = std::max(S.zMax, x[2]);
  }

  // Setup rAU from the zero-velocity assembly (h8_main_base.cu:1917-1922);
  // with the DG inlet from the plug on every free vertex (:1918).
  std::array<std::vector<double>, 3> U0;
  for (auto& u : U0) u.assign((std::size_t)n, 0.0);
  if (S.dgInlet) {
    for (int v = 0; v < np; ++v) {
      const int g = S.g2free[(std::size_t)v];
      if (g < 0) continue;
      for (int d = 0; d < 3; ++d) U0[(std::size_t)d][(std::size_t)g] = S.pipe.inletVelocity[d];
    }
    S.plug.assign((std::size_t)n * 3, 0.0);
    for (int g = 0; g < n; ++g)
      for (int d = 0; d < 3; ++d) S.plug[(std::size_t)g * 3 + d] = U0[(std::size_t)d][(std::size_t)g];
  }
  std::vector<double> a0;
  std::array<std::vector<double>, 3> c0;
  hostAssembleCentral(S, U0, a0, c0);
  S.initRau = hostFinalizeRelax(S, a0, prm.alphaU, &S.initDelta);
  if (prm.rauScale != 1.0)
    for (double& q : S.initRau) q *= prm.rauScale;
  if (S.duct) {
    for (int pass = 0; pass < 2; ++pass) {
      const Patch& pt = M.patches[(std::size_t)(pass ? S.pipe.outlet : S.pipe.inlet)];
      for (int f = pt.start; f < pt.start + pt.count; ++f) {
        const auto& F = M.faces[(std::size_t)f];
        const V3 X[3] = {M.points[(std::size_t)F[0]], M.points[(std::size_t)F[1]], M.points[(std::size_t)F[2]]};
        (pass ? S.outletCells : S.inletCells).push_back(M.owner[(std::size_t)f]);
        (pass ? S.outletAreas : S.inletAreas).push_back(0.5 * norm(cross(sub(X[1], X[0]), sub(X[2], X[0]))));
      }
    }
    S.wallDistance = wallDistanceField(M, S.pipe.wall, nth);
  }
  return S;
}

double pressureDropPatches(const H8Setup& S, const std::vector<double>& p) {
  if (p.size() != (std::size_t)S.nc || S.inletCells.empty() || S.outletCells.empty())
    return std::numeric_limits<double>::quiet_NaN();
  auto mean = [&](const std::vector<std::int32_t>& c, const std::vector<double>& a) {
    long double sa = 0, sp = 0;
    for (std::size_t i = 0; i < c.size(); ++i) {
      sa += a[i];
      sp += a[i] * p[(std::size_t)c[i]];
    }
    return (double)(sp / sa);
  };
  return mean(S.inletCells, S.inletAreas) - mean(S.outletCells, S.outletAreas);
}

namespace {

// Distance from point p to triangle (a, b, c): the closest point by the
// Voronoi-region tests of Ericson, Real-Time Collision Detection, 5.1.5.
double pointTriangleDistance(const V3& p, const V3& a, const V3& b, const V3& c) {
  const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  auto dot3 = [](const V3& u, const V3& v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  auto at = [&](double v, double w) {
    const V3 q = {a[0] + v * ab[0] + w * ac[0], a[1] + v * ab[1] + w * ac[1], a[2] + v * ab[2] + w * ac[2]};
    return norm(sub(p, q));
  };
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0 && d2 <= 0) return norm(ap);
  const V3 bp = sub(p, b);
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0 && d4 <= d3) return norm(bp);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) return at(d1 / (d1 - d3), 0.0);
  const V3 cp = sub(p, c);
  const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0 && d5 <= d6) return norm(cp);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) return at(0.0, d2 / (d2 - d6));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return at(1.0 - w, w);
  }
  const double den = 1.0 / (va + vb + vc);
  return at(vb * den, vc * den);
}

}  // namespace

std::vector<double> wallDistanceField(const TetMesh& M, int patch, int threads) {
  require(patch >= 0 && patch < (int)M.patches.size(), "wall patch not found for the distance field");
  const Patch& pt = M.patches[(std::size_t)patch];
  require(pt.count > 0, "empty wall patch");
  // uniform grid over the wall triangles' bounding boxes
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  double meanEdge = 0;
  for (int f = pt.start; f < pt.start + pt.count; ++f)
    for (int j = 0; j < 3; ++j) {
      const V3& x = M.points[(std::size_t)M.faces[(std::size_t)f][(std::size_t)j]];
      const V3& y = M.points[(std::size_t)M.faces[(std::size_t)f][(std::size_t)((j + 1) % 3)]];
      meanEdge += norm(sub(x, y));
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], x[(std::size_t)d]);
        hi[d] = std::max(hi[d], x[(std::size_t)d]);
      }
    }
  meanEdge /= 3.0 * pt.count;
  const double h = std::max(2.0 * meanEdge, 1e-12);
  int dim[3];
  for (int d = 0; d < 3; ++d) dim[d] = std::max(1, std::min(512, (int)std::ceil((hi[d] - lo[d]) / h) + 1));
  auto cellOf = [&](double x, int d) {
    return std::max(0, std::min(dim[d] - 1, (int)std::floor((x - lo[d]) / h)));
  };
  const std::size_t ncell = (std::size_t)dim[0] * dim[1] * dim[2];
  std::vector<std::vector<int>> bins(ncell);
  for (int f = pt.start; f < pt.start + pt.count; ++f) {
    int b0[3] = {1 << 30, 1 << 30, 1 << 30}, b1[3] = {-1, -1, -1};
    for (int j = 0; j < 3; ++j) {
      const V3& x = M.points[(std::size_t)M.faces[(std::size_t)f][(std::size_t)j]];
      for (int d = 0; d < 3; ++d) {
        b0[d] = std::min(b0[d], cellOf(x[(std::size_t)d], d));
        b1[d] = std::max(b1[d], cellOf(x[(std::size_t)d], d));
      }
    }
    for (int i = b0[0]; i <= b1[0]; ++i)
      for (int j = b0[1]; j <= b1[1]; ++j)
        for (int k = b0[2]; k <= b1[2]; ++k) bins[((std::size_t)i * dim[1] + j) * dim[2] + k].push_back(f);
  }
  const int np = M.nPoints();
  std::vector<double> dist((std::size_t)np, 0.0);
  parallelFor(np, hostThreads(threads), 256, [&](int vb, int ve) {
    std::vector<int> seen;
    for (int v = vb; v < ve; ++v) {
      const V3& p = M.points[(std::size_t)v];
      const int c[3] = {cellOf(p[0], 0), cellOf(p[1], 1), cellOf(p[2], 2)};
      double best = 1e300;
      // rings of grid cells; every triangle within distance r*h of p lies in
      // ring <= r + 1, so the search stops once best <= r * h
      for (int r = 0; r < 1024; ++r) {
        for (int i = c[0] - r; i <= c[0] + r; ++i)
          for (int j = c[1] - r; j <= c[1] + r; ++j)
            for (int k = c[2] - r; k <= c[2] + r; ++k) {
              if (std::max(std::abs(i - c[0]), std::max(std::abs(j - c[1]), std::abs(k - c[2]))) != r) continue;
              if (i < 0 || j < 0 || k < 0 || i >= dim[0] || j >= dim[1] || k >= dim[2]) continue;
              for (int f : bins[((std::size_t)i * dim[1] + j) * dim[2] + k]) {
                const auto& F = M.faces[(std::size_t)f];
                best = std::min(best, pointTriangleDistance(p, M.points[(std::size_t)F[0]], M.points[(std::size_t)F[1]],
                                                            M.points[(std::size_t)F[2]]));
              }
            }
        if (best <= r * h) break;
        if (r > std::max(dim[0], std::max(dim[1], dim[2]))) break;
      }
      dist[(std::size_t)v] = best;
    }
  });
  return dist;
}

void hostAssembleCentral(const H8Setup& S, const std::array<std::vector<double>, 3>& U,
                         std::vector<double>& aval, std::array<std::vector<double>, 3>& conv) {
  const Tensors& T = referenceTensors();
  const CsrD& A = S.mom;
  aval = A.val;
  for (auto& r : conv) r.assign((std::size_t)A.n, 0.0);
  for (const CellPlan& cp : S.cells) {
    double I[3][3];
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) I[j][d] = cp.invJ[3 * j + d];
    double cf[3][8] = {{0}};
    for (int m = 0; m < 8; ++m) {
      const int r = cp.ref[m];
      for (int d = 0; d < 3; ++d)
        cf[d][m] = r >= 0 ? U[(std::size_t)d][(std::size_t)r] : S.fixedValue[(std::size_t)(-r - 1)][(std::size_t)d];
    }
    double ur[8][3] = {{0}};
    for (int m = 0; m < 8; ++m)
      for (int j = 0; j < 3; ++j)
        for (int d = 0; d < 3; ++d) ur[m][j] += cf[d][m] * I[j][d];
    for (int a = 0; a < 8; ++a) {
      const int r = cp.ref[a];
      if (r < 0) continue;
      for (int b = 0; b < 8; ++b) {
        double v = 0;
        for (int m = 0; m < 8; ++m)
          for (int j = 0; j < 3; ++j) v += ur[m][j] * T.cent[a][m][b][j];
        const double C = cp.det * v;
        if (cp.ref[b] >= 0) {
          aval[(std::size_t)(A.row[(std::size_t)r] + cp.rowSlot[8 * a + b])] += C;
        } else {
          const auto& fv = S.fixedValue[(std::size_t)(-cp.ref[b] - 1)];
          for (int d = 0; d < 3; ++d) conv[(std::size_t)d][(std::size_t)r] -= C * fv[(std::size_t)d];
        }
      }
    }
  }
}

std::vector<double> hostFinalizeRelax(const H8Setup& S, std::vector<double>& a, double alphaU,
                                      std::vector<double>* delta) {
  const CsrD& A = S.mom;
  std::vector<double> rau((std::size_t)A.n), del((std::size_t)A.n);
  const double fac = 1.0 / alphaU - 1.0;
  for (int i = 0; i < A.n; ++i) {
    const std::size_t dp = (std::size_t)S.momDiagPos[(std::size_t)i];
    double m = 0;
    for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k) m += std::abs(a[(std::size_t)k]);
    const double de = fac * m;
    del[(std::size_t)i] = de;
    a[dp] += de;
    require(a[dp] > 0, "relaxed momentum diagonal invalid");
    rau[(std::size_t)i] = 1.0 / a[dp];
  }
  if (delta) *delta = std::move(del);
  return rau;
}

FineTopo buildFineTopo(const H8Setup& S) {
  FineTopo H;
  H.n = S.nc;
  H.nv = S.nv;
  require((std::uint64_t)H.n < (1ull << 29), "packed cell index exceeds 29 bits");
  require((std::uint64_t)H.nv < (1ull << 26), "free velocity count exceeds 26 bits");
  // Free velocity DOF -> (cell, basis) incidence in ascending cell order.
  H.incOff.assign((std::size_t)H.nv + 1, 0);
  for (const CellPlan& cp : S.cells)
    for (int a = 0; a < 8; ++a)
      if (cp.ref[a] >= 0) ++H.incOff[(std::size_t)cp.ref[a] + 1];
  for (int g = 0; g < H.nv; ++g) H.incOff[(std::size_t)g + 1] += H.incOff[(std::size_t)g];
  H.packed.resize((std::size_t)H.incOff[(std::size_t)H.nv]);
  {
    std::vector<std::int32_t> cur(H.incOff.begin(), H.incOff.end() - 1);
    for (int c = 0; c < H.n; ++c)
      for (int a = 0; a < 8; ++a) {
        const int g = S.cells[(std::size_t)c].ref[a];
        if (g >= 0) H.packed[(std::size_t)cur[(std::size_t)g]++] = ((std::uint32_t)c << 3) | (std::uint32_t)a;
      }
  }
  // Per pressure row: directed contributions (a, incident cell) in the order
  // of the refresh loop, their unique sorted targets and uint8 slots
  // (h2b_fine_csr.cuh:103-148).
  std::vector<std::int32_t> rowCount((std::size_t)H.n), contribCount((std::size_t)H.n);
  const int nth = hostThreads(0);
  auto target = [&](int c, std::vector<std::int32_t>& t) {
    t.clear();
    const CellPlan& cp = S.cells[(std::size_t)c];
    for (int a = 0; a < 8; ++a) {
      const int g = cp.ref[a];
      if (g < 0) continue;
      for (std::int32_t k = H.incOff[(std::size_t)g]; k < H.incOff[(std::size_t)g + 1]; ++k)
        t.push_back((std::int32_t)(H.packed[(std::size_t)k] >> 3));
    }
  };
  parallelFor(H.n, nth, 256, [&](int cb, int ce) {
    std::vector<std::int32_t> t, u;
    for (int c = cb; c < ce; ++c) {
      target(c, t);
      u = t;
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      require(!u.empty(), "empty pressure CSR row");
      require(u.size() <= 255, "pressure row wider than 255 entries");
      require(std::binary_search(u.begin(), u.end(), c), "pressure diagonal absent");
      rowCount[(std::size_t)c] = (std::int32_t)u.size();
      contribCount[(std::size_t)c] = (std::int32_t)t.size();
    }
  });
  H.row.assign((std::size_t)H.n + 1, 0);
  H.contribOff.assign((std::size_t)H.n + 1, 0);
  std::int64_t nnz = 0, dir = 0;
  for (int c = 0; c < H.n; ++c) {
    H.row[(std::size_t)c] = (std::int32_t)nnz;
    H.contribOff[(std::size_t)c] = (std::int32_t)dir;
    nnz += rowCount[(std::size_t)c];
    dir += contribCount[(std::size_t)c];
    H.rowMax = std::max<std::uint32_t>(H.rowMax, (std::uint32_t)rowCount[(std::size_t)c]);
    require(nnz <= std::numeric_limits<std::int32_t>::max() && dir <= std::numeric_limits<std::int32_t>::max(),
            "fine pressure CSR exceeds int32");
  }
  H.row[(std::size_t)H.n] = (std::int32_t)nnz;
  H.contribOff[(std::size_t)H.n] = (std::int32_t)dir;
  H.col.resize((std::size_t)nnz);
  H.slot.resize((std::size_t)dir);
  H.diagPos.assign((std::size_t)H.n, -1);
  // Gather layout: terms per nonzero, counted then filled.
  std::vector<std::int32_t> termCount((std::size_t)nnz, 0);
  parallelFor(H.n, nth, 256, [&](int cb, int ce) {
    std::vector<std::int32_t> t, u;
    for (int c = cb; c < ce; ++c) {
      target(c, t);
      u = t;
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      const std::int32_t rb = H.row[(std::size_t)c];
      std::copy(u.begin(), u.end(), H.col.begin() + rb);
      H.diagPos[(std::size_t)c] = rb + (std::int32_t)(std::lower_bound(u.begin(), u.end(), c) - u.begin());
      const std::int32_t qb = H.contribOff[(std::size_t)c];
      for (std::size_t q = 0; q < t.size(); ++q) {
        const auto s = std::lower_bound(u.begin(), u.end(), t[q]) - u.begin();
        H.slot[(std::size_t)qb + q] = (std::uint8_t)s;
        ++termCount[(std::size_t)(rb + s)];
      }
    }
  });
  H.nzOff.assign((std::size_t)nnz + 1, 0);
  for (std::int64_t k = 0; k < nnz; ++k) H.nzOff[(std::size_t)k + 1] = H.nzOff[(std::size_t)k] + termCount[(std::size_t)k];
  H.nzTerm.resize((std::size_t)dir);
  parallelFor(H.n, nth, 256, [&](int cb, int ce) {
    for (int c = cb; c < ce; ++c) {
      const CellPlan& cp = S.cells[(std::size_t)c];
      const std::int32_t rb = H.row[(std::size_t)c];
      const std::int32_t re = H.row[(std::size_t)c + 1];
      std::vector<std::int32_t> fill(H.nzOff.begin() + rb, H.nzOff.begin() + re);
      std::int32_t q = H.contribOff[(std::size_t)c];
      for (int a = 0; a < 8; ++a) {
        const int g = cp.ref[a];
        if (g < 0) continue;
        for (std::int32_t k = H.incOff[(std::size_t)g]; k < H.incOff[(std::size_t)g + 1]; ++k, ++q) {
          const std::uint32_t b = H.packed[(std::size_t)k] & 7u;
          const std::int32_t s = H.slot[(std::size_t)q];
          H.nzTerm[(std::size_t)fill[(std::size_t)s]++] = ((std::uint32_t)g << 6) | ((std::uint32_t)a << 3) | b;
        }
      }
    }
  });
  return H;
}

double pressureDropFit(const H8Setup& S, const std::vector<double>& p) {
  if (p.size() != (std::size_t)S.nc) return std::numeric_limits<double>::quiet_NaN();
  long double sz = 0, sp = 0, szz = 0, szp = 0;
  const long double n = (long double)p.size();
  for (std::size_t c = 0; c < p.size(); ++c) {
    const double z = S.cellZ[c];
    sz += z;
    sp += p[c];
    szz += z * z;
    szp += z * p[c];
  }
  const long double den = n * szz - sz * sz;
  if (std::abs((double)den) < 1e-300) return 0;
  const long double slope = (n * szp - sz * sp) / den;
  return (double)(-slope * (S.zMax - S.zMin));
}

}  // namespace nodals
