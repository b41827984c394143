// SA AMG setup.  See amg_setup.hpp.  Compile with -ffp-contract=off.
#include "amg_setup.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <unordered_map>
#include <utility>

#include "error.hpp"

namespace nodals {

namespace {

using Entry = std::pair<int, double>;

// Merge duplicate columns, keep the own aggregate plus the largest-magnitude
// entries up to maxNnz, then scale the row to unit sum
// (g5_sa_host.hpp:81-108).
void normalizePrune(std::vector<Entry>& row, int own, int maxNnz) {
  if (row.empty()) row.push_back({own, 1.0});
  std::sort(row.begin(), row.end(), [](const Entry& a, const Entry& b) { return a.first < b.first; });
  std::size_t w = 0;
  for (std::size_t k = 0; k < row.size();) {
    const int c = row[k].first;
    double v = 0.0;
    std::size_t j = k;
    for (; j < row.size() && row[j].first == c; ++j) v += row[j].second;
    row[w++] = {c, v};
    k = j;
  }
  row.resize(w);
  if (maxNnz > 0 && (int)row.size() > maxNnz) {
    std::vector<Entry> keep;
    keep.reserve((std::size_t)maxNnz);
    auto it = std::find_if(row.begin(), row.end(), [&](const Entry& e) { return e.first == own; });
    keep.push_back(it != row.end() ? *it : Entry{own, 0.0});
    std::vector<Entry> others;
    others.reserve(row.size());
    for (const Entry& e : row)
      if (e.first != own) others.push_back(e);
    std::sort(others.begin(), others.end(), [](const Entry& a, const Entry& b) {
      const double aa = std::abs(a.second), bb = std::abs(b.second);
      return aa > bb || (aa == bb && a.first < b.first);
    });
    for (const Entry& e : others)
      if ((int)keep.size() < maxNnz) keep.push_back(e);
    std::sort(keep.begin(), keep.end(), [](const Entry& a, const Entry& b) { return a.first < b.first; });
    row.swap(keep);
  }
  double sum = 0.0;
  for (const Entry& e : row) sum += e.second;
  if (!std::isfinite(sum) || std::abs(sum) < 1e-14) {
    row.clear();
    row.push_back({own, 1.0});
    return;
  }
  for (Entry& e : row) e.second /= sum;
}

CsrD mapsToCsr(std::vector<std::unordered_map<int, double>>& rows) {
  // Exact zeros are dropped; columns sorted (g5_sa_host.hpp:504-521).
  CsrD A;
  A.n = (int)rows.size();
  A.row.assign((std::size_t)A.n + 1, 0);
  A.diag.assign((std::size_t)A.n, 0.0);
  std::vector<Entry> tmp;
  for (int i = 0; i < A.n; ++i) {
    tmp.clear();
    for (const auto& kv : rows[(std::size_t)i])
      if (kv.second != 0.0) tmp.push_back(kv);
    std::sort(tmp.begin(), tmp.end(), [](const Entry& a, const Entry& b) { return a.first < b.first; });
    for (const Entry& e : tmp) {
      A.col.push_back(e.first);
      A.val.push_back(e.second);
      if (e.first == i) A.diag[(std::size_t)i] = e.second;
    }
    A.row[(std::size_t)i + 1] = (std::int64_t)A.col.size();
    require(A.diag[(std::size_t)i] > 0.0 && std::isfinite(A.diag[(std::size_t)i]),
            "SA coarse matrix non-positive diagonal");
  }
  return A;
}

// Matrix-free y = B diag(rau) B^T x on the host (g5_sa_host.hpp:110-132).
void schurApply(const H8Setup& S, const std::vector<double>& rau, const std::vector<double>& x,
                std::vector<double>& y, std::vector<double>& v0, std::vector<double>& v1,
                std::vector<double>& v2) {
  std::fill(v0.begin(), v0.end(), 0.0);
  std::fill(v1.begin(), v1.end(), 0.0);
  std::fill(v2.begin(), v2.end(), 0.0);
  const double* B = S.bcoef.data();
  for (int c = 0; c < S.nc; ++c) {
    const CellPlan& cp = S.cells[(std::size_t)c];
    const double p = x[(std::size_t)c];
    for (int a = 0; a < 8; ++a) {
      const int g = cp.ref[a];
      if (g < 0) continue;
      const double* b = B + (std::size_t)c * 24 + 3 * a;
      v0[(std::size_t)g] += b[0] * p;
      v1[(std::size_t)g] += b[1] * p;
      v2[(std::size_t)g] += b[2] * p;
    }
  }
  for (int g = 0; g < S.nv; ++g) {
    const double r = rau[(std::size_t)g];
    v0[(std::size_t)g] *= r;
    v1[(std::size_t)g] *= r;
    v2[(std::size_t)g] *= r;
  }
  y.assign((std::size_t)S.nc, 0.0);
  for (int c = 0; c < S.nc; ++c) {
    const CellPlan& cp = S.cells[(std::size_t)c];
    double s = 0.0;
    for (int a = 0; a < 8; ++a) {
      const int g = cp.ref[a];
      if (g < 0) continue;
      const double* b = B + (std::size_t)c * 24 + 3 * a;
      s += b[0] * v0[(std::size_t)g] + b[1] * v1[(std::size_t)g] + b[2] * v2[(std::size_t)g];
    }
    y[(std::size_t)c] = s;
  }
}

// Symmetric power iteration on D^-1/2 A D^-1/2 from the sin/cos start vector
// (g5_sa_host.hpp:134-171).
template <class Apply>
double powerIteration(int n, const std::vector<double>& diag, int its, double safety, Apply&& apply) {
  std::vector<double> v((std::size_t)n), u((std::size_t)n), au, w((std::size_t)n);
  long double ss = 0.0L;
  for (int i = 0; i < n; ++i) {
    const double g = (double)(i + 1);
    v[(std::size_t)i] = std::sin(.731 * g) + .27 * std::cos(1.117 * g);
    ss += (long double)v[(std::size_t)i] * v[(std::size_t)i];
  }
  const double vn = std::sqrt((double)ss);
  for (auto& q : v) q /= vn;
  double ray = 0.0;
  for (int it = 0; it < its; ++it) {
    for (int i = 0; i < n; ++i) u[(std::size_t)i] = v[(std::size_t)i] / std::sqrt(diag[(std::size_t)i]);
    apply(u, au);
    long double num = 0.0L, den = 0.0L, nw = 0.0L;
    for (int i = 0; i < n; ++i) {
      w[(std::size_t)i] = au[(std::size_t)i] / std::sqrt(diag[(std::size_t)i]);
      num += (long double)v[(std::size_t)i] * w[(std::size_t)i];
      den += (long double)v[(std::size_t)i] * v[(std::size_t)i];
      nw += (long double)w[(std::size_t)i] * w[(std::size_t)i];
    }
    ray = (double)(num / std::max(den, (long double)1e-300));
    const double nrm = std::sqrt((double)nw);
    require(nrm > 0.0 && std::isfinite(nrm) && ray > 0.0 && std::isfinite(ray), "SA power iteration failed");
    for (int i = 0; i < n; ++i) v[(std::size_t)i] = w[(std::size_t)i] / nrm;
  }
  return safety * ray;
}

void csrApply(const CsrD& A, const std::vector<double>& x, std::vector<double>& y) {
  y.assign((std::size_t)A.n, 0.0);
  for (int i = 0; i < A.n; ++i) {
    double s = 0.0;
    for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k)
      s += A.val[(std::size_t)k] * x[(std::size_t)A.col[(std::size_t)k]];
    y[(std::size_t)i] = s;
  }
}

// Smoothed transfer of an explicit level: P = normalize(prune(I - w D^-1 A)).
Transfer explicitTransfer(const CsrD& A, const AggResult& G, double lambdaMax, double damping, int maxNnz) {
  const double omega = damping / lambdaMax;
  Transfer P;
  P.nFine = A.n;
  P.nCoarse = G.nagg;
  P.row.assign((std::size_t)A.n + 1, 0);
  std::vector<Entry> row;
  for (int i = 0; i < A.n; ++i) {
    row.clear();
    const int own = G.id[(std::size_t)i];
    row.push_back({own, 1.0});
    const double invd = 1.0 / A.diag[(std::size_t)i];
    for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k)
      row.push_back({G.id[(std::size_t)A.col[(std::size_t)k]], -omega * invd * A.val[(std::size_t)k]});
    normalizePrune(row, own, maxNnz);
    for (const Entry& e : row) {
      P.col.push_back(e.first);
      P.val.push_back(e.second);
    }
    P.row[(std::size_t)i + 1] = (std::int64_t)P.col.size();
  }
  return P;
}

// Galerkin product P^T A P (g5_sa_host.hpp:940-971).
CsrD galerkin(const CsrD& A, const Transfer& P) {
  require(P.nFine == A.n, "SA PtAP transfer size mismatch");
  std::vector<std::unordered_map<int, double>> rows((std::size_t)P.nCoarse);
  std::vector<Entry> q;
  for (int i = 0; i < A.n; ++i) {
    q.clear();
    for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k) {
      const int j = A.col[(std::size_t)k];
      const double a = A.val[(std::size_t)k];
      for (std::int64_t kp = P.row[(std::size_t)j]; kp < P.row[(std::size_t)j + 1]; ++kp)
        q.push_back({P.col[(std::size_t)kp], a * P.val[(std::size_t)kp]});
    }
    std::sort(q.begin(), q.end(), [](const Entry& a, const Entry& b) { return a.first < b.first; });
    std::size_t w = 0;
    for (std::size_t k = 0; k < q.size();) {
      const int c = q[k].first;
      double v = 0.0;
      std::size_t j = k;
      for (; j < q.size() && q[j].first == c; ++j) v += q[j].second;
      q[w++] = {c, v};
      k = j;
    }
    q.resize(w);
    for (std::int64_t kp = P.row[(std::size_t)i]; kp < P.row[(std::size_t)i + 1]; ++kp) {
      const int I = P.col[(std::size_t)kp];
      const double pi = P.val[(std::size_t)kp];
      for (const Entry& e : q) {
        const double v = pi * e.second;
        if (v != 0.0) rows[(std::size_t)I][e.first] += v;
      }
    }
  }
  return mapsToCsr(rows);
}

void finishHierarchy(AmgHierarchy& H, const AmgOptions& o) {
  while (H.A.back().n > o.terminal) {
    const CsrD& A = H.A.back();
    const double lam = powerCsr(A, o.powerIts, o.safety);
    H.levelLambda.resize(H.A.size(), 0.0);
    H.levelLambda[H.A.size() - 1] = lam;
    AggResult G = aggregateGraph(csrGraph(A), o.target, o.minSize, o.softMax);
    H.agg.push_back(G);
    H.P.push_back(explicitTransfer(A, G, lam, o.damping, o.maxPnnz));
    CsrD C = galerkin(A, H.P.back());
    require(C.n < A.n, "SA coarsening stalled");
    H.A.push_back(std::move(C));
  }
  H.levelLambda.resize(H.A.size(), 0.0);
  H.terminalN = H.A.back().n;
  H.terminalInv = denseInverseCholesky(H.A.back());
}

}  // namespace

Graph meshFaceGraph(const TetMesh& M) {
  Graph g((std::size_t)M.nCells());
  for (int f = 0; f < M.nInternal(); ++f) {
    const int a = M.owner[(std::size_t)f], b = M.neighbour[(std::size_t)f];
    g[(std::size_t)a].push_back(b);
    g[(std::size_t)b].push_back(a);
  }
  for (auto& v : g) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  return g;
}

Graph csrGraph(const CsrD& A) {
  Graph g((std::size_t)A.n);
  for (int i = 0; i < A.n; ++i)
    for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k) {
      const int j = A.col[(std::size_t)k];
      if (j != i) g[(std::size_t)i].push_back(j);
    }
  return g;
}

AggResult aggregateGraph(const Graph& g, int target, int minSize, int softMax) {
  // Greedy BFS aggregates of up to `target` rows seeded in natural order, then
  // aggregates smaller than minSize merged into the smallest neighbouring
  // aggregate, preferring merges that stay within softMax
  // (g2_host.hpp:84-86).
  const int n = (int)g.size();
  std::vector<int> id((std::size_t)n, -1);
  std::vector<std::vector<int>> members;
  std::vector<int> q;
  q.reserve((std::size_t)target);
  for (int seed = 0; seed < n; ++seed) {
    if (id[(std::size_t)seed] >= 0) continue;
    const int a = (int)members.size();
    q.clear();
    q.push_back(seed);
    id[(std::size_t)seed] = a;
    for (std::size_t h = 0; h < q.size() && (int)q.size() < target; ++h) {
      const int i = q[h];
      for (int j : g[(std::size_t)i])
        if (id[(std::size_t)j] < 0 && (int)q.size() < target) {
          id[(std::size_t)j] = a;
          q.push_back(j);
        }
    }
    members.push_back(q);
  }
  for (int a = 0; a < (int)members.size(); ++a) {
    auto& ma = members[(std::size_t)a];
    if (ma.empty() || (int)ma.size() >= minSize) continue;
    int best = -1, bestSize = INT_MAX;
    bool bestFit = false;
    for (int i : ma)
      for (int j : g[(std::size_t)i]) {
        const int b = id[(std::size_t)j];
        if (b < 0 || b == a || members[(std::size_t)b].empty()) continue;
        const int sz = (int)members[(std::size_t)b].size();
        const bool fit = (int)ma.size() + sz <= softMax;
        if (best < 0 || (fit && !bestFit) || (fit == bestFit && sz < bestSize)) {
          best = b;
          bestSize = sz;
          bestFit = fit;
        }
      }
    if (best >= 0) {
      for (int i : ma) {
        id[(std::size_t)i] = best;
        members[(std::size_t)best].push_back(i);
      }
      ma.clear();
    }
  }
  std::vector<int> compact(members.size(), -1);
  int na = 0;
  for (std::size_t a = 0; a < members.size(); ++a)
    if (!members[a].empty()) compact[a] = na++;
  AggResult R;
  R.nagg = na;
  R.id.resize((std::size_t)n);
  for (int i = 0; i < n; ++i) {
    R.id[(std::size_t)i] = compact[(std::size_t)id[(std::size_t)i]];
    require(R.id[(std::size_t)i] >= 0, "aggregation lost a row", kErrInternal);
  }
  return R;
}

double powerCsr(const CsrD& A, int its, double safety) {
  return powerIteration(A.n, A.diag, its, safety,
                        [&](const std::vector<double>& x, std::vector<double>& y) { csrApply(A, x, y); });
}

std::vector<double> denseInverseCholesky(const CsrD& A) {
  // Cholesky factor, column-by-column triangular solves, then symmetrise
  // (g2_host.hpp:137-138).
  const int n = A.n;
  const std::size_t N = (std::size_t)n;
  std::vector<double> M(N * N, 0.0), L(N * N, 0.0);
  for (int i = 0; i < n; ++i)
    for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k)
      M[(std::size_t)i * N + (std::size_t)A.col[(std::size_t)k]] = A.val[(std::size_t)k];
  double md = 0;
  for (int i = 0; i < n; ++i) md = std::max(md, std::abs(M[(std::size_t)i * N + (std::size_t)i]));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = M[(std::size_t)i * N + (std::size_t)j];
      for (int k = 0; k < j; ++k) s -= L[(std::size_t)i * N + (std::size_t)k] * L[(std::size_t)j * N + (std::size_t)k];
      if (i == j) {
        require(s > std::max(1e-30, md * 1e-20) && std::isfinite(s),
                "terminal Cholesky pivot failure at row " + std::to_string(i));
        L[(std::size_t)i * N + (std::size_t)j] = std::sqrt(s);
      } else {
        L[(std::size_t)i * N + (std::size_t)j] = s / L[(std::size_t)j * N + (std::size_t)j];
      }
    }
  std::vector<double> inv(N * N, 0.0), y(N), x(N);
  for (int col = 0; col < n; ++col) {
    for (int i = 0; i < n; ++i) {
      double s = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= L[(std::size_t)i * N + (std::size_t)k] * y[(std::size_t)k];
      y[(std::size_t)i] = s / L[(std::size_t)i * N + (std::size_t)i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[(std::size_t)i];
      for (int k = i + 1; k < n; ++k) s -= L[(std::size_t)k * N + (std::size_t)i] * x[(std::size_t)k];
      x[(std::size_t)i] = s / L[(std::size_t)i * N + (std::size_t)i];
    }
    for (int i = 0; i < n; ++i) inv[(std::size_t)i * N + (std::size_t)col] = x[(std::size_t)i];
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double v = 0.5 * (inv[(std::size_t)i * N + (std::size_t)j] + inv[(std::size_t)j * N + (std::size_t)i]);
      inv[(std::size_t)i * N + (std::size_t)j] = inv[(std::size_t)j * N + (std::size_t)i] = v;
    }
  return inv;
}

AmgHierarchy buildSaHierarchy(const TetMesh& M, const H8Setup& S, const FineTopo& F,
                              const std::vector<double>& rau, const AmgOptions& o) {
  AmgHierarchy H;
  const double* B = S.bcoef.data();
  // Level-0 aggregates of the cell face graph.
  AggResult G0 = aggregateGraph(meshFaceGraph(M), o.target, o.minSize, o.softMax);
  H.agg.push_back(G0);
  // Fine diagonal sum_a sum_d rAU_g b^2 (g2_host.hpp:78).
  H.fineDiag.assign((std::size_t)S.nc, 0.0);
  for (int c = 0; c < S.nc; ++c) {
    const CellPlan& cp = S.cells[(std::size_t)c];
    double s = 0;
    for (int a = 0; a < 8; ++a) {
      const int g = cp.ref[a];
      if (g < 0) continue;
      for (int k = 0; k < 3; ++k) {
        const double b = B[(std::size_t)c * 24 + (std::size_t)(3 * a + k)];
        s += rau[(std::size_t)g] * b * b;
      }
    }
    H.fineDiag[(std::size_t)c] = s;
  }
  {
    std::vector<double> v0((std::size_t)S.nv), v1((std::size_t)S.nv), v2((std::size_t)S.nv);
    H.fineLambda = powerIteration(S.nc, H.fineDiag, o.powerIts, o.safety,
                                  [&](const std::vector<double>& x, std::vector<double>& y) {
                                    schurApply(S, rau, x, y, v0, v1, v2);
                                  });
  }
  // Smoothed fine transfer (g5_sa_host.hpp:173-203).
  {
    const double omega = o.damping / H.fineLambda;
    Transfer P;
    P.nFine = S.nc;
    P.nCoarse = G0.nagg;
    P.row.assign((std::size_t)S.nc + 1, 0);
    std::vector<Entry> row;
    for (int i = 0; i < S.nc; ++i) {
      row.clear();
      const int own = G0.id[(std::size_t)i];
      row.push_back({own, 1.0});
      const CellPlan& ci = S.cells[(std::size_t)i];
      const double invd = 1.0 / H.fineDiag[(std::size_t)i];
      for (int a = 0; a < 8; ++a) {
        const int g = ci.ref[a];
        if (g < 0) continue;
        const double* bi = B + (std::size_t)i * 24 + 3 * a;
        for (std::int32_t k = F.incOff[(std::size_t)g]; k < F.incOff[(std::size_t)g + 1]; ++k) {
          const int j = (int)(F.packed[(std::size_t)k] >> 3);
          const int aj = (int)(F.packed[(std::size_t)k] & 7u);
          const double* bj = B + (std::size_t)j * 24 + 3 * aj;
          double dot = 0.0;
          for (int d = 0; d < 3; ++d) dot += bi[d] * bj[d];
          row.push_back({G0.id[(std::size_t)j], -omega * invd * rau[(std::size_t)g] * dot});
        }
      }
      normalizePrune(row, own, o.maxPnnz);
      for (const Entry& e : row) {
        P.col.push_back(e.first);
        P.val.push_back(e.second);
      }
      P.row[(std::size_t)i + 1] = (std::int64_t)P.col.size();
    }
    H.P.push_back(std::move(P));
  }
  // First coarse operator P^T B diag(rAU) B^T P, factorised per velocity DOF
  // (g5_sa_host.hpp:525-546).
  {
    const Transfer& P = H.P[0];
    struct QE {
      int coarse;
      double b[3];
    };
    std::vector<std::unordered_map<int, double>> rows((std::size_t)P.nCoarse);
    std::vector<QE> q;
    for (int g = 0; g < S.nv; ++g) {
      q.clear();
      for (std::int32_t k = F.incOff[(std::size_t)g]; k < F.incOff[(std::size_t)g + 1]; ++k) {
        const int c = (int)(F.packed[(std::size_t)k] >> 3);
        const int a = (int)(F.packed[(std::size_t)k] & 7u);
        const double* bb = B + (std::size_t)c * 24 + 3 * a;
        for (std::int64_t kp = P.row[(std::size_t)c]; kp < P.row[(std::size_t)c + 1]; ++kp) {
          const double pv = P.val[(std::size_t)kp];
          QE e;
          e.coarse = P.col[(std::size_t)kp];
          for (int d = 0; d < 3; ++d) e.b[d] = pv * bb[d];
          q.push_back(e);
        }
      }
      std::sort(q.begin(), q.end(), [](const QE& a, const QE& b) { return a.coarse < b.coarse; });
      std::size_t w = 0;
      for (std::size_t k = 0; k < q.size();) {
        const int c = q[k].coarse;
        double bs[3] = {0, 0, 0};
        std::size_t j = k;
        for (; j < q.size() && q[j].coarse == c; ++j)
          for (int d = 0; d < 3; ++d) bs[d] += q[j].b[d];
        q[w].coarse = c;
        for (int d = 0; d < 3; ++d) q[w].b[d] = bs[d];
        ++w;
        k = j;
      }
      q.resize(w);
      const double r = rau[(std::size_t)g];
      for (std::size_t a = 0; a < q.size(); ++a)
        for (std::size_t b = a; b < q.size(); ++b) {
          const double v = r * (q[a].b[0] * q[b].b[0] + q[a].b[1] * q[b].b[1] + q[a].b[2] * q[b].b[2]);
          if (v == 0.0) continue;
          rows[(std::size_t)q[a].coarse][q[b].coarse] += v;
          if (a != b) rows[(std::size_t)q[b].coarse][q[a].coarse] += v;
        }
    }
    H.A.push_back(mapsToCsr(rows));
  }
  finishHierarchy(H, o);
  return H;
}

AmgHierarchy buildSaHierarchyCsr(const CsrD& A0, const AmgOptions& o) {
  AmgHierarchy H;
  H.fineDiag = A0.diag;
  H.fineLambda = powerCsr(A0, o.powerIts, o.safety);
  AggResult G = aggregateGraph(csrGraph(A0), o.target, o.minSize, o.softMax);
  H.agg.push_back(G);
  H.P.push_back(explicitTransfer(A0, G, H.fineLambda, o.damping, o.maxPnnz));
  H.A.push_back(galerkin(A0, H.P.back()));
  finishHierarchy(H, o);
  return H;
}

CsrD galerkinProduct(const CsrD& A, const Transfer& P) { return galerkin(A, P); }

// ------------------------------------------------------------------ CF/PMIS

CsrD exactFineCsr(const H8Setup& S, const FineTopo& F, const std::vector<double>& rau) {
  // Per cell, per free basis a, per incidence entry of its DOF, add
  // rAU_g (B[c][a] . B[c2][b]) into the slot recorded for that contribution
  // (g8_cf_hierarchy_host.hpp:116-166).
  require((int)rau.size() == F.nv, "CF fine CSR rAU size mismatch");
  CsrD A;
  A.n = F.n;
  A.row.assign(F.row.begin(), F.row.end());
  A.col = F.col;
  A.val.assign(F.col.size(), 0.0);
  A.diag.assign((std::size_t)F.n, 0.0);
  const double* B = S.bcoef.data();
  for (int c = 0; c < F.n; ++c) {
    const int rb = F.row[(std::size_t)c];
    const CellPlan& cp = S.cells[(std::size_t)c];
    int q = F.contribOff[(std::size_t)c];
    for (int a = 0; a < 8; ++a) {
      const int g = cp.ref[a];
      if (g < 0) continue;
      const double rv = rau[(std::size_t)g];
      const double* ba = B + (std::size_t)c * 24 + 3 * a;
      for (std::int32_t k = F.incOff[(std::size_t)g]; k < F.incOff[(std::size_t)g + 1]; ++k, ++q) {
        const std::uint32_t pk = F.packed[(std::size_t)k];
        const int c2 = (int)(pk >> 3), b = (int)(pk & 7u);
        const double* bb = B + (std::size_t)c2 * 24 + 3 * b;
        const double dot = ba[0] * bb[0] + ba[1] * bb[1] + ba[2] * bb[2];
        A.val[(std::size_t)(rb + (int)F.slot[(std::size_t)q])] += rv * dot;
      }
    }
    const double d = A.val[(std::size_t)F.diagPos[(std::size_t)c]];
    require(d > 0.0 && std::isfinite(d), "CF fine CSR non-positive diagonal");
    A.diag[(std::size_t)c] = d;
  }
  return A;
}

void pinCsr(CsrD& A, int cell) {
  require(cell >= 0 && cell < A.n, "pinned cell out of range");
  for (int i = 0; i < A.n; ++i)
    for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k) {
      const int j = A.col[(std::size_t)k];
      if (i != j && (i == cell || j == cell)) A.val[(std::size_t)k] = 0.0;
    }
}

std::uint64_t cfHash64(std::uint64_t x) {
  // SplitMix64 output function (Steele, Lea and Flood, OOPSLA 2014): the
  // state advanced by the golden-ratio increment, then two xor-shift-multiply
  // rounds and a final xor-shift.
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

namespace {

double csrEntry(const CsrD& A, int i, int j) {
  const auto b = A.col.begin() + A.row[(std::size_t)i];
  const auto e = A.col.begin() + A.row[(std::size_t)i + 1];
  const auto it = std::lower_bound(b, e, j);
  return (it == e || *it != j) ? 0.0 : A.val[(std::size_t)(it - A.col.begin())];
}

}  // namespace

CfSplit pmisSplit(const CsrD& A, double theta) {
  // g8_cf_hierarchy_host.hpp:169-293.
  const int n = A.n;
  CfSplit R;
  R.n = n;
  R.strong.resize((std::size_t)n);
  R.graph.resize((std::size_t)n);
  for (int i = 0; i < n; ++i) {
    double mx = 0.0;
    for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k) {
      const int j = A.col[(std::size_t)k];
      const double a = A.val[(std::size_t)k];
      if (j != i && a < 0.0) mx = std::max(mx, -a);
    }
    const double thr = theta * mx;
    if (thr > 0.0) {
      for (std::int64_t k = A.row[(std::size_t)i]; k < A.row[(std::size_t)i + 1]; ++k) {
        const int j = A.col[(std::size_t)k];
        const double a = A.val[(std::size_t)k];
        if (j != i && a < 0.0 && -a >= thr) R.strong[(std::size_t)i].push_back(j);
      }
    }
    // CSR columns are sorted and unique, so each strong row is too.
  }
  // Symmetric union graph, sorted and unique.
  for (int i = 0; i < n; ++i)
    for (int j : R.strong[(std::size_t)i]) {
      R.graph[(std::size_t)i].push_back(j);
      R.graph[(std::size_t)j].push_back(i);
    }
  for (auto& g : R.graph) {
    std::sort(g.begin(), g.end());
    g.erase(std::unique(g.begin(), g.end()), g.end());
  }
  // Priority: larger union degree, then larger hash of (index + 0x63021),
  // then larger index.
  std::vector<std::uint64_t> h((std::size_t)n);
  for (int i = 0; i < n; ++i) h[(std::size_t)i] = cfHash64((std::uint64_t)i + 0x63021ull);
  auto higher = [&](int a, int b) {
    const std::size_t da = R.graph[(std::size_t)a].size(), db = R.graph[(std::size_t)b].size();
    if (da != db) return da > db;
    if (h[(std::size_t)a] != h[(std::size_t)b]) return h[(std::size_t)a] > h[(std::size_t)b];
    return a > b;
  };
  R.state.assign((std::size_t)n, 0);
  std::vector<int> winners;
  int undecided = n;
  while (undecided > 0) {
    winners.clear();
    for (int i = 0; i < n; ++i) {
      if (R.state[(std::size_t)i] != 0) continue;
      bool win = true;
      for (int j : R.graph[(std::size_t)i])
        if (R.state[(std::size_t)j] == 0 && higher(j, i)) {
          win = false;
          break;
        }
      if (win) winners.push_back(i);
    }
    require(!winners.empty(), "PMIS made no progress", kErrInternal);
    for (int i : winners) {
      R.state[(std::size_t)i] = 1;
      --undecided;
    }
    for (int i : winners)
      for (int j : R.graph[(std::size_t)i])
        if (R.state[(std::size_t)j] == 0) {
          R.state[(std::size_t)j] = 2;
          --undecided;
        }
    ++R.rounds;
    require(R.rounds <= n, "PMIS excessive rounds", kErrInternal);
  }
  R.coarseId.assign((std::size_t)n, -1);
  for (int i = 0; i < n; ++i)
    if (R.state[(std::size_t)i] == 1) R.coarseId[(std::size_t)i] = R.nC++;
  return R;
}

Transfer cfTransfer(const CsrD& A, const CfSplit& S, int pmax, const std::string& interp) {
  // g8_cf_hierarchy_host.hpp:625-785.  Weights accumulate per coarse target
  // in visiting order; candidates are then ordered by (weight desc, target
  // asc), cut to pmax, normalised to unit sum in that order and stored by
  // target.
  require(interp == "direct" || interp == "exti", "CF interpolation must be direct or exti");
  Transfer P;
  P.nFine = A.n;
  P.nCoarse = S.nC;
  P.row.assign((std::size_t)A.n + 1, 0);
  std::map<int, double> acc;
  std::vector<Entry> cand;
  const bool exti = interp == "exti";
  for (int i = 0; i < A.n; ++i) {
    cand.clear();
    if (S.state[(std::size_t)i] == 1) {
      cand.push_back({S.coarseId[(std::size_t)i], 1.0});
    } else {
      acc.clear();
      for (int j : S.strong[(std::size_t)i]) {
        if (S.state[(std::size_t)j] != 1) continue;
        const double a = csrEntry(A, i, j);
        const double w = a < 0.0 ? -a : std::abs(a);
        if (w > 0.0) acc[S.coarseId[(std::size_t)j]] += w;
      }
      if (exti) {
        for (int k : S.strong[(std::size_t)i]) {
          if (S.state[(std::size_t)k] != 2) continue;
          const double aik = csrEntry(A, i, k);
          const double path = aik < 0.0 ? -aik : 0.0;
          if (!(path > 0.0)) continue;
          double denom = 0.0;
          for (int j : S.strong[(std::size_t)k]) {
            if (S.state[(std::size_t)j] != 1) continue;
            const double akj = csrEntry(A, k, j);
            if (akj < 0.0) denom += -akj;
          }
          if (!(denom > 0.0)) continue;
          for (int j : S.strong[(std::size_t)k]) {
            if (S.state[(std::size_t)j] != 1) continue;
            const double akj = csrEntry(A, k, j);
            if (akj < 0.0) acc[S.coarseId[(std::size_t)j]] += path * ((-akj) / denom);
          }
        }
      }
      if (acc.empty()) {
        for (int j : S.graph[(std::size_t)i]) {
          if (S.state[(std::size_t)j] != 1) continue;
          const double w = std::abs(csrEntry(A, i, j));
          if (w > 0.0) acc[S.coarseId[(std::size_t)j]] += w;
        }
      }
      require(!acc.empty(), "CF interpolation row has no C target");
      for (const auto& kv : acc)
        if (kv.second > 0.0) cand.push_back(kv);
      require(!cand.empty(), "CF interpolation empty candidates");
      std::sort(cand.begin(), cand.end(), [](const Entry& a, const Entry& b) {
        return a.second != b.second ? a.second > b.second : a.first < b.first;
      });
      if (pmax > 0 && (int)cand.size() > pmax) cand.resize((std::size_t)pmax);
      double sum = 0.0;
      for (const Entry& e : cand) sum += e.second;
      require(sum > 0.0 && std::isfinite(sum), "CF interpolation invalid weight sum");
      for (Entry& e : cand) e.second /= sum;
      std::sort(cand.begin(), cand.end(), [](const Entry& a, const Entry& b) { return a.first < b.first; });
    }
    for (const Entry& e : cand) {
      P.col.push_back(e.first);
      P.val.push_back(e.second);
    }
    P.row[(std::size_t)i + 1] = (std::int64_t)P.col.size();
  }
  return P;
}

Transfer composeTransfers(const Transfer& P01, const Transfer& P12) {
  // Exact product P01 P12, exact zeros dropped, no pruning
  // (g8_cf_hierarchy_host.hpp:1100-1195).
  require(P01.nCoarse == P12.nFine, "transfer composition size mismatch");
  Transfer Q;
  Q.nFine = P01.nFine;
  Q.nCoarse = P12.nCoarse;
  Q.row.assign((std::size_t)Q.nFine + 1, 0);
  std::map<int, double> acc;
  for (int i = 0; i < Q.nFine; ++i) {
    acc.clear();
    for (std::int64_t a = P01.row[(std::size_t)i]; a < P01.row[(std::size_t)i + 1]; ++a) {
      const int mid = P01.col[(std::size_t)a];
      const double w0 = P01.val[(std::size_t)a];
      for (std::int64_t b = P12.row[(std::size_t)mid]; b < P12.row[(std::size_t)mid + 1]; ++b)
        acc[P12.col[(std::size_t)b]] += w0 * P12.val[(std::size_t)b];
    }
    require(!acc.empty(), "composed interpolation empty row");
    for (const auto& kv : acc)
      if (kv.second != 0.0) {
        Q.col.push_back(kv.first);
        Q.val.push_back(kv.second);
      }
    Q.row[(std::size_t)i + 1] = (std::int64_t)Q.col.size();
  }
  return Q;
}

namespace {

AmgHierarchy cfFromFine(CsrD A, const CfOptions& o) {
  // g8_cf_hierarchy_host.hpp:1584-1783.
  AmgHierarchy H;
  H.fineDiag = A.diag;
  H.fineLambda = o.powerIts > 0 ? powerCsr(A, o.powerIts, o.safety) : 1.0;
  auto split = [&](const CsrD& M) {
    CfSplit S = pmisSplit(M, o.theta);
    H.cfState.push_back(S.state);
    H.cfCoarseId.push_back(S.coarseId);
    H.cfRounds.push_back(S.rounds);
    return S;
  };
  auto lambdaOf = [&](const CsrD& M) { return o.powerIts > 0 ? powerCsr(M, o.powerIts, o.safety) : 1.0; };
  int logical = 0;
  if (o.aggressiveFirst && A.n > o.terminal) {
    const CfSplit s0 = split(A);
    Transfer P01 = cfTransfer(A, s0, o.pmax, o.interp);
    CsrD A1 = galerkin(A, P01);
    if (A1.n <= o.terminal) {
      H.P.push_back(std::move(P01));
      H.A.push_back(std::move(A1));
    } else {
      const CfSplit s1 = split(A1);
      Transfer P12 = cfTransfer(A1, s1, o.pmax, o.interp);
      CsrD A2 = galerkin(A1, P12);
      H.P.push_back(composeTransfers(P01, P12));
      H.A.push_back(std::move(A2));
    }
    if (H.A.back().n > o.terminal) {
      H.levelLambda.assign(H.A.size(), 0.0);
      H.levelLambda[0] = lambdaOf(H.A.back());
    }
    A = H.A.back();
    logical = 1;
  }
  while (H.A.empty() || H.A.back().n > o.terminal) {
    const CfSplit s = split(A);
    H.P.push_back(cfTransfer(A, s, o.pmax, o.interp));
    H.A.push_back(galerkin(A, H.P.back()));
    if (H.A.back().n <= o.terminal) break;
    if (H.levelLambda.size() < H.A.size()) H.levelLambda.resize(H.A.size(), 0.0);
    H.levelLambda[H.A.size() - 1] = lambdaOf(H.A.back());
    A = H.A.back();
    ++logical;
    require(logical <= 20, "CF hierarchy excessive levels");
  }
  if (H.levelLambda.size() < H.A.size()) H.levelLambda.resize(H.A.size(), 1.0);
  H.terminalN = H.A.back().n;
  H.terminalInv = denseInverseCholesky(H.A.back());
  return H;
}

}  // namespace

AmgHierarchy buildCfHierarchy(const H8Setup& S, const FineTopo& F, const std::vector<double>& rau,
                              const CfOptions& o) {
  return cfFromFine(exactFineCsr(S, F, rau), o);
}

AmgHierarchy buildCfHierarchyCsr(const CsrD& A0, const CfOptions& o) { return cfFromFine(A0, o); }

}  // namespace nodals
