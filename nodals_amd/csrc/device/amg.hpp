// Device AMG V-cycle over a host-built SA hierarchy.
//
// Cycle structure of G4AMG::apply / explicit_vcycle
// (gpu/serial_cuda/h8_main_base.cu:532-564): smoothers start from zero, the
// post-smoother is applied as a correction to the current residual, the
// terminal level is a dense inverse mat-vec.  Restriction is a gather SpMV
// with the stored transpose P^T (no atomics), prolongation adds P x_c.
//
// Template parameters: TF is the storage type of the finest operator values
// (double, float or bf16), TV the type of every vector and of the coarse
// levels (double, or float for the mixed-precision hierarchy).
#pragma once

#include <cmath>
#include <string>
#include <vector>

#include "../host/amg_setup.hpp"
#include "la_kernels.hpp"

namespace nodals {
namespace dev {

struct SmootherOptions {
  std::string kind = "jacobi";  // jacobi | l1jacobi | cheb2 | mcgs
  double jacobiOmega = 0.7;
  int jacobiFineSweeps = 1, jacobiCoarseSweeps = 1;
  int chebDegree = 2;
  double lambdaLowFraction = 0.05;
  double mcgsOmega = 1.0;
  int mcgsFineSweeps = 1, mcgsCoarseSweeps = 1;
  std::string mcgsOrder = "symmetric";  // forward | backward | symmetric
  // Coarse levels with at most this many nonzeros (and vectors that fit
  // kFuseLdsBytes of LDS) run each MCGS smoothing call as one workgroup
  // (k_mcgs_level_wg) instead of one launch per colour; 0 turns it off.
  // One CU streams a level's matrix once per colour pass: on the 207k helix's
  // CF hierarchy the 108k-nonzero level gains and the 873k-nonzero level
  // loses against launches spread over the chip.
  long long mcgsFuseNnz = 250000;
};

// LDS of the fused level smoother: iterate, right-hand side, diagonal and
// row pointers.
constexpr std::size_t kFuseLdsBytes = 96 * 1024;
template <class TV>
constexpr std::size_t fuseLdsBytes(int n) {
  return (std::size_t)n * (3 * sizeof(TV) + sizeof(int)) + sizeof(int);
}

// View of the finest operator, owned elsewhere (refreshed each SIMPLE step).
template <class TF, class TV>
struct FineView {
  int n = 0;
  int lanes = 16;
  double meanRow = 8.0;  // nonzeros per row (lanes of the per-colour passes)
  const int* row = nullptr;
  const int* col = nullptr;
  const TF* val = nullptr;
  const TV* diag = nullptr;
  const TV* l1 = nullptr;
};

template <class TV>
struct DevColoring {
  int ncolors = 0;
  std::vector<int> off;
  DBuf<int> rows;
  bool contiguous = false;  // the operator is numbered in colour order: colour c is rows [off[c], off[c+1])
  void upload(const Coloring& C) {
    ncolors = C.ncolors;
    off = C.offset;
    rows.upload(C.rows);
    contiguous = false;
  }
  void uploadContiguous(const Coloring& C) {
    ncolors = C.ncolors;
    off = C.offset;
    contiguous = true;
  }
  const int* rowList() const { return contiguous ? nullptr : rows.data(); }
};

template <class TF, class TV>
class DevAmg {
 public:
  struct Level {
    int n = 0, lanes = 8;
    double meanRow = 1.0;
    DBuf<int> row, col;
    DBuf<TV> val, diag, l1, b, x, r, tmp, corr;
    double lambda = 1.0;
    DevColoring<TV> color;
    bool fused = false;          // MCGS smoothing in one workgroup
    DBuf<int> colorOff, colorLanes;
  };
  struct Xfer {
    int nFine = 0, nCoarse = 0, lanesP = 8, lanesT = 8;
    bool aggregation = false;  // exactly one entry per fine row (row i holds entry i)
    DBuf<int> prow, pcol, trow, tcol;
    DBuf<TV> pval, tval;
  };

  void setup(const AmgHierarchy& H, int fineN, const SmootherOptions& o, const Coloring* fineColoring) {
    opt_ = o;
    fineN_ = fineN;
    fineLambda_ = H.fineLambda;
    const std::size_t nl = H.A.size();
    require(nl >= 1 && H.P.size() == nl, "AMG hierarchy shape mismatch", kErrInternal);
    L_.clear();
    P_.clear();
    L_.resize(nl);
    P_.resize(nl);
    // An MCGS-smoothed coarse level is stored renumbered in colour order
    // (the reference's greedy colouring, unchanged): every colour is then
    // a contiguous row range and a colour pass needs no row list.  Each
    // row keeps its entries in their order, and the transfers on both
    // sides are relabelled the same way, so every sum is formed in the
    // same order as in the original numbering and results are identical.
    // perm[l][new] = original row; empty = not renumbered.
    std::vector<std::vector<int>> perm(nl), inv(nl);
    for (std::size_t l = 0; l < nl; ++l) {
      const CsrD& A = H.A[l];
      Level& Lv = L_[l];
      Lv.n = A.n;
      Lv.meanRow = A.n ? (double)A.col.size() / A.n : 1.0;
      Lv.lanes = lanesFor(Lv.meanRow, A.n);
      if (opt_.kind == "mcgs" && l + 1 < nl) {
        const Coloring C = greedyColoring(A.n, A.row.data(), A.col.data());
        perm[l].assign(C.rows.begin(), C.rows.end());
        inv[l].assign((std::size_t)A.n, 0);
        for (int i = 0; i < A.n; ++i) inv[l][(std::size_t)perm[l][(std::size_t)i]] = i;
        Lv.color.uploadContiguous(C);
        if (o.mcgsFuseNnz > 0 && (long long)A.col.size() <= o.mcgsFuseNnz && fuseLdsBytes<TV>(A.n) <= kFuseLdsBytes) {
          std::vector<int> lanes((std::size_t)C.ncolors);
          for (int c = 0; c < C.ncolors; ++c)
            lanes[(std::size_t)c] = lanesFor(Lv.meanRow, C.offset[(std::size_t)c + 1] - C.offset[(std::size_t)c]);
          Lv.colorOff.upload(C.offset);
          Lv.colorLanes.upload(lanes);
          Lv.fused = true;
        }
      }
      const std::vector<int>& pm = perm[l];
      const std::vector<int>& iv = inv[l];
      auto orig = [&](int i) { return pm.empty() ? i : pm[(std::size_t)i]; };
      auto relabel = [&](int j) { return iv.empty() ? j : iv[(std::size_t)j]; };
      std::vector<int> r32((std::size_t)A.n + 1, 0), c32;
      std::vector<double> v64, d64((std::size_t)A.n), l1((std::size_t)A.n);
      c32.reserve(A.col.size());
      v64.reserve(A.col.size());
      for (int i = 0; i < A.n; ++i) {
        const int o = orig(i);
        double s = 0;
        for (std::int64_t k = A.row[(std::size_t)o]; k < A.row[(std::size_t)o + 1]; ++k) {
          c32.push_back(relabel(A.col[(std::size_t)k]));
          v64.push_back(A.val[(std::size_t)k]);
          s += std::abs(A.val[(std::size_t)k]);
        }
        r32[(std::size_t)i + 1] = (int)c32.size();
        d64[(std::size_t)i] = A.diag[(std::size_t)o];
        l1[(std::size_t)i] = s;
      }
      Lv.row.upload(r32);
      Lv.col.upload(c32);
      Lv.val.upload(castVec(v64));
      Lv.diag.upload(castVec(d64));
      Lv.l1.upload(castVec(l1));
      for (DBuf<TV>* v : {&Lv.b, &Lv.x, &Lv.r, &Lv.tmp, &Lv.corr}) v->alloc((std::size_t)A.n);
      Lv.x.upload(std::vector<TV>((std::size_t)A.n, TV(1)));
      Lv.lambda = l < H.levelLambda.size() ? H.levelLambda[l] : 1.0;
    }
    static const std::vector<int> none;
    for (std::size_t l = 0; l < nl; ++l) uploadTransfer(H.P[l], P_[l], l > 0 ? inv[l - 1] : none, inv[l]);
    termN_ = H.terminalN;
    tinv_.upload(castVec(H.terminalInv));
    for (DBuf<TV>* v : {&fr_, &ftmp_, &fcorr_}) v->alloc((std::size_t)fineN);
    if (opt_.kind == "mcgs") {
      require(fineColoring != nullptr, "MCGS smoother requires a fine coloring", kErrInternal);
      fineColor_.upload(*fineColoring);
    }
  }

  void setFine(const FineView<TF, TV>& v) { fine_ = v; }
  // tmp = A_l x on explicit level l (measurement of the coarse SpMV).
  void levelSpmv(hipStream_t s, int l) {
    Level& A = L_[(std::size_t)l];
    spmv<TV, TV, false, false>(s, A.lanes, A.n, A.row.data(), A.col.data(), A.val.data(), A.x.data(), nullptr,
                               A.tmp.data());
  }
  int levels() const { return (int)L_.size(); }
  const Level& level(int l) const { return L_[(std::size_t)l]; }

  // apply() replayed from a captured HIP graph (b and z must stay fixed
  // buffers; the finest values behind the fine view may change in place).
  void applyGraphed(hipStream_t s, const TV* b, TV* z) {
    graph_.run(s, graphKey({b, z, fine_.val, fine_.diag}), [&]() { apply(s, b, z); });
  }

  // z = M^{-1} b for FP64 vectors b, z, replayed from one captured graph per
  // (b, z) pair: with the fused Jacobi cycle the FP64 <-> TV roundings happen
  // inside its first and last kernels (identical values to separate casts);
  // otherwise through the fixed TV buffers.  Returns false (nothing done)
  // when the smoother is not the fused Jacobi cycle.
  bool applyWide(hipStream_t s, const double* b, double* z) {
    if (!fusedJacobi()) return false;
    wide_.run(s, graphKey({b, z, fine_.val, fine_.diag}), [&]() { fusedCycle<double, double>(s, b, z); });
    return true;
  }

  // z = M^{-1} b on the finest level.
  void apply(hipStream_t s, const TV* b, TV* z) {
    if (fusedJacobi()) {
      fusedCycle<TV, TV>(s, b, z);
      return;
    }
    smoothFine(s, b, z);
    spmv<TF, TV, true, false>(s, fine_.lanes, fine_.n, fine_.row, fine_.col, fine_.val, z, b, fr_.data());
    Level& C = L_[0];
    restrictTo(s, P_[0], fr_.data(), C.b.data());
    vcycle(s, 0, C.b.data(), C.x.data());
    prolongAdd(s, P_[0], C.x.data(), z);
    spmv<TF, TV, true, false>(s, fine_.lanes, fine_.n, fine_.row, fine_.col, fine_.val, z, b, fr_.data());
    smoothFine(s, fr_.data(), fcorr_.data());
    axpy<TV>(s, fine_.n, 1.0, fcorr_.data(), z);
  }

 private:
  template <class TB, class TZ>
  // Pre-smoothing + residual, coarse correction, residual + post-smoothing:
  // two fine passes instead of five kernels.
  void fusedCycle(hipStream_t s, const TB* b, TZ* z) {
    const TV* den = opt_.kind == "l1jacobi" ? fine_.l1 : fine_.diag;
    jacobiPreResid<TF, TV, TB>(s, fine_.lanes, fine_.n, fine_.row, fine_.col, fine_.val, b, den, opt_.jacobiOmega,
                               ftmp_.data(), fr_.data());
    Level& C = L_[0];
    restrictTo(s, P_[0], fr_.data(), C.b.data());
    vcycle(s, 0, C.b.data(), C.x.data());
    const Xfer& X = P_[0];
    // The fine level keeps the prolongation as its own pass: folded into the
    // post-smoothing gather (as on the coarse levels) its two-level gather
    // made the 40k-cell solve about 1% slower once the row loops kept
    // several loads in flight.
    prolongAdd(s, X, C.x.data(), ftmp_.data());
    residJacobiPost<TF, TV, TB, TZ>(s, fine_.lanes, fine_.n, fine_.row, fine_.col, fine_.val, b, den,
                                    opt_.jacobiOmega, ftmp_.data(), z);
  }

  template <class TS>
  static std::vector<TV> castVec(const std::vector<TS>& v) {
    std::vector<TV> o(v.size());
    for (std::size_t i = 0; i < v.size(); ++i) o[i] = (TV)v[i];
    return o;
  }

  // Upload transfer P (fine rows -> coarse columns) and its transpose, with
  // the fine side relabelled by invF and the coarse side by invC (original
  // -> stored index; empty = unchanged).  Rows keep their entries' order;
  // the transpose is formed in the original numbering (within a coarse row,
  // fine rows ascend) and relabelled afterwards, so its sums keep the
  // original order too.
  static void uploadTransfer(const Transfer& P, Xfer& X, const std::vector<int>& invF,
                             const std::vector<int>& invC) {
    X.nFine = P.nFine;
    X.nCoarse = P.nCoarse;
    auto relF = [&](int i) { return invF.empty() ? i : invF[(std::size_t)i]; };
    auto relC = [&](int c) { return invC.empty() ? c : invC[(std::size_t)c]; };
    std::vector<int> origF((std::size_t)P.nFine), origC((std::size_t)P.nCoarse);
    for (int i = 0; i < P.nFine; ++i) origF[(std::size_t)relF(i)] = i;
    for (int c = 0; c < P.nCoarse; ++c) origC[(std::size_t)relC(c)] = c;
    X.aggregation = (std::size_t)P.nFine == P.col.size();
    for (int i = 0; X.aggregation && i < P.nFine; ++i)
      X.aggregation = P.row[(std::size_t)i] == i && P.row[(std::size_t)i + 1] == i + 1;
    std::vector<int> pr((std::size_t)P.nFine + 1, 0), pc;
    std::vector<double> pv;
    pc.reserve(P.col.size());
    pv.reserve(P.col.size());
    for (int i = 0; i < P.nFine; ++i) {
      const int o = origF[(std::size_t)i];
      for (std::int64_t k = P.row[(std::size_t)o]; k < P.row[(std::size_t)o + 1]; ++k) {
        pc.push_back(relC(P.col[(std::size_t)k]));
        pv.push_back(P.val[(std::size_t)k]);
      }
      pr[(std::size_t)i + 1] = (int)pc.size();
    }
    X.prow.upload(pr);
    X.pcol.upload(pc);
    X.pval.upload(castVec(pv));
    X.lanesP = lanesFor(P.nFine ? (double)P.col.size() / P.nFine : 1.0, P.nFine);
    // Transpose for gather restriction, in the original numbering.
    std::vector<int> tr((std::size_t)P.nCoarse + 1, 0);
    for (int c : P.col) ++tr[(std::size_t)c + 1];
    for (int c = 0; c < P.nCoarse; ++c) tr[(std::size_t)c + 1] += tr[(std::size_t)c];
    std::vector<int> tc(P.col.size());
    std::vector<double> tv(P.col.size());
    std::vector<int> cur(tr.begin(), tr.end() - 1);
    for (int i = 0; i < P.nFine; ++i)
      for (std::int64_t k = P.row[(std::size_t)i]; k < P.row[(std::size_t)i + 1]; ++k) {
        const int c = P.col[(std::size_t)k];
        const int z = cur[(std::size_t)c]++;
        tc[(std::size_t)z] = i;
        tv[(std::size_t)z] = P.val[(std::size_t)k];
      }
    // ... then its rows reordered by the coarse relabelling and its columns
    // relabelled on the fine side.
    std::vector<int> tr2((std::size_t)P.nCoarse + 1, 0), tc2;
    std::vector<double> tv2;
    tc2.reserve(tc.size());
    tv2.reserve(tc.size());
    for (int c = 0; c < P.nCoarse; ++c) {
      const int o = origC[(std::size_t)c];
      for (int k = tr[(std::size_t)o]; k < tr[(std::size_t)o + 1]; ++k) {
        tc2.push_back(relF(tc[(std::size_t)k]));
        tv2.push_back(tv[(std::size_t)k]);
      }
      tr2[(std::size_t)c + 1] = (int)tc2.size();
    }
    X.trow.upload(tr2);
    X.tcol.upload(tc2);
    X.tval.upload(castVec(tv2));
    X.lanesT = lanesFor(P.nCoarse ? (double)tc.size() / P.nCoarse : 1.0, P.nCoarse);
  }

  void restrictTo(hipStream_t s, const Xfer& X, const TV* f, TV* c) {
    spmv<TV, TV, false, false>(s, X.lanesT, X.nCoarse, X.trow.data(), X.tcol.data(), X.tval.data(), f, nullptr, c);
  }
  void prolongAdd(hipStream_t s, const Xfer& X, const TV* c, TV* f) {
    spmv<TV, TV, false, true>(s, X.lanesP, X.nFine, X.prow.data(), X.pcol.data(), X.pval.data(), c, nullptr, f);
  }

  static double chebRoot(double hi, double lo, int k, int degree) {
    const double centre = .5 * (hi + lo), radius = .5 * (hi - lo);
    const double pi = 3.141592653589793238462643383279502884;
    return centre - radius * std::cos(pi * (2.0 * k + 1.0) / (2.0 * degree));
  }

  // Generic smoother on an operator (row, col, val, diag, l1) from x = 0.
  template <class TM>
  void smooth(hipStream_t s, int n, int lanes, double meanRow, const int* row, const int* col, const TM* val, const TV* diag,
              const TV* l1, const DevColoring<TV>* color, double lambda, int jacobiSweeps, int mcgsSweeps,
              const TV* b, TV* x, TV* tmp, TV* r) {
    const int grid = blocksFor(n);
    if (opt_.kind == "jacobi" || opt_.kind == "l1jacobi") {
      const TV* den = opt_.kind == "l1jacobi" ? l1 : diag;
      hipLaunchKernelGGL(k_jacobi_zero<TV>, grid, 256, 0, s, n, opt_.jacobiOmega, b, den, x);
      NODALS_LAUNCHED();
      for (int sw = 1; sw < jacobiSweeps; ++sw) {
        spmv<TM, TV, true, false>(s, lanes, n, row, col, val, x, b, r);
        hipLaunchKernelGGL(k_jacobi_add<TV>, grid, 256, 0, s, n, opt_.jacobiOmega, r, den, x);
        NODALS_LAUNCHED();
      }
    } else if (opt_.kind == "cheb2") {
      const double lo = opt_.lambdaLowFraction * lambda;
      require(lambda > lo && lo > 0, "Chebyshev interval invalid");
      hipLaunchKernelGGL(k_jacobi_zero<TV>, grid, 256, 0, s, n, 1.0 / chebRoot(lambda, lo, 0, opt_.chebDegree), b,
                         diag, x);
      NODALS_LAUNCHED();
      for (int k = 1; k < opt_.chebDegree; ++k) {
        spmv<TM, TV, true, false>(s, lanes, n, row, col, val, x, b, r);
        hipLaunchKernelGGL(k_jacobi_add<TV>, grid, 256, 0, s, n, 1.0 / chebRoot(lambda, lo, k, opt_.chebDegree), r,
                           diag, x);
        NODALS_LAUNCHED();
      }
    } else if (opt_.kind == "mcgs") {
      fill<TV>(s, n, x, TV(0));
      for (int sw = 0; sw < mcgsSweeps; ++sw) {
        if (opt_.mcgsOrder == "forward" || opt_.mcgsOrder == "symmetric")
          colorPass(s, meanRow, row, col, val, diag, b, x, *color, false);
        if (opt_.mcgsOrder == "backward" || opt_.mcgsOrder == "symmetric")
          colorPass(s, meanRow, row, col, val, diag, b, x, *color, true);
      }
    } else {
      fail("unknown AMG smoother " + opt_.kind);
    }
    (void)tmp;
  }

  template <class TM>
  void colorPass(hipStream_t s, double meanRow, const int* row, const int* col, const TM* val, const TV* diag, const TV* b,
                 TV* x, const DevColoring<TV>& C, bool reverse) {
    for (int q = 0; q < C.ncolors; ++q) {
      const int c = reverse ? C.ncolors - 1 - q : q;
      const int begin = C.off[(std::size_t)c], count = C.off[(std::size_t)c + 1] - begin;
      mcgsColor<TM, TV>(s, lanesFor(meanRow, count), begin, count, C.rowList(), row, col, val, diag, b, x,
                        opt_.mcgsOmega);
    }
  }

  void smoothFine(hipStream_t s, const TV* b, TV* x) {
    smooth<TF>(s, fine_.n, fine_.lanes, fine_.meanRow, fine_.row, fine_.col, fine_.val, fine_.diag, fine_.l1, &fineColor_,
               fineLambda_, opt_.jacobiFineSweeps, opt_.mcgsFineSweeps, b, x, ftmp_.data(), ftmp_.data());
  }

  void smoothLevel(hipStream_t s, Level& A, const TV* b, TV* x) {
    if (A.fused) {
      const int order = opt_.mcgsOrder == "forward" ? 0 : opt_.mcgsOrder == "backward" ? 1 : 2;
      const std::size_t lds = fuseLdsBytes<TV>(A.n);
      static const bool attr = [] {
        NODALS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mcgs_level_wg<TV, TV>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFuseLdsBytes));
        return true;
      }();
      (void)attr;
      hipLaunchKernelGGL((k_mcgs_level_wg<TV, TV>), 1, 1024, lds, s, A.n, A.row.data(), A.col.data(), A.val.data(),
                         A.diag.data(), b, x, A.color.ncolors, A.colorOff.data(), A.colorLanes.data(),
                         opt_.mcgsCoarseSweeps, order, opt_.mcgsOmega);
      NODALS_LAUNCHED();
      return;
    }
    smooth<TV>(s, A.n, A.lanes, A.meanRow, A.row.data(), A.col.data(), A.val.data(), A.diag.data(), A.l1.data(), &A.color,
               A.lambda, opt_.jacobiCoarseSweeps, opt_.mcgsCoarseSweeps, b, x, A.tmp.data(), A.tmp.data());
  }

  void vcycle(hipStream_t s, int l, const TV* b, TV* x) {
    Level& A = L_[(std::size_t)l];
    if (l == (int)L_.size() - 1) {
      hipLaunchKernelGGL(k_dense_mv<TV>, blocksFor((long long)A.n * 64), 256, 0, s, A.n, tinv_.data(), b, x);
      NODALS_LAUNCHED();
      return;
    }
    if (fusedJacobi()) {
      const TV* den = opt_.kind == "l1jacobi" ? A.l1.data() : A.diag.data();
      jacobiPreResid<TV, TV>(s, A.lanes, A.n, A.row.data(), A.col.data(), A.val.data(), b, den, opt_.jacobiOmega,
                             A.tmp.data(), A.r.data());
      Level& C = L_[(std::size_t)l + 1];
      const Xfer& X = P_[(std::size_t)l + 1];
      restrictTo(s, X, A.r.data(), C.b.data());
      vcycle(s, l + 1, C.b.data(), C.x.data());
      if (X.aggregation) {
        residJacobiPostAgg<TV, TV>(s, A.lanes, A.n, A.row.data(), A.col.data(), A.val.data(), b, den,
                                   opt_.jacobiOmega, A.tmp.data(), X.pcol.data(), X.pval.data(), C.x.data(), x);
        return;
      }
      prolongAdd(s, X, C.x.data(), A.tmp.data());
      residJacobiPost<TV, TV>(s, A.lanes, A.n, A.row.data(), A.col.data(), A.val.data(), b, den, opt_.jacobiOmega,
                              A.tmp.data(), x);
      return;
    }
    smoothLevel(s, A, b, x);
    spmv<TV, TV, true, false>(s, A.lanes, A.n, A.row.data(), A.col.data(), A.val.data(), x, b, A.r.data());
    Level& C = L_[(std::size_t)l + 1];
    restrictTo(s, P_[(std::size_t)l + 1], A.r.data(), C.b.data());
    vcycle(s, l + 1, C.b.data(), C.x.data());
    prolongAdd(s, P_[(std::size_t)l + 1], C.x.data(), x);
    spmv<TV, TV, true, false>(s, A.lanes, A.n, A.row.data(), A.col.data(), A.val.data(), x, b, A.r.data());
    smoothLevel(s, A, A.r.data(), A.corr.data());
    axpy<TV>(s, A.n, 1.0, A.corr.data(), x);
  }

  // One Jacobi sweep before and after on every level: the fused kernels.
  bool fusedJacobi() const {
    return (opt_.kind == "jacobi" || opt_.kind == "l1jacobi") && opt_.jacobiFineSweeps == 1 &&
           opt_.jacobiCoarseSweeps == 1;
  }

  SmootherOptions opt_;
  int fineN_ = 0;
  double fineLambda_ = 1.0;
  FineView<TF, TV> fine_;
  DevColoring<TV> fineColor_;
  std::vector<Level> L_;
  std::vector<Xfer> P_;
  DBuf<TV> tinv_;
  int termN_ = 0;
  DBuf<TV> fr_, ftmp_, fcorr_;
  GraphCache graph_;
  GraphSet wide_;
};

}  // namespace dev
}  // namespace nodals
