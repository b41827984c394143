// Wall-function momentum path of an H8 context (see h8_solver.hpp and
// wall_kernels.hpp): the k-omega SST SIMPLE iteration of the reference's
// spiral case (turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
// modules/app/sst_application.inc:2818-2978 with simple_momentum_assembly.inc,
// simple_pressure_predictor.inc and simple_pressure_corrector.inc).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "h8_kernels.hpp"
#include "h8_solver.hpp"
#include "turb_step.hpp"
#include "wall_kernels.hpp"
#include "wall_state.hpp"

namespace nodals {

using namespace dev;

WallParams parseWallArgs(int argc, const char* const* argv) {
  WallParams P;
  auto num = [](const std::string& flag, const char* v) {
    char* end = nullptr;
    const double x = std::strtod(v, &end);
    require(end && *end == '\0' && std::isfinite(x), "invalid value for " + flag + ": " + v);
    return x;
  };
  for (int i = 0; i < argc; ++i) {
    const std::string f = argv[i];
    require(f.rfind("--", 0) == 0 && i + 1 < argc, "expected --flag value, got " + f);
    const char* v = argv[++i];
    const std::string sv = v;
    if (f == "--gamma") P.gammaN = num(f, v);
    else if (f == "--wall-active-points") P.activePoints = (int)num(f, v);
    else if (f == "--wall-zero-limit") P.zeroLimit = num(f, v) != 0.0;
    else if (f == "--convection-form") {
      require(sv == "skew_symmetric" || sv == "advective", "--convection-form must be skew_symmetric or advective");
      P.skew = sv == "skew_symmetric";
    } else if (f == "--stress-form") {
      require(sv == "deviatoric" || sv == "vector_laplacian", "--stress-form must be deviatoric or vector_laplacian");
      P.deviatoric = sv == "deviatoric";
    } else if (f == "--m3-mode") {
      require(sv == "fgs" || sv == "sgs", "--m3-mode must be fgs or sgs");
      P.m3Symmetric = sv == "sgs";
    } else if (f == "--m3-sweeps") P.m3Sweeps = (int)num(f, v);
    else if (f == "--m3-omega") P.m3Omega = num(f, v);
    else if (f == "--initial-velocity") {
      require(sv == "potential" || sv == "plug", "--initial-velocity must be potential or plug");
      P.potential = sv == "potential";
    } else if (f == "--potential-tol") P.potentialTol = num(f, v);
    else fail("unknown wall option " + f);
  }
  require(P.gammaN >= 0.0, "--gamma must be >= 0");
  require(P.activePoints >= 1 && P.activePoints <= 5, "--wall-active-points must be in 1..5");
  require(P.m3Sweeps >= 1 && P.m3Sweeps <= 60, "--m3-sweeps must be in 1..60");
  require(P.m3Omega > 0.0 && P.m3Omega < 2.0, "--m3-omega must be in (0, 2)");
  require(P.potentialTol > 0.0 && P.potentialTol < 1.0, "--potential-tol must be in (0, 1)");
  return P;
}

void H8Solver::wallInit(const WallParams& P) {
  require(S_.wallFunction, "the wall-function path needs a context created with --wall-model spalding", kErrInvalid);
  require((bool)turb_, "the wall-function path needs the k-omega model (nodals_h8_turb_init) first", kErrInvalid);
  require(turb_->P.wall, "the wall-function path needs the k-omega wall faces (--sst-omega-wall-mode of_auto)",
          kErrInvalid);
  {
    // ray-traced placement covers only the slots that were traced
    const unsigned fitMask = sst::activeMask(P.activePoints);
    require((fitMask & ~turb_->T.rayMask) == 0,
            "the Spalding fit uses wall samples that were not ray-traced: give the k-omega setup the same "
            "active point count (--sst-wall-active-points)",
            kErrInvalid);
  }
  turb_->omegaFitPoints = P.activePoints;
  auto st = std::make_unique<WallState>();
  st->P = P;
  st->W = buildWallTopo(M_, S_, turb_->T);
  // Initial velocity of the free DOFs: the potential-flow start or the inlet
  // plug (the inlet velocity on every free vertex, zero on the face bubbles:
  // plug_initial, hxt5b_boundary.cuh:698-700).  Fixed DOFs keep their
  // Dirichlet values.
  st->u0.assign(3 * (std::size_t)nv_, 0.0);
  st->startReport.assign(7, 0.0);
  {
    const int np = M_.nPoints(), nent = np + M_.nFaces();
    if (P.potential) {
      const PotentialStart ps = potentialStart(M_, S_, P.potentialTol);
      for (int e = 0; e < nent; ++e) {
        const int g = S_.g2free[(std::size_t)e];
        if (g < 0) continue;
        for (int d = 0; d < 3; ++d) st->u0[3 * (std::size_t)g + d] = ps.value[3 * (std::size_t)e + d];
      }
      st->startReport = {(double)ps.iterations, ps.rel, ps.inletFlux, ps.outletFlux, ps.imbalanceRel, ps.wallFluxMax,
                         ps.maxSpeed};
    } else {
      for (int v = 0; v < np; ++v) {
        const int g = S_.g2free[(std::size_t)v];
        if (g < 0) continue;
        for (int d = 0; d < 3; ++d) st->u0[3 * (std::size_t)g + d] = S_.pipe.inletVelocity[d];
      }
    }
  }
  if (!hasDevice()) {
    wall_ = std::move(st);
    return;
  }
  require(turb_->vdistPtr() != nullptr, "the wall-function path needs the duct's wall-distance field", kErrInternal);
  WallState& w = *st;
  const WallTopo& W = w.W;
  const int nf = W.nf;
  auto dv = [&](int r) { return pinv_[(std::size_t)r]; };
  w.wcell.upload(W.cell);
  w.wlf.upload(W.lf);
  w.wn.upload(W.n);
  w.wh.upload(W.h);
  w.qw.upload(W.qw);
  w.ql.upload(W.ql);
  w.bw.upload(W.bw);
  w.bp.upload(W.bp);
  w.heights.upload(std::vector<double>(turb_->T.heights.begin(), turb_->T.heights.end()));
  std::vector<int> bub((std::size_t)nf), vert((std::size_t)nf * 3);
  for (int f = 0; f < nf; ++f) {
    bub[(std::size_t)f] = dv(W.bub[(std::size_t)f]);
    for (int a = 0; a < 3; ++a) vert[(std::size_t)f * 3 + a] = dv(W.vert[(std::size_t)f * 3 + a]);
  }
  w.bub.upload(bub);
  w.vert.upload(vert);
  std::vector<int> dof(W.dof.size()), bubFace((std::size_t)nv_, -1), wallDof((std::size_t)nv_, -1);
  for (std::size_t t = 0; t < W.dof.size(); ++t) {
    dof[t] = dv(W.dof[t]);
    wallDof[(std::size_t)dof[t]] = (int)t;
    if (W.dofIsBubble[t]) bubFace[(std::size_t)dof[t]] = W.dofFace[(std::size_t)W.dofOff[t]];
  }
  w.dof.upload(dof);
  w.dofOff.upload(W.dofOff);
  w.dofFace.upload(W.dofFace);
  w.isBub.upload(W.dofIsBubble);
  w.bubFace.upload(bubFace);
  w.wallDof.upload(wallDof);
  w.cellWallOff.upload(W.cellWallOff);
  w.cellWall.upload(W.cellWall.empty() ? std::vector<int>{0} : W.cellWall);
  // wall-block gather in device entry order
  {
    std::vector<long long> oldOfNew((std::size_t)momNnz_);
    for (long long k = 0; k < momNnz_; ++k) oldOfNew[(std::size_t)newOfOld_[(std::size_t)k]] = k;
    std::vector<int> off((std::size_t)momNnz_ + 1, 0), idx;
    idx.reserve(W.blkIdx.size());
    for (long long nk = 0; nk < momNnz_; ++nk) {
      const long long k = oldOfNew[(std::size_t)nk];
      for (int t = W.blkOff[(std::size_t)k]; t < W.blkOff[(std::size_t)k + 1]; ++t) idx.push_back(W.blkIdx[(std::size_t)t]);
      off[(std::size_t)nk + 1] = (int)idx.size();
    }
    if (idx.empty()) idx.push_back(0);
    w.blkOff.upload(off);
    w.blkIdx.upload(idx);
  }
  auto mapPos = [&](const std::vector<std::int32_t>& v) {
    std::vector<int> o(v.size());
    for (std::size_t i = 0; i < v.size(); ++i) o[i] = (int)newOfOld_[(std::size_t)v[i]];
    if (o.empty()) o.push_back(0);
    return o;
  };
  auto nonEmpty = [](std::vector<int> v) {
    if (v.empty()) v.push_back(0);
    return v;
  };
  w.nRSrc = (int)W.rSrcPos.size();
  w.nRTgt = (int)W.rTgtPos.size();
  w.nLSrc = (int)W.lSrcPos.size();
  w.nLTgt = (int)W.lTgtPos.size();
  w.rSrcFace.upload(nonEmpty(W.rSrcFace));
  w.rSrcPos.upload(mapPos(W.rSrcPos));
  w.rTgtPos.upload(mapPos(W.rTgtPos));
  w.rTgtOff.upload(nonEmpty(W.rTgtOff));
  w.rTgtSrc.upload(nonEmpty(W.rTgtSrc));
  w.lSrcFace.upload(nonEmpty(W.lSrcFace));
  w.lSrcPos.upload(mapPos(W.lSrcPos));
  w.lTgtPos.upload(mapPos(W.lTgtPos));
  w.lTgtOff.upload(nonEmpty(W.lTgtOff));
  w.lTgtSrc.upload(nonEmpty(W.lTgtSrc));
  const std::size_t nf1 = (std::size_t)std::max(nf, 1);
  w.fit.alloc(3 * nf1);
  w.blk.alloc(nf1 * 64 * 3);
  w.pen.alloc(nf1 * 64 * 3);
  w.cpl.alloc(nf1 * 64);
  w.nuq.alloc(nf1 * 12);
  w.bn.alloc(nf1);
  w.shear.alloc(nf1 * 4);
  w.K.alloc((std::size_t)momNnz_ * 9);
  w.D.alloc((std::size_t)std::max(1, std::max(w.nRSrc, w.nLSrc)) * 9);
  for (DBuf<double>* b : {&w.rau3, &w.r3, &w.delta3, &w.diag3, &w.y3, &w.uPrev}) b->alloc(3 * (std::size_t)nv_);
  w.kPrev.alloc((std::size_t)turb_->T.nS);
  w.wPrev.alloc((std::size_t)turb_->T.nS);
  if (P.deviatoric) w.dbuf.alloc((std::size_t)nc_ * 64 * 9);
  w.counters.alloc(4);
  // convection tensor of the momentum: skew form T[a][m][b][j] + T[a][b][m][j] / 2
  if (P.skew) {
    const Tensors& T = referenceTensors();
    std::vector<double> sk((std::size_t)8 * 8 * 8 * 3);
    for (int a = 0; a < 8; ++a)
      for (int m = 0; m < 8; ++m)
        for (int b = 0; b < 8; ++b)
          for (int j = 0; j < 3; ++j)
            sk[(((std::size_t)a * 8 + m) * 8 + b) * 3 + j] = T.cent[a][m][b][j] + 0.5 * T.cent[a][b][m][j];
    w.momT.upload(sk);
    momT_ = w.momT.data();
  } else {
    momT_ = nullptr;
  }
  rau3_ = w.rau3.data();
  {
    std::vector<double> u0((std::size_t)nv_ * 3, 0.0);
    for (int g = 0; g < nv_; ++g)
      for (int d = 0; d < 3; ++d) u0[3 * (std::size_t)dv(g) + d] = w.u0[3 * (std::size_t)g + d];
    u_.upload(u0);
  }
  wall_ = std::move(st);
}

int H8Solver::wallSweeps() const { return wall_ ? wall_->P.m3Sweeps * (wall_->P.m3Symmetric ? 2 : 1) : 0; }

int H8Solver::wallCounters(bool sweep) {
  std::vector<int> c = wall_->counters.host();
  require(c[0] == 0, "wall-function fit failed on " + std::to_string(c[0]) + " wall faces", kErrNonFinite);
  require(c[1] == 0, "wall-function effective viscosity non-finite at " + std::to_string(c[1]) + " points",
          kErrNonFinite);
  require(c[2] == 0, "wall-function momentum rows with a non-positive relaxed diagonal: " + std::to_string(c[2]),
          kErrNonFinite);
  if (sweep)
    require(c[3] == 0, "M3 block solves failed: " + std::to_string(c[3]), kErrNonFinite);
  return 0;
}

void H8Solver::wallAssemble() {
  requireDevice();
  require((bool)wall_, "wall-function path not initialised (nodals_h8_wall_init)", kErrInvalid);
  WallState& w = *wall_;
  TurbState& st = *turb_;
  const TurbParams& TP = st.P;
  const int nf = w.W.nf;
  hipStream_t s = stream_;
  hipLaunchKernelGGL(k_fill<int>, 1, 64, 0, s, 4ll, w.counters.data(), 0);
  NODALS_LAUNCHED();
  if (nf > 0) {  // the velocity enters on the constrained (physical) space
    hipLaunchKernelGGL(k_wall_t, blocksFor(nf), 256, 0, s, nf, w.bub.data(), w.vert.data(), w.wn.data(), u_.data(), 0);
    NODALS_LAUNCHED();
  }
  // convection (skew or advective) and the eddy-viscosity blocks, plus the
  // deviatoric tensors
  convLocal(s, 8192);
  hipLaunchKernelGGL(k_sst_volume, (nc_ + 3) / 4, 256, 0, s, nc_, cref_.data(), cinvJ_.data(), cdet_.data(),
                     sstXY_.data(), u_.data(), fixedVal_.data(), st.k.data(), st.k.data(), st.w.data(), st.w.data(),
                     sstQuad_.data(), S_.pipe.R, S_.pipe.nu, 2, TP.kFloor, TP.omegaFloor, TP.deviatoric ? 1 : 0,
                     cbuf_.data(), nullptr, st.sref.data(), 0.0, cbuf_.data(), st.vdistPtr(),
                     w.P.deviatoric ? w.dbuf.data() : nullptr);
  NODALS_LAUNCHED();
  if (nf > 0) {
    hipLaunchKernelGGL(k_wall_fit, blocksFor(nf), 256, 0, s, nf, w.wn.data(), w.qw.data(), w.bp.data(),
                       w.wcell.data(), cref_.data(), u_.data(), fixedVal_.data(), st.scell.data(), st.sphi.data(),
                       w.heights.data(), st.sy.data(), S_.pipe.nu, w.P.activePoints, w.P.zeroLimit ? 1 : 0, w.fit.data(),
                       w.counters.data());
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_wall_nitsche, (nf + 3) / 4, 256, 0, s, nf, w.wcell.data(), w.wn.data(), w.wh.data(),
                       w.qw.data(), w.ql.data(), w.bw.data(), w.fit.data(), cref_.data(), cinvJ_.data(), u_.data(),
                       fixedVal_.data(), st.sref.data(), st.k.data(), st.w.data(), st.vdistPtr(), S_.pipe.R,
                       S_.pipe.nu, TP.kFloor, TP.omegaFloor, TP.deviatoric ? 1 : 0, w.P.gammaN, w.blk.data(),
                       w.pen.data(), w.cpl.data(), w.nuq.data(), w.counters.data() + 1);
    NODALS_LAUNCHED();
  }
  WallRowsArgs a;
  a.mrow = mrow_.data();
  a.mdiagPos = mdiagPos_.data();
  a.mdiff = mdiff_.data();
  a.termOff = termOff_.data();
  a.termIdx = termIdx_.data();
  a.cbuf = cbuf_.data();
  a.dbuf = w.P.deviatoric ? w.dbuf.data() : nullptr;
  a.blkOff = w.blkOff.data();
  a.blkIdx = w.blkIdx.data();
  a.wblk = w.blk.data();
  a.wpen = w.pen.data();
  a.wcpl = w.cpl.data();
  a.wn = w.wn.data();
  a.fixOff = fixOff_.data();
  a.fixTerm = fixTerm_.data();
  a.fixedVal = fixedVal_.data();
  a.incOff = incOff_.data();
  a.incPacked = incPacked_.data();
  a.bcoef = bcoef_.data();
  a.p = p_.data();
  a.srhs = srhs_.data();
  a.u = u_.data();
  a.fac = 1.0 / opt_.alphaU - 1.0;
  a.rauScale = opt_.rauScale;
  a.K = w.K.data();
  a.brhs = brhs_.data();
  a.rau3 = w.rau3.data();
  a.r3 = w.r3.data();
  a.delta3 = w.delta3.data();
  a.diag3 = w.diag3.data();
  a.bad = w.counters.data() + 2;
  const int nbL = blocksFor((long long)nLong_ * kLongLanes), nbS = blocksFor((long long)nShort_ * kShortLanes);
  hipLaunchKernelGGL((k_wall_rows<kLongLanes, kShortLanes>), nbL + nbS, 256, 0, s, a, nLong_, longRows_.data(),
                     nShort_, shortRows_.data(), nbL);
  NODALS_LAUNCHED();
  // K = T^T M T, then T^T on the right-hand side
  if (w.nRSrc > 0) {
    hipLaunchKernelGGL(k_m3_source, blocksFor(w.nRSrc), 256, 0, s, w.nRSrc, w.rSrcFace.data(), w.rSrcPos.data(),
                       w.wn.data(), w.K.data(), 0, w.D.data());
    hipLaunchKernelGGL(k_m3_apply, blocksFor(w.nRTgt), 256, 0, s, w.nRTgt, w.rTgtPos.data(), w.rTgtOff.data(),
                       w.rTgtSrc.data(), w.D.data(), w.K.data());
    hipLaunchKernelGGL(k_m3_source, blocksFor(w.nLSrc), 256, 0, s, w.nLSrc, w.lSrcFace.data(), w.lSrcPos.data(),
                       w.wn.data(), w.K.data(), 1, w.D.data());
    hipLaunchKernelGGL(k_m3_apply, blocksFor(w.nLTgt), 256, 0, s, w.nLTgt, w.lTgtPos.data(), w.lTgtOff.data(),
                       w.lTgtSrc.data(), w.D.data(), w.K.data());
    NODALS_LAUNCHED();
  }
  if (nf > 0) {
    hipLaunchKernelGGL(k_wall_bn, blocksFor(nf), 256, 0, s, nf, w.bub.data(), w.wn.data(), brhs_.data(), w.bn.data());
    hipLaunchKernelGGL(k_wall_tt, blocksFor((long long)w.W.dof.size()), 256, 0, s, (int)w.W.dof.size(), w.dof.data(),
                       w.dofOff.data(), w.dofFace.data(), w.isBub.data(), w.wn.data(), w.bn.data(), brhs_.data());
    NODALS_LAUNCHED();
  }
}

void H8Solver::wallMomentum(int sweeps, double omega, int symmetric, double* rel0, double* rel1) {
  requireDevice();
  require((bool)wall_, "wall-function path not initialised (nodals_h8_wall_init)", kErrInvalid);
  WallState& w = *wall_;
  const int nf = w.W.nf;
  hipStream_t s = stream_;
  const int nSweeps = sweeps >= 0 ? sweeps : w.P.m3Sweeps;
  const double om = omega > 0.0 ? omega : w.P.m3Omega;
  const bool sym = symmetric >= 0 ? symmetric != 0 : w.P.m3Symmetric;
  if (nf > 0) {  // Q: the parameter vector carries no bubble-normal part
    hipLaunchKernelGGL(k_wall_t, blocksFor(nf), 256, 0, s, nf, w.bub.data(), w.vert.data(), w.wn.data(), u_.data(), 1);
    NODALS_LAUNCHED();
  }
  auto relResid = [&]() {
    hipLaunchKernelGGL(k_m3_resid, blocksFor(nv_), 256, 0, s, nv_, mrow_.data(), mcol_.data(), w.K.data(),
                       brhs_.data(), u_.data(), w.y3.data());
    NODALS_LAUNCHED();
    const double rr = red_.dot(s, 3 * nv_, w.y3.data(), w.y3.data());
    const double bb = red_.dot(s, 3 * nv_, brhs_.data(), brhs_.data());
    return std::sqrt(rr) / std::max(std::sqrt(bb), 1e-300);
  };
  if (rel0) *rel0 = relResid();
  const int ncol = (int)colorOff_.size() - 1;
  auto pass = [&](bool backward) {
    for (int cc = 0; cc < ncol; ++cc) {
      const int c = backward ? ncol - 1 - cc : cc;
      const int r0 = colorOff_[(std::size_t)c], r1 = colorOff_[(std::size_t)c + 1];
      if (r1 <= r0) continue;
      const int rs = colorSplit_[(std::size_t)c];
      const int nbL = blocksFor((long long)(rs - r0) * kLongLanes), nbS = blocksFor((long long)(r1 - rs) * kShortLanes);
      if (probeOn_ && probeKind_ == 2) {  // the launch probe times each colour from its own dispatch packet
        hipExtLaunchKernelGGL((k_m3_gs<kLongLanes, kShortLanes>), dim3(nbL + nbS), dim3(256), 0, s,
                              probeEv_[2 * (std::size_t)probeUsed_], probeEv_[2 * (std::size_t)probeUsed_ + 1], 0,
                              r0, rs, r1, nbL, mrow_.data(), mcol_.data(), mdiagPos_.data(), w.K.data(),
                              brhs_.data(), u_.data(), w.bubFace.data(), w.wn.data(), om, w.counters.data() + 3);
        NODALS_LAUNCHED();
        if (++probeUsed_ == kProbePairs) probeHarvest();
      } else {
        hipLaunchKernelGGL((k_m3_gs<kLongLanes, kShortLanes>), nbL + nbS, 256, 0, s, r0, rs, r1, nbL, mrow_.data(),
                           mcol_.data(), mdiagPos_.data(), w.K.data(), brhs_.data(), u_.data(), w.bubFace.data(),
                           w.wn.data(), om, w.counters.data() + 3);
        NODALS_LAUNCHED();
      }
    }
  };
  for (int k = 0; k < nSweeps; ++k) {
    pass(false);
    if (sym) pass(true);
  }
  if (rel1) *rel1 = relResid();
  if (nf > 0) {  // T: back to the physical velocity
    hipLaunchKernelGGL(k_wall_t, blocksFor(nf), 256, 0, s, nf, w.bub.data(), w.vert.data(), w.wn.data(), u_.data(), 0);
    NODALS_LAUNCHED();
  }
}

void H8Solver::wallSchur(hipStream_t s, const double* x, double* y) {
  WallState& w = *wall_;
  const int nf = w.W.nf;
  if (nf > 0) {
    hipLaunchKernelGGL(k_i3_bn, blocksFor(nf), 256, 0, s, nf, w.bub.data(), w.wn.data(), incOff_.data(),
                       incPacked_.data(), bcoef_.data(), x, w.bn.data());
    NODALS_LAUNCHED();
  }
  hipLaunchKernelGGL(k_i3_dof, blocksFor(nv_), 256, 0, s, nv_, incOff_.data(), incPacked_.data(), bcoef_.data(), x,
                     w.r3.data(), w.wallDof.data(), w.dofOff.data(), w.dofFace.data(), w.isBub.data(), w.wn.data(),
                     w.bn.data(), w.y3.data());
  NODALS_LAUNCHED();
  hipLaunchKernelGGL(k_i3_cell, blocksFor(nc_), 256, 0, s, nc_, cref_.data(), bcoef_.data(), w.y3.data(),
                     w.cellWallOff.data(), w.cellWall.data(), w.wlf.data(), w.bub.data(), w.vert.data(), w.wn.data(),
                     y);
  NODALS_LAUNCHED();
}

void H8Solver::wallCorrect(const double* dp) {
  requireDevice();
  require((bool)wall_, "wall-function path not initialised (nodals_h8_wall_init)", kErrInvalid);
  WallState& w = *wall_;
  const int nf = w.W.nf;
  hipStream_t s = stream_;
  if (nf > 0) {
    hipLaunchKernelGGL(k_i3_bn, blocksFor(nf), 256, 0, s, nf, w.bub.data(), w.wn.data(), incOff_.data(),
                       incPacked_.data(), bcoef_.data(), dp, w.bn.data());
    NODALS_LAUNCHED();
  }
  hipLaunchKernelGGL(k_i3_dof, blocksFor(nv_), 256, 0, s, nv_, incOff_.data(), incPacked_.data(), bcoef_.data(), dp,
                     w.r3.data(), w.wallDof.data(), w.dofOff.data(), w.dofFace.data(), w.isBub.data(), w.wn.data(),
                     w.bn.data(), w.y3.data());
  NODALS_LAUNCHED();
  if (nf > 0) {
    hipLaunchKernelGGL(k_wall_t, blocksFor(nf), 256, 0, s, nf, w.bub.data(), w.vert.data(), w.wn.data(), w.y3.data(),
                       0);
    NODALS_LAUNCHED();
  }
  axpy<double>(s, 3 * nv_, 1.0, w.y3.data(), u_.data());
}

std::vector<double> H8Solver::wallField(const std::string& n) {
  require((bool)wall_, "wall-function path not initialised (nodals_h8_wall_init)", kErrInvalid);
  WallState& w = *wall_;
  const WallTopo& W = w.W;
  if (n == "wall_normal") return W.n;
  if (n == "wall_area") return W.area;
  if (n == "wall_h") return W.h;
  if (n == "wall_bw") return W.bw;
  if (n == "wall_bp") return W.bp;
  if (n == "wall_qw") return W.qw;
  if (n == "wall_ql") return W.ql;
  if (n == "wall_cell") return std::vector<double>(W.cell.begin(), W.cell.end());
  if (n == "wall_lf") return std::vector<double>(W.lf.begin(), W.lf.end());
  if (n == "wall_bub") return std::vector<double>(W.bub.begin(), W.bub.end());
  if (n == "wall_vert") return std::vector<double>(W.vert.begin(), W.vert.end());
  if (n == "wall_u0") return w.u0;
  if (n == "wall_start") return w.startReport;
  requireDevice();
  sync();
  auto perDof = [&](const DBuf<double>& b) {  // device [nv][3] -> reference free order
    const std::vector<double> d = b.host();
    std::vector<double> o(d.size());
    for (int i = 0; i < nv_; ++i)
      for (int c = 0; c < 3; ++c) o[3 * (std::size_t)perm_[(std::size_t)i] + c] = d[3 * (std::size_t)i + c];
    return o;
  };
  if (n == "wall_K") {
    const std::vector<double> d = w.K.host();
    std::vector<double> o(d.size());
    for (long long k = 0; k < momNnz_; ++k)
      for (int j = 0; j < 9; ++j) o[9 * (std::size_t)k + j] = d[9 * (std::size_t)newOfOld_[(std::size_t)k] + j];
    return o;
  }
  if (n == "wall_rhs") return perDof(brhs_);
  if (n == "wall_rau3") return perDof(w.rau3);
  if (n == "wall_r3") return perDof(w.r3);
  if (n == "wall_delta3") return perDof(w.delta3);
  if (n == "wall_fit") {
    std::vector<double> v = w.fit.host();
    v.resize(3 * (std::size_t)W.nf);
    return v;
  }
  if (n == "wall_nuq") {
    std::vector<double> v = w.nuq.host();
    v.resize(12 * (std::size_t)W.nf);
    return v;
  }
  if (n == "wall_shear") {
    if (W.nf > 0) {
      hipLaunchKernelGGL(k_wall_shear, blocksFor(W.nf), 256, 0, stream_, W.nf, w.wcell.data(), w.wn.data(),
                         w.qw.data(), w.bp.data(), w.fit.data(), cref_.data(), u_.data(), fixedVal_.data(),
                         w.shear.data());
      NODALS_LAUNCHED();
      sync();  // the kernel runs on the context stream, the copy below does not
    }
    std::vector<double> v = w.shear.host();
    v.resize(4 * (std::size_t)W.nf);
    return v;
  }
  fail("unknown wall field " + n);
  return {};
}

// ------------------------------------------------------------------ loop
namespace {
int lastItOf(int firstIt, int count) { return firstIt + count - 1; }
}  // namespace

H8RunResult H8Solver::runWindow(int firstIt, int count) {
  requireDevice();
  require((bool)wall_, "run_window needs the wall-function path (nodals_h8_wall_init)", kErrInvalid);
  require(firstIt >= 1 && count >= 1, "run_window: first iteration and count must be >= 1", kErrInvalid);
  return runWall(firstIt, count);
}

H8RunResult H8Solver::runWall(int firstIt, int count) {
  const H8Options& o = opt_;
  WallState& w = *wall_;
  TurbState& st = *turb_;
  H8RunResult R;
  R.setupMs = setupMs_;
  const bool verbose = o.verbose > 0;
  const int nS = st.T.nS;
  hipEvent_t ev[6];
  for (auto& e : ev) NODALS_HIP(hipEventCreate(&e));
  auto destroyEvents = [&]() {
    for (auto& e : ev) (void)hipEventDestroy(e);
  };
  DevOp A = [this](hipStream_t s, const double* x, double* y) { wallSchur(s, x, y); };
  DevOp Mp = [this](hipStream_t s, const double* r, double* z) { precondApply(s, r, z); };
  double cont0 = 0.0, contRel = 1.0;
  const auto t0 = std::chrono::steady_clock::now();
  try {
    const int lastIt = firstIt + count - 1;
    for (int it = firstIt; it <= lastIt; ++it) {
      NODALS_HIP(hipMemcpyAsync(w.uPrev.data(), u_.data(), sizeof(double) * 3 * (std::size_t)nv_,
                                hipMemcpyDeviceToDevice, stream_));
      NODALS_HIP(hipEventRecord(ev[0], stream_));
      wallAssemble();
      NODALS_HIP(hipEventRecord(ev[1], stream_));
      wallMomentum();
      NODALS_HIP(hipEventRecord(ev[2], stream_));
      const double cn = continuity();
      refreshOn(stream_);
      NODALS_HIP(hipEventRecord(ev[3], stream_));
      wallCounters(true);
      KrylovResult pr;
      if (o.pressureSolver == "fgmres")
        pr = fgmres(stream_, red_, nc_, A, Mp, cont_.data(), o.pRtol, o.pAtol, o.pMax, o.pressureRestart, kw_,
                    nullptr, nullptr);
      else
        pr = pcg(stream_, red_, nc_, A, Mp, cont_.data(), o.pRtol, o.pAtol, o.pMax, kw_);
      require(pr.ok, "SST wall-function pressure solve failed to reach its target", kErrNotConverged);
      R.pressureIts += pr.its;
      axpy<double>(stream_, nc_, o.alphaP, kw_.x.data(), p_.data());
      wallCorrect(kw_.x.data());
      const double cnPost = continuity();
      NODALS_HIP(hipEventRecord(ev[4], stream_));
      NODALS_HIP(hipMemcpyAsync(w.kPrev.data(), st.k.data(), sizeof(double) * nS, hipMemcpyDeviceToDevice, stream_));
      NODALS_HIP(hipMemcpyAsync(w.wPrev.data(), st.w.data(), sizeof(double) * nS, hipMemcpyDeviceToDevice, stream_));
      turbAdvance();
      NODALS_HIP(hipEventRecord(ev[5], stream_));
      TurbStepResult tr;
      turbCounters(&tr);
      // relative changes of u, k and omega over the iteration
      axpy<double>(stream_, 3 * nv_, -1.0, u_.data(), w.uPrev.data());
      axpy<double>(stream_, nS, -1.0, st.k.data(), w.kPrev.data());
      axpy<double>(stream_, nS, -1.0, st.w.data(), w.wPrev.data());
      auto relOf = [&](int n, const double* d, const double* x) {
        return std::sqrt(red_.dot(stream_, n, d, d)) / std::max(std::sqrt(red_.dot(stream_, n, x, x)), 1e-300);
      };
      const double du = relOf(3 * nv_, w.uPrev.data(), u_.data());
      const double dk = relOf(nS, w.kPrev.data(), st.k.data());
      const double dw = relOf(nS, w.wPrev.data(), st.w.data());
      float ms[5];
      for (int j = 0; j < 5; ++j) NODALS_HIP(hipEventElapsedTime(&ms[j], ev[j], ev[j + 1]));
      R.assemblyMs += ms[0];
      R.momentumMs += ms[1];
      R.continuityMs += ms[2];
      R.pressureMs += ms[3];
      R.turbMs += ms[4];
      if (it == firstIt) cont0 = std::max(cnPost, 1e-300);
      contRel = cnPost / cont0;
      R.outer = it;
      R.finalRelCont = contRel;
      // the relative gate, or the SST application's gate (mass: absolute
      // continuity; turbulence: relative k and omega changes)
      const bool done = o.sstContAtol > 0.0
                            ? (cnPost <= o.sstContAtol && dk <= o.simpleTol && dw <= o.simpleTol)
                            : (it > firstIt && contRel <= o.simpleTol && du <= o.simpleTol && dk <= o.simpleTol &&
                               dw <= o.simpleTol);
      R.finalDu = du;
      R.finalDk = dk;
      R.finalDw = dw;
      R.finalCont = cnPost;
      if (verbose && (it == firstIt || it % o.printEvery == 0 || done || it == lastIt)) {
        std::printf("NODALS_SST_RESIDUALS it=%d contPre=%.6e contPost=%.6e contRel=%.6e dU=%.3e dK=%.3e "
                    "dOmega=%.3e pIts=%d kClip=%d omegaClip=%d status=PASS\n",
                    it, cn, cnPost, contRel, du, dk, dw, pr.its, tr.kClips, tr.omegaClips);
        std::fflush(stdout);
      }
      if (done) {
        R.converged = 1;
        break;
      }
    }
  } catch (...) {
    destroyEvents();
    throw;
  }
  destroyEvents();
  R.loopMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  R.avgPressureIts = R.outer >= firstIt ? (double)R.pressureIts / (R.outer - firstIt + 1) : 0.0;
  R.pressureDrop = pressureDropPatches(S_, p_.host());
  R.pass = (o.runMode == "fixed10") ? (R.outer == lastItOf(firstIt, count) || R.converged) : R.converged;
  return R;
}

}  // namespace nodals

namespace nodals {

void H8Solver::writeVtu(const std::string& path) {
  requireDevice();
  sync();
  const int np = M_.nPoints(), nc = M_.nCells();
  const std::vector<double> ud = u_.host(), pd = p_.host();
  std::vector<double> uv((std::size_t)np * 3, 0.0);
  for (int v = 0; v < np; ++v) {
    const int g = S_.g2free[(std::size_t)v];
    for (int d = 0; d < 3; ++d) {
      if (g >= 0) uv[3 * (std::size_t)v + d] = ud[3 * (std::size_t)pinv_[(std::size_t)g] + d];
      else if (S_.fixedSlot[(std::size_t)v] >= 0)
        uv[3 * (std::size_t)v + d] = S_.fixedValue[(std::size_t)S_.fixedSlot[(std::size_t)v]][(std::size_t)d];
    }
  }
  std::vector<double> kv, wv;
  if (turb_) {
    kv = turb_->k.host();
    wv = turb_->w.host();
  }
  std::FILE* f = std::fopen(path.c_str(), "w");
  require(f != nullptr, "cannot write VTU file " + path);
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n<Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n",
               np, nc);
  std::fprintf(f, "<PointData Vectors=\"U\">\n<DataArray type=\"Float64\" Name=\"U\" NumberOfComponents=\"3\" "
                  "format=\"ascii\">\n");
  for (int v = 0; v < np; ++v)
    std::fprintf(f, "%.17g %.17g %.17g\n", uv[3 * (std::size_t)v], uv[3 * (std::size_t)v + 1],
                 uv[3 * (std::size_t)v + 2]);
  std::fprintf(f, "</DataArray>\n");
  auto scalar = [&](const char* name, const std::vector<double>& a) {
    std::fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", name);
    for (int v = 0; v < np; ++v) std::fprintf(f, "%.17g\n", a[(std::size_t)v]);
    std::fprintf(f, "</DataArray>\n");
  };
  if (turb_) {
    scalar("k", kv);
    scalar("omega", wv);
  }
  if (!S_.wallDistance.empty()) scalar("wallDistance", S_.wallDistance);
  std::fprintf(f, "</PointData>\n<CellData Scalars=\"p\">\n<DataArray type=\"Float64\" Name=\"p\" format=\"ascii\">\n");
  for (int c = 0; c < nc; ++c) std::fprintf(f, "%.17g\n", pd[(std::size_t)c]);
  std::fprintf(f, "</DataArray>\n</CellData>\n<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" "
                  "format=\"ascii\">\n");
  for (const auto& x : M_.points) std::fprintf(f, "%.17g %.17g %.17g\n", x[0], x[1], x[2]);
  std::fprintf(f, "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"Int32\" Name=\"connectivity\" "
                  "format=\"ascii\">\n");
  for (const auto& t : M_.tets) std::fprintf(f, "%d %d %d %d\n", t[0], t[1], t[2], t[3]);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
  for (int c = 0; c < nc; ++c) std::fprintf(f, "%d\n", 4 * (c + 1));
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (int c = 0; c < nc; ++c) std::fprintf(f, "10\n");
  std::fprintf(f, "</DataArray>\n</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  const bool ok = std::fclose(f) == 0;
  require(ok, "error writing VTU file " + path);
}

}  // namespace nodals

namespace nodals {

KrylovResult H8Solver::momentumSolve(const std::string& mode, int sweeps, double omega) {
  requireDevice();
  require(sweeps >= 1 && sweeps <= 60 && omega > 0.0 && omega < 2.0, "momentum_solve: sweeps in 1..60, omega in (0,2)",
          kErrInvalid);
  KrylovResult r;
  r.its = sweeps;
  if (mode == "m3-fgs" || mode == "m3-sgs") {
    require((bool)wall_, "momentum_solve m3 modes need the wall-function path (nodals_h8_wall_init)", kErrInvalid);
    wallMomentum(sweeps, omega, mode == "m3-sgs" ? 1 : 0, &r.r0, &r.rel);
    wallCounters(true);
  } else {
    require(!wall_, "the wall-function path's momentum is solved by the m3-fgs / m3-sgs modes", kErrInvalid);
    const std::array<double, 3> a0 = momentumAudit();
    momentumWork(mode, omega, 0, sweeps);
    const std::array<double, 3> a1 = momentumAudit();
    r.r0 = std::max(a0[0], std::max(a0[1], a0[2]));
    r.rel = std::max(a1[0], std::max(a1[1], a1[2]));
  }
  r.ok = std::isfinite(r.rel);
  return r;
}

KrylovResult H8Solver::pcgAmg(const double* rhs, double* x, double rtol, double atol, int maxit) {
  requireDevice();
  require(rhs != nullptr && x != nullptr && maxit >= 1, "pcg_amg: null vectors or maxit < 1", kErrInvalid);
  DevOp A = [this](hipStream_t s, const double* v, double* y) {
    if (wall_) wallSchur(s, v, y);
    else fineApply(s, v, y);
  };
  DevOp M = [this](hipStream_t s, const double* v, double* z) { precondApply(s, v, z); };
  NODALS_HIP(hipMemcpyAsync(cont_.data(), rhs, sizeof(double) * (std::size_t)nc_, hipMemcpyHostToDevice, stream_));
  const KrylovResult r = pcg(stream_, red_, nc_, A, M, cont_.data(), rtol, atol, maxit, kw_);
  NODALS_HIP(hipMemcpyAsync(x, kw_.x.data(), sizeof(double) * (std::size_t)nc_, hipMemcpyDeviceToHost, stream_));
  sync();
  return r;
}

}  // namespace nodals
