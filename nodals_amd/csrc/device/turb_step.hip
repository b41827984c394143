// k-omega SST transport step of an H8 context (see turb_step.hpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "h8_kernels.hpp"
#include "h8_solver.hpp"
#include "turb_kernels.hpp"

namespace nodals {

using namespace dev;

TurbParams parseTurbArgs(int argc, const char* const* argv) {
  TurbParams P;
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
    if (f == "--sst-intensity") P.intensity = num(f, v);
    else if (f == "--sst-length-ratio") P.lengthRatio = num(f, v);
    else if (f == "--sst-alpha-k") P.alphaK = num(f, v);
    else if (f == "--sst-alpha-omega") P.alphaOmega = num(f, v);
    else if (f == "--sst-k-fgs-sweeps") P.kSweeps = (int)num(f, v);
    else if (f == "--sst-omega-fgs-sweeps") P.omegaSweeps = (int)num(f, v);
    else if (f == "--sst-gs-omega") P.gsOmega = num(f, v);
    else if (f == "--sst-k-floor") P.kFloor = num(f, v);
    else if (f == "--sst-omega-floor") P.omegaFloor = num(f, v);
    else if (f == "--sst-deviatoric-production") P.deviatoric = num(f, v) != 0.0;
    else if (f == "--sst-omega-wall-mode") {
      const std::string m = v;
      require(m == "of_auto" || m == "none", "--sst-omega-wall-mode must be of_auto or none");
      P.wall = m == "of_auto";
    } else if (f == "--sst-wall-penalty-gamma") P.wallGamma = num(f, v);
    else if (f == "--sst-omega-wall-target") {
      const std::string m = v;
      require(m == "of_auto" || m == "sample_log" || m == "trace_log",
              "--sst-omega-wall-target must be of_auto, sample_log or trace_log");
      P.omegaMode = m == "of_auto" ? 0 : m == "sample_log" ? 1 : 2;
    } else if (f == "--sst-omega-trace-gamma") P.traceGamma = num(f, v);
    else if (f == "--sst-sweep-precision") {
      const std::string m = v;
      require(m == "fp64" || m == "fp32", "--sst-sweep-precision must be fp64 or fp32");
      P.sweepFp32 = m == "fp32";
    }
    else if (f.size() == 20 && f.rfind("--sst-profile-fit-y", 0) == 0 && f[19] >= '1' && f[19] <= '5')
      P.heights[(std::size_t)(f[19] - '1')] = num(f, v);
    else if (f == "--sst-wall-sample-scale") P.sampleScale = num(f, v);
    else if (f == "--sst-of-kappa") P.ofKappa = num(f, v);
    else if (f == "--sst-of-E") P.ofE = num(f, v);
    else if (f == "--sst-of-beta1") P.ofBeta1 = num(f, v);
    else if (f == "--sst-of-re-blend") P.ofReBlend = num(f, v);
    else if (f == "--sst-of-production-limit") P.ofProdLimit = num(f, v);
    else if (f == "--sst-wall-sampling") {
      const std::string m = v;
      require(m == "ray_trace" || m == "legacy_endpoint", "--sst-wall-sampling must be ray_trace or legacy_endpoint");
      P.sampling = m == "ray_trace" ? 1 : 0;
    } else if (f == "--sst-wall-ray-max-crossings") P.rayMaxCrossings = (int)num(f, v);
    else if (f == "--sst-wall-ray-bary-tolerance") P.rayTol = num(f, v);
    else if (f == "--sst-wall-ray-local-adapt") P.rayLocalAdapt = num(f, v) != 0.0;
    else if (f == "--sst-wall-active-points") P.activePoints = (int)num(f, v);
    else if (f == "--sst-face-utau-csv") P.faceUtauCsv = v;
    else if (f == "--sst-face-yplus-sp") P.faceTargetSp = num(f, v);
    else if (f == "--sst-face-yplus-omega") P.faceTargetOmega = num(f, v);
    else if (f == "--sst-face-yplus-min-scale") P.faceMinScale = num(f, v);
    else if (f == "--sst-face-yplus-max-scale") P.faceMaxScale = num(f, v);
    else if (f == "--sst-spalding-sample-scale") P.baseScale = num(f, v);
    else fail("unknown k-omega option " + f);
  }
  require(P.gsOmega > 0 && P.gsOmega < 2, "--sst-gs-omega must be in (0, 2)");
  require(P.rayMaxCrossings >= 1, "--sst-wall-ray-max-crossings must be positive");
  require(P.rayTol > 0 && P.rayTol <= 1e-5, "--sst-wall-ray-bary-tolerance must be in (0, 1e-5]");
  require(P.activePoints >= 1 && P.activePoints <= 5, "--sst-wall-active-points must be in 1..5");
  require(P.wallGamma > 0, "--sst-wall-penalty-gamma must be positive");
  require(P.omegaMode != 2 || (std::isfinite(P.traceGamma) && P.traceGamma >= 10.0 && P.traceGamma <= 1e5),
          "trace_log trace_gamma must be in [10,100000]");
  require(P.intensity > 0 && P.lengthRatio > 0, "--sst-intensity and --sst-length-ratio must be positive");
  const bool anyY = P.heights[0] > 0 || P.heights[1] > 0 || P.heights[2] > 0 || P.heights[3] > 0 || P.heights[4] > 0;
  if (anyY)
    for (int m = 0; m < 5; ++m)
      require(P.heights[(std::size_t)m] > 0 && (m == 0 || P.heights[(std::size_t)m] > P.heights[(std::size_t)m - 1]),
              "the five --sst-profile-fit-y heights must all be given, positive and increasing");
  return P;
}

namespace {

Tri12 tri12() {
  Tri12 Q;
  for (int q = 0; q < 12; ++q)
    for (int j = 0; j < 4; ++j) Q.q[q][j] = kTri12[q][j];
  return Q;
}

}  // namespace

void H8Solver::turbInit(const TurbParams& P) {
  auto st = std::make_unique<TurbState>();
  st->P = P;
  st->T = buildTurbTopo(M_, S_, P);
  if (!hasDevice()) {  // host-only context: topology only
    turb_ = std::move(st);
    return;
  }
  sstGeometry();
  const TurbTopo& T = st->T;
  st->sref.upload(T.cellEnt);
  st->row.upload(T.row);
  st->col.upload(T.col);
  st->diagPos.upload(T.diagPos);
  st->entOff.upload(T.entOff);
  st->termOff.upload(T.termOff);
  st->term.upload(T.term);
  st->rhsOff.upload(T.rhsOff);
  st->rhsTerm.upload(T.rhsTerm);
  st->crows.upload(T.coloring.rows);
  st->colorOff = T.coloring.offset;
  std::vector<int> fc, fl;
  std::vector<double> fg;
  for (const TurbFace& f : T.inlet) {
    fc.push_back(f.cell);
    fl.push_back(f.lf);
    fg.insert(fg.end(), {f.n[0], f.n[1], f.n[2], f.area});
  }
  if (fc.empty()) {
    fc.push_back(0);
    fl.push_back(0);
    fg.assign(4, 0.0);
  }
  st->fcell.upload(fc);
  st->flf.upload(fl);
  st->fgeom.upload(fg);
  std::vector<double> wg, sl, sp, sy;
  std::vector<int> sc, wc, wl;
  for (const TurbFace& f : T.wall) {
    wg.insert(wg.end(), {f.n[0], f.n[1], f.n[2], f.area});
    wc.push_back(f.cell);
    wl.push_back(f.lf);
  }
  for (const TurbSample& s : T.samples) {
    sc.push_back(s.cell);
    sl.insert(sl.end(), s.l, s.l + 4);
    sp.insert(sp.end(), s.phi, s.phi + 8);
    sy.push_back(s.y);
  }
  if (wc.empty()) {
    wc.push_back(0);
    wl.push_back(0);
  }
  if (sc.empty()) {
    wg.assign(4, 0.0);
    sc.push_back(0);
    sl.assign(4, 0.0);
    sp.assign(8, 0.0);
    sy.assign(1, 0.0);
  }
  st->wgeom.upload(wg);
  st->scell.upload(sc);
  st->sl.upload(sl);
  st->sphi.upload(sp);
  st->sy.upload(sy);
  st->wcell.upload(wc);
  st->wlf.upload(wl);
  st->omegaFitPoints = P.activePoints;
  st->mu.alloc((std::size_t)std::max(T.nWallQ, 1));
  st->coef.alloc((std::size_t)std::max(T.nWallQ, 1));
  st->target.alloc((std::size_t)std::max(T.nWallQ, 1));
  st->conv.alloc((std::size_t)nc_ * 64);
  const std::size_t nb = (std::size_t)nc_ + T.nInlet + T.nWallQ;
  st->blk.alloc(nb * 64);
  st->brhs.alloc(nb * 8);
  st->val.alloc(T.col.size());
  if (P.sweepFp32) st->valF.alloc(T.col.size());
  st->diag.alloc((std::size_t)T.nvert);
  st->rhs.alloc((std::size_t)T.nvert);
  std::vector<double> k0((std::size_t)T.nS, 0.0), w0((std::size_t)T.nS, 0.0);
  for (int i = 0; i < T.nvert; ++i) {
    k0[(std::size_t)i] = T.kIn;
    w0[(std::size_t)i] = T.omegaIn;
  }
  st->k.upload(k0);
  st->w.upload(w0);
  st->counters.upload(std::vector<int>(4, 0));
  if (!S_.wallDistance.empty()) st->vdist.upload(S_.wallDistance);
  for (hipEvent_t& e : st->ev) NODALS_HIP(hipEventCreate(&e));
  turb_ = std::move(st);
}

int H8Solver::turbEquation(int eq, double* rel0, double* rel1) {
  TurbState& st = *turb_;
  const TurbTopo& T = st.T;
  const TurbParams& P = st.P;
  hipStream_t s = stream_;
  const TurbConsts K{S_.pipe.R, S_.pipe.nu, P.kFloor, P.omegaFloor, P.deviatoric ? 1 : 0, eq};
  double* x = eq == 0 ? st.k.data() : st.w.data();
  // element blocks: volume terms with the molecular diffusion, plus convection
  hipLaunchKernelGGL(k_sst_volume, (nc_ + 3) / 4, 256, 0, s, nc_, cref_.data(), cinvJ_.data(), cdet_.data(),
                     sstXY_.data(), u_.data(), fixedVal_.data(), st.k.data(), st.k.data(), st.w.data(), st.w.data(),
                     sstQuad_.data(), S_.pipe.R, S_.pipe.nu, eq, P.kFloor, P.omegaFloor, P.deviatoric ? 1 : 0,
                     st.blk.data(), st.brhs.data(), st.sref.data(), S_.pipe.nu, st.conv.data(), st.vdistPtr());
  NODALS_LAUNCHED();
  if (T.nInlet > 0) {
    hipLaunchKernelGGL(k_turb_inlet, (T.nInlet + 3) / 4, 256, 0, s, T.nInlet, st.fcell.data(), st.flf.data(),
                       st.fgeom.data(), cref_.data(), st.sref.data(), cinvJ_.data(), sstXY_.data(), u_.data(),
                       fixedVal_.data(), st.k.data(), st.w.data(), K, tri12(), S_.pipe.inletVelocity[0],
                       S_.pipe.inletVelocity[1], S_.pipe.inletVelocity[2], eq == 0 ? T.kIn : T.omegaIn,
                       st.blk.data() + (std::size_t)nc_ * 64, st.brhs.data() + (std::size_t)nc_ * 8, st.vdistPtr());
    NODALS_LAUNCHED();
  }
  int nb = nc_ + T.nInlet;
  if (eq == 1 && T.nWallQ > 0) {
    const WallConsts W{P.wallGamma, P.ofKappa,     P.ofE, P.ofBeta1, P.ofReBlend, P.ofProdLimit,
                       T.heights[4] - T.heights[0], st.omegaFitPoints, P.omegaMode};
    hipLaunchKernelGGL(k_turb_wall_coef, blocksFor(T.nWallQ), 256, 0, s, T.nWallQ, st.wgeom.data(), st.scell.data(),
                       st.sl.data(), st.sphi.data(), st.sy.data(), cref_.data(), st.sref.data(), cinvJ_.data(), sstXY_.data(),
                       u_.data(), fixedVal_.data(), st.k.data(), st.w.data(), K, W, tri12(), st.coef.data(),
                       st.target.data(), st.counters.data() + 2, st.vdistPtr(), st.mu.data());
    NODALS_LAUNCHED();
    if (P.omegaMode == 2)
      hipLaunchKernelGGL(k_turb_wall_trace, (T.nWallQ + 3) / 4, 256, 0, s, T.nWallQ, st.wcell.data(), st.wlf.data(),
                         st.wgeom.data(), cinvJ_.data(), cdet_.data(), st.coef.data(), st.target.data(),
                         st.mu.data(), P.traceGamma, tri12(), st.blk.data() + (std::size_t)nb * 64,
                         st.brhs.data() + (std::size_t)nb * 8, st.counters.data() + 2);
    else
      hipLaunchKernelGGL(k_turb_wall_block, (T.nWallQ + 3) / 4, 256, 0, s, T.nWallQ, st.sphi.data(),
                         st.coef.data(), st.target.data(), st.blk.data() + (std::size_t)nb * 64,
                         st.brhs.data() + (std::size_t)nb * 8);
    NODALS_LAUNCHED();
    nb += T.nWallQ;
  }
  const double alpha = eq == 0 ? P.alphaK : P.alphaOmega;
  // lanes per row from the entries per row (vertex + bubble columns, about
  // 50 on tetrahedral meshes): each lane's serial chain of indirect loads
  // stays short
  const double meanEnt = (double)T.entOff[(std::size_t)T.nvert] / std::max(T.nvert, 1);
  const int gl = meanEnt > 48 ? 32 : meanEnt > 12 ? 16 : 8;
#define NODALS_TURB_GATHER(L)                                                                                      \
  hipLaunchKernelGGL(k_turb_gather<L>, blocksFor((long long)T.nvert * L), 256, 0, s, T.nvert, st.row.data(),        \
                     st.diagPos.data(), st.entOff.data(), st.termOff.data(), st.term.data(), st.rhsOff.data(),      \
                     st.rhsTerm.data(), nb * 64, nb * 8, st.blk.data(), st.brhs.data(), x, alpha, st.val.data(),   \
                     st.diag.data(), st.rhs.data(), st.counters.data() + 3)
  if (gl == 32) NODALS_TURB_GATHER(32);
  else if (gl == 16) NODALS_TURB_GATHER(16);
  else NODALS_TURB_GATHER(8);
#undef NODALS_TURB_GATHER
  NODALS_LAUNCHED();
  auto relResid = [&]() {
    if (st.resid.size() < (std::size_t)T.nvert) st.resid.alloc((std::size_t)std::max(T.nvert, 1));
    hipLaunchKernelGGL(k_scalar_resid, blocksFor(T.nvert), 256, 0, s, T.nvert, st.row.data(), st.col.data(),
                       st.val.data(), st.diag.data(), st.rhs.data(), x, st.resid.data());
    NODALS_LAUNCHED();
    const double rr = red_.dot(s, T.nvert, st.resid.data(), st.resid.data());
    const double bb = red_.dot(s, T.nvert, st.rhs.data(), st.rhs.data());
    return std::sqrt(rr) / std::max(std::sqrt(bb), 1e-300);
  };
  if (rel0) *rel0 = relResid();
  // the fixed sweeps (sweeps x colours launches on fixed buffers) replay
  // from one captured graph per equation
  const double meanRow = (double)T.col.size() / std::max(T.nvert, 1);
  const int sweeps = eq == 0 ? P.kSweeps : P.omegaSweeps;
  if (P.sweepFp32) cast<float, double>(s, (long long)T.col.size(), st.val.data(), st.valF.data());
  st.sweepGraph[eq].run(s, graphKey({x, st.val.data(), st.rhs.data()}, sweeps * 4.0 + P.gsOmega), [&]() {
    for (int sw = 0; sw < sweeps; ++sw)
      for (int c = 0; c < T.coloring.ncolors; ++c) {
        const int b = st.colorOff[(std::size_t)c], n = st.colorOff[(std::size_t)c + 1] - b;
        if (P.sweepFp32)
          mcgsColor<float, double>(s, lanesFor(meanRow, n), b, n, st.crows.data(), st.row.data(), st.col.data(),
                                   st.valF.data(), st.diag.data(), st.rhs.data(), x, P.gsOmega);
        else
          mcgsColor<double, double>(s, lanesFor(meanRow, n), b, n, st.crows.data(), st.row.data(),
                                    st.col.data(), st.val.data(), st.diag.data(), st.rhs.data(), x, P.gsOmega);
      }
  });
  if (rel1) *rel1 = relResid();
  hipLaunchKernelGGL(k_turb_floor, blocksFor(T.nvert), 256, 0, s, T.nvert, eq == 0 ? P.kFloor : P.omegaFloor, x,
                     st.counters.data() + eq);
  NODALS_LAUNCHED();
  return nb;
}

KrylovResult H8Solver::scalarSolve(int equation) {
  requireDevice();
  require((bool)turb_, "scalar_solve needs the k-omega model (nodals_h8_turb_init)", kErrInvalid);
  require(equation == 0 || equation == 1, "scalar_solve: equation 0 (k) or 1 (omega)", kErrInvalid);
  TurbState& st = *turb_;
  hipLaunchKernelGGL(k_fill<int>, 1, 256, 0, stream_, 4ll, st.counters.data(), 0);
  NODALS_LAUNCHED();
  const int waves = std::min(nc_, 8192);
  hipLaunchKernelGGL(k_conv_local, blocksFor((long long)waves * 64), 256, 0, stream_, nc_, cref_.data(),
                     cinvJ_.data(), cdet_.data(), u_.data(), fixedVal_.data(), centT_.data(), st.conv.data());
  NODALS_LAUNCHED();
  KrylovResult r;
  turbEquation(equation, &r.r0, &r.rel);
  TurbStepResult tr;
  turbCounters(&tr);
  r.its = equation == 0 ? st.P.kSweeps : st.P.omegaSweeps;
  r.ok = std::isfinite(r.rel);
  return r;
}

void H8Solver::turbAdvance() {
  TurbState& st = *turb_;
  hipStream_t s = stream_;
  hipLaunchKernelGGL(k_fill<int>, 1, 256, 0, s, 4ll, st.counters.data(), 0);
  NODALS_LAUNCHED();
  const int waves = std::min(nc_, 8192);
  hipLaunchKernelGGL(k_conv_local, blocksFor((long long)waves * 64), 256, 0, s, nc_, cref_.data(), cinvJ_.data(),
                     cdet_.data(), u_.data(), fixedVal_.data(), centT_.data(), st.conv.data());
  NODALS_LAUNCHED();
  turbEquation(0);
  turbEquation(1);
  ++st.steps;
}

int H8Solver::turbCounters(TurbStepResult* R) {
  std::vector<int> cnt = turb_->counters.host();
  R->kClips = cnt[0];
  R->omegaClips = cnt[1];
  R->wallInvalid = cnt[2];
  require(cnt[3] == 0, "k-omega step: " + std::to_string(cnt[3]) + " relaxed rows with a non-positive or "
                       "non-finite diagonal", kErrNonFinite);
  return cnt[3];
}

TurbStepResult H8Solver::turbStep() {
  requireDevice();
  require((bool)turb_, "k-omega step not initialised (nodals_h8_turb_init)", kErrInvalid);
  TurbState& st = *turb_;
  NODALS_HIP(hipEventRecord(st.ev[0], stream_));
  turbAdvance();
  NODALS_HIP(hipEventRecord(st.ev[1], stream_));
  sync();
  TurbStepResult R;
  float ms = 0;
  NODALS_HIP(hipEventElapsedTime(&ms, st.ev[0], st.ev[1]));
  R.ms = ms;
  turbCounters(&R);
  return R;
}

void H8Solver::turbCouple(bool on) {
  requireDevice();
  require(!on || (bool)turb_, "k-omega step not initialised (nodals_h8_turb_init)", kErrInvalid);
  turbCouple_ = on;
}

// Eddy-viscosity diffusion of the momentum equation added to the cell
// blocks of the convection (cbuf_), which the momentum gather then
// assembles like the convection itself (Dirichlet columns to the rhs).
void H8Solver::eddyDiffusion() {
  TurbState& st = *turb_;
  const TurbParams& P = st.P;
  hipLaunchKernelGGL(k_sst_volume, (nc_ + 3) / 4, 256, 0, stream_, nc_, cref_.data(), cinvJ_.data(), cdet_.data(),
                     sstXY_.data(), u_.data(), fixedVal_.data(), st.k.data(), st.k.data(), st.w.data(), st.w.data(),
                     sstQuad_.data(), S_.pipe.R, S_.pipe.nu, 2, P.kFloor, P.omegaFloor, P.deviatoric ? 1 : 0,
                     cbuf_.data(), nullptr, st.sref.data(), 0.0, cbuf_.data(), st.vdistPtr());
  NODALS_LAUNCHED();
}

std::vector<double> H8Solver::turbField(const std::string& n) {
  require((bool)turb_, "k-omega step not initialised (nodals_h8_turb_init)", kErrInvalid);
  TurbState& st = *turb_;
  const TurbTopo& T = st.T;
  const bool devField = n == "k" || n == "omega" || n == "turb_val" || n == "turb_diag" || n == "turb_rhs" ||
                        n == "turb_wall_target";
  if (devField) requireDevice();
  if (hasDevice()) sync();
  auto toD = [](const std::vector<int>& v) { return std::vector<double>(v.begin(), v.end()); };
  if (n == "k") return st.k.host();
  if (n == "omega") return st.w.host();
  if (n == "turb_val") return st.val.host();
  if (n == "turb_diag") return st.diag.host();
  if (n == "turb_rhs") return st.rhs.host();
  if (n == "turb_wall_target") return st.target.host();  // omega target per wall point, last omega assembly
  if (n == "turb_row") return toD(T.row);
  if (n == "turb_col") return toD(T.col);
  if (n == "turb_color") return toD(T.coloring.color);
  if (n == "turb_cell_ent") return toD(T.cellEnt);
  if (n == "turb_sample_cell") {
    std::vector<double> v;
    for (const TurbSample& s : T.samples) v.push_back(s.cell);
    return v;
  }
  if (n == "turb_sample_l") {
    std::vector<double> v;
    for (const TurbSample& s : T.samples) v.insert(v.end(), s.l, s.l + 4);
    return v;
  }
  if (n == "turb_heights") return std::vector<double>(T.heights.begin(), T.heights.end());
  if (n == "turb_sample_y") {
    std::vector<double> v;
    for (const TurbSample& s : T.samples) v.push_back(s.y);
    return v;
  }
  if (n == "turb_face_y3") return T.faceY3;
  if (n == "turb_face_yplus_stats") return std::vector<double>(T.faceYplusStats.begin(), T.faceYplusStats.end());
  if (n == "turb_wall_faces") {
    std::vector<double> v;
    for (const TurbFace& f : T.wall) v.insert(v.end(), {(double)f.cell, (double)f.lf});
    return v;
  }
  if (n == "turb_sample_stats")
    return {(double)T.placed, (double)T.crossed, (double)T.maxCrossings, (double)T.shortened, T.minKeptFraction,
            (double)T.rayMask, (double)st.P.sampling};
  if (n == "turb_inlet_state") return {T.kIn, T.omegaIn, (double)T.nInlet, (double)T.wall.size(), (double)T.nS,
                                       (double)T.nvert, (double)T.coloring.ncolors};
  fail("unknown k-omega field " + n);
}

void H8Solver::setTurbField(const std::string& n, const std::vector<double>& v) {
  require((bool)turb_, "k-omega step not initialised (nodals_h8_turb_init)", kErrInvalid);
  TurbState& st = *turb_;
  requireDevice();
  require(n == "k" || n == "omega", "settable k-omega fields: k, omega");
  require((int)v.size() == st.T.nS, n + " must have one value per scalar DOF (vertices, then faces)");
  for (int i = st.T.nvert; i < st.T.nS; ++i)
    require(v[(std::size_t)i] == 0.0, n + ": face-bubble coefficients must be zero in the masked-P1 state space");
  sync();
  (n == "k" ? st.k : st.w).upload(v);
}

}  // namespace nodals
