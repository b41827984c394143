// H8 laminar SIMPLE solver context on one MI355X.
//
// Mirrors the device state and solver calls of the reference H8 driver
// (gpu/serial_cuda/h8_main_base.cu): assemble_live (:1137-1155),
// bt_state + momentum RHS (:2061-2062), fixed_momentum_work (:1700-1715),
// the continuity residual (:2070-2081), the fine CSR refresh (:2083-2086),
// the pressure Krylov solvers (:567-726, plus FGMRES) and the p update
// (:2102).  All device memory is owned by the context; one HIP stream per
// context; errors are thrown as nodals::Error and turned into status codes at
// the C ABI.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "../host/amg_setup.hpp"
#include "../host/h8_setup.hpp"
#include "../host/turb_setup.hpp"
#include "../host/wall_setup.hpp"
#include "amg.hpp"
#include "gnn.hpp"
#include "krylov.hpp"
#include "turb_step.hpp"

namespace nodals {

struct WallState;

struct H8Options {
  std::string mesh, tag = "h8", wall = "patch_0_0", inlet = "patch_2_0", outlet = "patch_1_0", vtuOut;
  std::string lid;  // extension: lid-driven cavity on this patch (no inlet/outlet, pressure pinned in cell 0)
  bool duct = false;  // extension: general duct (uniform inflow, hydraulic diameter, patch pressure drop)
  // extension: "spalding" makes the duct wall a wall-function wall (free
  // velocity DOFs, Nitsche/Spalding closure; needs the k-omega model)
  std::string wallModel = "no_slip";
  bool rauScaleGiven = false;
  double re = 20, bulk = 1, simpleTol = 1e-6;
  double alphaU = .5, alphaP = .5, rauScale = 1.0;
  double momRtol = 1e-6, momAtol = 1e-12, momDrop = .1, momOmega = 1.0;
  double pRtol = .5, pAtol = 1e-12;
  double snapshotTol = 5e-6;
  bool mixingLength = false, dgInlet = false, weakWall = false, supg = false;
  std::string wallTractionMode = "spalding", wallDiagCsv;
  double wallReferenceF = 0.0284003, mixlenScale = 1.0, supgTauScale = 0.05, supgMagic = 9.0;
  int supgQuadPoints = 64;
  long long amgMcgsFuseNnz = 250000;  // extension: coarse MCGS levels smoothed in one workgroup (0: off)
  std::string supgForm = "implicit";
  int maxOuter = 2500, momMax = 20000, pMax = 20;
  int fineCsrRefreshEvery = 1;
  std::string momentumWork = "fgs1", runMode = "fixed10", amgSmoother = "cheb2", amgHierarchy = "sa",
              cfInterp = "direct", pressureSolver = "pcg", cfCoarsening = "pmis", cfStrength = "classical-negative",
              amgSpectrumPolicy = "auto";
  double amgJacobiOmega = .7;
  int amgJacobiFineSweeps = 1, amgJacobiCoarseSweeps = 1;
  double amgMcgsOmega = 1.0;
  int amgMcgsFineSweeps = 1, amgMcgsCoarseSweeps = 1;
  std::string amgMcgsOrder = "symmetric";
  double pressureRichardsonOmega = 1.0, pressureChebLowFraction = .05, pressurePowerSafety = 1.15;
  int pressureChebDegree = 3, pressurePowerIts = 6;
  int cfAggressiveFirst = 0, amgTerminal = 1000, amgPowerIts = 16, amgChebDegree = 2;
  double amgLambdaSafety = 1.5, amgLambdaLowFraction = .05, cfTheta = .25;
  int cfPmax = 4;
  // Extensions of this implementation.
  std::string amgPrecision = "fp64";  // fp64 | fp32 | bf16 (finest level bf16, coarse fp32)
  int pressureRestart = 30;           // FGMRES restart length
  // momentum-work units per SIMPLE iteration (the SST path's fixed M3 sweep
  // count, NODALS_M3_SWEEPS; with no-slip walls every 3x3 block is a scalar
  // times the identity, so an M3 forward sweep is this coloured sweep)
  int momSweeps = 1;
  int verbose = 1;
  // Duct mode: the case's diameter for nu = bulk D / Re (0: the inlet's
  // hydraulic diameter) and an explicit inlet velocity vector (the SST
  // case's [inlet.main] velocity; unset: bulk along the inward inlet normal).
  double ductDiameter = 0.0;
  bool inletVelocityGiven = false;
  double inletVelocity[3] = {0, 0, 0};
  // Wall-function path: the SST application's convergence gate
  // (compiled_gate.cpp:6-13) when > 0 -- the continuity residual after the
  // correction at most this absolute tolerance and the relative k and omega
  // changes at most simpleTol -- instead of the relative gate; and the
  // cadence of the progress line.
  double sstContAtol = 0.0;
  int printEvery = 10;
};

// Parses the reference H8 argv (h8_main_base.cu:1787-1850) and validates it
// the same way (:1851-1911).  argv[0] is not examined.
H8Options parseH8Args(int argc, const char* const* argv);

struct H8RunResult {
  int outer = 0;
  int converged = 0;
  int pass = 0;
  double finalRelCont = 0;
  double avgPressureIts = 0;
  double pressureDrop = 0;
  double exactPressureDrop = 0;
  double loopMs = 0;
  double assemblyMs = 0, momentumMs = 0, continuityMs = 0, pressureMs = 0, pupdateMs = 0;
  double setupMs = 0;
  long long pressureIts = 0;  // total Krylov iterations of the run
  int warmStarts = 0, fallbacks = 0, recorded = 0;
  double turbMs = 0;          // k-omega steps of a coupled run
  double finalDu = 0, finalDk = 0, finalDw = 0, finalCont = 0;  // wall path: last iteration's changes
};

class H8Solver {
 public:
  // withDevice=false builds the host setup only (mesh, DOFs, CSR patterns,
  // colouring, AMG hierarchy); every device call then fails with kErrInvalid.
  explicit H8Solver(const H8Options& o, bool withDevice = true);
  ~H8Solver();
  bool hasDevice() const { return stream_ != nullptr; }
  H8Solver(const H8Solver&) = delete;
  H8Solver& operator=(const H8Solver&) = delete;

  const H8Options& options() const { return opt_; }
  hipStream_t stream() const { return stream_; }

  // convReady: cbuf_ already holds (or has queued on the second stream,
  // ending with evConv_) the convection blocks of the current velocity
  void assemble(double alphaU, double rauScale, bool convReady = false);
  const char* momentumWork(const std::string& mode, double omega, int outer, int sweeps = 1);
  std::array<double, 3> momentumAudit();
  // The audit and the continuity norm in two halves: the launch (sums land
  // in pinned host memory out[], no synchronisation) and the read, valid
  // once the stream has passed the launch.
  void momentumAuditLaunch(double* out);
  static std::array<double, 3> momentumAuditRead(const double* out);
  void continuityLaunch(double* out);
  double continuity();
  void refresh();
  dev::KrylovResult pressureSolve(const std::string& solver, double rtol, double atol, int maxit);
  // hxt5e::momentum_solve (hxt5e_fgmres.cuh:480-527) in its fixed-sweep
  // modes: "m3-fgs" / "m3-sgs" on the wall path's block system, "fgs1" /
  // "sgs1" / "altgs1" on the H8 momentum; returns sweeps and the relative
  // residuals before (r0) and after (rel).
  dev::KrylovResult momentumSolve(const std::string& mode, int sweeps, double omega);
  // hxt4c::pcg_amg (hybrid_amg.cuh:642-698): PCG with the AMG preconditioner
  // on the live pressure operator for a host right-hand side; x (host) in
  // and out.
  dev::KrylovResult pcgAmg(const double* rhs, double* x, double rtol, double atol, int maxit);
  // hxt5c::scalar_solve (scalar_linear_solver.inc:3-136) for k (0) or omega
  // (1) at the current velocity: the equation's assembly and its fixed
  // sweeps, with the relative residuals before and after the sweeps.
  dev::KrylovResult scalarSolve(int equation);
  // The same from an initial guess (device, cells; fgmres only).
  // x0: initial guess; z0: first Krylov direction (fgmres only).
  dev::KrylovResult pressureSolveFrom(const std::string& solver, double rtol, double atol, int maxit,
                                      const double* x0, const double* z0);
  void pressureUpdate(double alphaP);
  double rauSnapshotRel();
  H8RunResult run();
  // Wall-function path: `count` SIMPLE iterations numbered from `firstIt`,
  // continuing from the current state (chunks of one run, for checkpoints).
  H8RunResult runWindow(int firstIt, int count);
  // ASCII VTK unstructured grid of the current fields: velocity, k, omega
  // and wall distance at the vertices, pressure per cell.
  void writeVtu(const std::string& path);

  // Host access in reference numbering (free DOFs as in the reference).
  std::vector<double> getVelocity(int component);
  void setVelocity(int component, const std::vector<double>& v);
  std::vector<double> getCellField(const std::string& name);
  void setCellField(const std::string& name, const std::vector<double>& v);
  std::vector<double> getRowField(const std::string& name);
  std::vector<double> getMomentumValues();  // reference CSR order
  std::vector<double> getFineValues();      // fine CSR order (cell rows unchanged)
  std::vector<double> getBrhs(int component);

  const H8Setup& setup() const { return S_; }
  const FineTopo& fineTopo() const { return F_; }
  const AmgHierarchy& hierarchy() const { return H_; }
  const TetMesh& mesh() const { return M_; }
  int ncolors() const { return S_.coloring.ncolors; }
  int fineLanes() const { return fineLanes_; }

  // Kernel-level handles for measurement: the fine CSR SpMV on device data.
  void fineSpmv(const double* x, double* y);
  // z = M^-1 r with the AMG V-cycle in the configured precision (device
  // pointers, length = cells).
  void precond(const double* r, double* z);
  // Mean duration (ms) of `reps` launches of a hot kernel measured with HIP
  // events on the context stream; *bytes = algorithmic bytes per launch.
  double timeKernel(const std::string& kernel, int reps, double* bytes);
  // Launch probe: while on, every launch of the probed kernel the solver
  // makes is bracketed by HIP events on the context stream (kind 1: the fine
  // Schur SpMV, the Krylov operator; kind 2: the momentum Gauss-Seidel colour
  // passes, k_mom_gs or the wall path's k_m3_gs, launched outside the sweep
  // graph while probed); probeStop returns the launch count and the summed
  // launch durations.
  void probeStart(int kind = 1);
  void probeStop(long long* launches, double* totalMs);
  // Learned warm start: the cell graph and normalised wall distance the
  // network sees, and attaching a network to predict each pressure
  // correction (warm) and/or to record (features, converged correction)
  // samples every `recordEvery` SIMPLE iterations (0: off).  The network
  // must outlive its attachment; attach nullptr to detach.
  CellGraph cellGraph() const;
  // k-omega SST volume terms (equation 0: k, 1: omega) of every cell from
  // the current velocity and the given k, omega (free DOFs in reference
  // order, then fixed slots): mat[nc][64] element matrices, rhs[nc][8].
  void sstVolume(int equation, const std::vector<double>& kFree, const std::vector<double>& kFixed,
                 const std::vector<double>& wFree, const std::vector<double>& wFixed, double kFloor,
                 double omegaFloor, int deviatoric, std::vector<double>& mat, std::vector<double>& rhs);
  // mode 1: prediction as the initial guess at its residual-optimal
  // multiple; mode 2: as an added first Krylov direction.
  void attachGnn(GnnModel* g, int warmMode, int recordEvery);
  // k-omega SST transport step on the current velocity (masked-P1 state
  // space; scalar DOFs are mesh entities, vertices first).  turbInit sets
  // up the step and the inlet state (k, omega at their inlet values on every
  // vertex, zero on face bubbles); turbStep advances k and then omega once.
  void turbInit(const TurbParams& P);
  TurbStepResult turbStep();
  bool hasTurb() const { return (bool)turb_; }
  // Coupled mode: the momentum assembly adds the eddy-viscosity diffusion
  // of the current k, omega (SST nu_t at every Keast point) and run()
  // advances k and omega once after every pressure update.
  void turbCouple(bool on);
  // "k", "omega" (nS), "turb_val" (vertex operator values, relaxed),
  // "turb_diag", "turb_rhs" (vertex rows) of the last equation assembled,
  // topology arrays "turb_row", "turb_col", "turb_color", "turb_cell_ent",
  // "turb_sample_cell", "turb_sample_l", "turb_heights", "turb_inlet_state"
  // (kIn, omegaIn).
  std::vector<double> turbField(const std::string& name);
  void setTurbField(const std::string& name, const std::vector<double>& v);
  // Wall-function momentum path (a duct created with --wall-model spalding,
  // after turbInit): the walls carry the face-averaged Spalding traction and
  // the Nitsche normal penalty, the wall-face bubbles the exact BF3 mean-flux
  // constraint u = T x, and the momentum is one 3x3 block operator
  // K = T^T M T swept by block Gauss-Seidel (the reference's
  // m3_fixed_coupled_csr).  The pressure Krylov operator becomes
  // B T Q R Q T^T B^T (R the gamma-free rAU) and the SIMPLE update corrects
  // the velocity with the same operator; run() runs this loop.
  void wallInit(const WallParams& P);
  bool hasWall() const { return (bool)wall_; }
  // Fit, wall blocks, component rows and relaxation, K, T^T b.
  void wallAssemble();
  // Q u, the block sweeps, T u.
  // The M3 block sweeps; sweeps / omega / symmetric < 0 take the wall
  // options; rel0 / rel1, when given, receive ||b - K x|| / ||b|| of the
  // block system before and after the sweeps.
  void wallMomentum(int sweeps = -1, double omega = -1.0, int symmetric = -1, double* rel0 = nullptr,
                    double* rel1 = nullptr);
  // y = B T Q R Q T^T B^T x on device vectors of length cells.
  void wallSchur(hipStream_t s, const double* x, double* y);
  // u += T Q R Q T^T B^T dp for a device vector dp (cells).
  void wallCorrect(const double* dp);
  // "wall_K" [nnz][9] (reference entry order), "wall_rhs", "wall_rau3",
  // "wall_r3", "wall_delta3" [nv][3] (reference free order), "wall_fit"
  // [nf][3], "wall_shear" [nf][4], "wall_nuq" [nf][12], topology arrays
  // "wall_normal" [nf][3], "wall_area", "wall_bw", "wall_bp", "wall_qw".
  std::vector<double> wallField(const std::string& name);
  double* devScratchA() { return tmpA_.data(); }
  double* devScratchB() { return tmpB_.data(); }
  void sync();

  // algorithmic bytes of one application of each hot kernel
  struct Bytes {
    double fineSpmv = 0, refresh = 0, momGsSweep = 0, convLocal = 0, momRows = 0, continuity = 0;
    double m3Sweep = 0, wallSchur = 0;  // wall-function path
  };
  Bytes algorithmicBytes() const;

 private:
  void requireDevice() const {
    require(stream_ != nullptr, "H8 context was created without a device (host setup only)");
  }
  void upload();
  void setupAmg();
  void precondApply(hipStream_t s, const double* r, double* z);
  void fineApply(hipStream_t s, const double* x, double* y);
  double pressurePrecondPower(int its, double safety);
  void probeHarvest();
  // Features of the current pressure system into X (device, M x 16 FP32);
  // the prediction scale 1/rms(dp_{n-1}) lands in gnnScal_[2].
  void gnnFeatures(float* X);

  H8Options opt_;
  TetMesh M_;
  H8Setup S_;
  FineTopo F_;
  AmgHierarchy H_;
  hipStream_t stream_ = nullptr;
  hipStream_t stream2_ = nullptr;  // the refresh, overlapped with the momentum solve
  int cuCount_ = 1;
  hipEvent_t evAsm_ = nullptr, evRef_ = nullptr, evSweep_ = nullptr, evConv_ = nullptr;
  void convLocal(hipStream_t s, int waves);
  // pinned: [0] the continuity residual's square sum, [1..6] the momentum
  // audit's sums, read by the SIMPLE loop after its pressure solve
  double* loopHost_ = nullptr;
  void refreshOn(hipStream_t s);
  dev::Reducer red_;
  dev::KrylovWork kw_;

  std::vector<int> perm_, pinv_;  // device numbering: perm_[new] = reference row
  int nc_ = 0, nv_ = 0, nFix_ = 0, fineLanes_ = 16, refreshLanes_ = 32;
  double fineMeanRow_ = 8.0;
  long long momNnz_ = 0, fineNnz_ = 0;
  std::vector<int> colorOff_, colorSplit_;  // colour c: [off, split) long rows, [split, off+1) short rows
  // Momentum rows with more than kLongRow nonzeros (vertex rows) get
  // kLongLanes lanes each, the others (face-bubble rows) kShortLanes.
  static constexpr int kLongRow = 24, kLongLanes = 32, kShortLanes = 16;
  // The momentum audit (a reduction, no colour dependence) runs best with
  // fewer, fuller lanes: 8 per vertex row and 4 per face row took the
  // 40k-cell momentum stage from 121 to 115.5 us per SIMPLE iteration
  // (16/8: 117.5, 4/4: 116.8, 32/16: 121).
  static constexpr int kNormLongLanes = 8, kNormShortLanes = 4;
  int nLong_ = 0, nShort_ = 0;
  DBuf<int> longRows_, shortRows_;
  DBuf<double> normPart_;

  DBuf<int> cref_;
  DBuf<double> cinvJ_, cdet_, bcoef_, fixedVal_, centT_;
  DBuf<int> mrow_, mcol_, mdiagPos_, termOff_, termIdx_, fixOff_, fixTerm_, incOff_;
  DBuf<unsigned> incPacked_;
  DBuf<double> mdiff_, aval_, cbuf_;
  DBuf<double> u_, brhs_, srhs_, delta_, diag_, rau_, rauSnap_;
  DBuf<double> p_, fixedDiv_, cont_, pcx_;
  DBuf<int> frow_, fcol_, fdiagPos_, nzOff_;
  DBuf<unsigned> nzTerm_;
  DBuf<double> fval_, fdiag_, fl1_, fineW_;  // fineW_: rAU-weighted B per cell basis
  DBuf<float> fvalf_, fdiagf_, fl1f_, rf_, zf_;
  DBuf<bf16> fvalb_;
  DBuf<double> tmpA_, tmpB_, calibA_, calibB_, rd_, zd_;
  GraphCache sweepGraph_[2];  // [0] forward / symmetric sweep, [1] backward sweep
  std::unique_ptr<dev::DevAmg<double, double>> amg64_;
  std::unique_ptr<dev::DevAmg<float, float>> amg32_;
  std::unique_ptr<dev::DevAmg<bf16, float>> amgb_;
  double setupMs_ = 0;
  double chebLambda_ = 1.0;
  DBuf<double> sstXY_, sstQuad_;
  void sstGeometry();
  // Mixing-length and SUPG cell blocks (--mixing-length, --supg) added to
  // cbuf_ in the assembly; per-cell largest squared edge and mixing-length
  // statistics [nc][4].
  DBuf<double> ch2_, ransStats_;
  void ransBlocks();
  // DG weak inlet (--dg-inlet): inlet faces, their right-hand sides and the
  // rows those reach
  DBuf<int> dgCell_, dgLf_, dgRows_, dgOff_, dgList_;
  DBuf<double> dgGeom_, dgFaceRhs_;
  int dgNf_ = 0, dgNrows_ = 0;
  void dgInletBlocks();
 public:
  // Mixing-length summary of the last assembly: volume-weighted mean and
  // maximum nu_t / nu, maximum strain rate (zeros when the term is off).
  std::vector<double> ransStats();
 private:
  // One k (equation 0) or omega (1) update of the transport step.
  int turbEquation(int equation, double* rel0 = nullptr, double* rel1 = nullptr);
  void turbAdvance();       // convection + k + omega, queued on the stream
  void eddyDiffusion();     // adds the nu_t diffusion blocks to cbuf_
  int turbCounters(TurbStepResult* R);  // reads the step's counters (syncs)
  std::unique_ptr<TurbState> turb_;
  bool turbCouple_ = false;
  std::unique_ptr<WallState> wall_;
  H8RunResult runWall(int firstIt, int count);
  int wallSweeps() const;
  int wallCounters(bool sweep);   // reads and checks the wall counters (syncs)
  std::vector<long long> newOfOld_;  // reference momentum entry -> device entry
  const double* momT_ = nullptr;     // convection tensor of the momentum (null: centT_)
  const double* rau3_ = nullptr;     // anisotropic rAU of the fine Schur refresh (null: rau_)
  GnnModel* gnn_ = nullptr;
  bool gnnWarm_ = false, gnnAsGuess_ = false;
  int gnnEvery_ = 0;
  DBuf<double> dp1_, dp2_, wallD_, x0_, gnnScal_, gnnPart_;
  DBuf<float> gnnX_;
  static constexpr int kProbePairs = 512;
  std::vector<hipEvent_t> probeEv_;
  int probeUsed_ = 0;
  bool probeOn_ = false;
  int probeKind_ = 1;
  long long probeLaunches_ = 0;
  double probeMs_ = 0;
};

}  // namespace nodals
