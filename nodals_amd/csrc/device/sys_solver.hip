// Explicit SPD system context: FP64 CSR operator, SA hierarchy in FP64, FP32
// or BF16-finest storage, FGMRES / PCG outer solve in FP64.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <sstream>
#include <string>

#include "../host/amg_setup.hpp"
#include "amg.hpp"
#include "krylov.hpp"
#include "sys_solver.hpp"

namespace nodals {

using namespace dev;

SysSolver::SysSolver(int n, const std::int64_t* row, const std::int32_t* col, const double* val,
                     const std::string& options, bool withDevice) {
  require(n > 0 && row && col && val, "sys_create: empty system");
  std::istringstream is(options);
  std::string kv;
  while (is >> kv) {
    const auto e = kv.find('=');
    require(e != std::string::npos, "sys option must be key=value: " + kv);
    const std::string k = kv.substr(0, e), v = kv.substr(e + 1);
    if (k == "precision") precision_ = v;
    else if (k == "smoother") so_.kind = v;
    else if (k == "sweeps") so_.mcgsFineSweeps = so_.mcgsCoarseSweeps = so_.jacobiFineSweeps = so_.jacobiCoarseSweeps = std::stoi(v);
    else if (k == "omega") so_.jacobiOmega = so_.mcgsOmega = std::stod(v);
    else if (k == "order") so_.mcgsOrder = v;
    else if (k == "terminal") ao_.terminal = std::stoi(v);
    else if (k == "power_its") ao_.powerIts = std::stoi(v);
    else if (k == "cheb_degree") so_.chebDegree = std::stoi(v);
    else if (k == "hierarchy") hierarchy_ = v;
    else if (k == "theta") cf_.theta = std::stod(v);
    else if (k == "pmax") cf_.pmax = std::stoi(v);
    else if (k == "interp") cf_.interp = v;
    else if (k == "aggressive") cf_.aggressiveFirst = std::stoi(v) != 0;
    else fail("unknown sys option " + k);
  }
  require(precision_ == "fp64" || precision_ == "fp32" || precision_ == "bf16", "precision must be fp64|fp32|bf16");
  require(so_.kind == "jacobi" || so_.kind == "l1jacobi" || so_.kind == "cheb2" || so_.kind == "mcgs",
          "smoother must be jacobi|l1jacobi|cheb2|mcgs");
  A_.n = n;
  A_.row.assign(row, row + n + 1);
  require(A_.row[0] == 0 && A_.row[(std::size_t)n] > 0 && A_.row[(std::size_t)n] < (1ll << 31), "sys: bad row pointers");
  const std::size_t nnz = (std::size_t)A_.row[(std::size_t)n];
  A_.col.assign(col, col + nnz);
  A_.val.assign(val, val + nnz);
  A_.diag.assign((std::size_t)n, 0.0);
  std::vector<int> dpos((std::size_t)n, -1);
  for (int i = 0; i < n; ++i) {
    require(A_.row[(std::size_t)i + 1] >= A_.row[(std::size_t)i], "sys: row pointers decrease");
    for (std::int64_t k = A_.row[(std::size_t)i]; k < A_.row[(std::size_t)i + 1]; ++k) {
      require(A_.col[(std::size_t)k] >= 0 && A_.col[(std::size_t)k] < n, "sys: column out of range");
      if (A_.col[(std::size_t)k] == i) {
        A_.diag[(std::size_t)i] = A_.val[(std::size_t)k];
        dpos[(std::size_t)i] = (int)k;
      }
    }
    require(A_.diag[(std::size_t)i] > 0.0, "sys: non-positive or missing diagonal at row " + std::to_string(i));
  }
  const auto t0 = std::chrono::steady_clock::now();
  require(hierarchy_ == "sa" || hierarchy_ == "cf", "hierarchy must be sa or cf");
  if (hierarchy_ == "cf") {
    cf_.terminal = ao_.terminal;
    cf_.powerIts = so_.kind == "cheb2" ? ao_.powerIts : 0;
    cf_.safety = ao_.safety;
    H_ = buildCfHierarchyCsr(A_, cf_);
  } else {
    H_ = buildSaHierarchyCsr(A_, ao_);
  }
  const auto t1 = std::chrono::steady_clock::now();
  setupHostMs_ = std::chrono::duration<double, std::milli>(t1 - t0).count();
  n_ = n;
  if (!withDevice) return;

  NODALS_HIP(hipSetDevice(0));
  NODALS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  red_.init(32);
  n_ = n;
  meanRow_ = (double)nnz / n;
  lanes_ = lanesFor(meanRow_, n);
  std::vector<int> r32(A_.row.begin(), A_.row.end());
  row_.upload(r32);
  col_.upload(A_.col);
  val_.upload(A_.val);
  diag_.upload(A_.diag);
  std::vector<double> l1((std::size_t)n);
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (std::int64_t k = A_.row[(std::size_t)i]; k < A_.row[(std::size_t)i + 1]; ++k) s += std::abs(A_.val[(std::size_t)k]);
    l1[(std::size_t)i] = s;
  }
  l1_.upload(l1);
  b_.alloc((std::size_t)n);
  x_.alloc((std::size_t)n);
  y_.alloc((std::size_t)n);
  rd_.alloc((std::size_t)n);
  zd_.alloc((std::size_t)n);
  Coloring fc;
  if (so_.kind == "mcgs") fc = greedyColoring(n, A_.row.data(), A_.col.data());
  const Coloring* pfc = so_.kind == "mcgs" ? &fc : nullptr;
  if (precision_ == "fp64") {
    amg64_ = std::make_unique<DevAmg<double, double>>();
    amg64_->setup(H_, n, so_, pfc);
    FineView<double, double> v;
    v.n = n;
    v.lanes = lanes_;
    v.meanRow = meanRow_;
    v.row = row_.data();
    v.col = col_.data();
    v.val = val_.data();
    v.diag = diag_.data();
    v.l1 = l1_.data();
    amg64_->setFine(v);
  } else {
    std::vector<float> df(A_.diag.begin(), A_.diag.end()), lf(l1.begin(), l1.end());
    diagf_.upload(df);
    l1f_.upload(lf);
    rf_.alloc((std::size_t)n);
    zf_.alloc((std::size_t)n);
    if (precision_ == "fp32") {
      std::vector<float> vf(A_.val.begin(), A_.val.end());
      valf_.upload(vf);
      amg32_ = std::make_unique<DevAmg<float, float>>();
      amg32_->setup(H_, n, so_, pfc);
      FineView<float, float> v;
      v.n = n;
      v.lanes = lanes_;
      v.meanRow = meanRow_;
    v.meanRow = meanRow_;
      v.row = row_.data();
      v.col = col_.data();
      v.val = valf_.data();
      v.diag = diagf_.data();
      v.l1 = l1f_.data();
      amg32_->setFine(v);
    } else {
      std::vector<bf16> vb(nnz);
      for (std::size_t k = 0; k < nnz; ++k) vb[k] = floatToBf16((float)A_.val[k]);
      valb_.upload(vb);
      amgb_ = std::make_unique<DevAmg<bf16, float>>();
      amgb_->setup(H_, n, so_, pfc);
      FineView<bf16, float> v;
      v.n = n;
      v.lanes = lanes_;
      v.meanRow = meanRow_;
    v.meanRow = meanRow_;
      v.row = row_.data();
      v.col = col_.data();
      v.val = valb_.data();
      v.diag = diagf_.data();
      v.l1 = l1f_.data();
      amgb_->setFine(v);
    }
  }
  NODALS_HIP(hipStreamSynchronize(stream_));
}

SysSolver::~SysSolver() {
  if (stream_) (void)hipStreamDestroy(stream_);
}

void SysSolver::apply(hipStream_t s, const double* x, double* y) {
  spmv<double, double, false, false>(s, lanes_, n_, row_.data(), col_.data(), val_.data(), x, nullptr, y);
}

void SysSolver::precond(hipStream_t s, const double* r, double* z) {
  if (amg64_) {
    cast<double, double>(s, n_, r, rd_.data());
    amg64_->applyGraphed(s, rd_.data(), zd_.data());
    cast<double, double>(s, n_, zd_.data(), z);
  } else if (amg32_) {
    cast<float, double>(s, n_, r, rf_.data());
    amg32_->applyGraphed(s, rf_.data(), zf_.data());
    cast<double, float>(s, n_, zf_.data(), z);
  } else {
    cast<float, double>(s, n_, r, rf_.data());
    amgb_->applyGraphed(s, rf_.data(), zf_.data());
    cast<double, float>(s, n_, zf_.data(), z);
  }
}

std::vector<double> SysSolver::spmvHost(const std::vector<double>& x) {
  requireDevice();
  require((int)x.size() == n_, "spmv: size mismatch");
  x_.upload(x);
  apply(stream_, x_.data(), y_.data());
  NODALS_HIP(hipStreamSynchronize(stream_));
  return y_.host();
}

std::vector<double> SysSolver::precondHost(const std::vector<double>& r) {
  requireDevice();
  require((int)r.size() == n_, "precond: size mismatch");
  x_.upload(r);
  precond(stream_, x_.data(), y_.data());
  NODALS_HIP(hipStreamSynchronize(stream_));
  return y_.host();
}

void SysSolver::setRhs(const std::vector<double>& b) {
  requireDevice();
  require((int)b.size() == n_, "rhs: size mismatch");
  b_.upload(b);
}

KrylovResult SysSolver::solveResident(const std::string& solver, double rtol, int restart, int maxit) {
  requireDevice();
  DevOp A = [this](hipStream_t s, const double* x, double* y) { apply(s, x, y); };
  DevOp M = [this](hipStream_t s, const double* r, double* z) { precond(s, r, z); };
  if (solver == "fgmres") return fgmres(stream_, red_, n_, A, M, b_.data(), rtol, 0.0, maxit, restart, kw_);
  if (solver == "pcg") return pcg(stream_, red_, n_, A, M, b_.data(), rtol, 0.0, maxit, kw_);
  fail("sys solver must be fgmres or pcg");
}

KrylovResult SysSolver::solve(const std::vector<double>& b, std::vector<double>& x, const std::string& solver,
                              double rtol, int restart, int maxit) {
  setRhs(b);
  KrylovResult r = solveResident(solver, rtol, restart, maxit);
  NODALS_HIP(hipStreamSynchronize(stream_));
  x = kw_.x.host();
  return r;
}

double SysSolver::timedSolve(const std::string& solver, double rtol, int restart, int maxit, int reps,
                             KrylovResult* last) {
  hipEvent_t e0, e1;
  NODALS_HIP(hipEventCreate(&e0));
  NODALS_HIP(hipEventCreate(&e1));
  NODALS_HIP(hipEventRecord(e0, stream_));
  KrylovResult r;
  for (int k = 0; k < reps; ++k) r = solveResident(solver, rtol, restart, maxit);
  NODALS_HIP(hipEventRecord(e1, stream_));
  NODALS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  NODALS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (last) *last = r;
  return (double)ms / reps;
}

double SysSolver::timeSpmv(int level, int reps, double* bytes) {
  requireDevice();
  require(reps >= 1, "reps must be >= 1");
  std::function<void()> launch;
  double by = 0;
  const int lv = level;
  if (lv < 0) {
    fill<double>(stream_, n_, x_.data(), 1.0);
    launch = [this]() { apply(stream_, x_.data(), y_.data()); };
    by = (double)A_.col.size() * 12.0 + (n_ + 1) * 4.0 + n_ * 16.0;
  } else {
    require(lv < levels(), "level out of range");
    const CsrD& L = H_.A[(std::size_t)lv];
    const double s = precision_ == "fp64" ? 8.0 : 4.0;
    by = (double)L.col.size() * (s + 4.0) + (L.n + 1) * 4.0 + L.n * 2.0 * s;
    if (amg64_) launch = [this, lv]() { amg64_->levelSpmv(stream_, lv); };
    else if (amg32_) launch = [this, lv]() { amg32_->levelSpmv(stream_, lv); };
    else launch = [this, lv]() { amgb_->levelSpmv(stream_, lv); };
  }
  launch();
  hipEvent_t e0, e1;
  NODALS_HIP(hipEventCreate(&e0));
  NODALS_HIP(hipEventCreate(&e1));
  NODALS_HIP(hipEventRecord(e0, stream_));
  for (int k = 0; k < reps; ++k) launch();
  NODALS_HIP(hipEventRecord(e1, stream_));
  NODALS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  NODALS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (bytes) *bytes = by;
  return (double)ms / reps;
}

}  // namespace nodals
