// Explicit SPD system context (pressure-Poisson kernel bench and generic
// sparse solves): FP64 CSR operator, device SA hierarchy, FP64 outer Krylov.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../host/amg_setup.hpp"
#include "amg.hpp"
#include "krylov.hpp"

namespace nodals {

class SysSolver {
 public:
  // withDevice=false stops after the host hierarchy (no HIP call).
  SysSolver(int n, const std::int64_t* row, const std::int32_t* col, const double* val, const std::string& options,
            bool withDevice = true);
  ~SysSolver();
  SysSolver(const SysSolver&) = delete;
  SysSolver& operator=(const SysSolver&) = delete;

  int n() const { return n_; }
  long long nnz() const { return (long long)A_.col.size(); }
  int levels() const { return (int)H_.A.size(); }
  const AmgHierarchy& hierarchy() const { return H_; }
  double setupHostMs() const { return setupHostMs_; }

  std::vector<double> spmvHost(const std::vector<double>& x);
  std::vector<double> precondHost(const std::vector<double>& r);
  void setRhs(const std::vector<double>& b);
  dev::KrylovResult solve(const std::vector<double>& b, std::vector<double>& x, const std::string& solver,
                          double rtol, int restart, int maxit);
  double timedSolve(const std::string& solver, double rtol, int restart, int maxit, int reps,
                    dev::KrylovResult* last);
  double timeSpmv(int level, int reps, double* bytes);

 private:
  void requireDevice() const {
    require(stream_ != nullptr, "system context was created without a device (host setup only)");
  }
  dev::KrylovResult solveResident(const std::string& solver, double rtol, int restart, int maxit);
  void apply(hipStream_t s, const double* x, double* y);
  void precond(hipStream_t s, const double* r, double* z);

  std::string precision_ = "fp64", hierarchy_ = "sa";
  CfOptions cf_;
  dev::SmootherOptions so_;
  AmgOptions ao_;
  CsrD A_;
  AmgHierarchy H_;
  int n_ = 0, lanes_ = 8;
  double meanRow_ = 8.0;
  double setupHostMs_ = 0;
  hipStream_t stream_ = nullptr;
  dev::Reducer red_;
  dev::KrylovWork kw_;
  DBuf<int> row_, col_;
  DBuf<double> val_, diag_, l1_, b_, x_, y_, rd_, zd_;
  DBuf<float> valf_, diagf_, l1f_, rf_, zf_;
  DBuf<bf16> valb_;
  std::unique_ptr<dev::DevAmg<double, double>> amg64_;
  std::unique_ptr<dev::DevAmg<float, float>> amg32_;
  std::unique_ptr<dev::DevAmg<bf16, float>> amgb_;
};

}  // namespace nodals
