// Outer Krylov solvers on FP64 vectors with a device preconditioner.
//
// PCG, Richardson and outer Chebyshev follow pressure_pcg /
// pressure_richardson / pressure_cheb_poly (gpu/serial_cuda/h8_main_base.cu:
// 569-726): x starts at zero, convergence is ||r|| <= max(atol, rtol*||b||).
// FGMRES(m) is right-preconditioned and flexible (the preconditioner may be a
// mixed-precision V-cycle), orthogonalises with classical Gram-Schmidt done
// twice (two fused multi-dot kernels per step instead of j+1 single dots),
// and recomputes the true FP64 residual at every restart and at the end, so a
// lower-precision preconditioner cannot change the converged residual.
#pragma once

#include <cmath>
#include <functional>
#include <vector>

#include "la_kernels.hpp"

namespace nodals {
namespace dev {

struct KrylovResult {
  int its = 0;
  double rel = 1.0;
  bool ok = false;
  double r0 = 0.0;
};

// y = A x and z = M r on length-n FP64 device vectors.
using DevOp = std::function<void(hipStream_t, const double*, double*)>;

struct KrylovWork {
  int n = 0, restart = 0;
  DBuf<double> x, r, z, p, q, V, Z;
  void ensure(int n_, int restart_) {
    if (n_ != n) {
      n = n_;
      for (DBuf<double>* b : {&x, &r, &z, &p, &q}) b->alloc((std::size_t)n);
      restart = -1;
    }
    if (restart_ > 0 && restart_ != restart) {
      restart = restart_;
      V.alloc((std::size_t)n * (restart + 1));
      Z.alloc((std::size_t)n * restart);
    }
  }
};

inline KrylovResult pcg(hipStream_t s, Reducer& R, int n, const DevOp& A, const DevOp& M, const double* rhs, double rtol,
                        double atol, int maxit, KrylovWork& W) {
  W.ensure(n, 0);
  fill<double>(s, n, W.x.data(), 0.0);
  copy<double>(s, n, rhs, W.r.data());
  const double r0 = std::sqrt(R.dot(s, n, W.r.data(), W.r.data()));
  require(std::isfinite(r0), "pressure initial norm non-finite", kErrNonFinite);
  KrylovResult out;
  out.r0 = r0;
  if (r0 <= atol) {
    out.ok = true;
    out.rel = 0.0;
    return out;
  }
  M(s, W.r.data(), W.z.data());
  copy<double>(s, n, W.z.data(), W.p.data());
  double rho = R.dot(s, n, W.r.data(), W.z.data());
  require(rho > 0 && std::isfinite(rho), "PCG initial rho non-positive", kErrBreakdown);
  const double target = std::max(atol, rtol * r0);
  for (int k = 0; k < maxit; ++k) {
    A(s, W.p.data(), W.q.data());
    const double pq = R.dot(s, n, W.p.data(), W.q.data());
    require(pq > 0 && std::isfinite(pq), "PCG pAp non-positive", kErrBreakdown);
    const double alpha = rho / pq;
    hipLaunchKernelGGL(k_axpy2<double>, blocksFor(n), 256, 0, s, (long long)n, alpha, W.p.data(), W.q.data(),
                       W.x.data(), W.r.data());
    NODALS_LAUNCHED();
    const double rn = std::sqrt(R.dot(s, n, W.r.data(), W.r.data()));
    out.its = k + 1;
    out.rel = rn / std::max(r0, 1e-300);
    if (rn <= target) {
      out.ok = true;
      break;
    }
    M(s, W.r.data(), W.z.data());
    const double rhon = R.dot(s, n, W.r.data(), W.z.data());
    require(rhon > 0 && std::isfinite(rhon), "PCG rho non-positive", kErrBreakdown);
    hipLaunchKernelGGL(k_xpby<double>, blocksFor(n), 256, 0, s, (long long)n, W.z.data(), rhon / rho, W.p.data());
    NODALS_LAUNCHED();
    rho = rhon;
  }
  return out;
}

// Stationary x += omega_k M r with omega_k = omega (Richardson) or the
// reciprocal Chebyshev roots of [low*lambda, lambda] cycled with period
// `degree` (outer Chebyshev).
inline KrylovResult richardsonCheb(hipStream_t s, Reducer& R, int n, const DevOp& A, const DevOp& M,
                                   const double* rhs, double rtol, double atol, int maxit, KrylovWork& W, bool cheb,
                                   double omega, double lambdaMax, double lowFraction, int degree) {
  W.ensure(n, 0);
  fill<double>(s, n, W.x.data(), 0.0);
  copy<double>(s, n, rhs, W.r.data());
  const double r0 = std::sqrt(R.dot(s, n, W.r.data(), W.r.data()));
  require(std::isfinite(r0), "pressure initial norm non-finite", kErrNonFinite);
  KrylovResult out;
  out.r0 = r0;
  if (r0 <= atol) {
    out.ok = true;
    out.rel = 0.0;
    return out;
  }
  const double target = std::max(atol, rtol * r0);
  const double lo = lowFraction * lambdaMax;
  const double centre = .5 * (lambdaMax + lo), radius = .5 * (lambdaMax - lo);
  const double pi = 3.141592653589793238462643383279502884;
  for (int k = 0; k < maxit; ++k) {
    double w = omega;
    if (cheb) w = 1.0 / (centre - radius * std::cos(pi * (2.0 * (k % degree) + 1.0) / (2.0 * degree)));
    M(s, W.r.data(), W.z.data());
    A(s, W.z.data(), W.q.data());
    hipLaunchKernelGGL(k_axpy2<double>, blocksFor(n), 256, 0, s, (long long)n, w, W.z.data(), W.q.data(), W.x.data(),
                       W.r.data());
    NODALS_LAUNCHED();
    const double rn = std::sqrt(R.dot(s, n, W.r.data(), W.r.data()));
    out.its = k + 1;
    out.rel = rn / std::max(r0, 1e-300);
    require(std::isfinite(out.rel), "pressure residual non-finite", kErrNonFinite);
    if (rn <= target) {
      out.ok = true;
      break;
    }
  }
  return out;
}

inline KrylovResult fgmres(hipStream_t s, Reducer& R, int n, const DevOp& A, const DevOp& M, const double* rhs,
                           double rtol, double atol, int maxit, int restart, KrylovWork& W) {
  require(restart >= 1 && restart <= 31, "FGMRES restart must be in [1,31]");
  W.ensure(n, restart);
  const int m = restart;
  const long long N = n;
  fill<double>(s, n, W.x.data(), 0.0);
  copy<double>(s, n, rhs, W.r.data());
  double beta = std::sqrt(R.dot(s, n, W.r.data(), W.r.data()));
  require(std::isfinite(beta), "FGMRES initial norm non-finite", kErrNonFinite);
  KrylovResult out;
  out.r0 = beta;
  if (beta <= atol) {
    out.ok = true;
    out.rel = 0.0;
    return out;
  }
  const double r0 = beta;
  const double target = std::max(atol, rtol * r0);
  std::vector<double> H((std::size_t)(m + 1) * m), cs((std::size_t)m), sn((std::size_t)m), g((std::size_t)m + 1),
      h((std::size_t)m + 1), h2((std::size_t)m + 1), y((std::size_t)m);
  double* V = W.V.data();
  double* Z = W.Z.data();
  int its = 0;
  while (true) {
    hipLaunchKernelGGL(k_scale_into<double>, blocksFor(N), 256, 0, s, N, 1.0 / beta, W.r.data(), V);
    NODALS_LAUNCHED();
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    bool estConverged = false;
    for (int j = 0; j < m && its < maxit; ++j) {
      double* vj = V + (long long)j * N;
      double* zj = Z + (long long)j * N;
      double* w = V + (long long)(j + 1) * N;
      M(s, vj, zj);
      A(s, zj, w);
      // classical Gram-Schmidt, twice
      R.multidot(s, N, j + 1, V, w, h.data());
      Coef32 c{};
      for (int i = 0; i <= j; ++i) c.c[i] = h[(std::size_t)i];
      hipLaunchKernelGGL(k_multi_axpy, blocksFor(N), 256, 0, s, N, j + 1, V, c, w, -1.0);
      NODALS_LAUNCHED();
      // second pass fused with ||w||^2; after it ||w'||^2 = ||w||^2 - ||h2||^2
      // (V orthonormal, small correction), so no separate norm reduction.
      R.multidot(s, N, j + 1, V, w, h2.data(), true);
      double hh = 0.0;
      for (int i = 0; i <= j; ++i) {
        c.c[i] = h2[(std::size_t)i];
        h[(std::size_t)i] += h2[(std::size_t)i];
        hh += h2[(std::size_t)i] * h2[(std::size_t)i];
      }
      hipLaunchKernelGGL(k_multi_axpy, blocksFor(N), 256, 0, s, N, j + 1, V, c, w, -1.0);
      NODALS_LAUNCHED();
      const double hn = std::sqrt(std::max(h2[(std::size_t)j + 1] - hh, 0.0));
      require(std::isfinite(hn), "FGMRES Arnoldi norm non-finite", kErrNonFinite);
      // Hessenberg column j with the previous rotations applied.
      for (int i = 0; i <= j; ++i) H[(std::size_t)i * m + j] = h[(std::size_t)i];
      H[(std::size_t)(j + 1) * m + j] = hn;
      for (int i = 0; i < j; ++i) {
        const double a = H[(std::size_t)i * m + j], b = H[(std::size_t)(i + 1) * m + j];
        H[(std::size_t)i * m + j] = cs[(std::size_t)i] * a + sn[(std::size_t)i] * b;
        H[(std::size_t)(i + 1) * m + j] = -sn[(std::size_t)i] * a + cs[(std::size_t)i] * b;
      }
      const double a = H[(std::size_t)j * m + j], b = H[(std::size_t)(j + 1) * m + j];
      const double d = std::hypot(a, b);
      require(d > 0 && std::isfinite(d), "FGMRES Givens breakdown", kErrBreakdown);
      cs[(std::size_t)j] = a / d;
      sn[(std::size_t)j] = b / d;
      H[(std::size_t)j * m + j] = d;
      H[(std::size_t)(j + 1) * m + j] = 0.0;
      g[(std::size_t)j + 1] = -sn[(std::size_t)j] * g[(std::size_t)j];
      g[(std::size_t)j] = cs[(std::size_t)j] * g[(std::size_t)j];
      ++its;
      k = j + 1;
      if (std::abs(g[(std::size_t)j + 1]) <= target || hn == 0.0) {
        estConverged = true;
        break;
      }
      hipLaunchKernelGGL(k_scale_into<double>, blocksFor(N), 256, 0, s, N, 1.0 / hn, w, w);
      NODALS_LAUNCHED();
    }
    // x += Z y with H y = g (upper triangular k x k)
    for (int i = k - 1; i >= 0; --i) {
      double t = g[(std::size_t)i];
      for (int l = i + 1; l < k; ++l) t -= H[(std::size_t)i * m + l] * y[(std::size_t)l];
      y[(std::size_t)i] = t / H[(std::size_t)i * m + i];
    }
    Coef32 c{};
    for (int i = 0; i < k; ++i) c.c[i] = y[(std::size_t)i];
    hipLaunchKernelGGL(k_multi_axpy, blocksFor(N), 256, 0, s, N, k, Z, c, W.x.data(), 1.0);
    NODALS_LAUNCHED();
    // true FP64 residual
    A(s, W.x.data(), W.r.data());
    hipLaunchKernelGGL(k_xpby<double>, blocksFor(N), 256, 0, s, N, rhs, -1.0, W.r.data());
    NODALS_LAUNCHED();
    beta = std::sqrt(R.dot(s, N, W.r.data(), W.r.data()));
    require(std::isfinite(beta), "FGMRES residual non-finite", kErrNonFinite);
    out.its = its;
    out.rel = beta / r0;
    if (beta <= target) {
      out.ok = true;
      return out;
    }
    if (its >= maxit) return out;
    (void)estConverged;
  }
}

}  // namespace dev
}  // namespace nodals
