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
  bool warm = false;      // started from a given initial guess
  bool fallback = false;  // the guess did not reduce the residual and was dropped
};

// y = A x and z = M r on length-n FP64 device vectors.
using DevOp = std::function<void(hipStream_t, const double*, double*)>;

// Device-side FGMRES state (doubles): Hessenberg H[i*32+j] (33 x 32, rotated
// in place), Givens cos/sin, the rotated right-hand side g, the two
// Gram-Schmidt coefficient vectors, the next basis scaling and beta^2.
namespace fg {
constexpr int kH = 0, kCs = 1056, kSn = 1088, kG = 1120, kH1 = 1160, kH2 = 1200, kScal = 1240, kBeta2 = 1241,
              kSize = 1248;
// Host-visible (coherent pinned) mirror: residual estimate |g_{j+1}| after
// step j, lucky-breakdown and failure flags, r0, the rotated H and g.
constexpr int hEst = 0, hZero = 32, hBad = 64, hR0 = 96, hH = 100, hG = 100 + 1056, hSize = 100 + 1056 + 40;
}  // namespace fg

struct KrylovWork {
  int n = 0, restart = 0;
  DBuf<double> x, r, z, p, q, V, Z, fgDev;
  double* fgHost = nullptr;
  std::vector<hipEvent_t> stepEv;
  KrylovWork() = default;
  KrylovWork(const KrylovWork&) = delete;
  KrylovWork& operator=(const KrylovWork&) = delete;
  ~KrylovWork() {
    for (hipEvent_t e : stepEv) (void)hipEventDestroy(e);
    if (fgHost) (void)hipHostFree(fgHost);
  }
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
    if (restart_ > 0 && !fgHost) {
      fgDev.alloc(fg::kSize);
      NODALS_HIP(hipHostMalloc((void**)&fgHost, sizeof(double) * fg::hSize, hipHostMallocCoherent));
      stepEv.resize(33);
      for (hipEvent_t& e : stepEv) NODALS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
  }
};

// FGMRES cycle start: r0 = sqrt(beta^2), g = (r0, 0, ...), the scaling of the
// first basis vector (0 when r0 = 0 so nothing non-finite is produced).
// Cycle start over the rows (grid blocksFor(n)): every block finishes the
// reduction of |r|^2 from its kReduceParts block partials with
// k_reduce_final's tree (one partial per thread, wave butterflies, the four
// waves in order), so every block holds the value the separate final kernel
// would store, and writes its rows of v0 = r / beta (and zeroes its rows of
// the solution x when one is given: the first cycle from x = 0); block 0
// also resets the Givens right-hand side and writes beta to the device
// state and the pinned host mirror.
static __global__ void __launch_bounds__(256) k_fgmres_start(long long n, const double* __restrict__ part,
                                                             const double* __restrict__ r, double* __restrict__ v0,
                                                             double* __restrict__ x, double* __restrict__ st,
                                                             double* __restrict__ host, int first) {
  __shared__ double sh[4];
  __shared__ double scal;
  const double r2 = blockSum256(0.0 + part[threadIdx.x], sh);  // 0 + p: k_reduce_final's first add
  if (threadIdx.x == 0) {
    const double beta = sqrt(r2 > 0.0 ? r2 : 0.0);
    scal = beta > 0.0 ? 1.0 / beta : 0.0;
    if (blockIdx.x == 0) {
      st[fg::kBeta2] = r2;
      for (int i = 0; i < 33; ++i) st[fg::kG + i] = 0.0;
      st[fg::kG] = beta;
      st[fg::kScal] = scal;
      host[fg::hG] = beta;
      if (first) host[fg::hR0] = beta;
    }
  }
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    v0[i] = scal * r[i];
    if (x) x[i] = 0.0;
  }
}

// First Gram-Schmidt update fused with the second pass's dot products:
//   w -= V h1;  part[q][block] = V_q . w (q <= j), part[j+1][block] = w . w.
// Reads the basis once for both.  h1 (k values) is first reduced from the
// first pass's kReduceParts block partials partIn by every block, with
// k_reduce_final's tree (one partial per thread, wave butterflies, the four
// waves in order), so each block holds exactly the values the separate
// final kernel would store; block 0 stores them in h1.  partIn and part
// are distinct (other blocks may still read partIn while one writes).
static __global__ void __launch_bounds__(256) k_gs_update_dots(long long n, int k, const double* __restrict__ V,
                                                                const double* __restrict__ partIn,
                                                                double* __restrict__ h1, double* __restrict__ w,
                                                                double* __restrict__ part) {
  __shared__ double c[32];
  __shared__ double sh[4][33];
  {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < k; ++q) {
      const double v = groupSum(0.0 + partIn[(long long)q * kReduceParts + threadIdx.x], 64);
      if (lane == 0) sh[wv][q] = v;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 32) {
      const double h = t < k ? sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t] : 0.0;
      c[t] = h;
      if (blockIdx.x == 0 && t < k) h1[t] = h;
    }
    __syncthreads();
  }
  double sv[33];
#pragma unroll
  for (int q = 0; q < 33; ++q) sv[q] = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    double vq[32];
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 32; ++q) {
      vq[q] = q < k ? V[(long long)q * n + i] : 0.0;
      acc += c[q] * vq[q];
    }
    const double wi = w[i] - acc;
    w[i] = wi;
#pragma unroll
    for (int q = 0; q < 32; ++q) sv[q] += vq[q] * wi;
    sv[32] += wi * wi;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 33; ++q) {
    if (q < k || q == 32) {
      const double t = groupSum(sv[q], 64);
      if (lane == 0) sh[wv][q] = t;
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < k) part[(long long)t * gridDim.x + blockIdx.x] = sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t];
  if (t == k) part[(long long)k * gridDim.x + blockIdx.x] = sh[0][32] + sh[1][32] + sh[2][32] + sh[3][32];
}


// Conditional end of an FGMRES cycle after step k-1, queued by the host when
// it predicts that step converges, so the update runs while the host is
// still waiting on the step.  If the step's residual estimate |g_k| meets
// the target, or its Arnoldi vector vanished (scaling 0), every block solves
// the rotated triangular system H y = g (k x k) by back substitution in the
// host's order, contraction off, so y equals the host's, and applies
// x += Z y with k_multi_axpy's sum.  Otherwise it does nothing.  It reads
// the same values as the host mirror, so the host's check of the step comes
// to the same decision and skips its own update exactly when this one ran.
static __global__ void __launch_bounds__(256) k_fgmres_finish(long long n, int k, double target,
                                                               const double* __restrict__ Z,
                                                               const double* __restrict__ st, double* __restrict__ x) {
  __shared__ double hl[32 * 32], gl[32], yl[32];
  const int t = threadIdx.x;
  // every thread reads the same two values, so the exit is uniform
  if (!(fabs(st[fg::kG + k]) <= target || st[fg::kScal] == 0.0)) return;
  for (int e = t; e < k * 32; e += 256) hl[e] = st[fg::kH + e];
  if (t < k) gl[t] = st[fg::kG + t];
  __syncthreads();
  if (t == 0) {
#pragma clang fp contract(off)
    for (int i = k - 1; i >= 0; --i) {
      double v = gl[i];
      for (int l = i + 1; l < k; ++l) v -= hl[i * 32 + l] * yl[l];
      yl[i] = v / hl[i * 32 + i];
    }
  }
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + t;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = 0; j < k; ++j) acc += yl[j] * Z[(long long)j * n + i];
  x[i] += acc;
}

// Arnoldi step j after the two Gram-Schmidt passes: h = h1 + h2,
// h_{j+1,j} = sqrt(max(|w|^2 - |h2|^2, 0)), previous rotations applied to
// column j, the new rotation, g updated; the residual estimate and flags go
// to host memory, the scaling 1/h_{j+1,j} to device memory.
// One 256-thread block: it first finishes the second pass's reduction
// (h2[q] = sum of nparts block partials, q <= j+1, one wavefront per output
// in turn, fixed order) while other lanes stage h1, the previous rotations
// and g_j in LDS; thread 0 then does the Givens step on LDS operands (a
// serial chain of global-memory round trips otherwise), and the column is
// written back by j+1 lanes at once.
// Second Gram-Schmidt pass, Arnoldi column and normalisation in one launch
// over the basis rows (grid blocksFor(n)).  Every block reduces the second
// pass's coefficients h2[0..j] and |w|^2 (h2[j+1]) from the first update's
// kReduceParts block partials (wave wv takes coefficients wv, wv + 4, ...;
// each lane sums its partials lane, lane + 64, ... in order, then a wave
// butterfly), forms hn = sqrt(|w|^2 - |h2|^2) and applies
//   w = (w - V h2) / hn.
// Block 0 also completes the Hessenberg column h1 + h2, applies the earlier
// Givens rotations and the new one, and writes the rotated column, the
// rotation, the right-hand side and the residual estimate |g_{j+1}| to the
// device state and the pinned host mirror.  Every block forms h2 and hn
// with the same operations in the same order, so all agree bit for bit.
static __global__ void __launch_bounds__(256) k_arnoldi_update(long long n, int j, int nparts,
                                                                const double* __restrict__ part,
                                                                const double* __restrict__ V, double* __restrict__ w,
                                                                double* __restrict__ st, double* __restrict__ host) {
  __shared__ double h2s[34], hcol[34], css[32], sns[32], gj, scal;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int q = wv; q <= j + 1; q += 4) {
    double v = 0.0;
    for (int i = lane; i < nparts; i += 64) v += part[(long long)q * nparts + i];
    v = groupSum(v, 64);
    if (lane == 0) h2s[q] = v;
  }
  const bool lead = blockIdx.x == 0;
  if (lead) {
    if (t <= j) hcol[t] = st[fg::kH1 + t];
    if (t < j) {
      css[t] = st[fg::kCs + t];
      sns[t] = st[fg::kSn + t];
    }
    if (t == 255) gj = st[fg::kG + j];
  }
  __syncthreads();
  if (t == 0) {
    double hh = 0.0;
    for (int i = 0; i <= j; ++i) hh += h2s[i] * h2s[i];
    const double d2 = h2s[j + 1] - hh;
    const double hn = sqrt(d2 > 0.0 ? d2 : 0.0);
    scal = hn > 0.0 ? 1.0 / hn : 0.0;
    if (lead) {
      for (int i = 0; i <= j + 1; ++i) st[fg::kH2 + i] = h2s[i];
      for (int i = 0; i <= j; ++i) hcol[i] = hcol[i] + h2s[i];
      hcol[j + 1] = hn;
      for (int i = 0; i < j; ++i) {
        const double a = hcol[i], b = hcol[i + 1];
        hcol[i] = css[i] * a + sns[i] * b;
        hcol[i + 1] = -sns[i] * a + css[i] * b;
      }
      const double a = hcol[j], b = hcol[j + 1];
      const double d = hypot(a, b);
      const bool bad = !(d > 0.0) || !isfinite(d) || !isfinite(hn);
      const double c = bad ? 0.0 : a / d, sgn = bad ? 0.0 : b / d;
      st[fg::kCs + j] = c;
      st[fg::kSn + j] = sgn;
      hcol[j] = d;
      hcol[j + 1] = 0.0;
      const double g1 = -sgn * gj, g0 = c * gj;
      st[fg::kG + j + 1] = g1;
      st[fg::kG + j] = g0;
      st[fg::kScal] = scal;
      host[fg::hG + j] = g0;
      host[fg::hG + j + 1] = g1;
      host[fg::hZero + j] = hn == 0.0 ? 1.0 : 0.0;
      host[fg::hBad + j] = bad ? 1.0 : 0.0;
      host[fg::hEst + j] = fabs(g1);
    }
  }
  __syncthreads();
  if (lead) {
    if (t <= j + 1) st[fg::kH + t * 32 + j] = hcol[t];
    if (t <= j) host[fg::hH + t * 32 + j] = hcol[t];
  }
  const long long i = (long long)blockIdx.x * blockDim.x + t;
  if (i >= n) return;
  const int k = j + 1;
  double acc = 0.0;
  for (int q = 0; q < k; ++q) acc += h2s[q] * V[(long long)q * n + i];
  w[i] = (w[i] - acc) * scal;
}

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

// FGMRES(m), device-resident Arnoldi.  One step enqueues M, A and four
// kernels for classical Gram-Schmidt done twice (coefficients stay in device
// memory), the Givens step and the basis scaling; the host never waits
// inside a step.  It normally reads step j's residual estimate only after
// step j+1 is queued, so the GPU has the next step in flight; when the last
// contraction predicts that step j converges, it waits for step j first, so
// that usually no step runs past convergence.  A step queued past
// convergence is discarded: it writes only basis vector j+2, Z_{j+1}, column
// j+1 of H and g[j+1..j+2], none of which the k = j+1 solution uses.
// At the end of each cycle the host solves the k x k triangular system,
// applies x += Z y and recomputes the true FP64 residual.
// With an initial guess x0 the convergence target stays relative to ||b||
// (the same rtol as from zero), and a guess whose residual is not below
// ||b|| is dropped (the solve then starts from zero).
// The initial guess is used at its residual-minimising multiple
// alpha = (b, A x0) / (A x0, A x0), so it never raises the residual; a guess
// with alpha <= 0 is dropped.
// With a first direction z0, the first cycle's Arnoldi steps take z0, then
// M^-1 v_0, M^-1 v_1, ... (flexible GMRES admits any directions): the
// space is the plain one augmented by z0, whose optimal multiple the
// least-squares update picks; a poor z0 costs one step.
inline KrylovResult fgmres(hipStream_t s, Reducer& R, int n, const DevOp& A, const DevOp& M, const double* rhs,
                           double rtol, double atol, int maxit, int restart, KrylovWork& W,
                           const double* x0 = nullptr, const double* z0 = nullptr) {
  require(restart >= 1 && restart <= 31, "FGMRES restart must be in [1,31]");
  W.ensure(n, restart);
  const int m = restart;
  const long long N = n;
  const int nb = blocksFor(N);
  double* st = W.fgDev.data();
  double* hs = W.fgHost;
  double* V = W.V.data();
  double* Z = W.Z.data();
  // cycle start from r (first cycle r = rhs, x = 0)
  // zeroX: the cycle starts from x = 0 (the kernel clears x)
  auto start = [&](const double* r, bool first, bool zeroX) {
    R.dotPartials(s, N, r, r);
    hipLaunchKernelGGL(k_fgmres_start, nb, 256, 0, s, N, R.partials(), r, V, zeroX ? W.x.data() : nullptr, st, hs,
                       first ? 1 : 0);
    NODALS_LAUNCHED();
  };
  bool useZ0 = z0 != nullptr;
  auto step = [&](int j) {
    double* vj = V + (long long)j * N;
    double* zj = Z + (long long)j * N;
    double* w = V + (long long)(j + 1) * N;
    if (useZ0) {
      if (j == 0) copy<double>(s, n, z0, zj);
      else M(s, vj - N, zj);  // lag the preconditioned basis by one
    } else {
      M(s, vj, zj);
    }
    A(s, zj, w);
    // classical Gram-Schmidt twice; pass 2's dots share a kernel with pass 1's
    // update, and pass 2's update shares one with the normalisation
    R.multidotPartials(s, N, j + 1, V, w);
    hipLaunchKernelGGL(k_gs_update_dots, kReduceParts, 256, 0, s, N, j + 1, V, R.partials(), st + fg::kH1, w,
                       R.partials2());
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_arnoldi_update, nb, 256, 0, s, N, j, kReduceParts, R.partials2(), V, w, st, hs);
    NODALS_LAUNCHED();
    NODALS_HIP(hipEventRecord(W.stepEv[(std::size_t)j], s));
  };

  KrylovResult out;
  double r0 = -1.0, target = 0.0;
  int its = 0;
  bool firstCycle = true;
  std::vector<double> y((std::size_t)m);
  double prevEst = 0.0;  // residual at the cycle start (r0, then the true residual at each restart)
  if (x0) {
    out.warm = true;
    const double bn = std::sqrt(R.dot(s, N, rhs, rhs));
    require(std::isfinite(bn), "FGMRES right-hand side non-finite", kErrNonFinite);
    out.r0 = r0 = bn;
    target = std::max(atol, rtol * bn);
    // w = A x0; (b, w) and (w, w) in one pass
    A(s, x0, W.q.data());
    double bw[2];
    R.multidot(s, N, 1, rhs, W.q.data(), bw, true);
    const double alpha = bw[1] > 0.0 ? bw[0] / bw[1] : 0.0;
    double rn = bn;
    if (alpha > 0.0 && std::isfinite(alpha)) {
      hipLaunchKernelGGL(k_scale_into<double>, nb, 256, 0, s, N, alpha, x0, W.x.data());
      NODALS_LAUNCHED();
      copy<double>(s, n, rhs, W.r.data());
      axpy<double>(s, n, -alpha, W.q.data(), W.r.data());
      rn = std::sqrt(std::max(bn * bn - bw[0] * alpha, 0.0));
    } else {
      out.fallback = true;
      fill<double>(s, n, W.x.data(), 0.0);
      copy<double>(s, n, rhs, W.r.data());
    }
    if (bn <= atol || rn <= target) {
      NODALS_HIP(hipStreamSynchronize(s));
      out.ok = true;
      out.rel = bn > 0.0 ? rn / bn : 0.0;
      return out;
    }
    start(W.r.data(), false, false);
    prevEst = rn;
    firstCycle = false;
  } else {
    start(rhs, true, true);
    NODALS_HIP(hipEventRecord(W.stepEv[32], s));
  }
  while (true) {
    const int limit = std::min(m, maxit - its);
    int queued = 0, checked = -1, k = 0;
    int finishQueued = -1;  // step after which k_fgmres_finish was queued
    bool done = false;
    auto check = [&](int q) {
      NODALS_HIP(hipEventSynchronize(W.stepEv[(std::size_t)q]));
      require(hs[fg::hBad + q] == 0.0, "FGMRES Givens breakdown", kErrBreakdown);
      k = q + 1;
      checked = q;
      if (hs[fg::hEst + q] <= target || hs[fg::hZero + q] != 0.0) done = true;
    };
    // Residual estimate expected after the oldest unchecked step, from the
    // last observed contraction; when it is predicted to converge, the host
    // waits for it instead of queueing one more step.
    auto predictsConverged = [&]() {
      if (checked < 0) return false;
      const double e1 = hs[fg::hEst + checked];
      const double e0 = checked > 0 ? hs[fg::hEst + checked - 1] : prevEst;
      return e0 > 0.0 && e1 * (e1 / e0) <= target;
    };
    while (!done) {
      if (checked + 1 <= queued - 1 && predictsConverged()) {
        const int q = checked + 1;
        if (q == queued - 1) {
          // nothing is queued after step q: let the device apply the update
          // if the step converges while the host waits for it
          hipLaunchKernelGGL(k_fgmres_finish, nb, 256, 0, s, N, q + 1, target, Z, st, W.x.data());
          NODALS_LAUNCHED();
          finishQueued = q;
        }
        check(q);
        continue;
      }
      if (queued < limit) {
        step(queued);
        ++queued;
        if (firstCycle && queued == 1) {
          // r0 is known once the start kernels have run (step 0 is in flight)
          NODALS_HIP(hipEventSynchronize(W.stepEv[32]));
          r0 = hs[fg::hR0];
          require(std::isfinite(r0), "FGMRES initial norm non-finite", kErrNonFinite);
          out.r0 = r0;
          if (r0 <= atol) {
            NODALS_HIP(hipStreamSynchronize(s));
            out.ok = true;
            out.rel = 0.0;
            return out;
          }
          target = std::max(atol, rtol * r0);
          prevEst = r0;
        }
      } else if (checked == queued - 1) {
        break;
      }
      const int q = checked + 1;
      if (q > queued - 1) continue;
      // check the oldest unchecked step when a newer one is already queued
      // or when no more can be queued
      if (queued - 1 > q || queued == limit) check(q);
    }
    its += k;
    firstCycle = false;
    // x += Z y, H y = g (upper triangular k x k, rotated H in host memory),
    // unless k_fgmres_finish has applied it after the last checked step
    if (!(done && finishQueued == checked)) {
      for (int i = k - 1; i >= 0; --i) {
        double t = hs[fg::hG + i];
        for (int l = i + 1; l < k; ++l) t -= hs[fg::hH + i * 32 + l] * y[(std::size_t)l];
        y[(std::size_t)i] = t / hs[fg::hH + i * 32 + i];
      }
      Coef32 c{};
      for (int i = 0; i < k; ++i) c.c[i] = y[(std::size_t)i];
      hipLaunchKernelGGL(k_multi_axpy, nb, 256, 0, s, N, k, Z, c, W.x.data(), 1.0);
      NODALS_LAUNCHED();
    }
    // A cycle whose residual estimate met the target ends the solve: the
    // estimate is |b - A x| up to FP64 rounding (w = A z_j and its
    // orthogonalisation are FP64 whatever the preconditioner's precision),
    // so no residual pass is spent on it.
    if (done && k > 0 && hs[fg::hEst + k - 1] <= target) {
      out.its = its;
      out.rel = hs[fg::hEst + k - 1] / r0;
      out.ok = true;
      return out;
    }
    // true FP64 residual for the restart
    A(s, W.x.data(), W.r.data());
    hipLaunchKernelGGL(k_xpby<double>, nb, 256, 0, s, N, rhs, -1.0, W.r.data());
    NODALS_LAUNCHED();
    const double beta = std::sqrt(R.dot(s, N, W.r.data(), W.r.data()));
    require(std::isfinite(beta), "FGMRES residual non-finite", kErrNonFinite);
    out.its = its;
    out.rel = beta / r0;
    if (beta <= target) {
      out.ok = true;
      return out;
    }
    if (its >= maxit) return out;
    start(W.r.data(), false, false);
    prevEst = beta;
  }
}

}  // namespace dev
}  // namespace nodals
