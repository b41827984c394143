// Pointwise k-omega SST-2003 and wall-law physics for the SST solver path.
//
// Device functions, evaluated per quadrature point (SST local state) or per
// wall face (Spalding root, five-point profile fit, omega wall target).  The
// formulas are the published ones:
//   - Menter, Kuntz & Langtry (2003), "Ten years of industrial experience
//     with the SST turbulence model": blending F1/F2, nu_t = a1 k /
//     max(a1 omega, S F2), production limiter 10 beta* k omega, cross
//     diffusion CD_kw.
//   - Spalding (1961) single-formula wall law
//       y+ = u+ + e^{-kappa B} (e^{kappa u+} - 1 - kappa u+ - (kappa u+)^2/2
//            - (kappa u+)^3/6),  kappa = 0.4, B = 5.5.
// The reference's portable cores (modules/physics/sst2003m_core.hpp,
// modules/sst/*.inc, modules/wall/portable/spalding_core.hpp:11-365,
// of_auto_core.hpp:16-66) fix the floors, guards, iteration counts and
// operation order; they are followed so that results agree with the
// reference to rounding.
#pragma once

// The same functions also compile as plain host C++ (the checker builds
// them with g++ to compare against the reference cores bit for bit).
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NODALS_SST_FN __host__ __device__ inline
#else
#include <cmath>
#define NODALS_SST_FN inline
#endif

namespace nodals {
namespace sst {
#if !defined(__HIPCC__)
using std::exp;
using std::expm1;
using std::fabs;
using std::fmax;
using std::fmin;
using std::isfinite;
using std::log;
using std::sqrt;
using std::tanh;
#endif

// SST-2003 model constants (modules/sst/sst_constants.inc).
constexpr double kBetaStar = 0.09, kA1 = 0.31;
constexpr double kSigmaK1 = 0.85, kSigmaK2 = 1.00, kSigmaW1 = 0.50, kSigmaW2 = 0.856;
constexpr double kBeta1 = 0.075, kBeta2 = 0.0828, kGamma1 = 5.0 / 9.0, kGamma2 = 0.44;

struct Local {
  double nut, F1, F2, sigmaK, sigmaW, beta, gamma, strain;
  double pk, srcOmega, reactK, reactOmega, omegaProd, omegaCross;
  int kFloored, omegaFloored;
};

// Wall distance of the straight-pipe case: radius - r, floored
// (modules/sst/sst_pipe_distance.inc).
NODALS_SST_FN double pipeDistance(double x, double y, double radius) {
  const double r = sqrt(x * x + y * y);
  return fmax(radius - r, fmax(1e-10 * radius, 1e-12));
}

NODALS_SST_FN double floorDistance(double d, double radius) {
  return fmax(d, fmax(1e-10 * radius, 1e-12));
}

// Local SST state at one point from the velocity gradient gu (row-major
// du_i/dx_j), k, omega and their gradients, at wall distance d (already
// floored).
NODALS_SST_FN Local local(double d, double nu, const double* gu, double kRaw, double omRaw,
                                       const double* gk, const double* gw, double kFloor, double omFloor,
                                       bool deviatoric) {
  Local q{};
  q.kFloored = !(kRaw > kFloor);
  q.omegaFloored = !(omRaw > omFloor);
  const double k = fmax(kRaw, kFloor), om = fmax(omRaw, omFloor);
  // strain-rate magnitude S = sqrt(2 S_ij S_ij)
  double ss = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double sij = 0.5 * (gu[3 * i + j] + gu[3 * j + i]);
      ss += sij * sij;
    }
  const double S2 = fmax(0.0, 2.0 * ss), S = sqrt(S2);
  q.strain = S;
  // blending functions
  const double dotkw = gk[0] * gw[0] + gk[1] * gw[1] + gk[2] * gw[2];
  const double CD = fmax((2 * kSigmaW2) * dotkw / om, 1e-10);
  const double sqk = sqrt(k), d2 = d * d;
  const double t1 = sqk / (kBetaStar * om * d);
  const double t2 = 500.0 * nu / (d2 * om);
  const double t3 = (4 * kSigmaW2) * k / (CD * d2);
  const double arg1 = fmin(fmax(t1, t2), t3);
  const double arg2 = fmax(2.0 * t1, t2);
  const double a1sq = arg1 * arg1;
  q.F1 = tanh(a1sq * a1sq);
  q.F2 = tanh(arg2 * arg2);
  q.sigmaK = q.F1 * kSigmaK1 + (1.0 - q.F1) * kSigmaK2;
  q.sigmaW = q.F1 * kSigmaW1 + (1.0 - q.F1) * kSigmaW2;
  q.beta = q.F1 * kBeta1 + (1.0 - q.F1) * kBeta2;
  q.gamma = q.F1 * kGamma1 + (1.0 - q.F1) * kGamma2;
  // eddy viscosity
  q.nut = kA1 * k / fmax(kA1 * om, S * q.F2);
  q.nut = fmax(q.nut, 0.0);
  // sources: production (optionally of the deviatoric strain), limiter,
  // destruction rates, omega production and cross diffusion
  const double divU = gu[0] + gu[4] + gu[8];
  const double prodMetric = deviatoric ? fmax(0.0, S2 - (2.0 / 3.0) * divU * divU) : S2;
  const double prodRaw = q.nut * prodMetric;
  const double prodLim = (10 * kBetaStar) * k * om;
  q.pk = fmin(prodRaw, prodLim);
  q.reactK = kBetaStar * om;
  q.reactOmega = q.beta * om;
  const double nutTiny = fmax(1e-12 * nu, 1e-20);
  const double omegaProd = q.nut > nutTiny ? q.gamma * q.pk / q.nut : q.gamma * prodMetric;
  const double cross = (2 * kSigmaW2) * (1.0 - q.F1) * dotkw / om;
  q.omegaProd = omegaProd;
  q.omegaCross = cross;
  q.srcOmega = omegaProd + cross;
  return q;
}

// ---------------------------------------------------------------- Spalding
constexpr double kKappa = 0.4, kB = 5.5;
constexpr double kE = 9.025013499434122;  // exp(kappa B)

// y+ of Spalding's law at u+; the bracket e^x - 1 - x - x^2/2 - x^3/6 is
// taken from its Taylor series x^4/24 + x^5/120 + x^6/720 + x^7/5040 when
// |x| < 1e-3 (cancellation) and from expm1 otherwise.
NODALS_SST_FN double yPlusOf(double up, double kap, double B) {
  const double x = kap * up;
  double rem;
  if (fabs(x) < 1e-3) {
    const double x2 = x * x, x4 = x2 * x2;
    rem = x4 * (1.0 / 24 + x / 120 + x2 / 720 + x2 * x / 5040);
  } else {
    rem = expm1(x) - x - .5 * x * x - x * x * x / 6;
  }
  return up + exp(-kap * B) * rem;
}

// d y+ / d u+.
NODALS_SST_FN double dYPlusOf(double up, double kap, double B) {
  const double x = kap * up;
  double drem;
  if (fabs(x) < 1e-3) {
    const double x2 = x * x, x3 = x2 * x;
    drem = kap * x3 * (1.0 / 6.0 + x / 24.0 + x2 / 120.0 + x2 * x / 720.0 + x2 * x2 / 5040.0);
  } else {
    drem = kap * (exp(x) - 1.0 - x - 0.5 * x * x);
  }
  return 1.0 + exp(-kap * B) * drem;
}

// g(u+) = u+ y+(u+) - Re_y, whose root gives u+ for Re_y = y U / nu, and g'.
NODALS_SST_FN double spaldingResidual(double up, double reY) { return up * yPlusOf(up, kKappa, kB) - reY; }
NODALS_SST_FN double spaldingResidualDeriv(double up) {
  return yPlusOf(up, kKappa, kB) + up * dYPlusOf(up, kKappa, kB);
}

// Friction velocity from the slip speed at distance y: bracketed root of g,
// started from the viscous (sqrt Re_y) or log-law fixed-point guess, six
// safeguarded Newton steps (a step leaving the bracket falls back to its
// midpoint).  Returns false where the law has no valid root.
NODALS_SST_FN bool spaldingRoot(double slip, double y, double nu, double& ut, double& yp, double& beta) {
  const double U = fabs(slip);
  const double reY = y * U / nu;
  if (!(y > 0) || !(nu > 0) || !isfinite(U)) return false;
  if (reY <= 1e-14) {
    ut = 0;
    yp = 0;
    beta = nu / y;
    return true;
  }
  double lo = 0.0;
  double hi = fmax(1.0, sqrt(reY) + 1.0);
  for (int k = 0; k < 40 && spaldingResidual(hi, reY) < 0.0; ++k) hi *= 2.0;
  const double gh = spaldingResidual(hi, reY);
  if (!isfinite(gh) || gh < 0.0) return false;
  double up;
  if (reY < 50.0) {
    up = sqrt(reY);
  } else {
    constexpr double invK = 2.5;  // 1 / kappa
    up = invK * log(fmax(kE * reY, 1.0000001));
    for (int k = 0; k < 4; ++k) {
      const double arg = kE * reY / fmax(up, 1e-12);
      if (arg > 1.0) up = invK * log(arg);
    }
  }
  if (!(up > lo && up < hi) || !isfinite(up)) up = 0.5 * (lo + hi);
  for (int k = 0; k < 6; ++k) {
    const double g = spaldingResidual(up, reY);
    if (!isfinite(g)) return false;
    if (fabs(g) <= 1e-13 * fmax(reY, 1e-300)) break;
    if (g > 0.0) hi = up;
    else lo = up;
    const double dg = spaldingResidualDeriv(up);
    double cand = (isfinite(dg) && dg > 0.0) ? up - g / dg : 0.5 * (lo + hi);
    if (!(cand > lo && cand < hi) || !isfinite(cand)) cand = 0.5 * (lo + hi);
    up = cand;
  }
  const double gf = spaldingResidual(up, reY);
  if (!isfinite(gf) || !(up > 0.0)) return false;
  ut = U / up;
  yp = y * ut / nu;
  beta = ut * ut / U;
  return isfinite(beta) && beta > 0.0 && isfinite(yp);
}

// Mean squared Spalding residual of five samples at friction velocity ut.
NODALS_SST_FN double profileObjective(const double* U, const double* y, double nu, double ut) {
  if (!(ut > 0.0) || !(nu > 0.0) || !isfinite(ut)) return 1.0e300;
  double J = 0.0;
  for (int m = 0; m < 5; ++m) {
    const double r = yPlusOf(U[m] / ut, kKappa, kB) - y[m] * ut / nu;
    if (!isfinite(r)) return 1.0e300;
    J += r * r;
  }
  return J / 5.0;
}

// Objective, gradient and Gauss-Newton Hessian in one pass.
NODALS_SST_FN bool profileGN(const double* U, const double* y, double nu, double ut, double& J, double& g,
                                          double& h) {
  if (!(nu > 0.0 && ut > 0.0) || !isfinite(ut)) return false;
  double sum = 0.0;
  g = 0.0;
  h = 0.0;
  for (int m = 0; m < 5; ++m) {
    const double up = U[m] / ut;
    const double r = yPlusOf(up, kKappa, kB) - y[m] * ut / nu;
    const double dr = -(up / ut) * dYPlusOf(up, kKappa, kB) - y[m] / nu;
    if (!isfinite(r) || !isfinite(dr)) return false;
    sum += r * r;
    g += r * dr;
    h += dr * dr;
  }
  J = sum / 5.0;
  return isfinite(J) && isfinite(g) && isfinite(h);
}

struct Profile {
  double uTau, yPlus, beta, uMid, fitRms, span;
  int valid;
};

// Least-squares friction velocity of a five-point wall-normal profile
// (speeds U, signed tangential speeds Us, heights y increasing): Gauss-Newton
// in a bracket set from the pointwise Spalding roots, with a halving line
// search; a golden-section search in log(ut) if Gauss-Newton fails.
NODALS_SST_FN Profile fitProfile5(const double* U, const double* Us, const double* y, double nu) {
  Profile p{0, 0, 0, 0, 0, 0, 0};
  if (!(nu > 0.0)) return p;
  double minRoot = 1.0e300, maxRoot = 0.0, sumRoot = 0.0, Umax = 0.0;
  int nRoot = 0;
  for (int m = 0; m < 5; ++m) {
    Umax = fmax(Umax, U[m]);
    double r = 0.0, yp = 0.0, bt = 0.0;
    if (spaldingRoot(U[m], y[m], nu, r, yp, bt) && r > 0.0 && isfinite(r)) {
      minRoot = fmin(minRoot, r);
      maxRoot = fmax(maxRoot, r);
      sumRoot += r;
      ++nRoot;
    }
  }
  for (int m = 1; m < 5; ++m)
    if (!(y[m - 1] < y[m])) return p;
  if (Umax <= 1.0e-14) {
    p.uMid = Us[2];
    p.beta = nu / y[2];
    p.valid = isfinite(p.beta) && p.beta > 0.0;
    return p;
  }
  double scale = nRoot > 0 ? sumRoot / (double)nRoot : sqrt(nu * Umax / fmax(y[2], 1.0e-30));
  if (!(scale > 0.0) || !isfinite(scale)) scale = 1.0;
  double lo = fmax(0.05 * scale, 1.0e-10);
  double hi = fmax(20.0 * scale, 100.0 * lo);
  if (nRoot > 0) {
    lo = fmax(fmin(lo, 0.20 * minRoot), 1.0e-10);
    hi = fmax(hi, 5.0 * maxRoot);
  }
  double ut = fmin(hi, fmax(lo, scale));
  bool gnOK = true;
  for (int it = 0; it < 12; ++it) {
    double J0 = 0.0, g = 0.0, h = 0.0;
    if (!profileGN(U, y, nu, ut, J0, g, h) || J0 >= 1.0e299 || !(h > 0.0)) {
      gnOK = false;
      break;
    }
    double step = g / h;
    double trial = fmin(hi, fmax(lo, ut - step));
    double Jt = profileObjective(U, y, nu, trial);
    for (int ls = 0; ls < 10 && (!isfinite(Jt) || Jt >= 1.0e299 || Jt > J0); ++ls) {
      step *= 0.5;
      trial = fmin(hi, fmax(lo, ut - step));
      Jt = profileObjective(U, y, nu, trial);
    }
    if (!isfinite(Jt) || Jt >= 1.0e299 || Jt > J0) {
      gnOK = false;
      break;
    }
    const double rel = fabs(trial - ut) / fmax(ut, 1.0e-30);
    ut = trial;
    if (rel < 1.0e-10) break;
  }
  double J = profileObjective(U, y, nu, ut);
  if (!gnOK || !(ut > 0.0) || !isfinite(ut) || !isfinite(J) || J >= 1.0e299) {
    // golden-section search on log(ut) over [lo, hi]
    double xa = log(lo), xb = log(hi);
    const double phi = 0.6180339887498948482;
    double x1 = xb - phi * (xb - xa), x2 = xa + phi * (xb - xa);
    double f1 = profileObjective(U, y, nu, exp(x1));
    double f2 = profileObjective(U, y, nu, exp(x2));
    for (int it = 0; it < 28; ++it) {
      if (f1 <= f2) {
        xb = x2;
        x2 = x1;
        f2 = f1;
        x1 = xb - phi * (xb - xa);
        f1 = profileObjective(U, y, nu, exp(x1));
      } else {
        xa = x1;
        x1 = x2;
        f1 = f2;
        x2 = xa + phi * (xb - xa);
        f2 = profileObjective(U, y, nu, exp(x2));
      }
    }
    ut = exp(f1 <= f2 ? x1 : x2);
    J = f1 <= f2 ? f1 : f2;
  }
  if (!(ut > 0.0) || !isfinite(ut) || !isfinite(J) || J >= 1.0e299) return p;
  p.uTau = ut;
  p.yPlus = y[2] * ut / nu;
  p.uMid = Us[2];
  const double Ua = fabs(p.uMid);
  p.beta = Ua > 1.0e-14 ? ut * ut / Ua : nu / y[2];
  p.fitRms = sqrt(fmax(J, 0.0));
  p.span = (nRoot >= 2 && maxRoot >= minRoot && minRoot < 1.0e299) ? (maxRoot - minRoot) / fmax(ut, 1.0e-30) : 0.0;
  p.valid = isfinite(p.yPlus) && isfinite(p.beta) && p.beta > 0.0 && isfinite(p.fitRms) && isfinite(p.span);
  return p;
}

// Spalding least squares over n of the five wall-normal sample slots
// (modules/wall/portable/spalding_active.hpp:24-128): for n = 1..4 the
// physical slots {2}, {0,4}, {0,2,4}, {0,1,3,4}; for n = 5 the five-point
// fit above.  Same bracket, Gauss-Newton with halving line search and
// golden-section fallback as the five-point fit, on the n active points;
// the reference height and signed speed are those of point n/2.
NODALS_SST_FN int activeSlot(int n, int j) {
  const int ids[5][5] = {{2, -1, -1, -1, -1}, {0, 4, -1, -1, -1}, {0, 2, 4, -1, -1}, {0, 1, 3, 4, -1},
                         {0, 1, 2, 3, 4}};
  return (n >= 1 && n <= 5 && j >= 0 && j < n) ? ids[n - 1][j] : -1;
}
NODALS_SST_FN unsigned activeMask(int n) {
  return n == 1 ? 0x04u : n == 2 ? 0x11u : n == 3 ? 0x15u : n == 4 ? 0x1bu : n == 5 ? 0x1fu : 0u;
}
NODALS_SST_FN double profileObjectiveN(const double* U, const double* y, int n, double nu, double ut) {
  if (!(n >= 1 && n <= 5 && ut > 0.0 && nu > 0.0) || !isfinite(ut)) return 1.0e300;
  double J = 0.0;
  for (int m = 0; m < n; ++m) {
    const double r = yPlusOf(U[m] / ut, kKappa, kB) - y[m] * ut / nu;
    if (!isfinite(r)) return 1.0e300;
    J += r * r;
  }
  return J / (double)n;
}
NODALS_SST_FN bool profileGNN(const double* U, const double* y, int n, double nu, double ut, double& J, double& g,
                              double& h) {
  if (!(n >= 1 && n <= 5 && nu > 0.0 && ut > 0.0) || !isfinite(ut)) return false;
  double sum = 0.0;
  g = 0.0;
  h = 0.0;
  for (int m = 0; m < n; ++m) {
    const double up = U[m] / ut;
    const double r = yPlusOf(up, kKappa, kB) - y[m] * ut / nu;
    const double dr = -(up / ut) * dYPlusOf(up, kKappa, kB) - y[m] / nu;
    if (!isfinite(r) || !isfinite(dr)) return false;
    sum += r * r;
    g += r * dr;
    h += dr * dr;
  }
  J = sum / (double)n;
  return isfinite(J) && isfinite(g) && isfinite(h);
}
NODALS_SST_FN Profile fitActive(const double* Uall, const double* Sall, const double* yall, int n, double nu) {
  if (n == 5) return fitProfile5(Uall, Sall, yall, nu);
  Profile p{0, 0, 0, 0, 0, 0, 0};
  if (n < 1 || n > 4 || !(nu > 0.0)) return p;
  double U[5] = {0, 0, 0, 0, 0}, S[5] = {0, 0, 0, 0, 0}, y[5] = {0, 0, 0, 0, 0};
  double minRoot = 1.0e300, maxRoot = 0.0, sumRoot = 0.0, Umax = 0.0;
  int nRoot = 0;
  for (int m = 0; m < n; ++m) {
    const int id = activeSlot(n, m);
    U[m] = Uall[id];
    S[m] = Sall[id];
    y[m] = yall[id];
    if (!(y[m] > 0.0) || !isfinite(y[m]) || !isfinite(U[m]) || !isfinite(S[m]) || (m > 0 && !(y[m] > y[m - 1])))
      return p;
    Umax = fmax(Umax, U[m]);
    double r = 0.0, yp = 0.0, bt = 0.0;
    if (spaldingRoot(U[m], y[m], nu, r, yp, bt) && r > 0.0 && isfinite(r)) {
      minRoot = fmin(minRoot, r);
      maxRoot = fmax(maxRoot, r);
      sumRoot += r;
      ++nRoot;
    }
  }
  const int mid = n / 2;
  p.uMid = S[mid];
  if (Umax <= 1.0e-14) {
    p.beta = nu / y[mid];
    p.valid = isfinite(p.beta) && p.beta > 0.0;
    return p;
  }
  double scale = nRoot ? sumRoot / nRoot : sqrt(nu * Umax / fmax(y[mid], 1.0e-30));
  if (!(scale > 0.0) || !isfinite(scale)) scale = 1.0;
  double lo = fmax(0.05 * scale, 1.0e-10), hi = fmax(20.0 * scale, 100.0 * lo);
  if (nRoot) {
    lo = fmax(fmin(lo, 0.20 * minRoot), 1.0e-10);
    hi = fmax(hi, 5.0 * maxRoot);
  }
  double ut = fmin(hi, fmax(lo, scale));
  bool gnOK = true;
  for (int it = 0; it < 12; ++it) {
    double J0 = 0.0, g = 0.0, h = 0.0;
    if (!profileGNN(U, y, n, nu, ut, J0, g, h) || J0 >= 1.0e299 || !(h > 0.0)) {
      gnOK = false;
      break;
    }
    double step = g / h, trial = fmin(hi, fmax(lo, ut - step));
    double Jt = profileObjectiveN(U, y, n, nu, trial);
    for (int ls = 0; ls < 10 && (!isfinite(Jt) || Jt >= 1.0e299 || Jt > J0); ++ls) {
      step *= 0.5;
      trial = fmin(hi, fmax(lo, ut - step));
      Jt = profileObjectiveN(U, y, n, nu, trial);
    }
    if (!isfinite(Jt) || Jt >= 1.0e299 || Jt > J0) {
      gnOK = false;
      break;
    }
    const double rel = fabs(trial - ut) / fmax(ut, 1.0e-30);
    ut = trial;
    if (rel < 1.0e-10) break;
  }
  double J = profileObjectiveN(U, y, n, nu, ut);
  if (!gnOK || !(ut > 0.0) || !isfinite(ut) || !isfinite(J) || J >= 1.0e299) {
    double xa = log(lo), xb = log(hi);
    const double phi = 0.6180339887498948482;
    double x1 = xb - phi * (xb - xa), x2 = xa + phi * (xb - xa);
    double f1 = profileObjectiveN(U, y, n, nu, exp(x1)), f2 = profileObjectiveN(U, y, n, nu, exp(x2));
    for (int it = 0; it < 28; ++it) {
      if (f1 <= f2) {
        xb = x2;
        x2 = x1;
        f2 = f1;
        x1 = xb - phi * (xb - xa);
        f1 = profileObjectiveN(U, y, n, nu, exp(x1));
      } else {
        xa = x1;
        x1 = x2;
        f1 = f2;
        x2 = xa + phi * (xb - xa);
        f2 = profileObjectiveN(U, y, n, nu, exp(x2));
      }
    }
    ut = exp(f1 <= f2 ? x1 : x2);
    J = fmin(f1, f2);
  }
  if (!(ut > 0.0) || !isfinite(ut) || !isfinite(J) || J >= 1.0e299) return p;
  p.uTau = ut;
  p.yPlus = y[mid] * ut / nu;
  p.beta = fabs(p.uMid) > 1.0e-14 ? ut * ut / fabs(p.uMid) : nu / y[mid];
  p.fitRms = sqrt(fmax(J, 0.0));
  p.span = nRoot >= 2 && maxRoot >= minRoot ? (maxRoot - minRoot) / fmax(ut, 1.0e-30) : 0.0;
  p.valid = isfinite(p.yPlus) && isfinite(p.beta) && p.beta > 0.0 && isfinite(p.fitRms) && isfinite(p.span);
  return p;
}

// Tangential Spalding fit of a stationary wall with unit normal n
// (modules/wall/portable/spalding_vector_core.hpp:54-84 and
// spalding_active.hpp:129-154): the speeds are the norms of the samples'
// tangential projections; the Robin coefficient is u_tau^2 over the
// tangential speed of the Picard functional pic, or the fit's own
// coefficient when that speed is below 1e-6 max(|u_mid|, 1).  zeroLimit
// (the reference's STAGE27_SPALDING_ZERO_LIMIT) keeps the fit's molecular
// limit nu/y when an n < 5 point fit returns u_tau = 0.
struct VecFit {
  double beta, uTau, yPlus;
  int valid;
};
NODALS_SST_FN VecFit fitTangential(const double sx[5], const double sy[5], const double sz[5], const double y[5],
                                   const double n[3], const double pic[3], double nu, int count, bool zeroLimit) {
  VecFit out{0, 0, 0, 0};
  const double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (count < 1 || count > 5 || !(nu > 0.0) || !isfinite(nu) || !isfinite(n2) || !(fabs(n2 - 1.0) <= 1.e-9) ||
      !isfinite(pic[0]) || !isfinite(pic[1]) || !isfinite(pic[2]))
    return out;
  const unsigned mask = activeMask(count);
  double U[5] = {0, 0, 0, 0, 0}, Y[5] = {0, 0, 0, 0, 0};
  double prevY = 0.0;
  for (int i = 0; i < 5; ++i) {
    if (count < 5 && !(mask & (1u << i))) {
      Y[i] = y[i];
      continue;
    }
    if (!isfinite(sx[i]) || !isfinite(sy[i]) || !isfinite(sz[i]) || !isfinite(y[i]) || !(y[i] > 0.0)) return out;
    if (count < 5) {
      if (!(y[i] > prevY)) return out;
      prevY = y[i];
    } else if (i > 0 && !(y[i] > y[i - 1])) {
      return out;
    }
    const double vn = sx[i] * n[0] + sy[i] * n[1] + sz[i] * n[2];
    const double tx = sx[i] - vn * n[0], ty = sy[i] - vn * n[1], tz = sz[i] - vn * n[2];
    U[i] = sqrt(tx * tx + ty * ty + tz * tz);
    Y[i] = y[i];
    if (!isfinite(U[i])) return out;
  }
  const Profile p = fitActive(U, U, Y, count, nu);
  if (!p.valid) return out;
  const double pn = pic[0] * n[0] + pic[1] * n[1] + pic[2] * n[2];
  const double px = pic[0] - pn * n[0], py = pic[1] - pn * n[1], pz = pic[2] - pn * n[2];
  const double speed = sqrt(px * px + py * py + pz * pz);
  const double floor = 1.0e-6 * fmax(fabs(p.uMid), 1.0);
  if (count < 5 && zeroLimit && !(p.uTau > 0.0)) out.beta = p.beta;
  else out.beta = speed > floor ? p.uTau * p.uTau / speed : p.beta;
  if (!isfinite(out.beta) || !(out.beta > 0.0)) return out;
  out.uTau = p.uTau;
  out.yPlus = p.yPlus;
  out.valid = isfinite(p.fitRms) && isfinite(p.yPlus);
  return out;
}

// ---------------------------------------------------------------- omega wall target
struct OmegaWall {
  double omegaTarget, omegaVis, omegaLog, lamFrac, rey, yPlusK, pkAuto, nutSpalding;
};

// OpenFOAM-style automatic omega wall value: viscous 6 nu / (beta1 y^2) and
// log-layer u*/(sqrt(Cmu) kappa y) blended by exp(-Re_y / reBlend), with the
// matching production blend and limiter; in shared mode u* is the Spalding
// friction velocity and kappa, E those of the momentum wall law.
NODALS_SST_FN OmegaWall omegaWall(double yd, double nu, double Ut, double uTau, double kRaw, double omRaw,
                                               double pk, double kFloor, double omFloor, double kappa, double E,
                                               double beta1, double reBlend, double prodLimit, bool shared) {
  OmegaWall q{};
  const double kval = fmax(kRaw, kFloor), om = fmax(omRaw, omFloor);
  const double Cmu25 = sqrt(sqrt(0.09)), Cmu5 = sqrt(0.09);
  q.rey = yd * sqrt(kval) / nu;
  q.yPlusK = Cmu25 * q.rey;
  q.lamFrac = exp(-q.rey / fmax(reBlend, 1e-6));
  q.lamFrac = fmin(fmax(q.lamFrac, 0.0), 1.0);
  const double turbFrac = 1.0 - q.lamFrac;
  const double magGrad = Ut / fmax(yd, 1e-20);
  const double kap = shared ? kKappa : fmax(kappa, 1e-6);
  const double Espg = shared ? kE : E;
  const double uStar = shared ? fmax(uTau, 0.0) : sqrt(fmax(0.0, q.lamFrac * nu * magGrad + turbFrac * Cmu5 * kval));
  q.omegaVis = 6.0 * nu / (fmax(beta1, 1e-8) * yd * yd);
  q.omegaLog = uStar / (Cmu5 * kap * yd);
  q.omegaTarget = q.lamFrac * q.omegaVis + turbFrac * q.omegaLog;
  double gLog = pk;
  if (shared) {
    if (uTau > 1e-12 && yd > 1e-20) gLog = fmax(0.0, uTau * uTau * uTau / (kap * yd));
  } else {
    const double arg = fmax(Espg * q.yPlusK, 1.000001);
    const double uPlus = log(arg) / kap;
    if (q.yPlusK > 1e-8 && uPlus > 1e-8) {
      const double t = uStar * magGrad * yd / uPlus;
      gLog = t * t / (nu * kap * q.yPlusK);
    }
  }
  q.pkAuto = q.lamFrac * pk + turbFrac * gLog;
  const double lim = fmax(prodLimit, 0.0) * 0.09 * kval * om;
  q.pkAuto = fmin(fmax(q.pkAuto, 0.0), lim);
  q.nutSpalding = Ut > 1e-12 ? fmax(0.0, uTau * uTau / (Ut / yd) - nu) : 0.0;
  return q;
}

}  // namespace sst
}  // namespace nodals

#undef NODALS_SST_FN
