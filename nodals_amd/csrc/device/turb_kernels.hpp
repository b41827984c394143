// Kernels of the k-omega SST transport step (masked-P1 state space on the
// P1+BF3 tetrahedral space; scalar DOFs are mesh entities, vertices first,
// with every face-bubble coefficient zero).
//
// Per equation: element blocks (volume terms + Picard convection, one
// wavefront per cell: k_sst_volume), inlet flux blocks (k_turb_inlet), the
// omega wall condition (k_turb_wall_coef, then k_turb_wall_block or
// k_turb_wall_trace), then one
// gather per vertex row that sums its entries over the blocks, forms the
// row-L1 relaxation and writes the sweep operator, diagonal and right-hand
// side (k_turb_gather); the coloured sweeps are k_mcgs_color; k_turb_floor
// applies the floor.  No atomics in the assembly: every entry is summed by
// exactly one lane in a fixed order.
//
// Reference: turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
// modules/transport/tet_inlet_flux.inc:1-23 (sst_inlet_tet),
// modules/wall/of_auto_point3_kernel.inc:19-76 (sst_tet_of_auto_spalding5_point3
// with spalding_profile5_tet_fit.inc:40-95, of_auto_point3_target.inc,
// of_auto_point3_rank1_penalty.inc and of_auto_trace_log_nitsche.inc),
// hxt5c_sst_g1.cuh:438 (combine_scalar), modules/transport/
// scalar_relaxation.inc:8-33 (scalar_row_l1_relax_kernel) with row_l1_plan
// (modules/transport/portable/transport_core.hpp:26-31), :176
// (floor_vertex_kernel).
#pragma once

#include <hip/hip_runtime.h>

#include "la_kernels.hpp"
#include "sst_physics.hpp"
#include "sst_volume.hpp"

namespace nodals {
namespace dev {

// Local vertices of face lf (opposite vertex lf), reference order.
static __constant__ int kTurbFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

struct TurbConsts {
  double radius, nu, kFloor, omFloor;
  int deviatoric, equation;
};

// Twelve-point triangle rule rows (l on the face's three vertices, weight).
struct Tri12 {
  double q[12][4];
};

// Velocity gradient, velocity, k, omega and their gradients at barycentric
// point l of cell c; phi/gp receive the basis values and physical gradients.
__device__ inline void turbStateAt(int c, const double l[4], const int* __restrict__ cref,
                                   const int* __restrict__ sref, const double* __restrict__ cinvJ,
                                   const double* __restrict__ u, const double* __restrict__ ufix,
                                   const double* __restrict__ kf, const double* __restrict__ wf, double phi[8],
                                   double gp[8][3], double gu[9], double us[3], double& kv, double& wv, double gk[3],
                                   double gw[3]) {
  double gr[8][3];
  p1bf3Basis(l, phi, gr);
  const double* I = cinvJ + (long long)c * 9;
  for (int i = 0; i < 9; ++i) gu[i] = 0.0;
  for (int d = 0; d < 3; ++d) us[d] = gk[d] = gw[d] = 0.0;
  kv = wv = 0.0;
  for (int a = 0; a < 8; ++a) {
    const int r = cref[(long long)c * 8 + a];
    const double* U = r >= 0 ? u + 3ll * r : ufix + 3ll * (-r - 1);
    const int g = sref[(long long)c * 8 + a];
    const double Ka = kf[g], Wa = wf[g], pa = phi[a];
    kv += Ka * pa;
    wv += Wa * pa;
    for (int d = 0; d < 3; ++d) {
      us[d] += U[d] * pa;
      const double gd = gr[a][0] * I[d] + gr[a][1] * I[3 + d] + gr[a][2] * I[6 + d];
      gp[a][d] = gd;
      gk[d] += Ka * gd;
      gw[d] += Wa * gd;
      gu[d] += U[0] * gd;
      gu[3 + d] += U[1] * gd;
      gu[6 + d] += U[2] * gd;
    }
  }
}

// Wall distance at barycentric point l of cell c: the P1 interpolation of
// the vertex distance field when one is installed (floored at zero, the
// reference's stage20_distance, modules/geometry/wall_distance_device.cuh:
// 4-10), else the analytic distance of the straight pipe about its axis.
__device__ inline double pointDistance(int c, const double l[4], const int* __restrict__ sref,
                                       const double* __restrict__ vdist, const double* __restrict__ cxy,
                                       double radius) {
  if (vdist) {
    double d = 0.0;
    for (int a = 0; a < 4; ++a) d += l[a] * vdist[sref[(long long)c * 8 + a]];
    return sst::floorDistance(fmax(0.0, d), radius);
  }
  double x = 0.0, y = 0.0;
  for (int a = 0; a < 4; ++a) {
    x += l[a] * cxy[(long long)c * 8 + 2 * a];
    y += l[a] * cxy[(long long)c * 8 + 2 * a + 1];
  }
  return sst::pipeDistance(x, y, radius);
}

// Inlet flux of one face (one wavefront per face, four per block):
//   A_ab += sum_q w_q [ qin phi_a phi_b + nu_e (-phi_a dn phi_b + dn phi_a phi_b) ]
//   b_a  += sum_q w_q [ qin phi_a phiIn + nu_e dn phi_a phiIn ]
// with qin = max(0, -U_inlet . n), nu_e = nu + sigma nu_t at the point and
// phiIn the inlet value of the transported scalar.
static __global__ void __launch_bounds__(256) k_turb_inlet(
    int nf, const int* __restrict__ fcell, const int* __restrict__ flf, const double* __restrict__ fgeom,
    const int* __restrict__ cref, const int* __restrict__ sref, const double* __restrict__ cinvJ,
    const double* __restrict__ cxy, const double* __restrict__ u, const double* __restrict__ ufix,
    const double* __restrict__ kf, const double* __restrict__ wf, TurbConsts K, Tri12 Q, double inX, double inY,
    double inZ, double phiIn, double* __restrict__ blk, double* __restrict__ brhs, const double* __restrict__ vdist) {
  __shared__ double PH[4][12][8], DN[4][12][8], NUE[4][12], WQ[4][12];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int f = blockIdx.x * 4 + wv;
  const bool live = f < nf;
  const double* G = fgeom + 4ll * (live ? f : 0);
  if (live && lane < 12) {
    const int c = fcell[f], lf = flf[f];
    double l[4] = {0, 0, 0, 0};
    for (int j = 0; j < 3; ++j) l[kTurbFace[lf][j]] = Q.q[lane][j];
    double phi[8], gp[8][3], gu[9], us[3], kv, wvv, gk[3], gw[3];
    turbStateAt(c, l, cref, sref, cinvJ, u, ufix, kf, wf, phi, gp, gu, us, kv, wvv, gk, gw);
    const sst::Local s = sst::local(pointDistance(c, l, sref, vdist, cxy, K.radius), K.nu, gu, kv, wvv, gk, gw,
                                    K.kFloor, K.omFloor, K.deviatoric != 0);
    NUE[wv][lane] = K.nu + (K.equation == 0 ? s.sigmaK : s.sigmaW) * s.nut;
    WQ[wv][lane] = Q.q[lane][3] * G[3];
    for (int a = 0; a < 8; ++a) {
      PH[wv][lane][a] = phi[a];
      DN[wv][lane][a] = gp[a][0] * G[0] + gp[a][1] * G[1] + gp[a][2] * G[2];
    }
  }
  __syncthreads();
  if (!live) return;
  const double inflow = fmax(0.0, -(inX * G[0] + inY * G[1] + inZ * G[2]));
  {
    const int a = lane >> 3, b = lane & 7;
    double z = 0.0;
    for (int q = 0; q < 12; ++q)
      z += WQ[wv][q] * (inflow * PH[wv][q][a] * PH[wv][q][b] +
                        NUE[wv][q] * (-PH[wv][q][a] * DN[wv][q][b] + DN[wv][q][a] * PH[wv][q][b]));
    blk[(long long)f * 64 + lane] = z;
  }
  if (lane < 8) {
    double z = 0.0;
    for (int q = 0; q < 12; ++q)
      z += WQ[wv][q] * (inflow * PH[wv][q][lane] * phiIn + NUE[wv][q] * DN[wv][q][lane] * phiIn);
    brhs[(long long)f * 8 + lane] = z;
  }
}

struct WallConsts {
  double gamma, kappa, E, beta1, reBlend, prodLimit;
  double ell;          // y5 - y1 of the nominal sample heights
  int fitPoints, mode;  // active Spalding points; 0 of_auto, 1 sample_log, 2 trace_log
};

// Omega wall coefficient and target of one wall quadrature point (one
// thread each): Spalding fit of u_tau over the active samples from the
// magnitude of the velocity's component tangential to the wall face (the
// reference's vector form; the speed is invariant under a rigid rotation of
// velocity and normal), the SST state at the third sample, the automatic
// omega target with the shared Spalding u_tau (its blended value for
// of_auto, its log-layer value for sample_log and trace_log), the rank-one
// coefficient coeff = gamma nu_e / ell * (w_q area ell) / ell and nu_e =
// nu + sigma_w nu_t at the third sample (mu, the trace closure's
// viscosity).  Invalid points get coeff 0 and are counted in *bad.
static __global__ void __launch_bounds__(256) k_turb_wall_coef(
    int nq, const double* __restrict__ wgeom, const int* __restrict__ scell, const double* __restrict__ sl,
    const double* __restrict__ sphi, const double* __restrict__ sy, const int* __restrict__ cref,
    const int* __restrict__ sref, const double* __restrict__ cinvJ, const double* __restrict__ cxy,
    const double* __restrict__ u, const double* __restrict__ ufix, const double* __restrict__ kf,
    const double* __restrict__ wf, TurbConsts K, WallConsts W, Tri12 Q, double* __restrict__ coef,
    double* __restrict__ target, int* __restrict__ bad, const double* __restrict__ vdist,
    double* __restrict__ mu) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nq) return;
  const double* G = wgeom + 4ll * (w / 12);
  const unsigned mask = sst::activeMask(W.fitPoints);
  double U[5] = {0, 0, 0, 0, 0}, Y[5];
  for (int m = 0; m < 5; ++m) Y[m] = sy[(long long)w * 5 + m];
  bool ok = fabs(G[0] * G[0] + G[1] * G[1] + G[2] * G[2] - 1.0) < 1.0e-8;
  for (int m = 0; m < 5; ++m) {
    if (!(mask & (1u << m))) continue;
    const long long s = (long long)w * 5 + m;
    const int c = scell[s];
    double uv[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 8; ++a) {
      const int r = cref[(long long)c * 8 + a];
      const double* Ua = r >= 0 ? u + 3ll * r : ufix + 3ll * (-r - 1);
      const double ph = sphi[s * 8 + a];
      uv[0] += Ua[0] * ph;
      uv[1] += Ua[1] * ph;
      uv[2] += Ua[2] * ph;
    }
    const double un = uv[0] * G[0] + uv[1] * G[1] + uv[2] * G[2];
    const double t0 = uv[0] - un * G[0], t1 = uv[1] - un * G[1], t2 = uv[2] - un * G[2];
    const double mag = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    ok = ok && isfinite(uv[0]) && isfinite(uv[1]) && isfinite(uv[2]) && isfinite(mag);
    U[m] = mag;
  }
  const sst::Profile P = sst::fitActive(U, U, Y, W.fitPoints, K.nu);
  const long long s3 = (long long)w * 5 + 2;
  const int c = scell[s3];
  const double l[4] = {sl[s3 * 4], sl[s3 * 4 + 1], sl[s3 * 4 + 2], sl[s3 * 4 + 3]};
  double phi[8], gp[8][3], gu[9], us[3], kv, wvv, gk[3], gw[3];
  turbStateAt(c, l, cref, sref, cinvJ, u, ufix, kf, wf, phi, gp, gu, us, kv, wvv, gk, gw);
  const sst::Local ss = sst::local(pointDistance(c, l, sref, vdist, cxy, K.radius), K.nu, gu, kv, wvv, gk, gw,
                                   K.kFloor, K.omFloor, K.deviatoric != 0);
  const double normal = us[0] * G[0] + us[1] * G[1] + us[2] * G[2];
  const double Ut = sqrt(fmax(0.0, us[0] * us[0] + us[1] * us[1] + us[2] * us[2] - normal * normal));
  const sst::OmegaWall oa = sst::omegaWall(Y[2], K.nu, Ut, P.uTau, kv, wvv, ss.pk, K.kFloor, K.omFloor, W.kappa,
                                           W.E, W.beta1, W.reBlend, W.prodLimit, true);
  const double ell = W.ell;
  const double nuEff = K.nu + ss.sigmaW * ss.nut;
  const double vweight = Q.q[w % 12][3] * G[3] * ell;
  const double cf = W.gamma * nuEff / ell * vweight / ell;
  const double tgt = W.mode == 0 ? oa.omegaTarget : oa.omegaLog;
  ok = ok && P.valid && Y[2] > 0.0 && ell > 0.0 && isfinite(nuEff) && nuEff > 0.0 && isfinite(cf) && cf > 0.0 &&
       isfinite(tgt) && tgt > 0.0;
  coef[w] = ok ? cf : 0.0;
  target[w] = ok ? tgt : 0.0;
  mu[w] = ok ? nuEff : 0.0;
  if (!ok) atomicAdd(bad, 1);
}

// trace_log: the omega target g imposed on the wall face by symmetric
// Nitsche, one wavefront per wall quadrature point, on the face's own cell:
//   A_ab = w [ -mu dn phi_b phi_a - mu dn phi_a phi_b + tau phi_a phi_b ]
//   b_a  = w [ -mu dn phi_a + tau phi_a ] g
// at the face point, w = w_q area, tau = gammaT mu / h, h = det / (2 area)
// (three times the volume over the face area), n the outward normal.
// Points whose rank-one checks failed (coef 0) were counted already;
// a failed trace check is counted here.  Both give a zero block.
static __global__ void __launch_bounds__(256) k_turb_wall_trace(
    int nq, const int* __restrict__ wcell, const int* __restrict__ wlf, const double* __restrict__ wgeom,
    const double* __restrict__ cinvJ, const double* __restrict__ cdet, const double* __restrict__ coef,
    const double* __restrict__ target, const double* __restrict__ mu, double gammaT, Tri12 Q,
    double* __restrict__ blk, double* __restrict__ brhs, int* __restrict__ bad) {
  __shared__ double PH[4][8], DN[4][8];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + wv;
  const bool live = w < nq;
  const int f = live ? w / 12 : 0, q = live ? w % 12 : 0;
  const int c = wcell[f];
  const double* G = wgeom + 4ll * f;
  if (live && lane == 0) {
    double l[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < 3; ++j) l[kTurbFace[wlf[f]][j]] = Q.q[q][j];
    double phi[8], gr[8][3];
    p1bf3Basis(l, phi, gr);
    const double* I = cinvJ + (long long)c * 9;
    for (int a = 0; a < 8; ++a) {
      double dn = 0.0;
      for (int d = 0; d < 3; ++d) dn += (gr[a][0] * I[d] + gr[a][1] * I[3 + d] + gr[a][2] * I[6 + d]) * G[d];
      PH[wv][a] = phi[a];
      DN[wv][a] = dn;
    }
  }
  __syncthreads();
  if (!live) return;
  const double h = cdet[c] / (2.0 * G[3]);
  const double m = mu[w], g = target[w];
  const double tau = gammaT * m / h;
  const double ww = Q.q[q][3] * G[3];
  const bool okT = isfinite(h) && h > 0.0 && isfinite(m) && m > 0.0 && isfinite(tau) && tau > 0.0 && isfinite(ww) &&
                   ww > 0.0 && isfinite(g) && g > 0.0;
  const bool ok = coef[w] > 0.0 && okT;
  if (coef[w] > 0.0 && !okT && lane == 0) atomicAdd(bad, 1);
  const int a = lane >> 3, b = lane & 7;
  blk[(long long)w * 64 + lane] =
      ok ? ww * (-m * DN[wv][b] * PH[wv][a] - m * DN[wv][a] * PH[wv][b] + tau * PH[wv][a] * PH[wv][b]) : 0.0;
  if (lane < 8) brhs[(long long)w * 8 + lane] = ok ? ww * (-m * DN[wv][lane] + tau * PH[wv][lane]) * g : 0.0;
}

// Rank-one penalty block of each wall quadrature point on the cell of its
// third sample: A_ab = coeff phi_a phi_b, b_a = coeff phi_a target (zeros for
// the k equation, coef == nullptr).  One wavefront per point.
static __global__ void __launch_bounds__(256) k_turb_wall_block(int nq, const double* __restrict__ sphi,
                                                         const double* __restrict__ coef,
                                                         const double* __restrict__ target,
                                                         double* __restrict__ blk, double* __restrict__ brhs) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= nq) return;
  const double* ph = sphi + ((long long)w * 5 + 2) * 8;
  const double cf = coef ? coef[w] : 0.0;
  blk[(long long)w * 64 + lane] = cf * ph[lane >> 3] * ph[lane & 7];
  if (lane < 8) brhs[(long long)w * 8 + lane] = coef ? cf * ph[lane] * target[w] : 0.0;
}

// One group of L lanes per vertex row: every entry of the row (its vertex
// columns, then its bubble columns) summed over its block terms below
// termLimit; vertex entries go to val; l1 = sum |entry| over all entries;
// the relaxation delta = (1/alpha - 1) l1 raises the diagonal (val and
// diag) and adds delta x_i to the right-hand side.  A non-positive or
// non-finite relaxed row is counted in *bad.
template <int L>
static __global__ void __launch_bounds__(256) k_turb_gather(
    int nvert, const int* __restrict__ row, const int* __restrict__ diagPos, const int* __restrict__ entOff,
    const int* __restrict__ termOff, const int* __restrict__ term, const int* __restrict__ rhsOff,
    const int* __restrict__ rhsTerm, int termLimit, int rhsLimit, const double* __restrict__ blk,
    const double* __restrict__ brhs, const double* __restrict__ x, double alpha, double* __restrict__ val,
    double* __restrict__ diag, double* __restrict__ rhs, int* __restrict__ bad) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int i = gid / L, lane = gid % L;
  if (i >= nvert) return;
  const int r0 = row[i], nvc = row[i + 1] - r0, e0 = entOff[i], e1 = entOff[i + 1];
  const int dloc = diagPos[i] - r0;
  double l1 = 0.0, d0 = 0.0, rs = 0.0;
  bool finite = true;
  for (int e = e0 + lane; e < e1; e += L) {
    double v = 0.0;
    for (int t = termOff[e]; t < termOff[e + 1]; ++t) {
      const int tm = term[t];
      if (tm >= termLimit) break;
      v += blk[tm];
    }
    if (e - e0 < nvc) val[r0 + (e - e0)] = v;
    if (e - e0 == dloc) d0 = v;
    finite = finite && isfinite(v);
    l1 += fabs(v);
  }
  for (int t = rhsOff[i] + lane; t < rhsOff[i + 1]; t += L) {
    const int tm = rhsTerm[t];
    if (tm < rhsLimit) rs += brhs[tm];
  }
  l1 = groupSum(l1, L);
  d0 = groupSum(d0, L);
  rs = groupSum(rs, L);
  const int nonFinite = groupSum(finite ? 0 : 1, L);
  if (lane == 0) {
    const double delta = (1.0 / alpha - 1.0) * l1, dn = d0 + delta;
    if (nonFinite || !(l1 > 0.0) || !isfinite(delta) || !(dn > 0.0) || !isfinite(dn)) {
      atomicAdd(bad, 1);
      diag[i] = 1.0;
      rhs[i] = x[i];
      return;
    }
    val[r0 + dloc] = dn;
    diag[i] = dn;
    rhs[i] = rs + delta * x[i];
  }
}

static __global__ void k_turb_floor(int n, double floorv, double* __restrict__ x, int* __restrict__ clips) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (!(x[i] >= floorv) || !isfinite(x[i]))) {
    x[i] = floorv;
    atomicAdd(clips, 1);
  }
}

// Residual r = b - (offdiag(A) x + d x) of a scalar system in the form the
// multicolour sweep solves (the stored diagonal entry replaced by d), one
// thread per row.
static __global__ void k_scalar_resid(int n, const int* __restrict__ row, const int* __restrict__ col,
                                      const double* __restrict__ val, const double* __restrict__ d,
                                      const double* __restrict__ b, const double* __restrict__ x,
                                      double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = b[i] - d[i] * x[i];
  for (int k = row[i]; k < row[i + 1]; ++k)
    if (col[k] != i) s -= val[k] * x[col[k]];
  r[i] = s;
}

}  // namespace dev
}  // namespace nodals
