// Wall-function momentum closure of the k-omega SST path on gfx950: the
// face-averaged Spalding fit, the Nitsche/Spalding wall blocks, the
// component momentum rows with their row-L1 relaxation, the 3x3 block (M3)
// operator K = T^T M T with the exact BF3 wall-flux constraint T, its block
// Gauss-Seidel sweep and the constrained pressure Schur operator.
//
// Reference (turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/):
// modules/wall/stage22d4b_nitsche_live.inc:25-96 (face fit), :99-335
// (Nitsche face blocks, BF3 flux mode), stage22d4d_coupled_matrixfree.inc:
// 68-105 (frozen cross-component face coupling), :734-813 (M3 transforms),
// :843-889 (block colour sweep), modules/wall/stage22d4i_constraint.inc
// (T, Q, T^T), modules/wall/segregated_vector_tet_22c.inc:72-104
// (component combination and row-L1 relaxation),
// modules/coupling/stage22d4i3a_schur.cuh (gamma-free rAU, constrained Schur
// and velocity correction), modules/coupling/implicit_dev_volume_tet.inc
// (deviatoric 3x3 volume tensor).
//
// Differences in form, not in value: every reduction runs in a fixed order
// (face blocks sum their twelve points in order, transforms gather their
// sources through setup lists), so results do not depend on atomics.
#pragma once

#include "la_kernels.hpp"
#include "sst_physics.hpp"
#include "sst_volume.hpp"

namespace nodals {
namespace dev {

// Face-averaged Spalding fit, one thread per wall face: the face-quadrature
// means of the Picard functional, of the velocity and of the wall distance
// (sdist, per sample) at each active wall-normal sample slot, then one
// tangential fit.  fit[f] = (beta, u_tau,
// y+); beta = -1 and a raised counter for a face without a valid fit.
static __global__ void k_wall_fit(int nf, const double* __restrict__ wn, const double* __restrict__ qw,
                                  const double* __restrict__ bp, const int* __restrict__ wcell,
                                  const int* __restrict__ cref, const double* __restrict__ u,
                                  const double* __restrict__ ufix, const int* __restrict__ scell,
                                  const double* __restrict__ sphi, const double* __restrict__ heights,
                                  const double* __restrict__ sdist, double nu,
                                  int activePoints, int zeroLimit, double* __restrict__ fit, int* __restrict__ bad) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const unsigned mask = sst::activeMask(activePoints);
  double pic[3] = {0, 0, 0}, sx[5] = {0, 0, 0, 0, 0}, sy[5] = {0, 0, 0, 0, 0}, sz[5] = {0, 0, 0, 0, 0};
  double yy[5] = {0, 0, 0, 0, 0}, area = 0.0;
  bool ok = true;
  const int c = wcell[f];
  for (int q = 0; q < 12 && ok; ++q) {
    const double w = qw[f * 12 + q];
    if (!(w > 0.0) || !isfinite(w)) {
      ok = false;
      break;
    }
    area += w;
    for (int b = 0; b < 8; ++b) {
      const int r = cref[c * 8 + b];
      const double* U = r >= 0 ? u + 3ll * r : ufix + 3ll * (-r - 1);
      const double z = w * bp[((long long)f * 12 + q) * 8 + b];
      pic[0] += z * U[0];
      pic[1] += z * U[1];
      pic[2] += z * U[2];
    }
    for (int m = 0; m < 5; ++m) {
      if (!(mask & (1u << m))) continue;
      const long long si = ((long long)f * 12 + q) * 5 + m;
      yy[m] += w * sdist[si];
      const int sc = scell[si];
      for (int b = 0; b < 8; ++b) {
        const int r = cref[sc * 8 + b];
        const double* U = r >= 0 ? u + 3ll * r : ufix + 3ll * (-r - 1);
        const double z = w * sphi[si * 8 + b];
        sx[m] += z * U[0];
        sy[m] += z * U[1];
        sz[m] += z * U[2];
      }
    }
  }
  sst::VecFit F{0, 0, 0, 0};
  if (ok && area > 0.0 && isfinite(area)) {
    const double inv = 1.0 / area;
    for (int d = 0; d < 3; ++d) pic[d] *= inv;
    for (int m = 0; m < 5; ++m) {
      if (mask & (1u << m)) {
        sx[m] *= inv;
        sy[m] *= inv;
        sz[m] *= inv;
        yy[m] *= inv;
      } else {
        yy[m] = heights[m];
      }
    }
    const double n[3] = {wn[3 * f], wn[3 * f + 1], wn[3 * f + 2]};
    F = sst::fitTangential(sx, sy, sz, yy, n, pic, nu, activePoints, zeroLimit != 0);
  }
  if (!F.valid || !(F.beta > 0.0) || !isfinite(F.beta) || !(F.uTau >= 0.0) || !isfinite(F.yPlus)) {
    atomicAdd(bad, 1);
    fit[3 * f] = -1.0;
    fit[3 * f + 1] = 0.0;
    fit[3 * f + 2] = 0.0;
    return;
  }
  fit[3 * f] = F.beta;
  fit[3 * f + 1] = F.uTau;
  fit[3 * f + 2] = F.yPlus;
}

// Nitsche/Spalding wall blocks, one wavefront per wall face.  Lanes 0..11
// take the effective viscosity nu + nu_t of the SST local state at their
// quadrature point (the same provider as the volume terms); lane 8a+b then
// sums, in point order, the BF3-flux-mode entries
//   normalQ = w nu_eff gamma/h pa pb,  robin = w beta pa pb
// (pa, pb: wall traces) into
//   blk[f][ab][i] = n_i^2 normalQ + (1 - n_i^2) robin   (component i block)
//   pen[f][ab][i] = n_i^2 normalQ                      (Nitsche penalty part)
//   cpl[f][ab]    = normalQ - robin                    (cross coupling:
//                   component (i, j != i) gets n_i n_j cpl)
static __global__ void __launch_bounds__(256) k_wall_nitsche(
    int nf, const int* __restrict__ wcell, const double* __restrict__ wn, const double* __restrict__ wh,
    const double* __restrict__ qw, const double* __restrict__ ql, const double* __restrict__ bw,
    const double* __restrict__ fit, const int* __restrict__ cref, const double* __restrict__ cinvJ,
    const double* __restrict__ u, const double* __restrict__ ufix, const int* __restrict__ sref,
    const double* __restrict__ kf, const double* __restrict__ wf, const double* __restrict__ vdist, double radius,
    double nu, double kFloor, double omFloor, int deviatoric, double gammaN, double* __restrict__ blk,
    double* __restrict__ pen, double* __restrict__ cpl, double* __restrict__ nuq, int* __restrict__ bad) {
  __shared__ double NUE[4][12];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int f = blockIdx.x * 4 + wv;
  const bool live = f < nf;
  if (live && lane < 12) {
    const int c = wcell[f], q = lane;
    const double* L = ql + ((long long)f * 12 + q) * 4;
    const double l[4] = {L[0], L[1], L[2], L[3]};
    double phi[8], gr[8][3];
    p1bf3Basis(l, phi, gr);
    const double* I = cinvJ + (long long)c * 9;
    double gu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gk[3] = {0, 0, 0}, gw[3] = {0, 0, 0}, kv = 0, wvv = 0;
    for (int a = 0; a < 8; ++a) {
      const int r = cref[c * 8 + a];
      const double* U = r >= 0 ? u + 3ll * r : ufix + 3ll * (-r - 1);
      const int g = sref[c * 8 + a];
      const double Ka = kf[g], Wa = wf[g], pa = phi[a];
      kv += Ka * pa;
      wvv += Wa * pa;
      for (int d = 0; d < 3; ++d) {
        const double gd = gr[a][0] * I[d] + gr[a][1] * I[3 + d] + gr[a][2] * I[6 + d];
        gk[d] += Ka * gd;
        gw[d] += Wa * gd;
        gu[d] += U[0] * gd;
        gu[3 + d] += U[1] * gd;
        gu[6 + d] += U[2] * gd;
      }
    }
    double dd = 0;
    for (int a = 0; a < 4; ++a) dd += l[a] * vdist[sref[c * 8 + a]];
    const double dq = sst::floorDistance(fmax(0.0, dd), radius);
    const sst::Local s = sst::local(dq, nu, gu, kv, wvv, gk, gw, kFloor, omFloor, deviatoric != 0);
    const double ne = nu + s.nut;
    if (!(ne > 0.0) || !isfinite(ne)) atomicAdd(bad, 1);
    NUE[wv][q] = ne;
    nuq[f * 12 + q] = ne;
  }
  __syncthreads();
  if (!live) return;
  const int a = lane >> 3, b = lane & 7;
  const double beta = fit[3 * f];
  const double n0 = wn[3 * f], n1 = wn[3 * f + 1], n2 = wn[3 * f + 2];
  const double gh = gammaN / wh[f];
  double e0 = 0, e1 = 0, e2 = 0, p0 = 0, p1 = 0, p2 = 0, cc = 0;
  for (int q = 0; q < 12; ++q) {
    const double w = qw[f * 12 + q];
    const double pa = bw[((long long)f * 12 + q) * 8 + a], pb = bw[((long long)f * 12 + q) * 8 + b];
    const double normalQ = w * NUE[wv][q] * (gh * pa * pb);
    const double robin = w * beta * (pa * pb);
    e0 += n0 * n0 * normalQ + (1.0 - n0 * n0) * robin;
    e1 += n1 * n1 * normalQ + (1.0 - n1 * n1) * robin;
    e2 += n2 * n2 * normalQ + (1.0 - n2 * n2) * robin;
    p0 += n0 * n0 * w * NUE[wv][q] * gh * pa * pb;
    p1 += n1 * n1 * w * NUE[wv][q] * gh * pa * pb;
    p2 += n2 * n2 * w * NUE[wv][q] * gh * pa * pb;
    cc += normalQ - robin;
  }
  const long long o = (long long)f * 64 + lane;
  blk[3 * o] = e0;
  blk[3 * o + 1] = e1;
  blk[3 * o + 2] = e2;
  pen[3 * o] = p0;
  pen[3 * o + 1] = p1;
  pen[3 * o + 2] = p2;
  cpl[o] = cc;
}

// Component momentum rows of the wall-function path, L lanes per row.  Per
// entry: the shared value s = diffusion + convection (+ eddy diffusion)
// terms, the component values s + wall block_i, the penalty parts, the
// cross-component wall coupling n_i n_j cpl and the deviatoric 3x3 tensor.
// Per row and component (segregated_vector_tet_22c.inc:81-104): row-L1 L,
// delta = (1/alpha - 1) L, relaxed diagonal d + delta, rAU = s_r/(d +
// delta); and the gamma-free rAU of the constrained Schur
// (stage22d4i3a_schur.cuh:11-41), s_r/(d - pen_d + (1/alpha - 1) sum|v -
// pen|).  K[k][3i+j] = relaxed component value (i = j) + cross + deviatoric.
// Right-hand side: static Dirichlet terms, fixed-column convection and
// deviatoric terms, B^T p and delta u_old per component.
struct WallRowsArgs {
  const int *mrow, *mdiagPos;
  const double* mdiff;
  const int *termOff, *termIdx;
  const double* cbuf;
  const double* dbuf;  // [nc*64][9] deviatoric tensors, or null
  const int *blkOff, *blkIdx;
  const double *wblk, *wpen, *wcpl, *wn;
  const int *fixOff, *fixTerm;
  const double* fixedVal;
  const int* incOff;
  const unsigned* incPacked;
  const double *bcoef, *p, *srhs, *u;
  double fac, rauScale;
  double *K, *brhs, *rau3, *r3, *delta3, *diag3;
  int* bad;
};

template <int L>
__device__ __forceinline__ void wall_row(const WallRowsArgs& A, int i, int lane) {
  const int rb = A.mrow[i], re = A.mrow[i + 1], dp = A.mdiagPos[i];
  double l1[3] = {0, 0, 0}, l1p[3] = {0, 0, 0}, dg[3] = {0, 0, 0}, dpn[3] = {0, 0, 0}, ddev[3] = {0, 0, 0};
  for (int k = rb + lane; k < re; k += L) {
    double v = A.mdiff[k];
    double dev[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = A.termOff[k]; t < A.termOff[k + 1]; ++t) {
      const int q = A.termIdx[t];
      v += A.cbuf[q];
      if (A.dbuf)
        for (int j = 0; j < 9; ++j) dev[j] += A.dbuf[9ll * q + j];
    }
    double w3[3] = {0, 0, 0}, p3[3] = {0, 0, 0}, x01 = 0, x02 = 0, x12 = 0;
    for (int t = A.blkOff[k]; t < A.blkOff[k + 1]; ++t) {
      const int o = A.blkIdx[t], fc = o >> 6;
      for (int d = 0; d < 3; ++d) {
        w3[d] += A.wblk[3ll * o + d];
        p3[d] += A.wpen[3ll * o + d];
      }
      const double c = A.wcpl[o];
      const double n0 = A.wn[3 * fc], n1 = A.wn[3 * fc + 1], n2 = A.wn[3 * fc + 2];
      x01 += n0 * n1 * c;
      x02 += n0 * n2 * c;
      x12 += n1 * n2 * c;
    }
    double* Kk = A.K + 9ll * k;
    double vc[3];
    for (int d = 0; d < 3; ++d) {
      vc[d] = v + w3[d];
      l1[d] += fabs(vc[d]);
      l1p[d] += fabs(vc[d] - p3[d]);
    }
    Kk[0] = vc[0] + dev[0];
    Kk[1] = x01 + dev[1];
    Kk[2] = x02 + dev[2];
    Kk[3] = x01 + dev[3];
    Kk[4] = vc[1] + dev[4];
    Kk[5] = x12 + dev[5];
    Kk[6] = x02 + dev[6];
    Kk[7] = x12 + dev[7];
    Kk[8] = vc[2] + dev[8];
    if (k == dp)
      for (int d = 0; d < 3; ++d) {
        dg[d] = vc[d];
        dpn[d] = p3[d];
        ddev[d] = dev[4 * d];
      }
  }
  for (int d = 0; d < 3; ++d) {
    l1[d] = groupSum(l1[d], L);
    l1p[d] = groupSum(l1p[d], L);
    dg[d] = groupSum(dg[d], L);
    dpn[d] = groupSum(dpn[d], L);
    ddev[d] = groupSum(ddev[d], L);
  }
  // Dirichlet columns: convection (+ eddy diffusion) and deviatoric terms
  double c3[3] = {0, 0, 0};
  for (int t = A.fixOff[i] + lane; t < A.fixOff[i + 1]; t += L) {
    const int q = A.fixTerm[2 * t], slot = A.fixTerm[2 * t + 1];
    const double C = A.cbuf[q];
    const double* uf = A.fixedVal + 3ll * slot;
    for (int d = 0; d < 3; ++d) c3[d] -= C * uf[d];
    if (A.dbuf)
      for (int d = 0; d < 3; ++d)
        for (int j = 0; j < 3; ++j) c3[d] -= A.dbuf[9ll * q + 3 * d + j] * uf[j];
  }
  double t3[3] = {0, 0, 0};
  for (int k = A.incOff[i] + lane; k < A.incOff[i + 1]; k += L) {
    const unsigned pk = A.incPacked[k];
    const int c = (int)(pk >> 3), a = (int)(pk & 7u);
    const double pv = A.p[c];
    const double* B = A.bcoef + (long long)c * 24 + 3 * a;
    for (int d = 0; d < 3; ++d) t3[d] += B[d] * pv;
  }
  for (int d = 0; d < 3; ++d) {
    c3[d] = groupSum(c3[d], L);
    t3[d] = groupSum(t3[d], L);
  }
  // the lane that wrote the diagonal entry's block relaxes it
  if (dp >= rb + lane && ((dp - rb - lane) % L) == 0)
    for (int d = 0; d < 3; ++d) A.K[9ll * dp + 4 * d] = dg[d] + A.fac * l1[d] + ddev[d];
  if (lane == 0) {
    for (int d = 0; d < 3; ++d) {
      const double de = A.fac * l1[d];
      const double den = dg[d] + de;
      const double d0 = dg[d] - dpn[d] + A.fac * l1p[d];
      if (!(l1[d] > 0.0) || !(den > 0.0) || !isfinite(den) || !(d0 > 0.0) || !isfinite(d0)) atomicAdd(A.bad, 1);
      A.delta3[3ll * i + d] = de;
      A.diag3[3ll * i + d] = den;
      A.rau3[3ll * i + d] = A.rauScale / den;
      A.r3[3ll * i + d] = A.rauScale / d0;
      A.brhs[3ll * i + d] = A.srhs[3ll * i + d] + c3[d] + t3[d] + de * A.u[3ll * i + d];
    }
  }
}

template <int LL, int LS>
__global__ void __launch_bounds__(256) k_wall_rows(WallRowsArgs A, int nLong, const int* __restrict__ longRows,
                                                   int nShort, const int* __restrict__ shortRows, int nbLong) {
#define NODALS_WALL_ROW_BODY(LN, q, lane, isLong) wall_row<LN>(A, (isLong) ? longRows[q] : shortRows[q], lane)
  NODALS_TWO_CLASS(LL, LS, nLong, nShort, nbLong, NODALS_WALL_ROW_BODY)
#undef NODALS_WALL_ROW_BODY
}

// M3 transform sources, one thread per source entry (face f, position k):
// right (K T): D = -(K_k n) n^T; left (T^T K): D = -n (n^T K_k).
static __global__ void k_m3_source(int ns, const int* __restrict__ sface, const int* __restrict__ spos,
                                   const double* __restrict__ wn, const double* __restrict__ K, int left,
                                   double* __restrict__ D) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ns) return;
  const int f = sface[e];
  const double n[3] = {wn[3 * f], wn[3 * f + 1], wn[3 * f + 2]};
  const double* Kk = K + 9ll * spos[e];
  double* De = D + 9ll * e;
  if (!left) {
    for (int r = 0; r < 3; ++r) {
      const double mn = Kk[3 * r] * n[0] + Kk[3 * r + 1] * n[1] + Kk[3 * r + 2] * n[2];
      for (int c = 0; c < 3; ++c) De[3 * r + c] = -mn * n[c];
    }
  } else {
    for (int c = 0; c < 3; ++c) {
      const double nm = n[0] * Kk[c] + n[1] * Kk[3 + c] + n[2] * Kk[6 + c];
      for (int r = 0; r < 3; ++r) De[3 * r + c] = -n[r] * nm;
    }
  }
}

// M3 transform targets: K[pos] += sum of its sources, the entry's own source
// with coefficient 1 and face-vertex sources with 20/27, in source order.
static __global__ void k_m3_apply(int nt, const int* __restrict__ tpos, const int* __restrict__ toff,
                                  const int* __restrict__ tsrc, const double* __restrict__ D, double* __restrict__ K) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int s = toff[t]; s < toff[t + 1]; ++s) {
    const int code = tsrc[s];
    const double cf = (code & 1) ? 20.0 / 27.0 : 1.0;
    const double* De = D + 9ll * (code >> 1);
    for (int j = 0; j < 9; ++j) acc[j] += cf * De[j];
  }
  double* Kk = K + 9ll * tpos[t];
  for (int j = 0; j < 9; ++j) Kk[j] += acc[j];
}

// Solves the 3x3 system a x = r by Gaussian elimination with partial
// pivoting (the reference's m3_solve3); false for a pivot below 1e-27.
__device__ __forceinline__ bool solve3(double a[3][3], double r[3], double x[3]) {
  for (int q = 0; q < 3; ++q) {
    int pv = q;
    for (int j = q + 1; j < 3; ++j)
      if (fabs(a[j][q]) > fabs(a[pv][q])) pv = j;
    const double v = fabs(a[pv][q]);
    if (!(v > 1.e-27) || !isfinite(v)) return false;
    if (pv != q) {
      for (int j = q; j < 3; ++j) {
        const double t = a[q][j];
        a[q][j] = a[pv][j];
        a[pv][j] = t;
      }
      const double t = r[q];
      r[q] = r[pv];
      r[pv] = t;
    }
    for (int i = q + 1; i < 3; ++i) {
      const double fq = a[i][q] / a[q][q];
      for (int j = q; j < 3; ++j) a[i][j] -= fq * a[q][j];
      r[i] -= fq * r[q];
    }
  }
  for (int i = 2; i >= 0; --i) {
    double v = r[i];
    for (int j = i + 1; j < 3; ++j) v -= a[i][j] * x[j];
    x[i] = v / a[i][i];
  }
  return isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]);
}

// One colour of the M3 block Gauss-Seidel sweep: r = b - K x over the row
// (the diagonal block included), the 3x3 diagonal block solve, x += omega
// sol.  A wall-face bubble row carries only tangential unknowns: its block
// gets scale n n^T added, its residual and correction are projected onto
// the tangent plane (m3_block_colour).
template <int L>
__device__ __forceinline__ void m3_row(int i, int lane, const int* __restrict__ mrow, const int* __restrict__ mcol,
                                       const int* __restrict__ mdiagPos, const double* __restrict__ K,
                                       const double* __restrict__ b, double* __restrict__ x,
                                       const int* __restrict__ bubFace, const double* __restrict__ wn, double omega,
                                       int* __restrict__ bad) {
  const int ke = mrow[i + 1];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int k = mrow[i] + lane; k < ke; k += L) {
    const int j = mcol[k];
    const double* Kk = K + 9ll * k;
    const double x0 = x[3ll * j], x1 = x[3ll * j + 1], x2 = x[3ll * j + 2];
    s0 += Kk[0] * x0 + Kk[1] * x1 + Kk[2] * x2;
    s1 += Kk[3] * x0 + Kk[4] * x1 + Kk[5] * x2;
    s2 += Kk[6] * x0 + Kk[7] * x1 + Kk[8] * x2;
  }
  s0 = groupSum(s0, L);
  s1 = groupSum(s1, L);
  s2 = groupSum(s2, L);
  if (lane != 0) return;
  double r[3] = {b[3ll * i] - s0, b[3ll * i + 1] - s1, b[3ll * i + 2] - s2};
  const double* Kd = K + 9ll * mdiagPos[i];
  double a[3][3];
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) a[p][q] = Kd[3 * p + q];
  const int f = bubFace[i];
  double n[3] = {0, 0, 0};
  if (f >= 0) {
    for (int d = 0; d < 3; ++d) n[d] = wn[3 * f + d];
    const double sc = fmax(1.e-20, fmax(fabs(a[0][0]), fmax(fabs(a[1][1]), fabs(a[2][2]))));
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) a[p][q] += sc * n[p] * n[q];
    const double rn = n[0] * r[0] + n[1] * r[1] + n[2] * r[2];
    for (int d = 0; d < 3; ++d) r[d] -= rn * n[d];
  }
  double sol[3] = {0, 0, 0};
  if (!solve3(a, r, sol)) {
    atomicAdd(bad, 1);
    return;
  }
  if (f >= 0) {
    const double sn = n[0] * sol[0] + n[1] * sol[1] + n[2] * sol[2];
    for (int d = 0; d < 3; ++d) sol[d] -= sn * n[d];
  }
  for (int d = 0; d < 3; ++d) x[3ll * i + d] += omega * sol[d];
}

template <int LL, int LS>
__global__ void __launch_bounds__(256) k_m3_gs(int r0, int rs, int r1, int nbLong, const int* __restrict__ mrow,
                                               const int* __restrict__ mcol, const int* __restrict__ mdiagPos,
                                               const double* __restrict__ K, const double* __restrict__ b,
                                               double* __restrict__ x, const int* __restrict__ bubFace,
                                               const double* __restrict__ wn, double omega, int* __restrict__ bad) {
  if ((int)blockIdx.x < nbLong) {
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int i = r0 + gid / LL;
    if (i < rs) m3_row<LL>(i, gid % LL, mrow, mcol, mdiagPos, K, b, x, bubFace, wn, omega, bad);
  } else {
    const int gid = ((int)blockIdx.x - nbLong) * 256 + threadIdx.x;
    const int i = rs + gid / LS;
    if (i < r1) m3_row<LS>(i, gid % LS, mrow, mcol, mdiagPos, K, b, x, bubFace, wn, omega, bad);
  }
}

// bn[f] = n_f . v[bubble_f] (v interleaved [nv][3]).
static __global__ void k_wall_bn(int nf, const int* __restrict__ bub, const double* __restrict__ wn,
                                 const double* __restrict__ v, double* __restrict__ bn) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double* V = v + 3ll * bub[f];
  bn[f] = wn[3 * f] * V[0] + wn[3 * f + 1] * V[1] + wn[3 * f + 2] * V[2];
}

// T^T on a test-function vector (s22d4i_Tt_apply) from bn = n . v_bubble:
// a bubble loses its normal component, every face vertex gains
// -(20/27) bn n of each of its wall faces (in face order).
static __global__ void k_wall_tt(int nd, const int* __restrict__ dof, const int* __restrict__ dofOff,
                                 const int* __restrict__ dofFace, const unsigned char* __restrict__ isBub,
                                 const double* __restrict__ wn, const double* __restrict__ bn, double* __restrict__ v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nd) return;
  double* V = v + 3ll * dof[t];
  if (isBub[t]) {
    const int f = dofFace[dofOff[t]];
    for (int d = 0; d < 3; ++d) V[d] -= wn[3 * f + d] * bn[f];
  } else {
    double a[3] = {0, 0, 0};
    for (int s = dofOff[t]; s < dofOff[t + 1]; ++s) {
      const int f = dofFace[s];
      const double cf = -(20.0 / 27.0) * bn[f];
      for (int d = 0; d < 3; ++d) a[d] += cf * wn[3 * f + d];
    }
    for (int d = 0; d < 3; ++d) V[d] += a[d];
  }
}

// T on a parameter vector (s22d4i_T_apply): each wall bubble gets n delta,
// delta = -n.v_b - (20/27) sum_vertices n.v_v, so that the face's mean
// normal flux vanishes.  mode 1: Q only (remove the bubble's normal part).
static __global__ void k_wall_t(int nf, const int* __restrict__ bub, const int* __restrict__ vert,
                                const double* __restrict__ wn, double* __restrict__ v, int qOnly) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double n[3] = {wn[3 * f], wn[3 * f + 1], wn[3 * f + 2]};
  double* V = v + 3ll * bub[f];
  const double bnv = n[0] * V[0] + n[1] * V[1] + n[2] * V[2];
  double delta = -bnv;
  if (!qOnly) {
    double sn = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double* X = v + 3ll * vert[3 * f + a];
      sn += n[0] * X[0] + n[1] * X[1] + n[2] * X[2];
    }
    delta -= (20.0 / 27.0) * sn;
  }
  for (int d = 0; d < 3; ++d) V[d] += n[d] * delta;
}

// Constrained Schur operator, stage 1: bn[f] = n_f . (B^T x) at the wall
// bubble of face f (a wall bubble's incidence is its one cell).
static __global__ void k_i3_bn(int nf, const int* __restrict__ bub, const double* __restrict__ wn,
                               const int* __restrict__ incOff, const unsigned* __restrict__ incPacked,
                               const double* __restrict__ bcoef, const double* __restrict__ x,
                               double* __restrict__ bn) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int i = bub[f];
  double w[3] = {0, 0, 0};
  for (int k = incOff[i]; k < incOff[i + 1]; ++k) {
    const unsigned pk = incPacked[k];
    const int c = (int)(pk >> 3), a = (int)(pk & 7u);
    const double* B = bcoef + (long long)c * 24 + 3 * a;
    for (int d = 0; d < 3; ++d) w[d] += B[d] * x[c];
  }
  bn[f] = wn[3 * f] * w[0] + wn[3 * f + 1] * w[1] + wn[3 * f + 2] * w[2];
}

// Stage 2, one thread per free DOF: y = Q R (T^T (B^T x)), R = diag(r3).
static __global__ void k_i3_dof(int nv, const int* __restrict__ incOff, const unsigned* __restrict__ incPacked,
                                const double* __restrict__ bcoef, const double* __restrict__ x,
                                const double* __restrict__ r3, const int* __restrict__ wallDof,
                                const int* __restrict__ dofOff, const int* __restrict__ dofFace,
                                const unsigned char* __restrict__ isBub, const double* __restrict__ wn,
                                const double* __restrict__ bn, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  double w[3] = {0, 0, 0};
  for (int k = incOff[i]; k < incOff[i + 1]; ++k) {
    const unsigned pk = incPacked[k];
    const int c = (int)(pk >> 3), a = (int)(pk & 7u);
    const double* B = bcoef + (long long)c * 24 + 3 * a;
    for (int d = 0; d < 3; ++d) w[d] += B[d] * x[c];
  }
  const int t = wallDof[i];
  int bf = -1;
  if (t >= 0) {
    if (isBub[t]) {
      bf = dofFace[dofOff[t]];
      for (int d = 0; d < 3; ++d) w[d] -= wn[3 * bf + d] * bn[bf];
    } else {
      for (int s = dofOff[t]; s < dofOff[t + 1]; ++s) {
        const int f = dofFace[s];
        const double cf = -(20.0 / 27.0) * bn[f];
        for (int d = 0; d < 3; ++d) w[d] += cf * wn[3 * f + d];
      }
    }
  }
  for (int d = 0; d < 3; ++d) w[d] *= r3[3ll * i + d];
  if (bf >= 0) {
    const double wnn = wn[3 * bf] * w[0] + wn[3 * bf + 1] * w[1] + wn[3 * bf + 2] * w[2];
    for (int d = 0; d < 3; ++d) w[d] -= wn[3 * bf + d] * wnn;
  }
  for (int d = 0; d < 3; ++d) y[3ll * i + d] = w[d];
}

// Stage 3, one thread per cell: out = B (T y); the wall bubbles of the
// cell's wall faces are taken at their T-constrained values.
static __global__ void k_i3_cell(int nc, const int* __restrict__ cref, const double* __restrict__ bcoef,
                                 const double* __restrict__ y, const int* __restrict__ cellWallOff,
                                 const int* __restrict__ cellWall, const int* __restrict__ wlf,
                                 const int* __restrict__ bub, const int* __restrict__ vert,
                                 const double* __restrict__ wn, double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double* B = bcoef + (long long)c * 24;
  double s = 0.0;
  for (int a = 0; a < 8; ++a) {
    const int g = cref[c * 8 + a];
    if (g < 0) continue;
    s += B[3 * a] * y[3ll * g] + B[3 * a + 1] * y[3ll * g + 1] + B[3 * a + 2] * y[3ll * g + 2];
  }
  for (int t = cellWallOff[c]; t < cellWallOff[c + 1]; ++t) {
    const int f = cellWall[t];
    const double n[3] = {wn[3 * f], wn[3 * f + 1], wn[3 * f + 2]};
    const double* Yb = y + 3ll * bub[f];
    double delta = -(n[0] * Yb[0] + n[1] * Yb[1] + n[2] * Yb[2]);
    double sn = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double* X = y + 3ll * vert[3 * f + a];
      sn += n[0] * X[0] + n[1] * X[1] + n[2] * X[2];
    }
    delta -= (20.0 / 27.0) * sn;
    const double* Bb = B + 3 * (4 + wlf[f]);
    s += delta * (Bb[0] * n[0] + Bb[1] * n[1] + Bb[2] * n[2]);
  }
  out[c] = s;
}

// W[c][a][d] = rAU_d(g) B[c][a][d] (zero for fixed DOFs): the anisotropic
// weights of the fine Schur refresh B diag(rAU_x, rAU_y, rAU_z) B^T.
static __global__ void k_fine_weights3(int nc, const int* __restrict__ cref, const double* __restrict__ bcoef,
                                       const double* __restrict__ rau3, double* __restrict__ W) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nc * 8) return;
  const int g = cref[t];
  for (int d = 0; d < 3; ++d) W[3ll * t + d] = g >= 0 ? rau3[3ll * g + d] * bcoef[3ll * t + d] : 0.0;
}

// Wall shear of each face from its frozen fit: the traction
// -beta (I - n n^T) u_pic at each quadrature point, integrated:
// out[f] = (int t_x, int t_y, int t_z, int |t|) over the face.
static __global__ void k_wall_shear(int nf, const int* __restrict__ wcell, const double* __restrict__ wn,
                                    const double* __restrict__ qw, const double* __restrict__ bp,
                                    const double* __restrict__ fit, const int* __restrict__ cref,
                                    const double* __restrict__ u, const double* __restrict__ ufix,
                                    double* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int c = wcell[f];
  const double n[3] = {wn[3 * f], wn[3 * f + 1], wn[3 * f + 2]};
  const double beta = fit[3 * f];
  double t[4] = {0, 0, 0, 0};
  for (int q = 0; q < 12; ++q) {
    double pic[3] = {0, 0, 0};
    for (int b = 0; b < 8; ++b) {
      const int r = cref[c * 8 + b];
      const double* U = r >= 0 ? u + 3ll * r : ufix + 3ll * (-r - 1);
      const double z = bp[((long long)f * 12 + q) * 8 + b];
      for (int d = 0; d < 3; ++d) pic[d] += z * U[d];
    }
    const double pn = pic[0] * n[0] + pic[1] * n[1] + pic[2] * n[2];
    double tq[3], m = 0.0;
    for (int d = 0; d < 3; ++d) {
      tq[d] = -beta * (pic[d] - pn * n[d]);
      m += tq[d] * tq[d];
    }
    const double w = qw[f * 12 + q];
    for (int d = 0; d < 3; ++d) t[d] += w * tq[d];
    t[3] += w * sqrt(m);
  }
  for (int d = 0; d < 4; ++d) out[4 * f + d] = t[d];
}

// Residual r = b - K x of the M3 block system (every 3x3 block of the row,
// the diagonal one included), one thread per row; for residual norms.
static __global__ void k_m3_resid(int n, const int* __restrict__ mrow, const int* __restrict__ mcol,
                                  const double* __restrict__ K, const double* __restrict__ b,
                                  const double* __restrict__ x, double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s0 = b[3ll * i], s1 = b[3ll * i + 1], s2 = b[3ll * i + 2];
  for (int k = mrow[i]; k < mrow[i + 1]; ++k) {
    const long long j = mcol[k];
    const double* Kk = K + 9ll * k;
    const double x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
    s0 -= Kk[0] * x0 + Kk[1] * x1 + Kk[2] * x2;
    s1 -= Kk[3] * x0 + Kk[4] * x1 + Kk[5] * x2;
    s2 -= Kk[6] * x0 + Kk[7] * x1 + Kk[8] * x2;
  }
  r[3ll * i] = s0;
  r[3ll * i + 1] = s1;
  r[3ll * i + 2] = s2;
}

}  // namespace dev
}  // namespace nodals
