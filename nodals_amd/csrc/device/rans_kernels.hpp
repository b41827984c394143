// Optional momentum terms of the H8 binary, as cell blocks added to the
// convection blocks before the momentum row gather (so they reach the
// matrix and, for Dirichlet columns, the right-hand side exactly as the
// convection does):
//   mixing length  int nu_t grad(phi_a).grad(phi_b), nu_t = scale ell^2
//                  sqrt(2 S:S) of the lagged velocity, ell / R = max(0,
//                  0.14 - 0.08 eta^2 - 0.06 eta^4), eta = r / R the distance
//                  from the pipe axis (Nikuradse's smooth-pipe profile;
//                  gpu/serial_cuda/h9_rans_mixlen.cuh:8-16, 51-190);
//   SUPG           sum_q tau w_q det (a.grad phi_a)(a.grad phi_b - nu lap
//                  phi_b), a the lagged velocity at q, tau = tauScale /
//                  sqrt(4 |a|^2 / h^2 + magic (4 nu / h^2)^2), h^2 the
//                  cell's largest squared edge (implicit form;
//                  h7_momentum_assembly.cuh:101-251, h7_supg64_warp_cell_kernel).
// Both integrate with the 64-point collapsed (Stroud conical product) rule
// of the tetrahedron: 4-point Gauss-Jacobi in r (weight (1-r)^2) and s
// (weight 1-s), Gauss-Legendre in t, lambda = ((1-r)(1-s)(1-t), r, (1-r)s,
// (1-r)(1-s)t), weights summing to 1/6.
#pragma once

#include <hip/hip_runtime.h>

namespace nodals {
namespace dev {

// Nodes and weights on [0, 1] (computed from the Gauss-Jacobi and
// Gauss-Legendre rules; the host test recomputes them).
struct Collapsed64 {
  double rn[4], rw[4], sn[4], sw[4], tn[4], tw[4];
};

inline Collapsed64 collapsed64() {
  return {{0.048500549446997276, 0.23860073755186234, 0.5170472951043674, 0.7958514178967728},
          {0.11088841561127774, 0.14345878979921445, 0.0686338871729231, 0.010352240749918081},
          {0.057104196114517725, 0.2768430136381238, 0.5835904323689168, 0.8602401356562195},
          {0.13550691343148852, 0.2034645680102711, 0.12984754760823233, 0.031180970950008085},
          {0.06943184420297371, 0.33000947820757187, 0.6699905217924281, 0.9305681557970262},
          {0.1739274225687269, 0.3260725774312731, 0.3260725774312731, 0.1739274225687269}};
}

struct RansArgs {
  int nc;
  const int* cref;       // [nc][8] free DOF (device order) or -(fixed slot + 1)
  const double* cinvJ;   // [nc][9]
  const double* cdet;    // [nc]
  const double* ch2;     // [nc] largest squared edge
  const double* cxy;     // [nc][4][2] vertex x, y relative to the pipe axis
  const double* u;       // [nv][3]
  const double* ufix;    // [nfixed][3]
  double radius, nu;
  int mixing;            // mixing-length term on
  double mixScale;
  int supg;              // SUPG term on
  double tauScale, magic;
  double* cbuf;          // [nc][64] cell blocks (added to)
  double* stats;         // [nc][4] mixing length: sum nu_t/nu w, sum w, max nu_t/nu, max strain
};

// One wavefront per cell (blocks of 64 threads, so the barriers are
// wave-local); lane 8a + b owns block entry (a, b).
static __global__ void __launch_bounds__(64) k_rans64(RansArgs A, Collapsed64 Q) {
  __shared__ double U[8][3], G[8][3], PHI[8], LAP[8], GU[9], ADV[3], STR[8], STRONG[8], WM, WS, GL[4][3];
  const int lane = threadIdx.x;
  const int a = lane >> 3, b = lane & 7;
  for (int c = blockIdx.x; c < A.nc; c += gridDim.x) {
    const double* I = A.cinvJ + 9ll * c;
    const double det = A.cdet[c];
    if (lane < 8) {
      const int r = A.cref[8ll * c + lane];
      const double* v = r >= 0 ? A.u + 3ll * r : A.ufix + 3ll * (-r - 1);
      U[lane][0] = v[0];
      U[lane][1] = v[1];
      U[lane][2] = v[2];
    } else if (lane < 11) {  // gradients of lambda_1..3: rows of the inverse Jacobian
      const int i = lane - 8;
      GL[i + 1][0] = I[3 * i];
      GL[i + 1][1] = I[3 * i + 1];
      GL[i + 1][2] = I[3 * i + 2];
    }
    __syncthreads();
    if (lane < 3) GL[0][lane] = -(I[lane] + I[3 + lane] + I[6 + lane]);
    __syncthreads();
    double acc = 0.0, ratioW = 0.0, wsum = 0.0, ratioMax = 0.0, strainMax = 0.0;
    for (int iq = 0; iq < 64; ++iq) {
      const int ir = iq >> 4, is = (iq >> 2) & 3, it = iq & 3;
      const double r = Q.rn[ir], s = Q.sn[is], t = Q.tn[it];
      const double lam[4] = {(1.0 - r) * (1.0 - s) * (1.0 - t), r, (1.0 - r) * s, (1.0 - r) * (1.0 - s) * t};
      const double wq = Q.rw[ir] * Q.sw[is] * Q.tw[it] * det;
      if (lane < 8) {
        double g0, g1, g2, phi, lap = 0.0;
        if (lane < 4) {
          phi = lam[lane];
          g0 = GL[lane][0];
          g1 = GL[lane][1];
          g2 = GL[lane][2];
        } else {  // face bubble 27 l_i l_j l_k opposite vertex lane - 4
          const int o = lane - 4;
          const int i = o == 0 ? 1 : 0, j = o <= 1 ? 2 : 1, k = o <= 2 ? 3 : 2;
          phi = 27.0 * lam[i] * lam[j] * lam[k];
          g0 = 27.0 * (lam[j] * lam[k] * GL[i][0] + lam[i] * lam[k] * GL[j][0] + lam[i] * lam[j] * GL[k][0]);
          g1 = 27.0 * (lam[j] * lam[k] * GL[i][1] + lam[i] * lam[k] * GL[j][1] + lam[i] * lam[j] * GL[k][1]);
          g2 = 27.0 * (lam[j] * lam[k] * GL[i][2] + lam[i] * lam[k] * GL[j][2] + lam[i] * lam[j] * GL[k][2]);
          double dij = 0.0, dik = 0.0, djk = 0.0;
          for (int d = 0; d < 3; ++d) {
            dij += GL[i][d] * GL[j][d];
            dik += GL[i][d] * GL[k][d];
            djk += GL[j][d] * GL[k][d];
          }
          lap = 54.0 * (lam[k] * dij + lam[j] * dik + lam[i] * djk);
        }
        G[lane][0] = g0;
        G[lane][1] = g1;
        G[lane][2] = g2;
        PHI[lane] = phi;
        LAP[lane] = lap;
      }
      __syncthreads();
      if (lane < 9) {  // grad U, component lane / 3 along direction lane % 3
        const int comp = lane / 3, dir = lane % 3;
        double z = 0.0;
        for (int m = 0; m < 8; ++m) z += U[m][comp] * G[m][dir];
        GU[lane] = z;
      } else if (lane < 12) {  // advecting velocity
        const int d = lane - 9;
        double z = 0.0;
        for (int m = 0; m < 8; ++m) z += PHI[m] * U[m][d];
        ADV[d] = z;
      }
      __syncthreads();
      if (lane < 8) {
        const double st = ADV[0] * G[lane][0] + ADV[1] * G[lane][1] + ADV[2] * G[lane][2];
        STR[lane] = st;
        STRONG[lane] = st - A.nu * LAP[lane];
      } else if (lane == 8) {
        double wm = 0.0;
        if (A.mixing) {
          double ss = 0.0;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              const double sij = 0.5 * (GU[3 * i + j] + GU[3 * j + i]);
              ss += sij * sij;
            }
          const double strain = sqrt(fmax(0.0, 2.0 * ss));
          const double* xy = A.cxy + 8ll * c;
          double xq = 0.0, yq = 0.0;
          for (int i = 0; i < 4; ++i) {
            xq += lam[i] * xy[2 * i];
            yq += lam[i] * xy[2 * i + 1];
          }
          const double eta = fmin(1.0, fmax(0.0, sqrt(xq * xq + yq * yq) / A.radius));
          const double e2 = eta * eta;
          const double ell = A.radius * fmax(0.0, 0.14 - 0.08 * e2 - 0.06 * e2 * e2);
          const double nut = A.mixScale * ell * ell * strain;
          wm = nut * wq;
          const double ratio = nut / A.nu;
          ratioW += ratio * wq;
          wsum += wq;
          ratioMax = fmax(ratioMax, ratio);
          strainMax = fmax(strainMax, strain);
        }
        WM = wm;
      } else if (lane == 9) {
        double ws = 0.0;
        if (A.supg) {
          const double speed2 = ADV[0] * ADV[0] + ADV[1] * ADV[1] + ADV[2] * ADV[2];
          const double h2 = A.ch2[c];
          const double diff = 4.0 * A.nu / h2;
          const double den = fmax(4.0 * speed2 / h2 + A.magic * diff * diff, 1e-30);
          ws = A.tauScale / sqrt(den) * wq;
        }
        WS = ws;
      }
      __syncthreads();
      acc += WM * (G[a][0] * G[b][0] + G[a][1] * G[b][1] + G[a][2] * G[b][2]) + WS * STR[a] * STRONG[b];
      __syncthreads();
    }
    A.cbuf[64ll * c + lane] += acc;
    if (lane == 8 && A.stats) {
      double* st = A.stats + 4ll * c;
      st[0] = ratioW;
      st[1] = wsum;
      st[2] = ratioMax;
      st[3] = strainMax;
    }
    __syncthreads();
  }
}

// Twelve-point triangle rule rows (barycentrics on the face's three
// vertices in increasing local order, weight; weights sum to 1).
struct Tri12Rule {
  double q[12][4];
};

struct DgInletArgs {
  int nf;
  const int* cell;       // [nf] inlet cell (one inlet face per cell)
  const int* lf;         // [nf] local inlet face (opposite vertex lf)
  const double* geom;    // [nf][4] area, outward unit normal
  const int* cref;
  const double* cinvJ;
  const double* cxy;     // mixing length: vertex x, y relative to the axis
  const double* u;
  const double* ufix;
  double nu, radius, mixScale;
  int mixing;
  double uh[3];          // prescribed inlet velocity
  double* cbuf;          // [nc][64] cell blocks (added to)
  double* faceRhs;       // [nf][8][3]
};

// DG weak inlet, one thread per inlet face (gpu/serial_cuda/h10_dg_inlet.cuh:
// 94-133): on the face's 12-point rule, with beta = max(0, -uh.n) and
// nu_e = nu + nu_t (the mixing length when it is on),
//   G_ab += w (beta phi_a phi_b + nu_e (dn(phi_a) phi_b - phi_a dn(phi_b)))
//   R_a  += w (beta phi_a + nu_e dn(phi_a)) uh
// G joins the cell block (matrix and Dirichlet columns), R the momentum
// right-hand side through k_dg_rhs.
static __global__ void __launch_bounds__(64) k_dg_inlet(DgInletArgs A, Tri12Rule T) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.nf) return;
  const int c = A.cell[k], o = A.lf[k];
  const double* g = A.geom + 4ll * k;
  const double area = g[0], n0 = g[1], n1 = g[2], n2 = g[3];
  const double* I = A.cinvJ + 9ll * c;
  double GL[4][3];
  for (int d = 0; d < 3; ++d) {
    GL[1][d] = I[d];
    GL[2][d] = I[3 + d];
    GL[3][d] = I[6 + d];
    GL[0][d] = -(I[d] + I[3 + d] + I[6 + d]);
  }
  double U[8][3];
  for (int a = 0; a < 8; ++a) {
    const int r = A.cref[8ll * c + a];
    const double* v = r >= 0 ? A.u + 3ll * r : A.ufix + 3ll * (-r - 1);
    U[a][0] = v[0];
    U[a][1] = v[1];
    U[a][2] = v[2];
  }
  int fv[3], m = 0;
  for (int i = 0; i < 4; ++i)
    if (i != o) fv[m++] = i;
  const double bn = A.uh[0] * n0 + A.uh[1] * n1 + A.uh[2] * n2;
  const double inflow = fmax(0.0, -bn);
  double G[64], R[24];
  for (int i = 0; i < 64; ++i) G[i] = 0.0;
  for (int i = 0; i < 24; ++i) R[i] = 0.0;
  for (int q = 0; q < 12; ++q) {
    double lam[4] = {0.0, 0.0, 0.0, 0.0};
    lam[fv[0]] = T.q[q][0];
    lam[fv[1]] = T.q[q][1];
    lam[fv[2]] = T.q[q][2];
    double val[8], gr[8][3], dn[8];
    for (int a = 0; a < 8; ++a) {
      if (a < 4) {
        val[a] = lam[a];
        for (int d = 0; d < 3; ++d) gr[a][d] = GL[a][d];
      } else {
        const int b = a - 4;
        const int i = b == 0 ? 1 : 0, j = b <= 1 ? 2 : 1, l = b <= 2 ? 3 : 2;
        val[a] = 27.0 * lam[i] * lam[j] * lam[l];
        for (int d = 0; d < 3; ++d)
          gr[a][d] = 27.0 * (lam[j] * lam[l] * GL[i][d] + lam[i] * lam[l] * GL[j][d] + lam[i] * lam[j] * GL[l][d]);
      }
      dn[a] = gr[a][0] * n0 + gr[a][1] * n1 + gr[a][2] * n2;
    }
    double nut = 0.0;
    if (A.mixing) {
      double gu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) gu[i][j] += U[a][i] * gr[a][j];
      double ss = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const double sij = 0.5 * (gu[i][j] + gu[j][i]);
          ss += sij * sij;
        }
      const double strain = sqrt(fmax(0.0, 2.0 * ss));
      const double* xy = A.cxy + 8ll * c;
      double xq = 0.0, yq = 0.0;
      for (int i = 0; i < 4; ++i) {
        xq += lam[i] * xy[2 * i];
        yq += lam[i] * xy[2 * i + 1];
      }
      const double eta = fmin(1.0, fmax(0.0, sqrt(xq * xq + yq * yq) / A.radius));
      const double e2 = eta * eta;
      const double ell = A.radius * fmax(0.0, 0.14 - 0.08 * e2 - 0.06 * e2 * e2);
      nut = A.mixScale * ell * ell * strain;
    }
    const double nue = A.nu + nut, w = T.q[q][3] * area;
    for (int a = 0; a < 8; ++a) {
      for (int b = 0; b < 8; ++b) G[8 * a + b] += w * (inflow * val[a] * val[b] + nue * (-val[a] * dn[b] + dn[a] * val[b]));
      for (int d = 0; d < 3; ++d) R[3 * a + d] += w * (inflow * val[a] * A.uh[d] + nue * dn[a] * A.uh[d]);
    }
  }
  double* blk = A.cbuf + 64ll * c;
  for (int i = 0; i < 64; ++i) blk[i] += G[i];
  double* fr = A.faceRhs + 24ll * k;
  for (int i = 0; i < 24; ++i) fr[i] = R[i];
}

// Adds the inlet faces' right-hand sides to their free rows, one thread per
// row, terms in (face, basis) order: rows[r] (device numbering) sums
// faceRhs over list[off[r] .. off[r+1]) (face * 8 + basis).
static __global__ void k_dg_rhs(int nrows, const int* __restrict__ rows, const int* __restrict__ off,
                                const int* __restrict__ list, const double* __restrict__ faceRhs,
                                double* __restrict__ brhs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int t = off[r]; t < off[r + 1]; ++t) {
    const double* f = faceRhs + 3ll * list[t];
    s0 += f[0];
    s1 += f[1];
    s2 += f[2];
  }
  const long long i = rows[r];
  brhs[3 * i] += s0;
  brhs[3 * i + 1] += s1;
  brhs[3 * i + 2] += s2;
}

}  // namespace dev
}  // namespace nodals
