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

}  // namespace dev
}  // namespace nodals
