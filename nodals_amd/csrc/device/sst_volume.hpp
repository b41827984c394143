// k-omega SST volume terms of one tetrahedron for the scalar k or omega
// equation, on the P1+BF3 element (4 vertex + 4 face-bubble functions):
//   A_ab = sum_q w_q det [ D_q grad(phi_a).grad(phi_b) + R_q phi_a phi_b ]
//   b_a  = sum_q w_q det S_q phi_a
// with D = sigma_{k|w} nu_t, R = beta* omega (k) or beta omega (omega),
// S = P_k (k) or the omega production plus cross diffusion, all from the
// SST-2003 local state at each of the 24 points of Keast's degree-6 rule
// (reference kernel sst_tet_volume,
// turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/modules/transport/
// tet_volume_kernel.inc:1-30, with volume_coefficients,
// modules/transport/portable/transport_core.hpp:14-25).  One wavefront per
// cell: lanes 0..23 evaluate the points, then lane 8a+b forms A_ab and
// lanes 0..7 b_a from the per-point terms staged in LDS; every cell writes
// its own 8x8 block (no atomics).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <vector>

#include "sst_physics.hpp"

namespace nodals {

// Keast's 24-point rule on the reference tetrahedron (degree 6; weights sum
// to 1/6): three 4-point orbits (a,a,a,b) and one 12-point orbit of the
// distinct permutations of (a,a,b,c), in the point order of the reference
// table (hxt5a_mixlen.cuh:37-142).  Rows of (l0, l1, l2, l3, w).
inline std::vector<double> keast24() {
  std::vector<double> q;
  auto put = [&](const double l[4], double w) {
    for (int i = 0; i < 4; ++i) q.push_back(l[i]);
    q.push_back(w);
  };
  struct Orbit {
    double a, b, w;
    bool bFirst;  // b at slot i (else at slot 3 - i) for point i of the orbit
  };
  const Orbit o4[3] = {{0.214602871259151673, 0.356191386222544926, 0.00665379170969464511, false},
                       {0.0406739585346113375, 0.87797812439616596, 0.00167953517588677629, false},
                       {0.322337890142275652, 0.0329863295731730594, 0.00922619692394239903, true}};
  for (const Orbit& o : o4)
    for (int i = 0; i < 4; ++i) {
      double l[4] = {o.a, o.a, o.a, o.a};
      l[o.bFirst ? i : 3 - i] = o.b;
      put(l, o.w);
    }
  const double v[3] = {0.0636610018750175255, 0.269672331458315873, 0.603005664791649076};
  std::array<int, 4> rank = {0, 0, 1, 2};
  do {
    const double l[4] = {v[rank[0]], v[rank[1]], v[rank[2]], v[rank[3]]};
    put(l, 0.00803571428571428319);
  } while (std::next_permutation(rank.begin(), rank.end()));
  return q;
}

namespace dev {

constexpr int kSstQ = 24;

// P1 + face-bubble basis (bubble i opposite vertex i: 27 prod_{j != i} l_j)
// and its reference gradients.
__device__ inline void p1bf3Basis(const double l[4], double phi[8], double gr[8][3]) {
  const double gl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 4; ++i) {
    phi[i] = l[i];
    for (int d = 0; d < 3; ++d) gr[i][d] = gl[i][d];
  }
  for (int i = 0; i < 4; ++i) {
    int js[3], k = 0;
    for (int j = 0; j < 4; ++j)
      if (j != i) js[k++] = j;
    phi[4 + i] = 27.0 * l[js[0]] * l[js[1]] * l[js[2]];
    for (int d = 0; d < 3; ++d)
      gr[4 + i][d] = 27.0 * (gl[js[0]][d] * l[js[1]] * l[js[2]] + l[js[0]] * gl[js[1]][d] * l[js[2]] +
                             l[js[0]] * l[js[1]] * gl[js[2]][d]);
  }
}

// Inputs per cell: DOF refs (>= 0 device free DOF, < 0 -(fixed slot + 1)),
// invJ (row-major dxi_j/dx_d), det, the (x, y) of the four vertices relative
// to the pipe axis; fields: velocity (xyz interleaved, free and fixed) and
// k, omega (free and fixed).  Outputs: mat[c][64], rhs[c][8].
// Transport-step form: with sref (scalar DOF of every cell basis) k and
// omega are read as kf[sref], wf[sref]; nuBase is added to the diffusion
// coefficient (the molecular part, which Keast's rule integrates exactly
// for this basis) and conv[c][64], when given, is added to the matrix (conv
// may be mat itself: every lane reads and writes only its own entry).
// equation 2 is the momentum form: the eddy-viscosity diffusion block
// sum_q w_q det nu_t grad phi_a . grad phi_b only (no reaction, no rhs).
// With dmat (momentum form only) every cell also writes the deviatoric
// cross-component tensor of the implicit deviatoric stress
// (modules/coupling/implicit_dev_volume_tet.inc:1-68):
//   dmat[c][ab][3i+j] = sum_q w_q det (nu + nu_t)
//                       (d_j phi_a d_i phi_b - (2/3) d_i phi_a d_j phi_b).
static __global__ void __launch_bounds__(256) k_sst_volume(
    int nc, const int* __restrict__ cref, const double* __restrict__ cinvJ, const double* __restrict__ cdet,
    const double* __restrict__ cxy, const double* __restrict__ u, const double* __restrict__ ufix,
    const double* __restrict__ kf, const double* __restrict__ kfix, const double* __restrict__ wf,
    const double* __restrict__ wfix, const double* __restrict__ quad, double radius, double nu, int equation,
    double kFloor, double omFloor, int deviatoric, double* mat, double* __restrict__ rhs,
    const int* __restrict__ sref = nullptr, double nuBase = 0.0, const double* conv = nullptr,
    const double* __restrict__ vdist = nullptr, double* __restrict__ dmat = nullptr) {
  __shared__ double PH[4][kSstQ][8], GP[4][kSstQ][8][3], DW[4][kSstQ], RW[4][kSstQ], SW[4][kSstQ], NW[4][kSstQ];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + wv;
  const bool live = c < nc;
  if (live && lane < kSstQ) {
    const int q = lane;
    const double* Q = quad + 5 * q;
    const double l[4] = {Q[0], Q[1], Q[2], Q[3]};
    double phi[8], gr[8][3];
    p1bf3Basis(l, phi, gr);
    const double* I = cinvJ + (long long)c * 9;
    double gu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gk[3] = {0, 0, 0}, gw[3] = {0, 0, 0}, kv = 0, wvv = 0;
    for (int a = 0; a < 8; ++a) {
      const int r = cref[(long long)c * 8 + a];
      const double* U = r >= 0 ? u + 3ll * r : ufix + 3ll * (-r - 1);
      double Ka, Wa;
      if (sref) {
        const int g = sref[(long long)c * 8 + a];
        Ka = kf[g];
        Wa = wf[g];
      } else {
        Ka = r >= 0 ? kf[r] : kfix[-r - 1];
        Wa = r >= 0 ? wf[r] : wfix[-r - 1];
      }
      const double pa = phi[a];
      PH[wv][q][a] = pa;
      kv += Ka * pa;
      wvv += Wa * pa;
      for (int d = 0; d < 3; ++d) {
        const double gd = gr[a][0] * I[d] + gr[a][1] * I[3 + d] + gr[a][2] * I[6 + d];
        GP[wv][q][a][d] = gd;
        gk[d] += Ka * gd;
        gw[d] += Wa * gd;
        gu[d] += U[0] * gd;
        gu[3 + d] += U[1] * gd;
        gu[6 + d] += U[2] * gd;
      }
    }
    double dq;
    if (vdist) {  // P1 interpolation of the vertex wall-distance field
      double dd = 0;
      for (int a = 0; a < 4; ++a) dd += l[a] * vdist[sref[(long long)c * 8 + a]];
      dq = sst::floorDistance(fmax(0.0, dd), radius);
    } else {
      double xq = 0, yq = 0;
      for (int a = 0; a < 4; ++a) {
        xq += l[a] * cxy[(long long)c * 8 + 2 * a];
        yq += l[a] * cxy[(long long)c * 8 + 2 * a + 1];
      }
      dq = sst::pipeDistance(xq, yq, radius);
    }
    const sst::Local s = sst::local(dq, nu, gu, kv, wvv, gk, gw, kFloor, omFloor, deviatoric != 0);
    const double wt = cdet[c] * Q[4];
    const bool isK = equation == 0;
    if (equation == 2) {
      DW[wv][q] = wt * s.nut;
      RW[wv][q] = 0.0;
      SW[wv][q] = 0.0;
      NW[wv][q] = (nu + s.nut) * wt;
    } else {
      DW[wv][q] = wt * ((isK ? s.sigmaK : s.sigmaW) * s.nut + nuBase);
      RW[wv][q] = wt * (isK ? s.reactK : s.reactOmega);
      SW[wv][q] = wt * (isK ? s.pk : s.srcOmega);
    }
  }
  __syncthreads();
  if (!live) return;
  {
    const int a = lane >> 3, b = lane & 7;
    double z = 0.0;
    for (int q = 0; q < kSstQ; ++q) {
      const double gd = GP[wv][q][a][0] * GP[wv][q][b][0] + GP[wv][q][a][1] * GP[wv][q][b][1] +
                        GP[wv][q][a][2] * GP[wv][q][b][2];
      z += DW[wv][q] * gd + RW[wv][q] * PH[wv][q][a] * PH[wv][q][b];
    }
    mat[(long long)c * 64 + lane] = conv ? z + conv[(long long)c * 64 + lane] : z;
    if (dmat) {
      double* Dm = dmat + ((long long)c * 64 + lane) * 9;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double v = 0.0;
          for (int q = 0; q < kSstQ; ++q)
            v += NW[wv][q] * (GP[wv][q][a][j] * GP[wv][q][b][i] - (2.0 / 3.0) * GP[wv][q][a][i] * GP[wv][q][b][j]);
          Dm[3 * i + j] = v;
        }
    }
  }
  if (rhs && lane < 8) {
    double z = 0.0;
    for (int q = 0; q < kSstQ; ++q) z += SW[wv][q] * PH[wv][q][lane];
    rhs[(long long)c * 8 + lane] = z;
  }
}

}  // namespace dev
}  // namespace nodals
