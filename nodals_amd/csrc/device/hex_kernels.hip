// Q1+BF2 hexahedral element blocks on gfx950: the 14-DOF hexahedron of the
// SST application (8 trilinear vertex functions and 6 BF2 face bubbles,
// isoparametric Q1 geometry, 3 x 3 x 3 Gauss), its diffusion block
//   D_ab = sum_q nu det_q w_q grad(phi_a) . grad(phi_b)
// and its Picard convection block
//   C_ab = sum_q det_q w_q phi_a (u_q . grad(phi_b)),  u_q = sum_m u_m phi_m.
//
// Reference: turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
// modules/foundation/hxt2_foundation.inc:166-225 (hex_q1_ref, hex_basis_ref,
// hex_metric, phys_grad14, hex_volume_kernel) and hxt4b_foundation.inc:
// 246-345 (hex_conv_assemble_cached).
//
// One wavefront per hexahedron: lanes 0..26 evaluate the basis, the metric
// and the velocity at their Gauss point into LDS, then every lane sums its
// (a, b) pairs (196 per element, three or four per lane) over the 27 points
// in point order.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"

namespace nodals {
namespace dev {
namespace {

__constant__ double kG3x[3] = {-0.774596669241483377035853079956, 0.0, 0.774596669241483377035853079956};
__constant__ double kG3w[3] = {0.555555555555555555555555555556, 0.888888888888888888888888888889,
                               0.555555555555555555555555555556};

__device__ void hexBasis(double x, double y, double z, double phi[14], double gr[14][3]) {
  const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
               sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  for (int a = 0; a < 8; ++a) {
    const double ax = 1.0 + sx[a] * x, ay = 1.0 + sy[a] * y, az = 1.0 + sz[a] * z;
    phi[a] = 0.125 * ax * ay * az;
    gr[a][0] = 0.125 * sx[a] * ay * az;
    gr[a][1] = 0.125 * sy[a] * ax * az;
    gr[a][2] = 0.125 * sz[a] * ax * ay;
  }
  const double bx = 1.0 - x * x, by = 1.0 - y * y, bz = 1.0 - z * z;
  const double dx = -2.0 * x, dy = -2.0 * y, dz = -2.0 * z;
  // faces z-, z+, y-, x+, y+, x-
  double s = 0.5 * (1.0 - z);
  phi[8] = s * bx * by; gr[8][0] = s * dx * by; gr[8][1] = s * bx * dy; gr[8][2] = -0.5 * bx * by;
  s = 0.5 * (1.0 + z);
  phi[9] = s * bx * by; gr[9][0] = s * dx * by; gr[9][1] = s * bx * dy; gr[9][2] = 0.5 * bx * by;
  s = 0.5 * (1.0 - y);
  phi[10] = s * bx * bz; gr[10][0] = s * dx * bz; gr[10][1] = -0.5 * bx * bz; gr[10][2] = s * bx * dz;
  s = 0.5 * (1.0 + x);
  phi[11] = s * by * bz; gr[11][0] = 0.5 * by * bz; gr[11][1] = s * dy * bz; gr[11][2] = s * by * dz;
  s = 0.5 * (1.0 + y);
  phi[12] = s * bx * bz; gr[12][0] = s * dx * bz; gr[12][1] = 0.5 * bx * bz; gr[12][2] = s * bx * dz;
  s = 0.5 * (1.0 - x);
  phi[13] = s * by * bz; gr[13][0] = -0.5 * by * bz; gr[13][1] = s * dy * bz; gr[13][2] = s * by * dz;
}

__global__ void __launch_bounds__(256) k_hex_blocks(int n, const double* __restrict__ coords,
                                                    const double* __restrict__ u, double nu,
                                                    double* __restrict__ diff, double* __restrict__ conv,
                                                    int* __restrict__ bad) {
  __shared__ double GP[4][27][14][3];
  __shared__ double PH[4][27][14];
  __shared__ double DW[4][27];
  __shared__ double UQ[4][27][3];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + wv;
  const bool live = c < n;
  if (live && lane < 27) {
    const int ix = lane / 9, iy = (lane / 3) % 3, iz = lane % 3;
    const double x = kG3x[ix], y = kG3x[iy], z = kG3x[iz], w = kG3w[ix] * kG3w[iy] * kG3w[iz];
    double phi[14], gr[14][3];
    hexBasis(x, y, z, phi, gr);
    const double* P = coords + (long long)c * 24;
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 8; ++a)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) J[3 * i + j] += P[3 * a + i] * gr[a][j];
    const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                       J[2] * (J[3] * J[7] - J[4] * J[6]);
    double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (det > 0.0 && isfinite(det)) {
      const double q = 1.0 / det;
      I[0] = (J[4] * J[8] - J[5] * J[7]) * q; I[1] = (J[2] * J[7] - J[1] * J[8]) * q; I[2] = (J[1] * J[5] - J[2] * J[4]) * q;
      I[3] = (J[5] * J[6] - J[3] * J[8]) * q; I[4] = (J[0] * J[8] - J[2] * J[6]) * q; I[5] = (J[2] * J[3] - J[0] * J[5]) * q;
      I[6] = (J[3] * J[7] - J[4] * J[6]) * q; I[7] = (J[1] * J[6] - J[0] * J[7]) * q; I[8] = (J[0] * J[4] - J[1] * J[3]) * q;
      DW[wv][lane] = det * w;
    } else {
      atomicAdd(bad, 1);
      DW[wv][lane] = 0.0;
    }
    const double* U = u + (long long)c * 42;
    double uq[3] = {0, 0, 0};
    for (int a = 0; a < 14; ++a) {
      PH[wv][lane][a] = phi[a];
      for (int d = 0; d < 3; ++d) {
        GP[wv][lane][a][d] = gr[a][0] * I[d] + gr[a][1] * I[3 + d] + gr[a][2] * I[6 + d];
        uq[d] += U[3 * a + d] * phi[a];
      }
    }
    for (int d = 0; d < 3; ++d) UQ[wv][lane][d] = uq[d];
  }
  __syncthreads();
  if (!live) return;
  for (int p = lane; p < 196; p += 64) {
    const int a = p / 14, b = p % 14;
    double dab = 0.0, cab = 0.0;
    for (int q = 0; q < 27; ++q) {
      const double* ga = GP[wv][q][a];
      const double* gb = GP[wv][q][b];
      dab += DW[wv][q] * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
      cab += DW[wv][q] * PH[wv][q][a] * (UQ[wv][q][0] * gb[0] + UQ[wv][q][1] * gb[1] + UQ[wv][q][2] * gb[2]);
    }
    diff[(long long)c * 196 + p] = nu * dab;
    conv[(long long)c * 196 + p] = cab;
  }
}

}  // namespace

void hexElementBlocks(long long n, const double* coords, const double* u, double nu, double* diff, double* conv,
                      int* badOut) {
  require(n >= 0 && coords && u && diff && conv, "hex_element_blocks: null or bad arguments");
  require(n < (1ll << 31) / 196, "hex_element_blocks: too many elements for one call");
  if (n == 0) {
    if (badOut) *badOut = 0;
    return;
  }
  hipStream_t s = nullptr;
  NODALS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  DBuf<double> dc, du, dd, dv;
  DBuf<int> db;
  try {
    dc.alloc((std::size_t)n * 24);
    du.alloc((std::size_t)n * 42);
    dd.alloc((std::size_t)n * 196);
    dv.alloc((std::size_t)n * 196);
    db.alloc(1);
    NODALS_HIP(hipMemcpyAsync(dc.data(), coords, sizeof(double) * (std::size_t)n * 24, hipMemcpyHostToDevice, s));
    NODALS_HIP(hipMemcpyAsync(du.data(), u, sizeof(double) * (std::size_t)n * 42, hipMemcpyHostToDevice, s));
    NODALS_HIP(hipMemsetAsync(db.data(), 0, sizeof(int), s));
    hipLaunchKernelGGL(k_hex_blocks, (int)((n + 3) / 4), 256, 0, s, (int)n, dc.data(), du.data(), nu, dd.data(),
                       dv.data(), db.data());
    NODALS_LAUNCHED();
    NODALS_HIP(hipMemcpyAsync(diff, dd.data(), sizeof(double) * (std::size_t)n * 196, hipMemcpyDeviceToHost, s));
    NODALS_HIP(hipMemcpyAsync(conv, dv.data(), sizeof(double) * (std::size_t)n * 196, hipMemcpyDeviceToHost, s));
    int bad = 0;
    NODALS_HIP(hipMemcpyAsync(&bad, db.data(), sizeof(int), hipMemcpyDeviceToHost, s));
    NODALS_HIP(hipStreamSynchronize(s));
    if (badOut) *badOut = bad;
  } catch (...) {
    (void)hipStreamDestroy(s);
    throw;
  }
  NODALS_HIP(hipStreamDestroy(s));
}

}  // namespace dev
}  // namespace nodals
