// H8 SIMPLE hot-path kernels for gfx950.
//
// Free velocity DOFs are renumbered on the device in colour order, so every
// Gauss-Seidel colour is a contiguous row range and rows, right-hand sides
// and diagonals stream coalesced.  Velocities and momentum right-hand sides
// are stored xyz-interleaved so one gather fetches all three components.
// Assembly is atomic-free and deterministic: cell-local 8x8 convection blocks
// are written once per cell, and each momentum row gathers its terms in the
// fixed (cell, basis) order recorded at setup.
#pragma once

#include "la_kernels.hpp"

namespace nodals {
namespace dev {

__device__ __forceinline__ double readlaneD(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Cell-local Picard central convection block (reference kernel
// h4_assemble_convection_only_kernel, gpu/serial_cuda/h4_kernels.cuh:10-56):
//   ur[m][j] = sum_d u_d(m) invJ[j][d]
//   C[a][b]  = det * sum_{m,j} ur[m][j] T[a][m][b][j]
// One wavefront per cell, lane = 8a+b.  Each lane keeps its 24 tensor
// entries in registers for the whole grid-stride loop; the 24 ur values are
// broadcast with v_readlane, so the inner product runs on scalar operands.
static __global__ void __launch_bounds__(256) k_conv_local(int nc, const int* __restrict__ cref,
                                                    const double* __restrict__ cinvJ,
                                                    const double* __restrict__ cdet, const double* __restrict__ u,
                                                    const double* __restrict__ fixedVal,
                                                    const double* __restrict__ centT, double* __restrict__ cbuf) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int nwaves = gridDim.x * 4;
  const int a = lane >> 3, b = lane & 7;
  double T[24];
#pragma unroll
  for (int m = 0; m < 8; ++m)
#pragma unroll
    for (int j = 0; j < 3; ++j) T[m * 3 + j] = centT[((a * 8 + m) * 8 + b) * 3 + j];
  const int lm = lane < 24 ? lane / 3 : 0;
  const int lj = lane < 24 ? lane % 3 : 0;
  for (int c = wave; c < nc; c += nwaves) {
    double cf = 0.0;
    if (lane < 24) {
      const int r = cref[(long long)c * 8 + lm];
      cf = r >= 0 ? u[3ll * r + lj] : fixedVal[3ll * (-r - 1) + lj];
    }
    // lane 3m+j holds cf[d=j][m]; ur for lane (m,j)
    const double c0 = __shfl(cf, 3 * lm + 0);
    const double c1 = __shfl(cf, 3 * lm + 1);
    const double c2 = __shfl(cf, 3 * lm + 2);
    const double* I = cinvJ + (long long)c * 9 + 3 * lj;
    const double ur = lane < 24 ? (c0 * I[0] + c1 * I[1] + c2 * I[2]) : 0.0;
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 24; ++q) v += readlaneD(ur, q) * T[q];
    cbuf[(long long)c * 64 + lane] = cdet[c] * v;
  }
}

// Momentum rows come in two length classes: vertex rows (about 75-87
// nonzeros) and face-bubble rows (about 12).  One launch covers both: blocks
// [0, nbLong) give each long row LL lanes, the remaining blocks give each
// short row LS lanes, so neither class runs a long serial loop nor leaves
// most lanes idle.  The branch is uniform per block.
#define NODALS_TWO_CLASS(LL, LS, nLong, nShort, nbLong, BODY)            \
  if ((int)blockIdx.x < (nbLong)) {                                       \
    const int gid_ = blockIdx.x * 256 + threadIdx.x;                      \
    const int q_ = gid_ / (LL);                                           \
    if (q_ < (nLong)) BODY(LL, q_, gid_ % (LL), true);                   \
  } else {                                                                \
    const int gid_ = ((int)blockIdx.x - (nbLong)) * 256 + threadIdx.x;    \
    const int q_ = gid_ / (LS);                                           \
    if (q_ < (nShort)) BODY(LS, q_, gid_ % (LS), false);                 \
  }

// Per momentum row: gather the assembled matrix (persistent diffusion +
// convection terms), row-L1 under-relaxation delta = (1/alpha-1) sum|a_ik|
// added to the diagonal, rAU = 1/a_ii (g4_finalize_relax_kernel,
// g5e_fp32_kernels.cuh:78-85), the Dirichlet convection RHS, B^T p by the
// velocity incidence, and the momentum RHS b = s + c + B^T p + delta u
// (g4_momentum_rhs_kernel, g5e_fp32_kernels.cuh:97-104).
struct MomRowsArgs {
  const int *mrow, *mdiagPos;
  const double* mdiff;
  const int *termOff, *termIdx;
  const double* cbuf;
  const int *fixOff, *fixTerm;
  const double* fixedVal;
  const int* incOff;
  const unsigned* incPacked;
  const double *bcoef, *p, *srhs, *u;
  double fac;
  double *aval, *delta, *diag, *rau, *brhs;
};

template <int L>
__device__ __forceinline__ void mom_row(const MomRowsArgs& A, int i, int lane) {
  const int rb = A.mrow[i], re = A.mrow[i + 1], dp = A.mdiagPos[i];
  double sabs = 0.0, vdiag = 0.0;
  for (int k = rb + lane; k < re; k += L) {
    double v = A.mdiff[k];
    // the entry's convection terms four at a time: the index loads, then the
    // block loads, then the sums in term order (indices clamped in bounds,
    // terms past the end not accumulated)
    const int te = A.termOff[k + 1];
    for (int t0 = A.termOff[k]; t0 < te; t0 += 4) {
      int ix[4];
      double cv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) ix[q] = A.termIdx[min(t0 + q, te - 1)];
#pragma unroll
      for (int q = 0; q < 4; ++q) cv[q] = A.cbuf[ix[q]];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (t0 + q < te) v += cv[q];
    }
    A.aval[k] = v;
    sabs += fabs(v);
    if (k == dp) vdiag = v;
  }
  sabs = groupSum(sabs, L);
  vdiag = groupSum(vdiag, L);
  const double de = A.fac * sabs;
  const double dg = vdiag + de;
  double c0 = 0, c1 = 0, c2 = 0;
  for (int t = A.fixOff[i] + lane; t < A.fixOff[i + 1]; t += L) {
    const int q = A.fixTerm[2 * t], slot = A.fixTerm[2 * t + 1];
    const double C = A.cbuf[q];
    c0 -= C * A.fixedVal[3ll * slot + 0];
    c1 -= C * A.fixedVal[3ll * slot + 1];
    c2 -= C * A.fixedVal[3ll * slot + 2];
  }
  double t0 = 0, t1 = 0, t2 = 0;
  for (int k = A.incOff[i] + lane; k < A.incOff[i + 1]; k += L) {
    const unsigned pk = A.incPacked[k];
    const int c = (int)(pk >> 3), a = (int)(pk & 7u);
    const double pv = A.p[c];
    const double* B = A.bcoef + (long long)c * 24 + 3 * a;
    t0 += B[0] * pv;
    t1 += B[1] * pv;
    t2 += B[2] * pv;
  }
  c0 = groupSum(c0, L);
  c1 = groupSum(c1, L);
  c2 = groupSum(c2, L);
  t0 = groupSum(t0, L);
  t1 = groupSum(t1, L);
  t2 = groupSum(t2, L);
  if (dp >= rb + lane && ((dp - rb - lane) % L) == 0) A.aval[dp] = dg;
  if (lane == 0) {
    A.delta[i] = de;
    A.diag[i] = dg;
    A.rau[i] = 1.0 / dg;
    A.brhs[3ll * i + 0] = A.srhs[3ll * i + 0] + c0 + t0 + de * A.u[3ll * i + 0];
    A.brhs[3ll * i + 1] = A.srhs[3ll * i + 1] + c1 + t1 + de * A.u[3ll * i + 1];
    A.brhs[3ll * i + 2] = A.srhs[3ll * i + 2] + c2 + t2 + de * A.u[3ll * i + 2];
  }
}

// Rows by class lists (longRows[nLong], shortRows[nShort]).
template <int LL, int LS>
__global__ void __launch_bounds__(256) k_mom_rows(MomRowsArgs A, int nLong, const int* __restrict__ longRows,
                                                  int nShort, const int* __restrict__ shortRows, int nbLong) {
#define NODALS_MOM_ROW_BODY(LN, q, lane, isLong) mom_row<LN>(A, (isLong) ? longRows[q] : shortRows[q], lane)
  NODALS_TWO_CLASS(LL, LS, nLong, nShort, nbLong, NODALS_MOM_ROW_BODY)
#undef NODALS_MOM_ROW_BODY
}

// One colour of the three-component forward Gauss-Seidel sweep sharing A
// (g4_mcgs_color_kernel, g5e_fp32_kernels.cuh:106-116).  The colour's rows
// are [r0, r1), its long rows first: [r0, rs) long, [rs, r1) short.
template <int L>
__device__ __forceinline__ void mom_gs_row(int i, int lane, const int* __restrict__ mrow, const int* __restrict__ mcol,
                                           const double* __restrict__ aval, const double* __restrict__ diag,
                                           const double* __restrict__ brhs, double* __restrict__ u, double omega) {
  // The row's own diagonal, right-hand side and old value do not depend on
  // the gathers (no other row of the colour writes them), so their loads are
  // issued before the gather loop instead of after the reduction.
  double dg = 1, b0 = 0, b1 = 0, b2 = 0, o0 = 0, o1 = 0, o2 = 0;
  if (lane == 0) {
    dg = diag[i];
    b0 = brhs[3ll * i + 0];
    b1 = brhs[3ll * i + 1];
    b2 = brhs[3ll * i + 2];
    o0 = u[3ll * i + 0];
    o1 = u[3ll * i + 1];
    o2 = u[3ll * i + 2];
  }
  // Vertex rows (75-87 nonzeros on 32 lanes) take three trips: each lane
  // issues all its column, value and velocity loads before it accumulates,
  // in the one-at-a-time order, so the sums are bit-identical.  Positions
  // past the row end or on the diagonal load the row's own velocity (in
  // bounds) and are not accumulated.
  constexpr int U = L >= 32 ? 3 : 1;
  const int ke = mrow[i + 1];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int k0 = mrow[i] + lane; k0 < ke; k0 += U * L) {
    int jq[U];
    double aq[U], x0[U], x1[U], x2[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int k = k0 + q * L;
      jq[q] = k < ke ? mcol[k] : i;
      aq[q] = k < ke ? aval[k] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      x0[q] = u[3ll * jq[q] + 0];
      x1[q] = u[3ll * jq[q] + 1];
      x2[q] = u[3ll * jq[q] + 2];
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      if (k0 + q * L < ke && jq[q] != i) {
        s0 += aq[q] * x0[q];
        s1 += aq[q] * x1[q];
        s2 += aq[q] * x2[q];
      }
    }
  }
  s0 = groupSum(s0, L);
  s1 = groupSum(s1, L);
  s2 = groupSum(s2, L);
  if (lane == 0) {
    const double iv = 1.0 / dg;
    const double w1 = 1.0 - omega;
    u[3ll * i + 0] = w1 * o0 + omega * ((b0 - s0) * iv);
    u[3ll * i + 1] = w1 * o1 + omega * ((b1 - s1) * iv);
    u[3ll * i + 2] = w1 * o2 + omega * ((b2 - s2) * iv);
  }
}

template <int LL, int LS>
__global__ void __launch_bounds__(256) k_mom_gs(int r0, int rs, int r1, int nbLong, const int* __restrict__ mrow,
                                                const int* __restrict__ mcol, const double* __restrict__ aval,
                                                const double* __restrict__ diag, const double* __restrict__ brhs,
                                                double* __restrict__ u, double omega) {
#define NODALS_GS_BODY(LN, q, lane, isLong) \
  mom_gs_row<LN>((isLong) ? r0 + (q) : rs + (q), lane, mrow, mcol, aval, diag, brhs, u, omega)
  NODALS_TWO_CLASS(LL, LS, rs - r0, r1 - rs, nbLong, NODALS_GS_BODY)
#undef NODALS_GS_BODY
}

// Momentum residual audit partial sums: |b_d|^2 and |b_d - (A u)_d|^2
// (g4_mom_norms_kernel, g5e_fp32_kernels.cuh:117-127).  One row per lane
// group, rows by class lists; part[6][gridDim.x].
template <int L>
__device__ __forceinline__ void mom_norm_row(int i, int lane, const int* __restrict__ mrow,
                                             const int* __restrict__ mcol, const double* __restrict__ aval,
                                             const double* __restrict__ brhs, const double* __restrict__ u,
                                             double* acc) {
  // U trips' loads are issued before any accumulation, in the one-at-a-time
  // order (bit-identical sums); positions past the row end load the row's
  // own velocity and are not accumulated.  The right-hand side loads first.
  constexpr int U = L <= 4 ? 4 : 2;
  double b0 = 0, b1 = 0, b2 = 0;
  if (lane == 0) {
    b0 = brhs[3ll * i];
    b1 = brhs[3ll * i + 1];
    b2 = brhs[3ll * i + 2];
  }
  const int ke = mrow[i + 1];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int k0 = mrow[i] + lane; k0 < ke; k0 += U * L) {
    int jq[U];
    double aq[U], x0[U], x1[U], x2[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int k = k0 + q * L;
      jq[q] = k < ke ? mcol[k] : i;
      aq[q] = k < ke ? aval[k] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      x0[q] = u[3ll * jq[q] + 0];
      x1[q] = u[3ll * jq[q] + 1];
      x2[q] = u[3ll * jq[q] + 2];
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      if (k0 + q * L < ke) {
        s0 += aq[q] * x0[q];
        s1 += aq[q] * x1[q];
        s2 += aq[q] * x2[q];
      }
    }
  }
  s0 = groupSum(s0, L);
  s1 = groupSum(s1, L);
  s2 = groupSum(s2, L);
  if (lane == 0) {
    acc[0] += b0 * b0;
    acc[1] += b1 * b1;
    acc[2] += b2 * b2;
    acc[3] += (b0 - s0) * (b0 - s0);
    acc[4] += (b1 - s1) * (b1 - s1);
    acc[5] += (b2 - s2) * (b2 - s2);
  }
}

template <int LL, int LS>
__global__ void __launch_bounds__(256) k_mom_norms(int nLong, const int* __restrict__ longRows, int nShort,
                                                   const int* __restrict__ shortRows, int nbLong,
                                                   const int* __restrict__ mrow, const int* __restrict__ mcol,
                                                   const double* __restrict__ aval, const double* __restrict__ brhs,
                                                   const double* __restrict__ u, double* __restrict__ part) {
  __shared__ double sh[4][6];
  double acc[6] = {0, 0, 0, 0, 0, 0};
#define NODALS_NORM_BODY(LN, q, lane, isLong) \
  mom_norm_row<LN>((isLong) ? longRows[q] : shortRows[q], lane, mrow, mcol, aval, brhs, u, acc)
  NODALS_TWO_CLASS(LL, LS, nLong, nShort, nbLong, NODALS_NORM_BODY)
#undef NODALS_NORM_BODY
  // all six sums with cross-lane shuffles per wavefront, one barrier
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double t = groupSum(acc[q], 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][q] = t;
  }
  __syncthreads();
  if (threadIdx.x < 6)
    part[(long long)threadIdx.x * gridDim.x + blockIdx.x] =
        sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

// Continuity residual r_c = fixedDiv_c + sum_a B[c][a] . u(ref a)
// (h8_continuity_kernel, h8_precomputed_b.cuh:140-163), written negated as
// the pressure right-hand side, with block partials of r^2.  A pinned cell
// (pin >= 0, closed domains) gets a zero right-hand side; its residual still
// counts in the norm.
static __global__ void __launch_bounds__(256) k_continuity(int nc, const int* __restrict__ cref,
                                                    const double* __restrict__ bcoef,
                                                    const double* __restrict__ fixedDiv, const double* __restrict__ u,
                                                    int pin, double* __restrict__ negCont, double* __restrict__ part) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int c = blockIdx.x * 256 + threadIdx.x; c < nc; c += gridDim.x * 256) {
    double s = fixedDiv[c];
    const double* B = bcoef + (long long)c * 24;
    // all eight gathers are issued before the sum (a fixed DOF reads DOF 0,
    // in bounds, and is not accumulated); the sum keeps the basis order
    int gq[8];
    double xq[24];
#pragma unroll
    for (int a = 0; a < 8; ++a) gq[a] = cref[(long long)c * 8 + a];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const long long g = gq[a] < 0 ? 0 : gq[a];
      xq[3 * a] = u[3 * g];
      xq[3 * a + 1] = u[3 * g + 1];
      xq[3 * a + 2] = u[3 * g + 2];
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double t = B[3 * a] * xq[3 * a] + B[3 * a + 1] * xq[3 * a + 1] + B[3 * a + 2] * xq[3 * a + 2];
      if (gq[a] >= 0) s += t;
    }
    negCont[c] = c == pin ? 0.0 : -s;
    acc += s * s;
  }
  acc = blockSum256(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// W[c][a] = rAU_g B[c][a] for the free DOF g of cell c's basis a (zero for
// fixed ones): every term of the fine Schur refresh is then W[c][a].B[c2][b]
// with no per-term rAU gather.
static __global__ void k_fine_weights(int nc, const int* __restrict__ cref, const double* __restrict__ bcoef,
                                      const double* __restrict__ rau, double* __restrict__ W) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nc * 8) return;
  const int g = cref[t];
  const double r = g >= 0 ? rau[g] : 0.0;
  for (int d = 0; d < 3; ++d) W[3ll * t + d] = r * bcoef[3ll * t + d];
}

// Fine P0 Schur numeric refresh (L lanes per row, the grid looping over the
// rows), one nonzero per lane:
//   val[k] = sum_{shared free DOF g} rAU_g * (B[c][a] . B[c2][b])
//          = sum (W[c][a] . B[c2][b])
// (h8_fine_csr_refresh_warp_kernel, h8_precomputed_b.cuh:37-91, restated as
// a gather over the terms recorded at setup; no atomics).  Also writes the
// diagonal, optional lower-precision copies and the row-L1 sums.  With a
// pinned cell (pin >= 0) the off-diagonal entries of its row and column are
// zero.
template <int L, bool F32, bool B16, bool L1>
__global__ void __launch_bounds__(256) k_fine_refresh(
    int nc, const int* __restrict__ frow, const int* __restrict__ fcol, const int* __restrict__ fdiagPos,
    const int* __restrict__ nzOff, const unsigned* __restrict__ nzTerm, const double* __restrict__ bcoef,
    const double* __restrict__ W, int pin, double* __restrict__ val, float* __restrict__ valf,
    bf16* __restrict__ valb, double* __restrict__ diag, float* __restrict__ diagf, double* __restrict__ l1d,
    float* __restrict__ l1f) {
  const int lane = threadIdx.x % L;
  for (int c = (blockIdx.x * 256 + threadIdx.x) / L; c < nc; c += gridDim.x * (256 / L)) {
    const int rb = frow[c], re = frow[c + 1], dp = fdiagPos[c];
    const double* Wc = W + (long long)c * 24;
    double sabs = 0.0;
    for (int k = rb + lane; k < re; k += L) {
      const int c2 = fcol[k];
      const double* Bn = bcoef + (long long)c2 * 24;
      double v = 0.0;
      for (int t = nzOff[k]; t < nzOff[k + 1]; ++t) {
        const unsigned q = nzTerm[t];
        const int a = (int)((q >> 3) & 7u), b = (int)(q & 7u);
        v += Wc[3 * a] * Bn[3 * b] + Wc[3 * a + 1] * Bn[3 * b + 1] + Wc[3 * a + 2] * Bn[3 * b + 2];
      }
      if ((c == pin || c2 == pin) && c != c2) v = 0.0;
      val[k] = v;
      if (F32) valf[k] = (float)v;
      if (B16) valb[k] = floatToBf16((float)v);
      if (L1) sabs += fabs(v);
      if (k == dp) {
        diag[c] = v;
        if (F32 || B16) diagf[c] = (float)v;
      }
    }
    if (L1) {
      sabs = groupSum(sabs, L);
      if (lane == 0) {
        l1d[c] = sabs;
        if (F32 || B16) l1f[c] = (float)sabs;
      }
    }
  }
}

}  // namespace dev
}  // namespace nodals
