// Sparse and dense linear-algebra kernels for gfx950 (64-lane wavefronts).
//
// Row-parallel kernels map one row to a group of LANES consecutive lanes of
// a wavefront (LANES in {4,8,16,32,64}), chosen from the mean row length so a
// 64-lane wavefront keeps its loads coalesced on short rows without idling on
// long ones.  Group reductions use __shfl_xor inside the group, so every row's
// sum has a fixed order: results are deterministic run to run.  Global FP64
// reductions are two-stage (fixed grid of partial sums, then one block per
// output), also deterministic.  Nothing here uses atomics.
#pragma once

#include "common.hpp"

namespace nodals {
namespace dev {

constexpr int kReduceParts = 256;  // partial sums per reduction (grid size)

template <class T>
__device__ __forceinline__ T groupSum(T v, int lanes) {
  for (int o = lanes >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, lanes);
  return v;
}

// A lane's share of a CSR row product: s = sum of val(k) * gather(col[k])
// over k = k0, k0 + LANES, ... < e, accumulated in exactly that order (the
// order of the one-at-a-time loop, so results are bit-identical to it).  Each
// step first issues U column loads, U value loads and U gathers and only then
// accumulates, so a lane keeps U independent memory chains in flight instead
// of one (measured on gfx950: the 40k-cell P0 Schur product, 65 nnz/row,
// 8.7 -> 6.9 us; the 64^3 P1 Laplacian, 15 nnz/row, 17.2 -> 12.2 us at four
// lanes).  Slots past the row end are predicated off and never touch s.
// Entries in column `skip` are left out (-1: none).
template <int LANES>
constexpr int rowUnroll() {
  return LANES == 4 ? 4 : 2;
}
template <int LANES, class TS, class Val, class Gather>
__device__ __forceinline__ TS rowDot(int k0, int e, const int* __restrict__ col, int skip, Val val, Gather gather) {
  constexpr int U = rowUnroll<LANES>();
  TS s = 0;
  for (int k = k0; k < e; k += U * LANES) {
    int c[U];
    TS a[U], v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kk = k + u * LANES;
      c[u] = kk < e ? col[kk] : -1;
      a[u] = kk < e ? val(kk) : TS(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = c[u] >= 0 ? gather(c[u]) : TS(0);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (c[u] >= 0 && c[u] != skip) s += a[u] * v[u];
  }
  return s;
}

// Row-block index for XCD locality.  Blocks go to the eight XCDs round
// robin, and each XCD has its own L2; with the grid padded to a multiple
// of eight, XCD x takes virtual blocks [x*g/8, (x+1)*g/8), one contiguous
// eighth of the rows, so a vector written by one such kernel is read back
// by the next from the same L2.  A bijection on [0, gridDim.x): blocks past
// the rows exit in the caller's bounds check, whatever the hardware's
// placement.
__device__ __forceinline__ int xcdRowBlock() {
  return (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
}
inline int xcdGrid(int grid) { return (grid + 7) & ~7; }

// ---------------------------------------------------------------- BLAS-1
template <class T>
__global__ void k_fill(long long n, T* x, T a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = a;
}
template <class TD, class TS>
__global__ void k_cast(long long n, const TS* __restrict__ a, TD* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (TD)a[i];
}
template <class T>
__global__ void k_axpy(long long n, double a, const T* __restrict__ x, T* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += (T)a * x[i];
}
template <class T>
__global__ void k_scale_into(long long n, double a, const T* __restrict__ x, T* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (T)a * x[i];
}
template <class T>
__global__ void k_xpby(long long n, const T* __restrict__ x, double b, T* __restrict__ y) {
  // y = x + b*y
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + (T)b * y[i];
}
template <class T>
__global__ void k_axpy2(long long n, double a, const T* __restrict__ p, const T* __restrict__ q, T* __restrict__ x,
                        T* __restrict__ r) {
  // x += a p ; r -= a q
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const T aa = (T)a;
    x[i] += aa * p[i];
    r[i] -= aa * q[i];
  }
}
template <class T>
__global__ void k_jacobi_zero(int n, double omega, const T* __restrict__ b, const T* __restrict__ d, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (T)omega * b[i] / d[i];
}
template <class T>
__global__ void k_jacobi_add(int n, double omega, const T* __restrict__ r, const T* __restrict__ d, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (T)omega * r[i] / d[i];
}

// Multi-vector update y -= sum_k c_k V_k (V column-major, stride n).
struct Coef32 {
  double c[32];
};
static __global__ void k_multi_axpy(long long n, int k, const double* __restrict__ V, Coef32 c, double* __restrict__ y,
                             double sign) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += c.c[j] * V[(long long)j * n + i];
  y[i] += sign * s;
}
// Same with the coefficients read from device memory (k <= 32), so the
// update can follow a device-side reduction without a host round trip.
static __global__ void __launch_bounds__(256) k_multi_axpy_dev(long long n, int k, const double* __restrict__ V,
                                                                const double* __restrict__ coef,
                                                                double* __restrict__ y, double sign) {
  __shared__ double c[32];
  if (threadIdx.x < 32) c[threadIdx.x] = (int)threadIdx.x < k ? coef[threadIdx.x] : 0.0;
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += c[j] * V[(long long)j * n + i];
  y[i] += sign * s;
}
// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double blockSum256(double v, double* sh) {
  v = groupSum(v, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) t = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return t;
}

template <class TA, class TB>
__global__ void __launch_bounds__(256) k_dot_part(long long n, const TA* __restrict__ x, const TB* __restrict__ y,
                                                  double* __restrict__ part) {
  __shared__ double sh[4];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    s += (double)x[i] * (double)y[i];
  s = blockSum256(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// K dot products against one vector: part[k*gridDim + block] = sum V_k . w;
// with NORM, entry k holds w . w (one pass over w for Gram-Schmidt + norm).
// Each wavefront reduces all its sums with cross-lane shuffles, then one
// barrier and one LDS pass combine the four wavefronts.
template <int KMAX, bool NORM>
__global__ void __launch_bounds__(256) k_multidot_part(long long n, int k, const double* __restrict__ V,
                                                       const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[4][KMAX + 1];
  double s[KMAX + 1];
#pragma unroll
  for (int j = 0; j <= KMAX; ++j) s[j] = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double wi = w[i];
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
      if (j < k) s[j] += V[(long long)j * n + i] * wi;
    if (NORM) s[KMAX] += wi * wi;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j <= KMAX; ++j) {
    if (j < k || (NORM && j == KMAX)) {
      const double t = groupSum(s[j], 64);
      if (lane == 0) sh[wv][j] = t;
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < k) part[(long long)t * gridDim.x + blockIdx.x] = sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t];
  if (NORM && t == k)
    part[(long long)k * gridDim.x + blockIdx.x] = sh[0][KMAX] + sh[1][KMAX] + sh[2][KMAX] + sh[3][KMAX];
}

// out[b] = sum of part[b*nparts .. b*nparts+nparts)
static __global__ void __launch_bounds__(256) k_reduce_final(int nparts, const double* __restrict__ part, double* out) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) s += part[(long long)blockIdx.x * nparts + i];
  s = blockSum256(s, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ---------------------------------------------------------------- CSR kernels
// y = A x (RESID=false) or y = b - A x (RESID=true); ACC=true adds into y.
template <int LANES, class TM, class TV, bool RESID, bool ACC>
__global__ void __launch_bounds__(256) k_spmv(int n, const int* __restrict__ row, const int* __restrict__ col,
                                              const TM* __restrict__ val, const TV* __restrict__ x,
                                              const TV* __restrict__ b, TV* __restrict__ y) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int r = gid / LANES;
  const int lane = gid % LANES;
  if (r >= n) return;
  TV s = rowDot<LANES, TV>(
      row[r] + lane, row[r + 1], col, -1, [&](int k) { return ld<TV>(val, k); }, [&](int j) { return x[j]; });
  s = groupSum(s, LANES);
  if (lane == 0) {
    if (RESID) y[r] = b[r] - s;
    else if (ACC) y[r] += s;
    else y[r] = s;
  }
}

// y = A x of the P0 pressure Schur operator, the Krylov operator of the
// pressure solve (the body of k_spmv<LANES, double, double, false, false>
// under its own name, so profiles and launch probes see it apart from the
// AMG levels' products).
template <int LANES>
__global__ void __launch_bounds__(256) k_schur_spmv(int n, const int* __restrict__ row, const int* __restrict__ col,
                                                    const double* __restrict__ val, const double* __restrict__ x,
                                                    double* __restrict__ y) {
  const int gid = xcdRowBlock() * 256 + threadIdx.x;
  const int r = gid / LANES;
  const int lane = gid % LANES;
  if (r >= n) return;
  double s = rowDot<LANES, double>(
      row[r] + lane, row[r + 1], col, -1, [&](int k) { return val[k]; }, [&](int j) { return x[j]; });
  s = groupSum(s, LANES);
  if (lane == 0) y[r] = s;
}

// Single-sweep Jacobi smoothing fused into the V-cycle residuals.
// Pre-smoothing from zero and its residual in one pass:
//   t = w b / d,   r = b - A t
// (t_j of every neighbour is formed from b_j and d_j as the separate
// smoother would, so the values equal the unfused sequence).
// b may be stored wider than TV (an FP64 Krylov vector entering an FP32
// hierarchy): it is rounded to TV on load, exactly as a separate cast pass
// would round it.
template <int LANES, class TM, class TV, class TB = TV>
__global__ void __launch_bounds__(256) k_jacobi_pre_resid(int n, const int* __restrict__ row,
                                                          const int* __restrict__ col, const TM* __restrict__ val,
                                                          const TB* __restrict__ b, const TV* __restrict__ d,
                                                          double omega, TV* __restrict__ t, TV* __restrict__ r) {
  const int gid = xcdRowBlock() * 256 + threadIdx.x;
  const int i = gid / LANES;
  const int lane = gid % LANES;
  if (i >= n) return;
  TV s = rowDot<LANES, TV>(
      row[i] + lane, row[i + 1], col, -1, [&](int k) { return ld<TV>(val, k); },
      [&](int j) { return (TV)omega * (TV)b[j] / d[j]; });
  s = groupSum(s, LANES);
  if (lane == 0) {
    const TV bi = (TV)b[i];
    t[i] = (TV)omega * bi / d[i];
    r[i] = bi - s;
  }
}
// Residual and post-smoothing correction in one pass:
//   z = t + w (b - A t) / d      (z and t distinct)
// (b as above; z may be stored wider than TV and receives the TV result.)
template <int LANES, class TM, class TV, class TB = TV, class TZ = TV>
__global__ void __launch_bounds__(256) k_resid_jacobi_post(int n, const int* __restrict__ row,
                                                           const int* __restrict__ col, const TM* __restrict__ val,
                                                           const TB* __restrict__ b, const TV* __restrict__ d,
                                                           double omega, const TV* __restrict__ t,
                                                           TZ* __restrict__ z) {
  const int gid = xcdRowBlock() * 256 + threadIdx.x;
  const int i = gid / LANES;
  const int lane = gid % LANES;
  if (i >= n) return;
  TV s = rowDot<LANES, TV>(
      row[i] + lane, row[i + 1], col, -1, [&](int k) { return ld<TV>(val, k); }, [&](int j) { return t[j]; });
  s = groupSum(s, LANES);
  if (lane == 0) {
    const TV r = (TV)b[i] - s;
    z[i] = (TZ)(t[i] + (TV)omega * r / d[i]);
  }
}

// Product rounded on its own (never fused into a following add).
__device__ __forceinline__ float mulRn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mulRn(double a, double b) { return __dmul_rn(a, b); }

// The same with the coarse correction of an aggregation transfer (one
// entry per fine row: pcol[j], pval[j]) folded in: t'_j = t_j + pval_j
// xc[pcol_j] is formed on the fly for every row the pass reads, with the
// product rounded before the add, which is the value the separate
// prolongation kernel would have stored (its row sum is that single
// product).
template <int LANES, class TM, class TV, class TB = TV, class TZ = TV>
__global__ void __launch_bounds__(256) k_resid_jacobi_post_agg(
    int n, const int* __restrict__ row, const int* __restrict__ col, const TM* __restrict__ val,
    const TB* __restrict__ b, const TV* __restrict__ d, double omega, const TV* __restrict__ t,
    const int* __restrict__ pcol, const TV* __restrict__ pval, const TV* __restrict__ xc, TZ* __restrict__ z) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int i = gid / LANES;
  const int lane = gid % LANES;
  if (i >= n) return;
  TV s = rowDot<LANES, TV>(
      row[i] + lane, row[i + 1], col, -1, [&](int k) { return ld<TV>(val, k); },
      [&](int j) { return t[j] + mulRn(pval[j], xc[pcol[j]]); });
  s = groupSum(s, LANES);
  if (lane == 0) {
    const TV ti = t[i] + mulRn(pval[i], xc[pcol[i]]);
    const TV r = (TV)b[i] - s;
    z[i] = (TZ)(ti + (TV)omega * r / d[i]);
  }
}

// l1[r] = sum_k |a_rk|
template <int LANES, class TM, class TV>
__global__ void __launch_bounds__(256) k_row_l1(int n, const int* __restrict__ row, const TM* __restrict__ val,
                                                TV* __restrict__ l1) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int r = gid / LANES;
  const int lane = gid % LANES;
  if (r >= n) return;
  TV s = 0;
  for (int k = row[r] + lane; k < row[r + 1]; k += LANES) {
    const TV v = ld<TV>(val, k);
    s += v >= TV(0) ? v : -v;
  }
  s = groupSum(s, LANES);
  if (lane == 0) l1[r] = s;
}

// One multicolour Gauss-Seidel colour: rows crow[begin..begin+count), or
// rows begin..begin+count when crow is null (operator numbered in colour
// order).
// x_i <- x_i + omega ((b_i - sum_{j!=i} a_ij x_j)/d_i - x_i)
template <int LANES, class TM, class TV>
__global__ void __launch_bounds__(256) k_mcgs_color(int begin, int count, const int* __restrict__ crow,
                                                    const int* __restrict__ row, const int* __restrict__ col,
                                                    const TM* __restrict__ val, const TV* __restrict__ diag,
                                                    const TV* __restrict__ b, TV* __restrict__ x, double omega) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int q = gid / LANES;
  const int lane = gid % LANES;
  if (q >= count) return;
  const int i = crow ? crow[begin + q] : begin + q;  // no list: the colour is a contiguous row range
  // the row's own b, diagonal and value do not depend on the row product:
  // their loads are issued with the row's, not after the reduction
  const TV bi = b[i], di = diag[i], old = x[i];
  TV s = rowDot<LANES, TV>(
      row[i] + lane, row[i + 1], col, i, [&](int k) { return ld<TV>(val, k); }, [&](int j) { return x[j]; });
  s = groupSum(s, LANES);
  if (lane == 0) {
    const TV gs = (bi - s) / di;
    x[i] = old + (TV)omega * (gs - old);
  }
}

// A whole MCGS smoothing call of one small level in one workgroup, from x =
// 0: `sweeps` sweeps of forward (order 0), backward (1) or forward then
// backward (2) colour passes over the level numbered in colour order (colour
// c is rows [coff[c], coff[c+1])), with a barrier between colours instead of
// a launch per colour.  The iterate, the row pointers, the right-hand side
// and the diagonal are staged in LDS, so a row's chain is one global load
// (its columns and values) between LDS reads.  Colour c's rows use
// clanes[c] lanes each, the lanes of k_mcgs_color's launch for that colour,
// and the same rowDot/groupSum, so the result is bit-identical to the
// launch-per-colour passes.  Dynamic LDS: n (3 sizeof(TV) + 4) + 4 bytes.
template <int L, class TM, class TV>
__device__ __forceinline__ void mcgsRowsLds(int begin, int count, const int* rs, const int* __restrict__ col,
                                            const TM* __restrict__ val, const TV* ds, const TV* bs, TV* xs,
                                            double omega) {
  const int lane = threadIdx.x % L;
  for (int q = threadIdx.x / L; q < count; q += blockDim.x / L) {
    const int i = begin + q;
    const TV bi = bs[i], di = ds[i], old = xs[i];
    TV s = rowDot<L, TV>(
        rs[i] + lane, rs[i + 1], col, i, [&](int k) { return ld<TV>(val, k); }, [&](int j) { return xs[j]; });
    s = groupSum(s, L);
    if (lane == 0) {
      const TV gs = (bi - s) / di;
      xs[i] = old + (TV)omega * (gs - old);
    }
  }
}

template <class TM, class TV>
__global__ void __launch_bounds__(1024) k_mcgs_level_wg(int n, const int* __restrict__ row, const int* __restrict__ col,
                                                        const TM* __restrict__ val, const TV* __restrict__ diag,
                                                        const TV* __restrict__ b, TV* __restrict__ x, int ncolors,
                                                        const int* __restrict__ coff, const int* __restrict__ clanes,
                                                        int sweeps, int order, double omega) {
  extern __shared__ unsigned char ldsRaw[];
  TV* xs = reinterpret_cast<TV*>(ldsRaw);
  TV* bs = xs + n;
  TV* ds = bs + n;
  int* rs = reinterpret_cast<int*>(ds + n);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    xs[i] = TV(0);
    bs[i] = b[i];
    ds[i] = diag[i];
  }
  for (int i = threadIdx.x; i <= n; i += blockDim.x) rs[i] = row[i];
  __syncthreads();
  for (int sw = 0; sw < sweeps; ++sw)
    for (int pass = 0; pass < 2; ++pass) {
      if ((pass == 0 && order == 1) || (pass == 1 && order == 0)) continue;
      for (int q = 0; q < ncolors; ++q) {
        const int c = pass == 1 ? ncolors - 1 - q : q;
        const int begin = coff[c], count = coff[c + 1] - begin;
        switch (clanes[c]) {
          case 4: mcgsRowsLds<4, TM, TV>(begin, count, rs, col, val, ds, bs, xs, omega); break;
          case 8: mcgsRowsLds<8, TM, TV>(begin, count, rs, col, val, ds, bs, xs, omega); break;
          case 16: mcgsRowsLds<16, TM, TV>(begin, count, rs, col, val, ds, bs, xs, omega); break;
          case 32: mcgsRowsLds<32, TM, TV>(begin, count, rs, col, val, ds, bs, xs, omega); break;
          default: mcgsRowsLds<64, TM, TV>(begin, count, rs, col, val, ds, bs, xs, omega); break;
        }
        __syncthreads();
      }
    }
  for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = xs[i];
}

// x = T b for a dense row-major n x n matrix (terminal AMG level).
template <class TV>
__global__ void __launch_bounds__(256) k_dense_mv(int n, const TV* __restrict__ T, const TV* __restrict__ b,
                                                  TV* __restrict__ x) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int r = gid >> 6;
  const int lane = gid & 63;
  if (r >= n) return;
  TV s = 0;
  const TV* tr = T + (long long)r * n;
  for (int j = lane; j < n; j += 64) s += tr[j] * b[j];
  s = groupSum(s, 64);
  if (lane == 0) x[r] = s;
}

template <class T>
__global__ void k_diag_gather(int n, const int* __restrict__ diagPos, const T* __restrict__ val, T* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = val[diagPos[i]];
}

// ---------------------------------------------------------------- launch helpers
// Lanes per CSR row from the mean row length and the number of rows.  Four
// lanes (four loads in flight each) for rows up to 20 nonzeros, eight (two
// each) up to 128 (measured on gfx950: the 64^3 P1 Laplacian, 15 nnz/row,
// runs in 12.2 us at four lanes against 14.3 at eight; the 40k-cell P0 Schur
// operator, 65 nnz/row, in 6.9 us at eight against 9.7 at four and 7.5 at
// sixteen), wider groups for longer rows.  When the rows
// are too few to give the chip 2^17 threads (two waves per SIMD), the group
// widens further while every lane keeps at least two nonzeros: a 1,073-row
// coarse level with 88 nnz/row runs in 2.7 us at 32 lanes and 4.7 us at 8.
inline int lanesFor(double meanRow, long long rows) {
  int L = meanRow <= 20 ? 4 : meanRow <= 128 ? 8 : meanRow <= 256 ? 16 : meanRow <= 512 ? 32 : 64;
  while (L < 64 && rows * L < (1LL << 17) && 4.0 * L <= meanRow) L *= 2;
  return L;
}

template <class TM, class TV, bool RESID, bool ACC>
void spmv(hipStream_t s, int lanes, int n, const int* row, const int* col, const TM* val, const TV* x, const TV* b,
          TV* y, hipEvent_t start = nullptr, hipEvent_t stop = nullptr) {
  if (n <= 0) return;
  const int grid = blocksFor((long long)n * lanes);
  // With events, the launch goes through hipExtLaunchKernelGGL, which stamps
  // them from the kernel's own dispatch packet (its start and end), not from
  // separate marker packets.
#define NODALS_SPMV_LAUNCH(LN)                                                                             \
  if (start)                                                                                               \
    hipExtLaunchKernelGGL((k_spmv<LN, TM, TV, RESID, ACC>), dim3(grid), dim3(256), 0, s, start, stop, 0, n, \
                          row, col, val, x, b, y);                                                         \
  else                                                                                                     \
    hipLaunchKernelGGL((k_spmv<LN, TM, TV, RESID, ACC>), grid, 256, 0, s, n, row, col, val, x, b, y);
  switch (lanes) {
    case 4: NODALS_SPMV_LAUNCH(4) break;
    case 8: NODALS_SPMV_LAUNCH(8) break;
    case 16: NODALS_SPMV_LAUNCH(16) break;
    case 32: NODALS_SPMV_LAUNCH(32) break;
    default: NODALS_SPMV_LAUNCH(64) break;
  }
#undef NODALS_SPMV_LAUNCH
  NODALS_LAUNCHED();
}

inline void schurSpmv(hipStream_t s, int lanes, int n, const int* row, const int* col, const double* val,
                      const double* x, double* y, hipEvent_t start = nullptr, hipEvent_t stop = nullptr) {
  if (n <= 0) return;
  const int grid = xcdGrid(blocksFor((long long)n * lanes));  // the kernel maps blocks per XCD
#define NODALS_SCHUR_LAUNCH(LN)                                                                                  \
  if (start)                                                                                                    \
    hipExtLaunchKernelGGL((k_schur_spmv<LN>), dim3(grid), dim3(256), 0, s, start, stop, 0, n, row, col, val, x, y); \
  else                                                                                                          \
    hipLaunchKernelGGL((k_schur_spmv<LN>), grid, 256, 0, s, n, row, col, val, x, y);
  switch (lanes) {
    case 4: NODALS_SCHUR_LAUNCH(4) break;
    case 8: NODALS_SCHUR_LAUNCH(8) break;
    case 16: NODALS_SCHUR_LAUNCH(16) break;
    case 32: NODALS_SCHUR_LAUNCH(32) break;
    default: NODALS_SCHUR_LAUNCH(64) break;
  }
#undef NODALS_SCHUR_LAUNCH
  NODALS_LAUNCHED();
}

template <class TM, class TV, class TB = TV>
void jacobiPreResid(hipStream_t s, int lanes, int n, const int* row, const int* col, const TM* val, const TB* b,
                    const TV* d, double omega, TV* t, TV* r) {
  if (n <= 0) return;
  const int grid = xcdGrid(blocksFor((long long)n * lanes));  // the kernel maps blocks per XCD
#define NODALS_JPR(LN) \
  hipLaunchKernelGGL((k_jacobi_pre_resid<LN, TM, TV, TB>), grid, 256, 0, s, n, row, col, val, b, d, omega, t, r)
  switch (lanes) {
    case 4: NODALS_JPR(4); break;
    case 8: NODALS_JPR(8); break;
    case 16: NODALS_JPR(16); break;
    case 32: NODALS_JPR(32); break;
    default: NODALS_JPR(64); break;
  }
#undef NODALS_JPR
  NODALS_LAUNCHED();
}

template <class TM, class TV, class TB = TV, class TZ = TV>
void residJacobiPostAgg(hipStream_t s, int lanes, int n, const int* row, const int* col, const TM* val, const TB* b,
                        const TV* d, double omega, const TV* t, const int* pcol, const TV* pval, const TV* xc,
                        TZ* z) {
  if (n <= 0) return;
  const int grid = blocksFor((long long)n * lanes);
#define NODALS_RJPA(LN)                                                                                        \
  hipLaunchKernelGGL((k_resid_jacobi_post_agg<LN, TM, TV, TB, TZ>), grid, 256, 0, s, n, row, col, val, b, d, omega, \
                     t, pcol, pval, xc, z)
  switch (lanes) {
    case 4: NODALS_RJPA(4); break;
    case 8: NODALS_RJPA(8); break;
    case 16: NODALS_RJPA(16); break;
    case 32: NODALS_RJPA(32); break;
    default: NODALS_RJPA(64); break;
  }
#undef NODALS_RJPA
  NODALS_LAUNCHED();
}

template <class TM, class TV, class TB = TV, class TZ = TV>
void residJacobiPost(hipStream_t s, int lanes, int n, const int* row, const int* col, const TM* val, const TB* b,
                     const TV* d, double omega, const TV* t, TZ* z) {
  if (n <= 0) return;
  const int grid = xcdGrid(blocksFor((long long)n * lanes));  // the kernel maps blocks per XCD
#define NODALS_RJP(LN) \
  hipLaunchKernelGGL((k_resid_jacobi_post<LN, TM, TV, TB, TZ>), grid, 256, 0, s, n, row, col, val, b, d, omega, t, z)
  switch (lanes) {
    case 4: NODALS_RJP(4); break;
    case 8: NODALS_RJP(8); break;
    case 16: NODALS_RJP(16); break;
    case 32: NODALS_RJP(32); break;
    default: NODALS_RJP(64); break;
  }
#undef NODALS_RJP
  NODALS_LAUNCHED();
}

template <class TM, class TV>
void rowL1(hipStream_t s, int lanes, int n, const int* row, const TM* val, TV* l1) {
  if (n <= 0) return;
  const int grid = blocksFor((long long)n * lanes);
  switch (lanes) {
    case 4: hipLaunchKernelGGL((k_row_l1<4, TM, TV>), grid, 256, 0, s, n, row, val, l1); break;
    case 8: hipLaunchKernelGGL((k_row_l1<8, TM, TV>), grid, 256, 0, s, n, row, val, l1); break;
    case 16: hipLaunchKernelGGL((k_row_l1<16, TM, TV>), grid, 256, 0, s, n, row, val, l1); break;
    case 32: hipLaunchKernelGGL((k_row_l1<32, TM, TV>), grid, 256, 0, s, n, row, val, l1); break;
    default: hipLaunchKernelGGL((k_row_l1<64, TM, TV>), grid, 256, 0, s, n, row, val, l1); break;
  }
  NODALS_LAUNCHED();
}

template <class TM, class TV>
void mcgsColor(hipStream_t s, int lanes, int begin, int count, const int* crow, const int* row, const int* col,
               const TM* val, const TV* diag, const TV* b, TV* x, double omega) {
  if (count <= 0) return;
  const int grid = blocksFor((long long)count * lanes);
  switch (lanes) {
    case 4: hipLaunchKernelGGL((k_mcgs_color<4, TM, TV>), grid, 256, 0, s, begin, count, crow, row, col, val, diag, b, x, omega); break;
    case 8: hipLaunchKernelGGL((k_mcgs_color<8, TM, TV>), grid, 256, 0, s, begin, count, crow, row, col, val, diag, b, x, omega); break;
    case 16: hipLaunchKernelGGL((k_mcgs_color<16, TM, TV>), grid, 256, 0, s, begin, count, crow, row, col, val, diag, b, x, omega); break;
    case 32: hipLaunchKernelGGL((k_mcgs_color<32, TM, TV>), grid, 256, 0, s, begin, count, crow, row, col, val, diag, b, x, omega); break;
    default: hipLaunchKernelGGL((k_mcgs_color<64, TM, TV>), grid, 256, 0, s, begin, count, crow, row, col, val, diag, b, x, omega); break;
  }
  NODALS_LAUNCHED();
}

template <class T>
void fill(hipStream_t s, long long n, T* x, T a) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<T>, blocksFor(n), 256, 0, s, n, x, a);
  NODALS_LAUNCHED();
}
template <class TD, class TS>
void cast(hipStream_t s, long long n, const TS* a, TD* b) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_cast<TD, TS>), blocksFor(n), 256, 0, s, n, a, b);
  NODALS_LAUNCHED();
}
template <class T>
void copy(hipStream_t s, long long n, const T* a, T* b) {
  if (n <= 0) return;
  NODALS_HIP(hipMemcpyAsync(b, a, (std::size_t)n * sizeof(T), hipMemcpyDeviceToDevice, s));
}
template <class T>
void axpy(hipStream_t s, long long n, double a, const T* x, T* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy<T>, blocksFor(n), 256, 0, s, n, a, x, y);
  NODALS_LAUNCHED();
}

// Deterministic FP64 dot products; results land in pinned host memory.
class Reducer {
 public:
  Reducer() = default;
  ~Reducer() {
    if (host_) (void)hipHostFree(host_);
  }
  Reducer(const Reducer&) = delete;
  Reducer& operator=(const Reducer&) = delete;

  void init(int maxVectors) {
    maxVec_ = maxVectors;
    part_.alloc((std::size_t)kReduceParts * maxVectors);
    part2_.alloc((std::size_t)kReduceParts * maxVectors);
    if (host_) (void)hipHostFree(host_);
    // Coherent pinned memory: the final reduction kernel writes it directly
    // and the host reads it after the stream synchronisation.
    NODALS_HIP(hipHostMalloc((void**)&host_, sizeof(double) * (std::size_t)maxVectors, hipHostMallocCoherent));
  }

  template <class TA, class TB>
  void dotAsync(hipStream_t s, long long n, const TA* x, const TB* y) {
    hipLaunchKernelGGL((k_dot_part<TA, TB>), kReduceParts, 256, 0, s, n, x, y, part_.data());
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_reduce_final, 1, 256, 0, s, kReduceParts, part_.data(), host_);
    NODALS_LAUNCHED();
  }
  template <class TA, class TB>
  double dot(hipStream_t s, long long n, const TA* x, const TB* y) {
    dotAsync(s, n, x, y);
    NODALS_HIP(hipStreamSynchronize(s));
    return host_[0];
  }
  // out[j] = V_j . w for j < k (V column-major with stride n); with
  // withNorm, out[k] = w . w from the same pass.
  void multidot(hipStream_t s, long long n, int k, const double* V, const double* w, double* out,
                bool withNorm = false) {
    const int outs = k + (withNorm ? 1 : 0);
    if (outs > maxVec_ || k > 32) fail("multidot width exceeds reducer capacity", kErrInternal);
    if (withNorm)
      hipLaunchKernelGGL((k_multidot_part<32, true>), kReduceParts, 256, 0, s, n, k, V, w, part_.data());
    else
      hipLaunchKernelGGL((k_multidot_part<32, false>), kReduceParts, 256, 0, s, n, k, V, w, part_.data());
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_reduce_final, outs, 256, 0, s, kReduceParts, part_.data(), host_);
    NODALS_LAUNCHED();
    NODALS_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < outs; ++j) out[j] = host_[j];
  }
  double* partials() { return part_.data(); }
  // a second partials area, for a kernel that reads partials() and writes
  // its own in the same launch
  double* partials2() { return part2_.data(); }
  // Partial sums only (kReduceParts per output, output q at q*kReduceParts),
  // for a consumer kernel that finishes the reduction itself.
  template <class TA, class TB>
  void dotPartials(hipStream_t s, long long n, const TA* x, const TB* y) {
    hipLaunchKernelGGL((k_dot_part<TA, TB>), kReduceParts, 256, 0, s, n, x, y, part_.data());
    NODALS_LAUNCHED();
  }
  void multidotPartials(hipStream_t s, long long n, int k, const double* V, const double* w) {
    if (k > maxVec_ || k > 32) fail("multidot width exceeds reducer capacity", kErrInternal);
    hipLaunchKernelGGL((k_multidot_part<32, false>), kReduceParts, 256, 0, s, n, k, V, w, part_.data());
    NODALS_LAUNCHED();
  }
  double* hostOut() { return host_; }
  const double* hostOut() const { return host_; }

 private:
  DBuf<double> part_, part2_;
  double* host_ = nullptr;
  int maxVec_ = 0;
};

}  // namespace dev
}  // namespace nodals
