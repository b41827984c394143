// Kernels of the message-passing network that predicts the initial pressure
// correction: BF16 MFMA GEMMs with FP32 accumulation (forward, input
// gradient, weight gradient), neighbour-mean aggregation and its transpose,
// the decoder, the loss, and Adam.
//
// Matrices are row-major FP32 in HBM with the row count padded to a
// multiple of 32 (padding rows hold zeros in every gradient).  GEMM inputs
// are rounded to BF16 (round to nearest even) as they are loaded into the
// matrix cores; weights are kept as BF16 copies of the FP32 master values.
//
// v_mfma_f32_32x32x16_bf16 (gfx950): lane l supplies row l%32 of the 32x16
// A tile and column l%32 of the 16x32 B tile, each with the eight k values
// 8*(l/32) .. 8*(l/32)+7; accumulator register r of lane l holds output row
// 8*(r/4) + 4*(l/32) + r%4, column l%32.
#pragma once

#include "common.hpp"

namespace nodals {
namespace gnn {

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short bfBits(float f) { return (short)floatToBf16(f).bits; }
__device__ __forceinline__ float bfRound(float f) { return bf16ToFloat(floatToBf16(f)); }

__device__ __forceinline__ f32x16 mfma(const s16x8& a, const s16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// C = A B with A (M x KA, FP32) and B given transposed as Bt (NC x KA, BF16).
// One wavefront per 32 rows, all NC columns.  Epilogue:
//   EPI 0: Z = C (+ bias when given)
//   EPI 1: Z = C + bias, H = relu(Z)                 (encoder)
//   EPI 2: Z = C + bias, H = Hprev + relu(Z)          (residual layer)
template <int KA, int NC, int EPI>
__global__ void __launch_bounds__(256) k_gemm(int M, const float* __restrict__ A, const bf16* __restrict__ Bt,
                                              const float* __restrict__ bias, float* __restrict__ Z,
                                              const float* __restrict__ Hprev, float* __restrict__ H) {
  static_assert(KA % 16 == 0 && NC % 32 == 0, "tile shape");
  const int lane = threadIdx.x & 63;
  const int r0 = ((blockIdx.x * 256 + threadIdx.x) >> 6) * 32;
  if (r0 >= M) return;
  f32x16 acc[NC / 32];
#pragma unroll
  for (int j = 0; j < NC / 32; ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[j][r] = 0.0f;
  const int arow = r0 + (lane & 31);
  const int kh = (lane >> 5) * 8;
  const unsigned short* B = reinterpret_cast<const unsigned short*>(Bt);
#pragma unroll
  for (int k0 = 0; k0 < KA; k0 += 16) {
    const float4* ap = reinterpret_cast<const float4*>(A + (long long)arow * KA + k0 + kh);
    const float4 a0 = ap[0], a1 = ap[1];
    s16x8 av;
    av[0] = bfBits(a0.x);
    av[1] = bfBits(a0.y);
    av[2] = bfBits(a0.z);
    av[3] = bfBits(a0.w);
    av[4] = bfBits(a1.x);
    av[5] = bfBits(a1.y);
    av[6] = bfBits(a1.z);
    av[7] = bfBits(a1.w);
#pragma unroll
    for (int j = 0; j < NC / 32; ++j) {
      const int col = 32 * j + (lane & 31);
      const s16x8 bv = *reinterpret_cast<const s16x8*>(B + (long long)col * KA + k0 + kh);
      acc[j] = mfma(av, bv, acc[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < NC / 32; ++j) {
    const int col = 32 * j + (lane & 31);
    const float bb = bias ? bias[col] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = r0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const long long o = (long long)row * NC + col;
      const float v = acc[j][r] + bb;
      Z[o] = v;
      if (EPI == 1) H[o] = fmaxf(v, 0.0f);
      if (EPI == 2) H[o] = Hprev[o] + fmaxf(v, 0.0f);
    }
  }
}

// Weight gradient partials: part[c][k][n] = sum over rows of chunk c of
// bf16(A[m][k]) bf16(D[m][n]); chunks of 512 rows, one block per chunk, the
// (ceil(KA/32) x NC/32) output tiles shared by the four wavefronts.
constexpr int kChunk = 512;
template <int KA, int NC>
__global__ void __launch_bounds__(256) k_gemm_wgrad(int M, const float* __restrict__ A, const float* __restrict__ D,
                                                    float* __restrict__ part) {
  constexpr int KT = (KA + 31) / 32, NT = NC / 32;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m0 = blockIdx.x * kChunk;
  const int mEnd = min(M, m0 + kChunk);
  for (int t = wv; t < KT * NT; t += 4) {
    const int kt = t / NT, nt = t % NT;
    const int krow = 32 * kt + (lane & 31);  // feature index of this lane's X row
    const int ncol = 32 * nt + (lane & 31);
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int mb = m0; mb < mEnd; mb += 16) {
      const int mm = mb + (lane >> 5) * 8;
      s16x8 xv, yv;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        xv[q] = krow < KA ? bfBits(A[(long long)(mm + q) * KA + krow]) : (short)0;
        yv[q] = bfBits(D[(long long)(mm + q) * NC + ncol]);
      }
      acc = mfma(xv, yv, acc);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int k = 32 * kt + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (k < KA) part[((long long)blockIdx.x * KA + k) * NC + ncol] = acc[r];
    }
  }
}

// out[i] = sum_c part[c][i] (fixed order), i < len.
__global__ void k_sum_parts(int chunks, int len, const float* __restrict__ part, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int c = 0; c < chunks; ++c) s += part[(long long)c * len + i];
  out[i] = s;
}

// Column sums of D (M x 64) per 512-row chunk: part[c][n].
__global__ void __launch_bounds__(256) k_colsum64(int M, const float* __restrict__ D, float* __restrict__ part) {
  __shared__ float sh[4][64];
  const int n = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int m0 = blockIdx.x * kChunk, mEnd = min(M, m0 + kChunk);
  float s = 0.0f;
  for (int m = m0 + q; m < mEnd; m += 4) s += D[(long long)m * 64 + n];
  sh[q][n] = s;
  __syncthreads();
  if (q == 0) part[(long long)blockIdx.x * 64 + n] = sh[0][n] + sh[1][n] + sh[2][n] + sh[3][n];
}

// X = [H | mean of H over graph neighbours] (M x 128); rows >= n are zero.
__global__ void k_concat_mean(int M, int n, const float* __restrict__ H, const int* __restrict__ off,
                              const int* __restrict__ nbr, float* __restrict__ X) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M * 64) return;
  const int row = (int)(t >> 6), c = (int)(t & 63);
  float h = 0.0f, m = 0.0f;
  if (row < n) {
    h = H[t];
    const int b = off[row], e = off[row + 1];
    for (int k = b; k < e; ++k) m += H[(long long)nbr[k] * 64 + c];
    if (e > b) m /= (float)(e - b);
  }
  X[(long long)row * 128 + c] = h;
  X[(long long)row * 128 + 64 + c] = m;
}

// dHprev = dH + dX[:, :64] + sum_{j in N(row)} dX[j, 64:] / deg(j)  (the
// adjacency is symmetric); rows >= n are zero.
__global__ void k_concat_mean_back(int M, int n, const float* __restrict__ dH, const float* __restrict__ dX,
                                   const int* __restrict__ off, const int* __restrict__ nbr,
                                   float* __restrict__ dHprev) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M * 64) return;
  const int row = (int)(t >> 6), c = (int)(t & 63);
  float v = 0.0f;
  if (row < n) {
    v = dH[t] + dX[(long long)row * 128 + c];
    for (int k = off[row]; k < off[row + 1]; ++k) {
      const int j = nbr[k];
      v += dX[(long long)j * 128 + 64 + c] / (float)(off[j + 1] - off[j]);
    }
  }
  dHprev[t] = v;
}

// dZ = dH * (Z > 0), zero on padding rows.
__global__ void k_relu_back(int M, int n, const float* __restrict__ dH, const float* __restrict__ Z,
                            float* __restrict__ dZ) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M * 64) return;
  const int row = (int)(t >> 6);
  dZ[t] = (row < n && Z[t] > 0.0f) ? dH[t] : 0.0f;
}

// y[i] = sum_k bf16(H[i][k]) bf16(w[k]) + b + X[i][skip]: the network
// learns a correction to one input column (the previous pressure
// correction), so an untrained network already predicts that column.
__global__ void k_decode(int M, const float* __restrict__ H, const float* __restrict__ w, const float* __restrict__ b,
                         const float* __restrict__ X, int skip, float* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  float s = 0.0f;
  for (int k = 0; k < 64; ++k) s += bfRound(H[(long long)i * 64 + k]) * bfRound(w[k]);
  y[i] = s + b[0] + X[(long long)i * 16 + skip];
}

// Squared error and its gradient: dy = 2 (y - t) / n on rows < n, partial
// sums of (y - t)^2 per block.
__global__ void __launch_bounds__(256) k_mse(int M, int n, const float* __restrict__ y, const float* __restrict__ t,
                                             float* __restrict__ dy, float* __restrict__ part) {
  __shared__ float sh[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  float e = 0.0f;
  if (i < M) {
    float g = 0.0f;
    if (i < n) {
      const float d = y[i] - t[i];
      e = d * d;
      g = 2.0f * d / (float)n;
    }
    dy[i] = g;
  }
  sh[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Decoder backward: dH[i][k] = dy[i] bf16(w[k]); per-chunk partials of
// dw[k] = sum_i bf16(H[i][k]) dy[i] and db = sum_i dy[i]: part[c][65].
__global__ void __launch_bounds__(256) k_decode_back(int M, const float* __restrict__ H, const float* __restrict__ w,
                                                     const float* __restrict__ dy, float* __restrict__ dH,
                                                     float* __restrict__ part) {
  __shared__ float sh[4][65];
  const int k = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int m0 = blockIdx.x * kChunk, mEnd = min(M, m0 + kChunk);
  const float wk = bfRound(w[k]);
  float sw = 0.0f, sb = 0.0f;
  for (int m = m0 + q; m < mEnd; m += 4) {
    const float g = dy[m];
    dH[(long long)m * 64 + k] = g * wk;
    sw += bfRound(H[(long long)m * 64 + k]) * g;
    if (k == 0) sb += g;
  }
  sh[q][k] = sw;
  if (k == 0) sh[q][64] = sb;
  __syncthreads();
  if (q == 0) {
    part[(long long)blockIdx.x * 65 + k] = sh[0][k] + sh[1][k] + sh[2][k] + sh[3][k];
    if (k == 0) part[(long long)blockIdx.x * 65 + 64] = sh[0][64] + sh[1][64] + sh[2][64] + sh[3][64];
  }
}

// Adam step on the FP32 master parameters (Kingma & Ba 2015, bias-corrected).
__global__ void k_adam(int np, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of a K x N FP32 weight: W16 (K x N) and its transpose (N x K).
__global__ void k_weight_bf16(int K, int N, const float* __restrict__ W, bf16* __restrict__ W16,
                              bf16* __restrict__ Wt16) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * N) return;
  const int k = i / N, n = i % N;
  const bf16 b = floatToBf16(W[i]);
  W16[i] = b;
  Wt16[n * K + k] = b;
}

}  // namespace gnn
}  // namespace nodals
