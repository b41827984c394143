// Message-passing pressure-correction network: parameters, forward,
// backward, Adam and the training loop.  See gnn.hpp.
#include "gnn.hpp"

#include <algorithm>
#include <cmath>
#include <random>

#include "gnn_kernels.hpp"
#include "la_kernels.hpp"

namespace nodals {

using namespace gnn;
using dev::copy;
using dev::fill;

namespace {
inline int grid1(long long n) { return (int)((n + 255) / 256); }
}  // namespace

GnnModel::GnnModel(const CellGraph& g, unsigned seed) {
  require(g.n > 0 && (int)g.off.size() == g.n + 1, "GNN graph is empty or malformed");
  n_ = g.n;
  M_ = (n_ + 31) / 32 * 32;
  chunks_ = (M_ + kChunk - 1) / kChunk;
  // parameter layout (FP32 master)
  int o = 0;
  oW0_ = o;
  o += kIn * kHid;
  ob0_ = o;
  o += kHid;
  for (int l = 0; l < kLayers; ++l) {
    oW_[l] = o;
    o += kCat * kHid;
    ob_[l] = o;
    o += kHid;
  }
  ow_ = o;
  o += kHid;
  obout_ = o;
  o += 1;
  np_ = o;
  // He-uniform initialisation from a seeded generator; feature rows beyond
  // kFeatures and all biases start at zero.
  std::vector<float> P((std::size_t)np_, 0.0f);
  std::mt19937 rng(seed);
  auto initW = [&](int off, int K, int N, int activeK) {
    const float a = std::sqrt(6.0f / (float)activeK);
    std::uniform_real_distribution<float> U(-a, a);
    for (int k = 0; k < K; ++k)
      for (int n = 0; n < N; ++n) P[(std::size_t)off + (std::size_t)k * N + n] = k < activeK ? U(rng) : 0.0f;
  };
  initW(oW0_, kIn, kHid, kFeatures);
  for (int l = 0; l < kLayers; ++l) {
    initW(oW_[l], kCat, kHid, kCat);
    // residual branches start small so the initial network is close to the
    // encoder-decoder path
    for (int i = 0; i < kCat * kHid; ++i) P[(std::size_t)oW_[l] + i] *= 0.1f;
  }
  initW(ow_, kHid, 1, kHid);
  for (int i = 0; i < kHid; ++i) P[(std::size_t)ow_ + i] *= 0.1f;

  NODALS_HIP(hipSetDevice(0));
  NODALS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  off_.upload(g.off);
  std::vector<int> nb(g.nbr);
  if (nb.empty()) nb.push_back(0);
  nbr_.upload(nb);
  P_.upload(P);
  grad_.alloc((std::size_t)np_);
  m_.alloc((std::size_t)np_);
  v_.alloc((std::size_t)np_);
  fill<float>(stream_, np_, m_.data(), 0.0f);
  fill<float>(stream_, np_, v_.data(), 0.0f);
  o16W0_ = 0;
  int o16 = kIn * kHid;
  for (int l = 0; l < kLayers; ++l) {
    o16W_[l] = o16;
    o16 += kCat * kHid;
  }
  w16_.alloc((std::size_t)o16);
  wt16_.alloc((std::size_t)o16);
  const std::size_t M = (std::size_t)M_;
  Z0_.alloc(M * kHid);
  H0_.alloc(M * kHid);
  for (int l = 0; l < kLayers; ++l) {
    X_[l].alloc(M * kCat);
    Z_[l].alloc(M * kHid);
    H_[l].alloc(M * kHid);
  }
  y_.alloc(M);
  dHa_.alloc(M * kHid);
  dHb_.alloc(M * kHid);
  dZ_.alloc(M * kHid);
  dX_.alloc(M * kCat);
  dy_.alloc(M);
  part_.alloc((std::size_t)chunks_ * kCat * kHid);
  lossPart_.alloc((std::size_t)grid1(M_));
  loss_.alloc(1);
  NODALS_HIP(hipHostMalloc((void**)&hostLoss_, sizeof(float), hipHostMallocCoherent));
  refreshWeights(stream_);
  NODALS_HIP(hipStreamSynchronize(stream_));
}

GnnModel::~GnnModel() {
  if (hostLoss_) (void)hipHostFree(hostLoss_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void GnnModel::refreshWeights(hipStream_t s) {
  hipLaunchKernelGGL(k_weight_bf16, grid1(kIn * kHid), 256, 0, s, kIn, kHid, P_.data() + oW0_, w16_.data() + o16W0_,
                     wt16_.data() + o16W0_);
  NODALS_LAUNCHED();
  for (int l = 0; l < kLayers; ++l) {
    hipLaunchKernelGGL(k_weight_bf16, grid1(kCat * kHid), 256, 0, s, kCat, kHid, P_.data() + oW_[l],
                       w16_.data() + o16W_[l], wt16_.data() + o16W_[l]);
    NODALS_LAUNCHED();
  }
}

void GnnModel::getParams(float* host) const {
  NODALS_HIP(hipStreamSynchronize(stream_));
  NODALS_HIP(hipMemcpy(host, P_.data(), sizeof(float) * (std::size_t)np_, hipMemcpyDeviceToHost));
}

void GnnModel::setParams(const float* host) {
  NODALS_HIP(hipStreamSynchronize(stream_));
  NODALS_HIP(hipMemcpy(P_.data(), host, sizeof(float) * (std::size_t)np_, hipMemcpyHostToDevice));
  fill<float>(stream_, np_, m_.data(), 0.0f);
  fill<float>(stream_, np_, v_.data(), 0.0f);
  adamT_ = 0;
  refreshWeights(stream_);
  NODALS_HIP(hipStreamSynchronize(stream_));
}

void GnnModel::forward(hipStream_t s, const float* X, float* y) {
  const int gemmGrid = (M_ / 32 + 3) / 4;
  hipLaunchKernelGGL((k_gemm<kIn, kHid, 1>), gemmGrid, 256, 0, s, M_, X, wt16_.data() + o16W0_, P_.data() + ob0_,
                     Z0_.data(), nullptr, H0_.data());
  NODALS_LAUNCHED();
  const float* h = H0_.data();
  for (int l = 0; l < kLayers; ++l) {
    hipLaunchKernelGGL(k_concat_mean, grid1((long long)M_ * 64), 256, 0, s, M_, n_, h, off_.data(), nbr_.data(),
                       X_[l].data());
    NODALS_LAUNCHED();
    hipLaunchKernelGGL((k_gemm<kCat, kHid, 2>), gemmGrid, 256, 0, s, M_, X_[l].data(), wt16_.data() + o16W_[l],
                       P_.data() + ob_[l], Z_[l].data(), h, H_[l].data());
    NODALS_LAUNCHED();
    h = H_[l].data();
  }
  hipLaunchKernelGGL(k_decode, grid1(M_), 256, 0, s, M_, h, P_.data() + ow_, P_.data() + obout_, X, kSkip,
                     y_.data());
  NODALS_LAUNCHED();
  if (y && y != y_.data()) copy<float>(s, M_, y_.data(), y);
}

void GnnModel::backward(hipStream_t s, const float* X, const float* t, float* loss) {
  const int gemmGrid = (M_ / 32 + 3) / 4;
  hipLaunchKernelGGL(k_mse, grid1(M_), 256, 0, s, M_, n_, y_.data(), t, dy_.data(), lossPart_.data());
  NODALS_LAUNCHED();
  hipLaunchKernelGGL(k_sum_parts, 1, 256, 0, s, grid1(M_), 1, lossPart_.data(), loss);
  NODALS_LAUNCHED();
  // decoder
  hipLaunchKernelGGL(k_decode_back, chunks_, 256, 0, s, M_, H_[kLayers - 1].data(), P_.data() + ow_, dy_.data(),
                     dHa_.data(), part_.data());
  NODALS_LAUNCHED();
  hipLaunchKernelGGL(k_sum_parts, 1, 256, 0, s, chunks_, kHid + 1, part_.data(), grad_.data() + ow_);
  NODALS_LAUNCHED();
  float* dH = dHa_.data();
  float* dHn = dHb_.data();
  for (int l = kLayers - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_relu_back, grid1((long long)M_ * 64), 256, 0, s, M_, n_, dH, Z_[l].data(), dZ_.data());
    NODALS_LAUNCHED();
    // dW_l = X_l^T dZ, db_l = colsum dZ
    hipLaunchKernelGGL((k_gemm_wgrad<kCat, kHid>), chunks_, 256, 0, s, M_, X_[l].data(), dZ_.data(), part_.data());
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_sum_parts, grid1(kCat * kHid), 256, 0, s, chunks_, kCat * kHid, part_.data(),
                       grad_.data() + oW_[l]);
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_colsum64, chunks_, 256, 0, s, M_, dZ_.data(), part_.data());
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_sum_parts, 1, 256, 0, s, chunks_, kHid, part_.data(), grad_.data() + ob_[l]);
    NODALS_LAUNCHED();
    // dX_l = dZ W_l^T (B operand: W_l row-major = (W_l^T)^T)
    hipLaunchKernelGGL((k_gemm<kHid, kCat, 0>), gemmGrid, 256, 0, s, M_, dZ_.data(), w16_.data() + o16W_[l], nullptr,
                       dX_.data(), nullptr, nullptr);
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_concat_mean_back, grid1((long long)M_ * 64), 256, 0, s, M_, n_, dH, dX_.data(), off_.data(),
                       nbr_.data(), dHn);
    NODALS_LAUNCHED();
    std::swap(dH, dHn);
  }
  // encoder
  hipLaunchKernelGGL(k_relu_back, grid1((long long)M_ * 64), 256, 0, s, M_, n_, dH, Z0_.data(), dZ_.data());
  NODALS_LAUNCHED();
  hipLaunchKernelGGL((k_gemm_wgrad<kIn, kHid>), chunks_, 256, 0, s, M_, X, dZ_.data(), part_.data());
  NODALS_LAUNCHED();
  hipLaunchKernelGGL(k_sum_parts, grid1(kIn * kHid), 256, 0, s, chunks_, kIn * kHid, part_.data(),
                     grad_.data() + oW0_);
  NODALS_LAUNCHED();
  hipLaunchKernelGGL(k_colsum64, chunks_, 256, 0, s, M_, dZ_.data(), part_.data());
  NODALS_LAUNCHED();
  hipLaunchKernelGGL(k_sum_parts, 1, 256, 0, s, chunks_, kHid, part_.data(), grad_.data() + ob0_);
  NODALS_LAUNCHED();
}

void GnnModel::adamStep(hipStream_t s, float lr) {
  ++adamT_;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)adamT_), c2 = 1.0f - std::pow(b2, (float)adamT_);
  hipLaunchKernelGGL(k_adam, grid1(np_), 256, 0, s, np_, P_.data(), grad_.data(), m_.data(), v_.data(), lr, b1, b2,
                     eps, c1, c2);
  NODALS_LAUNCHED();
  refreshWeights(s);
}

void GnnModel::reserveSamples(int capacity) {
  require(capacity >= 1, "sample capacity must be >= 1");
  NODALS_HIP(hipStreamSynchronize(stream_));
  if (capacity != cap_) {
    dataX_.alloc((std::size_t)capacity * M_ * kIn);
    dataT_.alloc((std::size_t)capacity * M_);
    fill<float>(stream_, (long long)capacity * M_ * kIn, dataX_.data(), 0.0f);
    fill<float>(stream_, (long long)capacity * M_, dataT_.data(), 0.0f);
    NODALS_HIP(hipStreamSynchronize(stream_));
  }
  cap_ = capacity;
  nSamples_ = 0;
}

int GnnModel::takeSlot() {
  if (nSamples_ >= cap_) return -1;
  return nSamples_++;
}

double GnnModel::train(int epochs, float lr, unsigned seed, std::vector<double>* epochLoss) {
  require(nSamples_ >= 1, "GNN training needs recorded samples");
  require(epochs >= 1, "epochs must be >= 1");
  std::mt19937 rng(seed);
  std::vector<int> order((std::size_t)nSamples_);
  double last = 0.0;
  for (int e = 0; e < epochs; ++e) {
    for (int i = 0; i < nSamples_; ++i) order[(std::size_t)i] = i;
    std::shuffle(order.begin(), order.end(), rng);
    // the per-step squared-error sums accumulate on the device (one slot
    // per step, summed once per epoch)
    if (epochLoss_.size() < (std::size_t)nSamples_) epochLoss_.alloc((std::size_t)nSamples_);
    int q = 0;
    for (int i : order) {
      forward(stream_, sampleX(i), nullptr);
      backward(stream_, sampleX(i), sampleT(i), epochLoss_.data() + q++);
      adamStep(stream_, lr);
    }
    hipLaunchKernelGGL(k_sum_parts, 1, 256, 0, stream_, nSamples_, 1, epochLoss_.data(), loss_.data());
    NODALS_LAUNCHED();
    NODALS_HIP(hipMemcpyAsync(hostLoss_, loss_.data(), sizeof(float), hipMemcpyDeviceToHost, stream_));
    NODALS_HIP(hipStreamSynchronize(stream_));
    last = (double)*hostLoss_ / n_ / nSamples_;
    if (epochLoss) epochLoss->push_back(last);
  }
  return last;
}

double GnnModel::evaluate() {
  require(nSamples_ >= 1, "GNN evaluation needs recorded samples");
  double sum = 0.0;
  for (int i = 0; i < nSamples_; ++i) {
    forward(stream_, sampleX(i), nullptr);
    hipLaunchKernelGGL(k_mse, grid1(M_), 256, 0, stream_, M_, n_, y_.data(), sampleT(i), dy_.data(), lossPart_.data());
    NODALS_LAUNCHED();
    hipLaunchKernelGGL(k_sum_parts, 1, 256, 0, stream_, grid1(M_), 1, lossPart_.data(), loss_.data());
    NODALS_LAUNCHED();
    NODALS_HIP(hipMemcpyAsync(hostLoss_, loss_.data(), sizeof(float), hipMemcpyDeviceToHost, stream_));
    NODALS_HIP(hipStreamSynchronize(stream_));
    sum += (double)*hostLoss_ / n_;
  }
  return sum / nSamples_;
}

}  // namespace nodals
