// Message-passing network that predicts the pressure correction of a SIMPLE
// iteration from the continuity residual, the previous two corrections and
// the wall distance, on the graph of face-adjacent cells (the P0 pressure
// unknowns).
//
//   h0 = relu(x W0 + b0)                            x: 16 features (8 used)
//   h_l = h_{l-1} + relu([h_{l-1} | mean_N h_{l-1}] W_l + b_l),  l = 1..4
//   y = h4 w + b + x[2]        (x[2]: the previous correction, scaled)
// Width 64, BF16 weights and GEMM inputs, FP32 accumulation and FP32
// master parameters (Adam).  The same object holds the training set
// recorded from the solver's own SIMPLE iterations.
#pragma once

#include <vector>

#include "common.hpp"

namespace nodals {

struct CellGraph {
  int n = 0;
  std::vector<int> off, nbr;  // symmetric adjacency, CSR
};

class GnnModel {
 public:
  static constexpr int kIn = 16, kHid = 64, kCat = 128, kLayers = 4;
  static constexpr int kFeatures = 8;  // used input columns
  static constexpr int kSkip = 2;      // input column added to the output (previous correction)

  GnnModel(const CellGraph& g, unsigned seed);
  ~GnnModel();
  GnnModel(const GnnModel&) = delete;
  GnnModel& operator=(const GnnModel&) = delete;

  int n() const { return n_; }
  int paddedRows() const { return M_; }
  int paramCount() const { return np_; }
  hipStream_t stream() const { return stream_; }

  void getParams(float* host) const;
  void setParams(const float* host);

  // Inference / training on device feature rows X (M x kIn FP32, rows >= n
  // zero); y (M) device output.  The stream is the caller's.
  void forward(hipStream_t s, const float* X, float* y);
  // Gradient of mean((y - t)^2) over the n rows for the last forward;
  // the loss lands in device memory (loss[0]).
  void backward(hipStream_t s, const float* X, const float* t, float* loss);
  void adamStep(hipStream_t s, float lr);
  const float* gradDevice() const { return grad_.data(); }
  float* outputDevice() { return y_.data(); }

  // Training set: rows of kIn features and one target per cell, per sample.
  void reserveSamples(int capacity);
  int samples() const { return nSamples_; }
  int capacity() const { return cap_; }
  float* sampleX(int i) { return dataX_.data() + (long long)i * M_ * kIn; }
  float* sampleT(int i) { return dataT_.data() + (long long)i * M_; }
  int takeSlot();  // index of the next sample slot (or -1 when full)
  // Epochs of Adam over the recorded samples in a seeded shuffled order;
  // returns the mean loss of the last epoch.
  double train(int epochs, float lr, unsigned seed, std::vector<double>* epochLoss);
  // Mean loss over the recorded samples (no update).
  double evaluate();

 private:
  void refreshWeights(hipStream_t s);
  int n_ = 0, M_ = 0, np_ = 0, chunks_ = 0;
  int adamT_ = 0;
  hipStream_t stream_ = nullptr;
  DBuf<int> off_, nbr_;
  DBuf<float> P_, grad_, m_, v_;
  DBuf<bf16> w16_, wt16_;  // per GEMM layer: W (K x N) and W^T (N x K)
  DBuf<float> Z0_, H0_, X_[kLayers], Z_[kLayers], H_[kLayers], y_;
  DBuf<float> dHa_, dHb_, dZ_, dX_, dy_, part_, lossPart_, loss_, epochLoss_;
  DBuf<float> dataX_, dataT_;
  int cap_ = 0, nSamples_ = 0;
  float* hostLoss_ = nullptr;
  // parameter offsets
  int oW0_ = 0, ob0_ = 0, oW_[kLayers] = {}, ob_[kLayers] = {}, ow_ = 0, obout_ = 0;
  int o16W0_ = 0, o16W_[kLayers] = {};
};

}  // namespace nodals
