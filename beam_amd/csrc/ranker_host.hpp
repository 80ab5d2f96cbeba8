// Host-side owner of the ranker's device buffers (parameters, Adam moments,
// BF16 operand copies, normalisation and the training batch's activations).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ranker.hpp"

namespace beam {

constexpr int kNormCount = 2 * kFeat + 2;  // feature means, inverse stds, target mean, inverse std

struct RankerState {
  float* params = nullptr;
  float* grad = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  unsigned short* w1t = nullptr;
  unsigned short* w2t = nullptr;
  float* norm = nullptr;
  double* moments = nullptr;
  int batch = 0;
  int32_t* idx = nullptr;
  float* pred = nullptr;
  float* h1 = nullptr;
  float* h2 = nullptr;
  float* xn = nullptr;
  float* dy = nullptr;
  float* dh1 = nullptr;
  float* dh2 = nullptr;
  float* loss = nullptr;
  float* part = nullptr;    // per-block partial sums of the batch reductions
  double* dpart = nullptr;  // per-block partial moments
  int step = 0;

  RankerState() = default;
  RankerState(const RankerState&) = delete;
  RankerState& operator=(const RankerState&) = delete;
  ~RankerState();

  void Allocate(int batch);
  void Init(uint64_t seed, hipStream_t stream);
  void Upload(const float* h_params, const float* h_norm, hipStream_t stream);
  void Download(float* h_params, float* h_norm, hipStream_t stream) const;
  void Pack(hipStream_t stream);
  RankerView View() const;
  // Feature and target moments of n samples into the normalisation.
  void Normalize(const float* feats, const float* target, int n, hipStream_t stream);
  // Predictions of rows (sample rows[r], or r) into out; keep = store
  // activations for Gradient.
  void Forward(const float* feats, const int32_t* rows, int n, float* out, bool keep, hipStream_t stream);
  // Gradient of the mean squared error of one batch into grad (and loss).
  void Gradient(const float* feats, const float* target, const int32_t* rows, int n, hipStream_t stream);
  void AdamStep(float lr, hipStream_t stream);
  // Normalises, then runs `epochs` shuffled passes of Adam over full batches;
  // returns the last batch's loss.
  float Train(const float* feats, const float* target, int n, int epochs, int batch_size, float lr, uint64_t seed,
              hipStream_t stream);
};

}  // namespace beam
