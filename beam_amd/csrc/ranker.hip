// Training and batch inference of the learned pair ranker on gfx950: the
// forward runs on MFMA (BF16 operands, FP32 accumulation, ranker.hpp); the
// backward is FP32 (weight gradients as split-K reductions over the batch);
// the optimiser is Adam on the FP32 master parameters, after which the BF16
// operand copies are refreshed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "ranker.hpp"
#include "ranker_host.hpp"

namespace beam {

namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kRedBlocks = 512;      // blocks of the per-column batch reductions
constexpr int kMomentBlocks = 256;

// Rows r of a batch take sample idx[r] (or r without idx).
__global__ __launch_bounds__(kWavesPerBlock * kWave) void RankerForwardKernel(RankerView w, const float* feats,
                                                                             const int32_t* idx, int n, float* pred,
                                                                             float* h1, float* h2, float* xn) {
  const int base = (static_cast<int>(blockIdx.x) * kWavesPerBlock + WaveId()) * 32;
  if (base >= n) return;  // whole wave
  const int lane = LaneId(), col = lane & 31, half = lane >> 5;
  const int row = base + col;
  const bool valid = row < n;
  const long long src = valid ? (idx != nullptr ? idx[row] : row) : 0;
  float xk[12];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int f = 8 * s + 4 * half + j;
      const float raw = valid ? feats[src * kFeat + f] : 0.0f;
      const float x = valid ? (raw - w.norm[f]) * w.norm[kFeat + f] : 0.0f;
      xk[s * 4 + j] = x;
      if (xn != nullptr && valid) xn[static_cast<long long>(row) * kFeat + f] = x;
    }
  }
  const float y = MlpForward32(w, xk, h1, h2, base, min(32, n - base));
  if (half == 0 && valid) pred[row] = y;
}

// dy = 2 (pred - t) / n with t the standardised slog gain; loss += mean square.
__global__ void LossGradKernel(const float* pred, const float* target, const int32_t* idx, const float* norm, int n,
                               float* dy, float* loss) {
  __shared__ float s_sum;
  if (threadIdx.x == 0) s_sum = 0.0f;
  __syncthreads();
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.0f;
  if (r < n) {
    const float t = (target[idx != nullptr ? idx[r] : r] - norm[2 * kFeat]) * norm[2 * kFeat + 1];
    const float d = pred[r] - t;
    dy[r] = 2.0f * d / static_cast<float>(n);
    sq = d * d / static_cast<float>(n);
  }
  atomicAdd(&s_sum, sq);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, s_sum);
}

// Reductions over the batch write one partial row per block and are summed
// in block order afterwards (ReduceRowsKernel), so training is deterministic.

// Partial [blk][kHid + 1] of dW3 = sum_r dy_r h2_r and db3 = sum_r dy_r;
// dH2 = dy w3 [h2 > 0].
__global__ void OutputGradKernel(const float* dy, const float* h2, const float* params, int n, float* dh2,
                                 float* part) {
  const int tid = static_cast<int>(threadIdx.x);  // blockDim = kHid
  float acc = 0.0f, accb = 0.0f;
  const float w3 = params[kOffW3 + tid];
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    const float d = dy[r];
    const float h = h2[static_cast<long long>(r) * kHid + tid];
    acc += d * h;
    dh2[static_cast<long long>(r) * kHid + tid] = h > 0.0f ? d * w3 : 0.0f;
    if (tid == 0) accb += d;
  }
  part[blockIdx.x * (kHid + 1) + tid] = acc;
  if (tid == 0) part[blockIdx.x * (kHid + 1) + kHid] = accb;
}

// dH1[r][i] = [h1 > 0] sum_j dH2[r][j] W2[i][j], 32 rows per block.
__global__ __launch_bounds__(256) void HiddenGradKernel(const float* dh2, const float* params, const float* h1, int n,
                                                        float* dh1) {
  __shared__ float s_w2[kHid * (kHid + 1)];
  __shared__ float s_d[32][kHid];
  const int tid = static_cast<int>(threadIdx.x);
  for (int k = tid; k < kHid * kHid; k += 256) s_w2[(k / kHid) * (kHid + 1) + k % kHid] = params[kOffW2 + k];
  const int r0 = blockIdx.x * 32;
  for (int k = tid; k < 32 * kHid; k += 256) {
    const int r = r0 + k / kHid;
    s_d[k / kHid][k % kHid] = r < n ? dh2[static_cast<long long>(r) * kHid + k % kHid] : 0.0f;
  }
  __syncthreads();
  const int i = tid & (kHid - 1), rsel = tid >> 7;  // two row phases
  for (int rl = rsel; rl < 32; rl += 2) {
    const int r = r0 + rl;
    if (r >= n) break;
    float acc = 0.0f;
#pragma unroll 8
    for (int j = 0; j < kHid; ++j) acc += s_d[rl][j] * s_w2[i * (kHid + 1) + j];
    const float h = h1[static_cast<long long>(r) * kHid + i];
    dh1[static_cast<long long>(r) * kHid + i] = h > 0.0f ? acc : 0.0f;
  }
}

// Partial [chunk][m x kHid] of G[i][j] = sum_r A[r][i] D[r][j] over one
// chunk of rows per block (A: n x m, D: n x kHid), 32 x 32 tile per block.
constexpr int kGemmChunk = 1024;
__global__ __launch_bounds__(256) void GemmAtBKernel(const float* a, int m, const float* d, int n, float* g) {
  __shared__ float s_a[64][33];
  __shared__ float s_b[64][33];
  const int tid = static_cast<int>(threadIdx.x);
  const int ti = blockIdx.y / (kHid / 32), tj = blockIdx.y % (kHid / 32);
  const int i0 = ti * 32, j0 = tj * 32;
  const int rbeg = blockIdx.x * kGemmChunk, rend = min(n, rbeg + kGemmChunk);
  const int oi = tid / 8, oj = (tid % 8) * 4;  // 32 rows x 4 columns per thread group
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int r0 = rbeg; r0 < rend; r0 += 64) {
    for (int k = tid; k < 64 * 32; k += 256) {
      const int rr = k / 32, cc = k % 32;
      const int r = r0 + rr;
      s_a[rr][cc] = (r < rend && i0 + cc < m) ? a[static_cast<long long>(r) * m + i0 + cc] : 0.0f;
      s_b[rr][cc] = r < rend ? d[static_cast<long long>(r) * kHid + j0 + cc] : 0.0f;
    }
    __syncthreads();
    for (int rr = 0; rr < 64; ++rr) {
      const float av = s_a[rr][oi];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] += av * s_b[rr][oj + q];
    }
    __syncthreads();
  }
  if (i0 + oi < m) {
    float* out = g + static_cast<long long>(blockIdx.x) * m * kHid;
#pragma unroll
    for (int q = 0; q < 4; ++q) out[(i0 + oi) * kHid + j0 + oj + q] = acc[q];
  }
}

// Partial [blk][kHid] of sum_r D[r][j] (D: n x kHid).
__global__ void ColSumKernel(const float* d, int n, float* part) {
  const int j = static_cast<int>(threadIdx.x);  // blockDim = kHid
  float acc = 0.0f;
  for (int r = blockIdx.x; r < n; r += gridDim.x) acc += d[static_cast<long long>(r) * kHid + j];
  part[blockIdx.x * kHid + j] = acc;
}

// out[c] = sum over rows r (in order) of part[r][c].
template <typename T>
__global__ void ReduceRowsKernel(const T* part, int rows, int width, T* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= width) return;
  T acc = 0;
  for (int r = 0; r < rows; ++r) acc += part[static_cast<long long>(r) * width + c];
  out[c] = acc;
}

__global__ void AdamKernel(float* params, const float* grad, float* m, float* v, int n, float lr, float b1, float b2,
                           float eps, float bc1, float bc2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = grad[k];
  const float mk = b1 * m[k] + (1.0f - b1) * gk;
  const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  params[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
}

__global__ void PackWeightsKernel(const float* params, unsigned short* w1t, unsigned short* w2t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kFeat * kHid) {
    const int in = k / kHid, out = k % kHid;
    w1t[out * kFeat + in] = F32ToBf16(params[kOffW1 + k]);
  }
  if (k < kHid * kHid) {
    const int in = k / kHid, out = k % kHid;
    w2t[out * kHid + in] = F32ToBf16(params[kOffW2 + k]);
  }
}

// Partial [blk][2 kFeat + 2]: sums of x then x^2 per feature, the target's
// two last.
__global__ __launch_bounds__(256) void MomentsKernel(const float* feats, const float* target, int n, double* part) {
  __shared__ double s[8][2 * kFeat + 2];
  const int tid = static_cast<int>(threadIdx.x);
  const int f = tid & 31, rl = tid >> 5;  // 8 rows in flight per block
  double s1 = 0.0, s2 = 0.0;
  if (f <= kFeat) {
    for (long long r = static_cast<long long>(blockIdx.x) * 8 + rl; r < n; r += static_cast<long long>(gridDim.x) * 8) {
      const double x = f < kFeat ? feats[r * kFeat + f] : target[r];
      s1 += x;
      s2 += x * x;
    }
    s[rl][f < kFeat ? f : 2 * kFeat] = s1;
    s[rl][f < kFeat ? kFeat + f : 2 * kFeat + 1] = s2;
  }
  __syncthreads();
  if (tid < 2 * kFeat + 2) {
    double t = 0.0;
    for (int r = 0; r < 8; ++r) t += s[r][tid];
    part[blockIdx.x * (2 * kFeat + 2) + tid] = t;
  }
}

__global__ void NormFinalizeKernel(const double* acc, int n, float* norm) {
  const int f = static_cast<int>(threadIdx.x);
  if (f > kFeat) return;
  const double mean = (f < kFeat ? acc[f] : acc[2 * kFeat]) / n;
  const double sq = (f < kFeat ? acc[kFeat + f] : acc[2 * kFeat + 1]) / n;
  const double sd = sqrt(fmax(sq - mean * mean, 0.0));
  const float inv = sd > 1e-6 ? static_cast<float>(1.0 / sd) : 1.0f;
  if (f < kFeat) {
    norm[f] = static_cast<float>(mean);
    norm[kFeat + f] = inv;
  } else {
    norm[2 * kFeat] = static_cast<float>(mean);
    norm[2 * kFeat + 1] = inv;
  }
}

template <typename T>
T* Alloc(size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

}  // namespace

RankerState::~RankerState() {
  void* ptrs[] = {params, grad, m, v, w1t, w2t, norm, moments, idx, pred, h1, h2, xn, dy, dh1, dh2, loss, part, dpart};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

void RankerState::Allocate(int batch_) {
  if (params == nullptr) {
    params = Alloc<float>(kParams);
    grad = Alloc<float>(kParams);
    m = Alloc<float>(kParams);
    v = Alloc<float>(kParams);
    w1t = Alloc<unsigned short>(kFeat * kHid);
    w2t = Alloc<unsigned short>(kHid * kHid);
    norm = Alloc<float>(kNormCount);
    moments = Alloc<double>(2 * kFeat + 2);
    dpart = Alloc<double>(static_cast<size_t>(kMomentBlocks) * (2 * kFeat + 2));
    loss = Alloc<float>(1);
  }
  if (batch_ > batch) {
    void* ptrs[] = {idx, pred, h1, h2, xn, dy, dh1, dh2, part};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    batch = batch_;
    idx = Alloc<int32_t>(batch);
    pred = Alloc<float>(batch);
    dy = Alloc<float>(batch);
    xn = Alloc<float>(static_cast<size_t>(batch) * kFeat);
    h1 = Alloc<float>(static_cast<size_t>(batch) * kHid);
    h2 = Alloc<float>(static_cast<size_t>(batch) * kHid);
    dh1 = Alloc<float>(static_cast<size_t>(batch) * kHid);
    dh2 = Alloc<float>(static_cast<size_t>(batch) * kHid);
    const size_t chunks = (static_cast<size_t>(batch) + kGemmChunk - 1) / kGemmChunk;
    part = Alloc<float>(std::max<size_t>(chunks * kHid * kHid, static_cast<size_t>(kRedBlocks) * (kHid + 1)));
  }
}

// He-uniform initial weights from a seeded host generator; zero biases.
void RankerState::Init(uint64_t seed, hipStream_t stream) {
  Allocate(0);
  std::vector<float> h(kParams, 0.0f);
  std::mt19937_64 rng(seed ^ 0x5851f42d4c957f2dULL);
  auto fill = [&](int off, int fan_in, int count) {
    const float lim = std::sqrt(6.0f / static_cast<float>(fan_in));
    std::uniform_real_distribution<float> u(-lim, lim);
    for (int k = 0; k < count; ++k) h[static_cast<size_t>(off + k)] = u(rng);
  };
  fill(kOffW1, kFeat, kFeat * kHid);
  fill(kOffW2, kHid, kHid * kHid);
  fill(kOffW3, kHid, kHid);
  std::vector<float> hn(kNormCount, 0.0f);
  for (int f = 0; f < kFeat; ++f) hn[static_cast<size_t>(kFeat + f)] = 1.0f;
  hn[2 * kFeat + 1] = 1.0f;
  Upload(h.data(), hn.data(), stream);
  step = 0;
}

void RankerState::Upload(const float* h_params, const float* h_norm, hipStream_t stream) {
  Allocate(0);
  HIP_CHECK(hipMemcpyAsync(params, h_params, sizeof(float) * kParams, hipMemcpyHostToDevice, stream));
  HIP_CHECK(hipMemcpyAsync(norm, h_norm, sizeof(float) * kNormCount, hipMemcpyHostToDevice, stream));
  HIP_CHECK(hipMemsetAsync(m, 0, sizeof(float) * kParams, stream));
  HIP_CHECK(hipMemsetAsync(v, 0, sizeof(float) * kParams, stream));
  Pack(stream);
  HIP_CHECK(hipStreamSynchronize(stream));
}

void RankerState::Download(float* h_params, float* h_norm, hipStream_t stream) const {
  HIP_CHECK(hipMemcpyAsync(h_params, params, sizeof(float) * kParams, hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipMemcpyAsync(h_norm, norm, sizeof(float) * kNormCount, hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
}

void RankerState::Pack(hipStream_t stream) {
  PackWeightsKernel<<<(kHid * kHid + 255) / 256, 256, 0, stream>>>(params, w1t, w2t);
}

RankerView RankerState::View() const { return RankerView{w1t, w2t, params, norm}; }

void RankerState::Normalize(const float* feats, const float* target, int n, hipStream_t stream) {
  Allocate(0);
  if (n <= 0) return;
  MomentsKernel<<<kMomentBlocks, 256, 0, stream>>>(feats, target, n, dpart);
  ReduceRowsKernel<double><<<1, 64, 0, stream>>>(dpart, kMomentBlocks, 2 * kFeat + 2, moments);
  NormFinalizeKernel<<<1, 64, 0, stream>>>(moments, n, norm);
}

void RankerState::Forward(const float* feats, const int32_t* rows, int n, float* out, bool keep, hipStream_t stream) {
  if (n <= 0) return;
  if (keep) Allocate(n);
  const int waves = (n + 31) / 32;
  RankerForwardKernel<<<(waves + kWavesPerBlock - 1) / kWavesPerBlock, kWavesPerBlock * kWave, 0, stream>>>(
      View(), feats, rows, n, out, keep ? h1 : nullptr, keep ? h2 : nullptr, keep ? xn : nullptr);
}

void RankerState::Gradient(const float* feats, const float* target, const int32_t* rows, int n, hipStream_t stream) {
  Allocate(n);
  HIP_CHECK(hipMemsetAsync(grad, 0, sizeof(float) * kParams, stream));
  HIP_CHECK(hipMemsetAsync(loss, 0, sizeof(float), stream));
  Forward(feats, rows, n, pred, true, stream);
  LossGradKernel<<<(n + 255) / 256, 256, 0, stream>>>(pred, target, rows, norm, n, dy, loss);
  const int rblocks = std::min(kRedBlocks, n);
  OutputGradKernel<<<rblocks, kHid, 0, stream>>>(dy, h2, params, n, dh2, part);
  ReduceRowsKernel<float><<<1, kHid + 1, 0, stream>>>(part, rblocks, kHid + 1, grad + kOffW3);
  HiddenGradKernel<<<(n + 31) / 32, 256, 0, stream>>>(dh2, params, h1, n, dh1);
  const int chunks = (n + kGemmChunk - 1) / kGemmChunk;
  GemmAtBKernel<<<dim3(chunks, (kHid / 32) * (kHid / 32)), 256, 0, stream>>>(h1, kHid, dh2, n, part);
  ReduceRowsKernel<float><<<kHid * kHid / 256, 256, 0, stream>>>(part, chunks, kHid * kHid, grad + kOffW2);
  GemmAtBKernel<<<dim3(chunks, 1 * (kHid / 32)), 256, 0, stream>>>(xn, kFeat, dh1, n, part);
  ReduceRowsKernel<float><<<kFeat * kHid / 256, 256, 0, stream>>>(part, chunks, kFeat * kHid, grad + kOffW1);
  ColSumKernel<<<rblocks, kHid, 0, stream>>>(dh1, n, part);
  ReduceRowsKernel<float><<<1, kHid, 0, stream>>>(part, rblocks, kHid, grad + kOffB1);
  ColSumKernel<<<rblocks, kHid, 0, stream>>>(dh2, n, part);
  ReduceRowsKernel<float><<<1, kHid, 0, stream>>>(part, rblocks, kHid, grad + kOffB2);
}

void RankerState::AdamStep(float lr, hipStream_t stream) {
  ++step;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float bc1 = 1.0f - std::pow(b1, static_cast<float>(step));
  const float bc2 = 1.0f - std::pow(b2, static_cast<float>(step));
  AdamKernel<<<(kParams + 255) / 256, 256, 0, stream>>>(params, grad, m, v, kParams, lr, b1, b2, eps, bc1, bc2);
  Pack(stream);
}

float RankerState::Train(const float* feats, const float* target, int n, int epochs, int batch_size, float lr,
                         uint64_t seed, hipStream_t stream) {
  if (n <= 0) return 0.0f;
  const int bs = std::max(32, std::min(batch_size, n));
  Allocate(bs);
  Normalize(feats, target, n, stream);
  std::vector<int32_t> order(static_cast<size_t>(n));
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 rng(seed ^ 0x2545f4914f6cdd1dULL);
  int32_t* d_order = Alloc<int32_t>(static_cast<size_t>(n));
  float last = 0.0f;
  try {
    for (int ep = 0; ep < epochs; ++ep) {
      std::shuffle(order.begin(), order.end(), rng);
      HIP_CHECK(hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream));
      for (int b0 = 0; b0 + bs <= n; b0 += bs) {
        Gradient(feats, target, d_order + b0, bs, stream);
        AdamStep(lr, stream);
      }
      HIP_CHECK(hipMemcpyAsync(&last, loss, sizeof(float), hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
    }
  } catch (...) {
    (void)hipFree(d_order);
    throw;
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  (void)hipFree(d_order);
  return last;
}

}  // namespace beam

// ---------------------------------------------------------------- C-ABI (standalone checks)

namespace {

thread_local std::string g_ranker_error;

template <typename F>
int RankerGuard(F&& f) {
  try {
    return f();
  } catch (const std::exception& ex) {
    g_ranker_error = ex.what();
    return -1;
  }
}

struct Scratch {
  hipStream_t stream = nullptr;
  std::vector<void*> ptrs;
  ~Scratch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : ptrs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <typename T>
  T* Put(const T* h, size_t count) {
    T* d = beam::Alloc<T>(count);
    ptrs.push_back(d);
    if (h != nullptr) HIP_CHECK(hipMemcpyAsync(d, h, sizeof(T) * count, hipMemcpyHostToDevice, stream));
    return d;
  }
};

}  // namespace

extern "C" {

const char* beam_hip_ranker_last_error(void) { return g_ranker_error.c_str(); }

int beam_hip_ranker_forward(const float* params, const float* norm, const float* feats, int32_t n, float* out) {
  return RankerGuard([&] {
    if (n <= 0) return 0;
    Scratch s;
    HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    beam::RankerState r;
    r.Upload(params, norm, s.stream);
    const float* d_feats = s.Put(feats, static_cast<size_t>(n) * beam::kFeat);
    float* d_out = s.Put<float>(nullptr, static_cast<size_t>(n));
    r.Forward(d_feats, nullptr, n, d_out, false, s.stream);
    HIP_CHECK(hipMemcpyAsync(out, d_out, sizeof(float) * n, hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
  });
}

int beam_hip_ranker_gradient(const float* params, const float* norm, const float* feats, const float* target,
                             int32_t n, float* grad, float* loss) {
  return RankerGuard([&] {
    if (n <= 0) return 0;
    Scratch s;
    HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    beam::RankerState r;
    r.Upload(params, norm, s.stream);
    const float* d_feats = s.Put(feats, static_cast<size_t>(n) * beam::kFeat);
    const float* d_target = s.Put(target, static_cast<size_t>(n));
    r.Gradient(d_feats, d_target, nullptr, n, s.stream);
    HIP_CHECK(hipMemcpyAsync(grad, r.grad, sizeof(float) * beam::kParams, hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipMemcpyAsync(loss, r.loss, sizeof(float), hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
  });
}

int beam_hip_ranker_train(float* params, float* norm, const float* feats, const float* target, int32_t n,
                          int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss) {
  return RankerGuard([&] {
    if (n <= 0) return 0;
    Scratch s;
    HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    beam::RankerState r;
    r.Upload(params, norm, s.stream);
    const float* d_feats = s.Put(feats, static_cast<size_t>(n) * beam::kFeat);
    const float* d_target = s.Put(target, static_cast<size_t>(n));
    *loss = r.Train(d_feats, d_target, n, epochs, batch, lr, seed, s.stream);
    r.Download(params, norm, s.stream);
    return 0;
  });
}

int beam_hip_ranker_adam(float* params, const float* grad, float* m, float* v, int32_t step, float lr) {
  return RankerGuard([&] {
    Scratch s;
    HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    beam::RankerState r;
    std::vector<float> norm(beam::kNormCount, 0.0f);
    r.Upload(params, norm.data(), s.stream);
    HIP_CHECK(hipMemcpyAsync(r.m, m, sizeof(float) * beam::kParams, hipMemcpyHostToDevice, s.stream));
    HIP_CHECK(hipMemcpyAsync(r.v, v, sizeof(float) * beam::kParams, hipMemcpyHostToDevice, s.stream));
    HIP_CHECK(hipMemcpyAsync(r.grad, grad, sizeof(float) * beam::kParams, hipMemcpyHostToDevice, s.stream));
    r.step = step - 1;
    r.AdamStep(lr, s.stream);
    HIP_CHECK(hipMemcpyAsync(params, r.params, sizeof(float) * beam::kParams, hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipMemcpyAsync(m, r.m, sizeof(float) * beam::kParams, hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipMemcpyAsync(v, r.v, sizeof(float) * beam::kParams, hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
  });
}

}  // extern "C"
