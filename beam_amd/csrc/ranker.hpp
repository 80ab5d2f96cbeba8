// Learned pair ranker: a 24 -> 128 -> 128 -> 1 MLP (ReLU) that predicts a
// candidate pair's exact merge gain from cheap integer features, used to keep
// only the most promising candidates of each group for exact scoring.
//
// Parameters live in one FP32 vector (the Adam master copy) laid out as
//   W1[24][128] (in x out), b1[128], W2[128][128] (in x out), b2[128],
//   w3[128], b3
// and, for the MFMA forward, as BF16 transposes W1T[128][24] and W2T[128][128]
// (out x in).  Features are normalised with per-feature mean / inverse
// standard deviation taken from the training samples; the target is
// slog(gain) = sign(gain) * log1p(|gain|), standardised the same way.
#pragma once

#include "common.hpp"

namespace beam {

constexpr int kFeat = 24;
constexpr int kHid = 128;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kFeat * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kParams = kOffB3 + 1;

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// What the forward needs on the device.
struct RankerView {
  const unsigned short* w1t;  // BF16 [kHid][kFeat]
  const unsigned short* w2t;  // BF16 [kHid][kHid]
  const float* params;        // FP32 master (biases and w3 are read from here)
  const float* norm;          // [kFeat] mean, [kFeat] inverse std, then target mean, std
};

__device__ __forceinline__ unsigned short F32ToBf16(float f) {
  // Round to nearest even (inputs are finite).
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = u + 0x7fffU + ((u >> 16) & 1U);
  return static_cast<unsigned short>(r >> 16);
}

__device__ __forceinline__ float Bf16ToF32(unsigned short h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ float SLog(double x) {
  return static_cast<float>(x >= 0 ? log1p(x) : -log1p(-x));
}

// Forward of 32 pairs on one wave with v_mfma_f32_32x32x8_bf16.  Lane l
// carries pair (l % 32); xk[s * 4 + j] is that pair's normalised feature
// 8 s + 4 (l / 32) + j (s < 3).  Layers are computed transposed
// (H^T = W^T X^T), so a 32 x 32 accumulator tile is directly the B operand
// of the next layer: accumulator register 4 s' + j of hidden tile m holds
// hidden unit 32 m + 8 s' + 4 (l / 32) + j of pair l % 32.  Returns the
// prediction (standardised target units) in every lane of the pair; h1/h2,
// when given, receive the post-ReLU activations for training
// (h[(pair) * kHid + unit], pair = base + l % 32).
__device__ __forceinline__ float MlpForward32(const RankerView& w, const float* xk, float* h1, float* h2, int base,
                                              int npairs) {
  const int lane = LaneId();
  const int col = lane & 31, half = lane >> 5;
  bf16x4_t xb[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
#pragma unroll
    for (int j = 0; j < 4; ++j) xb[s][j] = static_cast<short>(F32ToBf16(xk[s * 4 + j]));
  }
  // Layer 1: H1^T[128 x 32] = W1T[128 x 24] X^T[24 x 32]; kept as packed
  // BF16 pairs (the next layer's operands) to hold register pressure down.
  uint32_t h1p[4][8];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    f32x16_t acc = {};
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const bf16x4_t wa =
          *reinterpret_cast<const bf16x4_t*>(w.w1t + (m * 32 + col) * kFeat + s * 8 + half * 4);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(wa, xb[s], acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i += 2) {
      const int unit = m * 32 + 8 * (i >> 2) + 4 * half + (i & 3);
      const float v0 = fmaxf(acc[i] + w.params[kOffB1 + unit], 0.0f);
      const float v1 = fmaxf(acc[i + 1] + w.params[kOffB1 + unit + 1], 0.0f);
      if (h1 != nullptr && col < npairs) {
        h1[static_cast<long long>(base + col) * kHid + unit] = v0;
        h1[static_cast<long long>(base + col) * kHid + unit + 1] = v1;
      }
      h1p[m][i >> 1] = static_cast<uint32_t>(F32ToBf16(v0)) | (static_cast<uint32_t>(F32ToBf16(v1)) << 16);
    }
  }
  // Layer 2: H2^T[128 x 32] = W2T[128 x 128] H1^T[128 x 32]; layer 3 folded
  // in.  The output tiles run one after another (not unrolled) so that only
  // one tile's operands are in flight.
  float y = 0.0f;
#pragma unroll 1
  for (int m = 0; m < 4; ++m) {
    f32x16_t acc = {};
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const bf16x4_t wa =
          *reinterpret_cast<const bf16x4_t*>(w.w2t + (m * 32 + col) * kHid + s * 8 + half * 4);
      const uint32_t lo = h1p[s >> 2][2 * (s & 3)], hi = h1p[s >> 2][2 * (s & 3) + 1];
      bf16x4_t hb;
      hb[0] = static_cast<short>(lo & 0xffffU);
      hb[1] = static_cast<short>(lo >> 16);
      hb[2] = static_cast<short>(hi & 0xffffU);
      hb[3] = static_cast<short>(hi >> 16);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(wa, hb, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int unit = m * 32 + 8 * (i >> 2) + 4 * half + (i & 3);
      const float v = fmaxf(acc[i] + w.params[kOffB2 + unit], 0.0f);
      if (h2 != nullptr && col < npairs) h2[static_cast<long long>(base + col) * kHid + unit] = v;
      y += v * w.params[kOffW3 + unit];
    }
  }
  y += __shfl_xor(y, 32, kWave);
  return y + w.params[kOffB3];
}

// Raw (unnormalised) features of a candidate pair.
struct PairFeatureInput {
  int64_t sa, sb;          // supernode sizes
  int64_t la, lb;          // cross entries (row lengths)
  int64_t ia, ib;          // internal edge counts
  int64_t aoc_a, aoc_b;    // approximate old costs (legacy proxy denominators)
  int64_t proxy;           // legacy EA proxy of the pair
  int64_t corr;            // common-neighbour correction of the exact gain
  int sig_match, sig_both; // one-permutation min-hash: equal bins, bins set in both
  int g;                   // group size
  double theta;            // threshold of the iteration
};

__device__ __forceinline__ float SLogf(float x) { return x >= 0.0f ? log1pf(x) : -log1pf(-x); }

// The 24 features (FP32): sizes, degrees, internal counts, approximate old
// costs, the legacy proxy and its ratio, the exact gain's common-neighbour
// correction, min-hash neighbour overlap, group size and threshold, and a few
// ratios of these.
__device__ __forceinline__ void PairFeatures(const PairFeatureInput& in, float* f) {
  const float sa = static_cast<float>(in.sa), sb = static_cast<float>(in.sb);
  const float la = static_cast<float>(in.la), lb = static_cast<float>(in.lb);
  const float ia = static_cast<float>(in.ia), ib = static_cast<float>(in.ib);
  const float aa = static_cast<float>(in.aoc_a), ab = static_cast<float>(in.aoc_b);
  const float px = static_cast<float>(in.proxy), cr = static_cast<float>(in.corr);
  const float denom = static_cast<float>(max(static_cast<int64_t>(1), in.aoc_a + in.aoc_b));
  const float jac = in.sig_both > 0 ? static_cast<float>(in.sig_match) / static_cast<float>(in.sig_both) : 0.0f;
  f[0] = SLogf(sa);
  f[1] = SLogf(sb);
  f[2] = SLogf(la);
  f[3] = SLogf(lb);
  f[4] = SLogf(ia);
  f[5] = SLogf(ib);
  f[6] = SLogf(aa);
  f[7] = SLogf(ab);
  f[8] = SLogf(px);
  f[9] = px / denom;
  f[10] = SLogf(cr);
  f[11] = jac;
  f[12] = SLogf(jac * fminf(la, lb));
  f[13] = SLogf(static_cast<float>(in.g));
  f[14] = static_cast<float>(in.theta);
  f[15] = SLogf(sa * sb);
  f[16] = la / fmaxf(1.0f, la + lb);
  f[17] = sa / (sa + sb);
  f[18] = SLogf(la + lb);
  f[19] = SLogf(static_cast<float>(in.proxy - in.corr));
  f[20] = ia / fmaxf(1.0f, sa * (sa - 1.0f) * 0.5f);
  f[21] = ib / fmaxf(1.0f, sb * (sb - 1.0f) * 0.5f);
  f[22] = SLogf(static_cast<float>(in.aoc_a + in.aoc_b));
  f[23] = px / fmaxf(1.0f, la + lb);
}

// One-permutation min-hash of a neighbour id (64 bins, 8-bit values).
__device__ __forceinline__ void SigSlot(int t, int* bin, unsigned* val) {
  const uint64_t h = SplitMix64(0x6a09e667f3bcc909ULL ^ static_cast<uint64_t>(static_cast<uint32_t>(t)));
  *bin = static_cast<int>(h & 63U);
  *val = static_cast<unsigned>(h >> 56);
}

}  // namespace beam
