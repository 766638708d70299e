// Learned-scheduler network: 12 -> 64 -> 64 -> 16, ReLU, softmax, FP32.
// Parameter layout (6032 floats): W1[64][12] b1[64] W2[64][64] b2[64]
// W3[16][64] b3[16], row-major (out, in).
#pragma once

#include <hip/hip_runtime.h>

namespace cg {

constexpr int kMlpIn = 12, kMlpH = 64, kMlpOut = 16;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;   // 768
constexpr int kOffW2 = kOffB1 + kMlpH;            // 832
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;    // 4928
constexpr int kOffW3 = kOffB2 + kMlpH;            // 4992
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH;  // 6016
constexpr int kMlpParams = kOffB3 + kMlpOut;      // 6032

// Forward pass of one feature vector by one workgroup of >= 64 threads.
// `sh` is LDS scratch of at least 2*64 + 16 floats; every thread of the block
// must call it (it synchronises).  On return probs[0..16) (in LDS) hold the
// softmax and *best the argmax (lowest index on ties).
__device__ inline void mlp_forward_block(const float* __restrict__ P, const float* x, float* sh,
                                         float* probs, int* best) {
  float* h1 = sh;
  float* h2 = sh + kMlpH;
  float* z = sh + 2 * kMlpH;
  const int t = threadIdx.x;
  if (t < kMlpH) {
    float a = P[kOffB1 + t];
    for (int i = 0; i < kMlpIn; ++i) a += P[kOffW1 + t * kMlpIn + i] * x[i];
    h1[t] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (t < kMlpH) {
    float a = P[kOffB2 + t];
    for (int i = 0; i < kMlpH; ++i) a += P[kOffW2 + t * kMlpH + i] * h1[i];
    h2[t] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (t < kMlpOut) {
    float a = P[kOffB3 + t];
    for (int i = 0; i < kMlpH; ++i) a += P[kOffW3 + t * kMlpH + i] * h2[i];
    z[t] = a;
  }
  __syncthreads();
  if (t == 0) {
    float m = z[0];
    int b = 0;
    for (int k = 1; k < kMlpOut; ++k)
      if (z[k] > m) { m = z[k]; b = k; }
    float s = 0.f;
    for (int k = 0; k < kMlpOut; ++k) { probs[k] = expf(z[k] - m); s += probs[k]; }
    for (int k = 0; k < kMlpOut; ++k) probs[k] /= s;
    *best = b;
  }
  __syncthreads();
}

// The same network with its weights held in LDS.  Rows are stored at an odd
// stride, so the lanes that each read one row at the same column hit
// distinct banks; every output is still one lane's sum over its inputs in
// order, so the results are those of mlp_forward_block bit for bit.
constexpr int kW1S = kMlpIn + 1, kW2S = kMlpH + 1;
constexpr int kLdsW1 = 0;
constexpr int kLdsB1 = kLdsW1 + kMlpH * kW1S;
constexpr int kLdsW2 = kLdsB1 + kMlpH;
constexpr int kLdsB2 = kLdsW2 + kMlpH * kW2S;
constexpr int kLdsW3 = kLdsB2 + kMlpH;
constexpr int kLdsB3 = kLdsW3 + kMlpOut * kW2S;
constexpr int kMlpLds = kLdsB3 + kMlpOut;  // 6176 floats

__device__ inline int mlp_lds_index(int j) {
  if (j < kOffB1) return kLdsW1 + (j / kMlpIn) * kW1S + j % kMlpIn;
  if (j < kOffW2) return kLdsB1 + (j - kOffB1);
  if (j < kOffB2) return kLdsW2 + ((j - kOffW2) / kMlpH) * kW2S + (j - kOffW2) % kMlpH;
  if (j < kOffW3) return kLdsB2 + (j - kOffB2);
  if (j < kOffB3) return kLdsW3 + ((j - kOffW3) / kMlpH) * kW2S + (j - kOffW3) % kMlpH;
  return kLdsB3 + (j - kOffB3);
}

// A workgroup of NT threads loads the parameters into registers (coalesced,
// all loads in flight at once) early, and stores them to LDS later, so the
// load latency hides behind whatever the workgroup does in between.
template <int NT>
struct MlpStage {
  static constexpr int kPer = (kMlpParams + NT - 1) / NT;
  float r[kPer];
  __device__ inline void load(const float* __restrict__ P) {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int j = (int)threadIdx.x + k * NT;
      r[k] = j < kMlpParams ? P[j] : 0.f;
    }
  }
  __device__ inline void store(float* w) const {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int j = (int)threadIdx.x + k * NT;
      if (j < kMlpParams) w[mlp_lds_index(j)] = r[k];
    }
  }
};

// mlp_forward_block over LDS weights w (after MlpStage::store and a barrier).
__device__ inline void mlp_forward_lds(const float* w, const float* x, float* sh, float* probs, int* best) {
  float* h1 = sh;
  float* h2 = sh + kMlpH;
  float* z = sh + 2 * kMlpH;
  const int t = threadIdx.x;
  if (t < kMlpH) {
    float a = w[kLdsB1 + t];
    for (int i = 0; i < kMlpIn; ++i) a += w[kLdsW1 + t * kW1S + i] * x[i];
    h1[t] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (t < kMlpH) {
    float a = w[kLdsB2 + t];
    for (int i = 0; i < kMlpH; ++i) a += w[kLdsW2 + t * kW2S + i] * h1[i];
    h2[t] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (t < kMlpOut) {
    float a = w[kLdsB3 + t];
    for (int i = 0; i < kMlpH; ++i) a += w[kLdsW3 + t * kW2S + i] * h2[i];
    z[t] = a;
  }
  __syncthreads();
  if (t == 0) {
    float m = z[0];
    int b = 0;
    for (int k = 1; k < kMlpOut; ++k)
      if (z[k] > m) { m = z[k]; b = k; }
    float s = 0.f;
    for (int k = 0; k < kMlpOut; ++k) { probs[k] = expf(z[k] - m); s += probs[k]; }
    for (int k = 0; k < kMlpOut; ++k) probs[k] /= s;
    *best = b;
  }
  __syncthreads();
}

}  // namespace cg
