// Learned scheduler: the network's weights, its training on the GPU, and the
// forward / gradient entry points the tests check against the CPU oracle.
//
// Network: 12 -> 64 -> 64 -> 16, ReLU, softmax (mlp_dev.hpp).  Inference runs
// inside k_stats (traverse.hip).  Training data are the engine's own probe
// traces: in CG_SCHED_PROBE mode every iteration times all 16 schedule
// choices from the same state, so each record carries the 12 features and 16
// measured times; the label is the fastest choice.  Training is full-batch
// cross-entropy with Adam; every reduction runs in a fixed order, so a given
// trace and seed always give the same weights.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <string>
#include <vector>

#include "engine.hpp"
#include "mlp_dev.hpp"

namespace cg {

namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 64;  // samples per workgroup in the gradient kernel
constexpr char kMagic[8] = {'C', 'G', 'M', 'L', 'P', '0', '0', '1'};

// Per-workgroup gradient of the summed cross-entropy over a chunk of samples.
// Each thread owns the parameters idx = t, t+256, ... so accumulation needs no
// atomics and runs in sample order.
__global__ __launch_bounds__(kBlock) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                     const int* __restrict__ Y, int n,
                                                     float* __restrict__ gpart, float* __restrict__ lpart) {
  __shared__ float x[kMlpIn], h1[kMlpH], h2[kMlpH], z[kMlpOut], dz[kMlpOut], dh1[kMlpH], dh2[kMlpH];
  __shared__ float loss_acc;
  const int t = threadIdx.x;
  float acc[(kMlpParams + kBlock - 1) / kBlock];
  for (int k = 0; k < (kMlpParams + kBlock - 1) / kBlock; ++k) acc[k] = 0.f;
  if (t == 0) loss_acc = 0.f;
  const int s0 = blockIdx.x * kChunk;
  const int s1 = min(n, s0 + kChunk);
  for (int smp = s0; smp < s1; ++smp) {
    if (t < kMlpIn) x[t] = X[(size_t)smp * kMlpIn + t];
    __syncthreads();
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
      for (int k = 1; k < kMlpOut; ++k) m = fmaxf(m, z[k]);
      float sum = 0.f;
      for (int k = 0; k < kMlpOut; ++k) sum += expf(z[k] - m);
      const int y = Y[smp];
      for (int k = 0; k < kMlpOut; ++k) dz[k] = expf(z[k] - m) / sum - (k == y ? 1.f : 0.f);
      loss_acc += (logf(sum) + m) - z[y];
    }
    __syncthreads();
    if (t < kMlpH) {
      float a = 0.f;
      for (int k = 0; k < kMlpOut; ++k) a += P[kOffW3 + k * kMlpH + t] * dz[k];
      dh2[t] = h2[t] > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (t < kMlpH) {
      float a = 0.f;
      for (int j = 0; j < kMlpH; ++j) a += P[kOffW2 + j * kMlpH + t] * dh2[j];
      dh1[t] = h1[t] > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int k = 0, idx = t; idx < kMlpParams; ++k, idx += kBlock) {
      float gv;
      if (idx < kOffB1) gv = dh1[idx / kMlpIn] * x[idx % kMlpIn];
      else if (idx < kOffW2) gv = dh1[idx - kOffB1];
      else if (idx < kOffB2) { const int r = idx - kOffW2; gv = dh2[r / kMlpH] * h1[r % kMlpH]; }
      else if (idx < kOffW3) gv = dh2[idx - kOffB2];
      else if (idx < kOffB3) { const int r = idx - kOffW3; gv = dz[r / kMlpH] * h2[r % kMlpH]; }
      else gv = dz[idx - kOffB3];
      acc[k] += gv;
    }
    __syncthreads();
  }
  for (int k = 0, idx = t; idx < kMlpParams; ++k, idx += kBlock)
    gpart[(size_t)blockIdx.x * kMlpParams + idx] = acc[k];
  if (t == 0) lpart[blockIdx.x] = loss_acc;
}

// Mean gradient (fixed block order); optionally one Adam step.
__global__ void k_mlp_reduce(float* __restrict__ P, const float* __restrict__ gpart,
                             const float* __restrict__ lpart, int nblocks, int n, float* __restrict__ grad_out,
                             float* __restrict__ m, float* __restrict__ v, float lr, int step, int do_update,
                             float* __restrict__ loss_out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx == 0) {
    float l = 0.f;
    for (int b = 0; b < nblocks; ++b) l += lpart[b];
    *loss_out = l / (float)n;
  }
  if (idx >= kMlpParams) return;
  float gsum = 0.f;
  for (int b = 0; b < nblocks; ++b) gsum += gpart[(size_t)b * kMlpParams + idx];
  const float gmean = gsum / (float)n;
  if (grad_out) grad_out[idx] = gmean;
  if (do_update) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[idx] + (1.f - b1) * gmean;
    const float vv = b2 * v[idx] + (1.f - b2) * gmean * gmean;
    m[idx] = mm;
    v[idx] = vv;
    const float mh = mm / (1.f - powf(b1, (float)step));
    const float vh = vv / (1.f - powf(b2, (float)step));
    P[idx] -= lr * mh / (sqrtf(vh) + eps);
  }
}

__global__ __launch_bounds__(kBlock) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                        int n, float* __restrict__ probs) {
  __shared__ float x[kMlpIn];
  __shared__ float sh[2 * kMlpH + kMlpOut];
  __shared__ float pr[kMlpOut];
  __shared__ int best;
  for (int smp = blockIdx.x; smp < n; smp += gridDim.x) {
    if (threadIdx.x < kMlpIn) x[threadIdx.x] = X[(size_t)smp * kMlpIn + threadIdx.x];
    __syncthreads();
    mlp_forward_block(P, x, sh, pr, &best);
    if (threadIdx.x < kMlpOut) probs[(size_t)smp * kMlpOut + threadIdx.x] = pr[threadIdx.x];
    __syncthreads();
  }
}

// Pulls every number out of a JSON array that follows `key` in `line`.
bool json_floats(const std::string& line, const char* key, std::vector<double>& out) {
  const size_t k = line.find(key);
  if (k == std::string::npos) return false;
  size_t p = line.find('[', k);
  const size_t e = line.find(']', p);
  if (p == std::string::npos || e == std::string::npos) return false;
  out.clear();
  const char* c = line.c_str() + p + 1;
  const char* end = line.c_str() + e;
  while (c < end) {
    char* nx = nullptr;
    const double v = std::strtod(c, &nx);
    if (nx == c) { ++c; continue; }
    out.push_back(v);
    c = nx;
  }
  return true;
}

// Relative margin within which two probe-timed choices count as equally fast.
constexpr double kLabelTolerance = 0.03;

// The number that follows `key` (e.g. "\"iter\":") in `line`; false when absent.
bool json_number(const std::string& line, const char* key, double* out) {
  const size_t k = line.find(key);
  if (k == std::string::npos) return false;
  const char* c = line.c_str() + k + std::strlen(key);
  char* nx = nullptr;
  const double v = std::strtod(c, &nx);
  if (nx == c) return false;
  *out = v;
  return true;
}

struct Gpu {
  DevBuf<float> P, X, gpart, lpart, grad, m, v, loss;
  DevBuf<int> Y;
};

int grad_pass(Gpu& d, int n, float lr, int step, int do_update, float* loss_host) {
  const int nblocks = (n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(kBlock), 0, 0, d.P.p, d.X.p, d.Y.p, n, d.gpart.p,
                     d.lpart.p);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, d.P.p,
                     d.gpart.p, d.lpart.p, nblocks, n, d.grad.p, d.m.p, d.v.p, lr, step, do_update, d.loss.p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(CG_ERR_HIP, "MLP kernels: %s", hipGetErrorString(e));
  if (loss_host) CG_HIP(hipMemcpy(loss_host, d.loss.p, sizeof(float), hipMemcpyDeviceToHost));
  return CG_OK;
}

int upload(Gpu& d, const float* params, const float* x, const int* y, int n) {
  const int nblocks = (n + kChunk - 1) / kChunk;
  CG_TRY(d.P.ensure(kMlpParams));
  CG_TRY(d.X.ensure((size_t)n * kMlpIn));
  CG_TRY(d.Y.ensure(n));
  CG_TRY(d.gpart.ensure((size_t)nblocks * kMlpParams));
  CG_TRY(d.lpart.ensure(nblocks));
  CG_TRY(d.grad.ensure(kMlpParams));
  CG_TRY(d.m.ensure(kMlpParams));
  CG_TRY(d.v.ensure(kMlpParams));
  CG_TRY(d.loss.ensure(1));
  CG_HIP(hipMemcpy(d.P.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  CG_HIP(hipMemcpy(d.X.p, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
  if (y) CG_HIP(hipMemcpy(d.Y.p, y, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  CG_HIP(hipMemset(d.m.p, 0, kMlpParams * sizeof(float)));
  CG_HIP(hipMemset(d.v.p, 0, kMlpParams * sizeof(float)));
  return CG_OK;
}

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// He-uniform weights, zero biases, from a counter-based stream.
void init_params(uint64_t seed, std::vector<float>& P) {
  P.assign(kMlpParams, 0.f);
  auto fill = [&](int off, int count, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; ++i) {
      const uint64_t r = mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)(off + i) + 1);
      const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
      P[off + i] = (float)((2.0 * u - 1.0) * a);
    }
  };
  fill(kOffW1, kMlpH * kMlpIn, kMlpIn);
  fill(kOffW2, kMlpH * kMlpH, kMlpH);
  fill(kOffW3, kMlpOut * kMlpH, kMlpH);
}

}  // namespace

int mlp_load(cg_engine* e) {
  if (e->mlp_valid && !e->mlp.empty()) return CG_OK;
  e->mlp_valid = false;
  if (!e->opt.mlp_path && !(e->opt.fault_flags & CG_FAULT_MLP_NAN)) return CG_OK;  // heuristic fallback
  e->mlp.assign(kMlpParams, 0.f);
  if (e->opt.fault_flags & CG_FAULT_MLP_NAN) {
    for (float& p : e->mlp) p = NAN;
  } else {
    FILE* f = std::fopen(e->opt.mlp_path, "rb");
    if (!f) return set_error(CG_ERR_IO, "cannot open scheduler weights %s", e->opt.mlp_path);
    char magic[8];
    const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
                    std::fread(e->mlp.data(), sizeof(float), kMlpParams, f) == (size_t)kMlpParams;
    std::fclose(f);
    if (!ok) return set_error(CG_ERR_IO, "%s is not a scheduler weight file", e->opt.mlp_path);
  }
  CG_TRY(e->d_mlp.ensure(kMlpParams));
  CG_HIP(hipMemcpy(e->d_mlp.p, e->mlp.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  e->mlp_valid = true;
  return CG_OK;
}

}  // namespace cg

using namespace cg;

extern "C" {

int cg_mlp_forward(const float* params, const float* x, int n, float* probs) {
  clear_error();
  if (!params || !x || !probs || n <= 0) return set_error(CG_ERR_INVALID, "bad MLP forward arguments");
  CG_TRY(require_device());
  Gpu d;
  CG_TRY(upload(d, params, x, nullptr, n));
  DevBuf<float> out;
  CG_TRY(out.ensure((size_t)n * kMlpOut));
  hipLaunchKernelGGL(k_mlp_forward, dim3(std::min(n, 4096)), dim3(kBlock), 0, 0, d.P.p, d.X.p, n, out.p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(CG_ERR_HIP, "MLP forward: %s", hipGetErrorString(e));
  CG_HIP(hipMemcpy(probs, out.p, (size_t)n * kMlpOut * sizeof(float), hipMemcpyDeviceToHost));
  return CG_OK;
}

int cg_mlp_loss_grad(const float* params, const float* x, const int* labels, int n, float* loss, float* grad) {
  clear_error();
  if (!params || !x || !labels || !loss || !grad || n <= 0)
    return set_error(CG_ERR_INVALID, "bad MLP gradient arguments");
  for (int i = 0; i < n; ++i)
    if (labels[i] < 0 || labels[i] >= kMlpOut) return set_error(CG_ERR_INVALID, "label %d out of range", labels[i]);
  CG_TRY(require_device());
  Gpu d;
  CG_TRY(upload(d, params, x, labels, n));
  CG_TRY(grad_pass(d, n, 0.f, 1, 0, loss));
  CG_HIP(hipMemcpy(grad, d.grad.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  return CG_OK;
}

int cg_mlp_train(const char* trace_path, const char* out_path, int epochs, float lr, uint64_t seed,
                 float* final_loss) {
  clear_error();
  if (!trace_path || !out_path || epochs <= 0 || !(lr > 0.f)) return set_error(CG_ERR_INVALID, "bad arguments");
  FILE* f = std::fopen(trace_path, "r");
  if (!f) return set_error(CG_ERR_IO, "cannot open trace %s", trace_path);
  // The network picks the schedule of the next iteration from the features
  // of the frontier it has just seen: a sample is the features of probe
  // record k and, as label, the fastest choice of record k + 1 of the same
  // run (same algorithm, graph and root, next iteration).
  std::vector<float> X;
  std::vector<int> Y;
  std::string line;
  std::vector<double> feat, us, prev_feat;
  double prev_key[5] = {-1, -1, -1, -1, -1};
  bool have_prev = false;
  char buf[8192];
  while (std::fgets(buf, sizeof(buf), f)) {
    line = buf;
    if (line.find("\"probe\"") == std::string::npos) continue;
    double key[5];
    if (!json_number(line, "\"alg\":", &key[0]) || !json_number(line, "\"V\":", &key[1]) ||
        !json_number(line, "\"E\":", &key[2]) || !json_number(line, "\"root\":", &key[3]) ||
        !json_number(line, "\"iter\":", &key[4]))
      continue;
    if (!json_floats(line, "\"features\"", feat) || feat.size() != kMlpIn) continue;
    if (!json_floats(line, "\"us_all\"", us) || us.size() != kMlpOut) continue;
    int best = -1;
    for (int k = 0; k < kMlpOut; ++k)
      if (us[k] > 0 && (best < 0 || us[k] < us[best])) best = k;
    // choices within kLabelTolerance of the fastest are the same within the
    // timing noise; the lowest-numbered one (the GPU rule's default tiers
    // come first) is the label, so that indifferent iterations agree
    if (best >= 0) {
      const double lim = us[best] * (1.0 + kLabelTolerance);
      for (int k = 0; k < kMlpOut; ++k)
        if (us[k] > 0 && us[k] <= lim) {
          best = k;
          break;
        }
    }
    const bool follows = have_prev && key[0] == prev_key[0] && key[1] == prev_key[1] && key[2] == prev_key[2] &&
                         key[3] == prev_key[3] && key[4] == prev_key[4] + 1;
    if (follows && best >= 0) {
      for (double v : prev_feat) X.push_back((float)v);
      Y.push_back(best);
    }
    prev_feat = feat;
    for (int k = 0; k < 5; ++k) prev_key[k] = key[k];
    have_prev = true;
  }
  std::fclose(f);
  const int n = (int)Y.size();
  if (n == 0) return set_error(CG_ERR_INVALID, "%s holds no probe records", trace_path);
  CG_TRY(require_device());
  std::vector<float> P;
  init_params(seed, P);
  Gpu d;
  CG_TRY(upload(d, P.data(), X.data(), Y.data(), n));
  float loss = 0.f;
  for (int ep = 1; ep <= epochs; ++ep) CG_TRY(grad_pass(d, n, lr, ep, 1, ep == epochs ? &loss : nullptr));
  // report the loss of the final weights
  CG_TRY(grad_pass(d, n, 0.f, 1, 0, &loss));
  CG_HIP(hipMemcpy(P.data(), d.P.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  FILE* o = std::fopen(out_path, "wb");
  if (!o) return set_error(CG_ERR_IO, "cannot create %s", out_path);
  const bool ok = std::fwrite(kMagic, 1, 8, o) == 8 &&
                  std::fwrite(P.data(), sizeof(float), kMlpParams, o) == (size_t)kMlpParams;
  std::fclose(o);
  if (!ok) return set_error(CG_ERR_IO, "short write to %s", out_path);
  if (final_loss) *final_loss = loss;
  return CG_OK;
}

}  // extern "C"
