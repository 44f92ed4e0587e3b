// skipnet.hip - the learned series-skip predictor: a 132-64-64-1 MLP (ReLU, FP32)
// with hand-written HIP forward, backward and Adam kernels, and its training
// loop.  It predicts, per tile, the first index at which the series approximation
// becomes unstable at the tile's probe point (as the scaled rank of that index
// among the candidate indices of the reference's thresholds); the render side
// validates every prediction with the reference's stability test and falls back
// to the plain probe search.
//
// Parameters, one flat FP32 array (kParams = 12737):
//   W1[64][132], b1[64], W2[64][64], b2[64], W3[64], b3
// Forward: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 . h2 + b3.
// Loss: mean over the batch of (y - t)^2.  Gradients are reduced over the batch by
// one thread per parameter in sample order (deterministic, no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "skipnet.hpp"

namespace nm {

namespace {

constexpr int kIn = SkipNet::kIn, kH = SkipNet::kHidden;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
static_assert(oB3 + 1 == SkipNet::kParams, "parameter layout");

// one wave per sample: lane j computes hidden unit j of both layers
__global__ __launch_bounds__(64) void k_forward(const float* __restrict__ X, const float* __restrict__ P, int n,
                                                float* __restrict__ H1, float* __restrict__ H2, float* __restrict__ Y) {
  const int b = blockIdx.x, j = threadIdx.x;
  if (b >= n) return;
  __shared__ float x[kIn];
  __shared__ float h1[kH];
  __shared__ float h2[kH];
  for (int k = j; k < kIn; k += 64) x[k] = X[(size_t)b * kIn + k];
  __syncthreads();
  float s = P[oB1 + j];
  for (int k = 0; k < kIn; k++) s += P[oW1 + j * kIn + k] * x[k];
  const float a1 = s > 0.0f ? s : 0.0f;
  h1[j] = a1;
  __syncthreads();
  float t = P[oB2 + j];
  for (int k = 0; k < kH; k++) t += P[oW2 + j * kH + k] * h1[k];
  const float a2 = t > 0.0f ? t : 0.0f;
  h2[j] = a2;
  __syncthreads();
  if (H1) H1[(size_t)b * kH + j] = a1;
  if (H2) H2[(size_t)b * kH + j] = a2;
  if (j == 0) {
    float y = P[oB3];
    for (int k = 0; k < kH; k++) y += P[oW3 + k] * h2[k];
    Y[b] = y;
  }
}

// per-sample deltas: dY = 2 (y - t) / n, dH2 = dY W3 (h2 > 0), dH1 = W2^T dH2 (h1 > 0)
__global__ __launch_bounds__(64) void k_deltas(const float* __restrict__ P, const float* __restrict__ H1,
                                               const float* __restrict__ H2, const float* __restrict__ Y,
                                               const float* __restrict__ T, int n, float* __restrict__ DY,
                                               float* __restrict__ D2, float* __restrict__ D1) {
  const int b = blockIdx.x, j = threadIdx.x;
  if (b >= n) return;
  __shared__ float d2[kH];
  const float dy = 2.0f * (Y[b] - T[b]) / (float)n;
  const float g2 = H2[(size_t)b * kH + j] > 0.0f ? dy * P[oW3 + j] : 0.0f;
  d2[j] = g2;
  __syncthreads();
  float s = 0.0f;
  for (int i = 0; i < kH; i++) s += P[oW2 + i * kH + j] * d2[i];
  const float g1 = H1[(size_t)b * kH + j] > 0.0f ? s : 0.0f;
  D2[(size_t)b * kH + j] = g2;
  D1[(size_t)b * kH + j] = g1;
  if (j == 0) DY[b] = dy;
}

// one thread per parameter: its gradient summed over the batch in sample order
__global__ __launch_bounds__(256) void k_param_grads(const float* __restrict__ X, const float* __restrict__ H1,
                                                     const float* __restrict__ H2, const float* __restrict__ DY,
                                                     const float* __restrict__ D2, const float* __restrict__ D1,
                                                     int n, float* __restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= SkipNet::kParams) return;
  float g = 0.0f;
  if (p < oB1) {
    const int j = (p - oW1) / kIn, k = (p - oW1) % kIn;
    for (int b = 0; b < n; b++) g += D1[(size_t)b * kH + j] * X[(size_t)b * kIn + k];
  } else if (p < oW2) {
    const int j = p - oB1;
    for (int b = 0; b < n; b++) g += D1[(size_t)b * kH + j];
  } else if (p < oB2) {
    const int i = (p - oW2) / kH, j = (p - oW2) % kH;
    for (int b = 0; b < n; b++) g += D2[(size_t)b * kH + i] * H1[(size_t)b * kH + j];
  } else if (p < oW3) {
    const int i = p - oB2;
    for (int b = 0; b < n; b++) g += D2[(size_t)b * kH + i];
  } else if (p < oB3) {
    const int k = p - oW3;
    for (int b = 0; b < n; b++) g += DY[b] * H2[(size_t)b * kH + k];
  } else {
    for (int b = 0; b < n; b++) g += DY[b];
  }
  G[p] = g;
}

// mean squared error of a batch (one workgroup)
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int n,
                                             float* __restrict__ out) {
  __shared__ float acc[256];
  float s = 0.0f;
  for (int b = threadIdx.x; b < n; b += blockDim.x) {
    const float d = Y[b] - T[b];
    s += d * d;
  }
  acc[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) acc[threadIdx.x] += acc[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = acc[0] / (float)n;
}

// Adam (Kingma & Ba, 2015) with bias correction; c1 = 1 - beta1^t, c2 = 1 - beta2^t
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ G,
                                              float* __restrict__ M, float* __restrict__ V, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= SkipNet::kParams) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.0f - b1) * g;
  const float v = b2 * V[p] + (1.0f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// rows of X selected by idx -> a contiguous batch
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ X, const float* __restrict__ T,
                                                const int* __restrict__ idx, int n, float* __restrict__ XB,
                                                float* __restrict__ TB) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kIn) return;
  const int b = i / kIn, k = i % kIn;
  const int src = idx[b];
  XB[i] = X[(size_t)src * kIn + k];
  if (k == 0) TB[b] = T[src];
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

#define SN_HIP(expr)                                                    \
  do {                                                                  \
    hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess) {                                             \
      err_ = std::string(#expr) + ": " + hipGetErrorString(e_);         \
      return NM_ERR_HIP;                                                \
    }                                                                   \
  } while (0)

SkipNet::SkipNet() : params(kParams, 0.0f), mean(kIn, 0.0f), scale(kIn, 1.0f) {}

SkipNet::~SkipNet() { release(); }

void SkipNet::release() {
  for (void* p : {(void*)dP, (void*)dG, (void*)dM, (void*)dV, (void*)dX, (void*)dT, (void*)dXB, (void*)dTB,
                  (void*)dH1, (void*)dH2, (void*)dY, (void*)dDY, (void*)dD1, (void*)dD2, (void*)dIdx, (void*)dLoss})
    if (p) (void)hipFree(p);
  dP = dG = dM = dV = dX = dT = dXB = dTB = dH1 = dH2 = dY = dDY = dD1 = dD2 = dLoss = nullptr;
  dIdx = nullptr;
  capX = capB = 0;
}

// He-uniform initialisation from a seed (deterministic)
void SkipNet::init(uint64_t seed) {
  uint64_t s = seed;
  auto uni = [&](float a) { return (float)(((splitmix(s) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * a; };
  const float a1 = std::sqrt(6.0f / kIn), a2 = std::sqrt(6.0f / kH);
  for (int i = 0; i < kH * kIn; i++) params[oW1 + i] = uni(a1);
  for (int i = 0; i < kH; i++) params[oB1 + i] = 0.0f;
  for (int i = 0; i < kH * kH; i++) params[oW2 + i] = uni(a2);
  for (int i = 0; i < kH; i++) params[oB2 + i] = 0.0f;
  for (int i = 0; i < kH; i++) params[oW3 + i] = uni(a2);
  params[oB3] = 0.0f;
}

int SkipNet::reserve(size_t nx, size_t nb) {
  if (!dP) {
    SN_HIP(hipMalloc((void**)&dP, kParams * 4));
    SN_HIP(hipMalloc((void**)&dG, kParams * 4));
    SN_HIP(hipMalloc((void**)&dM, kParams * 4));
    SN_HIP(hipMalloc((void**)&dV, kParams * 4));
    SN_HIP(hipMalloc((void**)&dLoss, 4));
  }
  if (nx > capX) {
    if (dX) SN_HIP(hipFree(dX));
    if (dT) SN_HIP(hipFree(dT));
    dX = dT = nullptr;
    SN_HIP(hipMalloc((void**)&dX, nx * kIn * 4));
    SN_HIP(hipMalloc((void**)&dT, nx * 4));
    capX = nx;
  }
  if (nb > capB) {
    for (float** p : {&dXB, &dTB, &dH1, &dH2, &dY, &dDY, &dD1, &dD2})
      if (*p) SN_HIP(hipFree(*p));
    if (dIdx) SN_HIP(hipFree(dIdx));
    SN_HIP(hipMalloc((void**)&dXB, nb * kIn * 4));
    SN_HIP(hipMalloc((void**)&dTB, nb * 4));
    SN_HIP(hipMalloc((void**)&dH1, nb * kH * 4));
    SN_HIP(hipMalloc((void**)&dH2, nb * kH * 4));
    SN_HIP(hipMalloc((void**)&dY, nb * 4));
    SN_HIP(hipMalloc((void**)&dDY, nb * 4));
    SN_HIP(hipMalloc((void**)&dD1, nb * kH * 4));
    SN_HIP(hipMalloc((void**)&dD2, nb * kH * 4));
    SN_HIP(hipMalloc((void**)&dIdx, nb * 4));
    capB = nb;
  }
  return NM_OK;
}

void SkipNet::normalise(const float* raw, size_t n, std::vector<float>& out) const {
  out.resize(n * kIn);
  for (size_t b = 0; b < n; b++)
    for (int k = 0; k < kIn; k++) out[b * kIn + k] = (raw[b * kIn + k] - mean[k]) * scale[k];
}

void SkipNet::fit_normaliser(const float* raw, size_t n) {
  for (int k = 0; k < kIn; k++) {
    double s = 0, s2 = 0;
    for (size_t b = 0; b < n; b++) {
      const double v = raw[b * kIn + k];
      s += v;
      s2 += v * v;
    }
    const double mu = s / (double)n, var = std::max(0.0, s2 / (double)n - mu * mu);
    mean[k] = (float)mu;
    scale[k] = var > 1e-12 ? (float)(1.0 / std::sqrt(var)) : 1.0f;
  }
}

int SkipNet::predict(const float* raw, size_t n, float* y) {
  if (n == 0) return NM_OK;
  int st = reserve(n, n);
  if (st) return st;
  std::vector<float> xn;
  normalise(raw, n, xn);
  SN_HIP(hipMemcpy(dP, params.data(), kParams * 4, hipMemcpyHostToDevice));
  SN_HIP(hipMemcpy(dXB, xn.data(), n * kIn * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_forward, dim3((unsigned)n), dim3(64), 0, 0, (const float*)dXB, (const float*)dP, (int)n,
                     (float*)nullptr, (float*)nullptr, dY);
  SN_HIP(hipGetLastError());
  SN_HIP(hipMemcpy(y, dY, n * 4, hipMemcpyDeviceToHost));
  return NM_OK;
}

// loss and gradient of one batch (inputs already normalised), for checks
int SkipNet::gradient(const float* xn, const float* t, size_t n, float* grad, float* loss) {
  int st = reserve(n, n);
  if (st) return st;
  SN_HIP(hipMemcpy(dP, params.data(), kParams * 4, hipMemcpyHostToDevice));
  SN_HIP(hipMemcpy(dXB, xn, n * kIn * 4, hipMemcpyHostToDevice));
  SN_HIP(hipMemcpy(dTB, t, n * 4, hipMemcpyHostToDevice));
  st = batch_grad((int)n);
  if (st) return st;
  SN_HIP(hipMemcpy(grad, dG, kParams * 4, hipMemcpyDeviceToHost));
  if (loss) SN_HIP(hipMemcpy(loss, dLoss, 4, hipMemcpyDeviceToHost));
  return NM_OK;
}

// forward, loss, deltas and parameter gradients of the batch in dXB / dTB
int SkipNet::batch_grad(int n) {
  hipLaunchKernelGGL(k_forward, dim3(n), dim3(64), 0, 0, (const float*)dXB, (const float*)dP, n, dH1, dH2, dY);
  SN_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, 0, (const float*)dY, (const float*)dTB, n, dLoss);
  SN_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_deltas, dim3(n), dim3(64), 0, 0, (const float*)dP, (const float*)dH1, (const float*)dH2,
                     (const float*)dY, (const float*)dTB, n, dDY, dD2, dD1);
  SN_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_param_grads, dim3((kParams + 255) / 256), dim3(256), 0, 0, (const float*)dXB,
                     (const float*)dH1, (const float*)dH2, (const float*)dDY, (const float*)dD2, (const float*)dD1, n,
                     dG);
  SN_HIP(hipGetLastError());
  return NM_OK;
}

int SkipNet::train(const float* raw, const float* t, size_t n, const TrainOpts& o, std::vector<float>* curve) {
  if (n == 0 || o.batch <= 0 || o.epochs <= 0) return NM_ERR_ARG;
  fit_normaliser(raw, n);
  init(o.seed);
  const size_t nb = std::min<size_t>((size_t)o.batch, n);
  int st = reserve(n, nb);
  if (st) return st;
  std::vector<float> xn;
  normalise(raw, n, xn);
  SN_HIP(hipMemcpy(dX, xn.data(), n * kIn * 4, hipMemcpyHostToDevice));
  SN_HIP(hipMemcpy(dT, t, n * 4, hipMemcpyHostToDevice));
  SN_HIP(hipMemcpy(dP, params.data(), kParams * 4, hipMemcpyHostToDevice));
  SN_HIP(hipMemset(dM, 0, kParams * 4));
  SN_HIP(hipMemset(dV, 0, kParams * 4));
  std::vector<int> perm(n);
  for (size_t i = 0; i < n; i++) perm[i] = (int)i;
  uint64_t rs = o.seed ^ 0x5DEECE66Dull;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double p1 = 1.0, p2 = 1.0;
  if (curve) curve->clear();
  for (int ep = 0; ep < o.epochs; ep++) {
    for (size_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[splitmix(rs) % (i + 1)]);  // Fisher-Yates
    const float lr = o.lr * (ep < o.epochs / 2 ? 1.0f : (ep < 3 * o.epochs / 4 ? 0.3f : 0.1f));
    double ep_loss = 0.0;
    int nbatches = 0;
    for (size_t s = 0; s + nb <= n; s += nb) {
      SN_HIP(hipMemcpyAsync(dIdx, perm.data() + s, nb * 4, hipMemcpyHostToDevice, 0));
      hipLaunchKernelGGL(k_gather, dim3((unsigned)((nb * kIn + 255) / 256)), dim3(256), 0, 0, (const float*)dX,
                         (const float*)dT, (const int*)dIdx, (int)nb, dXB, dTB);
      SN_HIP(hipGetLastError());
      if ((st = batch_grad((int)nb))) return st;
      p1 *= b1;
      p2 *= b2;
      hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, 0, dP, (const float*)dG, dM, dV, lr, b1,
                         b2, eps, (float)(1.0 - p1), (float)(1.0 - p2));
      SN_HIP(hipGetLastError());
      if (curve && s == 0) {  // first batch loss of each epoch (before the step)
        float l = 0.0f;
        SN_HIP(hipMemcpy(&l, dLoss, 4, hipMemcpyDeviceToHost));
        ep_loss = l;
      }
      nbatches++;
    }
    if (curve) curve->push_back((float)ep_loss);
  }
  SN_HIP(hipMemcpy(params.data(), dP, kParams * 4, hipMemcpyDeviceToHost));
  trained = true;
  version++;
  return NM_OK;
}

}  // namespace nm
