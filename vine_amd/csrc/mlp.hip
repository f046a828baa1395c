// Training and batch inference of the bucket-width MLP on the GPU.
//
// k_mlp_grad: one workgroup per tile of 64 samples.  Each lane runs the
// forward and backward pass of its sample (activations and deltas to LDS),
// then the workgroup reduces the tile's parameter gradients out of LDS
// (dW = sum over samples of delta x activation^T) and adds them to the global
// accumulator with one atomic per parameter per tile.  k_mlp_adam applies an
// Adam step.  Loss: 0.5 * sum_k mask_k * (y_k - target_k)^2, averaged over the
// minibatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mlp.hpp"
#include "vine.h"

using namespace vine::mlp;

namespace {

constexpr int TILE = 64;

thread_local std::string g_mlp_err;

__global__ void __launch_bounds__(TILE) k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                   const float* __restrict__ T, const float* __restrict__ M,
                                                   const int32_t* __restrict__ idx, long long n,
                                                   float* __restrict__ grad, double* __restrict__ loss) {
  __shared__ float s_x[TILE][NIN];
  __shared__ float s_h1[TILE][NH];
  __shared__ float s_h2[TILE][NH];
  __shared__ float s_dy[TILE][NOUT];
  __shared__ float s_d2[TILE][NH];
  __shared__ float s_d1[TILE][NH];
  __shared__ float s_loss[TILE];
  const int t = threadIdx.x;
  const long long s = (long long)blockIdx.x * TILE + t;
  const bool valid = s < n;
  const long long row = valid ? (idx ? (long long)idx[s] : s) : 0;
  float xn[NIN], h1[NH], h2[NH], y[NOUT], dy[NOUT], d2[NH], d1[NH];
#pragma unroll
  for (int j = 0; j < NIN; j++) xn[j] = valid ? (X[row * NIN + j] - P[OFF_MU + j]) / P[OFF_SD + j] : 0.f;
#pragma unroll
  for (int i = 0; i < NH; i++) {
    float a = P[OFF_B1 + i];
#pragma unroll
    for (int j = 0; j < NIN; j++) a += P[OFF_W1 + i * NIN + j] * xn[j];
    h1[i] = fmaxf(a, 0.f);
  }
  for (int i = 0; i < NH; i++) {
    float a = P[OFF_B2 + i];
#pragma unroll
    for (int j = 0; j < NH; j++) a += P[OFF_W2 + i * NH + j] * h1[j];
    h2[i] = fmaxf(a, 0.f);
  }
  float l = 0.f;
#pragma unroll
  for (int k = 0; k < NOUT; k++) {
    float a = P[OFF_B3 + k];
#pragma unroll
    for (int j = 0; j < NH; j++) a += P[OFF_W3 + k * NH + j] * h2[j];
    y[k] = a;
    float m = valid ? M[row * NOUT + k] : 0.f;
    float e = valid ? (a - T[row * NOUT + k]) : 0.f;
    dy[k] = m * e;
    l += 0.5f * m * e * e;
  }
  for (int j = 0; j < NH; j++) {
    float a = 0.f;
#pragma unroll
    for (int k = 0; k < NOUT; k++) a += P[OFF_W3 + k * NH + j] * dy[k];
    d2[j] = h2[j] > 0.f ? a : 0.f;
  }
  for (int j = 0; j < NH; j++) {
    float a = 0.f;
    for (int i = 0; i < NH; i++) a += P[OFF_W2 + i * NH + j] * d2[i];
    d1[j] = h1[j] > 0.f ? a : 0.f;
  }
#pragma unroll
  for (int j = 0; j < NIN; j++) s_x[t][j] = xn[j];
  for (int j = 0; j < NH; j++) {
    s_h1[t][j] = h1[j];
    s_h2[t][j] = h2[j];
    s_d2[t][j] = d2[j];
    s_d1[t][j] = d1[j];
  }
#pragma unroll
  for (int k = 0; k < NOUT; k++) s_dy[t][k] = dy[k];
  s_loss[t] = l;
  __syncthreads();
  const int ns = (int)min((long long)TILE, n - (long long)blockIdx.x * TILE);
  for (int e = t; e < NTRAIN; e += TILE) {
    float g = 0.f;
    if (e < OFF_B1) {
      int i = e / NIN, j = e % NIN;
      for (int q = 0; q < ns; q++) g += s_d1[q][i] * s_x[q][j];
    } else if (e < OFF_W2) {
      int i = e - OFF_B1;
      for (int q = 0; q < ns; q++) g += s_d1[q][i];
    } else if (e < OFF_B2) {
      int i = (e - OFF_W2) / NH, j = (e - OFF_W2) % NH;
      for (int q = 0; q < ns; q++) g += s_d2[q][i] * s_h1[q][j];
    } else if (e < OFF_W3) {
      int i = e - OFF_B2;
      for (int q = 0; q < ns; q++) g += s_d2[q][i];
    } else if (e < OFF_B3) {
      int k = (e - OFF_W3) / NH, j = (e - OFF_W3) % NH;
      for (int q = 0; q < ns; q++) g += s_dy[q][k] * s_h2[q][j];
    } else {
      int k = e - OFF_B3;
      for (int q = 0; q < ns; q++) g += s_dy[q][k];
    }
    atomicAdd(&grad[e], g);
  }
  if (t == 0 && loss) {
    double acc = 0.0;
    for (int q = 0; q < ns; q++) acc += s_loss[q];
    atomicAdd(loss, acc);
  }
}

__global__ void k_mlp_adam(float* P, float* grad, float* m, float* v, float lr, float scale, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NTRAIN) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float g = grad[i] * scale;
  grad[i] = 0.f;
  float mi = b1 * m[i] + (1.f - b1) * g;
  float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / (1.f - powf(b1, (float)step));
  float vh = vi / (1.f - powf(b2, (float)step));
  P[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X, long long n, float* Y) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[NIN], y[NOUT];
#pragma unroll
  for (int j = 0; j < NIN; j++) x[j] = X[s * NIN + j];
  vine::mlp::apply(P, x, y);
#pragma unroll
  for (int k = 0; k < NOUT; k++) Y[s * NOUT + k] = y[k];
}

#define MLP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      g_mlp_err = std::string(#expr) + ": " + hipGetErrorString(e_);                       \
      return VINE_E_HIP;                                                                   \
    }                                                                                      \
  } while (0)

}  // namespace

struct vine_mlp {
  float* d_p = nullptr;
  float* d_grad = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  double* d_loss = nullptr;
  int step = 0;
};

extern "C" {

int32_t vine_mlp_param_count(void) { return NPARAM; }

const char* vine_mlp_last_error(void) { return g_mlp_err.c_str(); }

vine_status vine_mlp_destroy(vine_mlp* m) {
  if (!m) return VINE_OK;
  for (void* p : {(void*)m->d_p, (void*)m->d_grad, (void*)m->d_m, (void*)m->d_v, (void*)m->d_loss})
    if (p) (void)hipFree(p);
  delete m;
  return VINE_OK;
}

vine_status vine_mlp_create(vine_mlp** out) {
  if (!out) return VINE_E_ARG;
  auto* m = new vine_mlp();
  hipError_t e = hipMalloc(&m->d_p, sizeof(float) * NPARAM);
  if (e == hipSuccess) e = hipMalloc(&m->d_grad, sizeof(float) * NTRAIN);
  if (e == hipSuccess) e = hipMalloc(&m->d_m, sizeof(float) * NTRAIN);
  if (e == hipSuccess) e = hipMalloc(&m->d_v, sizeof(float) * NTRAIN);
  if (e == hipSuccess) e = hipMalloc(&m->d_loss, sizeof(double));
  if (e == hipSuccess) e = hipMemset(m->d_grad, 0, sizeof(float) * NTRAIN);
  if (e == hipSuccess) e = hipMemset(m->d_m, 0, sizeof(float) * NTRAIN);
  if (e == hipSuccess) e = hipMemset(m->d_v, 0, sizeof(float) * NTRAIN);
  if (e != hipSuccess) {
    g_mlp_err = hipGetErrorString(e);
    vine_mlp_destroy(m);
    return VINE_E_OOM;
  }
  *out = m;
  return VINE_OK;
}

vine_status vine_mlp_set_params(vine_mlp* m, const float* params) {
  if (!m || !params) return VINE_E_ARG;
  for (int i = 0; i < NPARAM; i++)
    if (!std::isfinite(params[i]) || (i >= OFF_SD && !(params[i] > 0.f))) {
      g_mlp_err = "parameters must be finite, standard deviations positive";
      return VINE_E_ARG;
    }
  MLP_TRY(hipMemcpy(m->d_p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  MLP_TRY(hipMemset(m->d_m, 0, sizeof(float) * NTRAIN));
  MLP_TRY(hipMemset(m->d_v, 0, sizeof(float) * NTRAIN));
  m->step = 0;
  return VINE_OK;
}

vine_status vine_mlp_get_params(const vine_mlp* m, float* params) {
  if (!m || !params) return VINE_E_ARG;
  MLP_TRY(hipMemcpy(params, m->d_p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return VINE_OK;
}

vine_status vine_mlp_forward(const vine_mlp* m, const float* x, int64_t n, float* y) {
  if (!m || n < 0 || (n > 0 && (!x || !y))) return VINE_E_ARG;
  if (n == 0) return VINE_OK;
  float *dx = nullptr, *dy = nullptr;
  MLP_TRY(hipMalloc(&dx, sizeof(float) * NIN * n));
  hipError_t e = hipMalloc(&dy, sizeof(float) * NOUT * n);
  if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(float) * NIN * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, m->d_p, dx, (long long)n, dy);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(y, dy, sizeof(float) * NOUT * n, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (e != hipSuccess) {
    g_mlp_err = hipGetErrorString(e);
    return VINE_E_HIP;
  }
  return VINE_OK;
}

// Mean loss and mean gradient over n samples (for checking and small sets).
vine_status vine_mlp_gradient(const vine_mlp* m, const float* x, const float* target, const float* mask,
                              int64_t n, float* grad, double* loss) {
  if (!m || n <= 0 || !x || !target || !mask || !grad) return VINE_E_ARG;
  float *dx = nullptr, *dt = nullptr, *dm = nullptr, *dg = nullptr;
  double* dl = nullptr;
  hipError_t e = hipMalloc(&dx, sizeof(float) * NIN * n);
  if (e == hipSuccess) e = hipMalloc(&dt, sizeof(float) * NOUT * n);
  if (e == hipSuccess) e = hipMalloc(&dm, sizeof(float) * NOUT * n);
  if (e == hipSuccess) e = hipMalloc(&dg, sizeof(float) * NTRAIN);
  if (e == hipSuccess) e = hipMalloc(&dl, sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(float) * NIN * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dt, target, sizeof(float) * NOUT * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dm, mask, sizeof(float) * NOUT * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dg, 0, sizeof(float) * NTRAIN);
  if (e == hipSuccess) e = hipMemset(dl, 0, sizeof(double));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)((n + TILE - 1) / TILE)), dim3(TILE), 0, 0, m->d_p, dx, dt, dm,
                       nullptr, (long long)n, dg, dl);
    e = hipGetLastError();
  }
  std::vector<float> g(NTRAIN);
  double l = 0.0;
  if (e == hipSuccess) e = hipMemcpy(g.data(), dg, sizeof(float) * NTRAIN, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&l, dl, sizeof(double), hipMemcpyDeviceToHost);
  for (void* p : {(void*)dx, (void*)dt, (void*)dm, (void*)dg, (void*)dl})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) {
    g_mlp_err = hipGetErrorString(e);
    return VINE_E_HIP;
  }
  for (int i = 0; i < NTRAIN; i++) grad[i] = g[i] / (float)n;
  if (loss) *loss = l / (double)n;
  return VINE_OK;
}

// Minibatch Adam on the GPU; data uploaded once, shuffled every epoch.
vine_status vine_mlp_train(vine_mlp* m, const float* x, const float* target, const float* mask, int64_t n,
                           int32_t epochs, int32_t batch, float lr, uint64_t seed, double* loss_curve) {
  if (!m || n <= 0 || !x || !target || !mask || epochs < 1 || batch < 1 || !(lr > 0.f)) return VINE_E_ARG;
  float *dx = nullptr, *dt = nullptr, *dm = nullptr;
  int32_t* di = nullptr;
  hipError_t e = hipMalloc(&dx, sizeof(float) * NIN * n);
  if (e == hipSuccess) e = hipMalloc(&dt, sizeof(float) * NOUT * n);
  if (e == hipSuccess) e = hipMalloc(&dm, sizeof(float) * NOUT * n);
  if (e == hipSuccess) e = hipMalloc(&di, sizeof(int32_t) * n);
  if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(float) * NIN * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dt, target, sizeof(float) * NOUT * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dm, mask, sizeof(float) * NOUT * n, hipMemcpyHostToDevice);
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  std::mt19937_64 rng(seed);
  for (int ep = 0; ep < epochs && e == hipSuccess; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    e = hipMemcpy(di, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->d_loss, 0, sizeof(double));
    for (int64_t b0 = 0; b0 < n && e == hipSuccess; b0 += batch) {
      long long nb = std::min<long long>(batch, n - b0);
      hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)((nb + TILE - 1) / TILE)), dim3(TILE), 0, 0, m->d_p, dx, dt, dm,
                         di + b0, nb, m->d_grad, m->d_loss);
      e = hipGetLastError();
      if (e != hipSuccess) break;
      m->step++;
      hipLaunchKernelGGL(k_mlp_adam, dim3((NTRAIN + 255) / 256), dim3(256), 0, 0, m->d_p, m->d_grad, m->d_m, m->d_v,
                         lr, 1.0f / (float)nb, m->step);
      e = hipGetLastError();
    }
    double l = 0.0;
    if (e == hipSuccess) e = hipMemcpy(&l, m->d_loss, sizeof(double), hipMemcpyDeviceToHost);
    if (loss_curve) loss_curve[ep] = l / (double)n;
  }
  for (void* p : {(void*)dx, (void*)dt, (void*)dm, (void*)di})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) {
    g_mlp_err = hipGetErrorString(e);
    return VINE_E_HIP;
  }
  return VINE_OK;
}

}  // extern "C"
