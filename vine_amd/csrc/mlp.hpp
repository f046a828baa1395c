// Bucket-width MLP: 8 inputs -> 32 -> 32 -> 2 outputs, fp32, ReLU.
//
// Replaces the reference's AR(3) frontier-size model (include/AR.hpp,
// algorithm/cspp.cuh:254-284) as the scheduler of the window width.  Output 0
// is log2 of the next window width (key units), output 1 a forecast of
// log2(1 + expansions) a few rounds ahead (the AR model's forecasting role).
// Inputs are normalised with stored per-feature mean and standard deviation.
//
// Parameter vector (floats, row-major W[out][in]):
//   W1[32*8] b1[32] W2[32*32] b2[32] W3[2*32] b3[2] xmu[8] xsd[8]
#pragma once

#include <hip/hip_runtime.h>

namespace vine {
namespace mlp {

constexpr int NIN = 8;
constexpr int NH = 32;
constexpr int NOUT = 2;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * NIN;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NOUT * NH;
constexpr int NTRAIN = OFF_B3 + NOUT;  // 1,410 trained parameters
constexpr int OFF_MU = NTRAIN;
constexpr int OFF_SD = OFF_MU + NIN;
constexpr int NPARAM = OFF_SD + NIN;  // 1,426 with the normalisation

// Forward pass for one sample by one lane.
__device__ __forceinline__ void apply(const float* __restrict__ P, const float* x, float* y) {
  float xn[NIN];
#pragma unroll
  for (int j = 0; j < NIN; j++) xn[j] = (x[j] - P[OFF_MU + j]) / P[OFF_SD + j];
  float h1[NH];
#pragma unroll
  for (int i = 0; i < NH; i++) {
    float a = P[OFF_B1 + i];
#pragma unroll
    for (int j = 0; j < NIN; j++) a += P[OFF_W1 + i * NIN + j] * xn[j];
    h1[i] = fmaxf(a, 0.f);
  }
  float h2[NH];
#pragma unroll 4
  for (int i = 0; i < NH; i++) {
    float a = P[OFF_B2 + i];
#pragma unroll
    for (int j = 0; j < NH; j++) a += P[OFF_W2 + i * NH + j] * h1[j];
    h2[i] = fmaxf(a, 0.f);
  }
#pragma unroll
  for (int k = 0; k < NOUT; k++) {
    float a = P[OFF_B3 + k];
#pragma unroll
    for (int j = 0; j < NH; j++) a += P[OFF_W3 + k * NH + j] * h2[j];
    y[k] = a;
  }
}

// The same network evaluated by a whole wave: lane i (mod NH) computes
// hidden unit i, so each lane keeps a handful of registers.  Every lane passes
// the same x and receives y.  Sums run in another order than apply(), so the
// outputs agree with it to rounding, not bit for bit.
__device__ __forceinline__ void apply_wave(const float* __restrict__ P, const float* x, float* y) {
  const int lane = (int)__lane_id();
  const int i = lane & (NH - 1);
  float a = P[OFF_B1 + i];
#pragma unroll
  for (int j = 0; j < NIN; j++) a += P[OFF_W1 + i * NIN + j] * ((x[j] - P[OFF_MU + j]) / P[OFF_SD + j]);
  const float h1 = fmaxf(a, 0.f);
  float b = P[OFF_B2 + i];
#pragma unroll 8
  for (int j = 0; j < NH; j++) b += P[OFF_W2 + i * NH + j] * __shfl(h1, j, 64);
  const float h2 = fmaxf(b, 0.f);
#pragma unroll
  for (int k = 0; k < NOUT; k++) {
    float t = lane < NH ? P[OFF_W3 + k * NH + i] * h2 : 0.f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    y[k] = P[OFF_B3 + k] + t;
  }
}

}  // namespace mlp
}  // namespace vine
