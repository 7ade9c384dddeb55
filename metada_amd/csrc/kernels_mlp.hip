// Learned inflation network (no counterpart in the reference): a per-column
// MLP 8 -> 64 -> 64 -> 1 with tanh hidden layers and rho = 1 + softplus(out),
// in FP32, with its forward pass, the gradient of the mean squared error
// against target inflations, and an Adam update.
//
// Parameter layout (4801 floats): W1[64][8], b1[64], W2[64][64], b2[64],
// W3[64], b3.
//
//   forward      one thread per sample, parameters in LDS (broadcast reads);
//   loss/grad    (a) per sample: forward keeping h1, h2 and the backward
//                    deltas g2 = dL/d(pre-activation 2), g1 likewise, ga;
//                (b) per chunk of 256 samples: every parameter's partial sum
//                    (outer products of deltas and activations);
//                (c) fixed-order sum of the chunk partials, so the gradient
//                    is deterministic;
//   Adam         elementwise with bias correction.
#include "mda_device.hpp"
#include "mda_internal.hpp"

namespace mda {

constexpr int MLP_IN = 8, MLP_H = 64;
constexpr int MLP_NP = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;  // 4801
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_H * MLP_IN, OFF_W2 = OFF_B1 + MLP_H,
              OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H;
static_assert(OFF_B3 + 1 == MLP_NP, "parameter layout");
constexpr int MLP_CHUNK = 256;

__device__ inline float softplus_f(float v) { return v > 30.f ? v : log1pf(expf(v)); }
__device__ inline float sigmoid_f(float v) { return 1.f / (1.f + expf(-v)); }

__device__ inline void load_params(float* P, const float* __restrict__ params) {
  for (int i = threadIdx.x; i < MLP_NP; i += blockDim.x) P[i] = params[i];
  __syncthreads();
}

// pre-activation of the output for one sample; h1 and h2 optionally kept
__device__ inline float mlp_eval(const float* P, const float* x, float* h1, float* h2) {
#pragma unroll
  for (int i = 0; i < MLP_H; ++i) {
    float a = P[OFF_B1 + i];
#pragma unroll
    for (int j = 0; j < MLP_IN; ++j) a = fmaf(P[OFF_W1 + i * MLP_IN + j], x[j], a);
    h1[i] = tanhf(a);
  }
  float out = P[OFF_B3];
  for (int i = 0; i < MLP_H; ++i) {
    float a = P[OFF_B2 + i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) a = fmaf(P[OFF_W2 + i * MLP_H + j], h1[j], a);
    const float h = tanhf(a);
    if (h2) h2[i] = h;
    out = fmaf(P[OFF_W3 + i], h, out);
  }
  return out;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ params,
                                                     const float* __restrict__ feat, int64_t m,
                                                     float* __restrict__ rho) {
  __shared__ float P[MLP_NP];
  load_params(P, params);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  float x[MLP_IN], h1[MLP_H];
#pragma unroll
  for (int j = 0; j < MLP_IN; ++j) x[j] = feat[s * MLP_IN + j];
  rho[s] = 1.f + softplus_f(mlp_eval(P, x, h1, nullptr));
}

// (a) forward with kept activations, and backward deltas, one thread per sample
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float* __restrict__ params,
                                                     const float* __restrict__ feat,
                                                     const float* __restrict__ target, int64_t m,
                                                     float* __restrict__ H1, float* __restrict__ H2,
                                                     float* __restrict__ G1, float* __restrict__ G2,
                                                     float* __restrict__ GA,
                                                     float* __restrict__ L2) {
  __shared__ float P[MLP_NP];
  load_params(P, params);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  float x[MLP_IN], h1[MLP_H], h2[MLP_H];
#pragma unroll
  for (int j = 0; j < MLP_IN; ++j) x[j] = feat[s * MLP_IN + j];
  const float a = mlp_eval(P, x, h1, h2);
  const float r = 1.f + softplus_f(a) - target[s];
  L2[s] = r * r;
  const float ga = 2.f * r * sigmoid_f(a) / (float)m;
  GA[s] = ga;
  float g2[MLP_H];
#pragma unroll
  for (int i = 0; i < MLP_H; ++i) {
    g2[i] = ga * P[OFF_W3 + i] * (1.f - h2[i] * h2[i]);
    H1[s * MLP_H + i] = h1[i];
    H2[s * MLP_H + i] = h2[i];
    G2[s * MLP_H + i] = g2[i];
  }
  for (int j = 0; j < MLP_H; ++j) {
    float t = 0.f;
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) t = fmaf(P[OFF_W2 + i * MLP_H + j], g2[i], t);
    G1[s * MLP_H + j] = t * (1.f - h1[j] * h1[j]);
  }
}

// (b) per-chunk partial sums of every parameter's gradient and of the loss
__global__ __launch_bounds__(256) void k_mlp_grad_partial(
    const float* __restrict__ feat, int64_t m, const float* __restrict__ H1,
    const float* __restrict__ H2, const float* __restrict__ G1, const float* __restrict__ G2,
    const float* __restrict__ GA, const float* __restrict__ L2, float* __restrict__ part,
    double* __restrict__ part_loss) {
  const int64_t s0 = (int64_t)blockIdx.x * MLP_CHUNK;
  const int64_t s1 = s0 + MLP_CHUNK < m ? s0 + MLP_CHUNK : m;
  float* out = part + (size_t)blockIdx.x * MLP_NP;
  for (int p = threadIdx.x; p < MLP_NP; p += blockDim.x) {
    float acc = 0.f;
    if (p < OFF_B1) {
      const int i = p / MLP_IN, j = p % MLP_IN;
      for (int64_t s = s0; s < s1; ++s) acc = fmaf(G1[s * MLP_H + i], feat[s * MLP_IN + j], acc);
    } else if (p < OFF_W2) {
      const int i = p - OFF_B1;
      for (int64_t s = s0; s < s1; ++s) acc += G1[s * MLP_H + i];
    } else if (p < OFF_B2) {
      const int i = (p - OFF_W2) / MLP_H, j = (p - OFF_W2) % MLP_H;
      for (int64_t s = s0; s < s1; ++s) acc = fmaf(G2[s * MLP_H + i], H1[s * MLP_H + j], acc);
    } else if (p < OFF_W3) {
      const int i = p - OFF_B2;
      for (int64_t s = s0; s < s1; ++s) acc += G2[s * MLP_H + i];
    } else if (p < OFF_B3) {
      const int j = p - OFF_W3;
      for (int64_t s = s0; s < s1; ++s) acc = fmaf(GA[s], H2[s * MLP_H + j], acc);
    } else {
      for (int64_t s = s0; s < s1; ++s) acc += GA[s];
    }
    out[p] = acc;
  }
  if (threadIdx.x == 0) {
    double l = 0.0;
    for (int64_t s = s0; s < s1; ++s) l += L2[s];
    part_loss[blockIdx.x] = l;
  }
}

// (c) fixed-order sum over chunks
__global__ __launch_bounds__(256) void k_mlp_grad_final(const float* __restrict__ part,
                                                        const double* __restrict__ part_loss,
                                                        int64_t nchunk, int64_t m,
                                                        float* __restrict__ grad,
                                                        double* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < MLP_NP) {
    float acc = 0.f;
    for (int64_t c = 0; c < nchunk; ++c) acc += part[(size_t)c * MLP_NP + p];
    grad[p] = acc;
  }
  if (p == 0) {
    double l = 0.0;
    for (int64_t c = 0; c < nchunk; ++c) l += part_loss[c];
    *loss = l / (double)m;
  }
}

__global__ void k_mlp_adam(float* __restrict__ params, float* __restrict__ m1,
                           float* __restrict__ m2, const float* __restrict__ grad, float lr,
                           float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NP) return;
  const float g = grad[p];
  const float a = b1 * m1[p] + (1.f - b1) * g;
  const float v = b2 * m2[p] + (1.f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  params[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

void launch_mlp_forward(hipStream_t s, const float* params, const float* feat, int64_t m,
                        float* rho) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, params,
                     feat, m, rho);
  MDA_LAUNCH_CHECK();
}

void launch_mlp_loss_grad(mda_ctx* c, hipStream_t s, const float* params, const float* feat,
                          const float* target, int64_t m, float* grad, double* loss) {
  MDA_REQUIRE(m > 0 && m < (int64_t)1 << 31, "bad number of samples");
  float* H1 = c->get<float>("mlp_h1", (size_t)m * MLP_H);
  float* H2 = c->get<float>("mlp_h2", (size_t)m * MLP_H);
  float* G1 = c->get<float>("mlp_g1", (size_t)m * MLP_H);
  float* G2 = c->get<float>("mlp_g2", (size_t)m * MLP_H);
  float* GA = c->get<float>("mlp_ga", m);
  float* L2 = c->get<float>("mlp_l2", m);
  const int64_t nchunk = (m + MLP_CHUNK - 1) / MLP_CHUNK;
  float* part = c->get<float>("mlp_part", (size_t)nchunk * MLP_NP);
  double* pl = c->get<double>("mlp_part_loss", nchunk);
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, params,
                     feat, target, m, H1, H2, G1, G2, GA, L2);
  MDA_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mlp_grad_partial, dim3((unsigned)nchunk), dim3(256), 0, s, feat, m, H1,
                     H2, G1, G2, GA, L2, part, pl);
  MDA_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mlp_grad_final, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, part, pl,
                     nchunk, m, grad, loss);
  MDA_LAUNCH_CHECK();
}

void launch_mlp_adam(hipStream_t s, float* params, float* m1, float* m2, const float* grad,
                     int64_t step, float lr, float b1, float b2, float eps) {
  MDA_REQUIRE(step >= 1, "Adam step counts from 1");
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, params, m1, m2,
                     grad, lr, b1, b2, eps, c1, c2);
  MDA_LAUNCH_CHECK();
}

}  // namespace mda
