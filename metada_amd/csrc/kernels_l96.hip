// Lorenz-96 forecast on the GPU, the model of the cycled twin experiments
// used to train and evaluate the learned inflation.  The tendency is the
// one of tests/python/lorenz95.py:21,
//   dx_i/dt = (x_{i+1} - x_{i-2}) x_{i-1} - x_i + F   (cyclic indices),
// stepped by classical RK4 or by the forward Euler step of lorenz95.py:28
// with its clipping to [-1e10, 1e10] (lorenz95.py:25,31).
//
// One workgroup advances one state (ensemble member or truth) through all
// steps.  Thread t owns elements i = t + 256 e (registers); the stage inputs
// that neighbouring elements need go through two LDS buffers that alternate
// between stages.
#include "mda_device.hpp"
#include "mda_internal.hpp"

namespace mda {

constexpr int L96_THREADS = 256;
constexpr int L96_EMAX = 24;  // N <= 6144

__device__ inline double l96_f(const double* __restrict__ b, int64_t N, int64_t i, double F) {
  const int64_t ip1 = i + 1 == N ? 0 : i + 1;
  const int64_t im1 = i == 0 ? N - 1 : i - 1;
  const int64_t im2 = i >= 2 ? i - 2 : i + N - 2;
  return (b[ip1] - b[im2]) * b[im1] - b[i] + F;
}

__global__ __launch_bounds__(L96_THREADS) void k_l96(double* __restrict__ X, int64_t N,
                                                     double F, double dt, int nsteps,
                                                     int scheme) {
  extern __shared__ double lds[];  // two buffers of N
  double* B0 = lds;
  double* B1 = lds + N;
  double* x = X + (size_t)blockIdx.x * N;
  const int t = threadIdx.x;
  double xr[L96_EMAX], acc[L96_EMAX];
#pragma unroll
  for (int e = 0; e < L96_EMAX; ++e) {
    const int64_t i = t + (int64_t)L96_THREADS * e;
    xr[e] = i < N ? x[i] : 0.0;
    if (i < N) B0[i] = xr[e];
  }
  __syncthreads();
  for (int step = 0; step < nsteps; ++step) {
    if (scheme == 1) {  // forward Euler with clipping
#pragma unroll
      for (int e = 0; e < L96_EMAX; ++e) {
        const int64_t i = t + (int64_t)L96_THREADS * e;
        if (i < N) {
          const double v = xr[e] + dt * l96_f(B0, N, i, F);
          xr[e] = fmin(fmax(v, -1e10), 1e10);
        }
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < L96_EMAX; ++e) {
        const int64_t i = t + (int64_t)L96_THREADS * e;
        if (i < N) B0[i] = xr[e];
      }
      __syncthreads();
      continue;
    }
    // RK4: stage s reads the stage input from one buffer and writes the next
    // stage input to the other; k-sums accumulate in registers
    const double* in = B0;
    double* out = B1;
    for (int st = 0; st < 4; ++st) {
      const double cw = (st == 0 || st == 3) ? 1.0 : 2.0;  // k1 + 2 k2 + 2 k3 + k4
      const double cs = st == 2 ? 1.0 : 0.5;                // stage input x + cs dt k
#pragma unroll
      for (int e = 0; e < L96_EMAX; ++e) {
        const int64_t i = t + (int64_t)L96_THREADS * e;
        if (i < N) {
          const double kk = l96_f(in, N, i, F);
          acc[e] = st == 0 ? kk : acc[e] + cw * kk;
          if (st < 3) out[i] = xr[e] + cs * dt * kk;
        }
      }
      __syncthreads();
      const double* tmp = in;
      in = out;
      out = const_cast<double*>(tmp);
    }
#pragma unroll
    for (int e = 0; e < L96_EMAX; ++e) {
      const int64_t i = t + (int64_t)L96_THREADS * e;
      if (i < N) {
        xr[e] += dt / 6.0 * acc[e];
        B0[i] = xr[e];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < L96_EMAX; ++e) {
    const int64_t i = t + (int64_t)L96_THREADS * e;
    if (i < N) x[i] = xr[e];
  }
}

void launch_l96(hipStream_t s, double* X, int64_t nstates, int64_t N, double F, double dt,
                int nsteps, int scheme) {
  MDA_REQUIRE(N >= 4 && N <= (int64_t)L96_THREADS * L96_EMAX,
              "Lorenz-96 forecast supports 4 <= N <= 6144");
  MDA_REQUIRE(nstates >= 1 && nstates < (int64_t)1 << 31, "bad number of states");
  const size_t shm = sizeof(double) * 2 * N;
  hipLaunchKernelGGL(k_l96, dim3((unsigned)nstates), dim3(L96_THREADS), shm, s, X, N, F, dt,
                     nsteps, scheme);
  MDA_LAUNCH_CHECK();
}

}  // namespace mda
