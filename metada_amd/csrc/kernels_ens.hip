// Streaming ensemble kernels and the global (all-observation) EnKF / ETKF
// pieces: ensemble mean and perturbations, H x with observation-space
// statistics, Y'^T R^-1 [Y' | d] reductions, the k x k solves, the global
// transform and Metrics.
#include "mda_device.hpp"
#include "mda_internal.hpp"

namespace mda {

__global__ void k_zero_words(uint32_t* p, int64_t nwords) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nwords;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

void launch_zero(hipStream_t s, void* ptr, size_t bytes) {
  int64_t nw = (int64_t)(bytes / 4);
  hipLaunchKernelGGL(k_zero_words, dim3(grid_for(nw, 256, 4096)), dim3(256), 0, s,
                     (uint32_t*)ptr, nw);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Ensemble::RecomputeMean / RecomputePerturbations (Ensemble.hpp:105-151):
// sequential sum over members, times the reciprocal 1/k, then (x - mean)*scale.
// Member-major [k][n]; each thread owns state indices, loads are coalesced.
// ---------------------------------------------------------------------------
__global__ void k_mean_pert(const double* __restrict__ X, int k, int64_t n, double scale,
                            double* __restrict__ mean, double* __restrict__ pert) {
  const double rk = 1.0 / (double)k;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += X[(size_t)j * n + i];
    double m = s * rk;
    if (mean) mean[i] = m;
    if (pert)
      for (int j = 0; j < k; ++j) pert[(size_t)j * n + i] = (X[(size_t)j * n + i] - m) * scale;
  }
}

void launch_mean_pert(hipStream_t s, const double* X, int k, int64_t n, double scale,
                      double* mean, double* pert) {
  hipLaunchKernelGGL(k_mean_pert, dim3(grid_for(n, 256, 65536)), dim3(256), 0, s, X, k, n,
                     scale, mean, pert);
  MDA_LAUNCH_CHECK();
}

// sum of squares, one partial per block (reduced on the host in block order)
__global__ void k_sumsq(const double* __restrict__ X, int64_t count, double* partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += X[i] * X[i];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

void launch_sumsq(hipStream_t s, const double* X, int64_t count, double* partial,
                  int nblocks) {
  hipLaunchKernelGGL(k_sumsq, dim3(nblocks), dim3(256), 0, s, X, count, partial);
  MDA_LAUNCH_CHECK();
}

// sum over members and state of ((x - mean) * scale)^2, per-block partials
// (EnKF.hpp:334 background spread, :246-253 analysis spread)
__global__ void k_dev_sumsq(const double* __restrict__ X, int k, int64_t n, double scale,
                            double* partial) {
  __shared__ double red[256];
  const double rk = 1.0 / (double)k;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += X[(size_t)j * n + i];
    double m = s * rk;
    for (int j = 0; j < k; ++j) {
      double v = (X[(size_t)j * n + i] - m) * scale;
      acc += v * v;
    }
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

void launch_dev_sumsq(hipStream_t s, const double* X, int k, int64_t n, double scale,
                      double* partial, int nblocks) {
  hipLaunchKernelGGL(k_dev_sumsq, dim3(nblocks), dim3(256), 0, s, X, k, n, scale, partial);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Observation space (EnKF.hpp:168-184, LETKF.hpp:197-211): hx_j = H x_j for
// every member (CSR stencil gather, or a caller-supplied member-major H(x)),
// ybar = mean over members, Y' = hx - ybar, d = y_o - ybar.  A block owns 64
// observations; the [64][k] tile is assembled in LDS so that the obs-major Y'
// rows go out as contiguous, coalesced stores.  With `order`, the block's
// observations are order[t] for its 64 positions t (for the LETKF, sorted by
// the grid position of their first stencil tap, groups of four levels
// together, so that neighbouring observations' state gathers share cache
// lines); without `opos` observation
// order[t] is written at position t, with `opos` at position opos[order[t]].
// Blocks map XCD-contiguously (block i -> slice i mod 8), so consecutive
// tiles share an L2.
// ---------------------------------------------------------------------------
constexpr int OBS_TILE = 64;
constexpr int OBS_TAPS = 8;  // stencils held in registers (trilinear: 8)

__global__ __launch_bounds__(256) void k_obs_stats(
    const double* __restrict__ X, int k, int64_t n, int64_t p,
    const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col,
    const double* __restrict__ w, const double* __restrict__ hx_mm,
    const double* __restrict__ yo, const uint32_t* __restrict__ order,
    const uint32_t* __restrict__ opos, double* __restrict__ ybar,
    double* __restrict__ yp_om, double* __restrict__ d, double* __restrict__ yp_mm,
    float* __restrict__ y32, int* __restrict__ status) {
  extern __shared__ double tile[];  // [OBS_TILE][k + 1] (padded rows)
  const int kp = k + 1;
  __shared__ double mean_s[OBS_TILE];
  __shared__ int64_t pos_s[OBS_TILE];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t nb = gridDim.x;  // a multiple of 8
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t o0 = vb * OBS_TILE;
  if (o0 >= p) return;
  const int nob = (int)((p - o0) < OBS_TILE ? (p - o0) : OBS_TILE);
  if (hx_mm) {
    // H x_j: entry (ob, j) with ob fastest so hx_mm reads are coalesced
    for (int e = tid; e < nob * k; e += nt) {
      int ob = e % nob, j = e / nob;
      const int64_t o = order ? (int64_t)order[o0 + ob] : o0 + ob;
      tile[ob * kp + j] = hx_mm[(size_t)j * p + o];
    }
  } else {
    // Stencil: lane ob of every wave owns observation ob, wave g members
    // g, g + 4, ...  A stencil of up to OBS_TAPS taps is read once into
    // registers and its gathers for one member go out together; the sum
    // runs over the taps in stencil order.
    const int ob = tid & 63, jg = tid >> 6;
    if (ob < nob) {
      const int64_t o = order ? (int64_t)order[o0 + ob] : o0 + ob;
      const int64_t t0 = row_ptr[o], t1 = row_ptr[o + 1];
      if (t1 - t0 <= OBS_TAPS) {
        const int nt_o = (int)(t1 - t0);
        int64_t cc[OBS_TAPS];
        double ww[OBS_TAPS];
#pragma unroll
        for (int q = 0; q < OBS_TAPS; ++q) {
          cc[q] = q < nt_o ? col[t0 + q] : 0;  // absent taps read entry 0, unused
          ww[q] = q < nt_o ? w[t0 + q] : 0.0;
        }
#pragma unroll 2
        for (int j = jg; j < k; j += nt / 64) {
          const double* xj = X + (size_t)j * n;
          double xv[OBS_TAPS];
#pragma unroll
          for (int q = 0; q < OBS_TAPS; ++q) xv[q] = xj[cc[q]];
          double v = 0.0;
#pragma unroll
          for (int q = 0; q < OBS_TAPS; ++q)
            if (q < nt_o) v += ww[q] * xv[q];
          tile[ob * kp + j] = v;
        }
      } else {
        for (int j = jg; j < k; j += nt / 64) {
          const double* xj = X + (size_t)j * n;
          double v = 0.0;
          for (int64_t t = t0; t < t1; ++t) v += w[t] * xj[col[t]];
          tile[ob * kp + j] = v;
        }
      }
    }
  }
  __syncthreads();
  for (int ob = tid; ob < nob; ob += nt) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += tile[ob * kp + j];
    double m = s / (double)k;
    mean_s[ob] = m;
    const int64_t o = order ? (int64_t)order[o0 + ob] : o0 + ob;
    const int64_t q = opos ? (int64_t)opos[o] : o0 + ob;  // output position
    pos_s[ob] = q;
    const double dv = yo[o] - m;
    if (ybar) ybar[q] = m;
    if (d) d[q] = dv;
    // a non-finite H(x) (hence mean, Y' row and innovation) or observation
    if (status && !isfinite(dv)) atomicOr(status, MDA_ST_NONFINITE_OBS);
  }
  __syncthreads();
  for (int e = tid; e < nob * k; e += nt) {
    int ob = e / k, j = e % k;
    double v = tile[ob * kp + j] - mean_s[ob];
    const size_t at = (size_t)pos_s[ob] * k + j;
    if (yp_om) yp_om[at] = v;
    if (y32) y32[at] = (float)v;
    tile[ob * kp + j] = v;
  }
  if (yp_mm) {
    __syncthreads();
    for (int e = tid; e < nob * k; e += nt) {
      int ob = e % nob, j = e / nob;
      yp_mm[(size_t)j * p + pos_s[ob]] = tile[ob * kp + j];
    }
  }
}

void launch_obs_stats(hipStream_t s, const double* X, int k, int64_t n, int64_t p,
                      const int64_t* row_ptr, const int64_t* col, const double* w,
                      const double* hx_mm, const double* yo, double* ybar, double* yp_om,
                      double* d, double* yp_mm, const uint32_t* order, float* y32,
                      const uint32_t* opos, int* status) {
  if (p <= 0) return;
  const int64_t tiles = (p + OBS_TILE - 1) / OBS_TILE;
  MDA_REQUIRE(tiles < ((int64_t)1 << 31) - 8, "too many observations");
  const unsigned nb = (unsigned)((tiles + 7) / 8 * 8);
  size_t shm = sizeof(double) * OBS_TILE * (k + 1);
  hipLaunchKernelGGL(k_obs_stats, dim3(nb), dim3(256), shm, s, X, k, n, p, row_ptr, col, w,
                     hx_mm, yo, order, opos, ybar, yp_om, d, yp_mm, y32, status);
  MDA_LAUNCH_CHECK();
}

// Processing order of the observation statistics: key = grid index of the
// first stencil tap (0 for an empty stencil), value = observation index.
__global__ void k_tap_keys(const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col,
                           int64_t p, int64_t nxy, uint32_t* __restrict__ keys,
                           uint32_t* __restrict__ vals) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < p;
       o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t0 = row_ptr[o];
    const int64_t c = t0 < row_ptr[o + 1] ? col[t0] : 0;
    // groups of four levels share one key range: the observations of those
    // levels in one grid row sort together, and a stencil's upper level is
    // the next observation's lower one (measured: groups of 1, 2, 4, 8
    // levels take 3.9, 3.5, 3.3, 3.4 ms at 1e6 observations, k = 64)
    const int64_t lev = c / nxy;
    keys[o] = (uint32_t)((lev >> 2) * nxy + (c - lev * nxy));
    vals[o] = (uint32_t)o;
  }
}

void launch_tap_keys(hipStream_t s, const int64_t* row_ptr, const int64_t* col, int64_t p,
                     int64_t nxy, uint32_t* keys, uint32_t* vals) {
  if (p <= 0) return;
  hipLaunchKernelGGL(k_tap_keys, dim3(grid_for(p, 256, 65536)), dim3(256), 0, s, row_ptr, col,
                     p, nxy, keys, vals);
  MDA_LAUNCH_CHECK();
}

// ipos[order[t]] = t
__global__ void k_invert_order(const uint32_t* __restrict__ order, int64_t p,
                               uint32_t* __restrict__ ipos) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < p;
       t += (int64_t)gridDim.x * blockDim.x)
    ipos[order[t]] = (uint32_t)t;
}

void launch_invert_order(hipStream_t s, const uint32_t* order, int64_t p, uint32_t* ipos) {
  if (p <= 0) return;
  hipLaunchKernelGGL(k_invert_order, dim3(grid_for(p, 256, 65536)), dim3(256), 0, s, order, p,
                     ipos);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Global Gram over all observations (ETKF.hpp:150-155; EnKF via Woodbury):
//   mode 0: G[a][c] = sum_o Y'[o][a] Y'[o][c] / var_o (c < k), and
//           G[a][k] = sum_o Y'[o][a] d_o / var_o
//   mode 1: G[a][j] = sum_o Y'[o][a] D[o][j] / var_o with the perturbed
//           innovation D[o][j] = d_o + sqrt(var_o) Z[o][j] - Y'[o][j]
//           (EnKF.hpp:212-224; Z is [p][k]).
// Fixed grid, per-block partials, then a fixed-order reduction.  Blocks of
// 256 threads take 32 observations at a time up to k (k+1) = 5120 entries;
// beyond (64 < k <= 128) blocks of 1024 threads take 16.
// ---------------------------------------------------------------------------
constexpr int GG_ACC = 20;  // entries per thread

__global__ __launch_bounds__(1024) void k_global_gram_partial(
    const double* __restrict__ yp_om, const double* __restrict__ d,
    const double* __restrict__ var, const double* __restrict__ Z, int k, int64_t p, int mode,
    int chunk, double* __restrict__ partial) {
  extern __shared__ double sh[];
  const int kq = mode == 0 ? k + 1 : k;
  double* ys = sh;                  // [chunk][k]
  double* qs = ys + chunk * k;      // [chunk][kq]
  double* rs = qs + chunk * kq;     // [chunk]
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ne = k * kq;
  double acc[GG_ACC];
  for (int t = 0; t < GG_ACC; ++t) acc[t] = 0.0;
  for (int64_t o0 = (int64_t)blockIdx.x * chunk; o0 < p; o0 += (int64_t)gridDim.x * chunk) {
    int nob = (int)((p - o0) < chunk ? (p - o0) : chunk);
    __syncthreads();
    for (int e = tid; e < nob * k; e += nt) ys[e] = yp_om[(size_t)o0 * k + e];
    for (int ob = tid; ob < nob; ob += nt) rs[ob] = 1.0 / var[o0 + ob];
    __syncthreads();
    for (int e = tid; e < nob * kq; e += nt) {
      int ob = e / kq, c = e % kq;
      int64_t o = o0 + ob;
      double v;
      if (mode == 0)
        v = (c < k) ? ys[ob * k + c] : d[o];
      else  // an invalid observation (var = +inf, weight 0) contributes nothing
        v = isinf(var[o]) ? 0.0 : d[o] + sqrt(var[o]) * Z[(size_t)o * k + c] - ys[ob * k + c];
      qs[e] = v;
    }
    __syncthreads();
    int t = 0;
    for (int e = tid; e < ne; e += nt, ++t) {
      int a = e / kq, c = e % kq;
      double s = 0.0;
      for (int ob = 0; ob < nob; ++ob) s += rs[ob] * ys[ob * k + a] * qs[ob * kq + c];
      acc[t] += s;
    }
  }
  int t = 0;
  for (int e = tid; e < ne; e += nt, ++t) partial[(size_t)blockIdx.x * ne + e] = acc[t];
}

__global__ void k_reduce_partials(const double* __restrict__ partial, int nblocks, int ne,
                                  double* __restrict__ out) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * ne + e];
    out[e] = s;
  }
}

void launch_global_gram(hipStream_t s, const double* yp_om, const double* d,
                        const double* var, const double* Z, int k, int64_t p, int mode,
                        double* partial, int nblocks, double* out) {
  const int kq = mode == 0 ? k + 1 : k;
  const int nt = k * kq <= 256 * GG_ACC ? 256 : 1024;
  const int chunk = nt == 256 ? 32 : 16;
  MDA_REQUIRE((k * kq + nt - 1) / nt <= GG_ACC, "global Gram supports k <= 128");
  size_t shm = sizeof(double) * (chunk * k + chunk * kq + chunk);
  hipLaunchKernelGGL(k_global_gram_partial, dim3(nblocks), dim3(nt), shm, s, yp_om, d, var, Z,
                     k, p, mode, chunk, partial);
  MDA_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(grid_for(k * kq, 256)), dim3(256), 0, s, partial,
                     nblocks, k * kq, out);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// ETKF k x k solve (ETKF.hpp:150-159), one workgroup, matrices in global
// scratch through flat pointers (EB is [k][k+1] = [E | b]):
//   Pa = (E + (k-1) I)^-1, wa = Pa b, Wa = sqrt(k-1) chol_L(Pa),
//   T = wa 1^T + Wa  so that X_a = xbar + X' T.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_etkf_solve(int k, const double* __restrict__ EB,
                                                    double* __restrict__ T,
                                                    double* __restrict__ scratch,
                                                    int* status) {
  __shared__ int flag;
  __shared__ double wa[256];
  const int tid = threadIdx.x, nt = blockDim.x;
  double* M = scratch;          // [k][k]
  double* W = scratch + k * k;  // [k][k]
  for (int idx = tid; idx < k * k; idx += nt) {
    int i = idx / k, j = idx % k;
    M[idx] = EB[i * (k + 1) + j] + (i == j ? (double)(k - 1) : 0.0);
  }
  __syncthreads();
  bool ok = block_cholesky(M, k, k, &flag);
  block_spd_inverse_from_chol(M, k, W, k, k);  // M <- Pa
  for (int i = tid; i < k; i += nt) {
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += M[i * k + c] * EB[c * (k + 1) + k];
    wa[i] = s;
  }
  __syncthreads();
  ok = block_cholesky(M, k, k, &flag) && ok;  // M <- chol_L(Pa)
  const double sq = sqrt((double)(k - 1));
  for (int idx = tid; idx < k * k; idx += nt) {
    int i = idx / k, j = idx % k;
    T[idx] = wa[i] + sq * M[idx];
  }
  if (tid == 0 && !ok) atomicOr(status, MDA_ST_NOT_SPD);
}

void launch_etkf_solve_ws(hipStream_t s, int k, const double* EB, double* T, double* scratch,
                          int* status) {
  MDA_REQUIRE(k <= 256, "ETKF supports k <= 256");
  hipLaunchKernelGGL(k_etkf_solve, dim3(1), dim3(256), 0, s, k, EB, T, scratch, status);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// EnKF k x k solve.  With M = (k-1) I + Y'^T R^-1 Y' the Woodbury identity
// gives Y'^T S^-1 / (k-1) = M^-1 Y'^T R^-1 for S = Y'Y'^T/(k-1) + R
// (EnKF.hpp:193-199), so K D = X' M^-1 F with F = Y'^T R^-1 D and the update
// x_a = xbar + X'_j + K D_j (EnKF.hpp:216-226) is X_a = xbar + X'(I + M^-1 F).
// Outputs T = I + M^-1 F and M^-1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_enkf_solve(int k, const double* __restrict__ EB,
                                                    const double* __restrict__ F,
                                                    double* __restrict__ T,
                                                    double* __restrict__ Minv,
                                                    double* __restrict__ scratch, int* status) {
  __shared__ int flag;
  const int tid = threadIdx.x, nt = blockDim.x;
  double* W = scratch;
  for (int idx = tid; idx < k * k; idx += nt) {
    int i = idx / k, j = idx % k;
    Minv[idx] = EB[i * (k + 1) + j] + (i == j ? (double)(k - 1) : 0.0);
  }
  __syncthreads();
  bool ok = block_cholesky(Minv, k, k, &flag);
  block_spd_inverse_from_chol(Minv, k, W, k, k);
  for (int idx = tid; idx < k * k; idx += nt) {
    int i = idx / k, j = idx % k;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += Minv[i * k + c] * F[c * k + j];
    T[idx] = (i == j ? 1.0 : 0.0) + s;
  }
  if (tid == 0 && !ok) atomicOr(status, MDA_ST_NOT_SPD);
}

void launch_enkf_solve_ws(hipStream_t s, int k, const double* EB, const double* F, double* T,
                          double* Minv, double* scratch, int* status) {
  hipLaunchKernelGGL(k_enkf_solve, dim3(1), dim3(256), 0, s, k, EB, F, T, Minv, scratch,
                     status);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Global transform: X_a[j][i] = xbar_i + sum_m (scale (x_m,i - xbar_i)) T[m][j]
// (ETKF.hpp:163-169, EnKF.hpp:216-227).  T in LDS; one thread per state index.
// Nothing is written when an earlier stage of the analysis failed (status bits
// MDA_ST_SKIP_UPDATE); a non-finite background value sets
// MDA_ST_NONFINITE_STATE (the analysis at that index is then non-finite, as
// the reference's arithmetic makes it).
// ---------------------------------------------------------------------------
template <int KMAX>
__global__ __launch_bounds__(256) void k_transform_global(double* __restrict__ X, int k,
                                                          int64_t n, double scale,
                                                          const double* __restrict__ T,
                                                          int* __restrict__ status) {
  if (status && (*status & MDA_ST_SKIP_UPDATE)) return;
  extern __shared__ double Ts[];
  for (int idx = threadIdx.x; idx < k * k; idx += blockDim.x) Ts[idx] = T[idx];
  __syncthreads();
  const double rk = 1.0 / (double)k;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double xp[KMAX];
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
      if (j < k) { xp[j] = X[(size_t)j * n + i]; s += xp[j]; }
    double m = s * rk;
    if (status && !isfinite(m)) atomicOr(status, MDA_ST_NONFINITE_STATE);
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
      if (j < k) xp[j] = (xp[j] - m) * scale;
    for (int j = 0; j < k; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < KMAX; ++a)
        if (a < k) acc += xp[a] * Ts[a * k + j];
      X[(size_t)j * n + i] = m + acc;
    }
  }
}

void launch_transform_global(hipStream_t s, double* X, int k, int64_t n, double scale,
                             const double* T, int* status) {
  size_t shm = sizeof(double) * k * k;
  unsigned g = grid_for(n, 256, 16384);
  if (k <= 32)
    hipLaunchKernelGGL(k_transform_global<32>, dim3(g), dim3(256), shm, s, X, k, n, scale, T, status);
  else if (k <= 64)
    hipLaunchKernelGGL(k_transform_global<64>, dim3(g), dim3(256), shm, s, X, k, n, scale, T, status);
  else if (k <= 128)
    hipLaunchKernelGGL(k_transform_global<128>, dim3(g), dim3(256), shm, s, X, k, n, scale, T, status);
  else
    throw mda_error(MDA_E_UNSUPPORTED, "global transform supports k <= 128");
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// EnKF diagnostics: max / min of K = X' G with G = M^-1 Y'^T R^-1 (k x p),
// EnKF.hpp:202-203.  G is formed first (obs-major [p][k]).
// ---------------------------------------------------------------------------
__global__ void k_enkf_G(const double* __restrict__ Minv, const double* __restrict__ yp_om,
                         const double* __restrict__ var, int k, int64_t p,
                         double* __restrict__ G) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < p * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t o = e / k;
    int a = (int)(e % k);
    double s = 0.0;
    for (int b = 0; b < k; ++b) s += Minv[a * k + b] * yp_om[(size_t)o * k + b];
    G[e] = s / var[o];
  }
}

__global__ __launch_bounds__(256) void k_enkf_gain_minmax(
    const double* __restrict__ X, int k, int64_t n, double scale,
    const double* __restrict__ G, int64_t p, double* __restrict__ partial) {
  __shared__ double rmax[256], rmin[256];
  double mx = -1.0e308, mn = 1.0e308;
  const double rk = 1.0 / (double)k;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double xp[128];
    double s = 0.0;
    for (int j = 0; j < k; ++j) { xp[j] = X[(size_t)j * n + i]; s += xp[j]; }
    double m = s * rk;
    for (int j = 0; j < k; ++j) xp[j] = (xp[j] - m) * scale;
    for (int64_t o = 0; o < p; ++o) {
      double v = 0.0;
      for (int a = 0; a < k; ++a) v += xp[a] * G[(size_t)o * k + a];
      mx = v > mx ? v : mx;
      mn = v < mn ? v : mn;
    }
  }
  rmax[threadIdx.x] = mx;
  rmin[threadIdx.x] = mn;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (threadIdx.x < s2) {
      rmax[threadIdx.x] = fmax(rmax[threadIdx.x], rmax[threadIdx.x + s2]);
      rmin[threadIdx.x] = fmin(rmin[threadIdx.x], rmin[threadIdx.x + s2]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = rmax[0];
    partial[2 * blockIdx.x + 1] = rmin[0];
  }
}

void launch_enkf_gain_minmax_ws(hipStream_t s, const double* X, int k, int64_t n, double scale,
                                const double* Minv, const double* yp_om, const double* var,
                                int64_t p, double* G, double* partial, int nblocks) {
  MDA_REQUIRE(k <= 128, "gain diagnostics support k <= 128");
  hipLaunchKernelGGL(k_enkf_G, dim3(grid_for(p * k, 256, 4096)), dim3(256), 0, s, Minv, yp_om,
                     var, k, p, G);
  MDA_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_enkf_gain_minmax, dim3(nblocks), dim3(256), 0, s, X, k, n, scale, G, p,
                     partial);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// cond(S), S = Y'Y'^T/(k-1) + diag(var), p <= 128 (EnKF.hpp:205-209): S is
// SPD, so its singular values are its eigenvalues (parallel Jacobi, no V).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cond_small(const double* __restrict__ yp_om,
                                                    const double* __restrict__ var, int k,
                                                    int p, double* __restrict__ S,
                                                    double* __restrict__ out) {
  __shared__ double rc[64], rs[64], red[256];
  __shared__ int rp[64], rq[64];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = (p + 1) & ~1;  // even size; the pad row/column is decoupled
  for (int idx = tid; idx < n * n; idx += nt) {
    int a = idx / n, b = idx % n;
    double v = 0.0;
    if (a < p && b < p) {
      for (int j = 0; j < k; ++j) v += yp_om[a * k + j] * yp_om[b * k + j];
      v = v / (double)(k - 1) + (a == b ? var[a] : 0.0);
    } else if (a == b) {
      v = 1.0;
    }
    S[idx] = v;
  }
  __syncthreads();
  block_jacobi(S, n, nullptr, 0, n, rc, rs, rp, rq, red, 60, 1e-16);
  if (tid == 0) {
    double mx = 0.0, mn = 1.0e308;
    for (int a = 0; a < p; ++a) {
      double v = fabs(S[a * n + a]);
      mx = v > mx ? v : mx;
      mn = v < mn ? v : mn;
    }
    out[0] = mx / mn;
  }
}

void launch_cond_small(hipStream_t s, const double* yp_om, const double* var, int k, int p,
                       double* S, double* out) {
  MDA_REQUIRE(p <= 128, "direct cond(S) supports p <= 128");
  hipLaunchKernelGGL(k_cond_small, dim3(1), dim3(256), 0, s, yp_om, var, k, p, S, out);
  MDA_LAUNCH_CHECK();
}

// largest eigenvalue of a k x k symmetric matrix (cond(S) for R = sigma^2 I)
__global__ __launch_bounds__(256) void k_sym_maxeig(const double* __restrict__ EB, int k,
                                                    double* __restrict__ S,
                                                    double* __restrict__ out) {
  __shared__ double rc[128], rs[128], red[256];
  __shared__ int rp[128], rq[128];
  const int n = (k + 1) & ~1;
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    int a = idx / n, b = idx % n;
    S[idx] = (a < k && b < k) ? EB[a * (k + 1) + b] : (a == b ? 0.0 : 0.0);
  }
  __syncthreads();
  block_jacobi(S, n, nullptr, 0, n, rc, rs, rp, rq, red, 60, 1e-16);
  if (threadIdx.x == 0) {
    double mx = 0.0;
    for (int a = 0; a < k; ++a) mx = fmax(mx, S[a * n + a]);
    out[0] = mx;
  }
}

void launch_sym_maxeig(hipStream_t s, const double* EB, int k, double* S, double* out) {
  MDA_REQUIRE(k <= 255, "k too large");
  hipLaunchKernelGGL(k_sym_maxeig, dim3(1), dim3(256), 0, s, EB, k, S, out);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Metrics::CalculateAll (Metrics.hpp:74-94): per-block partial sums of
// [rmse^2, bias, sum mean, sum truth, sum mean*truth, sum mean^2,
//  sum truth^2, spread, crps], reduced on the host in block order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_metrics(const double* __restrict__ X, int k, int64_t n,
                                                 const double* __restrict__ truth,
                                                 double* __restrict__ partial) {
  __shared__ double red[256];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double m = 0.0;
    for (int j = 0; j < k; ++j) m += X[(size_t)j * n + i];
    m /= (double)k;  // Metrics.hpp:111-113 divides
    double sp = 0.0;
    for (int j = 0; j < k; ++j) {
      double df = X[(size_t)j * n + i] - m;
      sp += df * df;
    }
    sp = sqrt(sp / (double)(k - 1));
    double t = truth[i], df = m - t;
    double sd = 0.0, st = 0.0;
    for (int a = 0; a < k; ++a) {
      double xa = X[(size_t)a * n + i];
      for (int b = 0; b < k; ++b) sd += fabs(xa - X[(size_t)b * n + i]);
      st += fabs(xa - t);
    }
    acc[0] += df * df; acc[1] += df; acc[2] += m; acc[3] += t; acc[4] += m * t;
    acc[5] += m * m; acc[6] += t * t; acc[7] += sp;
    acc[8] += st / k - sd / (2.0 * k * k);
  }
  for (int q = 0; q < 9; ++q) {
    double v = block_sum(acc[q], red);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * 9 + q] = v;
  }
}

void launch_metrics(hipStream_t s, const double* X, int k, int64_t n, const double* truth,
                    double* partial, int nblocks) {
  hipLaunchKernelGGL(k_metrics, dim3(nblocks), dim3(256), 0, s, X, k, n, truth, partial);
  MDA_LAUNCH_CHECK();
}

}  // namespace mda
