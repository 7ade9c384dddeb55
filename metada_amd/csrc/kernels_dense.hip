// Dense p x p linear algebra of the global filters, kept in device memory:
//   - the EnKF's S = Y'Y'^T/(k-1) + R and its condition number
//     (EnKF.hpp:188-209: JacobiSVD singular values),
//   - Householder tridiagonalisation of a symmetric matrix, whose extreme
//     eigenvalues (and the one nearest zero) are then found by bisection on
//     Sturm counts of the tridiagonal,
//   - Gauss-Jordan elimination with partial pivoting, A^-1 B for p x m B (the
//     S^-1 products of EnKF.hpp:199 and LWEnKF.hpp:273-276).
// Each elimination step is a short sequence of launches over the trailing
// matrix: O(p^3) work, as the reference's own dense algebra.  Sizes are
// bounded by MDA_DENSE_MAX.
#include <algorithm>
#include <cmath>
#include <limits>

#include "mda_device.hpp"
#include "mda_internal.hpp"

namespace mda {

// S[a][b] = (sum_j Y'[a][j] Y'[b][j]) / (k-1) + (a == b ? var[a] : 0), Y'
// obs-major [p][k]; the member sum runs in member order (as EnKF.hpp:193).
__global__ void k_form_enkf_S(const double* __restrict__ yp, const double* __restrict__ var,
                              int k, int64_t p, double* __restrict__ S) {
  const int64_t total = p * p;
  const double rk1 = 1.0 / (double)(k - 1);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = e / p, b = e % p;
    const double* ya = yp + (size_t)a * k;
    const double* yb = yp + (size_t)b * k;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += ya[j] * yb[j];
    S[e] = s * rk1 + (a == b ? var[a] : 0.0);
  }
}

void launch_form_enkf_S(hipStream_t s, const double* yp, const double* var, int k, int64_t p,
                        double* S) {
  hipLaunchKernelGGL(k_form_enkf_S, dim3(grid_for(p * p, 256, 8192)), dim3(256), 0, s, yp, var,
                     k, p, S);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Householder tridiagonalisation, step j (m = p - j - 1 trailing rows):
//   x = A[j+1:, j], alpha = -sign(x_0) ||x||, u = x - alpha e_1,
//   tau = 2 / u^T u, pv = tau A' u, w = pv - (tau/2)(pv^T u) u,
//   A' <- A' - u w^T - w u^T;  d_j = A[j][j], e_j = alpha.
// ---------------------------------------------------------------------------
__device__ inline double block_reduce_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

__global__ __launch_bounds__(1024) void k_hh_vec(int64_t p, int64_t j, const double* __restrict__ A,
                                                 double* __restrict__ u, double* __restrict__ sc,
                                                 double* __restrict__ d, double* __restrict__ e) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t r = j + 1 + threadIdx.x; r < p; r += blockDim.x) {
    const double x = A[(size_t)r * p + j];
    s += x * x;
  }
  const double nrm2 = block_reduce_sum(s, red);
  const double x0 = A[(size_t)(j + 1) * p + j];
  const bool nz = nrm2 > 0.0;
  const double alpha = nz ? (x0 > 0.0 ? -1.0 : 1.0) * sqrt(nrm2) : 0.0;
  const double tau = nz ? 2.0 / (nrm2 - 2.0 * alpha * x0 + alpha * alpha) : 0.0;
  for (int64_t r = threadIdx.x; r < p; r += blockDim.x) {
    double v = 0.0;
    if (r > j && nz) v = A[(size_t)r * p + j] - (r == j + 1 ? alpha : 0.0);
    u[r] = v;
  }
  if (threadIdx.x == 0) {
    sc[0] = tau;
    d[j] = A[(size_t)j * p + j];
    e[j] = alpha;
  }
}

// pv[r] = tau sum_{c > j} A[r][c] u[c] for r > j: one wavefront per row
__global__ __launch_bounds__(256) void k_hh_symv(int64_t p, int64_t j, const double* __restrict__ A,
                                                 const double* __restrict__ u,
                                                 const double* __restrict__ sc,
                                                 double* __restrict__ pv) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const double tau = sc[0];
  for (int64_t r = j + 1 + blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); r < p;
       r += nw) {
    const double* row = A + (size_t)r * p;
    double s = 0.0;
    for (int64_t c = j + 1 + lane; c < p; c += 64) s += row[c] * u[c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) pv[r] = tau * s;
  }
}

// w = pv - (tau/2)(pv^T u) u (one block)
__global__ __launch_bounds__(1024) void k_hh_w(int64_t p, int64_t j, const double* __restrict__ u,
                                               const double* __restrict__ pv,
                                               const double* __restrict__ sc,
                                               double* __restrict__ w) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t r = j + 1 + threadIdx.x; r < p; r += blockDim.x) s += pv[r] * u[r];
  const double K = 0.5 * sc[0] * block_reduce_sum(s, red);
  for (int64_t r = j + 1 + threadIdx.x; r < p; r += blockDim.x) w[r] = pv[r] - K * u[r];
}

// A' <- A' - u w^T - w u^T on the trailing block (two FMAs per entry)
__global__ __launch_bounds__(256) void k_hh_syr2(int64_t p, int64_t j, double* __restrict__ A,
                                                 const double* __restrict__ u,
                                                 const double* __restrict__ w) {
  const int64_t m = p - j - 1;
  const int64_t total = m * m;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = j + 1 + t / m, c = j + 1 + t % m;
    double& a = A[(size_t)r * p + c];
    a = fma(-u[r], w[c], fma(-w[r], u[c], a));
  }
}

// last 2 x 2 block of the tridiagonal
__global__ void k_hh_tail(int64_t p, const double* __restrict__ A, double* __restrict__ d,
                          double* __restrict__ e) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (p >= 2) {
    d[p - 2] = A[(size_t)(p - 2) * p + (p - 2)];
    e[p - 2] = A[(size_t)(p - 1) * p + (p - 2)];
  }
  d[p - 1] = A[(size_t)(p - 1) * p + (p - 1)];
  e[p - 1] = 0.0;
}

void dense_tridiagonalise(hipStream_t s, int64_t p, double* A, double* d, double* e, double* u,
                          double* pv, double* w, double* sc) {
  MDA_REQUIRE(p >= 1 && p <= MDA_DENSE_MAX, "dense matrix size out of range");
  for (int64_t j = 0; j + 2 < p; ++j) {
    const int64_t m = p - j - 1;
    hipLaunchKernelGGL(k_hh_vec, dim3(1), dim3(1024), 0, s, p, j, A, u, sc, d, e);
    MDA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hh_symv, dim3(grid_for(m, 4, 2048)), dim3(256), 0, s, p, j, A, u, sc, pv);
    MDA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hh_w, dim3(1), dim3(1024), 0, s, p, j, u, pv, sc, w);
    MDA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hh_syr2, dim3(grid_for(m * m, 256, 8192)), dim3(256), 0, s, p, j, A, u, w);
    MDA_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_hh_tail, dim3(1), dim3(64), 0, s, p, A, d, e);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Eigenvalues of a symmetric tridiagonal (d, e) by bisection on Sturm
// counts (host; O(p) per count).  count(x) = number of eigenvalues < x from
// the pivots of the LDL^T factorisation of T - x I.
// ---------------------------------------------------------------------------
static int64_t sturm_count(const std::vector<double>& d, const std::vector<double>& e, double x,
                           double pivmin) {
  int64_t cnt = 0;
  double q = d[0] - x;
  const int64_t p = (int64_t)d.size();
  for (int64_t i = 0;; ++i) {
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
    if (i + 1 >= p) break;
    q = (d[i + 1] - x) - e[i] * e[i] / q;
  }
  return cnt;
}

// eigenvalue number `idx` (ascending, 0-based) to full precision
static double tridiag_eigenvalue(const std::vector<double>& d, const std::vector<double>& e,
                                 int64_t idx, double lo, double hi, double pivmin) {
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_count(d, e, mid, pivmin) > idx) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

dense_extremes tridiag_extremes(const std::vector<double>& d, const std::vector<double>& e) {
  const int64_t p = (int64_t)d.size();
  double lo = std::numeric_limits<double>::infinity(), hi = -lo, emax = 0.0;
  for (int64_t i = 0; i < p; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < p ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
    emax = std::max(emax, std::fabs(e[i]));
  }
  const double span = std::max(std::fabs(lo), std::fabs(hi));
  lo -= 1e-14 * span + 1e-300;
  hi += 1e-14 * span + 1e-300;
  const double pivmin = std::max(1e-300, 1e-300 * emax * emax);
  dense_extremes r;
  r.lmin = tridiag_eigenvalue(d, e, 0, lo, hi, pivmin);
  r.lmax = tridiag_eigenvalue(d, e, p - 1, lo, hi, pivmin);
  // the eigenvalue nearest zero: the largest negative or the smallest non-negative
  const int64_t neg = sturm_count(d, e, 0.0, pivmin);
  double amin = std::numeric_limits<double>::infinity();
  if (neg > 0) amin = std::min(amin, std::fabs(tridiag_eigenvalue(d, e, neg - 1, lo, 0.0, pivmin)));
  if (neg < p) amin = std::min(amin, std::fabs(tridiag_eigenvalue(d, e, neg, 0.0, hi, pivmin)));
  r.absmin = amin;
  r.absmax = std::max(std::fabs(r.lmin), std::fabs(r.lmax));
  return r;
}

dense_extremes dense_sym_extremes(mda_ctx* c, double* A, int64_t p) {
  hipStream_t s = c->stream;
  double* d = c->get<double>("dense_d", p);
  double* e = c->get<double>("dense_e", p);
  double* u = c->get<double>("dense_u", p);
  double* pv = c->get<double>("dense_pv", p);
  double* w = c->get<double>("dense_w", p);
  double* sc = c->get<double>("dense_sc", 4);
  dense_tridiagonalise(s, p, A, d, e, u, pv, w, sc);
  std::vector<double> hd(p), he(p);
  d2h(c, hd.data(), d, sizeof(double) * p);
  d2h(c, he.data(), e, sizeof(double) * p);
  return tridiag_extremes(hd, he);
}

double dense_spd_condition(mda_ctx* c, const double* yp, const double* var, int k, int64_t p) {
  if (p > MDA_DENSE_MAX) return std::numeric_limits<double>::quiet_NaN();
  double* S = c->get<double>("dense_S", (size_t)p * p);
  launch_form_enkf_S(c->stream, yp, var, k, p, S);
  const dense_extremes x = dense_sym_extremes(c, S, p);
  return x.absmax / x.absmin;
}

// ---------------------------------------------------------------------------
// Gauss-Jordan elimination with partial pivoting on [A | B] (A p x p, B p x m,
// both row-major, destroyed): B <- A^-1 B.  Step j: the pivot row is the
// first row r >= j with the largest |A[r][j]| (the choice of partial-pivot
// LU); rows j and r swap; row j is divided by the pivot; every other row r
// takes -A[r][j] times row j (columns > j of A, all of B).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_gj_pivot(int64_t p, int64_t j, const double* __restrict__ A,
                                                   int64_t* __restrict__ piv,
                                                   double* __restrict__ pval,
                                                   int* __restrict__ status) {
  __shared__ double bv[1024];
  __shared__ int64_t bi[1024];
  double best = -1.0;
  int64_t bidx = j;
  for (int64_t r = j + threadIdx.x; r < p; r += blockDim.x) {
    const double v = fabs(A[(size_t)r * p + j]);
    if (v > best) { best = v; bidx = r; }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = bidx;
  __syncthreads();
  for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) {
      const double v2 = bv[threadIdx.x + s2];
      const int64_t i2 = bi[threadIdx.x + s2];
      if (v2 > bv[threadIdx.x] || (v2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
        bv[threadIdx.x] = v2;
        bi[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    piv[0] = bi[0];
    const double pv = A[(size_t)bi[0] * p + j];
    pval[0] = pv;
    if (!(pv != 0.0) || !isfinite(pv)) atomicOr(status, MDA_ST_NOT_SPD);
  }
}

// swap rows j and piv of [A | B] and divide row j by the pivot (columns >= j
// of A, all of B); one thread per column
__global__ void k_gj_swap_scale(int64_t p, int64_t m, int64_t j, double* __restrict__ A,
                                double* __restrict__ B, const int64_t* __restrict__ piv,
                                const double* __restrict__ pval) {
  const int64_t r = piv[0];
  const double inv = 1.0 / pval[0];
  const int64_t ncols = (p - j) + m;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ncols;
       t += (int64_t)gridDim.x * blockDim.x) {
    double* M;
    int64_t ld, cidx;
    if (t < p - j) { M = A; ld = p; cidx = j + t; } else { M = B; ld = m; cidx = t - (p - j); }
    double* rowj = M + (size_t)j * ld + cidx;
    double* rowr = M + (size_t)r * ld + cidx;
    const double vj = *rowj, vr = *rowr;
    *rowr = vj;
    *rowj = vr * inv;
  }
}

// row r != j:  X[r][c] -= A[r][j] X[j][c]  for c > j in A and every c of B
// (column j of A itself is left as it is: later steps never read it)
__global__ void k_gj_eliminate(int64_t p, int64_t m, int64_t j, double* __restrict__ A,
                               double* __restrict__ B) {
  const int64_t nca = p - j - 1, ncols = nca + m;
  const int64_t total = p * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / ncols, cc = t % ncols;
    if (r == j) continue;
    const double f = A[(size_t)r * p + j];
    if (cc < nca) {
      const int64_t c = j + 1 + cc;
      A[(size_t)r * p + c] = fma(-f, A[(size_t)j * p + c], A[(size_t)r * p + c]);
    } else {
      const int64_t c = cc - nca;
      B[(size_t)r * m + c] = fma(-f, B[(size_t)j * m + c], B[(size_t)r * m + c]);
    }
  }
}

void dense_gauss_jordan(hipStream_t s, int64_t p, int64_t m, double* A, double* B,
                        int64_t* piv, double* pval, int* status) {
  MDA_REQUIRE(p >= 1 && p <= MDA_DENSE_MAX && m >= 1, "dense solve size out of range");
  for (int64_t j = 0; j < p; ++j) {
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, s, p, j, A, piv, pval, status);
    MDA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gj_swap_scale, dim3(grid_for(p - j + m, 256, 1024)), dim3(256), 0, s, p,
                       m, j, A, B, piv, pval);
    MDA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gj_eliminate, dim3(grid_for(p * (p - j - 1 + m), 256, 8192)), dim3(256),
                       0, s, p, m, j, A, B);
    MDA_LAUNCH_CHECK();
  }
}

}  // namespace mda
