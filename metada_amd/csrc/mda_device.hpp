// Device-side helpers: distances, Gaspari-Cohn, and workgroup-cooperative
// dense kernels (Cholesky, SPD inverse, parallel-order Jacobi) that operate on
// matrices in LDS or in global memory through flat pointers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mda {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEarthRadiusKm = 6371.0;

// Location::haversine (Location.hpp:349-357), same formula, constants and
// operation order, with no FMA contraction (the reference's x86-64 build
// rounds every product and sum separately).
__device__ inline double haversine_km(double lat1, double lon1, double lat2, double lon2) {
#pragma clang fp contract(off)
  const double dlat = (lat2 - lat1) * kPi / 180.0;
  const double dlon = (lon2 - lon1) * kPi / 180.0;
  const double s1 = sin(dlat / 2), s2 = sin(dlon / 2);
  const double a = s1 * s1 + cos(lat1 * kPi / 180.0) * cos(lat2 * kPi / 180.0) * s2 * s2;
  const double c = 2 * atan2(sqrt(a), sqrt(1 - a));
  return kEarthRadiusKm * c;
}

// Squared GRID / CARTESIAN distance of Location::distance_to
// (Location.hpp:205-224) in its operation order, (dx dx + dy dy) + dz dz with
// every product rounded (no FMA), so that a distance equal to the radius in
// the reference's arithmetic is equal here too.  dz = 0 for GRID.
__device__ inline double dist2_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy + dz * dz;
}

// Relative half-width of the shell around the GEOGRAPHIC membership
// threshold chord2_r = 4 sin^2(r / 2R) inside which the chord test on unit
// vectors (equal to the haversine test in exact arithmetic) is not trusted:
// |c2 - chord2_r| <= shell_width(chord2_r).  The rounding of c2 (about
// 1.5e-15 |u - v| absolute) is covered 700 times over.
__host__ __device__ inline double shell_width(double chord2_r) {
  return 1e-9 * chord2_r + 1e-12 * sqrt(chord2_r);
}

// Gaspari & Cohn (1999) eq. (4.10) in z = d / c; zero for z >= 2.
__device__ inline double gaspari_cohn(double z) {
  z = fabs(z);
  if (z >= 2.0) return 0.0;
  if (z <= 1.0) return (((-0.25 * z + 0.5) * z + 0.625) * z - 5.0 / 3.0) * z * z + 1.0;
  return ((((z / 12.0 - 0.5) * z + 0.625) * z + 5.0 / 3.0) * z - 5.0) * z + 4.0 -
         2.0 / (3.0 * z);
}

// FP32 form of gaspari_cohn, for weights that feed FP32 / BF16 contractions.
__device__ inline float gaspari_cohn_f32(float z) {
  z = fabsf(z);
  if (z >= 2.f) return 0.f;
  if (z <= 1.f) return (((-0.25f * z + 0.5f) * z + 0.625f) * z - 5.f / 3.f) * z * z + 1.f;
  return ((((z / 12.f - 0.5f) * z + 0.625f) * z + 5.f / 3.f) * z - 5.f) * z + 4.f -
         2.f / (3.f * z);
}

// ---------------------------------------------------------------------------
// workgroup reductions
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double* red) {
  const int tid = threadIdx.x, nt = blockDim.x;
  red[tid] = v;
  __syncthreads();
  for (int s = nt / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Cholesky M = L L^T in place (lower triangle; strict upper set to 0).
// Returns false when a pivot is not positive.  All threads participate.
// ---------------------------------------------------------------------------
__device__ inline bool block_cholesky(double* A, int lda, int n, int* flag) {
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) *flag = 0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      double a = A[j * lda + j];
      if (!(a > 0.0)) *flag = 1;
      A[j * lda + j] = sqrt(a > 0.0 ? a : 1.0);
    }
    __syncthreads();
    double ljj = A[j * lda + j];
    for (int i = j + 1 + tid; i < n; i += nt) A[i * lda + j] /= ljj;
    __syncthreads();
    int m = n - j - 1;
    for (int idx = tid; idx < m * m; idx += nt) {
      int i = j + 1 + idx / m, l = j + 1 + idx % m;
      if (l <= i) A[i * lda + l] -= A[i * lda + j] * A[l * lda + j];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < n * n; idx += nt) {
    int i = idx / n, l = idx % n;
    if (l > i) A[i * lda + l] = 0.0;
  }
  __syncthreads();
  return *flag == 0;
}

// Given L (lower) in A, writes M^-1 = L^-T L^-1 into A, using W as scratch
// for L^-1 (both n x n with their leading dimensions).
__device__ inline void block_spd_inverse_from_chol(double* A, int lda, double* W, int ldw,
                                                   int n) {
  const int tid = threadIdx.x, nt = blockDim.x;
  // column c of L^-1 by forward substitution (one thread per column)
  for (int c = tid; c < n; c += nt) {
    for (int i = 0; i < c; ++i) W[i * ldw + c] = 0.0;
    for (int i = c; i < n; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int t = c; t < i; ++t) s -= A[i * lda + t] * W[t * ldw + c];
      W[i * ldw + c] = s / A[i * lda + i];
    }
  }
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += nt) {
    int i = idx / n, j = idx % n;
    int t0 = i > j ? i : j;
    double s = 0.0;
    for (int t = t0; t < n; ++t) s += W[t * ldw + i] * W[t * ldw + j];
    A[i * lda + j] = s;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Parallel-order cyclic Jacobi (round-robin "circle" ordering: n-1 rounds of
// n/2 disjoint rotations per sweep).  A (n x n symmetric, n even) is
// diagonalised in place; V (may be null) accumulates the rotations so that
// A_in = V diag(A) V^T.  Rotation: symmetric Schur decomposition (Golub & Van
// Loan, sec. 8.5).  Scratch: rc, rs [n/2] doubles, rp, rq [n/2] ints, red
// [blockDim] doubles.  Returns the number of sweeps.
// ---------------------------------------------------------------------------
__device__ inline int rr_player(int pos, int r, int n) {
  return pos == 0 ? 0 : 1 + (pos - 1 + r) % (n - 1);
}

__device__ inline int block_jacobi(double* A, int lda, double* V, int ldv, int n, double* rc,
                                   double* rs, int* rp, int* rq, double* red, int max_sweeps,
                                   double tol) {
  const int tid = threadIdx.x, nt = blockDim.x, h = n / 2;
  if (V)
    for (int idx = tid; idx < n * n; idx += nt) {
      int i = idx / n, j = idx % n;
      V[i * ldv + j] = (i == j) ? 1.0 : 0.0;
    }
  __syncthreads();
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int idx = tid; idx < n * n; idx += nt) {
      int i = idx / n, j = idx % n;
      double v = A[i * lda + j];
      tot += v * v;
      if (i != j) off += v * v;
    }
    off = block_sum(off, red);
    tot = block_sum(tot, red);
    if (off <= tol * tol * tot) break;
    for (int r = 0; r < n - 1; ++r) {
      for (int m = tid; m < h; m += nt) {
        int p = rr_player(m, r, n), q = rr_player(n - 1 - m, r, n);
        if (p > q) { int t = p; p = q; q = t; }
        double apq = A[p * lda + q], c = 1.0, s = 0.0;
        if (apq != 0.0) {
          double app = A[p * lda + p], aqq = A[q * lda + q];
          double theta = (aqq - app) / (2.0 * apq);
          double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
        }
        rc[m] = c; rs[m] = s; rp[m] = p; rq[m] = q;
      }
      __syncthreads();
      for (int blk = tid; blk < h * h; blk += nt) {
        int m1 = blk / h, m2 = blk % h;
        int p1 = rp[m1], q1 = rq[m1], p2 = rp[m2], q2 = rq[m2];
        double c1 = rc[m1], s1 = rs[m1], c2 = rc[m2], s2 = rs[m2];
        double a = A[p1 * lda + p2], b = A[p1 * lda + q2];
        double cc = A[q1 * lda + p2], d = A[q1 * lda + q2];
        double a2 = c2 * a - s2 * b, b2 = s2 * a + c2 * b;    // B J2
        double c3 = c2 * cc - s2 * d, d2 = s2 * cc + c2 * d;
        double na = c1 * a2 - s1 * c3, nc = s1 * a2 + c1 * c3;  // J1^T (B J2)
        double nb = c1 * b2 - s1 * d2, nd = s1 * b2 + c1 * d2;
        if (m1 == m2) { nb = 0.0; nc = 0.0; }
        A[p1 * lda + p2] = na; A[p1 * lda + q2] = nb;
        A[q1 * lda + p2] = nc; A[q1 * lda + q2] = nd;
      }
      if (V)
        for (int idx = tid; idx < n * h; idx += nt) {
          int m = idx / n, row = idx % n;
          int p = rp[m], q = rq[m];
          double c = rc[m], s = rs[m];
          double vp = V[row * ldv + p], vq = V[row * ldv + q];
          V[row * ldv + p] = c * vp - s * vq;
          V[row * ldv + q] = s * vp + c * vq;
        }
      __syncthreads();
    }
  }
  return sweep;
}

}  // namespace mda
