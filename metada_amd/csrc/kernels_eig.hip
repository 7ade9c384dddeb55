// Per-column symmetric eigensolver for the LETKF transform (the symmetric
// square root of LETKF.hpp:223-224 in its standard form, W = sqrt(k-1) A^-1/2).
//
//   E1  k_trd: Householder reduction A = Q T Q^T to tridiagonal form, one
//       wavefront per column with A in LDS.  d, e go to `de`; the Householder
//       vectors u_j and tau_j are parked in the column's slot of the T buffer
//       (that slot is only written with the final transform by the same
//       column's E3 wavefront, after it has consumed them).
//   E2  k_trd_eigvals: eigenvalues of T by implicit QL with Wilkinson shifts,
//       one LANE per column so the serial scalar recurrence of 64 columns runs
//       side by side.
//   E3  k_trd_vectors: eigenvector s_j of T for eigenvalue lambda_j by a
//       twisted factorisation of T - lambda_j I (one lane per eigenvalue),
//       residual and near-pair orthogonality checks, back-transformation
//       Z = Q S, then W = sqrt(k-1) sqrt(s) Z L^-1/2 Z^T on FP64 MFMA,
//       wbar = s Z L^-1 Z^T b, T = wbar 1^T + W (or the element-wise form).
// A column that fails a check (or QL non-convergence) is flagged and
// re-solved by the parallel-order Jacobi kernel.
#include "mda_device.hpp"
#include "mda_internal.hpp"

namespace mda {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int EIG_N = 64;  // columns handled by this path have k <= 64

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// ---------------------------------------------------------------------------
// E1: Householder tridiagonalisation.  For j = 0..n-3 with m = n-j-1:
//   x = A[j+1:, j],  alpha = -sign(x_0) ||x||,  u = x - alpha e_1,
//   tau = 2 / u^T u,  p = tau A' u,  w = p - (tau/2)(p^T u) u,
//   A' <- A' - u w^T - w u^T   (A' = A[j+1:, j+1:]),  e_j = alpha.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_trd(int n, int64_t ncol, const double* __restrict__ C,
                                            const int* __restrict__ nloc, double rho,
                                            const double* __restrict__ rho_col, int prior,
                                            double* __restrict__ de, double* __restrict__ Tslot) {
  __shared__ double A[EIG_N * (EIG_N + 1)];
  __shared__ double ub[EIG_N], wb[EIG_N];
  const int lane = threadIdx.x, ld = EIG_N + 1;
  for (int64_t g = blockIdx.x; g < ncol; g += gridDim.x) {
    if (nloc[g] == 0) continue;
    const double r = rho_col ? rho_col[g] : rho;
    const double sigma = prior ? (double)(n - 1) / r : (double)(n - 1);
    const double* Cg = C + (size_t)g * n * n;
    for (int idx = lane; idx < n * n; idx += 64) {
      int i = idx / n, c = idx % n;
      A[i * ld + c] = Cg[idx] + (i == c ? sigma : 0.0);
    }
    wave_lds_sync();
    double* U = Tslot + (size_t)g * n * n;
    double* dg = de + (size_t)g * 2 * n;
    for (int j = 0; j < n - 2; ++j) {
      const int m = n - j - 1, o = j + 1;
      double xi = lane < m ? A[(o + lane) * ld + j] : 0.0;
      double nrm2 = wave_sum(xi * xi);
      double x0 = __shfl(xi, 0);
      double alpha = 0.0, tau = 0.0, ui = 0.0;
      if (nrm2 > 0.0) {
        alpha = (x0 > 0.0 ? -1.0 : 1.0) * sqrt(nrm2);
        ui = xi - (lane == 0 ? alpha : 0.0);
        double unrm2 = nrm2 - 2.0 * alpha * x0 + alpha * alpha;
        tau = 2.0 / unrm2;
      }
      if (lane == 0) { dg[j] = A[j * ld + j]; dg[n + j] = alpha; }
      if (lane < m) U[(size_t)j * n + lane] = ui;
      if (lane == 0) U[(size_t)j * n + n - 1] = tau;
      if (tau == 0.0) continue;  // x = 0: H = I, A' unchanged
      ub[lane] = ui;
      wave_lds_sync();
      double pi = 0.0;
      if (lane < m) {
        const double* row = A + (o + lane) * ld + o;
        for (int c = 0; c < m; ++c) pi += row[c] * ub[c];
        pi *= tau;
      }
      double K = 0.5 * tau * wave_sum(pi * ui);
      double wi = pi - K * ui;
      wb[lane] = wi;
      wave_lds_sync();
      if (lane < m) {
        for (int rr = 0; rr < m; ++rr) {
          double* a = A + (o + rr) * ld + o + lane;
          *a -= ub[rr] * wi + wb[rr] * ui;
        }
      }
      wave_lds_sync();
    }
    if (lane == 0) {
      dg[n - 2] = A[(n - 2) * ld + n - 2];
      dg[n - 1] = A[(n - 1) * ld + n - 1];
      dg[n + n - 2] = A[(n - 1) * ld + n - 2];
      dg[n + n - 1] = 0.0;
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// E2: eigenvalues of the symmetric tridiagonal (d, e) by implicit QL with a
// Wilkinson-type shift (Golub & Van Loan, sec. 8.3, QL orientation), one lane
// per column, arrays in LDS as [index][lane].  Ascending output.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_trd_eigvals(int n, int64_t ncol,
                                                    const int* __restrict__ nloc,
                                                    const double* __restrict__ de,
                                                    double* __restrict__ lam,
                                                    int* __restrict__ flag) {
  __shared__ double sd[EIG_N * 64], se[EIG_N * 64];
  const int lane = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * 64 + lane;
  const bool active = g < ncol && nloc[g] != 0;
  double* d = sd + lane;  // d[i * 64]
  double* e = se + lane;
  if (active) {
    const double* dg = de + (size_t)g * 2 * n;
    for (int i = 0; i < n; ++i) { d[i * 64] = dg[i]; e[i * 64] = dg[n + i]; }
    e[(n - 1) * 64] = 0.0;
    bool fail = false;
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n && !fail; ++l) {
      int iter = 0;
      while (true) {
        int m = l;
        for (; m < n - 1; ++m) {
          double dd = fabs(d[m * 64]) + fabs(d[(m + 1) * 64]);
          if (fabs(e[m * 64]) <= eps * dd) break;
        }
        if (m == l) break;
        if (++iter > 60) { fail = true; break; }
        // shift from the leading 2x2 of the unreduced block
        double gg = (d[(l + 1) * 64] - d[l * 64]) / (2.0 * e[l * 64]);
        double rr = hypot(gg, 1.0);
        gg = d[m * 64] - d[l * 64] + e[l * 64] / (gg + (gg >= 0.0 ? rr : -rr));
        double s = 1.0, c = 1.0, p = 0.0;
        int i = m - 1;
        bool early = false;
        for (; i >= l; --i) {
          double f = s * e[i * 64], b = c * e[i * 64];
          rr = hypot(f, gg);
          e[(i + 1) * 64] = rr;
          if (rr == 0.0) {  // underflow: split and restart
            d[(i + 1) * 64] -= p;
            e[m * 64] = 0.0;
            early = true;
            break;
          }
          s = f / rr;
          c = gg / rr;
          gg = d[(i + 1) * 64] - p;
          rr = (d[i * 64] - gg) * s + 2.0 * c * b;
          p = s * rr;
          d[(i + 1) * 64] = gg + p;
          gg = c * rr - b;
        }
        if (early) continue;
        d[l * 64] -= p;
        e[l * 64] = gg;
        e[m * 64] = 0.0;
      }
    }
    // ascending insertion sort
    for (int a = 1; a < n; ++a) {
      double v = d[a * 64];
      int b = a - 1;
      while (b >= 0 && d[b * 64] > v) { d[(b + 1) * 64] = d[b * 64]; --b; }
      d[(b + 1) * 64] = v;
    }
    double* lg = lam + (size_t)g * n;
    for (int i = 0; i < n; ++i) lg[i] = d[i * 64];
    if (fail) flag[g] = 1;  // QL did not converge
  }
}

// ---------------------------------------------------------------------------
// E3: eigenvectors, back-transformation and the transform T.
// Twisted factorisation of T - lambda I (lane j owns eigenvalue j):
//   top-down   D+_0 = d_0 - l,  L_i = e_i / D+_i,  D+_{i+1} = d_{i+1} - l - L_i e_i
//   bottom-up  D-_{n-1} = d_{n-1} - l,  U_i = e_i / D-_{i+1},  D-_i = d_i - l - U_i e_i
//   gamma_k = D+_k + D-_k - (d_k - l);  twist k = argmin |gamma_k|
//   z_k = 1,  z_i = -L_i z_{i+1} (i < k),  z_{i+1} = -U_i z_i (i >= k)
// (T - lambda I) z = gamma_k e_k, so z is the eigenvector when |gamma_k| is at
// rounding level.  The z of lane j lives in registers (static indices).
// ---------------------------------------------------------------------------
__device__ inline double safe_pivot(double v, double tiny) {
  return fabs(v) < tiny ? (v < 0.0 ? -tiny : tiny) : v;
}

__global__ __launch_bounds__(64) void k_trd_vectors(
    int n, int64_t ncol, const int* __restrict__ nloc, const double* __restrict__ de,
    const double* __restrict__ lam, const double* __restrict__ bvec, double rho,
    const double* __restrict__ rho_col, int prior, int combine, double* __restrict__ T,
    int* __restrict__ flag) {
  __shared__ double Z[EIG_N * (EIG_N + 1)];
  __shared__ double sd[EIG_N], se[EIG_N], sl[EIG_N], sy[EIG_N], sb[EIG_N], swt[EIG_N];
  constexpr int BAND = 8;                 // near pairs (j, j+1..j+BAND) of the sorted spectrum
  __shared__ double dots[EIG_N][BAND];
  const int lane = threadIdx.x, ld = EIG_N + 1;
  for (int64_t g = blockIdx.x; g < ncol; g += gridDim.x) {
    if (nloc[g] == 0 || flag[g]) continue;  // empty columns / Jacobi fallback columns
    const double r = rho_col ? rho_col[g] : rho;
    double* Tg = T + (size_t)g * n * n;
    const double* dg = de + (size_t)g * 2 * n;
    sd[lane] = lane < n ? dg[lane] : 0.0;
    se[lane] = lane < n - 1 ? dg[n + lane] : 0.0;
    sl[lane] = lane < n ? lam[(size_t)g * n + lane] : 0.0;
    sb[lane] = lane < n ? bvec[(size_t)g * n + lane] : 0.0;
    wave_lds_sync();
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i) tnorm = fmax(tnorm, fabs(sd[i]) + 2.0 * fabs(se[i]));
    const double tiny = 1e-300 + 2.220446049250313e-16 * 1e-3 * tnorm;
    const double l = sl[lane < n ? lane : 0];
    // ---- twisted factorisation of T - l I ----
    double Lv[EIG_N], Uv[EIG_N];
    double dp = sd[0] - l;
#pragma unroll
    for (int i = 0; i < EIG_N - 1; ++i) {
      if (i < n - 1) {
        double piv = safe_pivot(dp, tiny);
        Lv[i] = se[i] / piv;
        dp = sd[i + 1] - l - Lv[i] * se[i];
      }
    }
    double gbest = fabs(dp);  // gamma_{n-1} = D+_{n-1}
    int kt = n - 1;
    double dm = sd[n - 1] - l;
#pragma unroll
    for (int i = EIG_N - 2; i >= 0; --i) {
      if (i < n - 1) {
        double piv = safe_pivot(dm, tiny);
        Uv[i] = se[i] / piv;
        dm = sd[i] - l - Uv[i] * se[i];
        double dpi = sd[i] - l;
        if (i > 0) dpi -= Lv[i > 0 ? i - 1 : 0] * se[i > 0 ? i - 1 : 0];
        double gam = fabs(dpi + dm - (sd[i] - l));
        if (gam < gbest) { gbest = gam; kt = i; }
      }
    }
    // z_kt = 1; downward z_i = -L_i z_{i+1} kept in Lv[i]; upward z_{i+1} = -U_i z_i in Uv[i]
    double zc = 1.0;
#pragma unroll
    for (int i = EIG_N - 2; i >= 0; --i) {
      if (i < kt) { zc = -Lv[i] * zc; Lv[i] = zc; }
    }
    zc = 1.0;
#pragma unroll
    for (int i = 0; i < EIG_N - 1; ++i) {
      if (i >= kt && i < n - 1) { zc = -Uv[i] * zc; Uv[i] = zc; }
    }
    double nrm = 0.0;
#pragma unroll
    for (int i = 0; i < EIG_N; ++i) {
      if (i < n) {
        double z = i < kt ? Lv[i < EIG_N - 1 ? i : 0] : (i == kt ? 1.0 : Uv[i > 0 ? i - 1 : 0]);
        nrm += z * z;
      }
    }
    const double inv = 1.0 / sqrt(nrm);
    // column `lane` of S into LDS (rows >= n and columns >= n stay zero)
#pragma unroll
    for (int i = 0; i < EIG_N; ++i) {
      double z = 0.0;
      if (i < n && lane < n)
        z = (i < kt ? Lv[i < EIG_N - 1 ? i : 0] : (i == kt ? 1.0 : Uv[i > 0 ? i - 1 : 0])) * inv;
      Z[i * ld + lane] = z;
    }
    wave_lds_sync();
    // ---- eigenpairs at sigma = (k-1)/rho (prior) or k-1 (posterior) ----
    // W and wbar are formed around sigma:
    //   W    = c [ f(sigma) I - sum_j g_j z_j z_j^T ],  g_j = sigma^-1/2 - lambda_j^-1/2
    //   wbar = s [ b / sigma - sum_j (1/sigma - 1/lambda_j)(z_j^T b) z_j ]
    // so eigenpairs with |lambda_j - sigma| <= 2e-12 sigma (the null space of
    // Y'^T R^-1 Y', where eigenvectors are ill-determined) need no vector:
    // each would change W by at most 1e-12 relative.
    const bool pri = prior != 0;
    const double sigma = pri ? (double)(n - 1) / r : (double)(n - 1);
    const bool skip = lane >= n || fabs(l - sigma) <= 2e-12 * sigma;
    // weight of eigenpair j in W and wbar: 1 - sigma/lambda_j (0 at sigma)
    const double wj = skip ? 0.0 : 1.0 - sigma / l;
    sy[lane] = skip ? 1.0 : 0.0;  // skip flags, shared
    swt[lane] = wj;
    wave_lds_sync();
    // ---- checks: eigen-residual of T; close pairs re-orthogonalised ----
    int why = 0;  // 2: residual, 4: near-pair orthogonality, 8: lambda < sigma
    if (!skip) {
      double res = 0.0;
      for (int i = 0; i < n; ++i) {
        double v = (sd[i] - l) * Z[i * ld + lane];
        if (i > 0) v += se[i - 1] * Z[(i - 1) * ld + lane];
        if (i < n - 1) v += se[i] * Z[(i + 1) * ld + lane];
        res = fmax(res, fabs(v));
      }
      if (!(res <= 1e-12 * tnorm)) why |= 2;
      if (!(l >= sigma)) why |= 8;
    }
    // Close eigenpairs from twisted factorisations are accurate except for a
    // small component of each along the other; a symmetric (Lowdin) step
    // z_j -= 1/2 sum_q (z_j.z_q) z_q removes the non-orthogonality to second
    // order and leaves only a rotation inside the pair, which changes W by
    // (g_j - g_q) times that rotation: negligible for close eigenvalues.
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
      for (int q = 0; q < BAND; ++q) dots[lane][q] = 0.0;
      if (!skip) {
        for (int q = lane + 1; q < n && (sl[q] - l) < 1e-3 * tnorm; ++q) {
          if (sy[q] != 0.0) continue;
          double dot = 0.0;
          for (int i = 0; i < n; ++i) dot += Z[i * ld + lane] * Z[i * ld + q];
          if (q - lane > BAND) {
            if (fabs(dot) * (wj + swt[q]) > 1e-12) why |= 4;  // cluster wider than the band
          } else {
            dots[lane][q - lane - 1] = dot;
            if (pass == 1 && fabs(dot) * (wj + swt[q]) > 1e-12) why |= 4;
          }
        }
      }
      wave_lds_sync();
      if (pass == 1) break;
      double znew[EIG_N];
#pragma unroll
      for (int i = 0; i < EIG_N; ++i) znew[i] = Z[i * ld + lane];
      if (!skip) {
        for (int q = 0; q < n; ++q) {
          double c = 0.0;
          if (q > lane && q - lane <= BAND) c = dots[lane][q - lane - 1];
          if (q < lane && lane - q <= BAND) c = dots[q][lane - q - 1];
          if (c != 0.0) {
#pragma unroll
            for (int i = 0; i < EIG_N; ++i) znew[i] -= 0.5 * c * Z[i * ld + q];
          }
        }
      }
      wave_lds_sync();
#pragma unroll
      for (int i = 0; i < EIG_N; ++i) Z[i * ld + lane] = znew[i];
      wave_lds_sync();
    }
    for (int off = 32; off > 0; off >>= 1) why |= __shfl_xor(why, off);
    if (why != 0) {
      if (lane == 0) flag[g] = why;
      wave_lds_sync();
      continue;
    }
    wave_lds_sync();
    // ---- Z = Q S: reflectors in reverse order (u_j, tau_j parked in Tg) ----
    for (int j = n - 3; j >= 0; --j) {
      const int m = n - j - 1, o = j + 1;
      const double* u = Tg + (size_t)j * n;
      const double tau = u[n - 1];
      if (tau == 0.0) continue;
      double dot = 0.0;
      for (int i = 0; i < m; ++i) dot += u[i] * Z[(o + i) * ld + lane];
      dot *= tau;
      for (int i = 0; i < m; ++i) Z[(o + i) * ld + lane] -= dot * u[i];
    }
    wave_lds_sync();
    // ---- wbar = s [ b/sigma - Z diag(1/sigma - 1/lambda) Z^T b ] ----
    const double sc = pri ? 1.0 : r;
    {
      double y = 0.0;
      if (!skip) {
        for (int i = 0; i < n; ++i) y += Z[i * ld + lane] * sb[i];
        y *= (1.0 / sigma - 1.0 / l);
      }
      sy[lane] = y;
    }
    wave_lds_sync();
    double wbar = 0.0;
    if (lane < n) {
      double acc = 0.0;
      for (int e2 = 0; e2 < n; ++e2) acc += Z[lane * ld + e2] * sy[e2];
      wbar = sc * (sb[lane] / sigma - acc);
    }
    wave_lds_sync();
    // ---- U_s = Z diag(sqrt(g)) (scale column `lane`) ----
    {
      double f = skip ? 0.0 : sqrt(1.0 / sqrt(sigma) - 1.0 / sqrt(l));
      for (int i = 0; i < n; ++i) Z[i * ld + lane] *= f;
    }
    wave_lds_sync();
    const double coef = sqrt((double)(n - 1)) * sqrt(sc);
    const double fs = 1.0 / sqrt(sigma);
    if (combine == MDA_COMBINE_ELEMENTWISE) {
      // t_i = wbar_i + coef sum_a U_s[i][a] (sum_j U_s[j][a])
      double cs = 0.0;
      for (int j = 0; j < n; ++j) cs += Z[j * ld + lane];
      sy[lane] = cs;
      wave_lds_sync();
      double t = wbar;
      if (lane < n) {
        double acc = 0.0;
        for (int a = 0; a < n; ++a) acc += Z[lane * ld + a] * sy[a];
        t += coef * (fs - acc);
      }
      for (int idx = lane; idx < n * n; idx += 64) Tg[idx] = 0.0;
      if (lane < n) Tg[(size_t)lane * n + lane] = t;
      wave_lds_sync();
      continue;
    }
    sy[lane] = wbar;
    wave_lds_sync();
    // ---- W = coef U_s U_s^T on FP64 MFMA (lower 16x16 tiles, mirrored) ----
    const int nt = (n + 15) / 16;
    const int i16 = lane & 15, kk = lane >> 4;
    for (int ti = 0; ti < nt; ++ti) {
      for (int tj = 0; tj <= ti; ++tj) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int s4 = 0; s4 < 16; ++s4) {
          double av = Z[(16 * ti + i16) * ld + 4 * s4 + kk];
          double bv = Z[(16 * tj + i16) * ld + 4 * s4 + kk];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int row = 16 * ti + kk + 4 * q, col = 16 * tj + i16;
          if (row < n && col < n) {
            double w = coef * ((row == col ? fs : 0.0) - acc[q]);
            Tg[(size_t)row * n + col] = sy[row] + w;
            if (ti != tj) Tg[(size_t)col * n + row] = sy[col] + w;
          }
        }
      }
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
void launch_letkf_eig(hipStream_t s, const letkf_solve_args& a, double* de, double* lam,
                      int* flag) {
  const int n = a.k;
  MDA_REQUIRE(n >= 2 && n <= EIG_N, "tridiagonal eigensolver path supports k <= 64");
  const int prior = a.inflation_form == MDA_INFLFORM_PRIOR;
  launch_zero(s, flag, sizeof(int) * a.ncol);
  unsigned nb = (unsigned)(a.ncol < 131072 ? a.ncol : 131072);
  hipLaunchKernelGGL(k_trd, dim3(nb), dim3(64), 0, s, n, a.ncol, a.C, a.nloc, a.rho, a.rho_col,
                     prior, de, a.T);
  MDA_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_trd_eigvals, dim3((unsigned)((a.ncol + 63) / 64)), dim3(64), 0, s, n,
                     a.ncol, a.nloc, de, lam, flag);
  MDA_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_trd_vectors, dim3(nb), dim3(64), 0, s, n, a.ncol, a.nloc, de, lam, a.b,
                     a.rho, a.rho_col, prior, a.combine, a.T, flag);
  MDA_LAUNCH_CHECK();
}

}  // namespace mda
