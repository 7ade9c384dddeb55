// Per-column symmetric eigensolver for the LETKF transform (the symmetric
// square root of LETKF.hpp:223-224 in its standard form, W = sqrt(k-1) A^-1/2).
//
//   E1  k_trd: Householder reduction A = Q T Q^T to tridiagonal form, one
//       wavefront per column with rows of A in registers.  d, e go to `de`;
//       the Householder vectors u_j and tau_j go to the reflector buffer U.
//   E2  k_trd_eigvals: eigenvalues of T by implicit QL with Wilkinson shifts,
//       one LANE per column so the serial scalar recurrence of 64 columns runs
//       side by side, d and e in registers.
//   E3  k_trd_vectors: eigenvector s_j of T for eigenvalue lambda_j by a
//       twisted factorisation of T - lambda_j I (one lane per eigenvalue),
//       residual and near-pair orthogonality checks, back-transformation
//       Z = Q S, then W = sqrt(k-1) sqrt(s) Z L^-1/2 Z^T on FP64 MFMA,
//       wbar = s Z L^-1 Z^T b, T = wbar 1^T + W (or the element-wise form).
// A column that fails a check (or QL non-convergence) is flagged and
// re-solved by the parallel-order Jacobi kernel.
#include "mda_device.hpp"
#include "mda_internal.hpp"
#include "mda_xform.hpp"

namespace mda {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int EIG_N = 64;  // largest ensemble size handled by this path

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Compiler-only memory barrier (emits no instruction).  Placed every few
// unrolled iterations it makes the compiler re-read broadcast LDS values
// instead of keeping all 2N of d, e (or a whole reflector row) live in
// registers across loops, which would otherwise cost occupancy and spills.
__device__ inline void lds_reload_point() { asm volatile("" ::: "memory"); }

// Waits until at most V vector memory operations of the wave are
// outstanding (vmcnt only; gfx9 s_waitcnt encoding: vmcnt bits 3:0 and
// 15:14, expcnt 6:4 and lgkmcnt 11:8 left at their no-wait maxima).  Vector
// memory operations complete in issue order, so this covers every operation
// issued before the last V.
template <int V>
__device__ inline void wait_vmcnt() {
  static_assert(V >= 0 && V < 64, "vmcnt is six bits");
  __builtin_amdgcn_s_waitcnt((V & 15) | ((V >> 4) << 14) | (7 << 4) | (15 << 8));
  lds_reload_point();
}

// One DPP move of a double (both halves); lanes whose source is out of range
// or whose row is masked off read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ inline double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes, returned wave-uniform: prefix sums within rows of
// 16 (row_shr 1, 2, 4, 8), row 15 -> rows 1, 3 (row_bcast:15), lane 31 ->
// rows 2, 3 (row_bcast:31); lane 63 then holds the total.
__device__ inline double wave_total(double v) {
  v += dpp_f64<0x111>(v);
  v += dpp_f64<0x112>(v);
  v += dpp_f64<0x114>(v);
  v += dpp_f64<0x118>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  return readlane_f64(v, 63);
}

// ---------------------------------------------------------------------------
// E1: Householder tridiagonalisation.  For j = 0..n-3 with m = n-j-1:
//   x = A[j+1:, j],  alpha = -sign(x_0) ||x||,  u = x - alpha e_1,
//   tau = 2 / u^T u,  p = tau A' u,  w = p - (tau/2)(p^T u) u,
//   A' <- A' - u w^T - w u^T   (A' = A[j+1:, j+1:]),  e_j = alpha.
// Lane r keeps row r of A in registers (static column index); lane r's
// entry of column j is picked from its registers in the step before.  The
// reflector j is stored in row U[j] from position b = 16 floor(j / 16) on:
// tau_j at r = b (never a reflector position), 0 at b < r <= j, u at r > j;
// positions r < b are not written (E3 reads each 16-row chunk of reflectors
// from the chunk's first row index on).
// ---------------------------------------------------------------------------
constexpr int E1_CH = 2;  // step chunk of k_trd (static column ranges start at the chunk)

template <int N>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
void k_trd(int64_t ncol, const double* __restrict__ C, int c32,
           const int* __restrict__ nloc, double rho, const double* __restrict__ rho_col,
           int prior, double* __restrict__ de, double* __restrict__ Uout) {
  constexpr int n = N;
  __shared__ double row[EIG_N], ub[EIG_N], wb[EIG_N];
  const int lane = threadIdx.x;
  // one column per block: no loop-carried state (a grid-stride loop costs
  // SGPR spills here)
  {
    const int64_t g = blockIdx.x;
    if (g >= ncol || nloc[g] == 0) return;
    const double r = rho_col ? rho_col[g] : rho;
    const double sigma = prior ? (double)(n - 1) / r : (double)(n - 1);
    const int rl = lane < n ? lane : n - 1;  // clamped row: loads stay in bounds
    double ar[N];
    if (c32) {  // FP32 entries, widened
      const float* Cg = reinterpret_cast<const float*>(C) + (size_t)g * n * n;
#pragma unroll
      for (int c = 0; c < N; ++c) ar[c] = (double)Cg[(size_t)rl * n + c];
    } else {
      const double* Cg = C + (size_t)g * n * n;
#pragma unroll
      for (int c = 0; c < N; ++c) ar[c] = Cg[(size_t)rl * n + c];
    }
#pragma unroll
    for (int c = 0; c < N; ++c) ar[c] = lane < n ? ar[c] + (lane == c ? sigma : 0.0) : 0.0;
    double* U = Uout + (size_t)g * n * n;
    double* dg = de + (size_t)g * 2 * n;
    // xc: column j of the current A, lane r holding A[r][j]; taken from the
    // registers during the previous step's update (static index, selected
    // on the uniform j), so no row travels through LDS.
    double xc = ar[0];
    // Steps j in chunks of E1_CH: within chunk ch every quantity indexed by a
    // column c <= j is zero or unused, so the static column ranges start at
    // E1_CH ch (fewer row writes, dot-product and update terms).
    constexpr int NCH = (N - 2 + E1_CH - 1) / E1_CH;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      const int c0 = E1_CH * ch;
      const int jend = (E1_CH * ch + E1_CH < n - 2) ? E1_CH * ch + E1_CH : n - 2;
#pragma unroll 1
      for (int j = c0; j < jend; ++j) {
        const double xi = (lane > j && lane < n) ? xc : 0.0;
        const double nrm2 = wave_total(xi * xi);
        const double x0 = readlane_f64(xc, j + 1);
        const bool nz = nrm2 > 0.0;
        const double alpha = nz ? (x0 > 0.0 ? -1.0 : 1.0) * sqrt(nrm2) : 0.0;
        const double ui = nz ? xi - (lane == j + 1 ? alpha : 0.0) : 0.0;
        const double tau = nz ? 2.0 / (nrm2 - 2.0 * alpha * x0 + alpha * alpha) : 0.0;
        if (lane == 0) { dg[j] = readlane_f64(xc, j); dg[n + j] = alpha; }
        {
          const int rb0 = 16 * (j / 16);
          if (lane < n && lane >= rb0) U[(size_t)j * n + lane] = lane == rb0 ? tau : ui;
        }
        // x = 0: tau = u = 0, the step below leaves A unchanged (H = I)
        ub[lane] = ui;
        wave_lds_sync();
        // the matrix-vector product runs on the rows of the trailing block
        // only (lanes > j; the rows above are final), so the finished rows
        // take no share of the LDS broadcast of u (masking the rank-2 update
        // as well costs 24 spilled registers at three waves per SIMD)
        const bool trail = lane > j;
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
        if (trail) {
#pragma unroll
          for (int c = c0; c < N; c += 4) {
            p0 += ar[c] * ub[c];
            if (c + 1 < N) p1 += ar[c + 1 < N ? c + 1 : c] * ub[c + 1 < N ? c + 1 : c];
            if (c + 2 < N) p2 += ar[c + 2 < N ? c + 2 : c] * ub[c + 2 < N ? c + 2 : c];
            if (c + 3 < N) p3 += ar[c + 3 < N ? c + 3 : c] * ub[c + 3 < N ? c + 3 : c];
          }
        }
        const double p = tau * ((p0 + p1) + (p2 + p3));
        const double K = 0.5 * tau * wave_total(p * ui);
        const double wi = (lane > j && lane < n) ? p - K * ui : 0.0;
        wb[lane] = wi;
        wave_lds_sync();
#pragma unroll
        for (int c = c0; c < N; ++c) {
          if ((c & 7) == 0) lds_reload_point();
          ar[c] = fma(-ui, wb[c], fma(-wi, ub[c], ar[c]));  // two FMAs
        }
#pragma unroll
        for (int c = c0 + 1; c <= c0 + E1_CH && c < N; ++c)
          if (c == j + 1) xc = ar[c];
        wave_lds_sync();
      }
    }
    // last 2x2 block: rows n-2 and n-1 through the LDS row
    if (lane == n - 2) {
#pragma unroll
      for (int c = 0; c < N; ++c) row[c] = ar[c];
    }
    wave_lds_sync();
    if (lane == 0) dg[n - 2] = row[n - 2];
    wave_lds_sync();
    if (lane == n - 1) {
#pragma unroll
      for (int c = 0; c < N; ++c) row[c] = ar[c];
    }
    wave_lds_sync();
    if (lane == 0) {
      dg[n - 1] = row[n - 1];
      dg[n + n - 2] = row[n - 2];
      dg[n + n - 1] = 0.0;
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// E2: eigenvalues of the symmetric tridiagonal (d, e) by implicit QL with a
// Wilkinson-type shift (Golub & Van Loan, sec. 8.3, QL orientation), one lane
// per column.  d and e live in registers; every lane runs the same static
// index sequence and applies only the steps inside its own unreduced block
// [l, m], so the arrays are addressed by compile-time indices (no LDS, one
// wavefront per SIMD).  Steps outside every lane's block are skipped by a
// wave-uniform branch.  Ascending output by an odd-even merge network.
// ---------------------------------------------------------------------------

// sqrt(f^2 + g^2) without the library hypot; rescaled only when the squares
// would leave the normal range.
__device__ inline double hyp2(double f, double g) {
  const double a = fabs(f), b = fabs(g);
  const double t = fmax(a, b);
  double r = sqrt(fma(f, f, g * g));
  if (t > 1e150 || (t < 1e-150 && t > 0.0)) {
    const double u = fmin(a, b) / t;
    r = t * sqrt(fma(u, u, 1.0));
  }
  return r;
}

// r = sqrt(f^2 + g^2) and 1/r from one reciprocal square root: the hardware
// estimate y (about 2^-24 relative) refined by one third-order step
// y <- y (1 + e/2 + 3e^2/8), e = 1 - t y^2 (the Taylor series of
// (1 - e)^(-1/2)), as accurate as two Newton steps (about 1 ulp, measured
// on the part) with a shorter dependency chain, and far shorter than a
// square root followed by a division.  Sums of squares outside the normal
// range (including 0) take the rescaled route.
__device__ inline void hyp_rinv(double f, double g, double& r, double& inv) {
  const double t2 = fma(f, f, g * g);
  if (t2 > 1e-290 && t2 < 1e290) {
    double y = __builtin_amdgcn_rsq(t2);
    const double e = fma(-t2 * y, y, 1.0);
    y = fma(y * e, fma(e, 0.375, 0.5), y);
    r = t2 * y;
    inv = y;
  } else {
    r = hyp2(f, g);
    inv = 1.0 / r;
  }
}

__device__ inline bool wave_any(bool v) { return __builtin_amdgcn_ballot_w64(v) != 0; }

// Smallest / largest v in [0, 64) over the lanes where p holds, wave-uniform,
// by bisection on ballots (6 steps, no cross-lane data movement); 63 / 0
// when p holds nowhere.
__device__ inline int wave_min_idx(int v, bool p) {
  int t = 0;
#pragma unroll
  for (int b = 32; b >= 1; b >>= 1)
    if (!wave_any(p && v < t + b)) t += b;
  return t;
}
__device__ inline int wave_max_idx(int v, bool p) {
  int t = 0;
#pragma unroll
  for (int b = 32; b >= 1; b >>= 1)
    if (wave_any(p && v >= t + b)) t += b;
  return t;
}

template <int N>
__device__ inline void cmp_swap(double (&d)[N], int i, int j) {
  const double a = d[i], b = d[j];
  d[i] = fmin(a, b);
  d[j] = fmax(a, b);
}

// Batcher's odd-even merge sort over the next power of two, comparators that
// touch an index >= N dropped (equivalent to padding with +inf).
template <int N>
__device__ inline void sort_network(double (&d)[N]) {
  constexpr int P = N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : 64;
#pragma unroll
  for (int p = 1; p < P; p <<= 1) {
#pragma unroll
    for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
      for (int j = k % p; j + k < P; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
          const int a = i + j, b = i + j + k;
          if (b < N && a / (2 * p) == b / (2 * p)) cmp_swap(d, a, b);
        }
      }
    }
  }
}

template <int N>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1)))
void k_trd_eigvals(int64_t ncol, const int* __restrict__ nloc, const double* __restrict__ de,
                   double* __restrict__ lam, int* __restrict__ flag) {
  constexpr int n = N;
  // the index bisections below cover [0, 64); wider problems use k_trd_eigvals_bisect
  static_assert(N >= 2 && N <= 64, "one lane per column holds at most 64 diagonal entries");
  const int lane = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * 64 + lane;
  const bool active = g < ncol && nloc[g] != 0;
  const double* dg = de + (size_t)(active ? g : 0) * 2 * n;
  double d[N], e[N];
#pragma unroll
  for (int i = 0; i < n; ++i) {
    d[i] = active ? dg[i] : 0.0;
    e[i] = (active && i < n - 1) ? dg[n + i] : 0.0;
  }
  // e_i is negligible below tol (|d_i| + |d_{i+1}|): deflating there moves
  // the eigenvalues by O(e_i^2 / gap), far inside the 1e-12 ||T|| residual
  // the eigenvector checks accept, and it can save the last sweep of an
  // eigenvalue's cubic convergence
  const double eps = 1e-13;
  // inactive lanes hold zeros: every e is negligible and l runs to n-1 at once
  int l = 0, iter = 0;
  bool fail = false, done = false;
  // smallest l and largest m over the lanes still iterating: the loops below
  // skip indices outside [lo, hi] by a scalar branch (per-lane conditions
  // still mask the work inside)
  int lo = 0;
  // sweeps of the wavefront: a lane runs at most 61 per value of l and l
  // takes at most n - 1 values, so more than 61 N means a broken invariant.
  // (The count also keeps the register allocator from spreading the
  // selection's values over the sweep: 291 registers instead of 414 at
  // N = 64, far fewer accumulator-register moves.)
  int wit = 0;
  while (true) {
    // negligible off-diagonals e_i (bit i), restricted to i >= the smallest l
    uint64_t small = 0;
#pragma unroll
    for (int i = 0; i < n - 1; ++i) {
      if (i < lo) continue;
      const bool z = fabs(e[i]) <= eps * (fabs(d[i]) + fabs(d[i + 1]));
      small |= z ? (1ull << i) : 0ull;
    }
    // leading converged entries: d_l is an eigenvalue while e_l is negligible
    if (!done) {
      const int lp = __builtin_ctzll(~small & (~0ull << l));  // bits >= n-1 of ~small are set
      if (lp != l) { l = lp; iter = 0; }
      if (l >= n - 1) done = true;
    }
    // end of the unreduced block starting at l
    const uint64_t ms = small & (done ? 0ull : (~0ull << l));
    const int m = ms ? __builtin_ctzll(ms) : n - 1;
    if (!done && ++iter > 60) { fail = true; done = true; }
    if (!wave_any(!done)) break;
    lo = wave_min_idx(l, !done);
    const int hi = wave_max_idx(m, !done);
    double dl = 0.0, dl1 = 0.0, el = 1.0, dm = 0.0;
#pragma unroll
    for (int i = 0; i < n; ++i) {
      if (i < lo || i > hi) continue;
      if (i == l) {
        dl = d[i];
        el = e[i];
        if (i + 1 < n) dl1 = d[i + 1];
      }
      if (i == m) dm = d[i];
    }
    ++wit;
    // shift from the leading 2x2 of the block
    double gg = (dl1 - dl) / (2.0 * el);
    double rr = hyp2(gg, 1.0);
    gg = dm - dl + el / (gg + (gg >= 0.0 ? rr : -rr));
    double s = 1.0, c = 1.0, p = 0.0;
    bool run = !done;  // cleared when this lane's sweep ends early (underflow)
#pragma unroll
    for (int i = n - 2; i >= 0; --i) {
      if (i < lo || i >= hi) continue;
      const bool act = run && l <= i && i < m;
      const double f = s * e[i], b = c * e[i];
      double r, inv;
      hyp_rinv(f, gg, r, inv);
      const bool under = act && r == 0.0;     // split and restart the block
      const bool rot = act && r != 0.0;
      if (act) e[i + 1] = (i == m - 1 || under) ? 0.0 : r;  // e_m ends at zero
      if (under) { d[i + 1] -= p; run = false; }
      const double sn = f * inv, cn = gg * inv;
      const double g2 = d[i + 1] - p;
      const double r2 = (d[i] - g2) * sn + 2.0 * cn * b;
      const double pn = sn * r2;
      if (rot) {
        d[i + 1] = g2 + pn;
        gg = cn * r2 - b;
        s = sn;
        c = cn;
        p = pn;
        if (i == l) {  // end of the sweep
          d[i] -= p;
          e[i] = gg;
        }
      }
    }
  }
  if (wit > 61 * N) fail = true;
  sort_network<N>(d);
  if (active) {
    double* lg = lam + (size_t)g * n;
#pragma unroll
    for (int i = 0; i < n; ++i) lg[i] = d[i];
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

// a / b for the recurrences of the twisted factorisation: the hardware
// reciprocal y (about 2^-24 relative) refined by one third-order step,
// a / b = a y (1 + e + e^2) with e = 1 - b y (about 1 ulp, measured on the
// part; three dependent operations after the estimate, against five for two
// Newton steps and far more for the IEEE division sequence).  b is a
// safeguarded pivot: |b| >= tiny > 0.
__device__ inline double fast_div(double a, double b) {
  const double y = __builtin_amdgcn_rcp(b);
  const double e = fma(-b, y, 1.0);
  const double q = a * y;
  return fma(q, fma(e, e, e), q);
}

constexpr int E3_SUB = 4;  // reflectors per static range in the back-transform

// global_load_lds instructions that stage chunk c of the Householder rows in
// E3, 64 dwords each: chunk 0 (rows 0 .. min(15, N - 3)) as one contiguous
// block, chunk c > 0 (rows 16c .. min(16c + 15, N - 3)) row by row from
// position 16c (0 for a chunk past the last)
template <int N>
constexpr int e3_stage_ops(int c) {
  const int jhi = 16 * c + 15 < N - 3 ? 16 * c + 15 : N - 3;
  const int rows = jhi - 16 * c + 1;
  if (rows <= 0) return 0;
  if (c == 0) return (rows * N * 8 + 255) / 256;
  return rows * ((2 * (N - 16 * c) + 63) / 64);
}

// LDS of one column in E3, in doubles: Z, then sd, se, sl, sy, sb, swt (64
// each) and the near-pair dot products [64][E3_BAND].
constexpr int E3_BAND = 8;  // near pairs (j, j+1..j+BAND) of the sorted spectrum
template <int N>
struct e3_shape {
  static constexpr int NP = (N + 15) / 16 * 16;  // MFMA tile padding
  static constexpr int ld = NP + 1;              // Z[i * ld + j]: component i of vector j
  static constexpr int ZS = NP * ld > N * 64 ? NP * ld : N * 64;
  static constexpr int SLICE = ZS + 6 * 64 + 64 * E3_BAND;
};

// E3 of column g by one wavefront, in the wavefront's own LDS slice of
// e3_shape<N>::SLICE doubles.  FUSED = false: T of the column goes to T.
// FUSED = true (N a multiple of 16, matrix combination): T stays in tb in
// the MFMA B layout, tb[s][tj] = T[4s + kk][16 tj + i16], and nothing is
// written to T; an empty column gets T = sqrt(rho) I there, and pre() is
// called once the eigenvectors have left the registers (the caller issues
// loads there whose latency the W products hide).  Returns true when tb
// holds the column's T (FUSED), false for a column left to the Jacobi
// fallback (flagged now or before) and, unfused, for an empty one.
template <int N, bool FUSED, class Pre>
__device__ __forceinline__ bool e3_column(
    int64_t g, double* __restrict__ slice, const int* __restrict__ nloc,
    const double* __restrict__ de, const double* __restrict__ lam,
    const double* __restrict__ bvec, double rho, const double* __restrict__ rho_col, int prior,
    int combine, double null_rel, const double* __restrict__ Ug, double* __restrict__ T,
    int* __restrict__ flag, double (&tb)[e3_shape<N>::NP / 4][e3_shape<N>::NP / 16],
    Pre&& pre) {
  constexpr int n = N;
  constexpr int NP = e3_shape<N>::NP;
  constexpr int ld = e3_shape<N>::ld;
  constexpr int BAND = E3_BAND;
  // Z doubles as the per-lane U_i store [i * 64 + lane] of the factorisation
  // and as the staging area of the Householder rows for the back-transform.
  double* Z = slice;
  double* sd = slice + e3_shape<N>::ZS;
  double* se = sd + 64;
  double* sl = se + 64;
  double* sy = sl + 64;
  double* sb = sy + 64;
  double* swt = sb + 64;
  double* dots = swt + 64;  // [64][BAND]
  const int lane = threadIdx.x & 63;
  const int i16 = lane & 15, kk = lane >> 4;
  const double r = rho_col ? rho_col[g] : rho;
  if (nloc[g] == 0) {
    if constexpr (FUSED) {
      const double sr = sqrt(r);
#pragma unroll
      for (int s4 = 0; s4 < NP / 4; ++s4)
#pragma unroll
        for (int tj = 0; tj < NP / 16; ++tj) tb[s4][tj] = 4 * s4 + kk == 16 * tj + i16 ? sr : 0.0;
      return true;
    }
    return false;
  }
  if (flag[g]) return false;  // QL did not converge: Jacobi fallback
  wave_lds_sync();            // the slice's previous readers are done
  double* Tg = T + (size_t)g * n * n;
  const double* dg = de + (size_t)g * 2 * n;
  sd[lane] = lane < n ? dg[lane] : 0.0;
  se[lane] = lane < n - 1 ? dg[n + lane] : 0.0;
  sl[lane] = lane < n ? lam[(size_t)g * n + lane] : 0.0;
  sb[lane] = lane < n ? bvec[(size_t)g * n + lane] : 0.0;
  wave_lds_sync();
  double tnorm = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if ((i & 7) == 0) lds_reload_point();
    tnorm = fmax(tnorm, fabs(sd[i]) + 2.0 * fabs(se[i]));
  }
  const double tiny = 1e-300 + 2.220446049250313e-16 * 1e-3 * tnorm;
  const double l = sl[lane < n ? lane : 0];
  // ---- eigenvector z of T for lambda_lane, in registers ----
  // L_i built in place in z[], U_i parked in LDS, then
  // z_kt = 1, z_i = -L_i z_{i+1} (i < kt), z_{i+1} = -U_i z_i (i >= kt)
  double z[N];
  double* Uv = Z;
  {
    double dp = sd[0] - l;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
      if ((i & 7) == 0) lds_reload_point();
      z[i] = fast_div(se[i], safe_pivot(dp, tiny));  // L_i
      dp = sd[i + 1] - l - z[i] * se[i];
    }
    z[N - 1] = 0.0;
    double gbest = fabs(dp);  // gamma_{n-1} = D+_{n-1}
    int kt = N - 1;
    double dm = sd[N - 1] - l;
#pragma unroll
    for (int i = N - 2; i >= 0; --i) {
      if ((i & 7) == 6) lds_reload_point();
      const double u = fast_div(se[i], safe_pivot(dm, tiny));
      Uv[i * 64 + lane] = u;
      dm = sd[i] - l - u * se[i];
      const double dpi = sd[i] - l - (i > 0 ? z[i > 0 ? i - 1 : 0] * se[i > 0 ? i - 1 : 0] : 0.0);
      const double gam = fabs(dpi + dm - (sd[i] - l));
      if (gam < gbest) { gbest = gam; kt = i; }
    }
    double zc = 1.0;
#pragma unroll
    for (int i = N - 2; i >= 0; --i)
      if (i < kt) { zc = -z[i] * zc; z[i] = zc; }
    double nrm = 0.0;
    zc = 1.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if ((i & 7) == 0) lds_reload_point();
      const double cand = i > 0 ? -Uv[(i > 0 ? i - 1 : 0) * 64 + lane] * zc : 0.0;
      const double v = i < kt ? z[i] : (i == kt ? 1.0 : cand);
      if (i >= kt) zc = v;
      z[i] = lane < n ? v : 0.0;
      nrm += z[i] * z[i];
    }
    const double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) z[i] *= inv;
  }
  wave_lds_sync();  // U_i reads done before Z is written
  // ---- eigenpairs at sigma = (k-1)/rho (prior) or k-1 (posterior) ----
  // W and wbar are formed around sigma:
  //   W    = c [ f(sigma) I - sum_j g_j z_j z_j^T ],  g_j = sigma^-1/2 - lambda_j^-1/2
  //   wbar = s [ b / sigma - sum_j (1/sigma - 1/lambda_j)(z_j^T b) z_j ]
  // so eigenpairs with |lambda_j - sigma| <= 2e-12 sigma (the null space of
  // Y'^T R^-1 Y', where eigenvectors are ill-determined) need no vector:
  // each would change W by at most 1e-12 relative.  With a reduced-precision
  // Gram the window widens to the Gram's own rounding level, null_rel ||A||.
  const bool pri = prior != 0;
  const double sigma = pri ? (double)(n - 1) / r : (double)(n - 1);
  const double null_tol = fmax(2e-12 * sigma, null_rel * tnorm);
  const bool skip = lane >= n || fabs(l - sigma) <= null_tol;
  // weight of eigenpair j in W and wbar: 1 - sigma/lambda_j (0 at sigma)
  const double wj = skip ? 0.0 : 1.0 - sigma / l;
  sy[lane] = skip ? 1.0 : 0.0;  // skip flags, shared
  swt[lane] = wj;
  int why = 0;  // 2: residual, 4: near-pair orthogonality, 8: lambda < sigma
  if (!skip) {
    double res = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if ((i & 7) == 0) lds_reload_point();
      double v = (sd[i] - l) * z[i];
      if (i > 0) v += se[i > 0 ? i - 1 : 0] * z[i > 0 ? i - 1 : 0];
      if (i < N - 1) v += se[i] * z[i < N - 1 ? i + 1 : 0];
      res = fmax(res, fabs(v));
    }
    if (!(res <= 1e-12 * tnorm)) why |= 2;
    if (!(l >= sigma - null_tol)) why |= 8;
  }
  if (lane < NP) {
#pragma unroll
    for (int i = 0; i < N; ++i) Z[i * ld + lane] = z[i];
#pragma unroll
    for (int i = N; i < NP; ++i) Z[i * ld + lane] = 0.0;
  }
  wave_lds_sync();
  // Close eigenpairs from twisted factorisations are accurate except for a
  // small component of each along the other; a symmetric (Lowdin) step
  // z_j -= 1/2 sum_q (z_j.z_q) z_q removes the non-orthogonality to second
  // order and leaves only a rotation inside the pair, which changes W by
  // (g_j - g_q) times that rotation: negligible for close eigenvalues.
  bool any_near = false;
  for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
    for (int q = 0; q < BAND; ++q) dots[lane * BAND + q] = 0.0;
    bool near = false;
    if (!skip) {
      for (int q = lane + 1; q < n && (sl[q] - l) < 1e-3 * tnorm; ++q) {
        if (sy[q] != 0.0) continue;
        near = true;
        double dot = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) dot += z[i] * Z[i * ld + q];
        if (q - lane > BAND) {
          if (fabs(dot) * (wj + swt[q]) > 1e-12) why |= 4;  // cluster wider than the band
        } else {
          dots[lane * BAND + q - lane - 1] = dot;
          if (pass == 1 && fabs(dot) * (wj + swt[q]) > 1e-12) why |= 4;
        }
      }
    }
    wave_lds_sync();
    if (pass == 0) any_near = __ballot(near) != 0ull;
    if (pass == 1 || !any_near) break;
    if (!skip) {
      for (int q = lane - BAND; q <= lane + BAND; ++q) {
        if (q < 0 || q >= n || q == lane) continue;
        const double c = q > lane ? dots[lane * BAND + q - lane - 1] : dots[q * BAND + lane - q - 1];
        if (c != 0.0) {
#pragma unroll
          for (int i = 0; i < N; ++i) z[i] -= 0.5 * c * Z[i * ld + q];
        }
      }
    }
    wave_lds_sync();
    if (lane < NP) {
#pragma unroll
      for (int i = 0; i < N; ++i) Z[i * ld + lane] = z[i];
    }
    wave_lds_sync();
  }
  for (int off = 32; off > 0; off >>= 1) why |= __shfl_xor(why, off);
  if (why != 0) {
    if (lane == 0) flag[g] = why;
    wave_lds_sync();
    return false;
  }
  // ---- z <- Q z with Q = H_0 ... H_{n-3} ----
  // The Householder rows (row j from position b = 16 floor(j / 16): tau_j at
  // b, 0 at b < r <= j, u at r > j) are the same for every lane.  They are
  // copied into the Z area (free between the near-pair step and U_s) by
  // global_load_lds, no registers held, two chunks of 16 rows ahead of their
  // use, and read back as LDS broadcasts; chunk c's rows from position 16c
  // on only (the entries before are zero for every reflector of the chunk),
  // one row at a time.  (The compiler makes the first LDS read of a chunk
  // wait for every copy still outstanding, so the copies issued before it
  // land first.)  Reflectors are applied in sub-chunks of E3_SUB so the
  // static inner range starts at the sub-chunk's first non-zero position.
  {
    const double* Ucol = Ug + (size_t)g * n * n;
    constexpr int NC = (N - 2 + 15) / 16;
    // rows 16c .. min(16c + 15, N - 3), positions [16c, N), to the same
    // places of Z (row stride N): per row ceil((N - 16c) / 32) instructions
    // of 64 dwords, all lanes (the LDS destination comes from one lane, so
    // the copies stay outside divergent code).  A row's last instruction may
    // run on into the next rows' first entries, which it copies to the same
    // places: entries below the next row's chunk start (never read), or the
    // row's own values; it stays inside the column's n x n block and Z.
    // Chunk 0 (whole rows) goes as one contiguous block of 256-byte
    // instructions, the last of which may run on up to 252 bytes, inside the
    // column's n x n block.
    auto stage = [&](int c) {
      const int jlo = 16 * c;
      const int jhi = (16 * c + 15 < N - 3) ? 16 * c + 15 : N - 3;
      if (c == 0) {
        const char* src = (const char*)Ucol + 4 * lane;
        char* dst = (char*)Z;
        for (int off = 0; off < (jhi + 1) * N * 8; off += 256)
          __builtin_amdgcn_global_load_lds((const void*)(src + off),
                                           (__attribute__((address_space(3))) void*)(dst + off),
                                           4, 0, 0);
        return;
      }
      const int dw = 2 * (N - jlo);  // dwords per row
      for (int j = jlo; j <= jhi; ++j) {
        const char* src = (const char*)(Ucol + (size_t)j * N + jlo) + 4 * lane;
        char* dst = (char*)(Z + j * N + jlo);
        for (int d0 = 0; d0 < dw; d0 += 64)
          __builtin_amdgcn_global_load_lds((const void*)(src + 4 * d0),
                                           (__attribute__((address_space(3))) void*)(dst + 4 * d0),
                                           4, 0, 0);
      }
    };
    stage(NC - 1);
    if (NC > 1) stage(NC - 2);
#pragma unroll
    for (int c = NC - 1; c >= 0; --c) {
      const int jlo = 16 * c;
      const int jhi = (16 * c + 15 < N - 3) ? 16 * c + 15 : N - 3;
      // chunk c has landed once at most chunk c-1's copies are outstanding
      if (c == 3) wait_vmcnt<e3_stage_ops<N>(2)>();
      else if (c == 2) wait_vmcnt<e3_stage_ops<N>(1)>();
      else if (c == 1) wait_vmcnt<e3_stage_ops<N>(0)>();
      else wait_vmcnt<0>();
      if (c >= 2) stage(c - 2);
      // sub-chunks of E3_SUB reflectors, each with its own static range
#pragma unroll
      for (int sc = 16 / E3_SUB - 1; sc >= 0; --sc) {
        const int slo = jlo + E3_SUB * sc;
        const int shi = (slo + E3_SUB - 1 < jhi) ? slo + E3_SUB - 1 : jhi;
        if (slo > jhi) continue;
#pragma unroll 1
        for (int j = shi; j >= slo; --j) {
          const double* u = Z + j * N;
          const double tau = u[jlo];
          double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
#pragma unroll
          for (int i = slo + 1; i < N; ++i) {
            const double t = u[i] * z[i];
            switch (i & 3) {
              case 0: p0 += t; break;
              case 1: p1 += t; break;
              case 2: p2 += t; break;
              default: p3 += t; break;
            }
          }
          const double dot = tau * ((p0 + p1) + (p2 + p3));
#pragma unroll
          for (int i = slo + 1; i < N; ++i) z[i] -= dot * u[i];
        }
      }
    }
  }
  // ---- U_s = Z diag(sqrt(g)) and y' for wbar ----
  const double sc = pri ? 1.0 : r;
  const double f = skip ? 0.0 : sqrt(1.0 / sqrt(sigma) - 1.0 / sqrt(l));
  double yb = 0.0;
  if (!skip) {
#pragma unroll
    for (int i = 0; i < N; ++i) yb += z[i] * sb[i];
    yb *= (1.0 / sigma - 1.0 / l);
  }
  sy[lane] = f > 0.0 ? yb / f : 0.0;  // Z y = U_s (y / f)
  if (lane < NP) {
#pragma unroll
    for (int i = 0; i < N; ++i) Z[i * ld + lane] = z[i] * f;
#pragma unroll
    for (int i = N; i < NP; ++i) Z[i * ld + lane] = 0.0;
  }
  wave_lds_sync();
  // wbar_r = s [ b_r / sigma - sum_j U_s[r][j] y'_j ]
  double wbar = 0.0;
  if (lane < n) {
    double acc = 0.0;
#pragma unroll
    for (int e2 = 0; e2 < N; ++e2) acc += Z[lane * ld + e2] * sy[e2];
    wbar = sc * (sb[lane] / sigma - acc);
  }
  const double coef = sqrt((double)(n - 1)) * sqrt(sc);
  const double fs = 1.0 / sqrt(sigma);
  if (combine == MDA_COMBINE_ELEMENTWISE) {
    // t_i = wbar_i + coef (f(sigma) - sum_a U_s[i][a] (sum_j U_s[j][a]))
    double cs = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) cs += z[i] * f;
    wave_lds_sync();
    sy[lane] = cs;
    wave_lds_sync();
    double t = wbar;
    if (lane < n) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < N; ++a) acc += Z[lane * ld + a] * sy[a];
      t += coef * (fs - acc);
    }
    for (int idx = lane; idx < n * n; idx += 64) Tg[idx] = 0.0;
    if (lane < n) Tg[(size_t)lane * n + lane] = t;
    wave_lds_sync();
    return false;
  }
  wave_lds_sync();
  sy[lane] = wbar;
  wave_lds_sync();
  // ---- W = coef (f(sigma) I - U_s U_s^T) on FP64 MFMA (lower tiles, mirrored) ----
  constexpr int nt = NP / 16;
  if constexpr (!FUSED) {
#pragma unroll 1
    for (int ti = 0; ti < nt; ++ti) {
#pragma unroll 1
      for (int tj = 0; tj <= ti; ++tj) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < NP / 4; ++s4) {
          const double av = Z[(16 * ti + i16) * ld + 4 * s4 + kk];
          const double bv = Z[(16 * tj + i16) * ld + 4 * s4 + kk];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = 16 * ti + kk + 4 * q, col = 16 * tj + i16;
          if (row < n && col < n) {
            // T = wbar + coef wd with one rounding (the fused path's form)
            const double wd = (row == col ? fs : 0.0) - acc[q];
            Tg[(size_t)row * n + col] = fma(coef, wd, sy[row]);
            if (ti != tj) Tg[(size_t)col * n + row] = fma(coef, wd, sy[col]);
          }
        }
      }
    }
    wave_lds_sync();
    return false;
  } else {
    static_assert(N % 16 == 0, "the fused transform needs N a multiple of 16");
    pre();
    // the lower tiles in registers until U_s has been read, then the full
    // symmetric W / coef through Z and T = wbar 1^T + W into the B layout:
    // the same values, rounded the same way, as the unfused path stores
    d4 wt[nt * (nt + 1) / 2];
#pragma unroll
    for (int ti = 0; ti < nt; ++ti) {
#pragma unroll
      for (int tj = 0; tj <= ti; ++tj) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < NP / 4; ++s4) {
          const double av = Z[(16 * ti + i16) * ld + 4 * s4 + kk];
          const double bv = Z[(16 * tj + i16) * ld + 4 * s4 + kk];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
        wt[ti * (ti + 1) / 2 + tj] = acc;
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int ti = 0; ti < nt; ++ti) {
#pragma unroll
      for (int tj = 0; tj <= ti; ++tj) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = 16 * ti + kk + 4 * q, col = 16 * tj + i16;
          const double wd = (row == col ? fs : 0.0) - wt[ti * (ti + 1) / 2 + tj][q];
          Z[row * ld + col] = wd;
          if (ti != tj) Z[col * ld + row] = wd;
        }
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int s4 = 0; s4 < NP / 4; ++s4)
#pragma unroll
      for (int tj = 0; tj < nt; ++tj) {
        const int row = 4 * s4 + kk, col = 16 * tj + i16;
        tb[s4][tj] = fma(coef, Z[row * ld + col], sy[row]);
      }
    wave_lds_sync();
    return true;
  }
}

template <int N>
__global__ __launch_bounds__(64) void k_trd_vectors(
    int64_t ncol, const int* __restrict__ nloc, const double* __restrict__ de,
    const double* __restrict__ lam, const double* __restrict__ bvec, double rho,
    const double* __restrict__ rho_col, int prior, int combine, double null_rel,
    const double* __restrict__ Ug, double* __restrict__ T, int* __restrict__ flag) {
  // one column per block: no loop-carried state (a grid-stride loop here
  // costs hundreds of SGPR spills)
  __shared__ double slice[e3_shape<N>::SLICE];
  double tb[e3_shape<N>::NP / 4][e3_shape<N>::NP / 16];
  const int64_t g = blockIdx.x;
  if (g < ncol)
    e3_column<N, false>(g, slice, nloc, de, lam, bvec, rho, rho_col, prior, combine, null_rel, Ug,
                        T, flag, tb, [] {});
}

constexpr int E3_PRE = 2;  // 16-level chunks of X loaded during E3's W products

// E3 with the transform fused: a block of XM_C = 4 wavefronts takes 4
// consecutive columns of the pass (one each, LDS slices side by side), and
// once every wavefront holds its T the block applies X_a = X T' to the 4
// columns (mda_xform.hpp) in place, so T never reaches HBM.  Columns left to
// the Jacobi fallback are not updated here (the transform restricted to the
// flagged columns does them after the fallback).  Blocks take groups in an
// XCD-contiguous order (gridDim.x a multiple of 8), so the groups that share
// X's cache lines run on one XCD at about the same time.
template <int N>
__global__ __launch_bounds__(256) void k_trd_vectors_x(
    int64_t ncol, const int* __restrict__ nloc, const double* __restrict__ de,
    const double* __restrict__ lam, const double* __restrict__ bvec, double rho,
    const double* __restrict__ rho_col, int prior, double null_rel,
    const double* __restrict__ Ug, int* __restrict__ flag, double* __restrict__ X,
    int64_t x_ncol, int64_t nlev, int64_t col0, int* __restrict__ status) {
  static_assert(N % 16 == 0 && N <= 64, "N must be a multiple of 16, at most 64");
  constexpr int SL = e3_shape<N>::SLICE;
  static_assert(xm_xs_size<N>() <= XM_C * SL, "X staging exceeds the block's LDS");
  __shared__ double lds[XM_C * SL];
  __shared__ int sok[XM_C];
  const int tid = threadIdx.x, wv = tid >> 6;
  const int64_t nb = gridDim.x;
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t g0 = vb * XM_C;  // within the pass
  if (g0 >= ncol) return;         // block-uniform
  const int ncl = (int)(ncol - g0 < XM_C ? ncol - g0 : XM_C);
  double tb[N / 4][N / 16];
#pragma unroll
  for (int s4 = 0; s4 < N / 4; ++s4)
#pragma unroll
    for (int tj = 0; tj < N / 16; ++tj) tb[s4][tj] = 0.0;
  // the first E3_PRE 16-level chunks of the block's X, each thread its
  // share, loaded by every wavefront during its W products (or at once by a
  // wavefront without them)
  double pre[E3_PRE][xm_nv<N>()];
  bool loaded = false;
  auto preload = [&]() {
    if (loaded) return;
#pragma unroll
    for (int ci = 0; ci < E3_PRE; ++ci)
      xm_load<N>(pre[ci], X, x_ncol, nlev, col0 + g0, ncl, (int64_t)XM_L * ci);
    loaded = true;
  };
  bool ok = false;
  if (wv < ncl)
    ok = e3_column<N, true>(g0 + wv, lds + wv * SL, nloc, de, lam, bvec, rho, rho_col, prior,
                            MDA_COMBINE_MATRIX, null_rel, Ug, nullptr, flag, tb, preload);
  preload();
  if ((tid & 63) == 0) sok[wv] = ok ? 1 : 0;
  __syncthreads();
  unsigned store = 0u;
  for (int c = 0; c < ncl; ++c) store |= sok[c] ? 1u << c : 0u;
  if (store == 0u || (*status & MDA_ST_SKIP_UPDATE)) return;  // block-uniform
  xm_tprime<N>(tb);
  xm_apply<N, E3_PRE>(X, x_ncol, nlev, col0 + g0, ncl, store, tb, lds, status, pre);
}

// ---------------------------------------------------------------------------
// Ensembles wider than one wavefront (k = 128): the same three stages with
// four wavefronts per column for E1 and E3 and, for the eigenvalues,
// bisection on Sturm counts.
// ---------------------------------------------------------------------------

constexpr int E1W_CH = 8;  // step chunk of the k = 128 E1 (static column ranges)

// E2 by bisection on Sturm counts (one wavefront per column, lane j owns
// eigenvalues j, j + 64, ...): count(x) = #{i : q_i < 0} with
// q_0 = d_0 - x, q_i = (d_i - x) - e_{i-1}^2 / q_{i-1}, |q_i| kept >= pivmin
// (Golub & Van Loan, sec. 8.4.1).  Starts from the Gershgorin interval cut
// into 65 cells by one multisection sweep and halves until the interval is
// within 1e-13 of the eigenvalue (relative).  Ascending by construction.
template <int N>
__global__ __launch_bounds__(256) void k_trd_eigvals_bisect(int64_t ncol,
                                                            const int* __restrict__ nloc,
                                                            const double* __restrict__ de,
                                                            double* __restrict__ lam,
                                                            int* __restrict__ flag) {
  constexpr int n = N, NQ = (N + 63) / 64;
  __shared__ double sd[4][N], se2[4][N], msx[4][64];
  __shared__ int msc[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t g = (int64_t)blockIdx.x * 4 + wv;
  if (g >= ncol || nloc[g] == 0) return;  // whole wavefront; no block barriers below
  const double* dg = de + (size_t)g * 2 * n;
  double* d = sd[wv];
  double* e2 = se2[wv];
  for (int i = lane; i < n; i += 64) {
    d[i] = dg[i];
    const double e = i < n - 1 ? dg[n + i] : 0.0;
    e2[i] = e * e;
  }
  wave_lds_sync();
  double lo = 0.0, hi = 0.0, emax = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = sqrt(e2[i]) + (i > 0 ? sqrt(e2[i - 1]) : 0.0);
    lo = i == 0 ? d[i] - r : fmin(lo, d[i] - r);
    hi = i == 0 ? d[i] + r : fmax(hi, d[i] + r);
    emax = fmax(emax, e2[i]);
  }
  const double eps = 2.220446049250313e-16;
  const double tnorm = fmax(fabs(lo), fabs(hi));
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
  lo -= 2.0 * eps * tnorm + pivmin;
  hi += 2.0 * eps * tnorm + pivmin;
  // one multisection sweep: lane L counts the eigenvalues below the L-th of
  // 64 interior points, and every eigenvalue starts from the cell between
  // the points whose counts bracket it (six halvings for one sweep)
  {
    const double xl = lo + (hi - lo) * ((double)(lane + 1) / 65.0);
    double qv = 1.0;
    int cnt = 0;
    for (int i = 0; i < n; ++i) {
      const double ei = i > 0 ? e2[i - 1] : 0.0;
      double rq = __builtin_amdgcn_rcp(qv);
      rq = fma(rq, fma(-qv, rq, 1.0), rq);
      double v = (d[i] - xl) - ei * rq;
      if (fabs(v) < pivmin) v = -pivmin;
      qv = v;
      cnt += v < 0.0;
    }
    msx[wv][lane] = xl;
    msc[wv][lane] = cnt;
  }
  wave_lds_sync();
  double a[NQ], b[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int j = lane + 64 * q;
    a[q] = lo;
    b[q] = hi;
    // rolled (by 4): fully unrolled, the 64 points and counts would sit in
    // registers (216 VGPRs, two waves per SIMD); rolled the kernel needs 44
#pragma unroll 4
    for (int m = 0; m < 64; ++m) {
      const double xm = msx[wv][m];
      if (msc[wv][m] <= j) a[q] = fmax(a[q], xm);
      else b[q] = fmin(b[q], xm);
    }
  }
  // converged at a relative width of 1e-13: the eigenvector checks ask for a
  // residual within 1e-12 of ||T|| and the transform for 1e-10, and the last
  // eight halvings to 2 eps would add 14 % to the sweeps
  const double rel = 1e-13;
  for (int it = 0; it < 128; ++it) {
    bool open = false;
    double x[NQ], qv[NQ];
    int cnt[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      x[q] = 0.5 * (a[q] + b[q]);
      open |= (lane + 64 * q < n) && b[q] - a[q] > rel * fmax(fabs(a[q]), fabs(b[q])) + pivmin
              && x[q] != a[q] && x[q] != b[q];
      qv[q] = 1.0;
      cnt[q] = 0;
    }
    if (!wave_any(open)) break;
    for (int i = 0; i < n; ++i) {
      const double di = d[i], ei = i > 0 ? e2[i - 1] : 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        // 1/q by the hardware reciprocal and one Newton step
        double rq = __builtin_amdgcn_rcp(qv[q]);
        rq = fma(rq, fma(-qv[q], rq, 1.0), rq);
        double v = (di - x[q]) - ei * rq;
        if (fabs(v) < pivmin) v = -pivmin;
        qv[q] = v;
        cnt[q] += v < 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (cnt[q] > lane + 64 * q) b[q] = x[q];
      else a[q] = x[q];
    }
  }
  double* lg = lam + (size_t)g * n;
  bool bad = false;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int j = lane + 64 * q;
    if (j < n) {
      const double v = 0.5 * (a[q] + b[q]);
      lg[j] = v;
      bad |= !(fabs(v) <= 1.7e308);
    }
  }
  if (wave_any(bad) && lane == 0) flag[g] = 1;
}

// Swaps a double between the two lanes of each even/odd pair (DPP
// quad_perm [1, 0, 3, 2]).
__device__ inline double pair_swap(double v) { return dpp_f64<0xB1>(v); }

// Sum over the block's NW wavefronts; `red` holds one slot per wavefront.
// Ends with a block barrier.
template <int NW>
__device__ inline double block_total(double v, double* red) {
  const double s = wave_total(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if constexpr (NW == 4) return (red[0] + red[1]) + (red[2] + red[3]);
  else return red[0] + red[1];
}

// E1 for N = 128 with four wavefronts per column: R consecutive rows belong
// to a group of G = 2R threads, thread h of the group holding those rows'
// entries of the columns G m + h (64 doubles in all).  At 1024 x 512 x 40:
// R = 1 (thread pairs) 260 ms, R = 2 219 ms, R = 4 239 ms per analysis (more
// row sums per step, no further triangle savings); R = 2 runs.  At N = 64
// (two wavefronts, four waves per SIMD) it took 26.7 ms against 19.8 ms for
// the one-wavefront k_trd, whose steps need no block barrier.
// Every u and w value read from LDS feeds R rows' FMAs, and two columns share
// a CU at two waves per SIMD.  The same Householder steps as the
// one-row-per-thread form: |x|^2 and p^T u by block reductions, A u by
// partial row sums combined across the group by DPP, the rank-2 update on
// each thread's share of the rows.
template <int N, int R>
__global__ __launch_bounds__(2 * N, 2) void k_trd_wr(int64_t ncol, const double* __restrict__ C,
                                                     int c32, const int* __restrict__ nloc,
                                                     double rho,
                                                     const double* __restrict__ rho_col,
                                                     int prior, double* __restrict__ de,
                                                     double* __restrict__ Uout) {
  static_assert(N == 128 && (R == 2 || R == 4), "four wavefronts per column");
  constexpr int n = N, G = 2 * R, M = N / G, NW = 2 * N / 64;
  static_assert(E1W_CH % G == 0, "chunks start on a column group");
  __shared__ double ub[N], wb[N], red[2][4], xs[2];
  const int t = threadIdx.x, h = t & (G - 1);
  const int rb = (t / G) * R;  // first row of the group
  const int64_t g = blockIdx.x;
  if (g >= ncol || nloc[g] == 0) return;  // block-uniform
  const double rr = rho_col ? rho_col[g] : rho;
  const double sigma = prior ? (double)(n - 1) / rr : (double)(n - 1);
  double a[R][M];  // a[i][m] = A[rb + i][G m + h]
  if (c32) {  // FP32 entries, widened
    const float* Cg = reinterpret_cast<const float*>(C) + (size_t)g * n * n;
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int m = 0; m < M; ++m) a[i][m] = (double)Cg[(size_t)(rb + i) * n + G * m + h];
  } else {
    const double* Cg = C + (size_t)g * n * n;
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int m = 0; m < M; ++m) a[i][m] = Cg[(size_t)(rb + i) * n + G * m + h];
  }
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int m = 0; m < M; ++m) a[i][m] += rb + i == G * m + h ? sigma : 0.0;
  double* U = Uout + (size_t)g * n * n;
  double* dg = de + (size_t)g * 2 * n;
  const int lbase = (t & 63) & ~(G - 1);  // the group's first lane
  // xc[i]: A[rb + i][j] of the current step, held by the whole group
  double xc[R];
#pragma unroll
  for (int i = 0; i < R; ++i) xc[i] = __shfl(a[i][0], lbase);
  // sum over the group's lanes, left in every lane
  auto group_sum = [&](double v) {
    v += dpp_f64<0xB1>(v);                   // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E>(v);                   // quad_perm [2, 3, 0, 1]
    if constexpr (G == 8) v += dpp_f64<0x141>(v);  // row_half_mirror
    return v;
  };
  constexpr int NCH = (N - 2 + E1W_CH - 1) / E1W_CH;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int c0 = E1W_CH * ch, m0 = c0 / G;
    const int jend = (E1W_CH * ch + E1W_CH < n - 2) ? E1W_CH * ch + E1W_CH : n - 2;
#pragma unroll 1
    for (int j = c0; j < jend; ++j) {
      double xi[R], s2 = 0.0;
#pragma unroll
      for (int i = 0; i < R; ++i) {
        xi[i] = rb + i > j ? xc[i] : 0.0;
        s2 = fma(xi[i], xi[i], s2);
        if (h == 0 && rb + i == j + 1) xs[0] = xc[i];
        if (h == 0 && rb + i == j) xs[1] = xc[i];
      }
      const double nrm2 = block_total<NW>(h == 0 ? s2 : 0.0, red[0]);
      const double x0 = xs[0];
      const bool nz = nrm2 > 0.0;
      const double alpha = nz ? (x0 > 0.0 ? -1.0 : 1.0) * sqrt(nrm2) : 0.0;
      const double tau = nz ? 2.0 / (nrm2 - 2.0 * alpha * x0 + alpha * alpha) : 0.0;
      double u[R];
#pragma unroll
      for (int i = 0; i < R; ++i) u[i] = nz ? xi[i] - (rb + i == j + 1 ? alpha : 0.0) : 0.0;
      if (t == 0) { dg[j] = xs[1]; dg[n + j] = alpha; }
      if (h == 0) {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          U[(size_t)j * n + rb + i] = rb + i == 0 ? tau : u[i];
          ub[rb + i] = u[i];
        }
      }
      __syncthreads();
      // groups whose rows are all final (rb + R - 1 <= j) skip the
      // matrix-vector product and the update: no share of the LDS
      // broadcasts of u and w (their diagonal entries were taken before)
      const bool live = rb + R - 1 > j;
      double p[R], q[R], pu = 0.0;  // two accumulation chains per row
#pragma unroll
      for (int i = 0; i < R; ++i) p[i] = q[i] = 0.0;
      if (live) {
#pragma unroll
        for (int m = m0; m < M; ++m) {
          const double uc = ub[G * m + h];
#pragma unroll
          for (int i = 0; i < R; ++i) {
            if (m & 1) q[i] = fma(a[i][m], uc, q[i]);
            else p[i] = fma(a[i][m], uc, p[i]);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < R; ++i) {
        p[i] = tau * group_sum(p[i] + q[i]);
        pu = fma(p[i], u[i], pu);
      }
      const double K = 0.5 * tau * block_total<NW>(h == 0 ? pu : 0.0, red[1]);
      double w[R];
#pragma unroll
      for (int i = 0; i < R; ++i) w[i] = rb + i > j ? p[i] - K * u[i] : 0.0;
      if (h == 0) {
#pragma unroll
        for (int i = 0; i < R; ++i) wb[rb + i] = w[i];
      }
      __syncthreads();
      double nx[R];  // A[rb + i][j + 1] if this thread holds column j + 1
#pragma unroll
      for (int i = 0; i < R; ++i) nx[i] = 0.0;
      if (live) {
#pragma unroll
        for (int m = m0; m < M; ++m) {
          if ((m & 3) == 0) lds_reload_point();
          const int c = G * m + h;
          const double wc = wb[c], uc = ub[c];
          const bool take = m <= m0 + E1W_CH / G && c == j + 1;
#pragma unroll
          for (int i = 0; i < R; ++i) {
            a[i][m] = fma(-u[i], wc, fma(-w[i], uc, a[i][m]));  // two FMAs per entry
            if (take) nx[i] = a[i][m];
          }
        }
      }
      const int src = lbase + ((j + 1) & (G - 1));
#pragma unroll
      for (int i = 0; i < R; ++i) xc[i] = __shfl(nx[i], src);
    }
  }
  // last 2x2 block from the owners' registers (row n-2 gives e_{n-2}):
  // rows n-2, n-1 are the last two rows of the last group, columns n-2 and
  // n-1 the lanes G-2 and G-1 at m = M-1
  if (rb + R - 2 == n - 2) {
    if (h == G - 2) dg[n - 2] = a[R - 2][M - 1];
    if (h == G - 1) {
      dg[2 * n - 2] = a[R - 2][M - 1];
      dg[n - 1] = a[R - 1][M - 1];
      dg[2 * n - 1] = 0.0;
    }
  }
}

// E3 for N = 128 with four wavefronts per column.  Eigenvector j belongs to
// the thread pair (2j, 2j + 1) of one wavefront; thread 2j + h holds the
// components i = 2m + h (m < N/2) in registers, so the pair shares every dot
// product and update evenly, nothing spills, and the block keeps all four
// SIMDs of its CU busy (one column per CU: 147 KB of LDS).
//   Twisted factorisation with its recurrences in Z (LDS): the even thread
//   runs the top-down recurrence D+_i, the odd one the bottom-up D-_i, side
//   by side; each stores the first half of its sequence and, over the
//   second half, forms gamma_i = D+_i + D-_i - (d_i - l) from the partner's
//   stored value.  After the twist kt is known, the one missing stretch of
//   D+ or D- is recomputed, then the vector is built from kt downwards (even
//   thread) and upwards (odd thread) at once.  The back-transform runs in
//   compact-WY blocks of 16 reflectors on FP64 MFMA.  Checks, near-pair
//   (Lowdin) step, W on FP64 MFMA and the T layout as in the one-wavefront
//   kernel.
template <int N>
__global__ __launch_bounds__(2 * N) void k_trd_vectors_w4(
    int64_t ncol, const int* __restrict__ nloc, const double* __restrict__ de,
    const double* __restrict__ lam, const double* __restrict__ bvec, double rho,
    const double* __restrict__ rho_col, int prior, int combine, double null_rel,
    const double* __restrict__ Ug, double* __restrict__ T, int* __restrict__ flag) {
  // N = 64 (two wavefronts, four columns per CU) measured slower than the
  // one-wavefront kernel at 1024 x 512 x 40: 33.6 against 29.0 ms
  static_assert(N == 128, "four wavefronts per column");
  // Z's row stride N + 6: 64-bit LDS accesses of an MFMA operand tile (16
  // rows x 4 columns) fall two lanes per bank pair, of an accumulator tile
  // (4 rows x 16 columns) at most three (N + 1: four and four), and the
  // vector-per-thread-pair rows and columns two
  constexpr int n = N, ld = N + 6, BAND = 8, H = N / 2, M = N / 2;
  constexpr int NW = 2 * N / 64;  // wavefronts per column
  __shared__ double Z[N * ld];
  // sk (column sums of U_s, after the back-transform) reuses swt (near-pair
  // weights)
  __shared__ double sd[N], se[N], sl[N], sb[N], swt[N], sy[N];
  double* sk = swt;
  __shared__ double dots[N][BAND];
  __shared__ int s_why, s_near;
  const int t = threadIdx.x, j = t >> 1, h = t & 1;
  const int64_t g = blockIdx.x;
  if (g >= ncol || nloc[g] == 0 || flag[g]) return;  // block-uniform
  const double r = rho_col ? rho_col[g] : rho;
  double* Tg = T + (size_t)g * n * n;
  const double* dg = de + (size_t)g * 2 * n;
  if (t < n) {
    sd[t] = dg[t];
    se[t] = t < n - 1 ? dg[n + t] : 0.0;
    sl[t] = lam[(size_t)g * n + t];
    sb[t] = bvec[(size_t)g * n + t];
  }
  if (t == 0) { s_why = 0; s_near = 0; }
  __syncthreads();
  double tnorm = 0.0;
#pragma unroll 8
  for (int i = 0; i < N; ++i) tnorm = fmax(tnorm, fabs(sd[i]) + 2.0 * fabs(se[i]));
  const double tiny = 1e-300 + 2.220446049250313e-16 * 1e-3 * tnorm;
  const double l = sl[j];
  double* Zj = Z + j;  // component i of vector j at Zj[i * ld]
  // ---- twisted factorisation ----
  // step s: i = s (h = 0, D+_i) or N-1-s (h = 1, D-_i),
  //   x_i = (d_i - l) - e (e / safe(x_prev)),  e = e_{i-1} (down) or e_i (up)
  double xp = 1.0, gbest = INFINITY;
  int kt = N - 1;
#pragma unroll 4
  for (int s = 0; s < H; ++s) {
    const int i = h ? N - 1 - s : s;
    const double ee = h ? se[i] : (i > 0 ? se[i > 0 ? i - 1 : 0] : 0.0);
    const double x = (sd[i] - l) - ee * fast_div(ee, safe_pivot(xp, tiny));
    Zj[i * ld] = x;
    xp = x;
  }
  wave_lds_sync();  // the partner's first halves are in Z
#pragma unroll 4
  for (int s = H; s < N; ++s) {
    const int i = h ? N - 1 - s : s;
    const double ee = h ? se[i] : se[i > 0 ? i - 1 : 0];
    const double x = (sd[i] - l) - ee * fast_div(ee, safe_pivot(xp, tiny));
    const double gam = fabs(x + Zj[i * ld] - (sd[i] - l));
    // ties go to the higher index (i ascends for h = 0, descends for h = 1)
    if (h ? gam < gbest : gam <= gbest) { gbest = gam; kt = i; }
    xp = x;
  }
  {
    const double og = pair_swap(gbest);
    const int ok = __builtin_amdgcn_mov_dpp(kt, 0xB1, 0xf, 0xf, true);
    if (og < gbest || (og == gbest && ok > kt)) { gbest = og; kt = ok; }
  }
  // the missing stretch: D-_i for kt < i < H (bottom-up from D-_H), or D+_i
  // for H <= i < kt (top-down from D+_{H-1}), by the even thread, over slots
  // whose stored values are no longer needed
  // Both loops below run in unrolled blocks of 8 steps with one wave-wide
  // exit test per block, so the LDS loads (and, for the vector, the
  // multipliers, which do not depend on the chain) are issued ahead.
  {
    const bool up = kt < H;
    double xq = Zj[(up ? H : H - 1) * ld];
    for (int r0 = 0; r0 < H; r0 += 8) {
      if (!wave_any(h == 0 && (up ? H - 1 - r0 >= kt + 1 : H + r0 <= kt - 1))) break;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = up ? H - 1 - (r0 + u) : H + r0 + u;
        const bool act = h == 0 && (up ? i >= kt + 1 : i <= kt - 1);
        const int ic = i < 1 ? 1 : (i > N - 1 ? N - 1 : i);  // in bounds when inactive
        const double ee = up ? se[ic] : se[ic - 1];
        const double dl = sd[ic] - l;
        if (act) {
          const double x = dl - ee * fast_div(ee, safe_pivot(xq, tiny));
          Zj[i * ld] = x;
          xq = x;
        }
      }
    }
  }
  wave_lds_sync();
  // z_kt = 1; even thread downwards z_i = -(e_i / safe(D+_i)) z_{i+1},
  // odd thread upwards z_{i+1} = -(e_i / safe(D-_{i+1})) z_i.  Each row is
  // read (its D value) before it is overwritten (its z), once.
  if (h == 0) Zj[kt * ld] = 1.0;
  {
    double zc = 1.0;
    for (int r0 = 0; r0 < N - 1; r0 += 8) {
      if (!wave_any(h ? kt + r0 <= N - 2 : kt - 1 - r0 >= 0)) break;
      double mlt[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = h ? kt + r0 + u : kt - 1 - r0 - u;
        const bool act = h ? i <= N - 2 : i >= 0;
        const int ic = act ? i : 0, rc = act ? (h ? i + 1 : i) : 0;
        const double m = -fast_div(se[ic], safe_pivot(Zj[rc * ld], tiny));
        mlt[u] = act ? m : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = h ? kt + r0 + u : kt - 1 - r0 - u;
        const bool act = h ? i <= N - 2 : i >= 0;
        if (act) {
          zc = mlt[u] * zc;
          Zj[(h ? i + 1 : i) * ld] = zc;
        }
      }
    }
  }
  wave_lds_sync();
  double z[M];
  {
    double nrm = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      z[m] = Zj[(2 * m + h) * ld];
      nrm += z[m] * z[m];
    }
    nrm += pair_swap(nrm);
    const double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) z[m] *= inv;
  }
  const bool pri = prior != 0;
  const double sigma = pri ? (double)(n - 1) / r : (double)(n - 1);
  const double null_tol = fmax(2e-12 * sigma, null_rel * tnorm);
  const bool skip = fabs(l - sigma) <= null_tol;
  const double wj = skip ? 0.0 : 1.0 - sigma / l;
  int why = 0;  // 2: residual, 4: near-pair orthogonality, 8: lambda < sigma
  {
    // residual (T - l I) z from the pair's registers: the neighbours of
    // component 2m + h are the partner's m - 1 + h and m + h
    double res = 0.0, pprev = 0.0, pcur = pair_swap(z[0]);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const double pnext = m + 1 < M ? pair_swap(z[m + 1 < M ? m + 1 : m]) : 0.0;
      const int i = 2 * m + h;
      const double left = h ? pcur : pprev;    // component i - 1
      const double right = h ? pnext : pcur;   // component i + 1
      double v = (sd[i] - l) * z[m];
      if (i > 0) v += se[i > 0 ? i - 1 : 0] * left;
      v += se[i] * right;  // se[N - 1] = 0
      res = fmax(res, fabs(v));
      pprev = pcur;
      pcur = pnext;
    }
    res = fmax(res, pair_swap(res));
    if (!skip) {
      if (!(res <= 1e-12 * tnorm)) why |= 2;
      if (!(l >= sigma - null_tol)) why |= 8;
    }
  }
  if (h == 0) {
    sy[j] = skip ? 1.0 : 0.0;  // skip flags, shared
    swt[j] = wj;
  }
#pragma unroll
  for (int m = 0; m < M; ++m) Zj[(2 * m + h) * ld] = z[m];
  __syncthreads();
  // near pairs: symmetric (Lowdin) correction, then a check
  for (int pass = 0; pass < 2; ++pass) {
    if (h == 0) {
#pragma unroll
      for (int q = 0; q < BAND; ++q) dots[j][q] = 0.0;
    }
    bool near = false;
    if (!skip) {
      for (int q = j + 1; q < n && (sl[q] - l) < 1e-3 * tnorm; ++q) {
        if (sy[q] != 0.0) continue;
        near = true;
        double dot = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) dot += z[m] * Z[(2 * m + h) * ld + q];
        dot += pair_swap(dot);
        if (q - j > BAND) {
          if (fabs(dot) * (wj + swt[q]) > 1e-12) why |= 4;
        } else {
          if (h == 0) dots[j][q - j - 1] = dot;
          if (pass == 1 && fabs(dot) * (wj + swt[q]) > 1e-12) why |= 4;
        }
      }
    }
    if (pass == 0 && near) atomicOr(&s_near, 1);
    __syncthreads();
    if (pass == 1 || s_near == 0) break;
    if (!skip) {
      for (int q = j - BAND; q <= j + BAND; ++q) {
        if (q < 0 || q >= n || q == j) continue;
        const double c = q > j ? dots[j][q - j - 1] : dots[q][j - q - 1];
        if (c != 0.0) {
#pragma unroll
          for (int m = 0; m < M; ++m) z[m] -= 0.5 * c * Z[(2 * m + h) * ld + q];
        }
      }
    }
    __syncthreads();  // every read of the uncorrected vectors is done
#pragma unroll
    for (int m = 0; m < M; ++m) Zj[(2 * m + h) * ld] = z[m];
    __syncthreads();
  }
  if (why) atomicOr(&s_why, why);
  __syncthreads();
  if (s_why != 0) {
    if (t == 0) flag[g] = s_why;
    return;  // block-uniform
  }
  // ---- z <- Q z in compact-WY blocks on FP64 MFMA ----
  // Q = Q_0 Q_1 ... Q_{N/16-1}, Q_b = H_16b ... H_16b+15 = I - V S V^T with V the
  // block's 16 Householder vectors (rows U[jj][i], i >= 1) and S upper
  // triangular, S^-1 = diag(1/tau) + strictly_upper(V^T V).  The vectors
  // move from Z into accumulator tiles (wavefront wv: vectors 32 wv .. 32 wv
  // + 31, all N components: zt[ci][vv][q] = Z[16 ci + kk + 4 q][32 wv +
  // 16 vv + i16], which is also the B layout of Z as a K = component
  // operand); per block, last block first:
  //   W1 = V^T Z (MFMA), S W1 by back-substitution (S^-1 W2 = W1, rows from
  //   the last reflector up; tau = 0 gives a zero row), Z -= V W2 (MFMA).
  // V rows at components <= 16 b are zero, so component tiles below b are
  // skipped.  Z's LDS area is free meanwhile: the V blocks (double
  // buffered, the next one loaded during this one), V^T V and W live there.
  {
    const int lane = t & 63, wv = t >> 6;
    const int i16 = lane & 15, kk = lane >> 4;
    constexpr int NB = N / 16, BW = 16;  // blocks of 16 reflectors (N - 2 real ones)
    const double* Ucol = Ug + (size_t)g * n * n;
    d4 zt[NB][2];
#pragma unroll
    for (int ci = 0; ci < NB; ++ci)
#pragma unroll
      for (int vv = 0; vv < 2; ++vv)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          zt[ci][vv][q] = Z[(16 * ci + kk + 4 * q) * ld + 32 * wv + 16 * vv + i16];
    __syncthreads();  // Z's area is reused below
    // rows of N + 6 (V, read both as the A operand V^T and as the B-side
    // operand of V W2: two and three lanes per bank pair; rows of N: sixteen
    // and four) and N + 16 (W, tiles written and read 16 columns wide)
    constexpr int LV = N + 6, LW = N + 16;
    static_assert(2 * BW * LV + BW * BW + BW * LW <= N * ld, "V, G and W exceed Z's area");
    double* vbuf0 = Z;               // [16][LV] block b even
    double* vbuf1 = Z + BW * LV;     // [16][LV] block b odd
    double* Gs = Z + 2 * BW * LV;    // [16][16] V^T V
    double* Ws = Gs + BW * BW;       // [16][LW] W1 -> W2
    constexpr int NPF = BW * N / (2 * N);  // staged doubles per thread
    double pf[NPF];
    auto fetch = [&](int b) {
#pragma unroll
      for (int p = 0; p < NPF; ++p) {
        const int idx = t + 2 * N * p, c = idx / N, i = idx % N, jj = BW * b + c;
        pf[p] = jj <= N - 3 ? Ucol[(size_t)jj * N + i] : 0.0;
      }
    };
    auto put = [&](double* dst) {
#pragma unroll
      for (int p = 0; p < NPF; ++p) {
        const int idx = t + 2 * N * p;
        dst[(idx / N) * LV + idx % N] = pf[p];
      }
    };
    fetch(NB - 1);
    put(((NB - 1) & 1) ? vbuf1 : vbuf0);
#pragma unroll
    for (int b = NB - 1; b >= 0; --b) {
      const double* V = (b & 1) ? vbuf1 : vbuf0;  // V[c * LV + i] = component i of vector 16b + c
      __syncthreads();  // block b staged; block b + 1's readers are done
      if (b > 0) fetch(b - 1);
      // W1 = V^T Z for the wavefront's two vector tiles; wavefront 0 also V^T V
      d4 w1[2] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
      d4 gg = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 4 * b; s < N / 4; ++s) {
        const int comp = 4 * s + kk;
        const double av = comp == 0 ? 0.0 : V[i16 * LV + comp];  // A = V^T, B = V for V^T V
#pragma unroll
        for (int vv = 0; vv < 2; ++vv)
          w1[vv] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, zt[s / 4][vv][s % 4], w1[vv], 0, 0, 0);
        if (wv == 0) gg = __builtin_amdgcn_mfma_f64_16x16x4f64(av, av, gg, 0, 0, 0);
      }
#pragma unroll
      for (int vv = 0; vv < 2; ++vv)
#pragma unroll
        for (int q = 0; q < 4; ++q) Ws[(kk + 4 * q) * LW + 32 * wv + 16 * vv + i16] = w1[vv][q];
      if (wv == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) Gs[(kk + 4 * q) * BW + i16] = gg[q];
      }
      __syncthreads();
      // W2 = S W1: W2[m] = tau_m (W1[m] - sum_{c > m} G[m][c] W2[c]), one
      // vector per thread
      if (t < N) {
        double a[BW];
#pragma unroll
        for (int c = 0; c < BW; ++c) a[c] = Ws[c * LW + t];
#pragma unroll
        for (int m = BW - 1; m >= 0; --m) {
          const double w = V[m * LV] * a[m];  // V[m][0] holds tau
          Ws[m * LW + t] = w;
#pragma unroll
          for (int c = 0; c < m; ++c) a[c] = fma(-Gs[c * BW + m], w, a[c]);
        }
      }
      __syncthreads();
      // Z -= V W2 on the component tiles at or above b
#pragma unroll
      for (int ci = b; ci < NB; ++ci) {
        const int comp = 16 * ci + i16;
#pragma unroll
        for (int s4 = 0; s4 < BW / 4; ++s4) {
          const double av = comp == 0 ? 0.0 : -V[(4 * s4 + kk) * LV + comp];
#pragma unroll
          for (int vv = 0; vv < 2; ++vv)
            zt[ci][vv] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                av, Ws[(4 * s4 + kk) * LW + 32 * wv + 16 * vv + i16], zt[ci][vv], 0, 0, 0);
        }
      }
      if (b > 0) put(((b - 1) & 1) ? vbuf1 : vbuf0);
    }
    __syncthreads();  // the last block's readers are done with Z's area
#pragma unroll
    for (int ci = 0; ci < NB; ++ci)
#pragma unroll
      for (int vv = 0; vv < 2; ++vv)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          Z[(16 * ci + kk + 4 * q) * ld + 32 * wv + 16 * vv + i16] = zt[ci][vv][q];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) z[m] = Zj[(2 * m + h) * ld];
  }
  // ---- U_s = Z diag(f), y' for wbar ----
  const double sc = pri ? 1.0 : r;
  const double f = skip ? 0.0 : sqrt(1.0 / sqrt(sigma) - 1.0 / sqrt(l));
  {
    double yb = 0.0, cs = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      yb += z[m] * sb[2 * m + h];
      Zj[(2 * m + h) * ld] = z[m] * f;
      cs += z[m] * f;
    }
    yb += pair_swap(yb);
    cs += pair_swap(cs);
    if (skip) yb = 0.0;
    else yb *= (1.0 / sigma - 1.0 / l);
    if (h == 0) {
      sy[j] = f > 0.0 ? yb / f : 0.0;
      sk[j] = cs;
    }
  }
  __syncthreads();
  const double coef = sqrt((double)(n - 1)) * sqrt(sc);
  const double fs = 1.0 / sqrt(sigma);
  // wbar_j = s [ b_j / sigma - sum_e U_s[j][e] y'_e ], row j by its pair
  double wbar;
  {
    double acc = 0.0, acs = 0.0;
#pragma unroll 8
    for (int m = 0; m < M; ++m) {
      const double us = Z[j * ld + 2 * m + h];
      acc += us * sy[2 * m + h];
      acs += us * sk[2 * m + h];
    }
    acc += pair_swap(acc);
    acs += pair_swap(acs);
    wbar = sc * (sb[j] / sigma - acc);
    if (combine == MDA_COMBINE_ELEMENTWISE) {
      // t_i = wbar_i + coef (f(sigma) - sum_a U_s[i][a] (sum_j U_s[j][a]))
      for (int idx = t; idx < n * n; idx += 2 * N) Tg[idx] = 0.0;
      __syncthreads();
      if (h == 0) Tg[(size_t)j * n + j] = wbar + coef * (fs - acs);
      return;
    }
  }
  __syncthreads();  // sy (y') reads done
  if (h == 0) sy[j] = wbar;
  __syncthreads();
  // ---- W = coef (f(sigma) I - U_s U_s^T) on FP64 MFMA, tiles round-robin ----
  constexpr int nt = N / 16;
  const int lane = t & 63, wv = t >> 6;
  const int i16 = lane & 15, kk = lane >> 4;
  int tile = 0;
  for (int ti = 0; ti < nt; ++ti) {
    for (int tj = 0; tj <= ti; ++tj, ++tile) {
      if ((tile % NW) != wv) continue;
      d4 am = {0.0, 0.0, 0.0, 0.0};
      const double* ra = Z + (16 * ti + i16) * ld + kk;
      const double* rb = Z + (16 * tj + i16) * ld + kk;
#pragma unroll 8
      for (int s4 = 0; s4 < N / 4; ++s4)
        am = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[4 * s4], rb[4 * s4], am, 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 16 * ti + kk + 4 * q, col = 16 * tj + i16;
        const double wd = (row == col ? fs : 0.0) - am[q];
        Tg[(size_t)row * n + col] = fma(coef, wd, sy[row]);
        if (ti != tj) Tg[(size_t)col * n + row] = fma(coef, wd, sy[col]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
template <int N>
static void launch_eig_n(hipStream_t s, const letkf_solve_args& a, double* de, double* lam,
                         int* flag) {
  const int prior = a.inflation_form == MDA_INFLFORM_PRIOR;
  launch_zero(s, flag, sizeof(int) * a.ncol);
  MDA_REQUIRE(a.ncol < (int64_t)1 << 31, "eigensolver grid too large");
  const unsigned nb = (unsigned)a.ncol;  // one column per block
  hipLaunchKernelGGL(k_trd<N>, dim3(nb), dim3(64), 0, s, a.ncol, a.C, a.c32, a.nloc, a.rho,
                     a.rho_col, prior, de, a.U);
  MDA_LAUNCH_CHECK();
  if (a.ev_stage[0]) MDA_HIP(hipEventRecord(a.ev_stage[0], s));
  hipLaunchKernelGGL(k_trd_eigvals<N>, dim3((unsigned)((a.ncol + 63) / 64)), dim3(64), 0, s,
                     a.ncol, a.nloc, de, lam, flag);
  MDA_LAUNCH_CHECK();
  if (a.ev_stage[1]) MDA_HIP(hipEventRecord(a.ev_stage[1], s));
  if constexpr (N % 16 == 0) {
    if (a.fuse) {
      MDA_REQUIRE(a.combine == MDA_COMBINE_MATRIX && a.X && a.status,
                  "fused eigenvector transform needs the matrix combination and X");
      const int64_t ngrp = (a.ncol + XM_C - 1) / XM_C;
      MDA_REQUIRE(ngrp < (int64_t)1 << 31, "eigenvector grid too large");
      const unsigned nbx = (unsigned)((ngrp + 7) / 8 * 8);
      hipLaunchKernelGGL(k_trd_vectors_x<N>, dim3(nbx), dim3(256), 0, s, a.ncol, a.nloc, de, lam,
                         a.b, a.rho, a.rho_col, prior, a.null_rel, a.U, flag, a.X, a.x_ncol,
                         a.x_nlev, a.x_col0, a.status);
      MDA_LAUNCH_CHECK();
      if (a.ev_stage[2]) MDA_HIP(hipEventRecord(a.ev_stage[2], s));
      return;
    }
  }
  MDA_REQUIRE(!a.fuse, "fused eigenvector transform: no kernel for this k");
  MDA_REQUIRE(a.ncol < (int64_t)1 << 31, "eigenvector grid too large");
  hipLaunchKernelGGL(k_trd_vectors<N>, dim3((unsigned)a.ncol), dim3(64), 0, s, a.ncol, a.nloc, de, lam, a.b,
                     a.rho, a.rho_col, prior, a.combine, a.null_rel, a.U, a.T, flag);
  MDA_LAUNCH_CHECK();
  if (a.ev_stage[2]) MDA_HIP(hipEventRecord(a.ev_stage[2], s));
}

// k = 128: two wavefronts per column for E1 / E3, bisection for E2
template <int N>
static void launch_eig_w(hipStream_t s, const letkf_solve_args& a, double* de, double* lam,
                         int* flag) {
  MDA_REQUIRE(!a.fuse, "fused eigenvector transform: no kernel for this k");
  const int prior = a.inflation_form == MDA_INFLFORM_PRIOR;
  launch_zero(s, flag, sizeof(int) * a.ncol);
  MDA_REQUIRE(a.ncol < (int64_t)1 << 31, "eigensolver grid too large");
  const unsigned nb = (unsigned)a.ncol;  // one column per block
  hipLaunchKernelGGL((k_trd_wr<N, 2>), dim3(nb), dim3(2 * N), 0, s, a.ncol, a.C, a.c32, a.nloc,
                     a.rho, a.rho_col, prior, de, a.U);
  MDA_LAUNCH_CHECK();
  if (a.ev_stage[0]) MDA_HIP(hipEventRecord(a.ev_stage[0], s));
  hipLaunchKernelGGL(k_trd_eigvals_bisect<N>, dim3((unsigned)((a.ncol + 3) / 4)), dim3(256), 0,
                     s, a.ncol, a.nloc, de, lam, flag);
  MDA_LAUNCH_CHECK();
  if (a.ev_stage[1]) MDA_HIP(hipEventRecord(a.ev_stage[1], s));
  MDA_REQUIRE(a.ncol < (int64_t)1 << 31, "eigenvector grid too large");
  hipLaunchKernelGGL(k_trd_vectors_w4<N>, dim3((unsigned)a.ncol), dim3(2 * N), 0, s, a.ncol, a.nloc, de, lam, a.b,
                     a.rho, a.rho_col, prior, a.combine, a.null_rel, a.U, a.T, flag);
  MDA_LAUNCH_CHECK();
  if (a.ev_stage[2]) MDA_HIP(hipEventRecord(a.ev_stage[2], s));
}

bool letkf_eig_fusable(int k) { return k == 16 || k == 32 || k == 48 || k == 64; }

bool letkf_eig_supported(int k) {
  switch (k) {
    case 8: case 16: case 20: case 32: case 40: case 48: case 64: case 128: return true;
    default: return false;
  }
}

void launch_letkf_eig(hipStream_t s, const letkf_solve_args& a, double* de, double* lam,
                      int* flag) {
  switch (a.k) {
    case 8: launch_eig_n<8>(s, a, de, lam, flag); break;
    case 16: launch_eig_n<16>(s, a, de, lam, flag); break;
    case 20: launch_eig_n<20>(s, a, de, lam, flag); break;
    case 32: launch_eig_n<32>(s, a, de, lam, flag); break;
    case 40: launch_eig_n<40>(s, a, de, lam, flag); break;
    case 48: launch_eig_n<48>(s, a, de, lam, flag); break;
    case 64: launch_eig_n<64>(s, a, de, lam, flag); break;
    case 128: launch_eig_w<128>(s, a, de, lam, flag); break;
    default: throw mda_error(MDA_E_UNSUPPORTED, "tridiagonal eigensolver: no kernel for this k");
  }
}

}  // namespace mda
