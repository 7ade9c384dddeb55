// Local weighted EnKF (LWEnKF.hpp:207-334, 413-668): member weights, a
// weighted observation-space covariance tapered by the index-distance
// localisation, the gain K = X'_rho Y'^T S^-1 / (k-1) tapered again
// element-wise, and the perturbed-observation update of every member.
//
//   S(a, b)  = L_o(a, b) * sum_i (w_i y'_ai) y'_bi + [a == b] r_a
//              L_o(a, b) = f(|a - b| / p / radius)                (:554-575)
//   K(s, b)  = L_g(s, b) * (sum_i X'_rho(s, i) G(b, i)) / (k - 1),
//              G = S^-T Y' (p x k), L_g(s, b) = f(|s - b| / max(n, p) / radius)
//   x_i^a    = xbar + (X'_rho(:, i) + K D(:, i)),
//              D(b, i) = (yo_b + sqrt(r_b) Z_bi) - Yb_bi              (:286-301)
//
// f is the reference's localisation function (:600-636): gaussian
// exp(-nd^2 / 2), exponential exp(-nd), cutoff [nd <= 1], or its
// "gaspari_cohn" polynomial -- which is not Gaspari & Cohn's function (4 at
// 0, -0.125 just below 1, a jump to 0.125, 1.0 just below 2) -- or, on
// request, Gaspari & Cohn (1999) eq. (4.10).  The taper and the weighted
// covariance are evaluated without FMA contraction in the reference's
// operation order, so S matches the reference's to the last bit up to exp().
#include "mda_device.hpp"
#include "mda_internal.hpp"

namespace mda {

__device__ inline double lw_taper(double dist, double radius, int fn, int gc_form) {
#pragma clang fp contract(off)
  const double nd = dist / radius;
  switch (fn) {
    case MDA_LW_LOC_EXPONENTIAL:
      return exp(-nd);
    case MDA_LW_LOC_CUTOFF:
      return nd <= 1.0 ? 1.0 : 0.0;
    case MDA_LW_LOC_GASPARI_COHN:
      if (gc_form == MDA_LW_GC_STANDARD) return gaspari_cohn(nd);
      if (nd >= 2.0) return 0.0;
      if (nd >= 1.0) {
        const double z = nd - 1.0;
        return ((-0.25 * z + 0.5) * z + 0.625) * z + 0.125;
      } else {
        const double z = nd;
        return (((-0.25 * z + 0.5) * z + 0.625) * z - 5.0) * z + 4.0;
      }
    default:  // gaussian
      return exp(-0.5 * nd * nd);
  }
}

// sum over the state of (x_ji - mean_i)^2 per member j: partial[j][blk]
__global__ __launch_bounds__(256) void k_lw_member_sumsq(const double* __restrict__ X,
                                                         const double* __restrict__ mean,
                                                         int64_t n, double* __restrict__ partial) {
  __shared__ double red[256];
  const int j = blockIdx.y;
  const double* xj = X + (size_t)j * n;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = xj[i] - mean[i];
    acc += v * v;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) partial[(size_t)j * gridDim.x + blockIdx.x] = acc;
}

// chi2_j = sum_b (yo_b - Yb_bj) / r_b * (yo_b - Yb_bj), Yb = ybar + Y'  (:475-500)
__global__ __launch_bounds__(256) void k_lw_chi2(const double* __restrict__ yo,
                                                 const double* __restrict__ ybar,
                                                 const double* __restrict__ yp_om,
                                                 const double* __restrict__ var, int k,
                                                 int64_t p, double* __restrict__ chi2) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int64_t b = threadIdx.x; b < p; b += blockDim.x) {
    const double inn = yo[b] - (ybar[b] + yp_om[(size_t)b * k + j]);
    acc += inn * (1.0 / var[b]) * inn;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) chi2[j] = acc;
}

// S and its transpose, 16 x 16 tiles with the Y' rows of both sides in LDS
constexpr int LW_ST = 16;

__global__ __launch_bounds__(256) void k_lw_form_S(const double* __restrict__ yp_om,
                                                   const double* __restrict__ w,
                                                   const double* __restrict__ var, int k,
                                                   int64_t p, double radius, int fn, int gc_form,
                                                   double* __restrict__ S,
                                                   double* __restrict__ St) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  double* ya = sm;                  // [LW_ST][k]
  double* yb = sm + LW_ST * k;      // [LW_ST][k]
  double* ws = sm + 2 * LW_ST * k;  // [k]
  const int64_t a0 = (int64_t)blockIdx.y * LW_ST, b0 = (int64_t)blockIdx.x * LW_ST;
  const int tid = threadIdx.x;
  for (int e = tid; e < LW_ST * k; e += 256) {
    const int r = e / k, i = e % k;
    ya[e] = a0 + r < p ? yp_om[(size_t)(a0 + r) * k + i] : 0.0;
    yb[e] = b0 + r < p ? yp_om[(size_t)(b0 + r) * k + i] : 0.0;
  }
  for (int i = tid; i < k; i += 256) ws[i] = w[i];
  __syncthreads();
  const int ra = tid / LW_ST, rb = tid % LW_ST;
  const int64_t a = a0 + ra, b = b0 + rb;
  if (a >= p || b >= p) return;
  double acc = 0.0;
  for (int i = 0; i < k; ++i) acc += (ws[i] * ya[ra * k + i]) * yb[rb * k + i];
  const double dist = (double)(a > b ? a - b : b - a) / (double)p;
  double v = acc * lw_taper(dist, radius, fn, gc_form);
  if (a == b) v = v + var[a];
  S[(size_t)a * p + b] = v;
  St[(size_t)b * p + a] = v;
}

// D(b, i) = (yo_b + sqrt(r_b) Z_bi) - (ybar_b + y'_bi), obs-major [p][k]
__global__ void k_lw_innovations(const double* __restrict__ yo, const double* __restrict__ ybar,
                                 const double* __restrict__ yp_om,
                                 const double* __restrict__ var, const double* __restrict__ Z,
                                 int k, int64_t p, double* __restrict__ D) {
#pragma clang fp contract(off)
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < p * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / k;
    D[e] = (yo[b] + sqrt(var[b]) * Z[e]) - (ybar[b] + yp_om[e]);
  }
}

// ---------------------------------------------------------------------------
// The gain and update, fused: a workgroup owns LW_TN state points of every
// member.  Their inflated perturbations stay in LDS; the observations stream
// through in chunks of LW_TB: the K tile (LW_TN x LW_TB) is formed from the
// perturbations and G, tapered, folded into the gain extremes, and applied to
// the chunk of D at once, so K never reaches HBM.  In place on X (every
// workgroup reads its points before it writes them).
// ---------------------------------------------------------------------------
constexpr int LW_TN = 64, LW_TB = 32;

template <int KQ>  // members per wave quarter: k <= 4 KQ
__global__ __launch_bounds__(256) void k_lw_update(
    double* __restrict__ X, const double* __restrict__ mean, double scale,
    const double* __restrict__ Gt, const double* __restrict__ D, int k, int64_t n, int64_t p,
    double radius, int fn, int gc_form, double* __restrict__ mm) {
  extern __shared__ double sm[];
  double* xs = sm;                      // [k][LW_TN]   X'_rho
  double* gs = xs + (size_t)k * LW_TN;  // [k][LW_TB]   G chunk, member-major
  double* ds = gs + (size_t)k * LW_TB;  // [LW_TB][k]   D chunk
  double* ks = ds + (size_t)LW_TB * k;  // [LW_TB][LW_TN] K tile
  __shared__ double red[256];
  const int tid = threadIdx.x, a = tid & 63, q = tid >> 6;
  const int64_t s0 = (int64_t)blockIdx.x * LW_TN;
  const int64_t sa = s0 + a;
  const bool live = sa < n;
  const double md = live ? mean[sa] : 0.0;
  for (int i = q; i < k; i += 4)
    xs[i * LW_TN + a] = live ? (X[(size_t)i * n + sa] - md) * scale : 0.0;
  const double kden = (double)(k - 1);
  const double gdim = (double)(n > p ? n : p);
  double acc[KQ];
#pragma unroll
  for (int t = 0; t < KQ; ++t) acc[t] = 0.0;
  double kmax = -INFINITY, kmin = INFINITY;
  for (int64_t b0 = 0; b0 < p; b0 += LW_TB) {
    const int nb = (int)(p - b0 < LW_TB ? p - b0 : LW_TB);
    __syncthreads();  // the previous chunk is consumed
    for (int e = tid; e < LW_TB * k; e += 256) {
      const int bb = e / k, i = e % k;
      const bool in = bb < nb;
      gs[i * LW_TB + bb] = in ? Gt[(size_t)(b0 + bb) * k + i] : 0.0;
      ds[bb * k + i] = in ? D[(size_t)(b0 + bb) * k + i] : 0.0;
    }
    __syncthreads();
    // K tile: this thread's point a against observations q*8 .. q*8+7
    double kv[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) kv[t] = 0.0;
    for (int i = 0; i < k; ++i) {
      const double xv = xs[i * LW_TN + a];
      const double* g = gs + i * LW_TB + q * 8;
#pragma unroll
      for (int t = 0; t < 8; ++t) kv[t] = fma(xv, g[t], kv[t]);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int bb = q * 8 + t;
      double kk = 0.0;
      if (live && bb < nb) {
        const int64_t ob = b0 + bb;
        const double dist = (double)(sa > ob ? sa - ob : ob - sa) / gdim;
        kk = (kv[t] / kden) * lw_taper(dist, radius, fn, gc_form);
        kmax = fmax(kmax, kk);
        kmin = fmin(kmin, kk);
      }
      ks[bb * LW_TN + a] = kk;
    }
    __syncthreads();
    // apply: acc[t] += sum_b K(a, b) D(b, q*KQ + t)
    for (int bb = 0; bb < nb; ++bb) {
      const double kk = ks[bb * LW_TN + a];
      const double* dr = ds + bb * k + q * KQ;
#pragma unroll
      for (int t = 0; t < KQ; ++t)
        if (q * KQ + t < k) acc[t] = fma(kk, dr[t], acc[t]);
    }
  }
  if (live) {
#pragma unroll
    for (int t = 0; t < KQ; ++t) {
      const int i = q * KQ + t;
      if (i < k) X[(size_t)i * n + sa] = md + (xs[i * LW_TN + a] + acc[t]);
    }
  }
  // gain extremes of this workgroup
  red[tid] = kmax;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
    __syncthreads();
  }
  if (tid == 0) mm[2 * blockIdx.x] = red[0];
  __syncthreads();
  red[tid] = kmin;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmin(red[tid], red[tid + s]);
    __syncthreads();
  }
  if (tid == 0) mm[2 * blockIdx.x + 1] = red[0];
}

// ---------------------------------------------------------------------------
// The same gain and update for k <= 64 with register blocking: the two
// products are K_tile = X'_rho^T G (64 points x 32 observations) and
// acc += K_tile D (64 points x 64 members), one after the other per chunk of
// observations, with the taper in between.  Each thread holds a 4 x 2 tile of
// K and a 4 x 4 tile of the accumulator, so every LDS value it reads feeds
// two to four FMAs (the one-value-per-FMA LDS broadcast is what bounds the
// unblocked form).  Members are padded to 64 with zeros.
//   phase 1: thread (ty, tx) = (tid / 16, tid % 16): points ty + 16 r,
//            observations tx + 16 t
//   phase 2: points tx + 16 r, members ty + 16 t (stores coalesce along points)
// ---------------------------------------------------------------------------
constexpr int LB_TN = 64, LB_TB = 32, LB_KP = 64;

// K tile entry (observation bb, point a) of the tile buffer: XOR-swizzled
// within 16-point groups, so the phase-1 stores (16 observations at one
// point per half-wave) and the phase-2 loads (16 points at one observation)
// both fall in distinct banks.  80 KB of LDS in all: two workgroups per CU.
__device__ inline int lb_ps(int bb, int a) { return bb * LB_TN + (a ^ (bb & 15)); }

__global__ __launch_bounds__(256) void k_lw_update_blk(
    double* __restrict__ X, const double* __restrict__ mean, double scale,
    const double* __restrict__ Gt, const double* __restrict__ D, int k, int64_t n, int64_t p,
    double radius, int fn, int gc_form, double* __restrict__ mm) {
  extern __shared__ double sm[];
  double* Qs = sm;                          // [LB_KP][LB_TN]  X'_rho, member-major
  double* Gs = Qs + LB_KP * LB_TN;          // [LB_KP][LB_TB]  G chunk, member-major
  double* Ds = Gs + LB_KP * LB_TB;          // [LB_TB][LB_KP]  D chunk
  double* Ps = Ds + LB_TB * LB_KP;          // [LB_TB][LB_TN]  tapered K tile (lb_ps)
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const int64_t s0 = (int64_t)blockIdx.x * LB_TN;
  for (int e = tid; e < LB_KP * LB_TN; e += 256) {
    const int i = e / LB_TN, a = e % LB_TN;
    const int64_t sa = s0 + a;
    Qs[e] = (i < k && sa < n) ? (X[(size_t)i * n + sa] - mean[sa]) * scale : 0.0;
  }
  const double kden = (double)(k - 1);
  const double gdim = (double)(n > p ? n : p);
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[r][t] = 0.0;
  double kmax = -INFINITY, kmin = INFINITY;
  for (int64_t b0 = 0; b0 < p; b0 += LB_TB) {
    const int nb = (int)(p - b0 < LB_TB ? p - b0 : LB_TB);
    __syncthreads();  // the previous chunk is consumed (and Qs is complete)
    for (int e = tid; e < LB_TB * LB_KP; e += 256) {
      const int bb = e / LB_KP, i = e % LB_KP;
      const bool in = bb < nb && i < k;
      const size_t at = (size_t)(b0 + bb) * k + i;
      Gs[i * LB_TB + bb] = in ? Gt[at] : 0.0;
      Ds[bb * LB_KP + i] = in ? D[at] : 0.0;
    }
    __syncthreads();
    // phase 1: K for points ty + 16 r and observations tx + 16 t
    double kt[4][2];
#pragma unroll
    for (int r = 0; r < 4; ++r) kt[r][0] = kt[r][1] = 0.0;
    for (int i = 0; i < k; ++i) {
      double q[4], g[2];
#pragma unroll
      for (int r = 0; r < 4; ++r) q[r] = Qs[i * LB_TN + ty + 16 * r];
#pragma unroll
      for (int t = 0; t < 2; ++t) g[t] = Gs[i * LB_TB + tx + 16 * t];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int t = 0; t < 2; ++t) kt[r][t] = fma(q[r], g[t], kt[r][t]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int a = ty + 16 * r, bb = tx + 16 * t;
        const int64_t sa = s0 + a, ob = b0 + bb;
        double kk = 0.0;
        if (sa < n && bb < nb) {
          const double dist = (double)(sa > ob ? sa - ob : ob - sa) / gdim;
          kk = (kt[r][t] / kden) * lw_taper(dist, radius, fn, gc_form);
          kmax = fmax(kmax, kk);
          kmin = fmin(kmin, kk);
        }
        Ps[lb_ps(bb, a)] = kk;
      }
    __syncthreads();
    // phase 2: acc += K D for points tx + 16 r and members ty + 16 t
    for (int bb = 0; bb < nb; ++bb) {
      double pk[4], dv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) pk[r] = Ps[lb_ps(bb, tx + 16 * r)];
#pragma unroll
      for (int t = 0; t < 4; ++t) dv[t] = Ds[bb * LB_KP + ty + 16 * t];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[r][t] = fma(pk[r], dv[t], acc[r][t]);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int a = tx + 16 * r;
    const int64_t sa = s0 + a;
    if (sa >= n) continue;
    const double md = mean[sa];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int m = ty + 16 * t;
      if (m < k) X[(size_t)m * n + sa] = md + (Qs[m * LB_TN + a] + acc[r][t]);
    }
  }
  double* red = Ps;  // the tile buffer is free after the last chunk
  __syncthreads();
  red[tid] = kmax;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) red[tid] = fmax(red[tid], red[tid + st]);
    __syncthreads();
  }
  if (tid == 0) mm[2 * blockIdx.x] = red[0];
  __syncthreads();
  red[tid] = kmin;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) red[tid] = fmin(red[tid], red[tid + st]);
    __syncthreads();
  }
  if (tid == 0) mm[2 * blockIdx.x + 1] = red[0];
}

// any +inf variance (an invalid observation) -> *flag = 1
__global__ void k_lw_flag_inf(const double* __restrict__ var, int64_t p, int* __restrict__ flag) {
  bool bad = false;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < p;
       o += (int64_t)gridDim.x * blockDim.x)
    bad = bad || isinf(var[o]);
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------
void launch_lw_member_sumsq(hipStream_t s, const double* X, const double* mean, int k, int64_t n,
                            double* partial, int nblocks) {
  hipLaunchKernelGGL(k_lw_member_sumsq, dim3(nblocks, k), dim3(256), 0, s, X, mean, n, partial);
  MDA_LAUNCH_CHECK();
}

void launch_lw_chi2(hipStream_t s, const double* yo, const double* ybar, const double* yp_om,
                    const double* var, int k, int64_t p, double* chi2) {
  hipLaunchKernelGGL(k_lw_chi2, dim3(k), dim3(256), 0, s, yo, ybar, yp_om, var, k, p, chi2);
  MDA_LAUNCH_CHECK();
}

void launch_lw_form_S(hipStream_t s, const double* yp_om, const double* w, const double* var,
                      int k, int64_t p, double radius, int fn, int gc_form, double* S,
                      double* St) {
  const unsigned nt = (unsigned)((p + LW_ST - 1) / LW_ST);
  const size_t shm = sizeof(double) * (2 * LW_ST * k + k);
  hipLaunchKernelGGL(k_lw_form_S, dim3(nt, nt), dim3(256), shm, s, yp_om, w, var, k, p, radius,
                     fn, gc_form, S, St);
  MDA_LAUNCH_CHECK();
}

void launch_lw_innovations(hipStream_t s, const double* yo, const double* ybar,
                           const double* yp_om, const double* var, const double* Z, int k,
                           int64_t p, double* D) {
  hipLaunchKernelGGL(k_lw_innovations, dim3(grid_for(p * k, 256, 4096)), dim3(256), 0, s, yo,
                     ybar, yp_om, var, Z, k, p, D);
  MDA_LAUNCH_CHECK();
}

static_assert(LB_TN == LW_TN, "both update forms cover the same points per block");
size_t lw_update_blocks(int64_t n) { return (size_t)((n + LW_TN - 1) / LW_TN); }

void launch_lw_update(hipStream_t s, double* X, const double* mean, double scale,
                      const double* Gt, const double* D, int k, int64_t n, int64_t p,
                      double radius, int fn, int gc_form, double* mm) {
  MDA_REQUIRE(k >= 2 && k <= 128, "LWEnKF supports 2 <= k <= 128");
  const unsigned nb = (unsigned)lw_update_blocks(n);
  if (k <= LB_KP) {
    const size_t shb = sizeof(double) * ((size_t)LB_KP * LB_TN + (size_t)LB_KP * LB_TB +
                                         (size_t)LB_TB * LB_KP + (size_t)LB_TB * LB_TN);
    hipLaunchKernelGGL(k_lw_update_blk, dim3(nb), dim3(256), shb, s, X, mean, scale, Gt, D, k, n,
                       p, radius, fn, gc_form, mm);
    MDA_LAUNCH_CHECK();
    return;
  }
  const size_t shm = sizeof(double) * ((size_t)k * LW_TN + (size_t)k * LW_TB +
                                       (size_t)LW_TB * k + (size_t)LW_TB * LW_TN);
  if (k <= 32) {
    hipLaunchKernelGGL(k_lw_update<8>, dim3(nb), dim3(256), shm, s, X, mean, scale, Gt, D, k, n,
                       p, radius, fn, gc_form, mm);
  } else if (k <= 64) {
    hipLaunchKernelGGL(k_lw_update<16>, dim3(nb), dim3(256), shm, s, X, mean, scale, Gt, D, k, n,
                       p, radius, fn, gc_form, mm);
  } else {
    hipLaunchKernelGGL(k_lw_update<32>, dim3(nb), dim3(256), shm, s, X, mean, scale, Gt, D, k, n,
                       p, radius, fn, gc_form, mm);
  }
  MDA_LAUNCH_CHECK();
}

void launch_lw_flag_inf(hipStream_t s, const double* var, int64_t p, int* flag) {
  if (p <= 0) return;
  hipLaunchKernelGGL(k_lw_flag_inf, dim3(grid_for(p, 256, 1024)), dim3(256), 0, s, var, p, flag);
  MDA_LAUNCH_CHECK();
}

}  // namespace mda
