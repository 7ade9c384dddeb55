// LETKF kernels (LETKF.hpp:152-243 re-designed per column):
//   1. observation bucketing: key = bucket of the observation location, stable
//      LSD radix sort, bucket start offsets, gather of the observation data
//      into bucket order;
//   2. per-column Gram C_g = sum_o w_go/var_o y'_o y'_o^T and b_g on FP64 MFMA
//      (one wavefront per column, candidates from the buckets within reach);
//   3. per-column k x k solve: parallel Jacobi (symmetric square root) or
//      Cholesky (reference square root) -> transform T_g;
//   4. X_a = xbar + X' T_g for every level of the column.
#include "mda_crmath.hpp"
#include "mda_device.hpp"
#include "mda_internal.hpp"
#include "mda_xform.hpp"

namespace mda {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// 1. bucketing
// ---------------------------------------------------------------------------
__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline double norm_lon(double lon) {
  double l = fmod(lon, 360.0);
  return l < 0.0 ? l + 360.0 : l;
}

__device__ inline uint32_t bucket_of(const bucket_grid& g, double r, double c) {
  if (g.coord == MDA_COORD_GEOGRAPHIC) c = norm_lon(c);
  int row = clampi((int)floor((r - g.a0) / g.da), 0, g.nrow - 1);
  int col = clampi((int)floor((c - g.b0) / g.db), 0, g.ncolb - 1);
  return (uint32_t)row * (uint32_t)g.ncolb + (uint32_t)col;
}

// rcoord: latitude (GEOGRAPHIC) or y; ccoord: longitude or x
__global__ void k_bucket_keys(const double* __restrict__ rcoord,
                              const double* __restrict__ ccoord, int64_t p, bucket_grid g,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < p;
       o += (int64_t)gridDim.x * blockDim.x) {
    keys[o] = bucket_of(g, rcoord[o], ccoord[o]);
    vals[o] = (uint32_t)o;
  }
}

void launch_bucket_keys(hipStream_t s, const double* rcoord, const double* ccoord, int64_t p,
                        bucket_grid g, uint32_t* keys, uint32_t* vals) {
  hipLaunchKernelGGL(k_bucket_keys, dim3(grid_for(p, 256, 8192)), dim3(256), 0, s, rcoord,
                     ccoord, p, g, keys, vals);
  MDA_LAUNCH_CHECK();
}

// Stable LSD radix sort of (key, value) pairs, 8-bit digits.  A block owns
// RS_TILE consecutive keys (RS_KPT sub-tiles of RS_BLOCK).  Per pass: block
// digit histograms (digit-major), one exclusive scan, stable scatter with
// in-wave ranks from ballots and per-digit running counts across the
// block's sub-tiles.  Deterministic.
constexpr int RS_BLOCK = 256;
constexpr int RS_KPT = 16;
constexpr int RS_TILE = RS_BLOCK * RS_KPT;

__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint32_t* __restrict__ keys,
                                                      int64_t p, int shift,
                                                      uint32_t* __restrict__ hist,
                                                      int64_t nblocks) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = blockIdx.x * (int64_t)RS_TILE + threadIdx.x;
#pragma unroll
  for (int st = 0; st < RS_KPT; ++st) {
    const int64_t i = base + (int64_t)st * RS_BLOCK;
    if (i < p) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// exclusive scan of `total` uint32 in place, one workgroup of 1024 threads
// walking the array in coalesced tiles of 4096 (4 consecutive values per
// thread), carrying the running total between tiles
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* __restrict__ h, int64_t total) {
  __shared__ uint32_t part[1024];
  __shared__ uint32_t carry_s;
  const int tid = threadIdx.x;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (int64_t t0 = 0; t0 < total; t0 += 4096) {
    const int64_t i0 = t0 + 4 * (int64_t)tid;
    uint32_t v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = i0 + q < total ? h[i0 + q] : 0u;
    const uint32_t s = v[0] + v[1] + v[2] + v[3];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const uint32_t u = tid >= off ? part[tid - off] : 0u;
      __syncthreads();
      part[tid] += u;
      __syncthreads();
    }
    const uint32_t carry = carry_s;
    uint32_t run = carry + part[tid] - s;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (i0 + q < total) h[i0 + q] = run;
      run += v[q];
    }
    __syncthreads();
    if (tid == 1023) carry_s = carry + part[1023];
    __syncthreads();
  }
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int64_t p, int shift,
    const uint32_t* __restrict__ hist, int64_t nblocks) {
  __shared__ uint32_t wcnt[RS_BLOCK / 64][256];
  __shared__ uint32_t run[256];  // keys of each digit in the block's earlier sub-tiles
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int t = tid; t < (RS_BLOCK / 64) * 256; t += RS_BLOCK) (&wcnt[0][0])[t] = 0;
  run[tid] = hist[(size_t)tid * nblocks + blockIdx.x];
  __syncthreads();
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int st = 0; st < RS_KPT; ++st) {
    const int64_t i = blockIdx.x * (int64_t)RS_TILE + (int64_t)st * RS_BLOCK + tid;
    if (blockIdx.x * (int64_t)RS_TILE + (int64_t)st * RS_BLOCK >= p) break;  // block-uniform
    const bool valid = i < p;
    const uint32_t key = valid ? kin[i] : 0u, val = valid ? vin[i] : 0u;
    const uint32_t dig = (key >> shift) & 255u;
    uint64_t match = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const uint64_t m = __ballot(valid && ((dig >> b) & 1u));
      match &= ((dig >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(match & below);
    if (valid && rank == 0) wcnt[wv][dig] = (uint32_t)__popcll(match);
    __syncthreads();
    if (valid) {
      uint32_t off = run[dig];
      for (int w = 0; w < wv; ++w) off += wcnt[w][dig];
      const uint32_t dst = off + rank;
      kout[dst] = key;
      vout[dst] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < RS_BLOCK / 64; ++w) {
      add += wcnt[w][tid];
      wcnt[w][tid] = 0;
    }
    run[tid] += add;
    __syncthreads();
  }
}

void radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp,
                      uint32_t* vals_tmp, int64_t p, int key_bits, uint32_t* hist,
                      int64_t hist_cap) {
  if (p <= 0) return;
  int64_t nblocks = (p + RS_TILE - 1) / RS_TILE;
  MDA_REQUIRE(nblocks * 256 <= hist_cap, "radix histogram workspace too small");
  uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
  int passes = 0;
  for (int shift = 0; shift < key_bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nblocks), dim3(RS_BLOCK), 0, s, ki, p, shift,
                       hist, nblocks);
    MDA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(1024), 0, s, hist, nblocks * 256);
    MDA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nblocks), dim3(RS_BLOCK), 0, s, ki, vi, ko,
                       vo, p, shift, hist, nblocks);
    MDA_LAUNCH_CHECK();
    uint32_t* t = ki; ki = ko; ko = t;
    t = vi; vi = vo; vo = t;
    ++passes;
  }
  if (passes & 1) {  // result sits in the tmp buffers: copy back
    MDA_HIP(hipMemcpyAsync(keys, ki, p * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    MDA_HIP(hipMemcpyAsync(vals, vi, p * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

// start[b] = first sorted position whose key >= b, for b in [0, nbuckets]
__global__ void k_bucket_ranges(const uint32_t* __restrict__ sk, int64_t p, int nbuckets,
                                int* __restrict__ start) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= p;
       t += (int64_t)gridDim.x * blockDim.x) {
    int lo = (t == 0) ? 0 : (int)sk[t - 1] + 1;
    int hi = (t == p) ? nbuckets : (int)sk[t];
    for (int b = lo; b <= hi; ++b) start[b] = (int)t;
  }
}

void launch_bucket_ranges(hipStream_t s, const uint32_t* sorted_keys, int64_t p, int nbuckets,
                          int* start) {
  hipLaunchKernelGGL(k_bucket_ranges, dim3(grid_for(p + 1, 256, 8192)), dim3(256), 0, s,
                     sorted_keys, p, nbuckets, start);
  MDA_LAUNCH_CHECK();
}

// Per-observation data in bucket order: 1/var (or 1), the coordinates, and
// for the FP32 / BF16x3 contractions of GEOGRAPHIC problems the unit vector
// of each location (ux, uy, uz may be null).
__global__ void k_gather_obs(const uint32_t* __restrict__ order, int64_t p,
                             const double* __restrict__ var, const double* __restrict__ oa,
                             const double* __restrict__ ob, const double* __restrict__ oc,
                             int rweight, double* __restrict__ s_rinv, double* __restrict__ s_a,
                             double* __restrict__ s_b, double* __restrict__ s_c,
                             double* __restrict__ ux, double* __restrict__ uy,
                             double* __restrict__ uz) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < p;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t o = order[t];
    const double a = oa[o], b = ob[o];
    s_rinv[t] = rweight == MDA_RW_IDENTITY ? 1.0 : 1.0 / var[o];
    s_a[t] = a;
    s_b[t] = b;
    s_c[t] = oc ? oc[o] : 0.0;
    if (ux) {
      const double la = a * kPi / 180.0, lo = b * kPi / 180.0;
      ux[t] = cos(la) * cos(lo); uy[t] = cos(la) * sin(lo); uz[t] = sin(la);
    }
  }
}

void launch_gather_obs(hipStream_t s, const uint32_t* order, int64_t p, const double* var,
                       const double* oa, const double* ob, const double* oc, int rweight,
                       double* s_rinv, double* s_a, double* s_b, double* s_c, double* ux,
                       double* uy, double* uz) {
  if (p <= 0) return;
  hipLaunchKernelGGL(k_gather_obs, dim3(grid_for(p, 256, 65536)), dim3(256), 0, s, order, p,
                     var, oa, ob, oc, rweight, s_rinv, s_a, s_b, s_c, ux, uy, uz);
  MDA_LAUNCH_CHECK();
}

// Y' rows and d moved into bucket order (the path where the caller supplies
// H(x) member-major and the observation statistics run in input order), with
// the FP32 copy of Y' for the reduced-precision contractions (y32 may be null).
__global__ void k_gather_rows(const uint32_t* __restrict__ order, int64_t p, int k,
                              const double* __restrict__ yp_om, const double* __restrict__ d,
                              double* __restrict__ s_yp, double* __restrict__ s_d,
                              float* __restrict__ y32) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < p * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e / k;
    const int j = (int)(e % k);
    const uint32_t o = order[t];
    const double v = yp_om[(size_t)o * k + j];
    s_yp[e] = v;
    if (y32) y32[e] = (float)v;
    if (j == 0) s_d[t] = d[o];
  }
}

void launch_gather_rows(hipStream_t s, const uint32_t* order, int64_t p, int k,
                        const double* yp_om, const double* d, double* s_yp, double* s_d,
                        float* y32) {
  if (p <= 0) return;
  hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(p * k, 256, 65536)), dim3(256), 0, s, order,
                     p, k, yp_om, d, s_yp, s_d, y32);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// 2. per-column Gram on FP64 MFMA.  One wavefront per column.  For each
// bucket range in reach (from the column's setup record), 64 candidates at a
// time: distance (Location.hpp metric; GEOGRAPHIC from the chord of unit
// vectors, equal to the haversine distance), localisation weight (GC taper
// or cutoff) times 1/var, ballot-compaction of the nonzero weights into LDS,
// then K-steps of 4 observations through v_mfma_f64_16x16x4f64 on the lower
// 16x16 tiles of C.
//   f64 16x16x4 lane layout: A[l%16][l/16], B[l/16][l%16],
//   D[l/16 + 4r][l%16] (r = 0..3).
// ---------------------------------------------------------------------------
struct seg_list {
  int n;
  int j0[2], j1[2];
};

__device__ inline seg_list col_segments(const bucket_grid& g, double c_lo, double c_hi,
                                        bool full) {
  seg_list s;
  if (full || (g.wrap && c_hi - c_lo >= 360.0 - g.db)) {
    s.n = 1; s.j0[0] = 0; s.j1[0] = g.ncolb - 1;
    return s;
  }
  const int last = g.ncolb - 1;
  if (!g.wrap) {
    s.n = 1;
    s.j0[0] = clampi((int)floor((c_lo - g.b0) / g.db), 0, last);
    s.j1[0] = clampi((int)floor((c_hi - g.b0) / g.db), 0, last);
    return s;
  }
  // longitude buckets tile [0, 360) exactly (db = 360 / ncolb): the part of
  // the range beyond either end is bucketed from its wrapped longitude
  if (c_lo < 0.0) {
    s.n = 2;
    s.j0[0] = clampi((int)floor((c_lo + 360.0) / g.db), 0, last); s.j1[0] = last;
    s.j0[1] = 0; s.j1[1] = clampi((int)floor(c_hi / g.db), 0, last);
  } else if (c_hi >= 360.0) {
    s.n = 2;
    s.j0[0] = clampi((int)floor(c_lo / g.db), 0, last); s.j1[0] = last;
    s.j0[1] = 0; s.j1[1] = clampi((int)floor((c_hi - 360.0) / g.db), 0, last);
  } else {
    s.n = 1;
    s.j0[0] = clampi((int)floor(c_lo / g.db), 0, last);
    s.j1[0] = clampi((int)floor(c_hi / g.db), 0, last);
  }
  return s;
}

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Candidate window of column g: its location and the bucket rows / column
// segments that can hold observations within the radius.
struct col_window {
  double ca, cb, cc;
  int row0, row1;
  seg_list sl;  // bucket columns of the whole window (widest row)
  bool full;    // every longitude (GEOGRAPHIC window reaching a pole)
};

__device__ inline col_window column_window(const letkf_gram_args& a, int64_t g) {
  const bucket_grid& G = a.grid;
  col_window w;
  w.cc = 0.0;
  if (a.col_a) { w.ca = a.col_a[g]; w.cb = a.col_b[g]; }
  else { w.ca = (double)((g + a.col0) % a.nx); w.cb = (double)((g + a.col0) / a.nx); }
  if (a.col_c) w.cc = a.col_c[g];
  double r_lo, r_hi, c_lo, c_hi;
  bool full = false;
  if (a.coord == MDA_COORD_GEOGRAPHIC) {
    const double dd = a.radius / kEarthRadiusKm;
    const double ddeg = dd * 180.0 / kPi + 1e-9;
    r_lo = w.ca - ddeg; r_hi = w.ca + ddeg;
    const double cos_lat = cos(w.ca * kPi / 180.0);
    full = r_lo <= -90.0 || r_hi >= 90.0 || sin(dd) >= cos_lat;
    const double dl = full ? 180.0 : asin(fmin(1.0, sin(dd) / cos_lat)) * 180.0 / kPi + 1e-9;
    const double lon = norm_lon(w.cb);
    c_lo = lon - dl; c_hi = lon + dl;
  } else {  // GRID / CARTESIAN: rows along y, bucket columns along x
    r_lo = w.cb - a.radius - 1e-9; r_hi = w.cb + a.radius + 1e-9;
    c_lo = w.ca - a.radius - 1e-9; c_hi = w.ca + a.radius + 1e-9;
  }
  w.row0 = clampi((int)floor((r_lo - G.a0) / G.da), 0, G.nrow - 1);
  w.row1 = clampi((int)floor((r_hi - G.a0) / G.da), 0, G.nrow - 1);
  w.sl = col_segments(G, c_lo, c_hi, full);
  w.full = full;
  return w;
}

// Bucket columns of one bucket row of the window, trimmed to the circle's
// extent inside that row's band (the window's own segments are those of its
// widest row).  GEOGRAPHIC: the cap of angular radius d around (lat0, lon0)
// spans longitudes lon0 +- dl(phi) at latitude phi, with
//   cos d = sin lat0 sin phi + cos lat0 cos phi cos dl,
// largest at sin phi_m = sin lat0 / cos d; over a band the largest extent is
// at phi_m clamped to the band (and a band that the cap does not reach has
// none).  GRID / CARTESIAN: the half-width sqrt(r^2 - dy^2) at the band's
// nearest y.  Margins keep every observation that the exact distance test
// accepts: the trimmed range only drops observations outside the circle, so
// the in-circle observations keep their order.
__device__ inline seg_list row_segments(const letkf_gram_args& a, const col_window& w, int row) {
  const bucket_grid& G = a.grid;
  seg_list none;
  none.n = 0;
  if (w.full) return w.sl;
  const double band_lo = G.a0 + row * G.da, band_hi = band_lo + G.da;
  if (a.coord == MDA_COORD_GEOGRAPHIC) {
    const double dd = a.radius / kEarthRadiusKm;
    const double lat0 = w.ca * kPi / 180.0;
    const double sm = sin(lat0) / cos(dd);
    const double phm = sm >= 1.0 ? 90.0 : sm <= -1.0 ? -90.0 : asin(sm) * 180.0 / kPi;
    const double ph = fmin(fmax(phm, band_lo), band_hi) * kPi / 180.0;
    const double cph = cos(ph), cl0 = cos(lat0);
    if (cph <= 1e-9 || cl0 <= 1e-9) return w.sl;
    const double cdl = (cos(dd) - sin(lat0) * sin(ph)) / (cl0 * cph) - 1e-12;
    if (cdl > 1.0) return none;
    if (cdl <= -1.0) return w.sl;
    const double dl = acos(cdl) * 180.0 / kPi + 1e-9;
    const double lon = norm_lon(w.cb);
    return col_segments(G, lon - dl, lon + dl, false);
  }
  const double dy = w.cb < band_lo ? band_lo - w.cb : (w.cb > band_hi ? w.cb - band_hi : 0.0);
  const double r = a.radius;
  if (dy > r * (1.0 + 1e-12) + 1e-9) return none;
  const double dx = sqrt(fmax(0.0, r * r - dy * dy)) * (1.0 + 1e-12) + 1e-9;
  return col_segments(G, w.ca - dx, w.ca + dx, false);
}

// Per-row candidate ranges of a window, computed before the candidate loop
// into a per-wave LDS table, one lane per row (the trigonometry's
// temporaries are dead during the loop, and the bucket-start loads of all
// rows are in flight together instead of one row at a time).  Entry i < RT_ROWS
// (row row0 + i): n, then [s0, s1) observation ranges of up to two
// segments.  Entry RT_ROWS holds the window's own bucket columns (n, j0[0],
// j1[0], j0[1], j1[1]) for rows past RT_ROWS, whose ranges are read in the
// loop.
constexpr int RT_ROWS = 16;
struct row_ranges {
  int n;
  int s0[2], s1[2];
};

__device__ inline void fill_row_table(const letkf_gram_args& a, const col_window& cw,
                                      int (*tab)[5]) {
  const int lane = threadIdx.x & 63;
  const int nr = cw.row1 - cw.row0 + 1;
  const bucket_grid& G = a.grid;
  if (lane < RT_ROWS && lane < nr) {
    const int row = cw.row0 + lane;
    const seg_list rs = row_segments(a, cw, row);
    tab[lane][0] = rs.n;
    tab[lane][1] = rs.n > 0 ? a.bstart[row * G.ncolb + rs.j0[0]] : 0;
    tab[lane][2] = rs.n > 0 ? a.bstart[row * G.ncolb + rs.j1[0] + 1] : 0;
    tab[lane][3] = rs.n > 1 ? a.bstart[row * G.ncolb + rs.j0[1]] : 0;
    tab[lane][4] = rs.n > 1 ? a.bstart[row * G.ncolb + rs.j1[1] + 1] : 0;
  }
  if (lane == RT_ROWS) {
    tab[lane][0] = cw.sl.n;
    tab[lane][1] = cw.sl.j0[0];
    tab[lane][2] = cw.sl.j1[0];
    tab[lane][3] = cw.sl.n > 1 ? cw.sl.j0[1] : 0;
    tab[lane][4] = cw.sl.n > 1 ? cw.sl.j1[1] : 0;
  }
  wave_sync_lds();
}

__device__ inline row_ranges row_from_table(const letkf_gram_args& a, int row0, int row,
                                            const int (*tab)[5]) {
  const int i = row - row0;
  row_ranges rr;
  if (i < RT_ROWS) {
    rr.n = tab[i][0];
    rr.s0[0] = tab[i][1];
    rr.s1[0] = tab[i][2];
    rr.s0[1] = tab[i][3];
    rr.s1[1] = tab[i][4];
  } else {
    const int* e = tab[RT_ROWS];
    const int base = row * a.grid.ncolb;
    rr.n = e[0];
    rr.s0[0] = a.bstart[base + e[1]];
    rr.s1[0] = a.bstart[base + e[2] + 1];
    rr.s0[1] = rr.n > 1 ? a.bstart[base + e[3]] : 0;
    rr.s1[1] = rr.n > 1 ? a.bstart[base + e[4] + 1] : 0;
  }
  return rr;
}

// A GEOGRAPHIC candidate inside the shell around the membership threshold
// (shell_width): (column, observation, the chord test's decision) goes to
// the pass's boundary list, and k_letkf_shell_fixup decides it exactly after
// the Gram.  Rare: about one pair per analysis at 1e6 observations.
__device__ inline void shell_push(const letkf_gram_args& a, int64_t g, int64_t t, bool in) {
  const unsigned slot = atomicAdd(a.shell_count, 1u);
  if (slot < (unsigned)a.shell_cap)
    a.shell[slot] = make_int4((int)g, (int)(t & 0xffffffff), (int)(t >> 32), in ? 1 : 0);
}

// Column iteration of the Gram kernels: a grid-stride walk, or (work-queue
// mode) lane 0 of each wave takes the next position of a cost-sorted order
// with a global atomic counter; every wave exits once the counter passes ncol.
__device__ inline int64_t queue_next(uint32_t* q) {
  uint32_t v = 0;
  if ((threadIdx.x & 63) == 0) v = atomicAdd(q, 1u);
  return (int64_t)__shfl(v, 0);
}

// Cost of a column = observations in the buckets of its candidate window
// (an upper bound of its local count, proportional to its Gram work); the
// key sorts columns by descending cost.
__global__ void k_column_cost(letkf_gram_args a, uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ cols) {
  const bucket_grid& G = a.grid;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < a.ncol;
       g += (int64_t)gridDim.x * blockDim.x) {
    const col_window cw = column_window(a, g);
    int64_t cost = 0;
    for (int row = cw.row0; row <= cw.row1; ++row) {
      const seg_list rs = row_segments(a, cw, row);
      const int sj00 = rs.j0[0], sj01 = rs.j1[0];
      const int sj10 = rs.n > 1 ? rs.j0[1] : 0, sj11 = rs.n > 1 ? rs.j1[1] : 0;
      for (int sg = 0; sg < rs.n; ++sg)
        cost += a.bstart[row * G.ncolb + (sg ? sj11 : sj01) + 1] -
                a.bstart[row * G.ncolb + (sg ? sj10 : sj00)];
    }
    keys[g] = 0xffffffffu - (uint32_t)(cost < 0xffffffffLL ? cost : 0xffffffffLL);
    cols[g] = (uint32_t)g;
  }
}

void launch_column_cost(hipStream_t s, const letkf_gram_args& a, uint32_t* keys,
                        uint32_t* cols) {
  hipLaunchKernelGGL(k_column_cost, dim3(grid_for(a.ncol, 256, 8192)), dim3(256), 0, s, a, keys,
                     cols);
  MDA_LAUNCH_CHECK();
}

// NPART > 1 (k > 64): the lower tiles are dealt round-robin to NPART waves
// per column (tile t to wave t % NPART), each repeating the candidate search,
// so the accumulators of one wave stay within its registers.
template <int KT, int NPART>
__device__ __forceinline__ void gram_f64_body(const letkf_gram_args& a) {
  constexpr int NT = KT * (KT + 1) / 2;
  constexpr int NA = (NT + NPART - 1) / NPART;
  __shared__ int q_off[4][64];
  __shared__ double q_w[4][64];
  // the column's setup record (k_column_setup)
  __shared__ __attribute__((aligned(16))) int crec[4][LETKF_CREC_INTS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i16 = lane & 15, kk = lane >> 4;
  const int k = a.k;
  const double cgc = a.radius / 2.0;
  // chord threshold for GEOGRAPHIC: |u - v|^2 <= 4 sin^2(r / 2R), the
  // haversine test d <= r on unit vectors
  const double half_ang = a.radius / (2.0 * kEarthRadiusKm);
  const double chord2_r = half_ang >= kPi / 2.0 ? 5.0 : 4.0 * sin(half_ang) * sin(half_ang);
  const double shellw = shell_width(chord2_r);
  const int64_t nitems = a.ncol * NPART;
  int64_t it = a.queue ? queue_next(a.queue) : (int64_t)blockIdx.x * 4 + wv;
  for (; it < nitems; it = a.queue ? queue_next(a.queue) : it + (int64_t)gridDim.x * 4) {
    const int part = (int)(it % NPART);
    const int64_t g = a.order ? (int64_t)a.order[it / NPART] : it / NPART;
    d4 acc[NA];
#pragma unroll
    for (int t = 0; t < NA; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    double bp[KT];
#pragma unroll
    for (int t = 0; t < KT; ++t) bp[t] = 0.0;
    int nl = 0;
    {
      const int2* src = (const int2*)(a.crec + (size_t)g * LETKF_CREC_INTS);
      if (lane < LETKF_CREC_INTS / 2) ((int2*)crec[wv])[lane] = src[lane];
      wave_sync_lds();
    }
    const int (*rtab)[5] = (const int (*)[5])(crec[wv] + 8);
    const int row0 = crec[wv][6], row1 = crec[wv][7];
    const double rx = ((const double*)crec[wv])[0];
    const double ry = ((const double*)crec[wv])[1];
    const double rz = ((const double*)crec[wv])[2];
    for (int row = row0; row <= row1; ++row) {
      const row_ranges rr = row_from_table(a, row0, row, rtab);
      for (int sg = 0; sg < rr.n; ++sg) {
        const int64_t s0 = sg ? rr.s0[1] : rr.s0[0];
        const int64_t s1 = sg ? rr.s1[1] : rr.s1[0];
        for (int64_t base = s0; base < s1; base += 64) {
          int64_t t = base + lane;
          double w = 0.0;
          bool in = false;
          if (t < s1) {
            double dist;
            if (a.coord == MDA_COORD_GEOGRAPHIC) {
              // great-circle distance from the chord of unit vectors,
              // d = 2R asin(|u - v| / 2), equal to the haversine distance
              const double dx = rx - a.ux[t], dy = ry - a.uy[t], dz = rz - a.uz[t];
              const double c2 = dx * dx + dy * dy + dz * dz;
              in = c2 <= chord2_r;
              if (fabs(c2 - chord2_r) <= shellw && part == 0) shell_push(a, g, t, in);
              dist = 2.0 * kEarthRadiusKm * asin(fmin(1.0, 0.5 * sqrt(c2)));
            } else {
              const double dx = rx - a.oa[t], dy = ry - a.ob[t];
              const double dz = a.coord == MDA_COORD_CARTESIAN ? rz - a.oc[t] : 0.0;
              dist = sqrt(dist2_exact(dx, dy, dz));
              in = dist <= a.radius;
            }
            if (in)
              w = (a.localization == MDA_LOC_GASPARI_COHN ? gaspari_cohn(dist / cgc) : 1.0) *
                  a.rinv[t];
          }
          nl += __popcll(__ballot(in));
          uint64_t wm = __ballot(w != 0.0);
          int cnt = __popcll(wm);
          if (cnt == 0) continue;
          uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
          int rank = __popcll(wm & below);
          if (w != 0.0) { q_off[wv][rank] = lane; q_w[wv][rank] = w; }
          wave_sync_lds();
          for (int s = 0; s < cnt; s += 4) {
            int tq = s + kk;
            double wt = 0.0;
            int64_t oi = -1;
            if (tq < cnt) { wt = q_w[wv][tq]; oi = base + q_off[wv][tq]; }
            double yv[KT];
#pragma unroll
            for (int ta = 0; ta < KT; ++ta) {
              int e = 16 * ta + i16;
              yv[ta] = (oi >= 0 && e < k) ? a.yp[(size_t)oi * k + e] : 0.0;
            }
            double dv = oi >= 0 ? a.d[oi] : 0.0;
            int tile = 0;
#pragma unroll
            for (int ta = 0; ta < KT; ++ta) {
              double aw = wt * yv[ta];
              bp[ta] += aw * dv;
#pragma unroll
              for (int tb = 0; tb <= ta; ++tb) {
                if (tile % NPART == part)
                  acc[tile / NPART] =
                      __builtin_amdgcn_mfma_f64_16x16x4f64(aw, yv[tb], acc[tile / NPART], 0, 0, 0);
                ++tile;
              }
            }
          }
          wave_sync_lds();
        }
      }
    }
    // b: sum the four observation slots (lanes i16, i16+16, i16+32, i16+48)
#pragma unroll
    for (int ta = 0; ta < KT; ++ta) {
      bp[ta] += __shfl_xor(bp[ta], 16);
      bp[ta] += __shfl_xor(bp[ta], 32);
    }
    double* Cg = a.C + (size_t)g * k * k;
    // k and the lane coordinates through opaque copies: the element offsets
    // below are recomputed per column, not kept live across the search
    int ko = k, kk_e = kk, i16_e = i16;
    __asm__ volatile("" : "+s"(ko), "+v"(kk_e), "+v"(i16_e));
    int tile = 0;
#pragma unroll
    for (int ta = 0; ta < KT; ++ta) {
#pragma unroll
      for (int tb = 0; tb <= ta; ++tb) {
        if (tile % NPART == part) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            int row = 16 * ta + kk_e + 4 * r, col = 16 * tb + i16_e;
            if (row < ko && col < ko) {
              Cg[row * ko + col] = acc[tile / NPART][r];
              if (ta != tb) Cg[col * ko + row] = acc[tile / NPART][r];
            }
          }
        }
        ++tile;
      }
    }
    if (part != 0) continue;  // b and the local count: first wave of the column
    if (kk == 0) {
#pragma unroll
      for (int ta = 0; ta < KT; ++ta) {
        int e = 16 * ta + i16;
        if (e < k) a.b[(size_t)g * k + e] = bp[ta];
      }
    }
    if (lane == 0) a.nloc[g] = nl;
  }
}

// k <= 64: three waves per SIMD; 64 < k <= 128 (two waves per column, 18
// tiles each): one.
template <int KT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_letkf_gram_f64(letkf_gram_args a) {
  gram_f64_body<KT, 1>(a);
}

__global__ __launch_bounds__(256) void k_letkf_gram_f64_w(letkf_gram_args a) {
  gram_f64_body<8, 2>(a);
}

// ---------------------------------------------------------------------------
// 2'. per-column Gram with FP32 / BF16x3 MFMA contractions (FP32 accumulate).
// Same candidate search and weights as the FP64 kernel (FP64 geometry; the
// GEOGRAPHIC distance from the chord of unit vectors, d = 2R asin(|u-v|/2),
// equal to the haversine distance).  C = sum_o (sqrt(w_o) y'_o)(sqrt(w_o) y'_o)^T
// so one scaled copy serves as both MFMA operands.  Compacted observations
// are queued in LDS and consumed 32 at a time: their FP32 rows are copied
// into the wave's LDS row buffer by global_load_lds, and each lane builds its
// operand fragments straight from that buffer (scaled by sqrt(w_o));
//   BF16x3: a = hi + lo (bf16, round to nearest even),
//           C += hi hi^T + hi lo^T + lo hi^T on v_mfma_f32_16x16x32_bf16
//           (A[l%16][8(l/16)+e], B[8(l/16)+e][l%16], D[4(l/16)+r][l%16]);
//   FP32:   C += a a^T on v_mfma_f32_16x16x4f32 (A[l%16][l/16]).
// b = sum_o w_o d_o y'_o: FP32 partial sums per lane and batch, accumulated in
// FP64.  The column's window comes from its setup record (k_column_setup).
// Blocks are mapped XCD-contiguously: the dispatcher deals blocks to the 8
// XCDs round-robin, so block i runs columns from slice (i % 8) and columns
// that share observations share an L2.
// ---------------------------------------------------------------------------
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float fl2 __attribute__((ext_vector_type(2)));

constexpr int GM_Q = 96;   // observation queue per wave (31 pending + 64 new)

__device__ inline double readlane_f64(double v, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Per-column setup of the FP32 / BF16x3 Gram, done in a
// kernel of its own so that the window's trigonometry (and its constants)
// stays out of the Gram's registers.  Record of column g, CT_INTS ints:
//   [0, 6)   reference point, 3 doubles: the unit vector of the column's
//            location (GEOGRAPHIC) or its coordinates (GRID / CARTESIAN)
//   [6], [7] first and last bucket row of the window
//   [8, 93)  the row table (RT_ROWS + 1 entries of 5 ints, as fill_row_table)
constexpr int CT_INTS = LETKF_CREC_INTS;

// One thread per record entry (RT_ROWS row entries, the window entry, the
// reference point and rows): each thread works out its column's window and
// then only its own entry.
constexpr int CT_EPC = RT_ROWS + 2;

__global__ __launch_bounds__(256) void k_column_setup(letkf_gram_args a, int* __restrict__ rec) {
  const bucket_grid& G = a.grid;
  const int64_t total = a.ncol * CT_EPC;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * 256) {
    const int64_t g = t / CT_EPC;
    const int e = (int)(t % CT_EPC);
    const col_window cw = column_window(a, g);
    int* r = rec + (size_t)g * CT_INTS;
    int* tab = r + 8;
    if (e < RT_ROWS) {
      // entry e: bucket row row0 + e (entries past the window stay unwritten,
      // the Gram reads entries of the window's rows only)
      if (e < cw.row1 - cw.row0 + 1) {
        const int row = cw.row0 + e;
        const seg_list rs = row_segments(a, cw, row);
        tab[5 * e + 0] = rs.n;
        tab[5 * e + 1] = rs.n > 0 ? a.bstart[row * G.ncolb + rs.j0[0]] : 0;
        tab[5 * e + 2] = rs.n > 0 ? a.bstart[row * G.ncolb + rs.j1[0] + 1] : 0;
        tab[5 * e + 3] = rs.n > 1 ? a.bstart[row * G.ncolb + rs.j0[1]] : 0;
        tab[5 * e + 4] = rs.n > 1 ? a.bstart[row * G.ncolb + rs.j1[1] + 1] : 0;
      }
    } else if (e == RT_ROWS) {
      tab[5 * RT_ROWS + 0] = cw.sl.n;
      tab[5 * RT_ROWS + 1] = cw.sl.j0[0];
      tab[5 * RT_ROWS + 2] = cw.sl.j1[0];
      tab[5 * RT_ROWS + 3] = cw.sl.n > 1 ? cw.sl.j0[1] : 0;
      tab[5 * RT_ROWS + 4] = cw.sl.n > 1 ? cw.sl.j1[1] : 0;
    } else {
      double* ref = (double*)r;
      if (a.coord == MDA_COORD_GEOGRAPHIC) {
        const double la = cw.ca * kPi / 180.0, lo = cw.cb * kPi / 180.0;
        ref[0] = cos(la) * cos(lo); ref[1] = cos(la) * sin(lo); ref[2] = sin(la);
      } else {
        ref[0] = cw.ca; ref[1] = cw.cb; ref[2] = cw.cc;
      }
      r[6] = cw.row0;
      r[7] = cw.row1;
    }
  }
}

// Row buffer of one batch: row j (observation j of the batch, KM members in
// FP32) starts at word gm_row_base(j).  The operand reads of a step take
// lanes (i16, kq) to rows 8kq + e (BF16x3) or 4s + kq (FP32) at member
// 16t + i16; the row shifts put those 64 reads in 64 different banks.
template <int KM, int PREC>
__host__ __device__ constexpr int gm_row_base(int j) {
  return PREC == MDA_PREC_BF16X3 ? j * KM + 16 * (j >> 3) : j * (KM + 16);
}
template <int KM, int PREC>
__host__ __device__ constexpr int gm_row_words() {
  return gm_row_base<KM, PREC>(31) + KM;
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// KT > 4 (64 < k <= 128): rows carry 128 members (two global_load_lds per
// row), the 36 lower tiles stay in one wave's registers (one wave per SIMD).
template <int KT, int PREC>
__device__ __forceinline__ void gram_mixed_body(const letkf_gram_args& a) {
  constexpr int NT = KT * (KT + 1) / 2;
  constexpr int MB = KT > 4 ? 2 : 1;  // 64-member blocks per row
  constexpr int KM = 64 * MB;         // member capacity
  constexpr int RW = gm_row_words<KM, PREC>();
  __shared__ int q_off[4][GM_Q];
  __shared__ float q_wd[4][GM_Q];   // w_o d_o (for b)
  __shared__ float q_sw[4][GM_Q];   // sqrt(w_o) (row scale)
  // the column's setup record (k_column_setup)
  __shared__ __attribute__((aligned(16))) int crec[4][CT_INTS];
  // scale and weight of the batch in flight, entry j at [j]
  __shared__ __attribute__((aligned(16))) float bt_sw[4][32];
  __shared__ __attribute__((aligned(16))) float bt_wd[4][32];
  // rows of the batch in flight, filled by global_load_lds (no registers
  // held while the rows are in flight)
  __shared__ __attribute__((aligned(16))) float rows_raw[4][RW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int k = a.k;
  const bucket_grid& G = a.grid;
  const float cgcf = (float)(a.radius / 2.0);
  // chord threshold for GEOGRAPHIC: |u - v|^2 <= 4 sin^2(r / 2R)
  const double half_ang = a.radius / (2.0 * kEarthRadiusKm);
  const double chord2_r = half_ang >= kPi / 2.0 ? 5.0 : 4.0 * sin(half_ang) * sin(half_ang);
  const double shellw = shell_width(chord2_r);
  // XCD-contiguous virtual block index (gridDim.x is a multiple of 8)
  const int64_t nb = gridDim.x;
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  float* rows = rows_raw[wv];
  int64_t it = a.queue ? queue_next(a.queue) : vb * 4 + wv;
  for (; it < a.ncol; it = a.queue ? queue_next(a.queue) : it + nb * 4) {
    const int64_t g = a.order ? (int64_t)a.order[it] : it;
    {
      const int2* src = (const int2*)(a.crec + (size_t)g * CT_INTS);
      if (lane < CT_INTS / 2) ((int2*)crec[wv])[lane] = src[lane];
      wave_sync_lds();
    }
    const int (*rtab)[5] = (const int (*)[5])(crec[wv] + 8);
    const int row0 = crec[wv][6], row1 = crec[wv][7];
    const double rx = ((const double*)crec[wv])[0];
    const double ry = ((const double*)crec[wv])[1];
    const double rz = ((const double*)crec[wv])[2];
    f4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    // b partials: lane (i16, kq) sums member 16t + i16 over its observation
    // slots; the four kq slots are added at the end.  Row-buffer entries of
    // members >= k hold member 0's values (the row loads of lanes past k read
    // member 0): their products only reach entries of C and b that the
    // epilogue leaves out.
    double bacc[KT];
#pragma unroll
    for (int t = 0; t < KT; ++t) bacc[t] = 0.0;
    int nl = 0, pend = 0;
    // Batches of 32 queued observations.  issue(cnt): the rows of queue
    // entries [0, cnt) start towards the wave's LDS row buffer (one
    // global_load_lds per row and 64-member block: lane m moves member m),
    // their scale and weight go to the batch arrays.  consume(): waits for
    // the rows, then scales, splits and multiplies them.  The rows of a batch
    // are in flight while the search goes on; entries j >= cnt carry zero
    // scale and weight.
    bool inflight = false;
    auto issue = [&](int cnt) {
      const int my_o = lane < cnt ? q_off[wv][lane] : 0;
      if (lane < 32) {
        bt_wd[wv][lane] = lane < cnt ? q_wd[wv][lane] : 0.f;
        bt_sw[wv][lane] = lane < cnt ? q_sw[wv][lane] : 0.f;
      }
      // row base uniform (scalar address), member offset per lane
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const unsigned lm = 64 * mb + lane < k ? 64 * mb + lane : 0;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          const int o = __builtin_amdgcn_readlane(my_o, j);
          const float* yrow = a.y32 + (size_t)o * k;
          __builtin_amdgcn_global_load_lds(
              (const void*)(yrow + lm),
              (__attribute__((address_space(3))) void*)(rows + gm_row_base<KM, PREC>(j) + 64 * mb),
              4, 0, 0);
        }
      }
      inflight = true;
    };
    auto consume = [&]() {
      // the wave's outstanding vector memory operations, the row copies
      // among them, have completed
      __builtin_amdgcn_s_waitcnt(0);
      wave_sync_lds();
      inflight = false;
      if (PREC == MDA_PREC_BF16X3) {
        // A operand of v_mfma_f32_16x16x32_bf16: lane (i16, kq) holds
        // A[i16][8kq + e], here member 16t + i16 of observation 8kq + e
        float sw8[8], wd8[8];
        {
          const float4 s0 = *(const float4*)&bt_sw[wv][8 * kq];
          const float4 s1 = *(const float4*)&bt_sw[wv][8 * kq + 4];
          const float4 w0 = *(const float4*)&bt_wd[wv][8 * kq];
          const float4 w1 = *(const float4*)&bt_wd[wv][8 * kq + 4];
          sw8[0] = s0.x; sw8[1] = s0.y; sw8[2] = s0.z; sw8[3] = s0.w;
          sw8[4] = s1.x; sw8[5] = s1.y; sw8[6] = s1.z; sw8[7] = s1.w;
          wd8[0] = w0.x; wd8[1] = w0.y; wd8[2] = w0.z; wd8[3] = w0.w;
          wd8[4] = w1.x; wd8[5] = w1.y; wd8[6] = w1.z; wd8[7] = w1.w;
        }
        bf8 fh[KT], fl[KT];
#pragma unroll
        for (int t = 0; t < KT; ++t) {
          // one tile's reads at a time (bounded live range of the row values)
          __asm__ volatile("" ::: "memory");
          const int m = 16 * t + i16;
          float av[8];
          float bp = 0.f;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const float y = rows[gm_row_base<KM, PREC>(8 * kq + e) + m];
            bp = fmaf(wd8[e], y, bp);
            av[e] = sw8[e] * y;
          }
          bacc[t] += (double)bp;
          // a = hi + lo with hardware round-to-nearest-even conversions
          u32x4 ph, pl;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const fl2 f = {av[2 * q], av[2 * q + 1]};
            const bf2 h = __builtin_convertvector(f, bf2);
            const fl2 r = f - __builtin_convertvector(h, fl2);
            const bf2 lw = __builtin_convertvector(r, bf2);
            ph[q] = __builtin_bit_cast(uint32_t, h);
            pl[q] = __builtin_bit_cast(uint32_t, lw);
          }
          fh[t] = __builtin_bit_cast(bf8, ph);
          fl[t] = __builtin_bit_cast(bf8, pl);
        }
        int tile = 0;
#pragma unroll
        for (int ta = 0; ta < KT; ++ta) {
#pragma unroll
          for (int tb = 0; tb <= ta; ++tb) {
            acc[tile] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fh[ta], fh[tb], acc[tile], 0, 0, 0);
            acc[tile] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fh[ta], fl[tb], acc[tile], 0, 0, 0);
            acc[tile] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fl[ta], fh[tb], acc[tile], 0, 0, 0);
            ++tile;
          }
        }
      } else {
        // A operand of v_mfma_f32_16x16x4f32: lane (i16, kq) holds A[i16][kq],
        // here member 16t + i16 of observation 4s + kq
        float bp[KT];
#pragma unroll
        for (int t = 0; t < KT; ++t) bp[t] = 0.f;
#pragma unroll
        for (int s4 = 0; s4 < 8; ++s4) {
          const int j = 4 * s4 + kq;
          const float sw = bt_sw[wv][j], wd = bt_wd[wv][j];
          float fr[KT];
#pragma unroll
          for (int t = 0; t < KT; ++t) {
            const int m = 16 * t + i16;
            const float y = rows[gm_row_base<KM, PREC>(j) + m];
            bp[t] = fmaf(wd, y, bp[t]);
            fr[t] = sw * y;
          }
          int tile = 0;
#pragma unroll
          for (int ta = 0; ta < KT; ++ta) {
#pragma unroll
            for (int tb = 0; tb <= ta; ++tb) {
              acc[tile] = __builtin_amdgcn_mfma_f32_16x16x4f32(fr[ta], fr[tb], acc[tile], 0, 0, 0);
              ++tile;
            }
          }
        }
#pragma unroll
        for (int t = 0; t < KT; ++t) bacc[t] += (double)bp[t];
      }
      wave_sync_lds();
    };
    // every full batch of the queue: its rows go out, the batch before is
    // consumed, the remainder moves to the front of the queue
    auto drain = [&]() {
      while (pend >= 32) {
        if (inflight) consume();
        issue(32);
        const int rest = pend - 32;
        int mo = 0;
        float mwd = 0.f, msw = 0.f;
        if (lane < rest) { mo = q_off[wv][32 + lane]; mwd = q_wd[wv][32 + lane]; msw = q_sw[wv][32 + lane]; }
        wave_sync_lds();
        if (lane < rest) { q_off[wv][lane] = mo; q_wd[wv][lane] = mwd; q_sw[wv][lane] = msw; }
        wave_sync_lds();
        pend = rest;
      }
    };
    for (int row = row0; row <= row1; ++row) {
      // this row's observation ranges (no dynamically indexed arrays)
      const row_ranges rr = row_from_table(a, row0, row, rtab);
      for (int sg = 0; sg < rr.n; ++sg) {
        const int64_t s0 = sg ? rr.s0[1] : rr.s0[0];
        const int64_t s1 = sg ? rr.s1[1] : rr.s1[0];
        // Candidate data one batch ahead: the location (unit vector, or
        // the GRID / CARTESIAN coordinates), 1/var and d of candidate
        // base + 64 + lane load while batch `base` is tested.
        const bool geo = a.coord == MDA_COORD_GEOGRAPHIC;
        const double* pa = geo ? a.ux : a.oa;
        const double* pb = geo ? a.uy : a.ob;
        const double* pc = geo ? a.uz : (a.coord == MDA_COORD_CARTESIAN ? a.oc : a.ob);
        double na = 0.0, nbv = 0.0, nc = 0.0, nr = 0.0, nd = 0.0;
        if (s0 + lane < s1) {
          const int64_t t0 = s0 + lane;
          na = pa[t0]; nbv = pb[t0]; nc = pc[t0]; nr = a.rinv[t0]; nd = a.d[t0];
        }
        for (int64_t base = s0; base < s1; base += 64) {
          // full batches queued by the previous candidates go out here, before
          // the next candidates' loads: consume()'s wait then also covers
          // this batch's candidate data, which is needed next anyway
          drain();
          const int64_t t = base + lane;
          const double ca = na, cb = nbv, cc = nc, crinv = nr, cd = nd;
          if (t + 64 < s1) {
            na = pa[t + 64]; nbv = pb[t + 64]; nc = pc[t + 64]; nr = a.rinv[t + 64]; nd = a.d[t + 64];
          }
          float w = 0.f;
          bool in = false;
          if (t < s1) {
            // membership in FP64 (the chord test c2 <= 4 sin^2(r/2R) is the
            // haversine test d <= r); the weight in FP32, the precision of
            // the contraction it feeds
            float dist;
            if (geo) {
              const double dx = rx - ca, dy = ry - cb, dz = rz - cc;
              const double c2 = dx * dx + dy * dy + dz * dz;
              in = c2 <= chord2_r;
              if (fabs(c2 - chord2_r) <= shellw) shell_push(a, g, t, in);
              dist = 2.f * (float)kEarthRadiusKm * asinf(fminf(1.f, 0.5f * sqrtf((float)c2)));
            } else {
              const double dx = rx - ca, dy = ry - cb;
              const double dz = a.coord == MDA_COORD_CARTESIAN ? rz - cc : 0.0;
              const double d2 = dist2_exact(dx, dy, dz);
              in = sqrt(d2) <= a.radius;
              dist = (float)sqrt(d2);
            }
            // the GC polynomial rounds to tiny negative values just inside its
            // support; those weights are zero here (sqrt(w) scales the rows)
            if (in)
              w = fmaxf((a.localization == MDA_LOC_GASPARI_COHN ? gaspari_cohn_f32(dist / cgcf)
                                                                : 1.f) *
                            (float)crinv,
                        0.f);
          }
          nl += __popcll(__ballot(in));
          const uint64_t wm = __ballot(w != 0.f);
          const int cnt = __popcll(wm);
          if (cnt == 0) continue;
          const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
          if (w != 0.f) {
            const int slot = pend + __popcll(wm & below);
            q_off[wv][slot] = (int)t;
            q_wd[wv][slot] = w * (float)cd;
            q_sw[wv][slot] = sqrtf(w);
          }
          pend += cnt;
          wave_sync_lds();
        }
      }
    }
    drain();
    if (inflight) consume();
    if (pend > 0) {
      issue(pend);
      consume();
    }
    // Projection onto the complement of the ensemble-mean direction:
    // Y'1 = 0 exactly, so C1 = 0 and 1^T b = 0 in exact arithmetic; the
    // reduced-precision products break this at their rounding level and the
    // eigenvalue (k-1)/rho of A along 1 would drift below the null window.
    // C <- PCP, b <- Pb with P = I - 11^T/k.  Row sums in FP32 (their error,
    // ~1e-7 ||C||, is far inside the null window):
    //   row i of tile (ta, tb <= ta): the 16 lanes of one kq group, summed by
    //   xor-shuffles over i16; the mirrored upper part of row i is column i
    //   of the tiles below the diagonal, summed over kq and r.
    double* rsum = (double*)rows;  // [KM] row sums (the row buffer is free now)
    // lane coordinates through opaque copies: the element indices of this
    // epilogue are recomputed per column instead of being kept live across
    // the candidate loop
    int kq_e = kq, i16_e = i16;
    __asm__ volatile("" : "+v"(kq_e), "+v"(i16_e));
    {
      float rowp[KT][4], colp[KT];
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        colp[t] = 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) rowp[t][r] = 0.f;
      }
      int tile = 0;
#pragma unroll
      for (int ta = 0; ta < KT; ++ta) {
#pragma unroll
        for (int tb = 0; tb <= ta; ++tb) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // entries of members >= k hold products of member 0's copy: left out
            const bool cv = 16 * tb + i16_e < k, rv = 16 * ta + 4 * kq_e + r < k;
            rowp[ta][r] += cv ? acc[tile][r] : 0.f;
            if (tb != ta) colp[tb] += rv ? acc[tile][r] : 0.f;
          }
          ++tile;
        }
      }
#pragma unroll
      for (int t = 0; t < KT; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          for (int off = 1; off < 16; off <<= 1) rowp[t][r] += __shfl_xor(rowp[t][r], off);
        colp[t] += __shfl_xor(colp[t], 16);
        colp[t] += __shfl_xor(colp[t], 32);
      }
      if (i16_e == 0) {
#pragma unroll
        for (int t = 0; t < KT; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) rsum[16 * t + 4 * kq_e + r] = (double)rowp[t][r];
      }
      wave_sync_lds();
      if (kq_e == 0) {
#pragma unroll
        for (int t = 0; t < KT; ++t) rsum[16 * t + i16_e] += (double)colp[t];
      }
      wave_sync_lds();
    }
    double tsum = 0.0;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
      if (64 * mb + lane < k) tsum += rsum[64 * mb + lane];
    for (int off = 32; off > 0; off >>= 1) tsum += __shfl_xor(tsum, off);
    const double rk = 1.0 / (double)k;
    const double tot = tsum * rk * rk;
    // b_m: the four observation slots of member m (lanes i16 + 16 kq) added;
    // members >= k left out
    double bsum = 0.0;
#pragma unroll
    for (int t = 0; t < KT; ++t) {
      bacc[t] += __shfl_xor(bacc[t], 16);
      bacc[t] += __shfl_xor(bacc[t], 32);
      if (16 * t + i16_e < k) bsum += bacc[t];
    }
    for (int off = 8; off > 0; off >>= 1) bsum += __shfl_xor(bsum, off);
    if (kq_e == 0) {
#pragma unroll
      for (int t = 0; t < KT; ++t)
        if (16 * t + i16_e < k) a.b[(size_t)g * k + 16 * t + i16_e] = bacc[t] - bsum * rk;
    }
    double* Cg = a.C + (size_t)g * k * k;
    float* Cf = reinterpret_cast<float*>(a.C) + (size_t)g * k * k;
    const bool c32 = a.c32 != 0;
    // k through an opaque copy: the element offsets below are recomputed per
    // column instead of being kept live across the candidate loop
    int ko = k;
    __asm__ volatile("" : "+s"(ko));
    int tile = 0;
#pragma unroll
    for (int ta = 0; ta < KT; ++ta) {
#pragma unroll
      for (int tb = 0; tb <= ta; ++tb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * ta + 4 * kq_e + r, col = 16 * tb + i16_e;
          if (row < ko && col < ko) {
            const double v = (double)acc[tile][r] - (rsum[row] + rsum[col]) * rk + tot;
            if (c32) {
              Cf[row * ko + col] = (float)v;
              if (ta != tb) Cf[col * ko + row] = (float)v;
            } else {
              Cg[row * ko + col] = v;
              if (ta != tb) Cg[col * ko + row] = v;
            }
          }
        }
        ++tile;
      }
    }
    wave_sync_lds();  // rsum reads done before the next column's batches
    if (lane == 0) a.nloc[g] = nl;
  }
}

template <int KT, int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_letkf_gram_mixed(letkf_gram_args a) {
  gram_mixed_body<KT, PREC>(a);
}

template <int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1)))
void k_letkf_gram_mixed_w(letkf_gram_args a) {
  gram_mixed_body<8, PREC>(a);
}

// ---------------------------------------------------------------------------
// Exact decisions for the boundary list of one pass (shell_push): each listed
// (column, observation) pair is decided by the haversine distance of the
// reference (Location.hpp:349-357) with correctly rounded sin / cos / atan2.
// Where that differs from the chord test the Gram used, the column's local
// count moves by one and its C and b take (+ or -) the pair's weighted
// contribution w y'y'^T and w y' d (w = GC(d/c) / var, which this close to
// the support is below 1e-27 / var, or 1 / var with the hard cutoff).  One
// wavefront per pair; a list longer than its capacity sets
// MDA_ST_BOUNDARY_OVERFLOW.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_letkf_shell_fixup(letkf_gram_args a,
                                                           int* __restrict__ status) {
  const unsigned cnt = *a.shell_count;
  if (cnt > (unsigned)a.shell_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(status, MDA_ST_BOUNDARY_OVERFLOW);
    return;
  }
  const int lane = threadIdx.x & 63;
  const int k = a.k;
  for (int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); e < (int64_t)cnt;
       e += (int64_t)gridDim.x * 4) {
    const int4 v = a.shell[e];
    const int64_t g = v.x;
    const int64_t t = (int64_t)(uint32_t)v.y | ((int64_t)v.z << 32);
    const bool chord_in = v.w != 0;
    const double dist = cr::haversine_km_cr(a.col_a[g], a.col_b[g], a.oa[t], a.ob[t]);
    const bool exact_in = dist <= a.radius;
    if (exact_in == chord_in) continue;
    if (lane == 0) atomicAdd(&a.nloc[g], exact_in ? 1 : -1);
    const double w =
        (a.localization == MDA_LOC_GASPARI_COHN ? gaspari_cohn(dist / (a.radius / 2.0)) : 1.0) *
        a.rinv[t];
    const double sg = exact_in ? w : -w;
    if (sg == 0.0) continue;
    const double* y = a.yp + (size_t)t * k;
    if (a.c32) {
      float* Cf = reinterpret_cast<float*>(a.C) + (size_t)g * k * k;
      for (int idx = lane; idx < k * k; idx += 64)
        atomicAdd(&Cf[idx], (float)(sg * y[idx / k] * y[idx % k]));
    } else {
      double* Cg = a.C + (size_t)g * k * k;
      for (int idx = lane; idx < k * k; idx += 64)
        atomicAdd(&Cg[idx], sg * y[idx / k] * y[idx % k]);
    }
    for (int i = lane; i < k; i += 64) atomicAdd(&a.b[(size_t)g * k + i], sg * y[i] * a.d[t]);
  }
}

// ---------------------------------------------------------------------------
// Local innovation and spread statistics of each column, the inputs of the
// learned inflation.  Same candidate search and weights as the FP64 Gram
// (w_o = localisation x 1/var_o); per column
//   Sw = sum w,  Sd2 = sum w d^2,  Sd = sum w d,  Ssp = sum w s_o
// with s_o = |y'_o|^2 / (k-1) the background variance of observation o, and
// the column's background variance v (member variance averaged over levels).
// Features (FP32): log(Sd2/Sw), log(Ssp/Sw), log(Sd2/(Ssp+Sw)),
// Sd/sqrt(Sw Sd2), log(Sw), log(1+n_local), log(v), rho_prev - 1.
// ---------------------------------------------------------------------------
__global__ void k_obs_spread(const double* __restrict__ s_yp, int64_t p, int k,
                             double* __restrict__ sp) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < p;
       t += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int m = 0; m < k; ++m) acc += s_yp[(size_t)t * k + m] * s_yp[(size_t)t * k + m];
    sp[t] = acc / (double)(k - 1);
  }
}

__global__ __launch_bounds__(256) void k_letkf_features(letkf_gram_args a,
                                                        const double* __restrict__ sp,
                                                        const double* __restrict__ X,
                                                        int64_t nlev,
                                                        const double* __restrict__ rho_prev,
                                                        float* __restrict__ feat) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int k = a.k;
  const bucket_grid& G = a.grid;
  const double cgc = a.radius / 2.0;
  const int64_t n = a.ncol * nlev;
  for (int64_t g = (int64_t)blockIdx.x * 4 + wv; g < a.ncol; g += (int64_t)gridDim.x * 4) {
    const col_window cw = column_window(a, g);

    double Sw = 0.0, Sd2 = 0.0, Sd = 0.0, Ssp = 0.0;
    int nl = 0;
    for (int row = cw.row0; row <= cw.row1; ++row) {
      const seg_list rs = row_segments(a, cw, row);
      const int sj00 = rs.j0[0], sj01 = rs.j1[0];
      const int sj10 = rs.n > 1 ? rs.j0[1] : 0, sj11 = rs.n > 1 ? rs.j1[1] : 0;
      for (int sg = 0; sg < rs.n; ++sg) {
        const int64_t s0 = a.bstart[row * G.ncolb + (sg ? sj10 : sj00)];
        const int64_t s1 = a.bstart[row * G.ncolb + (sg ? sj11 : sj01) + 1];
        for (int64_t t = s0 + lane; t < s1; t += 64) {
          double dist;
          if (a.coord == MDA_COORD_GEOGRAPHIC) {
            dist = haversine_km(cw.ca, cw.cb, a.oa[t], a.ob[t]);
            // near the radius, the reference's distance to the last bit
            if (fabs(dist - a.radius) <= 1e-9 * a.radius)
              dist = cr::haversine_km_cr(cw.ca, cw.cb, a.oa[t], a.ob[t]);
          } else {
            const double dx = cw.ca - a.oa[t], dy = cw.cb - a.ob[t];
            const double dz = a.coord == MDA_COORD_CARTESIAN ? cw.cc - a.oc[t] : 0.0;
            dist = sqrt(dist2_exact(dx, dy, dz));
          }
          if (dist <= a.radius) {
            ++nl;
            const double w = fmax((a.localization == MDA_LOC_GASPARI_COHN ? gaspari_cohn(dist / cgc)
                                                                           : 1.0) *
                                      a.rinv[t],
                                  0.0);
            const double d = a.d[t];
            Sw += w;
            Sd2 += w * d * d;
            Sd += w * d;
            Ssp += w * sp[t];
          }
        }
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      Sw += __shfl_xor(Sw, off);
      Sd2 += __shfl_xor(Sd2, off);
      Sd += __shfl_xor(Sd, off);
      Ssp += __shfl_xor(Ssp, off);
      nl += __shfl_xor(nl, off);
    }
    // background variance of the column, averaged over levels
    double var = 0.0;
    for (int64_t l = 0; l < nlev; ++l) {
      double sx = 0.0, sxx = 0.0;
      for (int m = lane; m < k; m += 64) {
        const double v = X[(size_t)m * n + l * a.ncol + g];
        sx += v;
        sxx += v * v;
      }
      for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sxx += __shfl_xor(sxx, off);
      }
      var += fmax(sxx - sx * sx / k, 0.0) / (double)(k - 1);
    }
    var /= (double)nlev;
    if (lane == 0) {
      const double tiny = 1e-300;
      float* f = feat + (size_t)g * 8;
      const bool has = Sw > 0.0;
      f[0] = has ? (float)log(Sd2 / Sw + tiny) : 0.f;
      f[1] = has ? (float)log(Ssp / Sw + tiny) : 0.f;
      f[2] = has ? (float)log(Sd2 / (Ssp + Sw) + tiny) : 0.f;
      f[3] = has && Sd2 > 0.0 ? (float)(Sd / sqrt(Sw * Sd2)) : 0.f;
      f[4] = has ? (float)log(Sw) : 0.f;
      f[5] = (float)log1p((double)nl);
      f[6] = (float)log(var + tiny);
      f[7] = rho_prev ? (float)(rho_prev[g] - 1.0) : 0.f;
    }
  }
}

void launch_letkf_features(hipStream_t s, const letkf_gram_args& a, double* sp_work,
                           const double* X, int64_t nlev, const double* rho_prev, float* feat) {
  hipLaunchKernelGGL(k_letkf_features, dim3(grid_for(a.ncol, 4, 262144)), dim3(256), 0, s, a,
                     sp_work, X, nlev, rho_prev, feat);
  MDA_LAUNCH_CHECK();
}

void launch_obs_spread(hipStream_t s, const double* s_yp, int64_t p, int k, double* sp) {
  if (p <= 0) return;
  hipLaunchKernelGGL(k_obs_spread, dim3(grid_for(p, 256, 65536)), dim3(256), 0, s, s_yp, p, k,
                     sp);
  MDA_LAUNCH_CHECK();
}

void launch_letkf_shell_fixup(hipStream_t s, const letkf_gram_args& a, int* status) {
  if (a.coord != MDA_COORD_GEOGRAPHIC) return;
  MDA_REQUIRE(a.shell && a.shell_count, "boundary list missing");
  hipLaunchKernelGGL(k_letkf_shell_fixup, dim3(64), dim3(256), 0, s, a, status);
  MDA_LAUNCH_CHECK();
}

void launch_column_setup(hipStream_t s, const letkf_gram_args& a) {
  MDA_REQUIRE(a.crec, "column setup records missing");
  if (a.ncol <= 0) return;
  hipLaunchKernelGGL(k_column_setup, dim3(grid_for(a.ncol * CT_EPC, 256, 262144)), dim3(256), 0,
                     s, a, a.crec);
  MDA_LAUNCH_CHECK();
}

void launch_letkf_gram(hipStream_t s, const letkf_gram_args& a) {
  MDA_REQUIRE(a.k >= 1 && a.k <= 128, "LETKF Gram supports k <= 128");
  // FP32 entries of C exactly for the FP32 / BF16x3 contractions (the FP64
  // kernels write doubles)
  MDA_REQUIRE((a.c32 != 0) == (a.precision != MDA_PREC_FP64), "Gram storage precision mismatch");
  const int npart = a.k > 64 ? 2 : 1;  // waves per column (FP64 Gram)
  unsigned nb = (unsigned)((a.ncol * (a.precision == MDA_PREC_FP64 ? npart : 1) + 3) / 4);
  if (nb > 262144) nb = 262144;
  if (a.nblocks > 0 && (unsigned)a.nblocks < nb) nb = (unsigned)a.nblocks;
  MDA_REQUIRE(!a.queue || a.ncol * npart < ((int64_t)1 << 32) - 262144 * 4,
              "too many columns for the queue");
  int kt = (a.k + 15) / 16;
  if (a.precision != MDA_PREC_FP64) {
    MDA_REQUIRE(a.y32 && (a.coord != MDA_COORD_GEOGRAPHIC || a.ux) && a.crec,
                "mixed Gram inputs missing");
    nb = (nb + 7) / 8 * 8;  // XCD-contiguous mapping needs a multiple of 8
    const bool bf = a.precision == MDA_PREC_BF16X3;
#define MDA_GM(KT)                                                                         \
  if (bf) hipLaunchKernelGGL((k_letkf_gram_mixed<KT, MDA_PREC_BF16X3>), dim3(nb), dim3(256), \
                             0, s, a);                                                     \
  else hipLaunchKernelGGL((k_letkf_gram_mixed<KT, MDA_PREC_FP32>), dim3(nb), dim3(256), 0, s, a);
    switch (kt) {
      case 1: MDA_GM(1); break;
      case 2: MDA_GM(2); break;
      case 3: MDA_GM(3); break;
      case 4: MDA_GM(4); break;
      default:  // 64 < k <= 128
        if (bf) hipLaunchKernelGGL((k_letkf_gram_mixed_w<MDA_PREC_BF16X3>), dim3(nb), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_letkf_gram_mixed_w<MDA_PREC_FP32>), dim3(nb), dim3(256), 0, s, a);
        break;
    }
#undef MDA_GM
    MDA_LAUNCH_CHECK();
    return;
  }
  MDA_REQUIRE(a.crec && (a.coord != MDA_COORD_GEOGRAPHIC || a.ux), "FP64 Gram inputs missing");
  switch (kt) {
    case 1: hipLaunchKernelGGL((k_letkf_gram_f64<1>), dim3(nb), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_letkf_gram_f64<2>), dim3(nb), dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_letkf_gram_f64<3>), dim3(nb), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_letkf_gram_f64<4>), dim3(nb), dim3(256), 0, s, a); break;
    default:  // 64 < k <= 128: 8 tile rows, two waves per column
      hipLaunchKernelGGL(k_letkf_gram_f64_w, dim3(nb), dim3(256), 0, s, a);
      break;
  }
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// 3. per-column solve.  With sigma = (k-1)/rho (prior inflation, Hunt et al.
// 2007) or k-1 (posterior: Pa *= rho, LETKF.hpp:219) and s = 1 or rho:
//   M = C + sigma I,  Pa = s M^-1,  wbar = Pa b,
//   W = sqrt(k-1) Pa^(1/2)  (symmetric: eigen-decomposition of M by Jacobi;
//                            Cholesky: chol_L(Pa), LETKF.hpp:223-224),
//   T = wbar 1^T + W  (matrix combination) or diag(wbar + W 1) (the
//   element-wise combination of LETKF.hpp:235-238).
// A column without local observations gets T = sqrt(rho) I (LETKF.hpp:167-190).
// ---------------------------------------------------------------------------
// Columns the fallback solve works on (flagged by the eigensolver, or without
// local observations), compacted into a list with wave-aggregated atomics.
// The list's order varies from run to run; every column is solved on its own,
// so the results do not.
__global__ void k_solve_count_zero(unsigned* cnt) { *cnt = 0u; }

__global__ void k_solve_worklist(int64_t ncol, const int* __restrict__ only,
                                 const int* __restrict__ nloc, int* __restrict__ list,
                                 unsigned* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ncol;
       g += (int64_t)gridDim.x * blockDim.x) {
    const bool need = only[g] != 0 || nloc[g] == 0;
    const uint64_t m = __ballot(need);
    if (m == 0ull) continue;
    const int leader = __builtin_ctzll(m);
    unsigned base = 0u;
    if (lane == leader) base = atomicAdd(cnt, (unsigned)__popcll(m));
    base = (unsigned)__shfl((int)base, leader);
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (need) list[base + __popcll(m & below)] = (int)g;
  }
}

__global__ __launch_bounds__(256) void k_letkf_solve(letkf_solve_args a, double* gscratch,
                                                     int use_lds, const int* only,
                                                     const int* list, const unsigned* cnt) {
  extern __shared__ double sh[];
  __shared__ double rc[64], rs[64], red[256], bv[128], yv[128], wb[128];
  __shared__ int rp[64], rq[64], flag;
  const int k = a.k, tid = threadIdx.x, nt = blockDim.x;
  const int kp = (k + 1) & ~1, ld = kp + 1;
  double* A = use_lds ? sh : gscratch + (size_t)blockIdx.x * 2 * kp * ld;
  double* V = A + kp * ld;
  // fallback mode: the listed columns (flagged, or empty: sqrt(rho) I);
  // otherwise every column
  const int64_t nwork = list ? (int64_t)*cnt : a.ncol;
  for (int64_t it = blockIdx.x; it < nwork; it += gridDim.x) {
    const int64_t g = list ? (int64_t)list[it] : it;
    if (only && only[g] == 0 && a.nloc[g] != 0) continue;
    const double rho = a.rho_col ? a.rho_col[g] : a.rho;
    double* Tg = a.T + (size_t)g * k * k;
    if (a.nloc[g] == 0) {
      const double sr = sqrt(rho);
      for (int idx = tid; idx < k * k; idx += nt) Tg[idx] = (idx / k == idx % k) ? sr : 0.0;
      if (tid == 0 && a.sweeps) a.sweeps[g] = 0;
      continue;
    }
    const bool prior = a.inflation_form == MDA_INFLFORM_PRIOR;
    const double sigma = prior ? (double)(k - 1) / rho : (double)(k - 1);
    const double sc = prior ? 1.0 : rho;
    const double* Cg = a.C + (size_t)g * k * k;
    const float* Cf = reinterpret_cast<const float*>(a.C) + (size_t)g * k * k;
    for (int idx = tid; idx < kp * kp; idx += nt) {
      int i = idx / kp, j = idx % kp;
      double v;
      if (i < k && j < k)
        v = (a.c32 ? (double)Cf[i * k + j] : Cg[i * k + j]) + (i == j ? sigma : 0.0);
      else v = (i == j) ? 1.0 : 0.0;  // decoupled pad index (odd k)
      A[i * ld + j] = v;
    }
    for (int i = tid; i < k; i += nt) bv[i] = a.b[(size_t)g * k + i];
    __syncthreads();
    const double sq = sqrt((double)(k - 1));
    if (a.sqrt_form == MDA_SQRT_SYMMETRIC) {
      int sw = block_jacobi(A, ld, V, ld, kp, rc, rs, rp, rq, red, 40, 1e-15);
      if (tid == 0 && a.sweeps) a.sweeps[g] = sw;
      // y = L^-1 V^T b ; wbar = s V y
      for (int e = tid; e < k; e += nt) {
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += V[i * ld + e] * bv[i];
        yv[e] = s / A[e * ld + e];
      }
      __syncthreads();
      for (int i = tid; i < k; i += nt) {
        double s = 0.0;
        for (int e = 0; e < k; ++e) s += V[i * ld + e] * yv[e];
        wb[i] = sc * s;
      }
      __syncthreads();
      // U = V L^-1/4 so that W = sqrt(k-1) sqrt(s) U U^T
      for (int idx = tid; idx < k * k; idx += nt) {
        int i = idx / k, e = idx % k;
        V[i * ld + e] *= 1.0 / sqrt(sqrt(A[e * ld + e]));
      }
      __syncthreads();
      const double coef = sq * sqrt(sc);
      if (a.combine == MDA_COMBINE_MATRIX) {
        for (int idx = tid; idx < k * k; idx += nt) {
          int i = idx / k, j = idx % k;
          double s = 0.0;
          for (int e = 0; e < k; ++e) s += V[i * ld + e] * V[j * ld + e];
          Tg[idx] = wb[i] + coef * s;
        }
      } else {
        for (int i = tid; i < k; i += nt) {
          double rsum = 0.0;
          for (int j = 0; j < k; ++j) {
            double s = 0.0;
            for (int e = 0; e < k; ++e) s += V[i * ld + e] * V[j * ld + e];
            rsum += coef * s;
          }
          yv[i] = wb[i] + rsum;
        }
        __syncthreads();
        for (int idx = tid; idx < k * k; idx += nt)
          Tg[idx] = (idx / k == idx % k) ? yv[idx / k] : 0.0;
      }
    } else {
      bool ok = block_cholesky(A, ld, k, &flag);
      block_spd_inverse_from_chol(A, ld, V, ld, k);  // A <- M^-1
      for (int idx = tid; idx < k * k; idx += nt) A[(idx / k) * ld + idx % k] *= sc;
      __syncthreads();
      for (int i = tid; i < k; i += nt) {
        double s = 0.0;
        for (int c = 0; c < k; ++c) s += A[i * ld + c] * bv[c];
        wb[i] = s;
      }
      __syncthreads();
      ok = block_cholesky(A, ld, k, &flag) && ok;  // A <- chol_L(Pa)
      if (tid == 0 && a.status && !ok) atomicOr(a.status, 1);
      if (a.combine == MDA_COMBINE_MATRIX) {
        for (int idx = tid; idx < k * k; idx += nt) {
          int i = idx / k, j = idx % k;
          Tg[idx] = wb[i] + sq * A[i * ld + j];
        }
      } else {
        for (int i = tid; i < k; i += nt) {
          double rsum = 0.0;
          for (int j = 0; j < k; ++j) rsum += sq * A[i * ld + j];
          yv[i] = wb[i] + rsum;
        }
        __syncthreads();
        for (int idx = tid; idx < k * k; idx += nt)
          Tg[idx] = (idx / k == idx % k) ? yv[idx / k] : 0.0;
      }
      if (tid == 0 && a.sweeps) a.sweeps[g] = 0;
    }
    __syncthreads();
  }
}

void launch_letkf_solve(hipStream_t s, const letkf_solve_args& a) {
  MDA_REQUIRE(a.k >= 2 && a.k <= 128, "LETKF solve supports 2 <= k <= 128");
  const int kp = (a.k + 1) & ~1, ld = kp + 1;
  size_t shm = sizeof(double) * 2 * kp * ld;
  int use_lds = shm <= 96 * 1024;
  unsigned nb = (unsigned)(a.ncol < 65536 ? a.ncol : 65536);
  if (nb < 1) nb = 1;
  double* gscratch = nullptr;
  if (!use_lds) {  // A and V in a global per-block scratch
    MDA_REQUIRE(a.jscratch && a.jblocks > 0, "Jacobi scratch missing");
    gscratch = a.jscratch;
    if (nb > (unsigned)a.jblocks) nb = (unsigned)a.jblocks;
  }
  const int* only = nullptr;
  const int* list = nullptr;
  const unsigned* cnt = nullptr;
  if (a.sqrt_form == MDA_SQRT_SYMMETRIC && letkf_eig_supported(a.k) && a.method == 0) {
    // tridiagonal eigensolver; flagged columns fall back to Jacobi below
    launch_letkf_eig(s, a, a.de, a.lam, a.flag);
    only = a.flag;
    if (a.wlist && a.wcount) {
      // the columns to solve, listed, so that an analysis without fallback
      // columns costs a scan instead of a pass over every column
      hipLaunchKernelGGL(k_solve_count_zero, dim3(1), dim3(1), 0, s, a.wcount);
      MDA_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_solve_worklist, dim3(grid_for(a.ncol, 256, 4096)), dim3(256), 0, s,
                         a.ncol, a.flag, a.nloc, a.wlist, a.wcount);
      MDA_LAUNCH_CHECK();
      list = a.wlist;
      cnt = a.wcount;
    }
  }
  hipLaunchKernelGGL(k_letkf_solve, dim3(nb), dim3(256), use_lds ? shm : 0, s, a, gscratch,
                     use_lds, only, list, cnt);
  MDA_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// 4. X_a = xbar + X' T_g for a tile of 16 consecutive columns x 8 levels: the
// tile is staged through LDS with coalesced 128-byte rows, the mean is formed
// per (level, column) as in Ensemble::RecomputeMean, and the product runs out
// of LDS with T_g streamed from global memory.
// ---------------------------------------------------------------------------
constexpr int TR_C = 16, TR_L = 8;

__global__ __launch_bounds__(256) void k_letkf_transform(double* __restrict__ X, int k,
                                                         int64_t ncol, int64_t nlev,
                                                         const double* __restrict__ T,
                                                         int64_t col0, int64_t nc,
                                                         int* __restrict__ status) {
  if (*status & MDA_ST_SKIP_UPDATE) return;
  extern __shared__ double xs[];  // [TR_L][TR_C][k+1]
  __shared__ double xm[TR_L][TR_C];
  const int tid = threadIdx.x, nt = blockDim.x, ld = k + 1;
  const int64_t n = ncol * nlev;
  // XCD-contiguous tile order (gridDim.x is a multiple of 8): the level tiles
  // of one column tile run back to back on one XCD, so T_g is read from HBM
  // once and from that XCD's L2 afterwards
  const int64_t nb = gridDim.x;
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t nlt = (nlev + TR_L - 1) / TR_L;
  if (vb >= ((nc + TR_C - 1) / TR_C) * nlt) return;
  // g0: first column of the tile within the chunk [col0, col0 + nc); T is
  // indexed within the chunk, X by the global column
  const int64_t g0 = (vb / nlt) * TR_C, l0 = (vb % nlt) * TR_L;
  const int ncl = (int)(nc - g0 < TR_C ? nc - g0 : TR_C);
  const int nll = (int)(nlev - l0 < TR_L ? nlev - l0 : TR_L);
  for (int idx = tid; idx < k * TR_L * TR_C; idx += nt) {
    int c = idx % TR_C, l = (idx / TR_C) % TR_L, m = idx / (TR_C * TR_L);
    if (c < ncl && l < nll)
      xs[(l * TR_C + c) * ld + m] = X[(size_t)m * n + (l0 + l) * ncol + col0 + g0 + c];
  }
  __syncthreads();
  const double rk = 1.0 / (double)k;
  for (int idx = tid; idx < TR_L * TR_C; idx += nt) {
    int l = idx / TR_C, c = idx % TR_C;
    if (c < ncl && l < nll) {
      double* row = xs + (l * TR_C + c) * ld;
      double s = 0.0;
      for (int m = 0; m < k; ++m) s += row[m];
      double mean = s * rk;
      if (!isfinite(mean)) atomicOr(status, MDA_ST_NONFINITE_STATE);
      xm[l][c] = mean;
      for (int m = 0; m < k; ++m) row[m] -= mean;
    }
  }
  __syncthreads();
  constexpr int QMAX = 8;
  double acc[QMAX][TR_L];
  const int nq = (TR_C * k + nt - 1) / nt;
#pragma unroll
  for (int q = 0; q < QMAX; ++q)
#pragma unroll
    for (int l = 0; l < TR_L; ++l) acc[q][l] = 0.0;
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    if (q >= nq) break;
    int e = tid + q * nt;
    int c = e / k, j = e % k;
    if (c < ncl) {
      const double* Tg = T + (size_t)(g0 + c) * k * k;
      for (int m = 0; m < k; ++m) {
        double t = Tg[(size_t)m * k + j];
#pragma unroll
        for (int l = 0; l < TR_L; ++l) acc[q][l] += xs[(l * TR_C + c) * ld + m] * t;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    if (q >= nq) break;
    int e = tid + q * nt;
    int c = e / k, j = e % k;
    if (c < ncl)
#pragma unroll
      for (int l = 0; l < TR_L; ++l)
        if (l < nll) xs[(l * TR_C + c) * ld + j] = xm[l][c] + acc[q][l];
  }
  __syncthreads();
  for (int idx = tid; idx < k * TR_L * TR_C; idx += nt) {
    int c = idx % TR_C, l = (idx / TR_C) % TR_L, m = idx / (TR_C * TR_L);
    if (c < ncl && l < nll)
      X[(size_t)m * n + (l0 + l) * ncol + col0 + g0 + c] = xs[(l * TR_C + c) * ld + m];
  }
}

// 4'. X_a = xbar + X' T_g with T_g held in registers: thread (c, j) keeps
// column j of T for column g0 + c (K doubles) and sweeps the levels, so T
// is read from memory exactly once per analysis.  A block covers XT_C = 4
// consecutive columns and all K members; levels are staged through LDS XT_L
// at a time as [level][column][member].  One pass over a staged row gives
// both the member sum s and the product p = sum_m x_m T[m][j]; with
// xbar = s / k and sum_m (x_m - xbar) T[m][j] = p - xbar colsum_j,
//   X_a[j] = xbar (1 - colsum_j) + p.
// Lane layout: c = lane % 4, j = 16 * wave + lane / 4, so a store
// instruction writes 16 members x 32 B.
constexpr int XT_C = 4, XT_L = 20;

template <int K>
__global__ __launch_bounds__(256) void k_letkf_transform_reg(double* __restrict__ X,
                                                             int64_t ncol, int64_t nlev,
                                                             const double* __restrict__ T,
                                                             int64_t col0, int64_t nc,
                                                             int* __restrict__ status) {
  if (*status & MDA_ST_SKIP_UPDATE) return;
  constexpr int KP = K + 2;  // row pad: 16-byte aligned, banks shift by 4 per row
  extern __shared__ double xs[];  // [XT_L][XT_C][KP]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c = lane % XT_C, j = 16 * wv + lane / XT_C;
  const int64_t n = ncol * nlev;
  const int64_t nb = gridDim.x;
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t g0 = vb * XT_C;  // within the chunk [col0, col0 + nc)
  if (g0 >= nc) return;
  const int ncl = (int)(nc - g0 < XT_C ? nc - g0 : XT_C);
  const bool own = j < K && c < ncl;
  double t[K];
  double colsum = 0.0;
  {
    const double* Tg = T + (size_t)(g0 + (c < ncl ? c : 0)) * K * K + (j < K ? j : 0);
#pragma unroll
    for (int m = 0; m < K; ++m) {
      t[m] = own ? Tg[(size_t)m * K] : 0.0;
      colsum += t[m];
    }
  }
  const double rk = 1.0 / (double)K;
  const double cmean = 1.0 - colsum;
  for (int64_t l0 = 0; l0 < nlev; l0 += XT_L) {
    const int nll = (int)(nlev - l0 < XT_L ? nlev - l0 : XT_L);
    __syncthreads();  // previous chunk's readers are done with xs
    for (int idx = tid; idx < K * XT_L * XT_C; idx += 256) {
      const int cc = idx % XT_C, l = (idx / XT_C) % XT_L, m = idx / (XT_C * XT_L);
      if (cc < ncl && l < nll)
        xs[(l * XT_C + cc) * KP + m] = X[(size_t)m * n + (l0 + l) * ncol + col0 + g0 + cc];
    }
    __syncthreads();
    if (own) {
      for (int l = 0; l < nll; ++l) {
        const double* row = xs + (l * XT_C + c) * KP;
        double a0 = 0.0, a1 = 0.0, s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int m = 0; m < K; m += 2) {
          a0 += row[m] * t[m];
          s0 += row[m];
          if (m + 1 < K) {
            a1 += row[m + 1 < K ? m + 1 : m] * t[m + 1 < K ? m + 1 : m];
            s1 += row[m + 1 < K ? m + 1 : m];
          }
        }
        const double xa = (s0 + s1) * rk * cmean + (a0 + a1);
        if (!isfinite(xa)) atomicOr(status, MDA_ST_NONFINITE_STATE);
        X[(size_t)j * n + (l0 + l) * ncol + col0 + g0 + c] = xa;
      }
    }
  }
}

// 4''. X_a = xbar (1 - colsum) + X' T as one FP64 MFMA product per column
// (mda_xform.hpp): a block covers XM_C = 4 consecutive columns, one
// wavefront each, T' of the column in registers.  `only` (optional, indexed
// within the pass) restricts the update to the columns with only[g] != 0:
// the columns the eigenvector kernel left to the Jacobi fallback.
template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_letkf_transform_mfma(double* __restrict__ X, int64_t ncol, int64_t nlev,
                            const double* __restrict__ T, int64_t col0, int64_t nc,
                            int* __restrict__ status, const int* __restrict__ only) {
  static_assert(K % 16 == 0 && K <= 64, "K must be a multiple of 16, at most 64");
  if (*status & MDA_ST_SKIP_UPDATE) return;
  constexpr int NTJ = K / 16, NS = K / 4;
  __shared__ double xs[xm_xs_size<K>()];  // staging of xm_apply
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i16 = lane & 15, kk = lane >> 4;
  const int64_t nb = gridDim.x;
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t g0 = vb * XM_C;  // within the pass [col0, col0 + nc)
  if (g0 >= nc) return;
  const int ncl = (int)(nc - g0 < XM_C ? nc - g0 : XM_C);
  unsigned store = (1u << ncl) - 1u;
  if (only) {
    unsigned m = 0u;
    for (int c = 0; c < ncl; ++c) m |= only[g0 + c] != 0 ? 1u << c : 0u;
    store = m;
    if (store == 0u) return;  // block-uniform
  }
  const bool own = (store >> wv) & 1u;
  // T of column g0 + wv in the B layout: tb[s][tj] = T[4s + kk][16 tj + i16]
  double tb[NS][NTJ];
  {
    const double* Tg = T + (size_t)(g0 + (own ? wv : 0)) * K * K;
#pragma unroll
    for (int s4 = 0; s4 < NS; ++s4)
#pragma unroll
      for (int tj = 0; tj < NTJ; ++tj)
        tb[s4][tj] = own ? Tg[(size_t)(4 * s4 + kk) * K + 16 * tj + i16] : 0.0;
  }
  xm_tprime<K>(tb);
  xm_apply<K>(X, ncol, nlev, col0 + g0, ncl, store, tb, xs, status);
}

// 4''''. X_a = X T' for K = 128 on FP64 MFMA: a block of 8 wavefronts takes
// X2_C = 2 consecutive columns, four wavefronts per column, each holding T'
// for two of the eight 16-member output tiles in registers (B layout,
// tb[s][tt] = T'[4s + kk][16 (2 tp + tt) + i16], 64 doubles per lane).  X is
// staged through LDS as [column][member][16 levels] from 16-byte column
// segments, X_a leaves the same way.  `only` as in the 4-column kernel.
constexpr int X2_C = 2;

template <int K>
__global__ __launch_bounds__(512) void k_letkf_transform_mfma2(
    double* __restrict__ X, int64_t ncol, int64_t nlev, const double* __restrict__ T,
    int64_t col0, int64_t nc, int* __restrict__ status, const int* __restrict__ only) {
  static_assert(K == 128, "eight 16-member tiles over four wavefronts");
  if (*status & MDA_ST_SKIP_UPDATE) return;
  constexpr int NS = K / 4, NT = 2, NV = X2_C * XM_L * K / 512;
  // X in as [column][member][level], X_a out as [column][level][j] with rows
  // of K + 16; each column's block starts 16 bank pairs further on, so the 32
  // lanes of each half wavefront fall on 32 different 64-bit bank pairs in
  // the staging writes, the product reads, the write-out and the store-side
  // reads
  constexpr int LJ = K + 16, CSI = K * XM_L + 16, CSO = XM_L * LJ + 16;
  constexpr int XS = X2_C * (CSI > CSO ? CSI : CSO);
  __shared__ double xs[XS];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c = wv >> 2, tp = wv & 3;
  const int i16 = lane & 15, kk = lane >> 4;
  const int64_t n = ncol * nlev;
  const int64_t nb = gridDim.x;
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t g0 = vb * X2_C;  // within the pass [col0, col0 + nc)
  if (g0 >= nc) return;
  const int ncl = (int)(nc - g0 < X2_C ? nc - g0 : X2_C);
  unsigned store = (1u << ncl) - 1u;
  if (only) {
    unsigned m = 0u;
    for (int cc = 0; cc < ncl; ++cc) m |= only[g0 + cc] != 0 ? 1u << cc : 0u;
    store = m;
    if (store == 0u) return;  // block-uniform
  }
  const bool own = (store >> c) & 1u;
  double tb[NS][NT];
  {
    const double* Tg = T + (size_t)(g0 + (own ? c : 0)) * K * K;
    double cs[NT] = {0.0, 0.0};
#pragma unroll
    for (int s4 = 0; s4 < NS; ++s4)
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        tb[s4][tt] = own ? Tg[(size_t)(4 * s4 + kk) * K + 16 * (2 * tp + tt) + i16] : 0.0;
        cs[tt] += tb[s4][tt];
      }
    const double rk = 1.0 / (double)K;
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) {
      cs[tt] += __shfl_xor(cs[tt], 16);
      cs[tt] += __shfl_xor(cs[tt], 32);
      const double dc = 1.0 - cs[tt];
#pragma unroll
      for (int s4 = 0; s4 < NS; ++s4) tb[s4][tt] = fma(dc, rk, tb[s4][tt]);
    }
  }
  double pv[NV];
  auto load_chunk = [&](int64_t l0) {
    const int nll = (int)(nlev - l0 < XM_L ? nlev - l0 : XM_L);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int idx = tid + 512 * q;
      const int cc = idx % X2_C, l = (idx / X2_C) % XM_L, m = idx / (X2_C * XM_L);
      double v = 0.0;
      if (cc < ncl && l < nll) v = X[(size_t)m * n + (l0 + l) * ncol + col0 + g0 + cc];
      pv[q] = v;
    }
  };
  load_chunk(0);
  for (int64_t l0 = 0; l0 < nlev; l0 += XM_L) {
    const int nll = (int)(nlev - l0 < XM_L ? nlev - l0 : XM_L);
    __syncthreads();  // the previous chunk's output has left xs
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int idx = tid + 512 * q;
      const int cc = idx % X2_C, l = (idx / X2_C) % XM_L, m = idx / (X2_C * XM_L);
      xs[cc * CSI + m * XM_L + l] = pv[q];
    }
    __syncthreads();
    if (l0 + XM_L < nlev) load_chunk(l0 + XM_L);
    d4 acc[NT];
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) acc[tt] = d4{0.0, 0.0, 0.0, 0.0};
    const double* xa = xs + c * CSI + kk * XM_L + i16;  // A[i16][kk] of step 0
#pragma unroll
    for (int s4 = 0; s4 < NS; ++s4) {
      const double av = xa[4 * s4 * XM_L];
#pragma unroll
      for (int tt = 0; tt < NT; ++tt)
        acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, tb[s4][tt], acc[tt], 0, 0, 0);
    }
    __syncthreads();  // every wavefront has read its X
    if (own) {
      bool fin = true;
#pragma unroll
      for (int tt = 0; tt < NT; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r) fin = fin && isfinite(acc[tt][r]);
      if (__ballot(!fin) != 0ull && lane == 0) atomicOr(status, MDA_ST_NONFINITE_STATE);
    }
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        xs[c * CSO + (kk + 4 * r) * LJ + 16 * (2 * tp + tt) + i16] = acc[tt][r];
    __syncthreads();
    // element q of thread tid: column idx % 2, member 32 (w % 4) + idx / 2 % 32,
    // level w / 4, with idx = tid + 512 q and w = idx / 64
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int idx = tid + 512 * q;
      const int cc = idx % X2_C, w = idx / 64;
      const int j = 32 * (w % (K / 32)) + (idx / X2_C) % 32, l = w / (K / 32);
      if (cc < ncl && l < nll && ((store >> cc) & 1u))
        X[(size_t)j * n + (l0 + l) * ncol + col0 + g0 + cc] = xs[cc * CSO + l * LJ + j];
    }
  }
}

// 4'''. X_a = xbar + X' T with split-precision products, the transform of
// the BF16x3 form: X' = X - xbar in FP64 (xbar over the members per level),
// X' = hi + lo and T = hi + lo in bf16 (round to nearest even; lo from the
// FP64 residual), X' T ~ hi hi + hi lo + lo hi on v_mfma_f32_16x16x32_bf16
// with FP32 accumulation (A[l%16][8(l/16)+e], B[8(l/16)+e][l%16],
// D[4(l/16)+r][l%16]), xbar added back in FP64.  Block shape as in the FP64
// form: XM_C columns per block, one wavefront each, 16-level chunks staged
// through LDS.  Members are padded to a multiple of 32 with zeros.
__device__ inline void split_bf16(const double (&v)[8], bf8& hi, bf8& lo) {
  u32x4 ph, pl;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const fl2 f = {(float)v[2 * q], (float)v[2 * q + 1]};
    const bf2 h = __builtin_convertvector(f, bf2);
    const fl2 hf = __builtin_convertvector(h, fl2);
    const fl2 r = {(float)(v[2 * q] - (double)hf[0]), (float)(v[2 * q + 1] - (double)hf[1])};
    const bf2 lw = __builtin_convertvector(r, bf2);
    ph[q] = __builtin_bit_cast(uint32_t, h);
    pl[q] = __builtin_bit_cast(uint32_t, lw);
  }
  hi = __builtin_bit_cast(bf8, ph);
  lo = __builtin_bit_cast(bf8, pl);
}

template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_letkf_transform_bf16x3(double* __restrict__ X, int64_t ncol, int64_t nlev,
                              const double* __restrict__ T, int64_t col0, int64_t nc,
                              int* __restrict__ status) {
  static_assert(K % 16 == 0 && K <= 64, "K must be a multiple of 16, at most 64");
  if (*status & MDA_ST_SKIP_UPDATE) return;
  constexpr int KS = (K + 31) / 32, NTJ = K / 16;
  __shared__ double xs[XM_C * K * XM_L];  // [column][member or j][level]
  __shared__ double xb[XM_C][XM_L];       // xbar of the chunk's levels, per column
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int64_t n = ncol * nlev;
  const int64_t nb = gridDim.x;
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t g0 = vb * XM_C;
  if (g0 >= nc) return;
  const int ncl = (int)(nc - g0 < XM_C ? nc - g0 : XM_C);
  const bool own = wv < ncl;
  // T of column g0 + wv, split: B[8 kq + e][i16] of k-step s4, tile tj =
  // T[32 s4 + 8 kq + e][16 tj + i16]
  bf8 th[KS][NTJ], tl[KS][NTJ];
  {
    const double* Tg = T + (size_t)(g0 + (own ? wv : 0)) * K * K;
#pragma unroll
    for (int s4 = 0; s4 < KS; ++s4)
#pragma unroll
      for (int tj = 0; tj < NTJ; ++tj) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int m = 32 * s4 + 8 * kq + e;
          v[e] = (own && m < K) ? Tg[(size_t)m * K + 16 * tj + i16] : 0.0;
        }
        split_bf16(v, th[s4][tj], tl[s4][tj]);
      }
  }
  constexpr int NV = XM_C * XM_L * K / 256;
  double pv[NV];
  auto load_chunk = [&](int64_t l0, int nll) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int idx = tid + 256 * q;
      const int c = idx % XM_C, l = (idx / XM_C) % XM_L, m = idx / (XM_C * XM_L);
      double v = 0.0;
      if (c < ncl && l < nll) v = X[(size_t)m * n + (l0 + l) * ncol + col0 + g0 + c];
      pv[q] = v;
    }
  };
  load_chunk(0, (int)(nlev < XM_L ? nlev : XM_L));
  for (int64_t l0 = 0; l0 < nlev; l0 += XM_L) {
    const int nll = (int)(nlev - l0 < XM_L ? nlev - l0 : XM_L);
    __syncthreads();  // the previous chunk's output has left xs
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int idx = tid + 256 * q;
      const int c = idx % XM_C, l = (idx / XM_C) % XM_L, m = idx / (XM_C * XM_L);
      xs[(c * K + m) * XM_L + l] = pv[q];
    }
    __syncthreads();
    if (l0 + XM_L < nlev)
      load_chunk(l0 + XM_L, (int)(nlev - l0 - XM_L < XM_L ? nlev - l0 - XM_L : XM_L));
    // xbar of level i16: lane (i16, kq) sums members kq K/4 .. (kq + 1) K/4 - 1,
    // then the four partial sums are combined
    const double* xc = xs + (size_t)wv * K * XM_L;
    double part = 0.0;
#pragma unroll
    for (int mm = 0; mm < K / 4; ++mm) part += xc[(kq * (K / 4) + mm) * XM_L + i16];
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    const double xbar = part * (1.0 / (double)K);
    if (kq == 0) xb[wv][i16] = xbar;
    f4 acc[NTJ];
#pragma unroll
    for (int tj = 0; tj < NTJ; ++tj) acc[tj] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s4 = 0; s4 < KS; ++s4) {
      double v[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int m = 32 * s4 + 8 * kq + e;
        v[e] = m < K ? xc[m * XM_L + i16] - xbar : 0.0;
      }
      bf8 ah, al;
      split_bf16(v, ah, al);
#pragma unroll
      for (int tj = 0; tj < NTJ; ++tj) {
        acc[tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, th[s4][tj], acc[tj], 0, 0, 0);
        acc[tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, tl[s4][tj], acc[tj], 0, 0, 0);
        acc[tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, th[s4][tj], acc[tj], 0, 0, 0);
      }
    }
    __syncthreads();  // every wavefront has read its X (and xb is complete)
    double out[NTJ][4];
    bool fin = true;
#pragma unroll
    for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        out[tj][r] = xb[wv][4 * kq + r] + (double)acc[tj][r];
        fin = fin && isfinite(out[tj][r]);
      }
    if (__ballot(!fin) != 0ull && lane == 0) atomicOr(status, MDA_ST_NONFINITE_STATE);
    // D[4 kq + r][i16] of tile tj = X_a[level 4 kq + r][j = 16 tj + i16] -> xs[c][j][l]
#pragma unroll
    for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) xs[(wv * K + 16 * tj + i16) * XM_L + 4 * kq + r] = out[tj][r];
    __syncthreads();
    for (int idx = tid; idx < XM_C * XM_L * K; idx += 256) {
      const int c = idx % XM_C, l = (idx / XM_C) % XM_L, j = idx / (XM_C * XM_L);
      if (c < ncl && l < nll) X[(size_t)j * n + (l0 + l) * ncol + col0 + g0 + c] = xs[(c * K + j) * XM_L + l];
    }
  }
}

bool letkf_transform_reg_supported(int k) {
  switch (k) {
    case 8: case 16: case 20: case 32: case 40: case 48: case 64: return true;
    default: return false;
  }
}

void launch_letkf_transform(hipStream_t s, double* X, int k, int64_t ncol, int64_t nlev,
                            const double* T, int* status, int64_t col0, int64_t nc,
                            int precision, const int* only) {
  MDA_REQUIRE(status != nullptr, "transform needs the analysis status word");
  MDA_REQUIRE(k <= 128, "transform supports k <= 128");
  if (nc < 0) nc = ncol;
  MDA_REQUIRE(col0 >= 0 && col0 + nc <= ncol, "transform column range out of bounds");
  if (nc == 0) return;
  MDA_REQUIRE(!only || (precision != MDA_PREC_BF16X3 && ((k % 16 == 0 && k <= 64) || k == 128)),
              "a column-restricted transform needs the FP64 MFMA path");
  if (k == 128 && precision != MDA_PREC_BF16X3) {
    MDA_REQUIRE((nc + X2_C - 1) / X2_C < (int64_t)1 << 31, "transform grid too large");
    const unsigned nb2 = (unsigned)(((nc + X2_C - 1) / X2_C + 7) / 8 * 8);
    hipLaunchKernelGGL(k_letkf_transform_mfma2<128>, dim3(nb2), dim3(512), 0, s, X, ncol, nlev, T,
                       col0, nc, status, only);
    MDA_LAUNCH_CHECK();
    return;
  }
  if (precision == MDA_PREC_BF16X3 && k % 16 == 0 && k <= 64) {
    MDA_REQUIRE((nc + XM_C - 1) / XM_C < (int64_t)1 << 31, "transform grid too large");
    const unsigned nbm = (unsigned)(((nc + XM_C - 1) / XM_C + 7) / 8 * 8);
    switch (k) {
      case 16: hipLaunchKernelGGL(k_letkf_transform_bf16x3<16>, dim3(nbm), dim3(256), 0, s, X, ncol, nlev, T, col0, nc, status); break;
      case 32: hipLaunchKernelGGL(k_letkf_transform_bf16x3<32>, dim3(nbm), dim3(256), 0, s, X, ncol, nlev, T, col0, nc, status); break;
      case 48: hipLaunchKernelGGL(k_letkf_transform_bf16x3<48>, dim3(nbm), dim3(256), 0, s, X, ncol, nlev, T, col0, nc, status); break;
      default: hipLaunchKernelGGL(k_letkf_transform_bf16x3<64>, dim3(nbm), dim3(256), 0, s, X, ncol, nlev, T, col0, nc, status); break;
    }
    MDA_LAUNCH_CHECK();
    return;
  }
  if (k % 16 == 0 && k <= 64) {
    MDA_REQUIRE((nc + XM_C - 1) / XM_C < (int64_t)1 << 31, "transform grid too large");
    const unsigned nbm = (unsigned)(((nc + XM_C - 1) / XM_C + 7) / 8 * 8);
    switch (k) {
      case 16: hipLaunchKernelGGL(k_letkf_transform_mfma<16>, dim3(nbm), dim3(256), 0, s, X, ncol, nlev, T, col0, nc, status, only); break;
      case 32: hipLaunchKernelGGL(k_letkf_transform_mfma<32>, dim3(nbm), dim3(256), 0, s, X, ncol, nlev, T, col0, nc, status, only); break;
      case 48: hipLaunchKernelGGL(k_letkf_transform_mfma<48>, dim3(nbm), dim3(256), 0, s, X, ncol, nlev, T, col0, nc, status, only); break;
      default: hipLaunchKernelGGL(k_letkf_transform_mfma<64>, dim3(nbm), dim3(256), 0, s, X, ncol, nlev, T, col0, nc, status, only); break;
    }
    MDA_LAUNCH_CHECK();
    return;
  }
  if (letkf_transform_reg_supported(k)) {
    MDA_REQUIRE((nc + XT_C - 1) / XT_C < (int64_t)1 << 31, "transform grid too large");
    const unsigned nbr = (unsigned)(((nc + XT_C - 1) / XT_C + 7) / 8 * 8);
    const size_t shr = sizeof(double) * XT_L * XT_C * (k + 2);
#define MDA_XT(KK)                                                                          \
  case KK:                                                                                 \
    hipLaunchKernelGGL(k_letkf_transform_reg<KK>, dim3(nbr), dim3(256), shr, s, X, ncol, nlev, T, \
                       col0, nc, status);                                                    \
    break;
    switch (k) { MDA_XT(8) MDA_XT(16) MDA_XT(20) MDA_XT(32) MDA_XT(40) MDA_XT(48) MDA_XT(64) }
#undef MDA_XT
    MDA_LAUNCH_CHECK();
    return;
  }
  size_t shm = sizeof(double) * TR_L * TR_C * (k + 1);
  const int64_t tiles = ((nc + TR_C - 1) / TR_C) * ((nlev + TR_L - 1) / TR_L);
  MDA_REQUIRE(tiles < (int64_t)1 << 31, "transform grid too large");
  const unsigned nb = (unsigned)((tiles + 7) / 8 * 8);
  hipLaunchKernelGGL(k_letkf_transform, dim3(nb), dim3(256), shm, s, X, k, ncol, nlev, T, col0,
                     nc, status);
  MDA_LAUNCH_CHECK();
}

// nloc statistics: out[0] = sum, out[1] = max, out[2] = #zeros
__global__ void k_reduce_int(const int* __restrict__ v, int64_t n, int64_t* out) {
  __shared__ long long s_sum[256], s_max[256], s_zero[256];
  long long su = 0, mx = 0, ze = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    long long x = v[i];
    su += x; mx = x > mx ? x : mx; ze += (x == 0);
  }
  s_sum[threadIdx.x] = su; s_max[threadIdx.x] = mx; s_zero[threadIdx.x] = ze;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
      s_max[threadIdx.x] = s_max[threadIdx.x] > s_max[threadIdx.x + s] ? s_max[threadIdx.x]
                                                                       : s_max[threadIdx.x + s];
      s_zero[threadIdx.x] += s_zero[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long*)&out[0], (unsigned long long)s_sum[0]);
    atomicMax((long long*)&out[1], s_max[0]);
    atomicAdd((unsigned long long*)&out[2], (unsigned long long)s_zero[0]);
  }
}

void launch_reduce_int(hipStream_t s, const int* v, int64_t n, int64_t* out3,
                       int64_t* /*unused*/) {
  hipLaunchKernelGGL(k_reduce_int, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, v, n, out3);
  MDA_LAUNCH_CHECK();
}

}  // namespace mda
