// X_a = X T' on FP64 MFMA for blocks of XM_C = 4 consecutive columns (one
// wavefront per column), shared by the stand-alone transform kernel and the
// eigenvector kernel that applies its own transforms.  Both go through the
// functions below, so a column gets bit-identical X_a either way.
//
// With T'[m][j] = T[m][j] + (1 - sum_m T[m][j]) / K,
//   X_a[lev][j] = sum_m X[lev][m] T'[m][j]
// (xbar = sum_m X[lev][m] / K), i.e. X_a = xbar + X' T = X T'.  The wavefront
// keeps T' of its column in registers in the MFMA B layout (tb[s][tj] =
// T'[4s + kk][16 tj + i16], lane = i16 + 16 kk) and runs
// v_mfma_f64_16x16x4f64 over 16-level chunks.  X is staged through LDS as
// [column][member][16 levels] from 32-byte column segments (coalesced across
// the block's columns), and X_a leaves through LDS as [column][level][member].
// Strides in doubles (the LDS serves a 64-bit access from 32 double-wide
// bank pairs, half a wavefront at a time): a column's block starts 8 pairs
// further on (xm_csi, xm_cso), and an output level row is K + 16 long when K
// is a multiple of 32 (xm_lj), so in the staging writes (lanes over 4
// columns and 8 levels per half), the product reads (16 levels x 2
// members), the accumulator write-out (16 members x 2 levels) and the
// store-side reads (4 columns x 8 members) the 32 lanes of each half fall on
// 32 different pairs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mda_internal.hpp"

namespace mda {

constexpr int XM_C = 4, XM_L = 16;

template <int K>
constexpr int xm_lj() { return K % 32 == 0 ? K + 16 : K; }
template <int K>
constexpr int xm_csi() { return XM_L * K + 8; }
template <int K>
constexpr int xm_cso() { return XM_L * xm_lj<K>() + 8; }
// LDS doubles of xm_apply's staging buffer
template <int K>
constexpr int xm_xs_size() {
  return XM_C * (xm_csi<K>() > xm_cso<K>() ? xm_csi<K>() : xm_cso<K>());
}

typedef double xm_d4 __attribute__((ext_vector_type(4)));

// T -> T' in place: column sums over the member index (each lane's K/4
// entries in order, then the four kk groups by xor-shuffles) and the
// correction (1 - colsum) / K.
template <int K>
__device__ __forceinline__ void xm_tprime(double (&tb)[K / 4][K / 16]) {
  constexpr int NTJ = K / 16, NS = K / 4;
  double cs[NTJ];
#pragma unroll
  for (int tj = 0; tj < NTJ; ++tj) cs[tj] = 0.0;
#pragma unroll
  for (int s4 = 0; s4 < NS; ++s4)
#pragma unroll
    for (int tj = 0; tj < NTJ; ++tj) cs[tj] += tb[s4][tj];
  const double rk = 1.0 / (double)K;
#pragma unroll
  for (int tj = 0; tj < NTJ; ++tj) {
    cs[tj] += __shfl_xor(cs[tj], 16);
    cs[tj] += __shfl_xor(cs[tj], 32);
    // one explicit rounding, so that every kernel using this rounds alike
    const double dc = 1.0 - cs[tj];
#pragma unroll
    for (int s4 = 0; s4 < NS; ++s4) tb[s4][tj] = fma(dc, rk, tb[s4][tj]);
  }
}

// Doubles of one 16-level chunk of the block's 4 columns per thread.
template <int K>
constexpr int xm_nv() { return XM_C * XM_L * K / 256; }

// Loads chunk c (levels 16c .. 16c + 15 of X[m][level][colbase + col]) into
// v: element q of thread tid is (col, level, member) = (idx % 4, idx / 4 %
// 16, idx / 64) with idx = tid + 256 q; zeros outside the columns and levels.
template <int K>
__device__ __forceinline__ void xm_load(double (&v)[xm_nv<K>()], const double* __restrict__ X,
                                        int64_t ncol, int64_t nlev, int64_t colbase, int ncl,
                                        int64_t l0) {
  const int tid = threadIdx.x;
  const int64_t n = ncol * nlev;
  const int nll = (int)(nlev - l0 < XM_L ? (nlev - l0 > 0 ? nlev - l0 : 0) : XM_L);
#pragma unroll
  for (int q = 0; q < xm_nv<K>(); ++q) {
    const int idx = tid + 256 * q;
    const int c = idx % XM_C, l = (idx / XM_C) % XM_L, m = idx / (XM_C * XM_L);
    double v0 = 0.0;
    if (c < ncl && l < nll) v0 = X[(size_t)m * n + (l0 + l) * ncol + colbase + c];
    v[q] = v0;
  }
}

// X_a = X T' for the columns colbase + c, c < ncl, of X (ncol columns of
// nlev levels, member-major), written for the columns whose bit is set in
// `store`.  Block of 256 threads (wavefront wv: column wv); xs holds
// xm_xs_size<K>() doubles of LDS.  Every thread of the block calls this (it
// synchronises the block).  The first NPC chunks come from `pre` (loaded by
// xm_load earlier, so their latency is hidden behind other work); later
// chunks are loaded one ahead of their products.  A non-finite X_a sets
// MDA_ST_NONFINITE_STATE.
template <int K, int NPC = 0>
__device__ __forceinline__ void xm_apply(double* __restrict__ X, int64_t ncol, int64_t nlev,
                                         int64_t colbase, int ncl, unsigned store,
                                         const double (&tb)[K / 4][K / 16], double* xs,
                                         int* __restrict__ status,
                                         const double (*pre)[xm_nv<K>()] = nullptr) {
  static_assert(K % 16 == 0 && K <= 64, "K must be a multiple of 16, at most 64");
  constexpr int NTJ = K / 16, NS = K / 4, NV = xm_nv<K>();
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i16 = lane & 15, kk = lane >> 4;
  constexpr int LJ = xm_lj<K>(), CSI = xm_csi<K>(), CSO = xm_cso<K>();
  const int64_t n = ncol * nlev;
  double pv[NV];
  // one chunk: src staged through xs (src may be pv: it is read before the
  // next chunk's loads overwrite pv), products, X_a out through xs
  auto chunk = [&](const double (&src)[NV], int64_t l0, bool load_next) {
    const int nll = (int)(nlev - l0 < XM_L ? nlev - l0 : XM_L);
    __syncthreads();  // the previous chunk's output has left xs
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int idx = tid + 256 * q;
      const int c = idx % XM_C, l = (idx / XM_C) % XM_L, m = idx / (XM_C * XM_L);
      xs[c * CSI + m * XM_L + l] = src[q];
    }
    __syncthreads();
    if (load_next) xm_load<K>(pv, X, ncol, nlev, colbase, ncl, l0 + XM_L);
    xm_d4 acc[NTJ];
#pragma unroll
    for (int tj = 0; tj < NTJ; ++tj) acc[tj] = xm_d4{0.0, 0.0, 0.0, 0.0};
    const double* xa = xs + wv * CSI + kk * XM_L + i16;  // A[i16][kk] of step 0
#pragma unroll
    for (int s4 = 0; s4 < NS; ++s4) {
      const double av = xa[4 * s4 * XM_L];
#pragma unroll
      for (int tj = 0; tj < NTJ; ++tj)
        acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, tb[s4][tj], acc[tj], 0, 0, 0);
    }
    __syncthreads();  // every wavefront has read its X
    if ((store >> wv) & 1u) {
      bool fin = true;
#pragma unroll
      for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) fin = fin && isfinite(acc[tj][r]);
      if (__ballot(!fin) != 0ull && lane == 0) atomicOr(status, MDA_ST_NONFINITE_STATE);
    }
    // D[kk + 4r][i16] of tile tj = X_a[level kk + 4r][j = 16 tj + i16] -> xs[c][l][j]
#pragma unroll
    for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) xs[wv * CSO + (kk + 4 * r) * LJ + 16 * tj + i16] = acc[tj][r];
    __syncthreads();
    // element q of thread tid: column idx % 4, member 16 (w % (K/16)) + idx / 4 % 16,
    // level w / (K/16), with idx = tid + 256 q and w = idx / 64
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const int idx = tid + 256 * q;
      const int c = idx % XM_C, w = idx / 64;
      const int j = 16 * (w % (K / 16)) + (idx / XM_C) % 16, l = w / (K / 16);
      if (c < ncl && l < nll && ((store >> c) & 1u))
        X[(size_t)j * n + (l0 + l) * ncol + colbase + c] = xs[c * CSO + l * LJ + j];
    }
  };
#pragma unroll
  for (int ci = 0; ci < NPC; ++ci) {
    const int64_t l0 = (int64_t)XM_L * ci;
    if (l0 >= nlev) return;  // block-uniform
    chunk(pre[ci], l0, ci == NPC - 1 && l0 + XM_L < nlev);
  }
  if (NPC == 0) xm_load<K>(pv, X, ncol, nlev, colbase, ncl, 0);
  for (int64_t l0 = (int64_t)XM_L * NPC; l0 < nlev; l0 += XM_L) chunk(pv, l0, l0 + XM_L < nlev);
}

}  // namespace mda
