// IdentityObsOperator on the device (IdentityObsOperator.hpp:154-281,
// 483-676): the four-point inverse-distance stencil of every observation as
// a CSR row over state linear indices, so that H x is the library's CSR
// apply.
//
//   GRID obs        (x, y, z) grid coordinates, clamped to the grid
//                   (find4NearestGridPoints, :594-638)
//   GEOGRAPHIC obs  first mapped to a grid point by convertGeographicToGrid
//                   (:483-587): with column coordinates, the nearest column
//                   in (lat, lon) degrees (the first one in j-major, i-minor
//                   order on ties) and the nearest level of vcoord (level 0
//                   without it); with a geometry that holds no geographic
//                   coordinates (SimpleGeometry), (0, 0, 0), as the
//                   reference's fallback search leaves it
//   invalid obs     (variance +inf) an empty row: H x = 0 (:164-167)
// Weights w = 1/dist, 1e12 at dist = 0 (:656-658), normalised by their sum
// (the reference divides sum(w x) by sum(w)).  Distances in the reference's
// operation order without FMA contraction.
#include "mda_device.hpp"
#include "mda_internal.hpp"

namespace mda {

__device__ inline double plain_sqrt2(double a, double b) {
#pragma clang fp contract(off)
  return sqrt(a * a + b * b);
}

__device__ inline double plain_sqrt3(double a, double b, double c) {
#pragma clang fp contract(off)
  return sqrt(a * a + b * b + c * c);
}

// find4NearestGridPoints + IDW weights at (x, y, z); writes 4 taps
__device__ inline void idw4_taps(int64_t nx, int64_t ny, int64_t nz, double x, double y,
                                 double z, int64_t* col, double* w) {
  x = fmax(0.0, fmin((double)(nx - 1), x));
  y = fmax(0.0, fmin((double)(ny - 1), y));
  z = fmax(0.0, fmin((double)(nz - 1), z));
  const int64_t i0 = (int64_t)floor(x), j0 = (int64_t)floor(y), k0 = (int64_t)floor(z);
  const int64_t i1 = i0 + 1 < nx - 1 ? i0 + 1 : nx - 1;
  const int64_t j1 = j0 + 1 < ny - 1 ? j0 + 1 : ny - 1;
  const int64_t k1 = k0 + 1 < nz - 1 ? k0 + 1 : nz - 1;
  double d[4];
  if (nz == 1) {
    const int64_t ii[4] = {i0, i1, i0, i1}, jj[4] = {j0, j0, j1, j1};
    for (int t = 0; t < 4; ++t) {
      d[t] = plain_sqrt2(x - (double)ii[t], y - (double)jj[t]);
      col[t] = jj[t] * nx + ii[t];
    }
  } else {
    // the 8 corners in the reference's order, the 4 closest by a stable
    // insertion sort (what std::sort does on 8 elements)
    const int64_t ci[8] = {i0, i1, i0, i1, i0, i1, i0, i1};
    const int64_t cj[8] = {j0, j0, j1, j1, j0, j0, j1, j1};
    const int64_t ck[8] = {k0, k0, k0, k0, k1, k1, k1, k1};
    double cd[8];
    int order[8];
    for (int t = 0; t < 8; ++t) {
      cd[t] = plain_sqrt3(x - (double)ci[t], y - (double)cj[t], z - (double)ck[t]);
      order[t] = t;
    }
    for (int a = 1; a < 8; ++a) {
      const int v = order[a];
      int b = a - 1;
      while (b >= 0 && cd[order[b]] > cd[v]) {
        order[b + 1] = order[b];
        --b;
      }
      order[b + 1] = v;
    }
    for (int t = 0; t < 4; ++t) {
      const int c = order[t];
      d[t] = cd[c];
      col[t] = ck[c] * nx * ny + cj[c] * nx + ci[c];
    }
  }
  double ws = 0.0;
  for (int t = 0; t < 4; ++t) {
    w[t] = d[t] == 0.0 ? 1e12 : 1.0 / d[t];
    ws += w[t];
  }
  for (int t = 0; t < 4; ++t) w[t] /= ws;
}

// One wavefront per observation (the nearest-column search of a GEOGRAPHIC
// observation is a wave-wide argmin over the columns).
__global__ __launch_bounds__(256) void k_idw4_stencil(
    int64_t p, int coord, const double* __restrict__ oa, const double* __restrict__ ob,
    const double* __restrict__ oc, const double* __restrict__ var, int64_t nx, int64_t ny,
    int64_t nlev, const double* __restrict__ clat, const double* __restrict__ clon,
    const double* __restrict__ vcoord, const int64_t* __restrict__ row_ptr,
    int64_t* __restrict__ col, double* __restrict__ weight) {
  const int lane = threadIdx.x & 63;
  const int64_t ncol = nx * ny;
  for (int64_t o = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); o < p;
       o += (int64_t)gridDim.x * 4) {
    if (isinf(var[o])) continue;  // invalid: empty row
    double x, y, z;
    if (coord == MDA_COORD_GEOGRAPHIC) {
      int64_t best = 0, bestk = 0;
      if (clat) {
        const double lat = oa[o], lon = ob[o];
        double bd = INFINITY;
        int64_t bi = ncol;
        for (int64_t g = lane; g < ncol; g += 64) {
          const double dist = plain_sqrt2(lon - clon[g], lat - clat[g]);
          if (dist < bd) { bd = dist; bi = g; }  // lanes visit g in increasing order
        }
        for (int off = 32; off > 0; off >>= 1) {
          const double od = __shfl_xor(bd, off);
          const int64_t oi = __shfl_xor(bi, off);
          if (od < bd || (od == bd && oi < bi)) { bd = od; bi = oi; }
        }
        best = bi < ncol ? bi : 0;
        if (vcoord) {
          const double lev = oc ? oc[o] : 0.0;
          double bv = INFINITY;
          for (int64_t kk = 0; kk < nlev; ++kk) {
            const double dv = fabs(lev - vcoord[kk]);
            if (dv < bv) { bv = dv; bestk = kk; }
          }
        }
      }
      x = (double)(best % nx);
      y = (double)(best / nx);
      z = (double)bestk;
    } else {
      x = oa[o];
      y = ob[o];
      z = oc ? oc[o] : 0.0;
    }
    if (lane == 0) {
      int64_t c4[4];
      double w4[4];
      idw4_taps(nx, ny, nlev, x, y, z, c4, w4);
      const int64_t r0 = row_ptr[o];
      for (int t = 0; t < 4; ++t) {
        col[r0 + t] = c4[t];
        weight[r0 + t] = w4[t];
      }
    }
  }
}

void launch_idw4_stencil(hipStream_t s, int64_t p, int coord, const double* oa, const double* ob,
                         const double* oc, const double* var, int64_t nx, int64_t ny,
                         int64_t nlev, const double* clat, const double* clon,
                         const double* vcoord, const int64_t* row_ptr, int64_t* col,
                         double* weight) {
  if (p <= 0) return;
  hipLaunchKernelGGL(k_idw4_stencil, dim3(grid_for(p, 4, 65536)), dim3(256), 0, s, p, coord, oa,
                     ob, oc, var, nx, ny, nlev, clat, clon, vcoord, row_ptr, col, weight);
  MDA_LAUNCH_CHECK();
}

}  // namespace mda
