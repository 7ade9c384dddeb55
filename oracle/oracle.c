/*
 * oracle.c -- CPU FP64 restatement of metada's ensemble Kalman analysis path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker that the HIP library is held
 * against.  It is loaded by tests/, by __graft_entry__.smoke() and by the CPU
 * baseline leg of bench.py, never by the product path (metada_amd/).
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the open-x-da/metada source tree).  How each is pinned (tests/):
 *   orc_cholesky_lower, orc_lu_inverse, orc_jacobi_eigh, orc_singular_values
 *       -> numpy/scipy fixtures (tests/golden) + reconstruction invariants
 *   orc_haversine_km      -> quarter meridian = pi*R/2, antipodes, symmetry
 *   orc_grid_distance, orc_cart_distance
 *                         -> Pythagorean KATs (3-4-5, 1-2-2, 3-4-12), symmetry
 *   orc_gaspari_cohn      -> GC(0)=1, GC(2)=0, continuity at 1, closed-form
 *                            integral check
 *   orc_idw4 / stencil    -> brute-force literal formula, exact-hit KAT
 *                            (2e12 x_i + 2 x_{i+1})/(2e12+2) on 1-row grids
 *   orc_enkf_analyse      -> p=1,k=2 closed form; R->inf; Woodbury identity vs
 *                            the ETKF mean weights; cond(S) of a diagonal S
 *   orc_etkf_analyse      -> Wa Wa^T = (k-1) Pa; mean update vs Kalman gain
 *   orc_letkf_standard    -> W^2 = (k-1) A^-1, W = W^T, W 1 = sqrt(rho) 1,
 *                            empty-column closed form, infinite-radius limit
 *                            equals a global ETKF (symmetric form); one
 *                            observation at d = r/2, r and beyond r against
 *                            the Sherman-Morrison closed form of the mean
 *                            update (pins the taper scale c = r/2)
 *   orc_letkf_refcompat_* -> at rho=1 and infinite radius equals ETKF compat
 *   orc_metrics           -> MetricsTest.cpp:64-195 known answers
 *   orc_l96_*             -> stationary point x=F, a hand-computed 5-variable
 *                            tendency, energy conservation of the advection
 *                            term (sum x_i a_i = 0), RK4 order check
 *   orc_mlp_*             -> central finite differences
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_E_ARG (-1)
#define ORC_E_NOTSPD (-2)
#define ORC_E_SINGULAR (-3)
#define ORC_E_NOMEM (-4)

/* ======================================================================== */
/* small dense linear algebra                                                 */
/* ======================================================================== */

/* A = L L^T, L lower triangular (the factor Eigen's LLT::matrixL() returns,
 * used at LETKF.hpp:223, ETKF.hpp:158, EnKF.hpp:359). Row-major. */
int orc_cholesky_lower(int n, const double* A, double* L) {
  if (n <= 0) return ORC_E_ARG;
  memset(L, 0, sizeof(double) * (size_t)n * n);
  for (int j = 0; j < n; ++j) {
    double s = A[(size_t)j * n + j];
    for (int t = 0; t < j; ++t) s -= L[(size_t)j * n + t] * L[(size_t)j * n + t];
    if (!(s > 0.0)) return ORC_E_NOTSPD;
    double d = sqrt(s);
    L[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double v = A[(size_t)i * n + j];
      for (int t = 0; t < j; ++t) v -= L[(size_t)i * n + t] * L[(size_t)j * n + t];
      L[(size_t)i * n + j] = v / d;
    }
  }
  return 0;
}

/* Inverse by LU with partial (row) pivoting, the algorithm behind Eigen's
 * MatrixXd::inverse() for sizes > 4 (called at EnKF.hpp:199, ETKF.hpp:150-152,
 * LETKF.hpp:215-218). */
int orc_lu_inverse(int n, const double* A, double* Ainv) {
  if (n <= 0) return ORC_E_ARG;
  double* LU = (double*)malloc(sizeof(double) * (size_t)n * n);
  int* piv = (int*)malloc(sizeof(int) * (size_t)n);
  if (!LU || !piv) { free(LU); free(piv); return ORC_E_NOMEM; }
  memcpy(LU, A, sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; ++i) piv[i] = i;
  int rc = 0;
  for (int c = 0; c < n; ++c) {
    int pr = c;
    double best = fabs(LU[(size_t)c * n + c]);
    for (int r = c + 1; r < n; ++r) {
      double v = fabs(LU[(size_t)r * n + c]);
      if (v > best) { best = v; pr = r; }
    }
    if (best == 0.0) { rc = ORC_E_SINGULAR; break; }
    if (pr != c) {
      for (int t = 0; t < n; ++t) {
        double tmp = LU[(size_t)c * n + t];
        LU[(size_t)c * n + t] = LU[(size_t)pr * n + t];
        LU[(size_t)pr * n + t] = tmp;
      }
      int tp = piv[c]; piv[c] = piv[pr]; piv[pr] = tp;
    }
    double inv = 1.0 / LU[(size_t)c * n + c];
    for (int r = c + 1; r < n; ++r) {
      double f = LU[(size_t)r * n + c] * inv;
      LU[(size_t)r * n + c] = f;
      for (int t = c + 1; t < n; ++t) LU[(size_t)r * n + t] -= f * LU[(size_t)c * n + t];
    }
  }
  if (rc == 0) {
    double* col = (double*)malloc(sizeof(double) * (size_t)n);
    for (int j = 0; j < n; ++j) {
      /* solve P A x = e_j: forward with unit L on permuted rhs, then U */
      for (int i = 0; i < n; ++i) col[i] = (piv[i] == j) ? 1.0 : 0.0;
      for (int i = 0; i < n; ++i) {
        double s = col[i];
        for (int t = 0; t < i; ++t) s -= LU[(size_t)i * n + t] * col[t];
        col[i] = s;
      }
      for (int i = n - 1; i >= 0; --i) {
        double s = col[i];
        for (int t = i + 1; t < n; ++t) s -= LU[(size_t)i * n + t] * col[t];
        col[i] = s / LU[(size_t)i * n + i];
      }
      for (int i = 0; i < n; ++i) Ainv[(size_t)i * n + j] = col[i];
    }
    free(col);
  }
  free(LU);
  free(piv);
  return rc;
}

/* Cyclic (row-by-row) Jacobi eigensolver for a symmetric matrix; eigenvalues
 * ascending, eigenvectors in the columns of V (row-major).  The rotation is
 * the textbook symmetric Schur decomposition (Golub & Van Loan, Matrix
 * Computations, sec. 8.5.2).  It is the standard-form replacement for the
 * Cholesky square root at LETKF.hpp:223-224. */
int orc_jacobi_eigh(int n, const double* Ain, double* evals, double* V, int max_sweeps,
                    int* sweeps_out) {
  if (n <= 0) return ORC_E_ARG;
  if (max_sweeps <= 0) max_sweeps = 100;
  double* A = (double*)malloc(sizeof(double) * (size_t)n * n);
  if (!A) return ORC_E_NOMEM;
  memcpy(A, Ain, sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[(size_t)i * n + j] = (i == j) ? 1.0 : 0.0;
  double fro = 0.0;
  for (size_t t = 0; t < (size_t)n * n; ++t) fro += A[t] * A[t];
  fro = sqrt(fro);
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        if (i != j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    if (sqrt(off) <= 1e-16 * fro || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int r = 0; r < n; ++r) { /* A <- A J (columns p, q) */
          double arp = A[(size_t)r * n + p], arq = A[(size_t)r * n + q];
          A[(size_t)r * n + p] = c * arp - s * arq;
          A[(size_t)r * n + q] = s * arp + c * arq;
        }
        for (int r = 0; r < n; ++r) { /* A <- J^T A (rows p, q) */
          double apr = A[(size_t)p * n + r], aqr = A[(size_t)q * n + r];
          A[(size_t)p * n + r] = c * apr - s * aqr;
          A[(size_t)q * n + r] = s * apr + c * aqr;
        }
        A[(size_t)p * n + q] = 0.0;
        A[(size_t)q * n + p] = 0.0;
        for (int r = 0; r < n; ++r) {
          double vrp = V[(size_t)r * n + p], vrq = V[(size_t)r * n + q];
          V[(size_t)r * n + p] = c * vrp - s * vrq;
          V[(size_t)r * n + q] = s * vrp + c * vrq;
        }
      }
    }
  }
  for (int i = 0; i < n; ++i) evals[i] = A[(size_t)i * n + i];
  /* selection sort ascending, permuting the columns of V */
  for (int i = 0; i < n - 1; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (evals[j] < evals[m]) m = j;
    if (m != i) {
      double t = evals[i]; evals[i] = evals[m]; evals[m] = t;
      for (int r = 0; r < n; ++r) {
        double tv = V[(size_t)r * n + i];
        V[(size_t)r * n + i] = V[(size_t)r * n + m];
        V[(size_t)r * n + m] = tv;
      }
    }
  }
  if (sweeps_out) *sweeps_out = sweep;
  free(A);
  return sweep < max_sweeps ? 0 : 1;
}

/* Singular values by one-sided (Hestenes) Jacobi on the columns of an m x n
 * row-major matrix, descending.  Replaces JacobiSVD::singularValues() at
 * EnKF.hpp:206-209 (two-sided there; the values agree to rounding). */
int orc_singular_values(int m, int n, const double* Ain, double* s) {
  if (m <= 0 || n <= 0) return ORC_E_ARG;
  double* G = (double*)malloc(sizeof(double) * (size_t)m * n);
  if (!G) return ORC_E_NOMEM;
  memcpy(G, Ain, sizeof(double) * (size_t)m * n);
  for (int sweep = 0; sweep < 100; ++sweep) {
    int rotated = 0;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double a = 0, b = 0, c = 0;
        for (int r = 0; r < m; ++r) {
          double gp = G[(size_t)r * n + p], gq = G[(size_t)r * n + q];
          a += gp * gp; b += gq * gq; c += gp * gq;
        }
        if (fabs(c) <= 1e-17 * sqrt(a * b) || c == 0.0) continue;
        rotated = 1;
        double zeta = (b - a) / (2.0 * c);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for (int r = 0; r < m; ++r) {
          double gp = G[(size_t)r * n + p], gq = G[(size_t)r * n + q];
          G[(size_t)r * n + p] = cs * gp - sn * gq;
          G[(size_t)r * n + q] = sn * gp + cs * gq;
        }
      }
    if (!rotated) break;
  }
  for (int j = 0; j < n; ++j) {
    double a = 0;
    for (int r = 0; r < m; ++r) a += G[(size_t)r * n + j] * G[(size_t)r * n + j];
    s[j] = sqrt(a);
  }
  for (int i = 0; i < n - 1; ++i)
    for (int j = i + 1; j < n; ++j)
      if (s[j] > s[i]) { double t = s[i]; s[i] = s[j]; s[j] = t; }
  free(G);
  return 0;
}

/* ======================================================================== */
/* locations and localisation                                                */
/* ======================================================================== */

/* Location::haversine, Location.hpp:349-357 (R = 6371 km, Location.hpp:325). */
double orc_haversine_km(double lat1, double lon1, double lat2, double lon2) {
  const double pi = 3.14159265358979323846;
  double dlat = (lat2 - lat1) * pi / 180.0;
  double dlon = (lon2 - lon1) * pi / 180.0;
  double a = sin(dlat / 2) * sin(dlat / 2) +
             cos(lat1 * pi / 180.0) * cos(lat2 * pi / 180.0) * sin(dlon / 2) * sin(dlon / 2);
  double c = 2 * atan2(sqrt(a), sqrt(1 - a));
  return 6371.0 * c;
}

/* Location::distance_to GRID branch, Location.hpp:205-211. */
double orc_grid_distance(double x1, double y1, double x2, double y2) {
  double dx = x1 - x2, dy = y1 - y2;
  return sqrt(dx * dx + dy * dy);
}

/* Location::distance_to CARTESIAN branch, Location.hpp:217-224. */
double orc_cart_distance(double x1, double y1, double z1, double x2, double y2, double z2) {
  double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

/* Gaspari & Cohn (1999), eq. (4.10): compactly supported fifth-order
 * piecewise rational function of z = d / c, zero for z >= 2.  (The reference
 * has no standard GC; the one in LWEnKF.hpp:627-636 is not this function.) */
double orc_gaspari_cohn(double z) {
  z = fabs(z);
  if (z >= 2.0) return 0.0;
  if (z <= 1.0)
    return (((-0.25 * z + 0.5) * z + 0.625) * z - 5.0 / 3.0) * z * z + 1.0;
  return ((((z / 12.0 - 0.5) * z + 0.625) * z + 5.0 / 3.0) * z - 5.0) * z + 4.0 -
         2.0 / (3.0 * z);
}

/* ======================================================================== */
/* observation operator                                                       */
/* ======================================================================== */

/* The four (i,j,k) neighbours find4NearestGridPoints picks, with their IDW
 * weights w = 1/dist (1e12 on an exact hit): IdentityObsOperator.hpp:594-638
 * (selection) and :643-676 (weights).  Returns the 4 linear indices
 * k*nx*ny + j*nx + i and the un-normalised weights. */
int orc_idw4_stencil(int64_t nx, int64_t ny, int64_t nz, double x, double y, double z,
                     int64_t* idx4, double* w4) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return ORC_E_ARG;
  x = fmax(0.0, fmin((double)(nx - 1), x));
  y = fmax(0.0, fmin((double)(ny - 1), y));
  z = fmax(0.0, fmin((double)(nz - 1), z));
  int64_t i0 = (int64_t)floor(x), j0 = (int64_t)floor(y), k0 = (int64_t)floor(z);
  int64_t i1 = i0 + 1 < nx - 1 ? i0 + 1 : nx - 1;
  int64_t j1 = j0 + 1 < ny - 1 ? j0 + 1 : ny - 1;
  int64_t k1 = k0 + 1 < nz - 1 ? k0 + 1 : nz - 1;
  if (nz == 1) {
    int64_t ii[4] = {i0, i1, i0, i1}, jj[4] = {j0, j0, j1, j1};
    for (int t = 0; t < 4; ++t) {
      double dx = x - (double)ii[t], dy = y - (double)jj[t];
      double d = sqrt(dx * dx + dy * dy);
      idx4[t] = jj[t] * nx + ii[t];
      w4[t] = (d == 0.0) ? 1e12 : 1.0 / d;
    }
    return 0;
  }
  /* 3-D: 8 corners ordered as in the reference, stable sort by distance,
   * keep the 4 closest (std::sort is not stable; ties are resolved here in
   * corner order, which is what std::sort does for these 8-element inputs
   * with distinct distances -- equal distances give equal weights either way
   * only when the tied corners hold the same value). */
  int64_t ci[8] = {i0, i1, i0, i1, i0, i1, i0, i1};
  int64_t cj[8] = {j0, j0, j1, j1, j0, j0, j1, j1};
  int64_t ck[8] = {k0, k0, k0, k0, k1, k1, k1, k1};
  double cd[8];
  int order[8];
  for (int t = 0; t < 8; ++t) {
    double dx = x - (double)ci[t], dy = y - (double)cj[t], dz = z - (double)ck[t];
    cd[t] = sqrt(dx * dx + dy * dy + dz * dz);
    order[t] = t;
  }
  for (int a = 1; a < 8; ++a) { /* insertion sort, stable */
    int v = order[a];
    int b = a - 1;
    while (b >= 0 && cd[order[b]] > cd[v]) { order[b + 1] = order[b]; --b; }
    order[b + 1] = v;
  }
  for (int t = 0; t < 4; ++t) {
    int c = order[t];
    idx4[t] = ck[c] * nx * ny + cj[c] * nx + ci[c];
    w4[t] = (cd[c] == 0.0) ? 1e12 : 1.0 / cd[c];
  }
  return 0;
}

/* idw4Interpolation, IdentityObsOperator.hpp:643-676: sum(w x) / sum(w). */
double orc_idw4(const double* state, int64_t nx, int64_t ny, int64_t nz, double x, double y,
                double z) {
  int64_t idx[4];
  double w[4];
  if (orc_idw4_stencil(nx, ny, nz, x, y, z, idx, w) != 0) return NAN;
  double ws = 0.0, sw = 0.0;
  for (int t = 0; t < 4; ++t) {
    ws += w[t] * state[idx[t]];
    sw += w[t];
  }
  return ws / sw;
}

/* y = H x for a linear H in CSR form (the device form of the stencils). */
int orc_csr_apply(int64_t p, const int64_t* row_ptr, const int64_t* col, const double* w,
                  const double* state, double* out) {
  for (int64_t o = 0; o < p; ++o) {
    double s = 0.0;
    for (int64_t t = row_ptr[o]; t < row_ptr[o + 1]; ++t) s += w[t] * state[col[t]];
    out[o] = s;
  }
  return 0;
}

/* ======================================================================== */
/* ensemble statistics                                                        */
/* ======================================================================== */

/* Ensemble::RecomputeMean, Ensemble.hpp:105-114: sequential sum over members,
 * then multiply by the reciprocal 1.0/k. */
void orc_ensemble_mean(int k, int64_t n, const double* X, double* mean) {
  double rk = 1.0 / (double)k;
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += X[(size_t)j * n + i];
    mean[i] = s * rk;
  }
}

/* Ensemble::RecomputePerturbations, Ensemble.hpp:142-151, times an optional
 * inflation scale (EnKF.hpp:319 uses sqrt(rho), ETKF.hpp:125 uses rho). */
void orc_ensemble_perturbations(int k, int64_t n, const double* X, const double* mean,
                                double scale, double* Xp) {
  for (int j = 0; j < k; ++j)
    for (int64_t i = 0; i < n; ++i)
      Xp[(size_t)j * n + i] = (X[(size_t)j * n + i] - mean[i]) * scale;
}

/* observation-space mean and anomalies, EnKF.hpp:175-181 / ETKF.hpp:135-141:
 * ybar = rowwise mean, Y' = Yb - ybar, d = yo - ybar.  Yp is [k][p]. */
static void obs_stats(int k, int64_t p, const double* Yb, const double* yo, double* ybar,
                      double* Yp, double* d) {
  for (int64_t o = 0; o < p; ++o) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += Yb[(size_t)j * p + o];
    ybar[o] = s / (double)k;
    d[o] = yo[o] - ybar[o];
    for (int j = 0; j < k; ++j) Yp[(size_t)j * p + o] = Yb[(size_t)j * p + o] - ybar[o];
  }
}

/* ======================================================================== */
/* EnKF                                                                       */
/* ======================================================================== */

/* EnKF::Analyse, EnKF.hpp:139-256, with the N(0,1) draws Z injected (they come
 * from std::random_device at EnKF.hpp:346-356).  Z is [p][k] row-major, the
 * (obs, member) fill order of EnKF.hpp:352-356.  All inflation methods are
 * multiplicative sqrt(rho) on X' (EnKF.hpp:316-331). */
int orc_enkf_analyse(int k, int64_t n, int64_t p, double* X, const double* Yb,
                     const double* yo, const double* var, const double* Z,
                     double inflation, orc_enkf_diag* diag) {
  if (k < 2 || n <= 0 || p <= 0) return ORC_E_ARG;
  size_t nk = (size_t)n * k, pk = (size_t)p * k, pp = (size_t)p * p;
  double* mean = (double*)malloc(sizeof(double) * n);
  double* Xp = (double*)malloc(sizeof(double) * nk);
  double* ybar = (double*)malloc(sizeof(double) * p);
  double* Yp = (double*)malloc(sizeof(double) * pk);
  double* d = (double*)malloc(sizeof(double) * p);
  double* S = (double*)malloc(sizeof(double) * pp);
  double* Sinv = (double*)malloc(sizeof(double) * pp);
  double* YtSi = (double*)malloc(sizeof(double) * pk); /* Y'^T S^-1, [k][p] */
  double* K = (double*)malloc(sizeof(double) * (size_t)n * p);
  double* sv = (double*)malloc(sizeof(double) * p);
  double* eps = (double*)malloc(sizeof(double) * pk);
  double* di = (double*)malloc(sizeof(double) * p);
  int rc = 0;
  if (!mean || !Xp || !ybar || !Yp || !d || !S || !Sinv || !YtSi || !K || !sv || !eps || !di) {
    rc = ORC_E_NOMEM;
    goto done;
  }
  /* 1. mean, perturbations, inflation (EnKF.hpp:149-165, :316-335) */
  orc_ensemble_mean(k, n, X, mean);
  orc_ensemble_perturbations(k, n, X, mean, sqrt(inflation), Xp);
  {
    double ss = 0.0;
    for (size_t t = 0; t < nk; ++t) ss += Xp[t] * Xp[t];
    diag->background_spread = sqrt(ss / (double)nk);
  }
  /* 2-3. observation space, innovation (EnKF.hpp:168-184) */
  obs_stats(k, p, Yb, yo, ybar, Yp, d);
  {
    double ss = 0.0;
    for (int64_t o = 0; o < p; ++o) ss += d[o] * d[o];
    diag->innovation_norm = sqrt(ss);
  }
  /* 4. S = Y'Y'^T/(k-1) + R, K = X' Y'^T S^-1 / (k-1) (EnKF.hpp:188-199) */
  for (int64_t a = 0; a < p; ++a)
    for (int64_t b = 0; b < p; ++b) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += Yp[(size_t)j * p + a] * Yp[(size_t)j * p + b];
      S[(size_t)a * p + b] = s / (double)(k - 1) + (a == b ? var[a] : 0.0);
    }
  rc = orc_lu_inverse((int)p, S, Sinv);
  if (rc) goto done;
  for (int j = 0; j < k; ++j)
    for (int64_t b = 0; b < p; ++b) {
      double s = 0.0;
      for (int64_t a = 0; a < p; ++a) s += Yp[(size_t)j * p + a] * Sinv[(size_t)a * p + b];
      YtSi[(size_t)j * p + b] = s;
    }
  diag->max_kalman_gain = -DBL_MAX;
  diag->min_kalman_gain = DBL_MAX;
  for (int64_t i = 0; i < n; ++i)
    for (int64_t b = 0; b < p; ++b) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += Xp[(size_t)j * n + i] * YtSi[(size_t)j * p + b];
      s /= (double)(k - 1);
      K[(size_t)i * p + b] = s;
      if (s > diag->max_kalman_gain) diag->max_kalman_gain = s;
      if (s < diag->min_kalman_gain) diag->min_kalman_gain = s;
    }
  /* cond(S) = s_max / s_min (EnKF.hpp:206-209) */
  rc = orc_singular_values((int)p, (int)p, S, sv);
  if (rc) goto done;
  diag->condition_number = sv[0] / sv[p - 1];
  /* 5. eps = chol_L(R) Z (EnKF.hpp:358-360); R diagonal => L = diag(sqrt var) */
  for (int64_t o = 0; o < p; ++o)
    for (int j = 0; j < k; ++j) eps[(size_t)j * p + o] = sqrt(var[o]) * Z[(size_t)o * k + j];
  /* 6-7. x_a,i = xbar + X'_i + K (yo + eps_i - Yb_i) (EnKF.hpp:216-234) */
  for (int j = 0; j < k; ++j) {
    for (int64_t o = 0; o < p; ++o)
      di[o] = (yo[o] + eps[(size_t)j * p + o]) - Yb[(size_t)j * p + o];
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t o = 0; o < p; ++o) s += K[(size_t)i * p + o] * di[o];
      X[(size_t)j * n + i] = mean[i] + (Xp[(size_t)j * n + i] + s);
    }
  }
  /* 8. recompute mean and perturbations, analysis spread (EnKF.hpp:237-253) */
  orc_ensemble_mean(k, n, X, mean);
  {
    double ss = 0.0;
    for (int j = 0; j < k; ++j)
      for (int64_t i = 0; i < n; ++i) {
        double v = X[(size_t)j * n + i] - mean[i];
        ss += v * v;
      }
    diag->analysis_spread = sqrt(ss / (double)((size_t)k * n));
  }
  diag->analysis_increment_norm = 0.0; /* never assigned, EnKF.hpp:398 */
  diag->ensemble_size = k;
  diag->observation_count = p;
done:
  free(mean); free(Xp); free(ybar); free(Yp); free(d); free(S); free(Sinv);
  free(YtSi); free(K); free(sv); free(eps); free(di);
  return rc;
}

/* ======================================================================== */
/* LWEnKF                                                                     */
/* ======================================================================== */

/* computeLocalizationFunction / computeGaspariCohnFunction, LWEnKF.hpp:600-636.
 * The reference's "gaspari_cohn" polynomial is kept as written (it is 4 at
 * 0); gc_standard = 1 evaluates Gaspari & Cohn (1999) eq. (4.10) instead. */
double orc_lw_taper(double dist, double radius, int fn, int gc_standard) {
  double nd = dist / radius;
  switch (fn) {
    case ORC_LW_EXPONENTIAL:
      return exp(-nd);
    case ORC_LW_CUTOFF:
      return nd <= 1.0 ? 1.0 : 0.0;
    case ORC_LW_GASPARI_COHN:
      if (gc_standard) return orc_gaspari_cohn(nd);
      if (nd >= 2.0) return 0.0;
      if (nd >= 1.0) {
        double z = nd - 1.0;
        return ((-0.25 * z + 0.5) * z + 0.625) * z + 0.125;
      } else {
        double z = nd;
        return (((-0.25 * z + 0.5) * z + 0.625) * z - 5.0) * z + 4.0;
      }
    default:
      return exp(-0.5 * nd * nd);
  }
}

/* LWEnKF::Analyse, LWEnKF.hpp:207-334, with the N(0,1) draws Z [p][k]
 * injected (they come from std::random_device at LWEnKF.hpp:690-706).
 *   weights      computeWeights and its helpers (:413-500): uniform 1/k;
 *                adaptive 1/(||x'_i|| + 1e-8) and inverse_var
 *                1/(||x'_i||^2 + 1e-8), each normalised twice; likelihood
 *                exp(-0.5 (yo - Hx_i)^T R^-1 (yo - Hx_i)), normalised once
 *   Pf           sum_i (w_i y'_i) y'_i^T over un-inflated Y' (:502-513)
 *   S            Pf o L_o + R, L_o(a,b) = f(|a-b| / p) (:515-533)
 *   K            (X'_rho Y'^T S^-1) / (k-1) o L_g, L_g(s,b) = f(|s-b| / max(n,p))
 *                (:272-279, :535-553)
 *   update       x_i = xbar + (X'_rho,i + K (yo + sqrt(r) z_i - Hx_i)) (:286-301)
 * Yb is [k][p]; X [k][n] is updated in place. */
int orc_lwenkf_analyse(int k, int64_t n, int64_t p, double* X, const double* Yb,
                       const double* yo, const double* var, const double* Z, double inflation,
                       double radius, int fn, int gc_standard, int weighting, double* weights,
                       double* Kout, int want_cond, orc_lwenkf_diag* diag) {
  if (k < 2 || n <= 0 || p <= 0 || !(radius > 0.0)) return ORC_E_ARG;
  size_t nk = (size_t)n * k, pk = (size_t)p * k, pp = (size_t)p * p;
  double* mean = (double*)malloc(sizeof(double) * n);
  double* Xp = (double*)malloc(sizeof(double) * nk);
  double* ybar = (double*)malloc(sizeof(double) * p);
  double* Yp = (double*)malloc(sizeof(double) * pk);
  double* d = (double*)malloc(sizeof(double) * p);
  double* S = (double*)malloc(sizeof(double) * pp);
  double* Sinv = (double*)malloc(sizeof(double) * pp);
  double* XY = (double*)malloc(sizeof(double) * (size_t)n * p);
  double* K = (double*)malloc(sizeof(double) * (size_t)n * p);
  double* sv = (double*)malloc(sizeof(double) * p);
  double* di = (double*)malloc(sizeof(double) * p);
  double* w = (double*)malloc(sizeof(double) * k);
  int rc = 0;
  if (!mean || !Xp || !ybar || !Yp || !d || !S || !Sinv || !XY || !K || !sv || !di || !w) {
    rc = ORC_E_NOMEM;
    goto done;
  }
  /* 1. mean and un-inflated perturbations (:215-229) */
  orc_ensemble_mean(k, n, X, mean);
  orc_ensemble_perturbations(k, n, X, mean, 1.0, Xp);
  /* 2. weights (:413-430) */
  if (weighting == ORC_LW_ADAPTIVE || weighting == ORC_LW_INVERSE_VAR) {
    double sum_inv = 0.0;
    for (int j = 0; j < k; ++j) {
      double ss = 0.0;
      for (int64_t i = 0; i < n; ++i) ss += Xp[(size_t)j * n + i] * Xp[(size_t)j * n + i];
      double m = weighting == ORC_LW_ADAPTIVE ? sqrt(ss) : ss;
      w[j] = 1.0 / (m + 1e-8);
      sum_inv += w[j];
    }
    for (int j = 0; j < k; ++j) w[j] /= sum_inv;
  } else if (weighting == ORC_LW_LIKELIHOOD) {
    for (int j = 0; j < k; ++j) {
      double chi2 = 0.0;
      for (int64_t o = 0; o < p; ++o) {
        double inn = yo[o] - Yb[(size_t)j * p + o];
        chi2 += inn * (1.0 / var[o]) * inn;
      }
      w[j] = exp(-0.5 * chi2);
    }
  } else {
    for (int j = 0; j < k; ++j) w[j] = 1.0 / k;
  }
  {
    double sum = 0.0;
    for (int j = 0; j < k; ++j) sum += w[j];
    if (!(sum > 0.0)) { rc = ORC_E_SINGULAR; goto done; }
    for (int j = 0; j < k; ++j) w[j] /= sum;
    double mx = w[0], mn = w[0], mw = 1.0 / k, wv = 0.0;
    for (int j = 0; j < k; ++j) {
      if (w[j] > mx) mx = w[j];
      if (w[j] < mn) mn = w[j];
      wv += (w[j] - mw) * (w[j] - mw);
    }
    diag->max_weight = mx;
    diag->min_weight = mn;
    diag->weight_variance = wv / k;
    if (weights) memcpy(weights, w, sizeof(double) * k);
  }
  /* 3. inflation of X' only (:648-668) */
  {
    double sc = sqrt(inflation), ss = 0.0;
    for (size_t t = 0; t < nk; ++t) {
      Xp[t] *= sc;
      ss += Xp[t] * Xp[t];
    }
    diag->background_spread = sqrt(ss / (double)nk);
  }
  /* 4-5. observation space, innovation (:240-255) */
  obs_stats(k, p, Yb, yo, ybar, Yp, d);
  {
    double ss = 0.0;
    for (int64_t o = 0; o < p; ++o) ss += d[o] * d[o];
    diag->innovation_norm = sqrt(ss);
  }
  /* 6. S = Pf o L_o + R (:258-270) */
  for (int64_t a = 0; a < p; ++a)
    for (int64_t b = 0; b < p; ++b) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += (w[j] * Yp[(size_t)j * p + a]) * Yp[(size_t)j * p + b];
      double dist = (double)(a > b ? a - b : b - a) / (double)p;
      S[(size_t)a * p + b] = s * orc_lw_taper(dist, radius, fn, gc_standard) +
                             (a == b ? var[a] : 0.0);
    }
  rc = orc_lu_inverse((int)p, S, Sinv);
  if (rc) goto done;
  /* K = (X' Y'^T) S^-1 / (k-1), then o L_g (:272-279) */
  for (int64_t i = 0; i < n; ++i)
    for (int64_t b = 0; b < p; ++b) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += Xp[(size_t)j * n + i] * Yp[(size_t)j * p + b];
      XY[(size_t)i * p + b] = s;
    }
  diag->max_kalman_gain = -DBL_MAX;
  diag->min_kalman_gain = DBL_MAX;
  {
    double gdim = (double)(n > p ? n : p);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t b = 0; b < p; ++b) {
        double s = 0.0;
        for (int64_t a = 0; a < p; ++a) s += XY[(size_t)i * p + a] * Sinv[(size_t)a * p + b];
        double dist = (double)(i > b ? i - b : b - i) / gdim;
        s = (s / (double)(k - 1)) * orc_lw_taper(dist, radius, fn, gc_standard);
        K[(size_t)i * p + b] = s;
        if (s > diag->max_kalman_gain) diag->max_kalman_gain = s;
        if (s < diag->min_kalman_gain) diag->min_kalman_gain = s;
      }
  }
  if (Kout) memcpy(Kout, K, sizeof(double) * (size_t)n * p);
  /* cond(S) = s_max / s_min (:280-284) */
  if (want_cond) {
    rc = orc_singular_values((int)p, (int)p, S, sv);
    if (rc) goto done;
    diag->condition_number = sv[0] / sv[p - 1];
  } else {
    diag->condition_number = NAN;
  }
  /* 7-8. perturbed observations and the member updates (:286-301) */
  for (int j = 0; j < k; ++j) {
    for (int64_t o = 0; o < p; ++o)
      di[o] = (yo[o] + sqrt(var[o]) * Z[(size_t)o * k + j]) - Yb[(size_t)j * p + o];
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t o = 0; o < p; ++o) s += K[(size_t)i * p + o] * di[o];
      X[(size_t)j * n + i] = mean[i] + (Xp[(size_t)j * n + i] + s);
    }
  }
  /* 10. analysis spread (:304-320) */
  orc_ensemble_mean(k, n, X, mean);
  {
    double ss = 0.0;
    for (int j = 0; j < k; ++j)
      for (int64_t i = 0; i < n; ++i) {
        double v = X[(size_t)j * n + i] - mean[i];
        ss += v * v;
      }
    diag->analysis_spread = sqrt(ss / (double)((size_t)k * n));
  }
done:
  free(mean); free(Xp); free(ybar); free(Yp); free(d); free(S); free(Sinv);
  free(XY); free(K); free(sv); free(di); free(w);
  return rc;
}

/* ======================================================================== */
/* ETKF                                                                       */
/* ======================================================================== */

/* ETKF::Analyse, ETKF.hpp:100-178: X' *= rho, Pa = (Y'^T R^-1 Y' + (k-1)I)^-1,
 * wa = Pa Y'^T R^-1 d, Wa = sqrt(k-1) chol_L(Pa), Xa = (xbar + X'wa)1^T + X'Wa. */
int orc_etkf_analyse(int k, int64_t n, int64_t p, double* X, const double* Yb,
                     const double* yo, const double* var, double inflation) {
  if (k < 2 || n <= 0 || p <= 0) return ORC_E_ARG;
  size_t kk = (size_t)k * k;
  double* mean = (double*)malloc(sizeof(double) * n);
  double* Xp = (double*)malloc(sizeof(double) * (size_t)n * k);
  double* ybar = (double*)malloc(sizeof(double) * p);
  double* Yp = (double*)malloc(sizeof(double) * (size_t)p * k);
  double* d = (double*)malloc(sizeof(double) * p);
  double* M = (double*)malloc(sizeof(double) * kk);
  double* Pa = (double*)malloc(sizeof(double) * kk);
  double* L = (double*)malloc(sizeof(double) * kk);
  double* b = (double*)malloc(sizeof(double) * k);
  double* wa = (double*)malloc(sizeof(double) * k);
  int rc = 0;
  if (!mean || !Xp || !ybar || !Yp || !d || !M || !Pa || !L || !b || !wa) {
    rc = ORC_E_NOMEM;
    goto done;
  }
  orc_ensemble_mean(k, n, X, mean);
  orc_ensemble_perturbations(k, n, X, mean, inflation, Xp);
  obs_stats(k, p, Yb, yo, ybar, Yp, d);
  for (int a = 0; a < k; ++a) {
    for (int c = 0; c < k; ++c) {
      double s = 0.0;
      for (int64_t o = 0; o < p; ++o)
        s += Yp[(size_t)a * p + o] * Yp[(size_t)c * p + o] / var[o];
      M[(size_t)a * k + c] = s + (a == c ? (double)(k - 1) : 0.0);
    }
    double s = 0.0;
    for (int64_t o = 0; o < p; ++o) s += Yp[(size_t)a * p + o] * d[o] / var[o];
    b[a] = s;
  }
  rc = orc_lu_inverse(k, M, Pa);
  if (rc) goto done;
  for (int a = 0; a < k; ++a) {
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += Pa[(size_t)a * k + c] * b[c];
    wa[a] = s;
  }
  rc = orc_cholesky_lower(k, Pa, L);
  if (rc) goto done;
  {
    double sq = sqrt((double)(k - 1));
    for (int64_t i = 0; i < n; ++i) {
      double xm = mean[i];
      for (int c = 0; c < k; ++c) xm += Xp[(size_t)c * n + i] * wa[c];
      for (int j = 0; j < k; ++j) {
        double s = 0.0;
        for (int c = 0; c < k; ++c) s += Xp[(size_t)c * n + i] * (sq * L[(size_t)c * k + j]);
        X[(size_t)j * n + i] = xm + s;
      }
    }
  }
done:
  free(mean); free(Xp); free(ybar); free(Yp); free(d); free(M); free(Pa); free(L);
  free(b); free(wa);
  return rc;
}

/* ======================================================================== */
/* LETKF                                                                      */
/* ======================================================================== */

static double letkf_distance(const orc_letkf_problem* P, int64_t g, int64_t o) {
  if (P->coord == ORC_COORD_GEOGRAPHIC)
    return orc_haversine_km(P->col_a[g], P->col_b[g], P->obs_a[o], P->obs_b[o]);
  if (P->coord == ORC_COORD_CARTESIAN)
    return orc_cart_distance(P->col_a[g], P->col_b[g], P->col_c ? P->col_c[g] : 0.0,
                             P->obs_a[o], P->obs_b[o], P->obs_c ? P->obs_c[o] : 0.0);
  return orc_grid_distance(P->col_a[g], P->col_b[g], P->obs_a[o], P->obs_b[o]);
}

/* A spatial index over observations so the standard LETKF can be timed at
 * full size: obs sorted by the second coordinate (latitude / grid y); a
 * column scans only the band within the radius.  Result sets are identical to
 * the full scan of LETKF.hpp:159-165. */
typedef struct {
  int64_t* order;
  double* key;
  int64_t p;
} obs_index;

/* qsort has no context argument; the key array is passed through a file
 * static.  build_index is only called outside parallel regions. */
static const double* g_sort_key;
static int cmp_key_idx(const void* a, const void* b) {
  int64_t ia = *(const int64_t*)a, ib = *(const int64_t*)b;
  if (g_sort_key[ia] < g_sort_key[ib]) return -1;
  if (g_sort_key[ia] > g_sort_key[ib]) return 1;
  return ia < ib ? -1 : (ia > ib);
}

static int build_index(const orc_letkf_problem* P, obs_index* ix) {
  ix->p = P->p;
  ix->order = (int64_t*)malloc(sizeof(int64_t) * (P->p > 0 ? P->p : 1));
  ix->key = (double*)malloc(sizeof(double) * (P->p > 0 ? P->p : 1));
  if (!ix->order || !ix->key) return ORC_E_NOMEM;
  const double* kv = (P->coord == ORC_COORD_GRID || P->coord == ORC_COORD_CARTESIAN) ? P->obs_b
                                                                                     : P->obs_a;
  for (int64_t o = 0; o < P->p; ++o) ix->order[o] = o;
  g_sort_key = kv;
  qsort(ix->order, (size_t)P->p, sizeof(int64_t), cmp_key_idx);
  for (int64_t t = 0; t < P->p; ++t) ix->key[t] = kv[ix->order[t]];
  return 0;
}

static int64_t lower_bound(const double* a, int64_t n, double v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* Local observations of column g: indices and distances, in ascending
 * observation index order (the order LETKF.hpp:160-165 produces). */
static int64_t local_obs(const orc_letkf_problem* P, const obs_index* ix, int64_t g,
                         int64_t* lidx, double* ldist) {
  int64_t lo = 0, hi = P->p;
  if (P->coord == ORC_COORD_GEOGRAPHIC) {
    double dlat = P->radius / 6371.0 * 180.0 / 3.14159265358979323846 * (1.0 + 1e-9) + 1e-9;
    lo = lower_bound(ix->key, ix->p, P->col_a[g] - dlat);
    hi = lower_bound(ix->key, ix->p, P->col_a[g] + dlat + 1e-300);
    while (hi < ix->p && ix->key[hi] <= P->col_a[g] + dlat) ++hi;
  } else if (P->coord == ORC_COORD_GRID) {
    lo = lower_bound(ix->key, ix->p, P->col_b[g] - P->radius - 1e-9);
    hi = lower_bound(ix->key, ix->p, P->col_b[g] + P->radius + 1e-9);
    while (hi < ix->p && ix->key[hi] <= P->col_b[g] + P->radius + 1e-9) ++hi;
  }
  int64_t m = 0;
  for (int64_t t = lo; t < hi; ++t) {
    int64_t o = ix->order[t];
    double dist = letkf_distance(P, g, o);
    if (dist <= P->radius) { lidx[m] = o; ldist[m] = dist; ++m; }
  }
  /* restore observation-index order */
  for (int64_t a = 1; a < m; ++a) {
    int64_t vi = lidx[a]; double vd = ldist[a]; int64_t b = a - 1;
    while (b >= 0 && lidx[b] > vi) { lidx[b + 1] = lidx[b]; ldist[b + 1] = ldist[b]; --b; }
    lidx[b + 1] = vi; ldist[b + 1] = vd;
  }
  return m;
}

/* Standard LETKF for one column: builds T (k x k, row-major) such that
 * X_a = xbar 1^T + X' T, i.e. T = wbar 1^T + W. Returns #local obs used. */
static int64_t standard_column_T(const orc_letkf_problem* P, const obs_index* ix, int64_t g,
                                 const double* Yp, const double* d, int64_t* lidx,
                                 double* ldist, double* A, double* V, double* ev, double* b,
                                 double* T) {
  const int k = P->k;
  const int64_t p = P->p;
  double rho = P->rho_col ? P->rho_col[g] : P->inflation;
  int64_t m = local_obs(P, ix, g, lidx, ldist);
  if (m == 0) { /* LETKF.hpp:167-190: xbar + sqrt(rho)(x - xbar) */
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) T[(size_t)i * k + j] = (i == j) ? sqrt(rho) : 0.0;
    return 0;
  }
  double c = P->radius / 2.0;
  for (int i = 0; i < k * k; ++i) A[i] = 0.0;
  for (int i = 0; i < k; ++i) b[i] = 0.0;
  for (int64_t t = 0; t < m; ++t) {
    int64_t o = lidx[t];
    double wgt = orc_gaspari_cohn(ldist[t] / c) / P->var[o];
    for (int i = 0; i < k; ++i) {
      double yi = Yp[(size_t)i * p + o] * wgt;
      b[i] += yi * d[o];
      for (int j = 0; j < k; ++j) A[(size_t)i * k + j] += yi * Yp[(size_t)j * p + o];
    }
  }
  for (int i = 0; i < k; ++i) A[(size_t)i * k + i] += (double)(k - 1) / rho;
  orc_jacobi_eigh(k, A, ev, V, 100, NULL);
  /* wbar = V L^-1 V^T b ; W = sqrt(k-1) V L^-1/2 V^T */
  double vb[512];
  double* tmp = vb;
  double* heap = NULL;
  if (k > 512) { heap = (double*)malloc(sizeof(double) * k); tmp = heap; }
  for (int a = 0; a < k; ++a) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += V[(size_t)i * k + a] * b[i];
    tmp[a] = s / ev[a];
  }
  double sq = sqrt((double)(k - 1));
  for (int i = 0; i < k; ++i) {
    double wbar = 0.0;
    for (int a = 0; a < k; ++a) wbar += V[(size_t)i * k + a] * tmp[a];
    for (int j = 0; j < k; ++j) {
      double s = 0.0;
      for (int a = 0; a < k; ++a) s += V[(size_t)i * k + a] * V[(size_t)j * k + a] / sqrt(ev[a]);
      T[(size_t)i * k + j] = wbar + sq * s;
    }
  }
  free(heap);
  return m;
}

int orc_letkf_standard(const orc_letkf_problem* P, double* X, const double* Yb,
                       const int64_t* cols, int64_t ncols_sel, double* T_out,
                       int64_t* nloc_out, int threads) {
  const int k = P->k;
  if (k < 2 || P->nx <= 0 || P->ny <= 0 || P->nlev <= 0 || P->p < 0) return ORC_E_ARG;
  const int64_t ncol = P->nx * P->ny, n = ncol * P->nlev, p = P->p;
  double* ybar = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  double* Yp = (double*)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1) * k);
  double* d = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  obs_index ix = {0, 0, 0};
  if (!ybar || !Yp || !d) { free(ybar); free(Yp); free(d); return ORC_E_NOMEM; }
  if (p > 0) obs_stats(k, p, Yb, P->yo, ybar, Yp, d);
  if (build_index(P, &ix)) { free(ybar); free(Yp); free(d); return ORC_E_NOMEM; }
  int64_t nsel = cols ? ncols_sel : ncol;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  int err = 0;
#pragma omp parallel
  {
    int64_t* lidx = (int64_t*)malloc(sizeof(int64_t) * (p > 0 ? p : 1));
    double* ldist = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
    double* A = (double*)malloc(sizeof(double) * k * k);
    double* V = (double*)malloc(sizeof(double) * k * k);
    double* T = (double*)malloc(sizeof(double) * k * k);
    double* ev = (double*)malloc(sizeof(double) * k);
    double* b = (double*)malloc(sizeof(double) * k);
    double* xl = (double*)malloc(sizeof(double) * k);
    if (!lidx || !ldist || !A || !V || !T || !ev || !b || !xl) {
#pragma omp atomic write
      err = 1;
    } else {
#pragma omp for schedule(dynamic, 16)
      for (int64_t s = 0; s < nsel; ++s) {
        int64_t g = cols ? cols[s] : s;
        int64_t m = standard_column_T(P, &ix, g, Yp, d, lidx, ldist, A, V, ev, b, T);
        if (nloc_out) nloc_out[s] = m;
        if (T_out) memcpy(T_out + (size_t)s * k * k, T, sizeof(double) * k * k);
        for (int64_t l = 0; l < P->nlev; ++l) {
          int64_t i = l * ncol + g;
          double xm = 0.0;
          for (int j = 0; j < k; ++j) { xl[j] = X[(size_t)j * n + i]; xm += xl[j]; }
          xm *= 1.0 / (double)k;
          for (int j = 0; j < k; ++j) xl[j] -= xm;
          for (int j = 0; j < k; ++j) {
            double acc = 0.0;
            for (int a = 0; a < k; ++a) acc += xl[a] * T[(size_t)a * k + j];
            X[(size_t)j * n + i] = xm + acc;
          }
        }
      }
    }
    free(lidx); free(ldist); free(A); free(V); free(T); free(ev); free(b); free(xl);
  }
  free(ybar); free(Yp); free(d); free(ix.order); free(ix.key);
  return err ? ORC_E_NOMEM : 0;
}

/* One reference grid-point update, LETKF.hpp:197-242, given the local
 * obs-space ensemble H_local [m][k] (row-major) and the local y_o. */
static int refcompat_point(int k, int64_t m, const double* Hl, const double* yol,
                           double inflation, double* xloc, double* tvec) {
  double* yb = (double*)malloc(sizeof(double) * m);
  double* Yp = (double*)malloc(sizeof(double) * (size_t)m * k);
  double* dl = (double*)malloc(sizeof(double) * m);
  double* M = (double*)malloc(sizeof(double) * k * k);
  double* Pa = (double*)malloc(sizeof(double) * k * k);
  double* L = (double*)malloc(sizeof(double) * k * k);
  double* b = (double*)malloc(sizeof(double) * k);
  int rc = 0;
  if (!yb || !Yp || !dl || !M || !Pa || !L || !b) { rc = ORC_E_NOMEM; goto out; }
  for (int64_t i = 0; i < m; ++i) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += Hl[(size_t)i * k + j];
    yb[i] = s / (double)k;
    for (int j = 0; j < k; ++j) Yp[(size_t)i * k + j] = Hl[(size_t)i * k + j] - yb[i];
    dl[i] = yol[i] - yb[i];
  }
  /* Pa = rho (Y'^T I^-1 Y' + (k-1) I)^-1   (LETKF.hpp:214-219) */
  for (int a = 0; a < k; ++a)
    for (int c = 0; c < k; ++c) {
      double s = 0.0;
      for (int64_t i = 0; i < m; ++i) s += Yp[(size_t)i * k + a] * Yp[(size_t)i * k + c];
      M[(size_t)a * k + c] = s + (a == c ? (double)(k - 1) : 0.0);
    }
  rc = orc_lu_inverse(k, M, Pa);
  if (rc) goto out;
  for (int t = 0; t < k * k; ++t) Pa[t] *= inflation;
  /* wa = Pa Y'^T d ; Wa = sqrt(k-1) chol_L(Pa)   (LETKF.hpp:221-224) */
  for (int a = 0; a < k; ++a) {
    double s = 0.0;
    for (int64_t i = 0; i < m; ++i) s += Yp[(size_t)i * k + a] * dl[i];
    b[a] = s;
  }
  rc = orc_cholesky_lower(k, Pa, L);
  if (rc) goto out;
  {
    double xm = 0.0;
    for (int j = 0; j < k; ++j) xm += xloc[j];
    xm /= (double)k;
    double sq = sqrt((double)(k - 1));
    for (int i = 0; i < k; ++i) {
      double wa = 0.0;
      for (int c = 0; c < k; ++c) wa += Pa[(size_t)i * k + c] * b[c];
      double rs = 0.0;
      for (int j = 0; j < k; ++j) rs += sq * L[(size_t)i * k + j];
      if (tvec) tvec[i] = wa + rs;
      double xp = xloc[i] - xm;
      /* LETKF.hpp:235-238: xbar + x'_i wa_i + sum_j x'_i Wa(i,j) */
      xloc[i] = (xm + xp * wa) + xp * rs;
    }
  }
out:
  free(yb); free(Yp); free(dl); free(M); free(Pa); free(L); free(b);
  return rc;
}

int orc_letkf_refcompat_snapshot(const orc_letkf_problem* P, double* X, const double* Yb,
                                 double* t_out) {
  const int k = P->k;
  if (k < 2) return ORC_E_ARG;
  const int64_t ncol = P->nx * P->ny, n = ncol * P->nlev, p = P->p;
  obs_index ix = {0, 0, 0};
  if (build_index(P, &ix)) return ORC_E_NOMEM;
  int64_t* lidx = (int64_t*)malloc(sizeof(int64_t) * (p > 0 ? p : 1));
  double* ldist = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  double* Hl = (double*)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1) * k);
  double* yol = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  double* xl = (double*)malloc(sizeof(double) * k);
  double* tv = (double*)malloc(sizeof(double) * k);
  int rc = 0;
  for (int64_t g = 0; g < ncol && rc == 0; ++g) {
    int64_t m = local_obs(P, &ix, g, lidx, ldist);
    for (int64_t t = 0; t < m; ++t) {
      yol[t] = P->yo[lidx[t]];
      for (int j = 0; j < k; ++j) Hl[(size_t)t * k + j] = Yb[(size_t)j * p + lidx[t]];
    }
    for (int64_t l = 0; l < P->nlev; ++l) {
      int64_t i = l * ncol + g;
      for (int j = 0; j < k; ++j) xl[j] = X[(size_t)j * n + i];
      if (m == 0) {
        double xm = 0.0;
        for (int j = 0; j < k; ++j) xm += xl[j];
        xm /= (double)k;
        double sr = sqrt(P->rho_col ? P->rho_col[g] : P->inflation);
        for (int j = 0; j < k; ++j) { xl[j] = xm + (xl[j] - xm) * sr; tv[j] = sr; }
      } else {
        rc = refcompat_point(k, m, Hl, yol, P->rho_col ? P->rho_col[g] : P->inflation, xl, tv);
        if (rc) break;
      }
      if (t_out && l == 0) memcpy(t_out + (size_t)g * k, tv, sizeof(double) * k);
      for (int j = 0; j < k; ++j) X[(size_t)j * n + i] = xl[j];
    }
  }
  free(lidx); free(ldist); free(Hl); free(yol); free(xl); free(tv);
  free(ix.order); free(ix.key);
  return rc;
}

int orc_letkf_refcompat_inplace(const orc_letkf_problem* P, double* X,
                                const int64_t* row_ptr, const int64_t* col, const double* w) {
  const int k = P->k;
  if (k < 2) return ORC_E_ARG;
  const int64_t ncol = P->nx * P->ny, n = ncol * P->nlev, p = P->p;
  obs_index ix = {0, 0, 0};
  if (build_index(P, &ix)) return ORC_E_NOMEM;
  int64_t* lidx = (int64_t*)malloc(sizeof(int64_t) * (p > 0 ? p : 1));
  double* ldist = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  double* Hl = (double*)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1) * k);
  double* yol = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  double* xl = (double*)malloc(sizeof(double) * k);
  int rc = 0;
  /* points in linear index order: lev-major, then y, then x (raster order of
   * SimpleGeometry.hpp:50-52 for a single level) */
  for (int64_t i = 0; i < n && rc == 0; ++i) {
    int64_t g = i % ncol;
    int64_t m = local_obs(P, &ix, g, lidx, ldist);
    for (int j = 0; j < k; ++j) xl[j] = X[(size_t)j * n + i];
    double rho = P->rho_col ? P->rho_col[g] : P->inflation;
    if (m == 0) {
      double xm = 0.0;
      for (int j = 0; j < k; ++j) xm += xl[j];
      xm /= (double)k;
      for (int j = 0; j < k; ++j) xl[j] = xm + (xl[j] - xm) * sqrt(rho);
    } else {
      /* H re-applied to the *current* members (LETKF.hpp:200-205) */
      for (int64_t t = 0; t < m; ++t) {
        int64_t o = lidx[t];
        yol[t] = P->yo[o];
        for (int j = 0; j < k; ++j) {
          double s = 0.0;
          for (int64_t q = row_ptr[o]; q < row_ptr[o + 1]; ++q)
            s += w[q] * X[(size_t)j * n + col[q]];
          Hl[(size_t)t * k + j] = s;
        }
      }
      rc = refcompat_point(k, m, Hl, yol, rho, xl, NULL);
    }
    for (int j = 0; j < k; ++j) X[(size_t)j * n + i] = xl[j];
  }
  free(lidx); free(ldist); free(Hl); free(yol); free(xl);
  free(ix.order); free(ix.key);
  return rc;
}

/* ======================================================================== */
/* metrics                                                                    */
/* ======================================================================== */

/* Metrics::CalculateAll and members, Metrics.hpp:74-294. */
int orc_metrics(int k, int64_t n, const double* X, const double* truth, double* mean_out,
                double* spread_out, orc_metrics_out* out) {
  if (k < 2 || n <= 0) return ORC_E_ARG;
  double* mean = mean_out ? mean_out : (double*)malloc(sizeof(double) * n);
  double* spread = spread_out ? spread_out : (double*)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; ++i) mean[i] = 0.0;
  for (int j = 0; j < k; ++j)
    for (int64_t i = 0; i < n; ++i) mean[i] += X[(size_t)j * n + i];
  for (int64_t i = 0; i < n; ++i) mean[i] /= (double)k;
  for (int64_t i = 0; i < n; ++i) spread[i] = 0.0;
  for (int j = 0; j < k; ++j)
    for (int64_t i = 0; i < n; ++i) {
      double df = X[(size_t)j * n + i] - mean[i];
      spread[i] += df * df;
    }
  for (int64_t i = 0; i < n; ++i) spread[i] = sqrt(spread[i] / (double)(k - 1));
  double rm = 0.0, bias = 0.0, sx = 0, sy = 0, sxy = 0, sx2 = 0, sy2 = 0, avs = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double df = mean[i] - truth[i];
    rm += df * df;
    bias += df;
    sx += mean[i]; sy += truth[i]; sxy += mean[i] * truth[i];
    sx2 += mean[i] * mean[i]; sy2 += truth[i] * truth[i];
    avs += spread[i];
  }
  double nn = (double)n;
  out->rmse = sqrt(rm / nn);
  out->bias = bias / nn;
  out->correlation = (nn * sxy - sx * sy) / sqrt((nn * sx2 - sx * sx) * (nn * sy2 - sy * sy));
  out->avg_spread = avs / nn;
  double crps = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double sd = 0.0, st = 0.0;
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b) sd += fabs(X[(size_t)a * n + i] - X[(size_t)b * n + i]);
    for (int a = 0; a < k; ++a) st += fabs(X[(size_t)a * n + i] - truth[i]);
    crps += st / k - sd / (2.0 * k * k);
  }
  out->crps = crps / nn;
  if (!mean_out) free(mean);
  if (!spread_out) free(spread);
  return 0;
}

/* ======================================================================== */
/* Lorenz-96                                                                  */
/* ======================================================================== */

/* tests/python/lorenz95.py:21: (x_{i+1} - x_{i-2}) x_{i-1} - x_i + F, cyclic. */
void orc_l96_tendency(int64_t N, const double* x, double F, double* dxdt) {
  for (int64_t i = 0; i < N; ++i) {
    double xp1 = x[(i + 1) % N], xm2 = x[(i + N - 2) % N], xm1 = x[(i + N - 1) % N];
    dxdt[i] = (xp1 - xm2) * xm1 - x[i] + F;
  }
}

/* lorenz95.py:3-31: forward Euler with clipping at +-1e10. */
void orc_l96_euler_step(int64_t N, double* x, double F, double dt) {
  double* d = (double*)malloc(sizeof(double) * N);
  orc_l96_tendency(N, x, F, d);
  for (int64_t i = 0; i < N; ++i) {
    double di = fmax(-1e10, fmin(1e10, d[i]));
    x[i] = fmax(-1e10, fmin(1e10, x[i] + dt * di));
  }
  free(d);
}

/* Classical fourth-order Runge-Kutta step of the same tendency. */
void orc_l96_rk4_step(int64_t N, double* x, double F, double dt) {
  double* k1 = (double*)malloc(sizeof(double) * N * 5);
  double *k2 = k1 + N, *k3 = k2 + N, *k4 = k3 + N, *xt = k4 + N;
  orc_l96_tendency(N, x, F, k1);
  for (int64_t i = 0; i < N; ++i) xt[i] = x[i] + 0.5 * dt * k1[i];
  orc_l96_tendency(N, xt, F, k2);
  for (int64_t i = 0; i < N; ++i) xt[i] = x[i] + 0.5 * dt * k2[i];
  orc_l96_tendency(N, xt, F, k3);
  for (int64_t i = 0; i < N; ++i) xt[i] = x[i] + dt * k3[i];
  orc_l96_tendency(N, xt, F, k4);
  for (int64_t i = 0; i < N; ++i)
    x[i] += dt / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
  free(k1);
}

/* ======================================================================== */
/* learned inflation network (no reference counterpart)                      */
/* ======================================================================== */

static double softplus(double v) { return v > 30.0 ? v : log1p(exp(v)); }
static double sigmoid(double v) { return 1.0 / (1.0 + exp(-v)); }

int orc_mlp_forward(int64_t m, const double* P, const double* feat, double* rho_out) {
  const double *W1 = P, *b1 = W1 + 64 * 8, *W2 = b1 + 64, *b2 = W2 + 64 * 64, *W3 = b2 + 64,
               *b3 = W3 + 64;
  for (int64_t s = 0; s < m; ++s) {
    double h1[64], h2[64];
    const double* x = feat + s * 8;
    for (int i = 0; i < 64; ++i) {
      double a = b1[i];
      for (int j = 0; j < 8; ++j) a += W1[i * 8 + j] * x[j];
      h1[i] = tanh(a);
    }
    for (int i = 0; i < 64; ++i) {
      double a = b2[i];
      for (int j = 0; j < 64; ++j) a += W2[i * 64 + j] * h1[j];
      h2[i] = tanh(a);
    }
    double a = b3[0];
    for (int j = 0; j < 64; ++j) a += W3[j] * h2[j];
    rho_out[s] = 1.0 + softplus(a);
  }
  return 0;
}

int orc_mlp_loss_grad(int64_t m, const double* P, const double* feat, const double* target,
                      double* loss, double* grad) {
  const double *W1 = P, *b1 = W1 + 64 * 8, *W2 = b1 + 64, *b2 = W2 + 64 * 64, *W3 = b2 + 64,
               *b3 = W3 + 64;
  double *gW1 = grad, *gb1 = gW1 + 64 * 8, *gW2 = gb1 + 64, *gb2 = gW2 + 64 * 64,
         *gW3 = gb2 + 64, *gb3 = gW3 + 64;
  memset(grad, 0, sizeof(double) * ORC_MLP_NPARAM);
  double L = 0.0;
  for (int64_t s = 0; s < m; ++s) {
    double h1[64], h2[64], g2[64], g1[64];
    const double* x = feat + s * 8;
    for (int i = 0; i < 64; ++i) {
      double a = b1[i];
      for (int j = 0; j < 8; ++j) a += W1[i * 8 + j] * x[j];
      h1[i] = tanh(a);
    }
    for (int i = 0; i < 64; ++i) {
      double a = b2[i];
      for (int j = 0; j < 64; ++j) a += W2[i * 64 + j] * h1[j];
      h2[i] = tanh(a);
    }
    double a = b3[0];
    for (int j = 0; j < 64; ++j) a += W3[j] * h2[j];
    double rho = 1.0 + softplus(a);
    double r = rho - target[s];
    L += r * r;
    double ga = 2.0 * r * sigmoid(a) / (double)m;
    gb3[0] += ga;
    for (int j = 0; j < 64; ++j) {
      gW3[j] += ga * h2[j];
      g2[j] = ga * W3[j] * (1.0 - h2[j] * h2[j]);
    }
    for (int i = 0; i < 64; ++i) {
      gb2[i] += g2[i];
      for (int j = 0; j < 64; ++j) gW2[i * 64 + j] += g2[i] * h1[j];
    }
    for (int j = 0; j < 64; ++j) {
      double s2 = 0.0;
      for (int i = 0; i < 64; ++i) s2 += W2[i * 64 + j] * g2[i];
      g1[j] = s2 * (1.0 - h1[j] * h1[j]);
    }
    for (int i = 0; i < 64; ++i) {
      gb1[i] += g1[i];
      for (int j = 0; j < 8; ++j) gW1[i * 8 + j] += g1[i] * x[j];
    }
  }
  *loss = L / (double)m;
  return 0;
}
