/*
 * oracle.h -- CPU FP64 restatement of metada's ensemble Kalman analysis path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library in
 * metada_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it.  The product path never calls into it.
 *
 * Conventions
 *   - Ensembles are member-major: X[j*n + i] is member j, state index i, with
 *     the reference linear index i = lev*nx*ny + y*nx + x
 *     (IdentityObsOperator.hpp:652, SimpleState.hpp:72-78).
 *   - Observation-space ensembles (HX) are member-major as well:
 *     Yb[j*p + o], i.e. one vector of p values per member, exactly what
 *     ObsOperator::apply returns per member (EnKF.hpp:168-172).
 *   - Small dense matrices are row-major, n x n.
 *   - All functions return 0 on success and a negative code on failure.
 *
 * Parity status of each function is listed in the header comment of oracle.c.
 */
#ifndef METADA_ORACLE_H
#define METADA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- small dense linear algebra (Eigen replacements) ------------------ */
int orc_cholesky_lower(int n, const double* A, double* L);
int orc_lu_inverse(int n, const double* A, double* Ainv);
int orc_jacobi_eigh(int n, const double* A, double* evals, double* V, int max_sweeps,
                    int* sweeps_out);
int orc_singular_values(int m, int n, const double* A, double* s);

/* ---- locations, localisation --------------------------------------------- */
double orc_haversine_km(double lat1, double lon1, double lat2, double lon2);
double orc_grid_distance(double x1, double y1, double x2, double y2);
double orc_cart_distance(double x1, double y1, double z1, double x2, double y2, double z2);
double orc_gaspari_cohn(double z);

/* ---- observation operator (IdentityObsOperator IDW4) ------------------- */
double orc_idw4(const double* state, int64_t nx, int64_t ny, int64_t nz,
                double x, double y, double z);
int orc_idw4_stencil(int64_t nx, int64_t ny, int64_t nz, double x, double y, double z,
                     int64_t* idx4, double* w4);
int orc_csr_apply(int64_t p, const int64_t* row_ptr, const int64_t* col, const double* w,
                  const double* state, double* out);

/* ---- ensemble statistics (Ensemble.hpp:105-151) ------------------------ */
void orc_ensemble_mean(int k, int64_t n, const double* X, double* mean);
void orc_ensemble_perturbations(int k, int64_t n, const double* X, const double* mean,
                                double scale, double* Xp);

/* ---- EnKF (EnKF.hpp:139-256, Z injected) ------------------------------- */
typedef struct {
  double innovation_norm;
  double analysis_increment_norm;
  double background_spread;
  double analysis_spread;
  double max_kalman_gain;
  double min_kalman_gain;
  double condition_number;
  int ensemble_size;
  int64_t observation_count;
} orc_enkf_diag;

int orc_enkf_analyse(int k, int64_t n, int64_t p, double* X, const double* Yb,
                     const double* yo, const double* var, const double* Z,
                     double inflation, orc_enkf_diag* diag);

/* ---- LWEnKF (LWEnKF.hpp:207-334, 413-668, Z injected) -------------------- */
enum { ORC_LW_GAUSSIAN = 0, ORC_LW_EXPONENTIAL = 1, ORC_LW_CUTOFF = 2, ORC_LW_GASPARI_COHN = 3 };
enum { ORC_LW_UNIFORM = 0, ORC_LW_ADAPTIVE = 1, ORC_LW_INVERSE_VAR = 2, ORC_LW_LIKELIHOOD = 3 };
typedef struct {
  double innovation_norm;
  double background_spread;
  double analysis_spread;
  double max_kalman_gain;
  double min_kalman_gain;
  double condition_number;
  double max_weight;
  double min_weight;
  double weight_variance;
} orc_lwenkf_diag;

/* the taper f(dist / radius); gc_standard selects Gaspari & Cohn (1999) for
 * "gaspari_cohn" instead of the reference's polynomial */
double orc_lw_taper(double dist, double radius, int fn, int gc_standard);
/* weights[k] out; K[n][p] out when not NULL (the tapered gain); cond(S) is
 * computed when want_cond (NaN otherwise) */
int orc_lwenkf_analyse(int k, int64_t n, int64_t p, double* X, const double* Yb,
                       const double* yo, const double* var, const double* Z, double inflation,
                       double radius, int fn, int gc_standard, int weighting, double* weights,
                       double* K, int want_cond, orc_lwenkf_diag* diag);

/* ---- ETKF (ETKF.hpp:100-178) ------------------------------------------- */
int orc_etkf_analyse(int k, int64_t n, int64_t p, double* X, const double* Yb,
                     const double* yo, const double* var, double inflation);

/* ---- LETKF ------------------------------------------------------------- */
enum { ORC_COORD_GRID = 0, ORC_COORD_GEOGRAPHIC = 1, ORC_COORD_CARTESIAN = 2 };

typedef struct {
  int k;
  int64_t nx, ny, nlev;     /* n = nx*ny*nlev, ncol = nx*ny */
  int coord;                /* ORC_COORD_* */
  const double* col_a;      /* per column: grid x | latitude deg  | cart x */
  const double* col_b;      /* per column: grid y | longitude deg | cart y */
  const double* col_c;      /* per column: unused | unused        | cart z */
  int64_t p;
  const double* obs_a;      /* per obs: grid x | latitude deg  | cart x */
  const double* obs_b;      /* per obs: grid y | longitude deg | cart y */
  const double* obs_c;      /* per obs: unused | unused        | cart z */
  const double* yo;         /* observation values */
  const double* var;        /* observation error variances (R diagonal) */
  double inflation;         /* rho */
  const double* rho_col;    /* optional per-column rho (NULL: use inflation) */
  double radius;            /* localisation radius (GC support / cutoff) */
} orc_letkf_problem;

/* Standard form (Hunt et al. 2007, R-localised with Gaspari-Cohn, symmetric
 * inverse square root by Jacobi).  X in/out [k][n].  Yb = snapshot HX [k][p].
 * If cols != NULL only the listed columns are updated (others untouched) and,
 * if T_out != NULL, the k x k transform of listed column c is written to
 * T_out[c*k*k] (row-major, X_a = xbar + X' T).  threads <= 0: all cores. */
int orc_letkf_standard(const orc_letkf_problem* P, double* X, const double* Yb,
                       const int64_t* cols, int64_t ncols_sel, double* T_out,
                       int64_t* nloc_out, int threads);

/* Reference semantics (LETKF.hpp:152-243), snapshot HX: R = I, hard cutoff,
 * Pa = rho (Y'^T Y' + (k-1) I)^-1 (PartialPivLU inverse), Wa = sqrt(k-1)
 * chol_L(Pa), element-wise combination.  t_out (optional) receives the
 * per-column combination vector t_i = wa_i + sum_j Wa(i,j) [ncol][k]. */
int orc_letkf_refcompat_snapshot(const orc_letkf_problem* P, double* X, const double* Yb,
                                 double* t_out);

/* Reference semantics, literal in-place order (LETKF.hpp:101-102,200-205,
 * 240-242): H is re-applied to the partially updated members at every grid
 * point; points are visited in linear index order.  H given as CSR. */
int orc_letkf_refcompat_inplace(const orc_letkf_problem* P, double* X,
                                const int64_t* row_ptr, const int64_t* col, const double* w);

/* ---- metrics (Metrics.hpp:74-294) -------------------------------------- */
typedef struct {
  double rmse, bias, correlation, crps, avg_spread;
} orc_metrics_out;
int orc_metrics(int k, int64_t n, const double* X, const double* truth, double* mean_out,
                double* spread_out, orc_metrics_out* out);

/* ---- Lorenz-96 (tests/python/lorenz95.py:3-37) ------------------------- */
void orc_l96_tendency(int64_t N, const double* x, double F, double* dxdt);
void orc_l96_euler_step(int64_t N, double* x, double F, double dt);
void orc_l96_rk4_step(int64_t N, double* x, double F, double dt);

/* ---- learned inflation network (8 -> 64 -> 64 -> 1, softplus + 1) ------- */
enum { ORC_MLP_IN = 8, ORC_MLP_H = 64, ORC_MLP_NPARAM = 8 * 64 + 64 + 64 * 64 + 64 + 64 + 1 };
/* params layout: W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3 */
int orc_mlp_forward(int64_t m, const double* params, const double* feat, double* rho_out);
/* loss = mean_i (rho_i - target_i)^2 ; grad has ORC_MLP_NPARAM entries */
int orc_mlp_loss_grad(int64_t m, const double* params, const double* feat,
                      const double* target, double* loss, double* grad);

#ifdef __cplusplus
}
#endif
#endif
