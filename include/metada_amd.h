/*
 * metada_amd.h -- C-ABI of the MI355X (gfx950) ensemble Kalman analysis library.
 *
 * A drop-in for the analysis step behind metada's algorithm interface:
 *   EnKF<Tag>::Analyse   src/framework/algorithms/EnKF.hpp:139-256
 *   ETKF<Tag>::Analyse   src/framework/algorithms/ETKF.hpp:100-178
 *   LETKF<Tag>::Analyse  src/framework/algorithms/LETKF.hpp:63-243
 * plus the ensemble statistics they call (Ensemble.hpp:105-151), the
 * observation-space quantities (EnKF.hpp:168-184) and Metrics (Metrics.hpp).
 *
 * Conventions (shared with the reference):
 *   - Ensembles are member-major: data[j*n + i] is member j at the linear
 *     state index i = lev*nx*ny + y*nx + x (SimpleState.hpp:72-78,
 *     IdentityObsOperator.hpp:652).  Members are updated IN PLACE, as the
 *     reference algorithms do (EnKF.hpp:230-234, LETKF.hpp:240-242).
 *   - Observation-space ensembles H(x_j) are member-major [k][p]: one vector
 *     per member, as ObsOperator::apply returns it (EnKF.hpp:168-172).
 *   - Observation error enters as the R diagonal, i.e. variances
 *     (GridObservation::getCovariance, GridObservation.hpp:239-252).
 *   - Every entry point returns MDA_OK (0) or a negative MDA_E_* code; no C++
 *     exception crosses the ABI (the reference's exceptions become codes;
 *     mda_last_error() gives the message).  The WRFDA bridge uses the same
 *     convention (WRFDAObsOperator_c_api.h:33,43-71).
 *   - Pointers flagged on_device are HIP device pointers on the context's
 *     device; otherwise they are host pointers and the library copies.
 *   - One context per host thread; each context owns one HIP stream and a
 *     device workspace that is reused across calls.
 */
#ifndef METADA_AMD_H
#define METADA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDA_OK 0
#define MDA_E_ARG (-1)         /* invalid argument / shape */
#define MDA_E_DEVICE (-2)      /* HIP runtime error */
#define MDA_E_NOMEM (-3)       /* device allocation failed */
#define MDA_E_NUMERIC (-4)     /* non-SPD matrix, NaN/Inf input */
#define MDA_E_UNSUPPORTED (-5) /* combination not implemented */

typedef struct mda_ctx mda_ctx;

/* ---- context ------------------------------------------------------------ */
/* Replaces the implicit single-threaded host state of the reference
 * (ApplicationContext.hpp:84-89 initialises config/logging only). */
int mda_ctx_create(int device, mda_ctx** out);
int mda_ctx_destroy(mda_ctx* ctx);
/* Message of the context's last failed call; with ctx == NULL, of this host
 * thread's last failed context-free call (the file functions accept a NULL
 * context). */
const char* mda_last_error(const mda_ctx* ctx);
const char* mda_last_io_error(void);
int mda_ctx_synchronize(mda_ctx* ctx);
/* hipStream_t of the context, as void* (callers that interleave their own
 * kernels order them against this stream). */
void* mda_ctx_stream(mda_ctx* ctx);
/* Device workspace currently held by the context, in bytes. */
int64_t mda_ctx_workspace_bytes(const mda_ctx* ctx);

/* ---- data descriptors ----------------------------------------------------- */
typedef struct {
  int32_t k;               /* ensemble size (Ensemble::Size, Ensemble.hpp:99) */
  int64_t nx, ny, nlev;    /* n = nx*ny*nlev */
  double* data;            /* [k][n] contiguous, or NULL when members is used */
  double* const* members;  /* host only: k separate member vectors (reference layout) */
  int32_t on_device;       /* data is a device pointer */
} mda_ensemble;

enum { MDA_COORD_GRID = 0, MDA_COORD_GEOGRAPHIC = 1, MDA_COORD_CARTESIAN = 2 };

typedef struct {
  int64_t p;
  const double* value;     /* y_o (GridObservation::getObservationValues) */
  const double* variance;  /* R diagonal, error^2; +inf marks an invalid obs */
  const double* loc_a;     /* GRID: x | GEOGRAPHIC: latitude deg | CARTESIAN: x */
  const double* loc_b;     /* GRID: y | GEOGRAPHIC: longitude deg | CARTESIAN: y */
  const double* loc_c;     /* CARTESIAN: z (may be NULL = 0) */
  int32_t on_device;
} mda_obs;

enum { MDA_HOP_CSR = 0, MDA_HOP_HX = 1 };

/* The observation operator: either a linear stencil in CSR form over state
 * linear indices (the form IdentityObsOperator's IDW4 and the trilinear
 * operator both take), or a caller-computed H(x_j) for operators that only
 * exist on the host (ObsOperator::apply, ObsOperator.hpp:255-259). */
typedef struct {
  int32_t kind;            /* MDA_HOP_CSR | MDA_HOP_HX */
  const int64_t* row_ptr;  /* CSR: [p+1] */
  const int64_t* col;      /* CSR: state linear indices */
  const double* weight;    /* CSR: stencil weights */
  int64_t nnz;
  const double* hx;        /* HX: [k][p] member-major */
  int32_t on_device;
} mda_hop;

/* Column (horizontal point) locations for localisation.  GRID with NULL
 * arrays means column g = y*nx + x sits at (x, y) (SimpleGeometry.hpp:42-55). */
typedef struct {
  int32_t coord;           /* MDA_COORD_* ; distances as Location::distance_to */
  const double* col_a;     /* [nx*ny]: x | latitude deg | x */
  const double* col_b;     /* [nx*ny]: y | longitude deg | y */
  const double* col_c;     /* [nx*ny]: CARTESIAN z (may be NULL) */
  int32_t on_device;
} mda_geometry;

/* ---- ensemble statistics --------------------------------------------------- */
/* Ensemble::RecomputeMean / RecomputePerturbations (Ensemble.hpp:105-151):
 * mean = (sum_j x_j) * (1/k); pert_j = (x_j - mean) * scale.  Either output
 * may be NULL.  Outputs live on the device when out_on_device. */
int mda_ensemble_mean_perturbations(mda_ctx* ctx, const mda_ensemble* ens, double scale,
                                    double* mean, double* pert, int32_t out_on_device);

/* Observation-space mean, anomalies and innovation (EnKF.hpp:168-184,
 * LETKF.hpp:209-211): hx = H x_j, ybar, Y' = hx - ybar ([k][p]), d = y_o - ybar.
 * Any output may be NULL. */
int mda_obs_innovations(mda_ctx* ctx, const mda_ensemble* ens, const mda_obs* obs,
                        const mda_hop* hop, double* ybar, double* yp, double* d,
                        int32_t out_on_device);

/* ---- EnKF (EnKF.hpp:105-256) ------------------------------------------------- */
enum { MDA_INFL_MULTIPLICATIVE = 0, MDA_INFL_ADDITIVE = 1, MDA_INFL_RELAXATION = 2 };

typedef struct { /* EnKF::AnalysisResults, EnKF.hpp:84-96 */
  double innovation_norm;
  double analysis_increment_norm; /* never assigned by the reference: 0 */
  double background_spread;
  double analysis_spread;
  double max_kalman_gain;
  double min_kalman_gain;
  double condition_number;        /* +inf with an invalid obs; NaN beyond p = 4096 */
  int32_t ensemble_size;
  int64_t observation_count;
  int32_t inflation_method;
  double inflation_factor;
} mda_enkf_diag;

/* Stochastic EnKF with perturbed observations.  Z is the N(0,1) draw matrix
 * the reference takes from std::random_device (EnKF.hpp:346-356), injected
 * here, [p][k] row-major in the reference fill order (obs, member). */
int mda_enkf_analyse(mda_ctx* ctx, double inflation, int32_t inflation_method,
                     mda_ensemble* ens, const mda_obs* obs, const mda_hop* hop,
                     const double* Z, int32_t z_on_device, mda_enkf_diag* diag);

/* ---- LWEnKF (LWEnKF.hpp:207-334) ----------------------------------------------- */
enum { /* localization_function (LWEnKF.hpp:148-162) */
  MDA_LW_LOC_GAUSSIAN = 0, MDA_LW_LOC_EXPONENTIAL = 1, MDA_LW_LOC_CUTOFF = 2,
  MDA_LW_LOC_GASPARI_COHN = 3
};
enum { /* what "gaspari_cohn" evaluates */
  MDA_LW_GC_REFERENCE = 0, /* the reference's polynomial (LWEnKF.hpp:627-636): 4 at 0 */
  MDA_LW_GC_STANDARD = 1   /* Gaspari & Cohn (1999) eq. (4.10) in nd = dist / radius */
};
enum { /* weighting_scheme (LWEnKF.hpp:164-178) */
  MDA_LW_W_UNIFORM = 0, MDA_LW_W_ADAPTIVE = 1, MDA_LW_W_INVERSE_VAR = 2,
  MDA_LW_W_LIKELIHOOD = 3
};

typedef struct {
  double inflation;           /* X' *= sqrt(inflation) for every method (LWEnKF.hpp:648-668) */
  double localization_radius; /* in the index distance |i - j| / dim (LWEnKF.hpp:554-598) */
  int32_t inflation_method;   /* MDA_INFL_*: reported only */
  int32_t localization;       /* MDA_LW_LOC_* */
  int32_t gc_form;            /* MDA_LW_GC_* */
  int32_t weighting;          /* MDA_LW_W_* */
} mda_lwenkf_params;

typedef struct { /* LWEnKF::AnalysisResults, LWEnKF.hpp:66-84 */
  double innovation_norm;
  double analysis_increment_norm; /* never assigned by the reference: 0 */
  double background_spread;
  double analysis_spread;
  double max_kalman_gain;
  double min_kalman_gain;
  double condition_number; /* sigma_max / sigma_min of S */
  double localization_radius;
  double max_weight;
  double min_weight;
  double weight_variance;
  int32_t ensemble_size;
  int64_t observation_count;
  int32_t inflation_method;
  int32_t localization;
  int32_t weighting;
  double inflation_factor;
  /* device time of this call's stages, ms (no counterpart in the reference) */
  double ms_gain_solve; /* S^-T Y' by Gauss-Jordan */
  double ms_update;     /* the fused gain and member update (one launch) */
} mda_lwenkf_diag;

/* Local weighted EnKF: member weights w, S = L_o o (sum_i w_i y'_i y'_i^T) + R,
 * K = L_g o (X'_rho Y'^T S^-1 / (k-1)), x_i <- xbar + X'_rho,i + K (yo + sqrt(R) z_i
 * - H x_i), with L_o and L_g the reference's index-distance tapers.  Z is the
 * N(0,1) draw matrix the reference takes from std::random_device
 * (LWEnKF.hpp:690-706), [p][k] row-major.  weights (k values, host) receives
 * w when not NULL.  2 <= k <= 128, 1 <= p <= 4096.  Errors: MDA_E_ARG for an
 * invalid observation (+inf variance: the reference's S^-1 would be NaN),
 * MDA_E_NUMERIC for non-finite input, a singular S or likelihood weights that
 * all underflow (the reference would divide 0 by 0). */
int mda_lwenkf_analyse(mda_ctx* ctx, const mda_lwenkf_params* params, mda_ensemble* ens,
                       const mda_obs* obs, const mda_hop* hop, const double* Z,
                       int32_t z_on_device, double* weights, mda_lwenkf_diag* diag);

/* ---- ETKF (ETKF.hpp:86-178) ---------------------------------------------------- */
int mda_etkf_analyse(mda_ctx* ctx, double inflation, mda_ensemble* ens, const mda_obs* obs,
                     const mda_hop* hop);

/* Allocation-free ETKF.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under hipGraph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same result as mda_etkf_analyse; the
 * ensemble, observations and H must be device-resident, and every other
 * buffer comes from a static 26 MiB device array per GPU (allocated at module
 * load), so k <= MDA_NOALLOC_MAX_K and p * (k + 1) <= MDA_NOALLOC_MAX_OBS_VALUES
 * (MDA_E_ARG beyond).  Calls share that array: calls that overlap on one GPU
 * race. */
#define MDA_NOALLOC_MAX_K 64
#define MDA_NOALLOC_MAX_OBS_VALUES 2097152
int mda_etkf_analyse_noalloc(mda_ctx* ctx, double inflation, mda_ensemble* ens,
                             const mda_obs* obs, const mda_hop* hop);

/* ---- LETKF (LETKF.hpp:47-243) --------------------------------------------------- */
enum { MDA_LOC_GASPARI_COHN = 0, MDA_LOC_CUTOFF = 1 };
enum { MDA_RW_REAL = 0, MDA_RW_IDENTITY = 1 };
enum { MDA_INFLFORM_PRIOR = 0, MDA_INFLFORM_POSTERIOR = 1 };
enum { MDA_SQRT_SYMMETRIC = 0, MDA_SQRT_CHOLESKY = 1 };
enum { MDA_COMBINE_MATRIX = 0, MDA_COMBINE_ELEMENTWISE = 1 };
enum { MDA_PREC_FP64 = 0, MDA_PREC_FP32 = 1, MDA_PREC_BF16X3 = 2 };

typedef struct {
  double inflation;          /* rho (LETKF.hpp:52, float-narrowed by the config layer) */
  double radius;             /* localisation radius, units of the distance (LETKF.hpp:53) */
  int32_t localization;      /* GC taper with support = radius | hard cutoff (LETKF.hpp:162) */
  int32_t rweight;           /* real R^-1 | R = I (LETKF.hpp:214) */
  int32_t inflation_form;    /* A=(k-1)I/rho+C | Pa = rho (C+(k-1)I)^-1 (LETKF.hpp:219) */
  int32_t sqrt_form;         /* symmetric A^-1/2 | chol_L(Pa) (LETKF.hpp:223) */
  int32_t combine;           /* X'(wa 1^T + Wa) | x'_i (wa_i + sum_j Wa_ij) (LETKF.hpp:235-238) */
  int32_t precision;         /* Y'^T R^-1 Y' in FP64 | FP32 | BF16x3 (eig always FP64) */
  const double* rho_column;  /* optional per-column rho [nx*ny] (learned inflation) */
  int32_t rho_on_device;
  int32_t eig_method;        /* symmetric square root: MDA_EIG_AUTO (Householder
                                tridiagonalisation, QL eigenvalues, twisted-
                                factorisation eigenvectors; columns failing the
                                residual/orthogonality checks are re-solved by
                                Jacobi) | MDA_EIG_JACOBI (parallel-order Jacobi) */
  int32_t transform_precision; /* X_a = xbar + X' T in FP64 (0, the default) | MDA_PREC_BF16X3
                                  (split-precision bf16 products, FP32 accumulation; k a
                                  multiple of 16, at most 64) */
  int32_t keep_transforms;     /* 1: keep every column's T for mda_letkf_last_transforms.
                                  0 (the default): at k in {16, 32, 48, 64} with the
                                  tridiagonal eigensolver, the matrix combination and an
                                  FP64 transform, the eigenvector kernel applies
                                  X_a = X T' itself and T is not stored */
} mda_letkf_params;

enum { MDA_EIG_AUTO = 0, MDA_EIG_JACOBI = 1 };

/* Presets: Hunt et al. (2007) with Gaspari-Cohn R-localisation, and the
 * reference's own LETKF semantics (snapshot H). */
void mda_letkf_params_standard(mda_letkf_params* prm, double inflation, double radius);
void mda_letkf_params_refcompat(mda_letkf_params* prm, double inflation, double radius);

typedef struct {
  int64_t n_columns;
  int64_t n_empty_columns;
  int64_t total_local_obs;
  int64_t max_local_obs;
  int32_t max_jacobi_sweeps;
  int64_t n_fallback_columns; /* columns the tridiagonal path handed to Jacobi */
  double ms_obs;        /* HX, obs-space mean/anomalies, innovations */
  double ms_bucket;     /* local-observation index */
  double ms_gram;       /* per-column Y'^T R^-1 Y', Y'^T R^-1 d */
  double ms_solve;      /* per-column eigensolve / Cholesky and T */
  double ms_transform;  /* X_a = xbar + X'T */
  double ms_total;
  /* parts of ms_solve: tridiagonal reduction, eigenvalues, eigenvectors and
   * T (zero when the tridiagonal path is not used), and the Jacobi /
   * Cholesky kernel (fallback columns, or every column without that path) */
  double ms_eig_tridiag;
  double ms_eig_values;
  double ms_eig_vectors;
  double ms_solve_jacobi;
} mda_letkf_diag;

int mda_letkf_analyse(mda_ctx* ctx, const mda_letkf_params* prm, mda_ensemble* ens,
                      const mda_geometry* geom, const mda_obs* obs, const mda_hop* hop,
                      mda_letkf_diag* diag);

/* Batched LETKF (the batching form): `nb` independent analyses -- any mix of
 * ensemble sizes, precisions and observation sets, e.g. k = 32 and 64 over
 * clustered observations -- given as arrays of nb structs.  Flags:
 *   MDA_BATCH_GRAPH     capture all analyses as concurrent branches of one
 *                       hipGraph on the first call and replay it when the
 *                       same buffers and parameters come again (inputs must
 *                       then be device-resident or are re-staged each call);
 *   MDA_BATCH_QUEUE     per-column Gram in work-queue mode: persistent waves
 *                       take columns in descending cost (observations in reach),
 *                       which balances clustered observation counts;
 *   MDA_BATCH_AUTOTUNE  on the first call of a workload (k, precision, columns,
 *                       observations) time candidate Gram launch shapes and keep
 *                       the fastest for the context.
 * Ensembles, observations and H must be device-resident (on_device).
 * Each analysis gives the same result as mda_letkf_analyse.  No counterpart in
 * the reference (its algorithms run one analysis at a time, LETKF.hpp:63). */
enum { MDA_BATCH_GRAPH = 1, MDA_BATCH_QUEUE = 2, MDA_BATCH_AUTOTUNE = 4 };
typedef struct {
  double ms_total;          /* the whole batch (HIP events around the launch) */
  int32_t graph_replayed;   /* 1 if an instantiated graph was replayed */
  int32_t n_graphs;         /* graphs held by the context */
  int32_t gram_blocks[16];  /* Gram launch blocks per analysis (0 = one wave per column) */
  int32_t gram_queue[16];   /* 1 if the Gram ran in work-queue mode */
} mda_batch_info;
int mda_letkf_analyse_batch(mda_ctx* ctx, int32_t nb, const mda_letkf_params* prm,
                            mda_ensemble* ens, const mda_geometry* geom, const mda_obs* obs,
                            const mda_hop* hop, mda_letkf_diag* diag, int32_t flags,
                            mda_batch_info* info);

/* Per-column k x k transforms T (X_a = xbar + X' T) of the last LETKF call,
 * for the listed columns, row-major [ncols][k][k] (host).  Diagnostics and
 * parity tests; valid until the next call on the context.  Kept when that
 * call set keep_transforms = 1 or ran a path that stores T; not kept when the
 * analysis ran its columns in passes (k > 64 on grids over 65,536 columns,
 * or a pass size set by the environment variable MDA_LETKF_CHUNK_COLS):
 * MDA_E_ARG then. */
int mda_letkf_last_transforms(mda_ctx* ctx, const int64_t* cols, int64_t ncols, double* T_out);

/* Number of local observations (within the localisation radius) of every
 * column in the last LETKF call, [nx*ny] (host).  The p_g histogram of a
 * workload; the reference builds the same sets at LETKF.hpp:158-165. */
int mda_letkf_last_local_counts(mda_ctx* ctx, int32_t* counts);

/* Per-column eigensolver status of the last LETKF call, [nx*ny] (host):
 * 0 = tridiagonal path, otherwise the column was re-solved by Jacobi and the
 * bits say why (1 QL non-convergence, 2 eigen-residual, 4 orthogonality of a
 * close eigenpair, 8 eigenvalue below (k-1)/rho). */
int mda_letkf_last_column_flags(mda_ctx* ctx, int32_t* flags);

/* ---- observation operator and files around the analysis ------------------------ */
/* IdentityObsOperator (IdentityObsOperator.hpp:154-281,483-676) built on the
 * device as a CSR stencil (MDA_HOP_CSR): four taps per observation, weights
 * w / sum(w) with w = 1/dist (1e12 at dist 0), the neighbours
 * find4NearestGridPoints selects (:594-638) on an nx x ny x nlev grid.
 *   obs_coord GRID        the observation's (x, y, z) grid coordinates;
 *   obs_coord GEOGRAPHIC  mapped to a grid point first (convertGeographicToGrid,
 *                         :483-587): with a GEOGRAPHIC geometry, the nearest
 *                         column in (lat, lon) degrees (first on ties, y-major
 *                         order) and the nearest of the nlev levels vcoord
 *                         (NULL: level 0); with any other geometry (a
 *                         SimpleGeometry), every observation maps to (0, 0, 0),
 *                         as the reference's fallback search leaves it.
 * An invalid observation (variance +inf) gets an empty row, so H x = 0 there
 * (:164-167).  row_ptr [p+1], col and weight [4 p] (at most), on the device
 * when out_on_device. */
int mda_hop_idw4_build(mda_ctx* ctx, const mda_obs* obs, int32_t obs_coord,
                       const mda_geometry* geom, int64_t nx, int64_t ny, int64_t nlev,
                       const double* vcoord, int64_t* row_ptr, int64_t* col, double* weight,
                       int32_t out_on_device);

/* The file functions below run on the host; ctx may be NULL (the error
 * message is then mda_last_error(NULL)).
 *
 * GridObservation::loadFromFile (GridObservation.hpp:401-476): the text
 * format "Time Z Y X ZNU XLONG_U XLAT_U Value" after its header and dashed
 * separator; values equal to missing_value are dropped; variance = error^2
 * (the config layer narrows error and missing_value through float,
 * GridObservation.hpp:85-86); locations (X, Y, Z) for MDA_COORD_GRID or
 * (XLAT_U, XLONG_U, ZNU) for MDA_COORD_GEOGRAPHIC.  The set owns host arrays,
 * viewed as an mda_obs until mda_obs_set_free. */
typedef struct mda_obs_set mda_obs_set;
int mda_obs_read_grid_file(mda_ctx* ctx, const char* path, double error, double missing_value,
                           int32_t coord, mda_obs_set** out);
int mda_obs_set_view(const mda_obs_set* set, mda_obs* view, int32_t* coord);
void mda_obs_set_free(mda_obs_set* set);

/* SimpleState::readFromFile (SimpleState.hpp:248-263): every number of the
 * file in order; *count = how many (out may be NULL to size the buffer). */
int mda_state_read_text(mda_ctx* ctx, const char* path, double* out, int64_t capacity,
                        int64_t* count);
/* SimpleState::saveToFile (SimpleState.hpp:162-188): nrows rows of nx values,
 * "%12.6f" separated by one space -- six decimals, so lossy (about 5e-7). */
int mda_state_write_text(mda_ctx* ctx, const char* path, const double* v, int64_t nx,
                         int64_t nrows);
/* FP64-exact sidecar of a state (no counterpart in the reference, whose text
 * output is lossy): a 64-byte header ("MDASTATE", version 1, nx, ny, nlev as
 * little-endian int64) and the n doubles. */
int mda_state_write_binary(mda_ctx* ctx, const char* path, const double* v, int64_t nx,
                           int64_t ny, int64_t nlev);
int mda_state_read_binary(mda_ctx* ctx, const char* path, double* out, int64_t capacity,
                          int64_t* nx, int64_t* ny, int64_t* nlev);
/* saveEnsemble (EnKF.hpp:261-282, LETKF.hpp:125-142): <base>_mean.<format>
 * (mean in Ensemble::RecomputeMean order) and <base>_member_<i>.<format>
 * in the text format above (ny*nlev rows of nx), each with a .f64 sidecar
 * beside it when sidecar != 0.  Host or device ensemble. */
int mda_ensemble_save(mda_ctx* ctx, const mda_ensemble* ens, const char* base,
                      const char* format, int32_t sidecar);

/* ---- Learned inflation (no counterpart in the reference) -------------------------
 * Cycled twin experiments and a per-column MLP that predicts the LETKF
 * inflation rho_g (fed to mda_letkf_params.rho_column).  The forecast model
 * is Lorenz-96 (tests/python/lorenz95.py:3-37).
 */

/* Advances `nstates` Lorenz-96 states of size N (row-major [nstates][N]) in
 * place by `nsteps` steps of dt: scheme 0 = classical RK4, 1 = the forward
 * Euler step of lorenz95.py:28 with its clipping to +-1e10.  4 <= N <= 6144. */
int mda_l96_forecast(mda_ctx* ctx, double* X, int64_t nstates, int64_t N, double forcing,
                     double dt, int32_t nsteps, int32_t scheme, int32_t on_device);

/* Eight local innovation and spread statistics of every column, FP32
 * [nx*ny][8], from the background ensemble and observations (same local
 * observation sets and weights as the LETKF with these params):
 * log(Sd2/Sw), log(Ssp/Sw), log(Sd2/(Ssp+Sw)), Sd/sqrt(Sw Sd2), log(Sw),
 * log(1+n_local), log(column background variance), rho_prev - 1, where
 * Sw = sum w, Sd2 = sum w d^2, Sd = sum w d, Ssp = sum w |y'|^2/(k-1) over the
 * column's local observations with w = localisation / variance.
 * rho_prev (device or host per prm->rho_on_device, may be NULL) is the
 * previous inflation of each column. */
int mda_letkf_inflation_features(mda_ctx* ctx, const mda_letkf_params* prm,
                                 const mda_ensemble* ens, const mda_geometry* geom,
                                 const mda_obs* obs, const mda_hop* hop, const double* rho_prev,
                                 float* features, int32_t features_on_device);

/* MLP 8 -> 64 -> 64 -> 1 (tanh, rho = 1 + softplus(out)), FP32.  Parameters:
 * MDA_MLP_NPARAM floats laid out W1[64][8], b1[64], W2[64][64], b2[64],
 * W3[64], b3.  All arrays on the device when on_device, else host. */
enum { MDA_MLP_NIN = 8, MDA_MLP_HIDDEN = 64, MDA_MLP_NPARAM = 4801 };
int mda_mlp_forward(mda_ctx* ctx, const float* params, const float* features, int64_t m,
                    float* rho, int32_t on_device);
/* Mean squared error of rho against `target` over m samples and its gradient
 * (grad: MDA_MLP_NPARAM floats); the loss is returned on the host. */
int mda_mlp_loss_grad(mda_ctx* ctx, const float* params, const float* features,
                      const float* target, int64_t m, double* loss, float* grad,
                      int32_t on_device);
/* One Adam step (step counts from 1) on params with moments m1, m2. */
int mda_mlp_adam_step(mda_ctx* ctx, float* params, float* m1, float* m2, const float* grad,
                      int64_t step, float lr, float beta1, float beta2, float eps,
                      int32_t on_device);

/* ---- Metrics (Metrics.hpp:74-294) ------------------------------------------------ */
typedef struct {
  double rmse, bias, correlation, crps, avg_spread;
} mda_metrics;
int mda_ensemble_metrics(mda_ctx* ctx, const mda_ensemble* ens, const double* truth,
                         int32_t truth_on_device, mda_metrics* out);

#ifdef __cplusplus
}
#endif
#endif
