// Internal declarations shared by the HIP sources of libmetada_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "metada_amd.h"

// ---------------------------------------------------------------------------
// errors: every entry point converts these into MDA_E_* codes
// ---------------------------------------------------------------------------
struct mda_error : std::runtime_error {
  int code;
  mda_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MDA_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw mda_error(e_ == hipErrorOutOfMemory ? MDA_E_NOMEM : MDA_E_DEVICE,      \
                      std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

#define MDA_REQUIRE(cond, msg)                                                     \
  do {                                                                             \
    if (!(cond)) throw mda_error(MDA_E_ARG, msg);                                  \
  } while (0)

#define MDA_LAUNCH_CHECK() MDA_HIP(hipGetLastError())

// Bits of an analysis's device status word (zeroed at the start of each
// analysis; kernels OR into it):
//   MDA_ST_NOT_SPD          a Cholesky factor did not exist (non-SPD matrix)
//   MDA_ST_NONFINITE_OBS    a non-finite H(x), observation-space mean or
//                           innovation
//   MDA_ST_NONFINITE_STATE  a non-finite background value reached the update
// The update kernels write nothing when a MDA_ST_SKIP_UPDATE bit is set, so
// an analysis that fails before its update leaves the ensemble unchanged.
//   MDA_ST_BOUNDARY_OVERFLOW more observations at the localisation radius
//                           than the boundary list holds
enum : int {
  MDA_ST_NOT_SPD = 1,
  MDA_ST_NONFINITE_OBS = 2,
  MDA_ST_NONFINITE_STATE = 4,
  MDA_ST_BOUNDARY_OVERFLOW = 8,
  MDA_ST_SKIP_UPDATE = MDA_ST_NOT_SPD | MDA_ST_NONFINITE_OBS | MDA_ST_BOUNDARY_OVERFLOW,
};

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct mda_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  struct buf {
    void* ptr = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, buf> ws;  // named device workspace, grown on demand
  hipEvent_t ev[12] = {};
  // batched analyses: extra streams (default priority) for concurrent
  // analyses, fork/join events, instantiated graphs by signature, and the
  // autotuned Gram launch (blocks, work queue) by workload key
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> fj;
  // timing events of column-chunked analyses (grown on demand)
  std::vector<hipEvent_t> evpool;
  std::map<std::string, hipGraphExec_t> graphs;
  std::map<std::string, std::pair<int, int>> tune;
  // state of the last LETKF call (for mda_letkf_last_transforms)
  int last_k = 0;
  int64_t last_ncol = 0;
  double* last_T = nullptr;  // device, [ncol][k][k]
  int* last_nloc = nullptr;  // device, [ncol]
  int* last_flag = nullptr;  // device, [ncol] (tridiagonal path only)

  // returns a device buffer of at least `bytes`; contents are unspecified.
  // Each name has one meaning; a call fetches a name once, since growing a
  // buffer frees the previous allocation.
  void* get(const std::string& name, size_t bytes) {
    if (bytes == 0) bytes = 256;
    buf& b = ws[name];
    if (b.bytes < bytes) {
      // a cached graph may hold the old address: drop them all (their keys
      // hold every address they use, so none would match again anyway)
      drop_graphs();
      if (b.ptr) MDA_HIP(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
      size_t want = bytes + bytes / 8;  // headroom: fewer re-allocations
      MDA_HIP(hipMalloc(&b.ptr, want));
      b.bytes = want;
    }
    return b.ptr;
  }
  template <class T>
  T* get(const std::string& name, size_t count) {
    return static_cast<T*>(get(name, count * sizeof(T)));
  }
  void drop_graphs() {
    for (auto& kv : graphs)
      if (kv.second) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }
};

// Scoped "make this context's device current".
struct device_guard {
  int prev = 0;
  explicit device_guard(int dev) {
    MDA_HIP(hipGetDevice(&prev));
    if (prev != dev) MDA_HIP(hipSetDevice(dev));
  }
  ~device_guard() { (void)hipSetDevice(prev); }
};

// Host <-> device transfers.  Host buffers are ordinary pageable memory, so
// the copy is synchronous: the context's stream is drained first (it is
// non-blocking, so the copy would not otherwise wait for it) and the copy has
// completed when the call returns.
inline void h2d(mda_ctx* c, void* dst, const void* src, size_t bytes) {
  MDA_HIP(hipStreamSynchronize(c->stream));
  MDA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
inline void d2h(mda_ctx* c, void* dst, const void* src, size_t bytes) {
  MDA_HIP(hipStreamSynchronize(c->stream));
  MDA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

// Host data staged to the device (no-op when already on the device).
template <class T>
const T* to_device(mda_ctx* c, const std::string& name, const T* src, size_t count,
                   bool on_device) {
  if (src == nullptr) return nullptr;
  if (on_device) return src;
  T* d = c->get<T>(name, count);
  h2d(c, d, src, count * sizeof(T));
  return d;
}

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// kernels launched by the API layer (defined in the kernel sources)
// ---------------------------------------------------------------------------
namespace mda {

// clears `bytes` (multiple of 4) with a kernel, so the clear is graph-safe
void launch_zero(hipStream_t s, void* ptr, size_t bytes);
// streaming ensemble statistics
void launch_mean_pert(hipStream_t s, const double* X, int k, int64_t n, double scale,
                      double* mean, double* pert);
void launch_sumsq(hipStream_t s, const double* X, int64_t count, double* partial,
                  int nblocks);
void launch_dev_sumsq(hipStream_t s, const double* X, int k, int64_t n, double scale,
                      double* partial, int nblocks);
// observation space: Y' obs-major [p][k], ybar, d; optional member-major Y'.
// With `order`, observations are processed in the order order[0], order[1],
// ...; observation order[t] is written at position t, or at opos[order[t]]
// when `opos` is given.  y32 (if not null) receives an FP32 copy of Y'.
void launch_obs_stats(hipStream_t s, const double* X, int k, int64_t n, int64_t p,
                      const int64_t* row_ptr, const int64_t* col, const double* w,
                      const double* hx_mm, const double* yo, double* ybar, double* yp_om,
                      double* d, double* yp_mm, const uint32_t* order = nullptr,
                      float* y32 = nullptr, const uint32_t* opos = nullptr,
                      int* status = nullptr);
// keys = grid position of each observation's first stencil tap (groups of
// four levels share a key range; nxy = points per level), vals = 0..p-1
void launch_tap_keys(hipStream_t s, const int64_t* row_ptr, const int64_t* col, int64_t p,
                     int64_t nxy, uint32_t* keys, uint32_t* vals);
// ipos[order[t]] = t
void launch_invert_order(hipStream_t s, const uint32_t* order, int64_t p, uint32_t* ipos);
// global (all-observation) Gram: G[a][c] = sum_o r_o Y'[o][a] Q[o][c] with Q =
// [Y' | d] (mode 0) or D = d + sigma Z - Y' (mode 1, EnKF); partials reduced.
void launch_global_gram(hipStream_t s, const double* yp_om, const double* d,
                        const double* var, const double* Z, int k, int64_t p, int mode,
                        double* partial, int nblocks, double* out);
// small dense solves on one workgroup (matrices in global scratch)
void launch_etkf_solve_ws(hipStream_t s, int k, const double* EB, double* T, double* scratch,
                          int* status);
void launch_enkf_solve_ws(hipStream_t s, int k, const double* EB, const double* F, double* T,
                          double* Minv, double* scratch, int* status);
// X_a = xbar + scale * X' T for all state points (global T); skipped when
// the status word holds a MDA_ST_SKIP_UPDATE bit
void launch_transform_global(hipStream_t s, double* X, int k, int64_t n, double scale,
                             const double* T, int* status);
// EnKF diagnostics
void launch_enkf_gain_minmax_ws(hipStream_t s, const double* X, int k, int64_t n, double scale,
                                const double* Minv, const double* yp_om, const double* var,
                                int64_t p, double* G, double* partial, int nblocks);
void launch_cond_small(hipStream_t s, const double* yp_om, const double* var, int k, int p,
                       double* S, double* out);
void launch_sym_maxeig(hipStream_t s, const double* EB, int k, double* S, double* out);
void launch_metrics(hipStream_t s, const double* X, int k, int64_t n, const double* truth,
                    double* partial, int nblocks);

// LETKF
struct bucket_grid {
  int coord;
  int nrow, ncolb;       // bucket rows (lat / y) and bucket columns (lon / x)
  double a0, b0;         // origin of the bucket grid
  double da, db;         // bucket sizes
  int wrap;              // longitude wrap-around
};
void launch_bucket_keys(hipStream_t s, const double* oa, const double* ob, int64_t p,
                        bucket_grid g, uint32_t* keys, uint32_t* vals);
void radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp,
                      uint32_t* vals_tmp, int64_t p, int key_bits, uint32_t* hist,
                      int64_t hist_cap);
void launch_bucket_ranges(hipStream_t s, const uint32_t* sorted_keys, int64_t p,
                          int nbuckets, int* start);
// per-observation data in bucket order (1/var, coordinates; unit vectors of
// GEOGRAPHIC locations when ux is not null)
void launch_gather_obs(hipStream_t s, const uint32_t* order, int64_t p, const double* var,
                       const double* oa, const double* ob, const double* oc, int rweight,
                       double* s_rinv, double* s_a, double* s_b, double* s_c, double* ux,
                       double* uy, double* uz);
// Y' rows and d in bucket order, plus the FP32 copy of Y' when y32 is not null
void launch_gather_rows(hipStream_t s, const uint32_t* order, int64_t p, int k,
                        const double* yp_om, const double* d, double* s_yp, double* s_d,
                        float* y32);
struct letkf_gram_args {
  int k;
  int64_t ncol, nx;
  int coord;
  const double* col_a;
  const double* col_b;
  const double* col_c;
  bucket_grid grid;
  const int* bstart;
  const double* yp;    // sorted obs-major Y'
  const double* d;     // sorted innovations
  const double* rinv;  // sorted 1/var (or 1)
  const double* oa;
  const double* ob;
  const double* oc;
  double radius;
  int localization;
  int precision;
  const float* y32;    // sorted obs-major Y' in FP32 (FP32 / BF16x3 contractions)
  const double* ux;    // sorted unit vectors of GEOGRAPHIC observation locations (any precision)
  const double* uy;
  const double* uz;
  double* C;    // out [ncol][k][k]
  // C holds FP32 entries (the FP32 / BF16x3 contractions, whose products are
  // accumulated in FP32): half the buffer is used, E1 widens them to FP64
  int c32;
  double* b;    // out [ncol][k]
  int* nloc;    // out [ncol]
  // per-column setup records [ncol][96] of the Gram (window rows, row table,
  // reference point), from launch_column_setup
  int* crec;
  // work-queue mode: persistent waves take columns order[0], order[1], ...
  // through the counter `queue` (zeroed before the launch)
  const uint32_t* order;
  uint32_t* queue;
  int nblocks;  // launch size (0: one wave per column)
  // GEOGRAPHIC boundary list of the pass: (column, observation, chord
  // decision) of candidates within shell_width of the threshold, decided
  // exactly by launch_letkf_shell_fixup; its counter is zeroed before the Gram
  int4* shell;
  uint32_t* shell_count;
  int64_t shell_cap;
  // Global index of column 0.  The per-column arrays (col_a/b/c, C, b, nloc)
  // start at that column; col0 only enters where the global index itself
  // matters (grid coordinates of a GRID geometry without coordinates).
  int64_t col0;
};
// setup records of columns [0, a.ncol) for the Gram kernels (a.crec),
// LETKF_CREC_INTS ints per column
constexpr int LETKF_CREC_INTS = 96;
void launch_column_setup(hipStream_t s, const letkf_gram_args& a);
void launch_letkf_gram(hipStream_t s, const letkf_gram_args& a);
// exact decisions of the pass's GEOGRAPHIC boundary list (no-op otherwise)
void launch_letkf_shell_fixup(hipStream_t s, const letkf_gram_args& a, int* status);
// per-column cost estimate (observations in the candidate buckets) as
// descending-order sort keys, and column ids
void launch_column_cost(hipStream_t s, const letkf_gram_args& a, uint32_t* keys,
                        uint32_t* cols);
// learned-inflation inputs: per-observation background variance of the
// sorted Y', then the 8 local statistics of every column
void launch_obs_spread(hipStream_t s, const double* s_yp, int64_t p, int k, double* sp);
void launch_letkf_features(hipStream_t s, const letkf_gram_args& a, double* sp,
                           const double* X, int64_t nlev, const double* rho_prev, float* feat);
// Lorenz-96 forecast of nstates states of size N in place (scheme 0 RK4, 1 Euler)
void launch_l96(hipStream_t s, double* X, int64_t nstates, int64_t N, double F, double dt,
                int nsteps, int scheme);
// learned inflation network
void launch_mlp_forward(hipStream_t s, const float* params, const float* feat, int64_t m,
                        float* rho);
void launch_mlp_loss_grad(mda_ctx* c, hipStream_t s, const float* params, const float* feat,
                          const float* target, int64_t m, float* grad, double* loss);
void launch_mlp_adam(hipStream_t s, float* params, float* m1, float* m2, const float* grad,
                     int64_t step, float lr, float b1, float b2, float eps);
struct letkf_solve_args {
  int k;
  int64_t ncol;
  const double* C;
  int c32;  // C holds FP32 entries (letkf_gram_args::c32)
  const double* b;
  const int* nloc;
  double rho;
  const double* rho_col;
  int inflation_form, sqrt_form, combine;
  double* T;
  int* sweeps;
  int* status;
  // tridiagonal eigensolver workspace (symmetric square root, k <= 64)
  double* de;    // [ncol][2k]
  double* lam;   // [ncol][k]
  double* U;     // [ncol][k][k] Householder reflectors
  int* flag;     // [ncol] columns re-solved by the Jacobi fallback
  // Jacobi in global memory (k too large for LDS): per-block scratch of
  // 2 kp (kp + 1) doubles for jblocks blocks
  double* jscratch;
  int jblocks;
  // fallback work list [ncol] and its length (set: only the flagged and empty
  // columns are visited after the tridiagonal eigensolver)
  int* wlist;
  unsigned* wcount;
  int method;    // 0: tridiagonal + Jacobi fallback, 1: Jacobi only
  hipEvent_t ev_stage[3];  // recorded after E1, E2, E3 when set (timing)
  // Reduced-precision Gram: eigenvalues within null_rel * ||A|| of (k-1)/rho
  // are treated as the null space of Y'^T R^-1 Y' (0 for FP64)
  double null_rel;
  // Fused eigenvectors + transform (tridiagonal path, k in {16, 32, 48, 64},
  // matrix combination, FP64 transform): E3 applies X_a = X T' to the pass's
  // columns [x_col0, x_col0 + ncol) of X (x_ncol columns of x_nlev levels)
  // and writes T only for the columns it leaves to the fallback (flag != 0)
  // and for the empty ones.
  int fuse;
  double* X;
  int64_t x_ncol, x_nlev, x_col0;
};
// true when the tridiagonal eigensolver can apply the transform itself
bool letkf_eig_fusable(int k);
void launch_letkf_solve(hipStream_t s, const letkf_solve_args& a);
// ensemble sizes with a compiled tridiagonal eigensolver (others use Jacobi)
bool letkf_eig_supported(int k);
void launch_letkf_eig(hipStream_t s, const letkf_solve_args& a, double* de, double* lam,
                      int* flag);
// X_a = xbar + X' T for the columns [col0, col0 + nc) of X (ncol columns in
// all; nc < 0: every column); T holds the nc columns' transforms.  Skipped
// when the status word holds a MDA_ST_SKIP_UPDATE bit.  MDA_PREC_BF16X3
// runs X' T as split-precision bf16 products with FP32 accumulation (k a
// multiple of 16, at most 64); otherwise FP64.  `only` ([nc], FP64 MFMA path
// only) restricts the update to the columns with only[g] != 0.
void launch_letkf_transform(hipStream_t s, double* X, int k, int64_t ncol, int64_t nlev,
                            const double* T, int* status, int64_t col0 = 0, int64_t nc = -1,
                            int precision = MDA_PREC_FP64, const int* only = nullptr);
void launch_reduce_int(hipStream_t s, const int* v, int64_t n, int64_t* out2 /* sum,max */,
                       int64_t* out_empty);

// runs f with the context's device current and converts its exceptions into
// MDA_E_* codes (message in ctx->err)
int guarded_call(mda_ctx* ctx, const std::function<void()>& f);

// ETKF (ETKF.hpp:100-178) on device buffers: Y' [p][k], d [p], Gram
// partials [ETKF_GG_BLOCKS][k][k+1], [E | b] [k][k+1], T [k][k], Cholesky
// scratch [2][k][k] and the status word.  etkf_run takes them from the
// context workspace when `bufs` is null.
constexpr int ETKF_GG_BLOCKS = 256;
struct etkf_bufs {
  double *yp, *dd, *part, *EB, *T, *scr;
  int* status;
};
void etkf_run(mda_ctx* c, double inflation, mda_ensemble* ens, const mda_obs* obs,
              const mda_hop* hop, const etkf_bufs* bufs);

// IDW4 observation operator on the device (kernels_hop.hip): 4 taps per
// row at row_ptr[o] (rows of invalid observations are empty and skipped)
void launch_idw4_stencil(hipStream_t s, int64_t p, int coord, const double* oa, const double* ob,
                         const double* oc, const double* var, int64_t nx, int64_t ny,
                         int64_t nlev, const double* clat, const double* clon,
                         const double* vcoord, const int64_t* row_ptr, int64_t* col,
                         double* weight);

// dense p x p algebra of the global filters (p <= MDA_DENSE_MAX)
constexpr int64_t MDA_DENSE_MAX = 4096;
struct dense_extremes {
  double lmin, lmax;      // smallest and largest eigenvalue
  double absmin, absmax;  // smallest and largest |eigenvalue| (singular values)
};
void launch_form_enkf_S(hipStream_t s, const double* yp, const double* var, int k, int64_t p,
                        double* S);
void dense_tridiagonalise(hipStream_t s, int64_t p, double* A, double* d, double* e, double* u,
                          double* pv, double* w, double* sc);
dense_extremes tridiag_extremes(const std::vector<double>& d, const std::vector<double>& e);
// extreme eigenvalues of the symmetric device matrix A (destroyed)
dense_extremes dense_sym_extremes(mda_ctx* c, double* A, int64_t p);
// cond(S) of S = Y'Y'^T/(k-1) + diag(var); NaN beyond MDA_DENSE_MAX
double dense_spd_condition(mda_ctx* c, const double* yp, const double* var, int k, int64_t p);
// B <- A^-1 B by Gauss-Jordan with partial pivoting (A destroyed); a zero
// pivot sets MDA_ST_NOT_SPD in *status
void dense_gauss_jordan(hipStream_t s, int64_t p, int64_t m, double* A, double* B,
                        int64_t* piv, double* pval, int* status);

// LWEnKF (kernels_lw.hip)
void launch_lw_member_sumsq(hipStream_t s, const double* X, const double* mean, int k, int64_t n,
                            double* partial, int nblocks);
void launch_lw_chi2(hipStream_t s, const double* yo, const double* ybar, const double* yp_om,
                    const double* var, int k, int64_t p, double* chi2);
void launch_lw_form_S(hipStream_t s, const double* yp_om, const double* w, const double* var,
                      int k, int64_t p, double radius, int fn, int gc_form, double* S,
                      double* St);
void launch_lw_innovations(hipStream_t s, const double* yo, const double* ybar,
                           const double* yp_om, const double* var, const double* Z, int k,
                           int64_t p, double* D);
size_t lw_update_blocks(int64_t n);
void launch_lw_update(hipStream_t s, double* X, const double* mean, double scale,
                      const double* Gt, const double* D, int k, int64_t n, int64_t p,
                      double radius, int fn, int gc_form, double* mm);
void launch_lw_flag_inf(hipStream_t s, const double* var, int64_t p, int* flag);

}  // namespace mda
