// C-ABI entry points (include/metada_amd.h).  Each one stages host inputs to
// the context's device workspace when needed, enqueues the kernels on the
// context's stream and converts every error into an MDA_E_* code.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <limits>

#include "mda_internal.hpp"

using namespace mda;

namespace {

template <class F>
int guarded(mda_ctx* ctx, F&& f) {
  if (!ctx) return MDA_E_ARG;
  try {
    device_guard dg(ctx->device);
    f();
    ctx->err.clear();
    return MDA_OK;
  } catch (const mda_error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MDA_E_DEVICE;
  }
}

struct dev_ensemble {
  double* X = nullptr;
  int k = 0;
  int64_t n = 0;
  bool staged = false;
};

dev_ensemble stage_ensemble(mda_ctx* c, const mda_ensemble* e) {
  MDA_REQUIRE(e != nullptr, "ensemble is null");
  MDA_REQUIRE(e->k >= 2, "ensemble size must be >= 2");
  MDA_REQUIRE(e->nx > 0 && e->ny > 0 && e->nlev > 0, "ensemble dimensions must be positive");
  dev_ensemble d;
  d.k = e->k;
  d.n = e->nx * e->ny * e->nlev;
  if (e->on_device) {
    MDA_REQUIRE(e->data != nullptr, "device ensemble needs data");
    d.X = e->data;
    return d;
  }
  d.X = c->get<double>("ens", (size_t)d.k * d.n);
  d.staged = true;
  if (e->data) {
    h2d(c, d.X, e->data, sizeof(double) * d.k * d.n);
  } else {
    MDA_REQUIRE(e->members != nullptr, "ensemble needs data or members");
    for (int j = 0; j < d.k; ++j) {
      MDA_REQUIRE(e->members[j] != nullptr, "member pointer is null");
      h2d(c, d.X + (size_t)j * d.n, e->members[j], sizeof(double) * d.n);
    }
  }
  return d;
}

void unstage_ensemble(mda_ctx* c, const mda_ensemble* e, const dev_ensemble& d) {
  if (!d.staged) return;
  if (e->data) {
    d2h(c, e->data, d.X, sizeof(double) * d.k * d.n);
  } else {
    for (int j = 0; j < d.k; ++j)
      d2h(c, e->members[j], d.X + (size_t)j * d.n, sizeof(double) * d.n);
  }
  MDA_HIP(hipStreamSynchronize(c->stream));
}

struct dev_obs {
  int64_t p = 0;
  const double* yo = nullptr;
  const double* var = nullptr;
  const double* a = nullptr;
  const double* b = nullptr;
  const double* c = nullptr;
};

dev_obs stage_obs(mda_ctx* c, const mda_obs* o, bool need_loc, bool allow_empty = false) {
  MDA_REQUIRE(o != nullptr, "observations are null");
  MDA_REQUIRE(o->p > 0 || (allow_empty && o->p == 0),
              "no observations (the global EnKF / ETKF need p >= 1)");
  if (o->p == 0) return dev_obs{};
  MDA_REQUIRE(o->value && o->variance, "observation values and variances are required");
  dev_obs d;
  d.p = o->p;
  bool od = o->on_device != 0;
  d.yo = to_device(c, "obs_value", o->value, o->p, od);
  d.var = to_device(c, "obs_var", o->variance, o->p, od);
  if (need_loc) {
    MDA_REQUIRE(o->loc_a && o->loc_b, "observation locations are required");
    d.a = to_device(c, "obs_a", o->loc_a, o->p, od);
    d.b = to_device(c, "obs_b", o->loc_b, o->p, od);
    d.c = to_device(c, "obs_c", o->loc_c, o->p, od);
  }
  return d;
}

struct dev_hop {
  const int64_t* row_ptr = nullptr;
  const int64_t* col = nullptr;
  const double* w = nullptr;
  const double* hx = nullptr;
};

dev_hop stage_hop(mda_ctx* c, const mda_hop* h, int k, int64_t p, int64_t n) {
  MDA_REQUIRE(h != nullptr, "observation operator is null");
  dev_hop d;
  bool od = h->on_device != 0;
  if (h->kind == MDA_HOP_HX) {
    MDA_REQUIRE(h->hx != nullptr, "HX operator needs hx");
    d.hx = to_device(c, "hop_hx", h->hx, (size_t)k * p, od);
    return d;
  }
  MDA_REQUIRE(h->kind == MDA_HOP_CSR, "unknown observation operator kind");
  MDA_REQUIRE(h->row_ptr && h->col && h->weight, "CSR operator needs row_ptr, col, weight");
  int64_t nnz = h->nnz;
  if (!od) {
    nnz = h->row_ptr[p] - h->row_ptr[0];
    MDA_REQUIRE(h->row_ptr[0] == 0, "CSR row_ptr must start at 0");
    for (int64_t t = 0; t < nnz; ++t)
      MDA_REQUIRE(h->col[t] >= 0 && h->col[t] < n, "CSR column index out of range");
  }
  MDA_REQUIRE(nnz >= 0, "CSR nnz must be given for device operators");
  d.row_ptr = to_device(c, "hop_rp", h->row_ptr, p + 1, od);
  d.col = to_device(c, "hop_col", h->col, nnz, od);
  d.w = to_device(c, "hop_w", h->weight, nnz, od);
  return d;
}

double host_sum(mda_ctx* c, const double* dev, int n) {
  std::vector<double> h(n);
  d2h(c, h.data(), dev, sizeof(double) * n);
  MDA_HIP(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// Observation values must be finite and variances > 0 (+inf marks an invalid
// observation, GridObservation.hpp:239-252); a violation sets
// MDA_ST_NONFINITE_OBS in the status word, so the update is skipped.
__global__ void k_check_obs(const double* __restrict__ yo, const double* __restrict__ var,
                            int64_t p, int* __restrict__ status) {
  bool bad = false;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < p;
       o += (int64_t)gridDim.x * blockDim.x)
    bad = bad || !isfinite(yo[o]) || !(var[o] > 0.0);
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(status, MDA_ST_NONFINITE_OBS);
}

void launch_check_obs(hipStream_t s, const dev_obs& o, int* status) {
  if (o.p <= 0) return;
  hipLaunchKernelGGL(k_check_obs, dim3(grid_for(o.p, 256, 2048)), dim3(256), 0, s, o.yo, o.var,
                     o.p, status);
  MDA_LAUNCH_CHECK();
}

// Rethrows the failure of a nested entry point (its message is in ctx->err).
void check_rc(mda_ctx* c, int rc) {
  if (rc != MDA_OK) throw mda_error(rc, c->err);
}

// The analysis's status word as an error (after the stream has drained).
void throw_status(int st, const char* what) {
  if (st & MDA_ST_NONFINITE_OBS)
    throw mda_error(MDA_E_NUMERIC,
                    std::string(what) + ": non-finite observation value, H(x) or innovation, "
                    "or an observation variance that is not > 0 (ensemble left unchanged)");
  if (st & MDA_ST_BOUNDARY_OVERFLOW)
    throw mda_error(MDA_E_UNSUPPORTED,
                    std::string(what) + ": more column-observation pairs at the localisation "
                    "radius (within 1e-9 relative) than the boundary list holds "
                    "(ensemble left unchanged)");
  if (st & MDA_ST_NOT_SPD)
    throw mda_error(MDA_E_NUMERIC, std::string(what) +
                                       ": a matrix to factor is not positive definite "
                                       "(ensemble left unchanged)");
  if (st & MDA_ST_NONFINITE_STATE)
    throw mda_error(MDA_E_NUMERIC,
                    std::string(what) + ": non-finite background values; the analysis is "
                    "non-finite at those state points");
}

}  // namespace

extern "C" {

int mda_ctx_create(int device, mda_ctx** out) {
  if (!out) return MDA_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MDA_E_DEVICE;
  if (device < 0 || device >= ndev) return MDA_E_ARG;
  mda_ctx* c = new mda_ctx();
  c->device = device;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MDA_E_DEVICE;
  }
  for (auto& e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return MDA_E_DEVICE;
    }
  }
  (void)hipSetDevice(prev);
  *out = c;
  return MDA_OK;
}

int mda_ctx_destroy(mda_ctx* c) {
  if (!c) return MDA_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->ws)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& kv : c->graphs)
    if (kv.second) (void)hipGraphExecDestroy(kv.second);
  for (auto st : c->side) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  for (auto e : c->fj)
    if (e) (void)hipEventDestroy(e);
  for (auto e : c->evpool)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  (void)hipSetDevice(prev);
  delete c;
  return MDA_OK;
}

const char* mda_last_error(const mda_ctx* c) { return c ? c->err.c_str() : mda_last_io_error(); }

int mda_ctx_synchronize(mda_ctx* c) {
  return guarded(c, [&] { MDA_HIP(hipStreamSynchronize(c->stream)); });
}

void* mda_ctx_stream(mda_ctx* c) { return c ? (void*)c->stream : nullptr; }

int64_t mda_ctx_workspace_bytes(const mda_ctx* c) {
  if (!c) return 0;
  int64_t s = 0;
  for (auto& kv : c->ws) s += (int64_t)kv.second.bytes;
  return s;
}

// ---------------------------------------------------------------------------
int mda_ensemble_mean_perturbations(mda_ctx* c, const mda_ensemble* ens, double scale,
                                    double* mean, double* pert, int32_t out_on_device) {
  return guarded(c, [&] {
    dev_ensemble d = stage_ensemble(c, ens);
    double* dm = mean ? (out_on_device ? mean : c->get<double>("out_mean", d.n)) : nullptr;
    double* dp = pert ? (out_on_device ? pert : c->get<double>("out_pert", d.k * d.n)) : nullptr;
    launch_mean_pert(c->stream, d.X, d.k, d.n, scale, dm, dp);
    if (!out_on_device) {
      if (mean) d2h(c, mean, dm, sizeof(double) * d.n);
      if (pert)
        d2h(c, pert, dp, sizeof(double) * d.k * d.n);
    }
    MDA_HIP(hipStreamSynchronize(c->stream));
  });
}

int mda_obs_innovations(mda_ctx* c, const mda_ensemble* ens, const mda_obs* obs,
                        const mda_hop* hop, double* ybar, double* yp, double* d,
                        int32_t out_on_device) {
  return guarded(c, [&] {
    dev_ensemble de = stage_ensemble(c, ens);
    dev_obs o = stage_obs(c, obs, false);
    dev_hop h = stage_hop(c, hop, de.k, o.p, de.n);
    const int64_t p = o.p;
    double* dyb = ybar ? (out_on_device ? ybar : c->get<double>("out_ybar", p)) : nullptr;
    double* dyp = yp ? (out_on_device ? yp : c->get<double>("out_yp", de.k * p)) : nullptr;
    double* dd = d ? (out_on_device ? d : c->get<double>("out_d", p)) : nullptr;
    launch_obs_stats(c->stream, de.X, de.k, de.n, p, h.row_ptr, h.col, h.w, h.hx, o.yo, dyb,
                     nullptr, dd, dyp);
    if (!out_on_device) {
      if (ybar) d2h(c, ybar, dyb, sizeof(double) * p);
      if (yp) d2h(c, yp, dyp, sizeof(double) * de.k * p);
      if (d) d2h(c, d, dd, sizeof(double) * p);
    }
    MDA_HIP(hipStreamSynchronize(c->stream));
  });
}

// ---------------------------------------------------------------------------
// cond(S) of the EnKF, S = Y'Y'^T/(k-1) + diag(var) (EnKF.hpp:205-209,
// JacobiSVD singular values; S is SPD, so they are its eigenvalues):
//   p <= 128                  Jacobi on the dense S (one workgroup);
//   constant var, p > k - 1   S = s^2 (I + Y'Y'^T/(s^2 (k-1))): the smallest
//                             eigenvalue is s^2 (rank deficiency) and the
//                             largest s^2 (1 + mu/(k-1)), mu the largest
//                             eigenvalue of E = Y'^T R^-1 Y' (from [E | b]);
//   otherwise                 Householder tridiagonalisation of the dense S
//                             and bisection for its extreme eigenvalues
//                             (p <= MDA_DENSE_MAX).
static double enkf_condition_number(mda_ctx* c, const double* yp, const double* var, int k,
                                    int64_t p, const double* EB) {
  hipStream_t s = c->stream;
  double* cnd = c->get<double>("enkf_cond", 1);
  std::vector<double> hv(p);
  d2h(c, hv.data(), var, sizeof(double) * p);
  // an invalid observation (var = +inf) is an infinite singular value of S
  if (std::any_of(hv.begin(), hv.end(), [](double v) { return std::isinf(v); }))
    return std::numeric_limits<double>::infinity();
  if (p <= 128) {
    double* S = c->get<double>("enkf_S", (size_t)130 * 130);
    launch_cond_small(s, yp, var, k, (int)p, S, cnd);
    double v = 0.0;
    d2h(c, &v, cnd, sizeof(double));
    return v;
  }
  const bool constant = std::all_of(hv.begin(), hv.end(), [&](double v) { return v == hv[0]; });
  if (constant) {
    double* S = c->get<double>("enkf_S", (size_t)130 * 130);
    launch_sym_maxeig(s, EB, k, S, cnd);
    double mu = 0.0;
    d2h(c, &mu, cnd, sizeof(double));
    return 1.0 + mu / (double)(k - 1);
  }
  return dense_spd_condition(c, yp, var, k, p);
}

}  // extern "C"

namespace mda {

int guarded_call(mda_ctx* ctx, const std::function<void()>& f) { return guarded(ctx, f); }

// ETKF (ETKF.hpp:100-178): Y' and d, the global Gram [E | b], the k x k
// solve and the update, on the device buffers `b` (from the context
// workspace when null).
void etkf_run(mda_ctx* c, double inflation, mda_ensemble* ens, const mda_obs* obs,
              const mda_hop* hop, const etkf_bufs* given) {
  MDA_REQUIRE(std::isfinite(inflation), "inflation must be finite");
  dev_ensemble de = stage_ensemble(c, ens);
  dev_obs o = stage_obs(c, obs, false);
  dev_hop h = stage_hop(c, hop, de.k, o.p, de.n);
  const int k = de.k;
  MDA_REQUIRE(k <= 128, "ETKF supports k <= 128");
  const int64_t p = o.p;
  etkf_bufs b;
  if (given) {
    b = *given;
  } else {
    b.yp = c->get<double>("yp_om", (size_t)p * k);
    b.dd = c->get<double>("d", p);
    b.part = c->get<double>("gg_part", (size_t)ETKF_GG_BLOCKS * k * (k + 1));
    b.EB = c->get<double>("gg_EB", (size_t)k * (k + 1));
    b.T = c->get<double>("gT", (size_t)k * k);
    b.scr = c->get<double>("g_scr", (size_t)2 * k * k);
    b.status = c->get<int>("g_status", 1);
  }
  hipStream_t s = c->stream;
  launch_zero(s, b.status, sizeof(int));
  launch_check_obs(s, o, b.status);
  launch_obs_stats(s, de.X, k, de.n, p, h.row_ptr, h.col, h.w, h.hx, o.yo, nullptr, b.yp, b.dd,
                   nullptr, nullptr, nullptr, nullptr, b.status);
  launch_global_gram(s, b.yp, b.dd, o.var, nullptr, k, p, 0, b.part, ETKF_GG_BLOCKS, b.EB);
  launch_etkf_solve_ws(s, k, b.EB, b.T, b.scr, b.status);
  launch_transform_global(s, de.X, k, de.n, inflation, b.T, b.status);
  int hs = 0;
  MDA_HIP(hipMemcpyAsync(&hs, b.status, sizeof(int), hipMemcpyDeviceToHost, s));
  MDA_HIP(hipStreamSynchronize(s));
  throw_status(hs, "ETKF");
  unstage_ensemble(c, ens, de);
}

}  // namespace mda

extern "C" {

int mda_etkf_analyse(mda_ctx* c, double inflation, mda_ensemble* ens, const mda_obs* obs,
                     const mda_hop* hop) {
  return guarded(c, [&] { etkf_run(c, inflation, ens, obs, hop, nullptr); });
}

int mda_enkf_analyse(mda_ctx* c, double inflation, int32_t method, mda_ensemble* ens,
                     const mda_obs* obs, const mda_hop* hop, const double* Z,
                     int32_t z_on_device, mda_enkf_diag* diag) {
  return guarded(c, [&] {
    MDA_REQUIRE(std::isfinite(inflation) && inflation >= 0.0, "inflation must be >= 0");
    MDA_REQUIRE(Z != nullptr, "EnKF needs the N(0,1) draws Z");
    dev_ensemble de = stage_ensemble(c, ens);
    dev_obs o = stage_obs(c, obs, false);
    dev_hop h = stage_hop(c, hop, de.k, o.p, de.n);
    const int k = de.k;
    const int64_t p = o.p, n = de.n;
    MDA_REQUIRE(k <= 128, "EnKF supports k <= 128 in this build");
    hipStream_t s = c->stream;
    const double* dZ = to_device(c, "enkf_Z", Z, (size_t)p * k, z_on_device != 0);
    // all inflation methods are multiplicative sqrt(rho) (EnKF.hpp:316-331)
    const double scale = std::sqrt(inflation);
    double* yp = c->get<double>("yp_om", (size_t)p * k);
    double* dd = c->get<double>("d", p);
    int* st = c->get<int>("g_status", 1);
    launch_zero(s, st, sizeof(int));
    launch_check_obs(s, o, st);
    launch_obs_stats(s, de.X, k, n, p, h.row_ptr, h.col, h.w, h.hx, o.yo, nullptr, yp, dd,
                     nullptr, nullptr, nullptr, nullptr, st);
    const int nb = 256;
    double* part = c->get<double>("gg_part", (size_t)nb * k * (k + 1));
    double* EB = c->get<double>("gg_EB", (size_t)k * (k + 1));
    double* F = c->get<double>("gg_F", (size_t)k * k);
    launch_global_gram(s, yp, dd, o.var, nullptr, k, p, 0, part, nb, EB);
    launch_global_gram(s, yp, dd, o.var, dZ, k, p, 1, part, nb, F);
    double* T = c->get<double>("gT", (size_t)k * k);
    double* Minv = c->get<double>("gMinv", (size_t)k * k);
    double* scr = c->get<double>("g_scr", (size_t)2 * k * k);
    launch_enkf_solve_ws(s, k, EB, F, T, Minv, scr, st);
    {
      // a failed factorisation or non-finite input ends the call before any
      // diagnostics or update
      int hs0 = 0;
      d2h(c, &hs0, st, sizeof(int));
      throw_status(hs0 & MDA_ST_SKIP_UPDATE, "EnKF");
    }
    mda_enkf_diag dg{};
    dg.ensemble_size = k;
    dg.observation_count = p;
    dg.inflation_method = method;
    dg.inflation_factor = inflation;
    dg.analysis_increment_norm = 0.0;
    double* part2 = c->get<double>("sumsq_part", 1024);
    if (diag) {
      launch_dev_sumsq(s, de.X, k, n, scale, part2, 512);
      dg.background_spread = std::sqrt(host_sum(c, part2, 512) / ((double)n * k));
      launch_sumsq(s, dd, p, part2, 512);
      dg.innovation_norm = std::sqrt(host_sum(c, part2, 512));
      // max / min Kalman gain over the n x p entries of K (EnKF.hpp:202-203);
      // O(n p k) work, as in the reference
      {
        double* G = c->get<double>("enkf_G", (size_t)p * k);
        double* mm = c->get<double>("enkf_mm", 2 * 512);
        launch_enkf_gain_minmax_ws(s, de.X, k, n, scale, Minv, yp, o.var, p, G, mm, 512);
        std::vector<double> hm(2 * 512);
        d2h(c, hm.data(), mm, sizeof(double) * 1024);
        MDA_HIP(hipStreamSynchronize(s));
        double mx = -std::numeric_limits<double>::infinity(), mn = -mx;
        for (int b = 0; b < 512; ++b) { mx = std::max(mx, hm[2 * b]); mn = std::min(mn, hm[2 * b + 1]); }
        dg.max_kalman_gain = mx;
        dg.min_kalman_gain = mn;
      }
      // cond(S) = ratio of the extreme singular values of the SPD
      // S = Y'Y'^T/(k-1) + R (EnKF.hpp:205-209)
      dg.condition_number = enkf_condition_number(c, yp, o.var, k, p, EB);
    }
    launch_transform_global(s, de.X, k, n, scale, T, st);
    if (diag) {
      launch_dev_sumsq(s, de.X, k, n, 1.0, part2, 512);
      dg.analysis_spread = std::sqrt(host_sum(c, part2, 512) / ((double)n * k));
    }
    int hs = 0;
    d2h(c, &hs, st, sizeof(int));
    MDA_HIP(hipStreamSynchronize(s));
    unstage_ensemble(c, ens, de);
    if (diag) *diag = dg;
    throw_status(hs, "EnKF");
  });
}

// ---------------------------------------------------------------------------
// LWEnKF (LWEnKF.hpp:207-334).  Member weights are k numbers formed on the
// host from device reductions in the reference's order (:413-500): the
// adaptive and inverse-variance forms normalise twice, as the reference does.
// ---------------------------------------------------------------------------
namespace {

std::vector<double> lw_weights(int scheme, const std::vector<double>& sumsq,
                               const std::vector<double>& chi2, int k) {
  std::vector<double> w(k);
  if (scheme == MDA_LW_W_ADAPTIVE || scheme == MDA_LW_W_INVERSE_VAR) {
    double sum_inv = 0.0;
    for (int i = 0; i < k; ++i) {
      const double m = scheme == MDA_LW_W_ADAPTIVE ? std::sqrt(sumsq[i]) : sumsq[i];
      w[i] = 1.0 / (m + 1e-8);
      sum_inv += w[i];
    }
    for (int i = 0; i < k; ++i) w[i] /= sum_inv;
  } else if (scheme == MDA_LW_W_LIKELIHOOD) {
    for (int i = 0; i < k; ++i) w[i] = std::exp(-0.5 * chi2[i]);
  } else {
    for (int i = 0; i < k; ++i) w[i] = 1.0 / k;
  }
  double sum = 0.0;
  for (int i = 0; i < k; ++i) sum += w[i];
  if (!(sum > 0.0) || !std::isfinite(sum))
    throw mda_error(MDA_E_NUMERIC,
                    "LWEnKF: the member weights sum to zero or a non-finite value (likelihood "
                    "weights underflow when every member is far from the observations)");
  for (int i = 0; i < k; ++i) w[i] /= sum;
  return w;
}

}  // namespace

int mda_lwenkf_analyse(mda_ctx* c, const mda_lwenkf_params* prm, mda_ensemble* ens,
                       const mda_obs* obs, const mda_hop* hop, const double* Z,
                       int32_t z_on_device, double* weights, mda_lwenkf_diag* diag) {
  return guarded(c, [&] {
    MDA_REQUIRE(prm != nullptr, "LWEnKF parameters are required");
    MDA_REQUIRE(std::isfinite(prm->inflation) && prm->inflation >= 0.0, "inflation must be >= 0");
    MDA_REQUIRE(prm->localization_radius > 0.0, "localization radius must be > 0");
    MDA_REQUIRE(prm->localization >= MDA_LW_LOC_GAUSSIAN &&
                    prm->localization <= MDA_LW_LOC_GASPARI_COHN,
                "unknown LWEnKF localization function");
    MDA_REQUIRE(prm->gc_form == MDA_LW_GC_REFERENCE || prm->gc_form == MDA_LW_GC_STANDARD,
                "unknown Gaspari-Cohn form");
    MDA_REQUIRE(prm->weighting >= MDA_LW_W_UNIFORM && prm->weighting <= MDA_LW_W_LIKELIHOOD,
                "unknown LWEnKF weighting scheme");
    MDA_REQUIRE(Z != nullptr, "LWEnKF needs the N(0,1) draws Z");
    dev_ensemble de = stage_ensemble(c, ens);
    dev_obs o = stage_obs(c, obs, false);
    dev_hop h = stage_hop(c, hop, de.k, o.p, de.n);
    const int k = de.k;
    const int64_t p = o.p, n = de.n;
    MDA_REQUIRE(k >= 2 && k <= 128, "LWEnKF supports 2 <= k <= 128");
    MDA_REQUIRE(p <= MDA_DENSE_MAX, "LWEnKF supports p <= 4096 (dense S)");
    hipStream_t s = c->stream;
    const double* dZ = to_device(c, "lw_Z", Z, (size_t)p * k, z_on_device != 0);
    const double scale = std::sqrt(prm->inflation);
    const double radius = prm->localization_radius;
    int* st = c->get<int>("lw_status", 2);  // [0] status word, [1] invalid-observation flag
    launch_zero(s, st, 2 * sizeof(int));
    launch_check_obs(s, o, st);
    launch_lw_flag_inf(s, o.var, p, st + 1);
    // forecast statistics (RecomputeMean / RecomputePerturbations, :215-229)
    double* mean = c->get<double>("lw_mean", n);
    launch_mean_pert(s, de.X, k, n, 1.0, mean, nullptr);
    double* ybar = c->get<double>("lw_ybar", p);
    double* yp = c->get<double>("lw_yp", (size_t)p * k);
    double* dd = c->get<double>("lw_d", p);
    launch_obs_stats(s, de.X, k, n, p, h.row_ptr, h.col, h.w, h.hx, o.yo, ybar, yp, dd, nullptr,
                     nullptr, nullptr, nullptr, st);
    int hst[2] = {0, 0};
    d2h(c, hst, st, sizeof(hst));
    MDA_HIP(hipStreamSynchronize(s));
    if (hst[1])
      throw mda_error(MDA_E_ARG, "LWEnKF: invalid observations (+inf variance) are not supported");
    throw_status(hst[0] & MDA_ST_SKIP_UPDATE, "LWEnKF");
    // member weights (:413-430)
    std::vector<double> sumsq(k, 0.0), chi2(k, 0.0);
    if (prm->weighting == MDA_LW_W_ADAPTIVE || prm->weighting == MDA_LW_W_INVERSE_VAR) {
      constexpr int NB = 64;
      double* part = c->get<double>("lw_part", (size_t)NB * k);
      launch_lw_member_sumsq(s, de.X, mean, k, n, part, NB);
      std::vector<double> hp((size_t)NB * k);
      d2h(c, hp.data(), part, sizeof(double) * NB * k);
      MDA_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < k; ++i)
        for (int b = 0; b < NB; ++b) sumsq[i] += hp[(size_t)i * NB + b];
    } else if (prm->weighting == MDA_LW_W_LIKELIHOOD) {
      double* dchi = c->get<double>("lw_chi2", k);
      launch_lw_chi2(s, o.yo, ybar, yp, o.var, k, p, dchi);
      d2h(c, chi2.data(), dchi, sizeof(double) * k);
      MDA_HIP(hipStreamSynchronize(s));
    }
    const std::vector<double> w = lw_weights(prm->weighting, sumsq, chi2, k);
    double* dw = c->get<double>("lw_w", k);
    h2d(c, dw, w.data(), sizeof(double) * k);
    // perturbed-observation innovations (:286-301) before Y' is overwritten
    double* D = c->get<double>("lw_D", (size_t)p * k);
    launch_lw_innovations(s, o.yo, ybar, yp, o.var, dZ, k, p, D);
    // S = L_o o Pf + R and its transpose (:258-270)
    double* S = c->get<double>("lw_S", (size_t)p * p);
    double* St = c->get<double>("lw_St", (size_t)p * p);
    launch_lw_form_S(s, yp, dw, o.var, k, p, radius, prm->localization, prm->gc_form, S, St);
    mda_lwenkf_diag dg{};
    if (diag) {
      // cond(S) = sigma_max / sigma_min (:280-284); S is symmetric, so its
      // singular values are the |eigenvalues|
      const dense_extremes x = dense_sym_extremes(c, S, p);
      dg.condition_number = x.absmax / x.absmin;
    }
    // G^T = S^-T Y' (p x k), in place of Y'
    int64_t* piv = c->get<int64_t>("lw_piv", 1);
    double* pval = c->get<double>("lw_pval", 1);
    MDA_HIP(hipEventRecord(c->ev[0], s));
    dense_gauss_jordan(s, p, k, St, yp, piv, pval, st);
    MDA_HIP(hipEventRecord(c->ev[1], s));
    double* part2 = c->get<double>("sumsq_part", 1024);
    if (diag) {
      launch_dev_sumsq(s, de.X, k, n, scale, part2, 512);
      dg.background_spread = std::sqrt(host_sum(c, part2, 512) / ((double)n * k));
      launch_sumsq(s, dd, p, part2, 512);
      dg.innovation_norm = std::sqrt(host_sum(c, part2, 512));
    }
    d2h(c, hst, st, sizeof(int));
    MDA_HIP(hipStreamSynchronize(s));
    if (hst[0] & MDA_ST_NOT_SPD)
      throw mda_error(MDA_E_NUMERIC, "LWEnKF: S = L o Pf + R is singular");
    // gain and update, fused (:272-301)
    const size_t nbu = lw_update_blocks(n);
    double* mm = c->get<double>("lw_mm", 2 * nbu);
    MDA_HIP(hipEventRecord(c->ev[2], s));
    launch_lw_update(s, de.X, mean, scale, yp, D, k, n, p, radius, prm->localization,
                     prm->gc_form, mm);
    MDA_HIP(hipEventRecord(c->ev[3], s));
    std::vector<double> hm(2 * nbu);
    d2h(c, hm.data(), mm, sizeof(double) * 2 * nbu);
    MDA_HIP(hipStreamSynchronize(s));
    double mx = -std::numeric_limits<double>::infinity(), mn = -mx;
    for (size_t b = 0; b < nbu; ++b) {
      mx = std::max(mx, hm[2 * b]);
      mn = std::min(mn, hm[2 * b + 1]);
    }
    if (diag) {
      launch_dev_sumsq(s, de.X, k, n, 1.0, part2, 512);
      dg.analysis_spread = std::sqrt(host_sum(c, part2, 512) / ((double)n * k));
      dg.max_kalman_gain = mx;
      dg.min_kalman_gain = mn;
      dg.localization_radius = radius;
      dg.max_weight = *std::max_element(w.begin(), w.end());
      dg.min_weight = *std::min_element(w.begin(), w.end());
      const double mw = 1.0 / (double)k;
      double wv = 0.0;
      for (double v : w) wv += (v - mw) * (v - mw);
      dg.weight_variance = wv / (double)k;
      dg.ensemble_size = k;
      dg.observation_count = p;
      dg.inflation_method = prm->inflation_method;
      dg.localization = prm->localization;
      dg.weighting = prm->weighting;
      dg.inflation_factor = prm->inflation;
      dg.analysis_increment_norm = 0.0;
      dg.ms_gain_solve = elapsed(c->ev[0], c->ev[1]);
      dg.ms_update = elapsed(c->ev[2], c->ev[3]);
    }
    unstage_ensemble(c, ens, de);
    if (weights) std::copy(w.begin(), w.end(), weights);
    if (diag) *diag = dg;
  });
}

// ---------------------------------------------------------------------------
void mda_letkf_params_standard(mda_letkf_params* p, double inflation, double radius) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->inflation = inflation;
  p->radius = radius;
  p->localization = MDA_LOC_GASPARI_COHN;
  p->rweight = MDA_RW_REAL;
  p->inflation_form = MDA_INFLFORM_PRIOR;
  p->sqrt_form = MDA_SQRT_SYMMETRIC;
  p->combine = MDA_COMBINE_MATRIX;
  p->precision = MDA_PREC_FP64;
}

void mda_letkf_params_refcompat(mda_letkf_params* p, double inflation, double radius) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->inflation = inflation;
  p->radius = radius;
  p->localization = MDA_LOC_CUTOFF;
  p->rweight = MDA_RW_IDENTITY;
  p->inflation_form = MDA_INFLFORM_POSTERIOR;
  p->sqrt_form = MDA_SQRT_CHOLESKY;
  p->combine = MDA_COMBINE_ELEMENTWISE;
  p->precision = MDA_PREC_FP64;
}

// ---------------------------------------------------------------------------
// LETKF pipeline (LETKF.hpp:63-243), in three parts:
//   plan     host-side validation, staging of host inputs, the bucket grid and
//            every workspace buffer (may synchronise; nothing is enqueued);
//   enqueue  the kernels only, on a given stream (no host synchronisation or
//            allocation, so it can be captured into a hipGraph);
//   finish   device statistics to the host and the diagnostics.
// Workspace names carry a prefix so the analyses of one batch do not share
// buffers.
// ---------------------------------------------------------------------------
struct letkf_plan {
  const mda_letkf_params* prm = nullptr;
  mda_ensemble* ens = nullptr;
  const mda_geometry* geom = nullptr;
  dev_ensemble de;
  dev_obs o;
  dev_hop h;
  int k = 0;
  int64_t p = 0, n = 0, nx = 0, ny = 0, nlev = 0, ncol = 0;
  // columns per pass of Gram -> solve -> transform: the k x k per-column
  // buffers (C, U, T) hold one pass; ncol when the whole grid fits
  int64_t chunk = 0;
  const double *rcoord = nullptr, *ccoord = nullptr;
  bucket_grid bg{};
  int nbuckets = 0, key_bits = 1;
  int64_t hist_cap = 0;
  double *yp = nullptr, *dd = nullptr;
  uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr, *hist = nullptr;
  // processing order of the observation statistics (first stencil tap) and
  // each observation's position in bucket order
  uint32_t *tkeys = nullptr, *tvals = nullptr, *ipos = nullptr;
  int tap_bits = 0;
  int* bstart = nullptr;
  int geo = 0;
  letkf_gram_args ga{};
  letkf_solve_args sa{};
  int* sweeps = nullptr;
  int64_t* stats = nullptr;
  bool eig_path = false;
  // work-queue Gram: cost keys, column order and the counter
  bool queue = false;
  uint32_t *ckeys = nullptr, *ccols = nullptr, *ckeys2 = nullptr, *ccols2 = nullptr,
           *chist = nullptr, *qcount = nullptr;
  int64_t chist_cap = 0;
};

static void check_letkf_params(const mda_letkf_params* prm, const mda_geometry* geom) {
  MDA_REQUIRE(prm && geom, "params and geometry are required");
  MDA_REQUIRE(std::isfinite(prm->inflation) && prm->inflation > 0.0, "inflation must be > 0");
  MDA_REQUIRE(prm->radius > 0.0, "radius must be > 0");
  MDA_REQUIRE(prm->precision == MDA_PREC_FP64 || prm->precision == MDA_PREC_FP32 ||
                  prm->precision == MDA_PREC_BF16X3,
              "precision must be FP64, FP32 or BF16X3");
  MDA_REQUIRE(prm->transform_precision == MDA_PREC_FP64 ||
                  prm->transform_precision == MDA_PREC_BF16X3,
              "transform_precision must be FP64 or BF16X3");
  MDA_REQUIRE(geom->coord >= 0 && geom->coord <= 2, "unknown coordinate system");
}

static letkf_plan plan_letkf(mda_ctx* c, const mda_letkf_params* prm, mda_ensemble* ens,
                             const mda_geometry* geom, const mda_obs* obs, const mda_hop* hop,
                             const std::string& pre, bool queue) {
  check_letkf_params(prm, geom);
  letkf_plan P;
  // argument blocks with zero padding: the batch's graph key holds their bytes
  std::memset(&P.ga, 0, sizeof(P.ga));
  std::memset(&P.sa, 0, sizeof(P.sa));
  std::memset(&P.bg, 0, sizeof(P.bg));
  P.prm = prm; P.ens = ens; P.geom = geom;
  P.de = stage_ensemble(c, ens);
  P.o = stage_obs(c, obs, true, /*allow_empty=*/true);
  if (P.o.p > 0) P.h = stage_hop(c, hop, P.de.k, P.o.p, P.de.n);
  const int k = P.k = P.de.k;
  const int64_t p = P.p = P.o.p;
  P.n = P.de.n; P.nx = ens->nx; P.ny = ens->ny; P.nlev = ens->nlev;
  const int64_t ncol = P.ncol = P.nx * P.ny;
  MDA_REQUIRE(k <= 128, "LETKF supports k <= 128 in this build");
  MDA_REQUIRE(prm->transform_precision != MDA_PREC_BF16X3 || (k % 16 == 0 && k <= 64),
              "a BF16x3 transform needs k a multiple of 16, at most 64");
  const bool gd = geom->on_device != 0;
  const double* ca = to_device(c, pre + "geom_a", geom->col_a, ncol, gd);
  const double* cb = to_device(c, pre + "geom_b", geom->col_b, ncol, gd);
  const double* cc = to_device(c, pre + "geom_c", geom->col_c, ncol, gd);
  MDA_REQUIRE(geom->coord == MDA_COORD_GRID || (ca && cb),
              "GEOGRAPHIC / CARTESIAN geometry needs column coordinates");
  const double* rho_col =
      to_device(c, pre + "rho_col", prm->rho_column, ncol, prm->rho_on_device != 0);
  if (P.h.hx) {  // input-order Y' and d before they move into bucket order
    P.yp = c->get<double>(pre + "yp_om", (size_t)p * k);
    P.dd = c->get<double>(pre + "d", p);
  }

  // bucket grid over the localisation coordinates
  bucket_grid& bg = P.bg;
  bg.coord = geom->coord;
  if (geom->coord == MDA_COORD_GEOGRAPHIC) {
    double ddeg = prm->radius / 6371.0 * 180.0 / 3.14159265358979323846;
    // a quarter of the radius: a window spans about 9 bucket rows (inside the
    // Gram's 16-row range table), each trimmed to the cap, and holds about
    // 1.3x the local observations (1.6x at half the radius); at least 1/16
    // degree: 2880 x 5760 buckets stay under the 2^24 key limit
    double step = std::min(10.0, std::max(0.0625, ddeg / 4.0));
    // bucket sizes divide the latitude and longitude ranges exactly
    bg.nrow = (int)std::ceil(180.0 / step);
    bg.ncolb = (int)std::ceil(360.0 / step);
    bg.a0 = -90.0; bg.b0 = 0.0; bg.da = 180.0 / bg.nrow; bg.db = 360.0 / bg.ncolb;
    bg.wrap = 1;
    P.rcoord = P.o.a; P.ccoord = P.o.b;
  } else {
    double xmin = 0, xmax = (double)(P.nx - 1), ymin = 0, ymax = (double)(P.ny - 1);
    if (geom->col_a) {
      std::vector<double> hx(ncol), hy(ncol);
      d2h(c, hx.data(), ca, sizeof(double) * ncol);
      d2h(c, hy.data(), cb, sizeof(double) * ncol);
      auto mx = std::minmax_element(hx.begin(), hx.end());
      auto my = std::minmax_element(hy.begin(), hy.end());
      xmin = *mx.first; xmax = *mx.second; ymin = *my.first; ymax = *my.second;
    }
    double r = std::min(prm->radius, 1e15);
    double x0 = xmin - r, x1 = xmax + r, y0 = ymin - r, y1 = ymax + r;
    double step = std::max(r / 2.0, std::max(x1 - x0, y1 - y0) / 2048.0);
    bg.a0 = y0; bg.b0 = x0; bg.da = step; bg.db = step; bg.wrap = 0;
    bg.nrow = std::max(1, (int)std::ceil((y1 - y0) / step) + 1);
    bg.ncolb = std::max(1, (int)std::ceil((x1 - x0) / step) + 1);
    P.rcoord = P.o.b; P.ccoord = P.o.a;
  }
  P.nbuckets = bg.nrow * bg.ncolb;
  MDA_REQUIRE(P.nbuckets > 0 && P.nbuckets < (1 << 24), "bucket grid too large");
  while ((1 << P.key_bits) < P.nbuckets) ++P.key_bits;
  P.keys = c->get<uint32_t>(pre + "bk_keys", p);
  P.vals = c->get<uint32_t>(pre + "bk_vals", p);
  P.keys2 = c->get<uint32_t>(pre + "bk_keys2", p);
  P.vals2 = c->get<uint32_t>(pre + "bk_vals2", p);
  P.hist_cap = ((p + 255) / 256) * 256;
  P.hist = c->get<uint32_t>(pre + "bk_hist", P.hist_cap);
  if (!P.h.hx && P.n < ((int64_t)1 << 32)) {
    P.tkeys = c->get<uint32_t>(pre + "tap_keys", p);
    P.tvals = c->get<uint32_t>(pre + "tap_vals", p);
    P.ipos = c->get<uint32_t>(pre + "tap_ipos", p);
    // sort keys below ((nlev - 1) / 4 + 1) nx ny (groups of four levels)
    const int64_t kmax = ((P.nlev + 3) / 4) * P.nx * P.ny;
    while (P.tap_bits < 32 && ((int64_t)1 << P.tap_bits) < kmax) ++P.tap_bits;
  }
  P.bstart = c->get<int>(pre + "bk_start", P.nbuckets + 1);

  letkf_gram_args& ga = P.ga;
  ga.k = k; ga.ncol = ncol; ga.nx = P.nx; ga.coord = geom->coord;
  ga.col_a = ca; ga.col_b = cb; ga.col_c = cc;
  ga.grid = bg; ga.bstart = P.bstart;
  ga.yp = c->get<double>(pre + "s_yp", (size_t)p * k);
  ga.d = c->get<double>(pre + "s_d", p);
  ga.rinv = c->get<double>(pre + "s_rinv", p);
  ga.oa = c->get<double>(pre + "s_a", p);
  ga.ob = c->get<double>(pre + "s_b", p);
  ga.oc = c->get<double>(pre + "s_c", p);
  ga.radius = prm->radius; ga.localization = prm->localization; ga.precision = prm->precision;
  P.geo = geom->coord == MDA_COORD_GEOGRAPHIC;
  if (prm->precision != MDA_PREC_FP64) ga.y32 = c->get<float>(pre + "s_y32", (size_t)p * k);
  if (P.geo) {  // unit vectors of the observation locations (chord distances)
    ga.ux = c->get<double>(pre + "s_ux", p);
    ga.uy = c->get<double>(pre + "s_uy", p);
    ga.uz = c->get<double>(pre + "s_uz", p);
  }
  // Column passes: at k > 64 the per-column k x k buffers of the whole grid
  // would not fit in HBM next to the ensemble (3 x 69 GB at k = 128 and
  // 524,288 columns), so the columns run in passes of 65,536.
  P.chunk = k > 64 ? std::min<int64_t>(ncol, 65536) : ncol;
  if (const char* ev = std::getenv("MDA_LETKF_CHUNK_COLS")) {
    const long long v = std::atoll(ev);
    if (v > 0) P.chunk = std::min<int64_t>(ncol, v);
  }
  const int64_t nck = P.chunk;
  ga.C = c->get<double>(pre + "col_C", (size_t)nck * k * k);
  ga.c32 = prm->precision != MDA_PREC_FP64;
  ga.crec = c->get<int>(pre + "col_rec", (size_t)ncol * LETKF_CREC_INTS);
  ga.b = c->get<double>(pre + "col_bvec", (size_t)ncol * k);
  ga.nloc = c->get<int>(pre + "col_nloc", ncol);
  if (geom->coord == MDA_COORD_GEOGRAPHIC) {
    ga.shell_cap = std::max<int64_t>(65536, 4 * nck);
    ga.shell = c->get<int4>(pre + "shell_list", (size_t)ga.shell_cap);
    ga.shell_count = c->get<uint32_t>(pre + "shell_count", 1);
  }
  P.queue = queue;
  if (queue) {
    P.ckeys = c->get<uint32_t>(pre + "cq_keys", ncol);
    P.ccols = c->get<uint32_t>(pre + "cq_cols", ncol);
    P.ckeys2 = c->get<uint32_t>(pre + "cq_keys2", ncol);
    P.ccols2 = c->get<uint32_t>(pre + "cq_cols2", ncol);
    P.chist_cap = ((ncol + 255) / 256) * 256;
    P.chist = c->get<uint32_t>(pre + "cq_hist", P.chist_cap);
    P.qcount = c->get<uint32_t>(pre + "cq_count", 1);
  }

  letkf_solve_args& sa = P.sa;
  sa.k = k; sa.ncol = ncol; sa.C = ga.C; sa.c32 = ga.c32; sa.b = ga.b; sa.nloc = ga.nloc;
  sa.rho = prm->inflation; sa.rho_col = rho_col;
  sa.inflation_form = prm->inflation_form; sa.sqrt_form = prm->sqrt_form;
  sa.combine = prm->combine;
  sa.T = c->get<double>(pre + "col_T", (size_t)nck * k * k);
  P.sweeps = sa.sweeps = c->get<int>(pre + "col_sweeps", ncol);
  sa.status = c->get<int>(pre + "col_status", 1);
  sa.de = c->get<double>(pre + "eig_de", (size_t)ncol * 2 * k);
  sa.lam = c->get<double>(pre + "eig_lam", (size_t)ncol * k);
  P.eig_path = prm->sqrt_form == MDA_SQRT_SYMMETRIC && letkf_eig_supported(k) &&
               prm->eig_method == MDA_EIG_AUTO;
  sa.U = P.eig_path ? c->get<double>(pre + "eig_U", (size_t)nck * k * k) : nullptr;
  sa.flag = c->get<int>(pre + "eig_flag", ncol);
  sa.wlist = c->get<int>(pre + "solve_list", ncol);
  sa.wcount = c->get<unsigned>(pre + "solve_count", 1);
  // Jacobi (fallback, or the Jacobi-only method) beyond the LDS capacity runs
  // in a global per-block scratch
  {
    const int64_t kp = (k + 1) & ~1;
    if (sizeof(double) * 2 * kp * (kp + 1) > 96 * 1024) {
      sa.jblocks = 1024;
      sa.jscratch = c->get<double>(pre + "jac_scratch", (size_t)sa.jblocks * 2 * kp * (kp + 1));
    }
  }
  sa.method = prm->eig_method;
  // eigenvalue noise of the Gram relative to ||A||: FP32 products with FP32
  // accumulation, or BF16x3 products (2^-16 relative) with FP32 accumulation
  sa.null_rel = prm->precision == MDA_PREC_FP32 ? 0x1p-22
                : prm->precision == MDA_PREC_BF16X3 ? 0x1p-17 : 0.0;
  // E3 applies the transform itself unless T is to be kept
  sa.fuse = P.eig_path && letkf_eig_fusable(k) && prm->combine == MDA_COMBINE_MATRIX &&
            prm->transform_precision == MDA_PREC_FP64 && !prm->keep_transforms;
  P.stats = c->get<int64_t>(pre + "col_stats", 9);
  return P;
}

// Stages before the per-column work (LETKF.hpp:152-211).  The bucket index
// comes first: the observation statistics then write Y', d (and the FP32 Y')
// straight into the bucket-ordered arrays the Gram reads.  With a stencil H
// they process the observations sorted by the grid position of their first
// tap, so the state gathers of consecutive observations share cache lines.  A
// caller-supplied member-major H(x) is read in input order (coalesced) and
// its rows are moved into bucket order afterwards.  Both idempotent.
static void enqueue_buckets(hipStream_t s, letkf_plan& P) {
  letkf_gram_args& ga = P.ga;
  // the analysis's status word, then the observations' validity
  launch_zero(s, P.sa.status, sizeof(int));
  launch_check_obs(s, P.o, P.sa.status);
  if (P.p == 0) return;  // every column takes the empty-column branch
  launch_bucket_keys(s, P.rcoord, P.ccoord, P.p, P.bg, P.keys, P.vals);
  radix_sort_pairs(s, P.keys, P.vals, P.keys2, P.vals2, P.p, P.key_bits, P.hist, P.hist_cap);
  launch_bucket_ranges(s, P.keys, P.p, P.nbuckets, P.bstart);
  launch_gather_obs(s, P.vals, P.p, P.o.var, P.o.a, P.o.b, P.o.c, P.prm->rweight,
                    const_cast<double*>(ga.rinv), const_cast<double*>(ga.oa),
                    const_cast<double*>(ga.ob), const_cast<double*>(ga.oc),
                    const_cast<double*>(ga.ux), const_cast<double*>(ga.uy),
                    const_cast<double*>(ga.uz));
  // window geometry of every column for the Gram
  launch_column_setup(s, ga);
}

static void enqueue_local(hipStream_t s, letkf_plan& P) {
  letkf_gram_args& ga = P.ga;
  if (P.p == 0) return;
  int* st = P.sa.status;
  float* y32 = P.prm->precision != MDA_PREC_FP64 ? const_cast<float*>(ga.y32) : nullptr;
  if (P.h.hx) {
    launch_obs_stats(s, P.de.X, P.k, P.n, P.p, P.h.row_ptr, P.h.col, P.h.w, P.h.hx, P.o.yo,
                     nullptr, P.yp, P.dd, nullptr, nullptr, nullptr, nullptr, st);
    launch_gather_rows(s, P.vals, P.p, P.k, P.yp, P.dd, const_cast<double*>(ga.yp),
                       const_cast<double*>(ga.d), y32);
  } else if (P.tkeys) {
    // in the order of the first stencil tap (neighbouring state gathers),
    // each observation's rows written at its bucket position
    launch_tap_keys(s, P.h.row_ptr, P.h.col, P.p, P.nx * P.ny, P.tkeys, P.tvals);
    radix_sort_pairs(s, P.tkeys, P.tvals, P.keys2, P.vals2, P.p, P.tap_bits, P.hist,
                     P.hist_cap);
    launch_invert_order(s, P.vals, P.p, P.ipos);
    launch_obs_stats(s, P.de.X, P.k, P.n, P.p, P.h.row_ptr, P.h.col, P.h.w, nullptr, P.o.yo,
                     nullptr, const_cast<double*>(ga.yp), const_cast<double*>(ga.d), nullptr,
                     P.tvals, y32, P.ipos, st);
  } else {
    launch_obs_stats(s, P.de.X, P.k, P.n, P.p, P.h.row_ptr, P.h.col, P.h.w, nullptr, P.o.yo,
                     nullptr, const_cast<double*>(ga.yp), const_cast<double*>(ga.d), nullptr,
                     P.vals, y32, nullptr, st);
  }
}

// The plan's per-column arguments restricted to the pass of columns
// [c0, c0 + nc): pointers into the whole-grid arrays move to column c0, the
// k x k buffers (C, U, T) are the pass's own.
static letkf_gram_args gram_pass(const letkf_plan& P, int64_t c0, int64_t nc) {
  letkf_gram_args ga = P.ga;
  ga.col0 = c0;
  ga.ncol = nc;
  if (ga.col_a) ga.col_a += c0;
  if (ga.col_b) ga.col_b += c0;
  if (ga.col_c) ga.col_c += c0;
  ga.b += c0 * P.k;
  ga.nloc += c0;
  if (ga.crec) ga.crec += c0 * LETKF_CREC_INTS;
  return ga;
}

static letkf_solve_args solve_pass(const letkf_plan& P, int64_t c0, int64_t nc) {
  letkf_solve_args sa = P.sa;
  sa.ncol = nc;
  sa.b += c0 * P.k;
  sa.nloc += c0;
  if (sa.rho_col) sa.rho_col += c0;
  sa.sweeps += c0;
  sa.de += c0 * 2 * P.k;
  sa.lam += c0 * P.k;
  sa.flag += c0;
  sa.X = P.de.X;
  sa.x_ncol = P.ncol;
  sa.x_nlev = P.nlev;
  sa.x_col0 = c0;
  return sa;
}

// Gram of one pass, in work-queue mode behind a descending-cost order of
// its columns.
static void enqueue_gram(hipStream_t s, letkf_plan& P, int64_t c0, int64_t nc, int nblocks,
                         bool queue) {
  letkf_gram_args ga = gram_pass(P, c0, nc);
  ga.nblocks = nblocks;
  if (queue) {
    MDA_REQUIRE(P.queue, "work queue was not planned");
    launch_column_cost(s, ga, P.ckeys, P.ccols);
    radix_sort_pairs(s, P.ckeys, P.ccols, P.ckeys2, P.ccols2, nc, 32, P.chist, P.chist_cap);
    launch_zero(s, P.qcount, sizeof(uint32_t));
    ga.order = P.ccols;  // the sort leaves the ordered pairs in the first arrays
    ga.queue = P.qcount;
  }
  if (ga.shell_count) launch_zero(s, ga.shell_count, sizeof(uint32_t));
  launch_letkf_gram(s, ga);
  launch_letkf_shell_fixup(s, ga, P.sa.status);
}

// Event marks of one pass: gram start, gram end, E1, E2, E3 (tridiagonal
// path), solve end, transform end.
struct pass_marks {
  hipEvent_t e[7];
};

// Gram -> solve -> transform over every pass of columns.  `marks` (optional)
// receives one set of events per pass.
static void enqueue_columns(hipStream_t s, letkf_plan& P, std::pair<int, int> cfg,
                            const std::vector<pass_marks>* marks) {
  launch_zero(s, P.sweeps, sizeof(int) * P.ncol);
  if (P.p == 0) launch_zero(s, P.ga.nloc, sizeof(int) * P.ncol);
  int pass = 0;
  for (int64_t c0 = 0; c0 < P.ncol; c0 += P.chunk, ++pass) {
    const int64_t nc = std::min(P.chunk, P.ncol - c0);
    const pass_marks* m = marks ? &(*marks)[pass] : nullptr;
    if (m) MDA_HIP(hipEventRecord(m->e[0], s));
    if (P.p > 0) enqueue_gram(s, P, c0, nc, cfg.first, cfg.second != 0);
    if (m) MDA_HIP(hipEventRecord(m->e[1], s));
    letkf_solve_args sa = solve_pass(P, c0, nc);
    for (int i = 0; i < 3; ++i) sa.ev_stage[i] = m ? m->e[2 + i] : nullptr;
    launch_letkf_solve(s, sa);
    if (m) MDA_HIP(hipEventRecord(m->e[5], s));
    // fused: only the columns E3 left to the fallback remain
    launch_letkf_transform(s, P.de.X, P.k, P.ncol, P.nlev, P.sa.T, P.sa.status, c0, nc,
                           P.prm->transform_precision, sa.fuse ? sa.flag : nullptr);
    if (m) MDA_HIP(hipEventRecord(m->e[6], s));
  }
}

static int64_t n_passes(const letkf_plan& P) { return (P.ncol + P.chunk - 1) / P.chunk; }

static void enqueue_stats(hipStream_t s, letkf_plan& P) {
  launch_zero(s, P.stats, sizeof(int64_t) * 9);
  launch_reduce_int(s, P.ga.nloc, P.ncol, P.stats, nullptr);
  launch_reduce_int(s, P.sweeps, P.ncol, P.stats + 3, nullptr);
  if (P.eig_path) launch_reduce_int(s, P.sa.flag, P.ncol, P.stats + 6, nullptr);
}

static void finish_letkf(mda_ctx* c, letkf_plan& P, mda_letkf_diag* diag) {
  int64_t hs[9] = {0};
  int hstat = 0;
  d2h(c, hs, P.stats, sizeof(hs));
  d2h(c, &hstat, P.sa.status, sizeof(int));
  // the transforms stay readable only when one pass covered every column
  c->last_T = P.chunk == P.ncol && !P.sa.fuse ? P.sa.T : nullptr;
  c->last_k = P.k; c->last_ncol = P.ncol; c->last_nloc = P.ga.nloc;
  c->last_flag = P.eig_path ? P.sa.flag : nullptr;
  throw_status(hstat & MDA_ST_SKIP_UPDATE, "LETKF");
  unstage_ensemble(c, P.ens, P.de);
  throw_status(hstat, "LETKF");
  if (diag) {
    diag->n_columns = P.ncol;
    diag->total_local_obs = hs[0];
    diag->max_local_obs = hs[1];
    diag->n_empty_columns = hs[2];
    diag->max_jacobi_sweeps = (int32_t)hs[4];
    diag->n_fallback_columns = P.eig_path ? P.ncol - hs[8] : 0;  // nonzero flags
  }
}

int mda_letkf_analyse(mda_ctx* c, const mda_letkf_params* prm, mda_ensemble* ens,
                      const mda_geometry* geom, const mda_obs* obs, const mda_hop* hop,
                      mda_letkf_diag* diag) {
  return guarded(c, [&] {
    hipStream_t s = c->stream;
    MDA_HIP(hipEventRecord(c->ev[0], s));
    letkf_plan P = plan_letkf(c, prm, ens, geom, obs, hop, "", false);
    const int64_t npass = n_passes(P);
    while ((int64_t)c->evpool.size() < 7 * npass) {
      hipEvent_t e = nullptr;
      MDA_HIP(hipEventCreate(&e));
      c->evpool.push_back(e);
    }
    std::vector<pass_marks> marks(npass);
    for (int64_t i = 0; i < npass; ++i)
      for (int j = 0; j < 7; ++j) marks[i].e[j] = c->evpool[7 * i + j];
    MDA_HIP(hipEventRecord(c->ev[1], s));
    enqueue_buckets(s, P);
    MDA_HIP(hipEventRecord(c->ev[2], s));
    enqueue_local(s, P);
    MDA_HIP(hipEventRecord(c->ev[3], s));
    enqueue_columns(s, P, {0, 0}, &marks);
    MDA_HIP(hipEventRecord(c->ev[6], s));
    enqueue_stats(s, P);
    finish_letkf(c, P, diag);
    if (diag) {
      diag->ms_bucket = elapsed(c->ev[1], c->ev[2]);
      diag->ms_obs = elapsed(c->ev[2], c->ev[3]);
      diag->ms_total = elapsed(c->ev[0], c->ev[6]);
      double g = 0, so = 0, tr = 0, e1 = 0, e2 = 0, e3 = 0, jf = 0;
      for (const pass_marks& m : marks) {
        g += elapsed(m.e[0], m.e[1]);
        so += elapsed(m.e[1], m.e[5]);
        tr += elapsed(m.e[5], m.e[6]);
        if (P.eig_path) {
          e1 += elapsed(m.e[1], m.e[2]);
          e2 += elapsed(m.e[2], m.e[3]);
          e3 += elapsed(m.e[3], m.e[4]);
          jf += elapsed(m.e[4], m.e[5]);
        }
      }
      diag->ms_gram = g;
      diag->ms_solve = so;
      diag->ms_transform = tr;
      if (P.eig_path) {
        diag->ms_eig_tridiag = e1;
        diag->ms_eig_values = e2;
        diag->ms_eig_vectors = e3;
        diag->ms_solve_jacobi = jf;
      } else {
        diag->ms_eig_tridiag = diag->ms_eig_values = diag->ms_eig_vectors = 0.0;
        diag->ms_solve_jacobi = diag->ms_solve;
      }
    }
  });
}

// Batched LETKF: independent analyses (any mix of k, precision and
// workload) run as concurrent branches of one hipGraph, the Gram of each in
// work-queue mode behind a descending-cost column order, with the Gram
// launch shape chosen by timing candidates on the first call of a workload.
static std::string tune_key(const letkf_plan& P) {
  return "k" + std::to_string(P.k) + ":prec" + std::to_string(P.prm->precision) + ":ncol" +
         std::to_string(P.ncol) + ":p" + std::to_string(P.p);
}

static std::pair<int, int> autotune_gram(mda_ctx* c, letkf_plan& P, bool allow_queue) {
  hipStream_t s = c->stream;
  enqueue_buckets(s, P);
  enqueue_local(s, P);
  const int cand_blocks[] = {0, 512, 1024, 2048};
  std::pair<int, int> best{0, 0};
  float best_ms = 1e30f;
  for (int q = 0; q <= (allow_queue ? 1 : 0); ++q) {
    for (int nb : cand_blocks) {
      if (q == 0 && nb != 0) continue;  // grid-stride mode: one wave per column
      // timed on the first pass of columns
      const int64_t nc = std::min(P.chunk, P.ncol);
      enqueue_gram(s, P, 0, nc, nb, q != 0);  // warm
      MDA_HIP(hipEventRecord(c->ev[10], s));
      enqueue_gram(s, P, 0, nc, nb, q != 0);
      MDA_HIP(hipEventRecord(c->ev[11], s));
      MDA_HIP(hipEventSynchronize(c->ev[11]));
      const float ms = elapsed(c->ev[10], c->ev[11]);
      if (ms < best_ms) { best_ms = ms; best = {nb, q}; }
    }
  }
  return best;
}

// Bytes of everything an enqueued analysis's kernels read from the plan
// (see mda_letkf_analyse_batch).  The argument blocks are value-initialised
// (padding zero) before their fields are set, so their raw bytes are
// deterministic.
extern "C++" {
template <class T>
static void sig_add(std::string& out, const T& v) {
  out.append(reinterpret_cast<const char*>(&v), sizeof(T));
}
}

static std::string plan_signature(const letkf_plan& P) {
  std::string sg;
  sig_add(sg, P.ga);
  sig_add(sg, P.sa);
  sig_add(sg, *P.prm);
  for (const void* ptr : {(const void*)P.de.X, (const void*)P.o.yo, (const void*)P.o.var,
                          (const void*)P.o.a, (const void*)P.o.b, (const void*)P.o.c,
                          (const void*)P.h.row_ptr, (const void*)P.h.col, (const void*)P.h.w,
                          (const void*)P.h.hx, (const void*)P.rcoord, (const void*)P.ccoord,
                          (const void*)P.yp, (const void*)P.dd, (const void*)P.keys,
                          (const void*)P.vals, (const void*)P.keys2, (const void*)P.vals2,
                          (const void*)P.hist, (const void*)P.tkeys, (const void*)P.tvals,
                          (const void*)P.ipos, (const void*)P.bstart, (const void*)P.sweeps,
                          (const void*)P.stats, (const void*)P.ckeys, (const void*)P.ccols,
                          (const void*)P.ckeys2, (const void*)P.ccols2, (const void*)P.chist,
                          (const void*)P.qcount})
    sig_add(sg, ptr);
  for (int64_t v : {(int64_t)P.k, P.p, P.n, P.nx, P.ny, P.nlev, P.ncol, P.chunk,
                    (int64_t)P.nbuckets, (int64_t)P.key_bits, P.hist_cap, (int64_t)P.tap_bits,
                    (int64_t)P.geo, (int64_t)P.eig_path, (int64_t)P.queue, P.chist_cap})
    sig_add(sg, v);
  return sg;
}

static void enqueue_plan(hipStream_t s, letkf_plan& P, std::pair<int, int> cfg) {
  enqueue_buckets(s, P);
  enqueue_local(s, P);
  enqueue_columns(s, P, cfg, nullptr);
  enqueue_stats(s, P);
}

int mda_letkf_analyse_batch(mda_ctx* c, int32_t nb, const mda_letkf_params* prm,
                            mda_ensemble* ens, const mda_geometry* geom, const mda_obs* obs,
                            const mda_hop* hop, mda_letkf_diag* diag, int32_t flags,
                            mda_batch_info* info) {
  return guarded(c, [&] {
    MDA_REQUIRE(nb >= 1 && nb <= 16, "batch size must be 1..16");
    MDA_REQUIRE(prm && ens && geom && obs && hop, "null batch arrays");
    MDA_REQUIRE((flags & ~(MDA_BATCH_GRAPH | MDA_BATCH_QUEUE | MDA_BATCH_AUTOTUNE)) == 0,
                "unknown batch flags");
    for (int i = 0; i < nb; ++i)
      MDA_REQUIRE(ens[i].on_device && obs[i].on_device && hop[i].on_device,
                  "batched analyses take device-resident ensembles, observations and H");
    const bool use_graph = flags & MDA_BATCH_GRAPH, use_queue = flags & MDA_BATCH_QUEUE;
    hipStream_t s = c->stream;
    std::vector<letkf_plan> plans;
    plans.reserve(nb);
    for (int i = 0; i < nb; ++i)
      plans.push_back(plan_letkf(c, &prm[i], &ens[i], &geom[i], &obs[i], &hop[i],
                                 "b" + std::to_string(i) + ":", use_queue));
    std::vector<std::pair<int, int>> cfg(nb, {0, use_queue ? 1 : 0});
    for (int i = 0; i < nb; ++i) {
      const std::string key = tune_key(plans[i]) + (use_queue ? ":q" : "");
      auto it = c->tune.find(key);
      if (it != c->tune.end()) {
        cfg[i] = it->second;
      } else if (flags & MDA_BATCH_AUTOTUNE) {
        cfg[i] = c->tune[key] = autotune_gram(c, plans[i], use_queue);
      }
    }
    // streams and fork/join events for concurrent branches
    while ((int)c->side.size() < nb - 1) {
      hipStream_t st = nullptr;
      MDA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      c->side.push_back(st);
    }
    while ((int)c->fj.size() < nb + 1) {
      hipEvent_t e = nullptr;
      MDA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      c->fj.push_back(e);
    }
    auto enqueue_all = [&] {
      MDA_HIP(hipEventRecord(c->fj[0], s));
      for (int i = 1; i < nb; ++i) MDA_HIP(hipStreamWaitEvent(c->side[i - 1], c->fj[0], 0));
      for (int i = 0; i < nb; ++i) enqueue_plan(i == 0 ? s : c->side[i - 1], plans[i], cfg[i]);
      for (int i = 1; i < nb; ++i) {
        MDA_HIP(hipEventRecord(c->fj[i], c->side[i - 1]));
        MDA_HIP(hipStreamWaitEvent(s, c->fj[i], 0));
      }
    };
    int replayed = 0;
    MDA_HIP(hipStreamSynchronize(s));
    if (use_graph) {
      // The graph bakes in every kernel argument, so its key holds every
      // device pointer and every value derived from the inputs that the
      // enqueued kernels take: the plan's argument blocks of the Gram and
      // the solve (pointers into the workspace and the caller's arrays,
      // bucket grid, radius, forms), its own buffers and sizes, the caller's
      // parameters and the Gram launch shape.
      std::string sig = std::to_string(nb);
      for (int i = 0; i < nb; ++i)
        sig += "|" + plan_signature(plans[i]) + std::to_string(cfg[i].first) + "," +
               std::to_string(cfg[i].second);
      auto it = c->graphs.find(sig);
      if (it == c->graphs.end()) {
        hipGraph_t graph = nullptr;
        MDA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          enqueue_all();
        } catch (...) {
          hipGraph_t g2 = nullptr;
          (void)hipStreamEndCapture(s, &g2);
          if (g2) (void)hipGraphDestroy(g2);
          throw;
        }
        MDA_HIP(hipStreamEndCapture(s, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        MDA_HIP(e);
        it = c->graphs.emplace(sig, exec).first;
      } else {
        replayed = 1;
      }
      MDA_HIP(hipEventRecord(c->ev[0], s));
      MDA_HIP(hipGraphLaunch(it->second, s));
      MDA_HIP(hipEventRecord(c->ev[6], s));
    } else {
      MDA_HIP(hipEventRecord(c->ev[0], s));
      enqueue_all();
      MDA_HIP(hipEventRecord(c->ev[6], s));
    }
    MDA_HIP(hipStreamSynchronize(s));
    const double ms = elapsed(c->ev[0], c->ev[6]);
    for (int i = 0; i < nb; ++i) {
      mda_letkf_diag* d = diag ? &diag[i] : nullptr;
      finish_letkf(c, plans[i], d);
      if (d) {
        d->ms_total = ms;
        d->ms_obs = d->ms_bucket = d->ms_gram = d->ms_solve = d->ms_transform = 0.0;
        d->ms_eig_tridiag = d->ms_eig_values = d->ms_eig_vectors = d->ms_solve_jacobi = 0.0;
      }
      if (info && i < 16) {
        info->gram_blocks[i] = cfg[i].first;
        info->gram_queue[i] = cfg[i].second;
      }
    }
    if (info) {
      info->ms_total = ms;
      info->graph_replayed = replayed;
      info->n_graphs = (int32_t)c->graphs.size();
    }
  });
}

int mda_hop_idw4_build(mda_ctx* c, const mda_obs* obs, int32_t obs_coord,
                       const mda_geometry* geom, int64_t nx, int64_t ny, int64_t nlev,
                       const double* vcoord, int64_t* row_ptr, int64_t* col, double* weight,
                       int32_t out_on_device) {
  return guarded(c, [&] {
    MDA_REQUIRE(obs && row_ptr && col && weight, "observations and CSR outputs are required");
    MDA_REQUIRE(nx > 0 && ny > 0 && nlev > 0, "grid dimensions must be positive");
    MDA_REQUIRE(obs_coord == MDA_COORD_GRID || obs_coord == MDA_COORD_GEOGRAPHIC,
                "IDW4 observations are GRID or GEOGRAPHIC");
    dev_obs o = stage_obs(c, obs, true);
    const int64_t p = o.p, ncol = nx * ny;
    const double *clat = nullptr, *clon = nullptr, *vc = nullptr;
    if (obs_coord == MDA_COORD_GEOGRAPHIC && geom && geom->coord == MDA_COORD_GEOGRAPHIC) {
      MDA_REQUIRE(geom->col_a && geom->col_b, "GEOGRAPHIC geometry needs column coordinates");
      clat = to_device(c, "hop_clat", geom->col_a, ncol, geom->on_device != 0);
      clon = to_device(c, "hop_clon", geom->col_b, ncol, geom->on_device != 0);
      vc = to_device(c, "hop_vcoord", vcoord, nlev, geom->on_device != 0);
    }
    // rows of 4 taps, none for an invalid observation (variance +inf)
    std::vector<double> hv(p);
    d2h(c, hv.data(), o.var, sizeof(double) * p);
    std::vector<int64_t> rp(p + 1, 0);
    for (int64_t t = 0; t < p; ++t) rp[t + 1] = rp[t] + (std::isinf(hv[t]) ? 0 : 4);
    const int64_t nnz = rp[p];
    int64_t* drp = out_on_device ? row_ptr : c->get<int64_t>("idw_rp", p + 1);
    int64_t* dcol = out_on_device ? col : c->get<int64_t>("idw_col", nnz);
    double* dw = out_on_device ? weight : c->get<double>("idw_w", nnz);
    h2d(c, drp, rp.data(), sizeof(int64_t) * (p + 1));
    launch_idw4_stencil(c->stream, p, obs_coord, o.a, o.b, o.c, o.var, nx, ny, nlev, clat, clon,
                        vc, drp, dcol, dw);
    if (!out_on_device) {
      std::memcpy(row_ptr, rp.data(), sizeof(int64_t) * (p + 1));
      d2h(c, col, dcol, sizeof(int64_t) * nnz);
      d2h(c, weight, dw, sizeof(double) * nnz);
    }
    MDA_HIP(hipStreamSynchronize(c->stream));
  });
}

int mda_ensemble_save(mda_ctx* c, const mda_ensemble* ens, const char* base, const char* format,
                      int32_t sidecar) {
  return guarded(c, [&] {
    MDA_REQUIRE(ens && base && format, "ensemble, base name and format are required");
    dev_ensemble d = stage_ensemble(c, ens);
    // Ensemble::RecomputeMean order (Ensemble.hpp:105-114) on the device
    double* dm = c->get<double>("save_mean", d.n);
    launch_mean_pert(c->stream, d.X, d.k, d.n, 1.0, dm, nullptr);
    std::vector<double> host(d.n);
    const int64_t nrows = ens->ny * ens->nlev;
    auto write = [&](const std::string& stem) {
      const std::string txt = stem + "." + format;
      check_rc(c, mda_state_write_text(c, txt.c_str(), host.data(), ens->nx, nrows));
      if (sidecar) {
        const std::string bin = stem + ".f64";
        check_rc(c, mda_state_write_binary(c, bin.c_str(), host.data(), ens->nx, ens->ny, ens->nlev));
      }
    };
    d2h(c, host.data(), dm, sizeof(double) * d.n);
    write(std::string(base) + "_mean");
    for (int j = 0; j < d.k; ++j) {
      d2h(c, host.data(), d.X + (size_t)j * d.n, sizeof(double) * d.n);
      write(std::string(base) + "_member_" + std::to_string(j));
    }
  });
}

int mda_letkf_last_transforms(mda_ctx* c, const int64_t* cols, int64_t ncols, double* T_out) {
  return guarded(c, [&] {
    MDA_REQUIRE(c->last_nloc != nullptr, "no LETKF analysis on this context yet");
    MDA_REQUIRE(c->last_T != nullptr,
                "the transforms of that analysis were not kept (set keep_transforms = 1; "
                "an analysis run in column passes never keeps them)");
    MDA_REQUIRE(cols && T_out, "cols and T_out are required");
    const size_t kk = (size_t)c->last_k * c->last_k;
    for (int64_t i = 0; i < ncols; ++i) {
      MDA_REQUIRE(cols[i] >= 0 && cols[i] < c->last_ncol, "column index out of range");
      d2h(c, T_out + i * kk, c->last_T + cols[i] * kk, sizeof(double) * kk);
    }
    MDA_HIP(hipStreamSynchronize(c->stream));
  });
}

int mda_letkf_last_local_counts(mda_ctx* c, int32_t* counts) {
  return guarded(c, [&] {
    MDA_REQUIRE(c->last_nloc != nullptr, "no LETKF analysis on this context yet");
    MDA_REQUIRE(counts, "counts is required");
    d2h(c, counts, c->last_nloc, sizeof(int32_t) * c->last_ncol);
  });
}

int mda_letkf_last_column_flags(mda_ctx* c, int32_t* flags) {
  return guarded(c, [&] {
    MDA_REQUIRE(c->last_nloc != nullptr, "no LETKF analysis on this context yet");
    MDA_REQUIRE(flags, "flags is required");
    if (c->last_flag)
      d2h(c, flags, c->last_flag, sizeof(int32_t) * c->last_ncol);
    else
      std::memset(flags, 0, sizeof(int32_t) * c->last_ncol);
  });
}

int mda_ensemble_metrics(mda_ctx* c, const mda_ensemble* ens, const double* truth,
                         int32_t truth_on_device, mda_metrics* out) {
  return guarded(c, [&] {
    MDA_REQUIRE(truth && out, "truth and out are required");
    dev_ensemble de = stage_ensemble(c, ens);
    const double* t = to_device(c, "truth", truth, de.n, truth_on_device != 0);
    const int nb = 512;
    double* part = c->get<double>("metrics_part", (size_t)nb * 9);
    launch_metrics(c->stream, de.X, de.k, de.n, t, part, nb);
    std::vector<double> h((size_t)nb * 9);
    d2h(c, h.data(), part, sizeof(double) * nb * 9);
    MDA_HIP(hipStreamSynchronize(c->stream));
    double a[9] = {0};
    for (int b = 0; b < nb; ++b)
      for (int q = 0; q < 9; ++q) a[q] += h[(size_t)b * 9 + q];
    const double nn = (double)de.n;
    out->rmse = std::sqrt(a[0] / nn);
    out->bias = a[1] / nn;
    out->correlation = (nn * a[4] - a[2] * a[3]) /
                       std::sqrt((nn * a[5] - a[2] * a[2]) * (nn * a[6] - a[3] * a[3]));
    out->avg_spread = a[7] / nn;
    out->crps = a[8] / nn;
  });
}


// ---------------------------------------------------------------------------
// learned inflation
// ---------------------------------------------------------------------------
int mda_l96_forecast(mda_ctx* c, double* X, int64_t nstates, int64_t N, double forcing,
                     double dt, int32_t nsteps, int32_t scheme, int32_t on_device) {
  return guarded(c, [&] {
    MDA_REQUIRE(X != nullptr, "X is null");
    MDA_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    MDA_REQUIRE(scheme == 0 || scheme == 1, "scheme is 0 (RK4) or 1 (Euler)");
    MDA_REQUIRE(std::isfinite(dt) && std::isfinite(forcing), "dt and forcing must be finite");
    double* d = X;
    if (!on_device) {
      d = c->get<double>("l96_states", (size_t)nstates * N);
      h2d(c, d, X, sizeof(double) * nstates * N);
    }
    launch_l96(c->stream, d, nstates, N, forcing, dt, nsteps, scheme);
    if (!on_device) d2h(c, X, d, sizeof(double) * nstates * N);
    MDA_HIP(hipStreamSynchronize(c->stream));
  });
}

int mda_letkf_inflation_features(mda_ctx* c, const mda_letkf_params* prm,
                                 const mda_ensemble* ens, const mda_geometry* geom,
                                 const mda_obs* obs, const mda_hop* hop, const double* rho_prev,
                                 float* features, int32_t features_on_device) {
  return guarded(c, [&] {
    MDA_REQUIRE(prm && geom && features, "params, geometry and features are required");
    MDA_REQUIRE(prm->radius > 0.0, "radius must be > 0");
    MDA_REQUIRE(geom->coord >= 0 && geom->coord <= 2, "unknown coordinate system");
    hipStream_t s = c->stream;
    mda_letkf_params fp = *prm;
    fp.precision = MDA_PREC_FP64;
    fp.rho_column = nullptr;
    letkf_plan P = plan_letkf(c, &fp, const_cast<mda_ensemble*>(ens), geom, obs, hop, "",
                              false);
    MDA_REQUIRE(P.p > 0, "inflation features need observations");
    enqueue_buckets(s, P);
    enqueue_local(s, P);
    const int64_t ncol = P.ncol;
    const dev_obs& o = P.o;
    const dev_ensemble& de = P.de;
    const letkf_gram_args& ga = P.ga;
    double* sp = c->get<double>("feat_sp", o.p);
    launch_obs_spread(s, ga.yp, o.p, de.k, sp);
    const double* rp = to_device(c, "feat_rho_prev", rho_prev, ncol, prm->rho_on_device != 0);
    float* f = features_on_device ? features : c->get<float>("feat_out", (size_t)ncol * 8);
    launch_letkf_features(s, ga, sp, de.X, ens->nlev, rp, f);
    if (!features_on_device) d2h(c, features, f, sizeof(float) * ncol * 8);
    MDA_HIP(hipStreamSynchronize(s));
  });
}

int mda_mlp_forward(mda_ctx* c, const float* params, const float* features, int64_t m,
                    float* rho, int32_t on_device) {
  return guarded(c, [&] {
    MDA_REQUIRE(params && features && rho, "params, features and rho are required");
    MDA_REQUIRE(m >= 0, "m must be >= 0");
    const bool od = on_device != 0;
    const float* P = to_device(c, "mlp_params_in", params, MDA_MLP_NPARAM, od);
    const float* F = to_device(c, "mlp_feat_in", features, (size_t)m * MDA_MLP_NIN, od);
    float* R = od ? rho : c->get<float>("mlp_rho_out", m);
    launch_mlp_forward(c->stream, P, F, m, R);
    if (!od) d2h(c, rho, R, sizeof(float) * m);
    MDA_HIP(hipStreamSynchronize(c->stream));
  });
}

int mda_mlp_loss_grad(mda_ctx* c, const float* params, const float* features,
                      const float* target, int64_t m, double* loss, float* grad,
                      int32_t on_device) {
  return guarded(c, [&] {
    MDA_REQUIRE(params && features && target && loss && grad, "null argument");
    MDA_REQUIRE(m > 0, "m must be > 0");
    const bool od = on_device != 0;
    const float* P = to_device(c, "mlp_params_in", params, MDA_MLP_NPARAM, od);
    const float* F = to_device(c, "mlp_feat_in", features, (size_t)m * MDA_MLP_NIN, od);
    const float* T = to_device(c, "mlp_target_in", target, m, od);
    float* G = od ? grad : c->get<float>("mlp_grad_out", MDA_MLP_NPARAM);
    double* Ld = c->get<double>("mlp_loss", 1);
    launch_mlp_loss_grad(c, c->stream, P, F, T, m, G, Ld);
    d2h(c, loss, Ld, sizeof(double));
    if (!od) d2h(c, grad, G, sizeof(float) * MDA_MLP_NPARAM);
  });
}

int mda_mlp_adam_step(mda_ctx* c, float* params, float* m1, float* m2, const float* grad,
                      int64_t step, float lr, float beta1, float beta2, float eps,
                      int32_t on_device) {
  return guarded(c, [&] {
    MDA_REQUIRE(params && m1 && m2 && grad, "null argument");
    MDA_REQUIRE(on_device != 0, "mda_mlp_adam_step works on device arrays");
    MDA_REQUIRE(lr > 0.f && beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps > 0.f,
                "bad Adam hyper-parameters");
    launch_mlp_adam(c->stream, params, m1, m2, grad, step, lr, beta1, beta2, eps);
  });
}

}  // extern "C"
