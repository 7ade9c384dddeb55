// C-ABI entry points (include/metada_amd.h).  Each one stages host inputs to
// the context's device workspace when needed, enqueues the kernels on the
// context's stream and converts every error into an MDA_E_* code.
#include <algorithm>
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

dev_obs stage_obs(mda_ctx* c, const mda_obs* o, bool need_loc) {
  MDA_REQUIRE(o != nullptr, "observations are null");
  MDA_REQUIRE(o->p > 0, "no observations");
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
  if (c->stream) (void)hipStreamDestroy(c->stream);
  (void)hipSetDevice(prev);
  delete c;
  return MDA_OK;
}

const char* mda_last_error(const mda_ctx* c) { return c ? c->err.c_str() : "null context"; }

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
int mda_etkf_analyse(mda_ctx* c, double inflation, mda_ensemble* ens, const mda_obs* obs,
                     const mda_hop* hop) {
  return guarded(c, [&] {
    MDA_REQUIRE(std::isfinite(inflation), "inflation must be finite");
    dev_ensemble de = stage_ensemble(c, ens);
    dev_obs o = stage_obs(c, obs, false);
    dev_hop h = stage_hop(c, hop, de.k, o.p, de.n);
    const int k = de.k;
    const int64_t p = o.p;
    double* yp = c->get<double>("yp_om", (size_t)p * k);
    double* dd = c->get<double>("d", p);
    launch_obs_stats(c->stream, de.X, k, de.n, p, h.row_ptr, h.col, h.w, h.hx, o.yo, nullptr, yp,
                     dd, nullptr);
    const int nb = 256;
    double* part = c->get<double>("gg_part", (size_t)nb * k * (k + 1));
    double* EB = c->get<double>("gg_EB", (size_t)k * (k + 1));
    launch_global_gram(c->stream, yp, dd, o.var, nullptr, k, p, 0, part, nb, EB);
    double* T = c->get<double>("gT", (size_t)k * k);
    double* scr = c->get<double>("g_scr", (size_t)2 * k * k);
    int* st = c->get<int>("g_status", 1);
    launch_etkf_solve_ws(c->stream, k, EB, T, scr, st);
    launch_transform_global(c->stream, de.X, k, de.n, inflation, T);
    int hs = 0;
    d2h(c, &hs, st, sizeof(int));
    MDA_HIP(hipStreamSynchronize(c->stream));
    if (hs) throw mda_error(MDA_E_NUMERIC, "ETKF: Pa is not positive definite");
    unstage_ensemble(c, ens, de);
  });
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
    MDA_REQUIRE(k <= 64, "EnKF supports k <= 64 in this build");
    const double* dZ = to_device(c, "enkf_Z", Z, (size_t)p * k, z_on_device != 0);
    // all inflation methods are multiplicative sqrt(rho) (EnKF.hpp:316-331)
    const double scale = std::sqrt(inflation);
    double* yp = c->get<double>("yp_om", (size_t)p * k);
    double* dd = c->get<double>("d", p);
    launch_obs_stats(c->stream, de.X, k, n, p, h.row_ptr, h.col, h.w, h.hx, o.yo, nullptr, yp, dd,
                     nullptr);
    const int nb = 256;
    double* part = c->get<double>("gg_part", (size_t)nb * k * (k + 1));
    double* EB = c->get<double>("gg_EB", (size_t)k * (k + 1));
    double* F = c->get<double>("gg_F", (size_t)k * k);
    launch_global_gram(c->stream, yp, dd, o.var, nullptr, k, p, 0, part, nb, EB);
    launch_global_gram(c->stream, yp, dd, o.var, dZ, k, p, 1, part, nb, F);
    double* T = c->get<double>("gT", (size_t)k * k);
    double* Minv = c->get<double>("gMinv", (size_t)k * k);
    double* scr = c->get<double>("g_scr", (size_t)2 * k * k);
    int* st = c->get<int>("g_status", 1);
    launch_enkf_solve_ws(c->stream, k, EB, F, T, Minv, scr, st);
    mda_enkf_diag dg{};
    dg.ensemble_size = k;
    dg.observation_count = p;
    dg.inflation_method = method;
    dg.inflation_factor = inflation;
    dg.analysis_increment_norm = 0.0;
    double* part2 = c->get<double>("sumsq_part", 1024);
    if (diag) {
      launch_dev_sumsq(c->stream, de.X, k, n, scale, part2, 512);
      dg.background_spread = std::sqrt(host_sum(c, part2, 512) / ((double)n * k));
      launch_sumsq(c->stream, dd, p, part2, 512);
      dg.innovation_norm = std::sqrt(host_sum(c, part2, 512));
      // max / min Kalman gain: n*p*k work, skipped (NaN) beyond 1e12
      if ((double)n * p * k <= 1e12) {
        double* G = c->get<double>("enkf_G", (size_t)p * k);
        double* mm = c->get<double>("enkf_mm", 2 * 512);
        launch_enkf_gain_minmax_ws(c->stream, de.X, k, n, scale, Minv, yp, o.var, p, G, mm, 512);
        std::vector<double> hm(2 * 512);
        d2h(c, hm.data(), mm, sizeof(double) * 1024);
        MDA_HIP(hipStreamSynchronize(c->stream));
        double mx = -std::numeric_limits<double>::infinity(), mn = -mx;
        for (int b = 0; b < 512; ++b) { mx = std::max(mx, hm[2 * b]); mn = std::min(mn, hm[2 * b + 1]); }
        dg.max_kalman_gain = mx;
        dg.min_kalman_gain = mn;
      } else {
        dg.max_kalman_gain = dg.min_kalman_gain = std::numeric_limits<double>::quiet_NaN();
      }
      // cond(S) (EnKF.hpp:205-209)
      double* cnd = c->get<double>("enkf_cond", 1);
      if (p <= 128) {
        double* S = c->get<double>("enkf_S", (size_t)130 * 130);
        launch_cond_small(c->stream, yp, o.var, k, (int)p, S, cnd);
        d2h(c, &dg.condition_number, cnd, sizeof(double));
        MDA_HIP(hipStreamSynchronize(c->stream));
      } else {
        std::vector<double> hv(p);
        d2h(c, hv.data(), o.var, sizeof(double) * p);
        MDA_HIP(hipStreamSynchronize(c->stream));
        bool constant = std::all_of(hv.begin(), hv.end(), [&](double v) { return v == hv[0]; });
        if (constant) {
          // R = s^2 I and p > k: eig(S) = s^2 (1 + mu(E)/(k-1)) and s^2 (rank deficiency)
          double* S = c->get<double>("enkf_S", (size_t)130 * 130);
          launch_sym_maxeig(c->stream, EB, k, S, cnd);
          double mu = 0.0;
          d2h(c, &mu, cnd, sizeof(double));
          MDA_HIP(hipStreamSynchronize(c->stream));
          dg.condition_number = 1.0 + mu / (double)(k - 1);
        } else {
          dg.condition_number = std::numeric_limits<double>::quiet_NaN();
        }
      }
    }
    launch_transform_global(c->stream, de.X, k, n, scale, T);
    int hs = 0;
    d2h(c, &hs, st, sizeof(int));
    if (diag) {
      launch_dev_sumsq(c->stream, de.X, k, n, 1.0, part2, 512);
      dg.analysis_spread = std::sqrt(host_sum(c, part2, 512) / ((double)n * k));
    }
    MDA_HIP(hipStreamSynchronize(c->stream));
    if (hs) throw mda_error(MDA_E_NUMERIC, "EnKF: (k-1)I + Y'^T R^-1 Y' is not positive definite");
    unstage_ensemble(c, ens, de);
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

// Geometry staging, observation-space statistics (HX, Y', d), the bucket
// index over the localisation coordinates and the gather of the observation
// data into bucket order: everything before the per-column work of the
// LETKF (LETKF.hpp:152-211).  Records ev[1..3].
struct local_obs {
  const double *ca = nullptr, *cb = nullptr, *cc = nullptr, *rho_col = nullptr;
  bucket_grid bg{};
  const int* bstart = nullptr;
  double *s_yp = nullptr, *s_d = nullptr, *s_r = nullptr, *s_a = nullptr, *s_b = nullptr,
         *s_c = nullptr;
  float* s_y32 = nullptr;
  double *s_ux = nullptr, *s_uy = nullptr, *s_uz = nullptr;
};

static local_obs prepare_local(mda_ctx* c, const mda_letkf_params* prm, const dev_ensemble& de,
                        const mda_ensemble* ens, const mda_geometry* geom, const dev_obs& o,
                        const dev_hop& h) {
  hipStream_t s = c->stream;
  local_obs L;
  const int k = de.k;
  const int64_t p = o.p, n = de.n, nx = ens->nx, ny = ens->ny;
  const int64_t ncol = nx * ny;
  const bool gd = geom->on_device != 0;
  const double* ca = L.ca = to_device(c, "geom_a", geom->col_a, ncol, gd);
  const double* cb = L.cb = to_device(c, "geom_b", geom->col_b, ncol, gd);
  L.cc = to_device(c, "geom_c", geom->col_c, ncol, gd);
  MDA_REQUIRE(geom->coord == MDA_COORD_GRID || (ca && cb),
              "GEOGRAPHIC / CARTESIAN geometry needs column coordinates");
  L.rho_col = to_device(c, "rho_col", prm->rho_column, ncol, prm->rho_on_device != 0);
  MDA_HIP(hipEventRecord(c->ev[1], s));

  // 1. observation space
  double* yp = c->get<double>("yp_om", (size_t)p * k);
  double* dd = c->get<double>("d", p);
  launch_obs_stats(s, de.X, k, n, p, h.row_ptr, h.col, h.w, h.hx, o.yo, nullptr, yp, dd,
                   nullptr);
  MDA_HIP(hipEventRecord(c->ev[2], s));

  // 2. bucket grid over the localisation coordinates
  bucket_grid bg{};
  bg.coord = geom->coord;
  const double* rcoord;
  const double* ccoord;
  if (geom->coord == MDA_COORD_GEOGRAPHIC) {
    double ddeg = prm->radius / 6371.0 * 180.0 / 3.14159265358979323846;
    double step = std::min(10.0, std::max(0.05, ddeg / 2.0));
    // bucket sizes divide the latitude and longitude ranges exactly
    bg.nrow = (int)std::ceil(180.0 / step);
    bg.ncolb = (int)std::ceil(360.0 / step);
    bg.a0 = -90.0; bg.b0 = 0.0; bg.da = 180.0 / bg.nrow; bg.db = 360.0 / bg.ncolb;
    bg.wrap = 1;
    rcoord = o.a; ccoord = o.b;
  } else {
    double xmin = 0, xmax = (double)(nx - 1), ymin = 0, ymax = (double)(ny - 1);
    if (geom->col_a) {
      std::vector<double> hx(ncol), hy(ncol);
      d2h(c, hx.data(), ca, sizeof(double) * ncol);
      d2h(c, hy.data(), cb, sizeof(double) * ncol);
      MDA_HIP(hipStreamSynchronize(s));
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
    rcoord = o.b; ccoord = o.a;
  }
  const int nbuckets = bg.nrow * bg.ncolb;
  MDA_REQUIRE(nbuckets > 0 && nbuckets < (1 << 24), "bucket grid too large");
  int key_bits = 1;
  while ((1 << key_bits) < nbuckets) ++key_bits;
  uint32_t* keys = c->get<uint32_t>("bk_keys", p);
  uint32_t* vals = c->get<uint32_t>("bk_vals", p);
  uint32_t* keys2 = c->get<uint32_t>("bk_keys2", p);
  uint32_t* vals2 = c->get<uint32_t>("bk_vals2", p);
  int64_t hist_cap = ((p + 255) / 256) * 256;
  uint32_t* hist = c->get<uint32_t>("bk_hist", hist_cap);
  int* bstart = c->get<int>("bk_start", nbuckets + 1);
  launch_bucket_keys(s, rcoord, ccoord, p, bg, keys, vals);
  radix_sort_pairs(s, keys, vals, keys2, vals2, p, key_bits, hist, hist_cap);
  launch_bucket_ranges(s, keys, p, nbuckets, bstart);
  double* s_yp = c->get<double>("s_yp", (size_t)p * k);
  double* s_d = c->get<double>("s_d", p);
  double* s_r = c->get<double>("s_rinv", p);
  double* s_a = c->get<double>("s_a", p);
  double* s_b = c->get<double>("s_b", p);
  double* s_c = c->get<double>("s_c", p);
  launch_gather_obs(s, vals, p, k, yp, dd, o.var, o.a, o.b, o.c, prm->rweight, s_yp, s_d, s_r,
                    s_a, s_b, s_c);
  float* s_y32 = nullptr;
  double *s_ux = nullptr, *s_uy = nullptr, *s_uz = nullptr;
  if (prm->precision != MDA_PREC_FP64) {
    const bool geo = geom->coord == MDA_COORD_GEOGRAPHIC;
    s_y32 = c->get<float>("s_y32", (size_t)p * k);
    if (geo) {
      s_ux = c->get<double>("s_ux", p);
      s_uy = c->get<double>("s_uy", p);
      s_uz = c->get<double>("s_uz", p);
    }
    launch_mixed_obs(s, p, k, s_yp, s_a, s_b, geo ? 1 : 0, s_y32, s_ux, s_uy, s_uz);
  }
  MDA_HIP(hipEventRecord(c->ev[3], s));
  L.bg = bg; L.bstart = bstart;
  L.s_yp = s_yp; L.s_d = s_d; L.s_r = s_r; L.s_a = s_a; L.s_b = s_b; L.s_c = s_c;
  L.s_y32 = s_y32; L.s_ux = s_ux; L.s_uy = s_uy; L.s_uz = s_uz;
  return L;
}

int mda_letkf_analyse(mda_ctx* c, const mda_letkf_params* prm, mda_ensemble* ens,
                      const mda_geometry* geom, const mda_obs* obs, const mda_hop* hop,
                      mda_letkf_diag* diag) {
  return guarded(c, [&] {
    MDA_REQUIRE(prm && geom, "params and geometry are required");
    MDA_REQUIRE(std::isfinite(prm->inflation) && prm->inflation > 0.0, "inflation must be > 0");
    MDA_REQUIRE(std::isfinite(prm->radius) || prm->radius > 0.0, "radius must be > 0");
    MDA_REQUIRE(prm->radius > 0.0, "radius must be > 0");
    MDA_REQUIRE(prm->precision == MDA_PREC_FP64 || prm->precision == MDA_PREC_FP32 ||
                    prm->precision == MDA_PREC_BF16X3,
                "precision must be FP64, FP32 or BF16X3");
    MDA_REQUIRE(geom->coord >= 0 && geom->coord <= 2, "unknown coordinate system");
    hipStream_t s = c->stream;
    MDA_HIP(hipEventRecord(c->ev[0], s));
    dev_ensemble de = stage_ensemble(c, ens);
    dev_obs o = stage_obs(c, obs, true);
    dev_hop h = stage_hop(c, hop, de.k, o.p, de.n);
    const int k = de.k;
    const int64_t p = o.p, n = de.n, nx = ens->nx, ny = ens->ny, nlev = ens->nlev;
    const int64_t ncol = nx * ny;
    MDA_REQUIRE(k <= 64, "LETKF supports k <= 64 in this build");
    const local_obs L = prepare_local(c, prm, de, ens, geom, o, h);
    const double* ca = L.ca;
    const double* cb = L.cb;
    const double* cc = L.cc;
    const double* rho_col = L.rho_col;
    const bucket_grid bg = L.bg;
    const int* bstart = L.bstart;
    double *s_yp = L.s_yp, *s_d = L.s_d, *s_r = L.s_r, *s_a = L.s_a, *s_b = L.s_b, *s_c = L.s_c;
    float* s_y32 = L.s_y32;
    double *s_ux = L.s_ux, *s_uy = L.s_uy, *s_uz = L.s_uz;

    // 3. per-column Gram
    double* C = c->get<double>("col_C", (size_t)ncol * k * k);
    double* bvec = c->get<double>("col_bvec", (size_t)ncol * k);
    int* nloc = c->get<int>("col_nloc", ncol);
    letkf_gram_args ga{};
    ga.k = k; ga.ncol = ncol; ga.nx = nx; ga.coord = geom->coord;
    ga.col_a = ca; ga.col_b = cb; ga.col_c = cc;
    ga.grid = bg; ga.bstart = bstart;
    ga.yp = s_yp; ga.d = s_d; ga.rinv = s_r; ga.oa = s_a; ga.ob = s_b; ga.oc = s_c;
    ga.radius = prm->radius; ga.localization = prm->localization; ga.precision = prm->precision;
    ga.y32 = s_y32; ga.ux = s_ux; ga.uy = s_uy; ga.uz = s_uz;
    ga.C = C; ga.b = bvec; ga.nloc = nloc;
    launch_letkf_gram(s, ga);
    MDA_HIP(hipEventRecord(c->ev[4], s));

    // 4. per-column solve
    double* T = c->get<double>("col_T", (size_t)ncol * k * k);
    int* sweeps = c->get<int>("col_sweeps", ncol);
    int* status = c->get<int>("col_status", 1);
    launch_zero(s, status, sizeof(int));
    letkf_solve_args sa{};
    sa.k = k; sa.ncol = ncol; sa.C = C; sa.b = bvec; sa.nloc = nloc;
    sa.rho = prm->inflation; sa.rho_col = rho_col;
    sa.inflation_form = prm->inflation_form; sa.sqrt_form = prm->sqrt_form;
    sa.combine = prm->combine; sa.T = T; sa.sweeps = sweeps; sa.status = status;
    sa.de = c->get<double>("eig_de", (size_t)ncol * 2 * k);
    sa.lam = c->get<double>("eig_lam", (size_t)ncol * k);
    sa.U = letkf_eig_supported(k) && prm->eig_method == MDA_EIG_AUTO &&
                   prm->sqrt_form == MDA_SQRT_SYMMETRIC
               ? c->get<double>("eig_U", (size_t)ncol * k * k)
               : nullptr;
    sa.flag = c->get<int>("eig_flag", ncol);
    sa.method = prm->eig_method;
    sa.ev_stage[0] = c->ev[7]; sa.ev_stage[1] = c->ev[8]; sa.ev_stage[2] = c->ev[9];
    // eigenvalue noise of the Gram relative to ||A||: FP32 products with FP32
    // accumulation, or BF16x3 products (2^-16 relative) with FP32 accumulation
    sa.null_rel = prm->precision == MDA_PREC_FP32 ? 0x1p-22
                  : prm->precision == MDA_PREC_BF16X3 ? 0x1p-17 : 0.0;
    launch_zero(s, sweeps, sizeof(int) * ncol);
    launch_letkf_solve(s, sa);
    MDA_HIP(hipEventRecord(c->ev[5], s));

    // 5. transform, in place
    launch_letkf_transform(s, de.X, k, ncol, nlev, T);
    MDA_HIP(hipEventRecord(c->ev[6], s));

    int64_t* stats = c->get<int64_t>("col_stats", 9);
    launch_zero(s, stats, sizeof(int64_t) * 9);
    launch_reduce_int(s, nloc, ncol, stats, nullptr);
    launch_reduce_int(s, sweeps, ncol, stats + 3, nullptr);
    const bool eig_path = prm->sqrt_form == MDA_SQRT_SYMMETRIC && letkf_eig_supported(k) &&
                          prm->eig_method == MDA_EIG_AUTO;
    if (eig_path) launch_reduce_int(s, sa.flag, ncol, stats + 6, nullptr);
    int64_t hs[9] = {0};
    int hstat = 0;
    d2h(c, hs, stats, sizeof(hs));
    d2h(c, &hstat, status, sizeof(int));
    MDA_HIP(hipStreamSynchronize(s));
    c->last_T = T; c->last_k = k; c->last_ncol = ncol; c->last_nloc = nloc;
    c->last_flag = eig_path ? sa.flag : nullptr;
    if (hstat) throw mda_error(MDA_E_NUMERIC, "LETKF: a local Pa is not positive definite");
    unstage_ensemble(c, ens, de);
    if (diag) {
      diag->n_columns = ncol;
      diag->total_local_obs = hs[0];
      diag->max_local_obs = hs[1];
      diag->n_empty_columns = hs[2];
      diag->max_jacobi_sweeps = (int32_t)hs[4];
      diag->n_fallback_columns = eig_path ? ncol - hs[8] : 0;  // nonzero flags
      diag->ms_obs = elapsed(c->ev[1], c->ev[2]);
      diag->ms_bucket = elapsed(c->ev[2], c->ev[3]);
      diag->ms_gram = elapsed(c->ev[3], c->ev[4]);
      diag->ms_solve = elapsed(c->ev[4], c->ev[5]);
      diag->ms_transform = elapsed(c->ev[5], c->ev[6]);
      diag->ms_total = elapsed(c->ev[0], c->ev[6]);
      if (eig_path) {
        diag->ms_eig_tridiag = elapsed(c->ev[4], c->ev[7]);
        diag->ms_eig_values = elapsed(c->ev[7], c->ev[8]);
        diag->ms_eig_vectors = elapsed(c->ev[8], c->ev[9]);
        diag->ms_solve_jacobi = elapsed(c->ev[9], c->ev[5]);
      } else {
        diag->ms_eig_tridiag = diag->ms_eig_values = diag->ms_eig_vectors = 0.0;
        diag->ms_solve_jacobi = diag->ms_solve;
      }
    }
  });
}

int mda_letkf_last_transforms(mda_ctx* c, const int64_t* cols, int64_t ncols, double* T_out) {
  return guarded(c, [&] {
    MDA_REQUIRE(c->last_T != nullptr, "no LETKF analysis on this context yet");
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
    MDA_REQUIRE(c->last_T != nullptr, "no LETKF analysis on this context yet");
    MDA_REQUIRE(counts, "counts is required");
    d2h(c, counts, c->last_nloc, sizeof(int32_t) * c->last_ncol);
  });
}

int mda_letkf_last_column_flags(mda_ctx* c, int32_t* flags) {
  return guarded(c, [&] {
    MDA_REQUIRE(c->last_T != nullptr, "no LETKF analysis on this context yet");
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
    dev_ensemble de = stage_ensemble(c, ens);
    dev_obs o = stage_obs(c, obs, true);
    dev_hop h = stage_hop(c, hop, de.k, o.p, de.n);
    const int64_t ncol = ens->nx * ens->ny;
    const local_obs L = prepare_local(c, &fp, de, ens, geom, o, h);
    letkf_gram_args ga{};
    ga.k = de.k; ga.ncol = ncol; ga.nx = ens->nx; ga.coord = geom->coord;
    ga.col_a = L.ca; ga.col_b = L.cb; ga.col_c = L.cc;
    ga.grid = L.bg; ga.bstart = L.bstart;
    ga.yp = L.s_yp; ga.d = L.s_d; ga.rinv = L.s_r; ga.oa = L.s_a; ga.ob = L.s_b; ga.oc = L.s_c;
    ga.radius = prm->radius; ga.localization = prm->localization;
    double* sp = c->get<double>("feat_sp", o.p);
    launch_obs_spread(s, L.s_yp, o.p, de.k, sp);
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
