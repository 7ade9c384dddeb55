// metada_amd.hpp -- header-only C++ adapter that gives the C-ABI of
// metada_amd.h the shape of metada's analysis algorithms, so a backend can
// switch its analysis step to the MI355X library without touching callers.
//
//   metada_amd::LETKF<Tag>   mirrors  LETKF<BackendTag>  (LETKF.hpp:47-142)
//   metada_amd::ETKF<Tag>    mirrors  ETKF<BackendTag>   (ETKF.hpp:84-200)
//   metada_amd::EnKF<Tag>    mirrors  EnKF<BackendTag>   (EnKF.hpp:84-301)
//   metada_amd::LWEnKF<Tag>  mirrors  LWEnKF<BackendTag> (LWEnKF.hpp:56-334)
//   metada_amd::Lorenz96Model is a ModelBackend for Model<BackendTag>
//                            (ModelConcepts.hpp:51-71, Model.hpp:57-177)
//
// Same constructor arguments (non-owning references; the analysis config,
// or for EnKF the root config whose "analysis" subsection it reads), same
// Analyse() that updates the ensemble members in place and refreshes the
// ensemble statistics as the reference does (LETKF: RecomputeMean,
// LETKF.hpp:116; EnKF: RecomputeMean and RecomputePerturbations,
// EnKF.hpp:237-238), same saveEnsemble() (EnKF.hpp:261-282, LETKF.hpp:125-142)
// and, for EnKF, getAnalysisResults() with the reference's AnalysisResults
// fields (EnKF.hpp:84-96).  Errors surface as std::runtime_error, the
// reference's error behaviour.
//
// The adapter is written against the reference's adapter API only:
//   Ensemble:    Size(), GetMember(i), RecomputeMean(), RecomputePerturbations(),
//                Mean()                                       (Ensemble.hpp)
//   State:       getDataPtr<double>(), size(), geometry(), saveToFile(name)
//                                                             (State.hpp:229-275,383)
//   Geometry:    range-for over Location                      (Geometry.hpp)
//   Location:    getCoordinateSystem(), getGridCoords(), getGeographicCoords(),
//                getCartesianCoords()                         (Location.hpp:100-180)
//   Observation: getObservationValues(), getCovariance(), size(), range-for
//                over points with .location                   (Observation.hpp:261-482)
//   ObsOperator: apply(member, obs) -> std::vector<double>    (ObsOperator.hpp:255-259)
//   Config:      Get(key) -> ConfigValue with asFloat() / asString(),
//                GetSubsection(key) -> Config                 (Config.hpp:203-313)
// H(x_j) comes from the backend's own ObsOperator (MDA_HOP_HX), so every
// observation operator the reference supports works unchanged.
#ifndef METADA_AMD_HPP
#define METADA_AMD_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "metada_amd.h"

namespace metada_amd {

inline void check(mda_ctx* ctx, int rc, const char* what) {
  if (rc != MDA_OK) {
    const char* msg = ctx ? mda_last_error(ctx) : "no context";
    throw std::runtime_error(std::string(what) + ": " + (msg ? msg : "") + " (code " +
                             std::to_string(rc) + ")");
  }
}

// One library context (device, stream, workspace) per host thread.
class Context {
 public:
  explicit Context(int device = 0) { check(nullptr, mda_ctx_create(device, &ctx_), "mda_ctx_create"); }
  ~Context() { mda_ctx_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  mda_ctx* get() const { return ctx_; }
  static Context& thread_default() {
    thread_local Context c(0);
    return c;
  }

 private:
  mda_ctx* ctx_ = nullptr;
};

// EnKF<BackendTag>::AnalysisResults (EnKF.hpp:84-96)
struct AnalysisResults {
  double innovation_norm;
  double analysis_increment_norm;
  double background_spread;
  double analysis_spread;
  double max_kalman_gain;
  double min_kalman_gain;
  double condition_number;
  int ensemble_size;
  int observation_count;
  std::string inflation_method;
  double inflation_factor;
};

namespace detail {

// The coordinate system of a Location as the C-ABI's MDA_COORD_*.  Compared
// by enumerator name, so the reference's enum class needs no include here.
template <class Loc>
inline int32_t coord_of(const Loc& l) {
  using CS = decltype(l.getCoordinateSystem());
  const CS cs = l.getCoordinateSystem();
  if (cs == CS::GEOGRAPHIC) return MDA_COORD_GEOGRAPHIC;
  if (cs == CS::CARTESIAN) return MDA_COORD_CARTESIAN;
  return MDA_COORD_GRID;
}

template <class Loc>
inline void coords_of(const Loc& l, double& a, double& b, double& c) {
  using CS = decltype(l.getCoordinateSystem());
  const CS cs = l.getCoordinateSystem();
  if (cs == CS::GEOGRAPHIC) {
    const auto t = l.getGeographicCoords();  // (lat, lon, level)
    a = std::get<0>(t); b = std::get<1>(t); c = std::get<2>(t);
  } else if (cs == CS::CARTESIAN) {
    const auto t = l.getCartesianCoords();
    a = std::get<0>(t); b = std::get<1>(t); c = std::get<2>(t);
  } else {
    const auto t = l.getGridCoords();  // (i, j, k)
    a = std::get<0>(t); b = std::get<1>(t); c = std::get<2>(t);
  }
}

// Host views of everything the C-ABI needs for one analysis.  The members
// stay where the State objects keep them (SimpleState.hpp:265): the library
// copies them to the device and writes the analysis back in place.
struct Staged {
  int k = 0;
  int64_t n = 0, p = 0;
  std::vector<double*> members;
  std::vector<double> yo, var, oa, ob, oc, hx;
  int32_t coord = MDA_COORD_GRID;
};

template <class Ens, class Obs, class ObsOp>
Staged stage(Ens& ens, Obs& obs, const ObsOp& op) {
  Staged s;
  s.k = static_cast<int>(ens.Size());
  for (int j = 0; j < s.k; ++j) {
    auto& member = ens.GetMember(j);
    double* data = member.template getDataPtr<double>();  // State.hpp:229-232
    const int64_t size = static_cast<int64_t>(member.size());
    if (!data) throw std::runtime_error("ensemble member has no data");
    if (j > 0 && size != s.n) throw std::runtime_error("ensemble members differ in size");
    s.members.push_back(data);
    s.n = size;
  }
  s.yo = obs.getObservationValues();
  s.var = obs.getCovariance();  // diagonal R: error^2 (+inf marks invalid)
  s.p = static_cast<int64_t>(s.yo.size());
  if (static_cast<int64_t>(s.var.size()) != s.p)
    throw std::runtime_error("observation values and covariance differ in size");
  bool first = true;
  for (const auto& pt : obs) {
    double a, b, c;
    coords_of(pt.location, a, b, c);
    if (first) { s.coord = coord_of(pt.location); first = false; }
    s.oa.push_back(a); s.ob.push_back(b); s.oc.push_back(c);
  }
  // H x_j through the backend's own operator, [k][p]
  s.hx.resize(static_cast<size_t>(s.k) * s.p);
  for (int j = 0; j < s.k; ++j) {
    const std::vector<double> y = op.apply(ens.GetMember(j), obs);
    if (static_cast<int64_t>(y.size()) != s.p)
      throw std::runtime_error("ObsOperator::apply returned the wrong size");
    std::copy(y.begin(), y.end(), s.hx.begin() + static_cast<size_t>(j) * s.p);
  }
  return s;
}

inline mda_ensemble ensemble_view(Staged& s, int64_t nx, int64_t ny, int64_t nlev) {
  mda_ensemble e{};
  e.k = s.k; e.nx = nx; e.ny = ny; e.nlev = nlev;
  e.data = nullptr;
  e.members = s.members.data();
  e.on_device = 0;
  return e;
}

inline mda_obs obs_view(Staged& s) {
  mda_obs o{};
  o.p = s.p; o.value = s.yo.data(); o.variance = s.var.data();
  o.loc_a = s.oa.data(); o.loc_b = s.ob.data(); o.loc_c = s.oc.data();
  o.on_device = 0;
  return o;
}

inline mda_hop hx_view(Staged& s) {
  mda_hop h{};
  h.kind = MDA_HOP_HX;
  h.hx = s.hx.data();
  h.on_device = 0;
  return h;
}

// saveEnsemble of the reference (EnKF.hpp:261-282, LETKF.hpp:125-142):
// <base>_mean.<format> and <base>_member_<i>.<format> through State::saveToFile
template <class Ens>
void save_ensemble(Ens& ens, const std::string& base, const std::string& format) {
  ens.Mean().saveToFile(base + "_mean." + format);
  for (size_t i = 0; i < ens.Size(); ++i)
    ens.GetMember(i).saveToFile(base + "_member_" + std::to_string(i) + "." + format);
}

}  // namespace detail

// LETKF<BackendTag> (LETKF.hpp:47-142).  `form` selects Hunt et al. (2007)
// with Gaspari-Cohn R-localisation (Standard) or the reference's own
// semantics (RefCompat: R = I, hard cutoff, posterior inflation, Cholesky
// square root, element-wise combination, snapshot H).  Grid columns come
// from the first member's geometry (LETKF.hpp:80-102) as nx * ny columns of
// nlev levels in the reference's linear order.
template <class BackendTag>
class LETKF {
 public:
  enum class Form { Standard, RefCompat };

  template <class Ens, class Obs, class ObsOp, class Cfg>
  LETKF(Ens& ensemble, Obs& obs, const ObsOp& obs_op, const Cfg& config,
        Form form = Form::Standard, Context& ctx = Context::thread_default())
      : ctx_(ctx),
        inflation_(config.Get("inflation").asFloat()),
        radius_(config.Get("localization_radius").asFloat()),
        output_base_file_(config.Get("output_base_file").asString()),
        format_(config.Get("format").asString()),
        form_(form) {
    run_ = [&ensemble, &obs, &obs_op, this] {
      this->analyse(ensemble, obs, obs_op);
      ensemble.RecomputeMean();  // LETKF.hpp:116
    };
    save_ = [&ensemble, this] { detail::save_ensemble(ensemble, output_base_file_, format_); };
  }

  void Analyse() { run_(); }
  void saveEnsemble() const { save_(); }
  const mda_letkf_diag& diagnostics() const { return diag_; }
  // per-column inflation (learned inflation), nx*ny values; empty = fixed
  void setColumnInflation(std::vector<double> rho) { rho_col_ = std::move(rho); }

 private:
  template <class Ens, class Obs, class ObsOp>
  void analyse(Ens& ens, Obs& obs, const ObsOp& op) {
    detail::Staged s = detail::stage(ens, obs, op);
    const auto* geom = ens.GetMember(0).geometry();
    if (!geom) throw std::runtime_error("Geometry pointer is null in LETKF");
    std::vector<double> ca, cb, cc;
    int32_t gcoord = MDA_COORD_GRID;
    bool first = true;
    for (const auto& loc : *geom) {
      double a, b, c;
      detail::coords_of(loc, a, b, c);
      if (first) { gcoord = detail::coord_of(loc); first = false; }
      ca.push_back(a); cb.push_back(b); cc.push_back(c);
    }
    const int64_t ncol = static_cast<int64_t>(ca.size());
    if (ncol == 0 || s.n % ncol != 0) throw std::runtime_error("geometry does not divide the state");
    mda_letkf_params prm;
    if (form_ == Form::Standard)
      mda_letkf_params_standard(&prm, inflation_, radius_);
    else
      mda_letkf_params_refcompat(&prm, inflation_, radius_);
    if (!rho_col_.empty()) {
      if (static_cast<int64_t>(rho_col_.size()) != ncol)
        throw std::runtime_error("column inflation has the wrong size");
      prm.rho_column = rho_col_.data();
      prm.rho_on_device = 0;
    }
    mda_ensemble e = detail::ensemble_view(s, ncol, 1, s.n / ncol);
    mda_geometry g{};
    g.coord = gcoord;
    g.col_a = ca.data(); g.col_b = cb.data();
    g.col_c = gcoord == MDA_COORD_CARTESIAN ? cc.data() : nullptr;
    g.on_device = 0;
    mda_obs o = detail::obs_view(s);
    mda_hop h = detail::hx_view(s);
    check(ctx_.get(), mda_letkf_analyse(ctx_.get(), &prm, &e, &g, &o, &h, &diag_),
          "LETKF::Analyse");
  }

  std::function<void()> run_, save_;
  Context& ctx_;
  double inflation_, radius_;
  std::string output_base_file_, format_;
  Form form_;
  std::vector<double> rho_col_;
  mda_letkf_diag diag_{};
};

// ETKF<BackendTag> (ETKF.hpp:84-200): reference semantics (X' *= rho,
// Cholesky square root, dense real R); the analysis config as LETKF's.
template <class BackendTag>
class ETKF {
 public:
  template <class Ens, class Obs, class ObsOp, class Cfg>
  ETKF(Ens& ensemble, Obs& obs, const ObsOp& obs_op, const Cfg& config,
       Context& ctx = Context::thread_default())
      : ctx_(ctx),
        inflation_(config.Get("inflation").asFloat()),
        output_base_file_(config.Get("output_base_file").asString()),
        format_(config.Get("format").asString()) {
    run_ = [&ensemble, &obs, &obs_op, this] {
      this->analyse(ensemble, obs, obs_op);
      ensemble.RecomputeMean();
    };
    save_ = [&ensemble, this] { detail::save_ensemble(ensemble, output_base_file_, format_); };
  }
  void Analyse() { run_(); }
  void saveEnsemble() const { save_(); }

 private:
  template <class Ens, class Obs, class ObsOp>
  void analyse(Ens& ens, Obs& obs, const ObsOp& op) {
    detail::Staged s = detail::stage(ens, obs, op);
    mda_ensemble e = detail::ensemble_view(s, s.n, 1, 1);
    mda_obs o = detail::obs_view(s);
    mda_hop h = detail::hx_view(s);
    check(ctx_.get(), mda_etkf_analyse(ctx_.get(), inflation_, &e, &o, &h), "ETKF::Analyse");
  }
  std::function<void()> run_, save_;
  Context& ctx_;
  double inflation_;
  std::string output_base_file_, format_;
};

// EnKF<BackendTag> (EnKF.hpp:84-301): takes the root config and reads its
// "analysis" subsection (EnKF.hpp:105-128), and the N(0,1) draws Z [p][k]
// that the reference takes from std::random_device (EnKF.hpp:346-360),
// injected here for reproducibility.
template <class BackendTag>
class EnKF {
 public:
  template <class Ens, class Obs, class ObsOp, class Cfg>
  EnKF(Ens& ensemble, Obs& obs, const ObsOp& obs_op, const Cfg& root_config,
       std::vector<double> Z, Context& ctx = Context::thread_default())
      : ctx_(ctx), Z_(std::move(Z)) {
    const auto analysis_config = root_config.GetSubsection("analysis");
    inflation_factor_ = analysis_config.Get("inflation").asFloat();
    output_base_file_ = analysis_config.Get("output_base_file").asString();
    format_ = analysis_config.Get("format").asString();
    // additive and relaxation fall back to multiplicative, as do unknown
    // names (EnKF.hpp:116-128, 316-331)
    const std::string m = analysis_config.Get("inflation_method").asString();
    if (m == "additive") {
      method_ = MDA_INFL_ADDITIVE; method_name_ = m;
    } else if (m == "relaxation") {
      method_ = MDA_INFL_RELAXATION; method_name_ = m;
    } else {
      method_ = MDA_INFL_MULTIPLICATIVE; method_name_ = "multiplicative";
    }
    run_ = [&ensemble, &obs, &obs_op, this] {
      this->analyse(ensemble, obs, obs_op);
      ensemble.RecomputeMean();  // EnKF.hpp:237-238
      ensemble.RecomputePerturbations();
    };
    save_ = [&ensemble, this] { detail::save_ensemble(ensemble, output_base_file_, format_); };
  }
  void Analyse() { run_(); }
  void saveEnsemble() const { save_(); }
  AnalysisResults getAnalysisResults() const {
    AnalysisResults r;
    r.innovation_norm = diag_.innovation_norm;
    r.analysis_increment_norm = diag_.analysis_increment_norm;
    r.background_spread = diag_.background_spread;
    r.analysis_spread = diag_.analysis_spread;
    r.max_kalman_gain = diag_.max_kalman_gain;
    r.min_kalman_gain = diag_.min_kalman_gain;
    r.condition_number = diag_.condition_number;
    r.ensemble_size = diag_.ensemble_size;
    r.observation_count = static_cast<int>(diag_.observation_count);
    r.inflation_method = method_name_;
    r.inflation_factor = inflation_factor_;
    return r;
  }

 private:
  template <class Ens, class Obs, class ObsOp>
  void analyse(Ens& ens, Obs& obs, const ObsOp& op) {
    detail::Staged s = detail::stage(ens, obs, op);
    if (static_cast<int64_t>(Z_.size()) != s.p * s.k)
      throw std::runtime_error("Z must hold p*k standard normal draws");
    mda_ensemble e = detail::ensemble_view(s, s.n, 1, 1);
    mda_obs o = detail::obs_view(s);
    mda_hop h = detail::hx_view(s);
    check(ctx_.get(),
          mda_enkf_analyse(ctx_.get(), inflation_factor_, method_, &e, &o, &h, Z_.data(), 0,
                           &diag_),
          "EnKF::Analyse");
  }
  std::function<void()> run_, save_;
  Context& ctx_;
  double inflation_factor_ = 1.0;
  std::string output_base_file_, format_, method_name_;
  int32_t method_ = MDA_INFL_MULTIPLICATIVE;
  std::vector<double> Z_;
  mda_enkf_diag diag_{};
};

// LWEnKF<BackendTag>::AnalysisResults (LWEnKF.hpp:66-84)
struct LWAnalysisResults {
  double innovation_norm;
  double analysis_increment_norm;
  double background_spread;
  double analysis_spread;
  double max_kalman_gain;
  double min_kalman_gain;
  double condition_number;
  double localization_radius;
  std::string localization_function;
  std::string weighting_scheme;
  double max_weight;
  double min_weight;
  double weight_variance;
  int ensemble_size;
  int observation_count;
  std::string inflation_method;
  double inflation_factor;
};

// LWEnKF<BackendTag> (LWEnKF.hpp:103-334): the root config's "analysis"
// subsection with inflation, localization_radius, output_base_file, format,
// inflation_method, localization_function and weighting_scheme; unknown
// names fall back to multiplicative / gaussian / uniform (:109-145).  The
// N(0,1) draws Z [p][k] are injected (the reference draws them from
// std::random_device, :690-706).  `standard_gc` replaces the reference's
// "gaspari_cohn" polynomial by Gaspari & Cohn (1999) eq. (4.10).
template <class BackendTag>
class LWEnKF {
 public:
  template <class Ens, class Obs, class ObsOp, class Cfg>
  LWEnKF(Ens& ensemble, Obs& obs, const ObsOp& obs_op, const Cfg& root_config,
         std::vector<double> Z, bool standard_gc = false,
         Context& ctx = Context::thread_default())
      : ctx_(ctx), Z_(std::move(Z)) {
    const auto ac = root_config.GetSubsection("analysis");
    prm_.inflation = ac.Get("inflation").asFloat();
    prm_.localization_radius = ac.Get("localization_radius").asFloat();
    output_base_file_ = ac.Get("output_base_file").asString();
    format_ = ac.Get("format").asString();
    const std::string m = ac.Get("inflation_method").asString();
    if (m == "additive") {
      prm_.inflation_method = MDA_INFL_ADDITIVE; method_name_ = m;
    } else if (m == "relaxation") {
      prm_.inflation_method = MDA_INFL_RELAXATION; method_name_ = m;
    } else {
      prm_.inflation_method = MDA_INFL_MULTIPLICATIVE; method_name_ = "multiplicative";
    }
    const std::string f = ac.Get("localization_function").asString();
    if (f == "exponential") {
      prm_.localization = MDA_LW_LOC_EXPONENTIAL; loc_name_ = f;
    } else if (f == "cutoff") {
      prm_.localization = MDA_LW_LOC_CUTOFF; loc_name_ = f;
    } else if (f == "gaspari_cohn") {
      prm_.localization = MDA_LW_LOC_GASPARI_COHN; loc_name_ = f;
    } else {
      prm_.localization = MDA_LW_LOC_GAUSSIAN; loc_name_ = "gaussian";
    }
    const std::string w = ac.Get("weighting_scheme").asString();
    if (w == "adaptive") {
      prm_.weighting = MDA_LW_W_ADAPTIVE; weight_name_ = w;
    } else if (w == "inverse_var") {
      prm_.weighting = MDA_LW_W_INVERSE_VAR; weight_name_ = w;
    } else if (w == "likelihood") {
      prm_.weighting = MDA_LW_W_LIKELIHOOD; weight_name_ = w;
    } else {
      prm_.weighting = MDA_LW_W_UNIFORM; weight_name_ = "uniform";
    }
    prm_.gc_form = standard_gc ? MDA_LW_GC_STANDARD : MDA_LW_GC_REFERENCE;
    run_ = [&ensemble, &obs, &obs_op, this] {
      this->analyse(ensemble, obs, obs_op);
      ensemble.RecomputeMean();  // LWEnKF.hpp:303-304
      ensemble.RecomputePerturbations();
    };
    save_ = [&ensemble, this] { detail::save_ensemble(ensemble, output_base_file_, format_); };
  }
  void Analyse() { run_(); }
  void saveEnsemble() const { save_(); }
  const std::vector<double>& weights() const { return weights_; }
  LWAnalysisResults getAnalysisResults() const {
    LWAnalysisResults r;
    r.innovation_norm = diag_.innovation_norm;
    r.analysis_increment_norm = diag_.analysis_increment_norm;
    r.background_spread = diag_.background_spread;
    r.analysis_spread = diag_.analysis_spread;
    r.max_kalman_gain = diag_.max_kalman_gain;
    r.min_kalman_gain = diag_.min_kalman_gain;
    r.condition_number = diag_.condition_number;
    r.localization_radius = prm_.localization_radius;
    r.localization_function = loc_name_;
    r.weighting_scheme = weight_name_;
    r.max_weight = diag_.max_weight;
    r.min_weight = diag_.min_weight;
    r.weight_variance = diag_.weight_variance;
    r.ensemble_size = diag_.ensemble_size;
    r.observation_count = static_cast<int>(diag_.observation_count);
    r.inflation_method = method_name_;
    r.inflation_factor = prm_.inflation;
    return r;
  }

 private:
  template <class Ens, class Obs, class ObsOp>
  void analyse(Ens& ens, Obs& obs, const ObsOp& op) {
    detail::Staged s = detail::stage(ens, obs, op);
    if (static_cast<int64_t>(Z_.size()) != s.p * s.k)
      throw std::runtime_error("Z must hold p*k standard normal draws");
    mda_ensemble e = detail::ensemble_view(s, s.n, 1, 1);
    mda_obs o = detail::obs_view(s);
    mda_hop h = detail::hx_view(s);
    weights_.assign(static_cast<size_t>(s.k), 0.0);
    check(ctx_.get(),
          mda_lwenkf_analyse(ctx_.get(), &prm_, &e, &o, &h, Z_.data(), 0, weights_.data(),
                             &diag_),
          "LWEnKF::Analyse");
  }
  std::function<void()> run_, save_;
  Context& ctx_;
  mda_lwenkf_params prm_{};
  std::string output_base_file_, format_, method_name_, loc_name_, weight_name_;
  std::vector<double> Z_, weights_;
  mda_lwenkf_diag diag_{};
};

// Lorenz-96 forecast as a metada model backend: it satisfies the
// ModelBackendImpl concept (ModelConcepts.hpp:51-71) -- constructed from the
// config backend, initialize / reset / finalize / isInitialized, run(initial,
// final) on state backends, no default constructor and no copies -- so
// BackendTraits<Tag>::ModelBackend = metada_amd::Lorenz96Model puts the
// library's forecast behind Model<Tag> (Model.hpp:57-177) unchanged.
//
// Config keys, each optional (YamlConfig Get / HasKey, YamlConfig.hpp:132-149;
// reals arrive float-narrowed as every config real does):
//   forcing    F (8.0, lorenz95.py:11)      dt         time step (0.05)
//   num_steps  steps per run (1)            scheme     "euler" (the forward
//              step of lorenz95.py:21-31 with its clipping) or "rk4"
// run() reads the state backend through getData() / size()
// (SimpleState.hpp:69,88); runEnsemble() advances every member of an
// Ensemble in one launch.  4 <= N <= 6144.
class Lorenz96Model {
 public:
  Lorenz96Model() = delete;
  Lorenz96Model(const Lorenz96Model&) = delete;
  Lorenz96Model& operator=(const Lorenz96Model&) = delete;
  Lorenz96Model(Lorenz96Model&& o) noexcept
      : ctx_(o.ctx_), forcing_(o.forcing_), dt_(o.dt_), steps_(o.steps_),
        scheme_(o.scheme_), initialized_(o.initialized_) {
    o.initialized_ = false;
  }
  ~Lorenz96Model() = default;

  template <class ConfigBackend>
  explicit Lorenz96Model(const ConfigBackend& config, Context& ctx = Context::thread_default())
      : ctx_(&ctx) {
    initialize(config);
  }

  template <class ConfigBackend>
  void initialize(const ConfigBackend& config) {
    if (initialized_) return;
    auto real = [&](const char* key, double def) -> double {
      if (!config.HasKey(key)) return def;
      const auto v = config.Get(key);
      return v.isInt() ? static_cast<double>(v.asInt()) : static_cast<double>(v.asFloat());
    };
    forcing_ = real("forcing", 8.0);
    dt_ = real("dt", 0.05);
    steps_ = config.HasKey("num_steps") ? config.Get("num_steps").asInt() : 1;
    const std::string s = config.HasKey("scheme") ? config.Get("scheme").asString() : "euler";
    if (s == "euler") {
      scheme_ = 1;
    } else if (s == "rk4") {
      scheme_ = 0;
    } else {
      throw std::runtime_error("Lorenz96Model: unknown scheme '" + s + "' (euler or rk4)");
    }
    if (steps_ < 0) throw std::runtime_error("Lorenz96Model: num_steps must be >= 0");
    if (!(dt_ > 0.0)) throw std::runtime_error("Lorenz96Model: dt must be positive");
    initialized_ = true;
  }
  bool isInitialized() const { return initialized_; }
  void reset() {
    if (!initialized_) throw std::runtime_error("Lorenz96Model not initialized");
  }
  void finalize() { initialized_ = false; }

  // final = M^num_steps(initial)
  template <class StateBackend>
  void run(const StateBackend& initial, StateBackend& final_state) {
    if (!initialized_) throw std::runtime_error("Lorenz96Model not initialized");
    const double* in = static_cast<const double*>(initial.getData());
    double* out = static_cast<double*>(final_state.getData());
    const size_t n = initial.size();
    if (final_state.size() != n) throw std::runtime_error("Lorenz96Model: state sizes differ");
    if (out != in) std::copy(in, in + n, out);
    advance(out, 1, static_cast<int64_t>(n));
  }

  // Every member of an Ensemble (Ensemble.hpp: Size, GetMember; State.hpp:
  // getDataPtr, size) forecast together, in place.
  template <class Ens>
  void runEnsemble(Ens& ens) {
    if (!initialized_) throw std::runtime_error("Lorenz96Model not initialized");
    const size_t k = ens.Size();
    if (k == 0) return;
    const size_t n = ens.GetMember(0).size();
    std::vector<double> buf(k * n);
    for (size_t j = 0; j < k; ++j) {
      auto& m = ens.GetMember(j);
      if (m.size() != n) throw std::runtime_error("ensemble members differ in size");
      const double* d = m.template getDataPtr<double>();
      std::copy(d, d + n, buf.begin() + j * n);
    }
    advance(buf.data(), static_cast<int64_t>(k), static_cast<int64_t>(n));
    for (size_t j = 0; j < k; ++j)
      std::copy(buf.begin() + j * n, buf.begin() + (j + 1) * n,
                ens.GetMember(j).template getDataPtr<double>());
  }

  double forcing() const { return forcing_; }
  double dt() const { return dt_; }
  int numSteps() const { return steps_; }

 private:
  void advance(double* x, int64_t nstates, int64_t n) {
    if (steps_ == 0) return;
    check(ctx_->get(),
          mda_l96_forecast(ctx_->get(), x, nstates, n, forcing_, dt_, steps_, scheme_, 0),
          "Lorenz96Model::run");
  }
  Context* ctx_;
  double forcing_ = 8.0, dt_ = 0.05;
  int steps_ = 1;
  int32_t scheme_ = 1;
  bool initialized_ = false;
};

}  // namespace metada_amd

#endif  // METADA_AMD_HPP
