// metada_amd.hpp -- header-only C++ adapter that gives the C-ABI of
// metada_amd.h the shape of metada's analysis algorithms, so a backend can
// switch its analysis step to the MI355X library without touching callers.
//
//   metada_amd::LETKF<Tag>   mirrors  LETKF<BackendTag>  (LETKF.hpp:47-142)
//   metada_amd::ETKF<Tag>    mirrors  ETKF<BackendTag>   (ETKF.hpp:84-200)
//   metada_amd::EnKF<Tag>    mirrors  EnKF<BackendTag>   (EnKF.hpp:84-301)
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

}  // namespace metada_amd

#endif  // METADA_AMD_HPP
