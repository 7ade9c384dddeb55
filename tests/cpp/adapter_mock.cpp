// Runs the C++ adapter (include/metada_amd.hpp) against stand-ins that
// reproduce the reference adapter API's signatures and behaviour where the
// adapter depends on them:
//   State::getData<T>() dereferences static_cast<T*>(backend getData()), a
//   double* for SimpleState (State.hpp:209-221, SimpleState.hpp:69) -- so
//   reading it as a std::vector would be undefined; getDataPtr<T>() returns
//   the pointer (State.hpp:229-241); size(); geometry(); saveToFile()
//   (SimpleState.hpp:162-188: rows of nx values, %12.6f);
//   Ensemble::RecomputeMean / RecomputePerturbations / Mean / GetMember
//   (Ensemble.hpp:71-185);
//   Config::Get -> ConfigValue (no Get of its own), Config::GetSubsection
//   (Config.hpp:203-313); ConfigValue reals are float and asFloat / asString
//   throw on a wrong type (ConfigValue.hpp:108-109);
//   Location with a private coordinate system behind getCoordinateSystem()
//   and integer grid coordinates (Location.hpp:100-180, 303);
//   Observation iteration over points with .location, getObservationValues,
//   getCovariance (Observation.hpp:261-482; GridObservation.hpp:197-252);
//   ObsOperator::apply(state, obs) (ObsOperator.hpp:255-259), here the IDW4
//   operator of IdentityObsOperator through the oracle (oracle.h);
//   Model<Tag> forwarding to its backend with the error wrapping of
//   Model.hpp:57-177, over metada_amd::Lorenz96Model.
//
// usage: adapter_mock <input file> <output prefix>
//   input: algo form nx ny nlev k p inflation radius  (or the l96 form of run_l96)
//          k*n member values, p lines "x y z value error", p*k Z values (EnKF)
//   output: <prefix>.analysis (k*n values, %.17g), <prefix>.mean, diagnostics
//           on stdout, and the files saveEnsemble() writes.
#include <cmath>
#include <concepts>
#include <cstdio>
#include <limits>
#include <fstream>
#include <iomanip>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <variant>
#include <vector>

#include "metada_amd.hpp"
#include "oracle.h"

namespace ref {
enum class CoordinateSystem { GRID, GEOGRAPHIC, CARTESIAN };

class Location {
 public:
  Location(int i, int j, int k) : system(CoordinateSystem::GRID), c(i, j, k) {}
  CoordinateSystem getCoordinateSystem() const { return system; }
  std::tuple<int, int, int> getGridCoords() const {
    if (system != CoordinateSystem::GRID) throw std::runtime_error("not a grid location");
    return c;
  }
  std::tuple<double, double, double> getGeographicCoords() const {
    throw std::runtime_error("not a geographic location");
  }
  std::tuple<double, double, double> getCartesianCoords() const {
    throw std::runtime_error("not a cartesian location");
  }

 private:
  CoordinateSystem system;
  std::tuple<int, int, int> c;
};

struct Geometry {  // SimpleGeometry: y-major list of grid points
  std::vector<Location> pts;
  Geometry(int nx, int ny) {
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) pts.emplace_back(x, y, 0);
  }
  auto begin() const { return pts.begin(); }
  auto end() const { return pts.end(); }
};

class SimpleState {  // the state backend: getData() is the raw buffer (SimpleState.hpp:69,88)
 public:
  explicit SimpleState(std::vector<double> v) : data_(std::move(v)) {}
  SimpleState(const SimpleState&) = delete;
  SimpleState& operator=(const SimpleState&) = delete;
  SimpleState(SimpleState&&) = default;
  SimpleState& operator=(SimpleState&&) = default;
  void* getData() { return data_.data(); }
  const void* getData() const { return data_.data(); }
  size_t size() const { return data_.size(); }
  std::vector<double> data_;
};

class State {  // State<Tag> over a SimpleState backend
 public:
  State(const Geometry* g, int nx, int ny, std::vector<double> v)
      : g_(g), nx_(nx), ny_(ny), backend_(std::move(v)) {}
  State clone() const { return State(g_, nx_, ny_, backend_.data_); }
  template <class T> T& getData() { return *static_cast<T*>(backend_.getData()); }
  template <class T> T* getDataPtr() { return static_cast<T*>(backend_.getData()); }
  template <class T> const T* getDataPtr() const { return static_cast<const T*>(backend_.getData()); }
  size_t size() const { return backend_.size(); }
  SimpleState& backend() { return backend_; }
  const SimpleState& backend() const { return backend_; }
  const Geometry* geometry() const { return g_; }
  void zero() { std::fill(backend_.data_.begin(), backend_.data_.end(), 0.0); }
  State& operator+=(const State& o) {
    for (size_t i = 0; i < backend_.data_.size(); ++i) backend_.data_[i] += o.backend_.data_[i];
    return *this;
  }
  State& operator-=(const State& o) {
    for (size_t i = 0; i < backend_.data_.size(); ++i) backend_.data_[i] -= o.backend_.data_[i];
    return *this;
  }
  State& operator*=(double s) {
    for (double& v : backend_.data_) v *= s;
    return *this;
  }
  void saveToFile(const std::string& name) const {
    std::ofstream f(name);
    if (!f) throw std::runtime_error("cannot open " + name);
    const size_t levels = backend_.data_.size() / ((size_t)nx_ * ny_);
    for (size_t r = 0; r < (size_t)ny_ * levels; ++r) {
      for (int x = 0; x < nx_; ++x) {
        if (x) f << " ";
        f << std::fixed << std::setprecision(6) << std::setw(12) << backend_.data_[r * nx_ + x];
      }
      f << "\n";
    }
  }
  const std::vector<double>& raw() const { return backend_.data_; }

 private:
  const Geometry* g_;
  int nx_, ny_;
  SimpleState backend_;
};

class Ensemble {
 public:
  explicit Ensemble(std::vector<State> m) : members_(std::move(m)) {}
  size_t Size() const { return members_.size(); }
  State& GetMember(size_t i) {
    if (i >= members_.size()) throw std::out_of_range("Ensemble member index out of range");
    return members_[i];
  }
  void RecomputeMean() {
    if (!mean_) mean_ = std::make_unique<State>(members_[0].clone());
    mean_->zero();
    for (auto& m : members_) *mean_ += m;
    *mean_ *= (1.0 / static_cast<double>(members_.size()));
  }
  State& Mean() {
    if (!mean_) RecomputeMean();
    return *mean_;
  }
  void RecomputePerturbations() {
    RecomputeMean();
    perts_.clear();
    for (auto& m : members_) {
      perts_.push_back(std::make_unique<State>(m.clone()));
      *perts_.back() -= *mean_;
    }
  }
  size_t perturbations() const { return perts_.size(); }

 private:
  std::vector<State> members_;
  std::unique_ptr<State> mean_;
  std::vector<std::unique_ptr<State>> perts_;
};

struct ObservationPoint {
  Location location;
  double value, error;
  bool is_valid;
};

class Observation {
 public:
  std::vector<ObservationPoint> pts;
  std::vector<double> getObservationValues() const {
    std::vector<double> r;
    for (const auto& p : pts) r.push_back(p.value);
    return r;
  }
  std::vector<double> getCovariance() const {  // GridObservation.hpp:239-252
    std::vector<double> r;
    for (const auto& p : pts)
      r.push_back(p.is_valid ? p.error * p.error : std::numeric_limits<double>::infinity());
    return r;
  }
  size_t size() const { return pts.size(); }
  auto begin() const { return pts.begin(); }
  auto end() const { return pts.end(); }
};

class ObsOperator {  // IdentityObsOperator's IDW4 (IdentityObsOperator.hpp:594-676)
 public:
  ObsOperator(int nx, int ny, int nz) : nx_(nx), ny_(ny), nz_(nz) {}
  std::vector<double> apply(const State& s, const Observation& o) const {
    std::vector<double> r;
    for (const auto& p : o.pts) {
      const auto [i, j, k] = p.location.getGridCoords();
      r.push_back(orc_idw4(s.getDataPtr<double>(), nx_, ny_, nz_, i, j, k));
    }
    return r;
  }

 private:
  int nx_, ny_, nz_;
};

class ConfigValue {  // reals are float (ConfigValue.hpp:77-108)
 public:
  ConfigValue() = default;
  ConfigValue(int i) : v_(i) {}
  ConfigValue(float f) : v_(f) {}
  ConfigValue(std::string s) : v_(std::move(s)) {}
  bool isInt() const { return std::holds_alternative<int>(v_); }
  int asInt() const { return std::get<int>(v_); }
  float asFloat() const { return std::get<float>(v_); }
  const std::string& asString() const { return std::get<std::string>(v_); }

 private:
  std::variant<std::monostate, int, float, std::string> v_;
};

class Config {
 public:
  std::vector<std::pair<std::string, ConfigValue>> values;
  std::vector<std::pair<std::string, Config>> subsections;
  ConfigValue Get(const std::string& key, const ConfigValue& def = ConfigValue()) const {
    for (const auto& kv : values)
      if (kv.first == key) return kv.second;
    return def;  // Config::Get swallows a missing key into the default
  }
  bool HasKey(const std::string& key) const {  // YamlConfig.hpp:149
    for (const auto& kv : values)
      if (kv.first == key) return true;
    return false;
  }
  Config GetSubsection(const std::string& key) const {
    for (const auto& kv : subsections)
      if (kv.first == key) return kv.second;
    throw std::runtime_error("Config subsection not found: " + key);
  }
};

// Model<Tag> (Model.hpp:57-177): forwards to the backend, wrapping errors
template <class Backend>
class Model {
 public:
  explicit Model(const Config& config) : backend_(config) {}
  void initialize(const Config& config) {
    if (backend_.isInitialized()) return;
    try {
      backend_.initialize(config);
    } catch (const std::exception& e) {
      throw std::runtime_error(std::string("Model initialization failed: ") + e.what());
    }
  }
  bool isInitialized() const { return backend_.isInitialized(); }
  void run(const State& initial, State& final_state) {
    if (!backend_.isInitialized()) throw std::runtime_error("Model not initialized");
    try {
      backend_.run(initial.backend(), final_state.backend());
    } catch (const std::exception& e) {
      throw std::runtime_error(std::string("Model run failed: ") + e.what());
    }
  }
  void reset() { backend_.reset(); }
  void finalize() {
    if (backend_.isInitialized()) backend_.finalize();
  }
  Backend& backend() { return backend_; }

 private:
  Backend backend_;
};

// The requirements ModelBackendImpl places on a model backend
// (ModelConcepts.hpp:51-71), checked against metada_amd::Lorenz96Model.
template <typename T, typename ConfigBackend, typename StateBackend>
concept ModelBackendImpl =
    requires(T& model, const ConfigBackend& config, const StateBackend& initialState,
             StateBackend& finalState) {
      { T(config) } -> std::same_as<T>;
      { model.initialize(config) } -> std::same_as<void>;
      { model.reset() } -> std::same_as<void>;
      { model.finalize() } -> std::same_as<void>;
      { model.run(initialState, finalState) } -> std::same_as<void>;
      requires !std::is_default_constructible_v<T>;
      requires !std::is_copy_constructible_v<T>;
      requires !std::is_copy_assignable_v<T>;
    };
static_assert(ModelBackendImpl<metada_amd::Lorenz96Model, Config, SimpleState>);

struct Tag {};
}  // namespace ref

// l96 <scheme> <N> <k> <num_steps> <dt> <forcing>, then k*N values: each
// member through Model<Lorenz96Model>::run into a fresh state, and the whole
// ensemble through runEnsemble; writes <prefix>.single and <prefix>.ensemble.
static int run_l96(std::ifstream& in, const std::string& out) {
  std::string scheme;
  int N, k, steps;
  float dt, forcing;
  in >> scheme >> N >> k >> steps >> dt >> forcing;
  std::vector<ref::State> members;
  ref::Geometry geom(N, 1);
  for (int j = 0; j < k; ++j) {
    std::vector<double> v(N);
    for (auto& x : v) in >> x;
    members.emplace_back(&geom, N, 1, std::move(v));
  }
  if (!in) throw std::runtime_error("input file is short");
  ref::Config cfg;
  cfg.values = {{"scheme", ref::ConfigValue(scheme)}, {"num_steps", ref::ConfigValue(steps)},
                {"dt", ref::ConfigValue(dt)}, {"forcing", ref::ConfigValue(forcing)}};
  ref::Model<metada_amd::Lorenz96Model> model(cfg);
  model.initialize(cfg);
  std::ofstream fs(out + ".single"), fe(out + ".ensemble");
  fs << std::setprecision(17);
  fe << std::setprecision(17);
  for (int j = 0; j < k; ++j) {
    ref::State fc(&geom, N, 1, std::vector<double>(N, 0.0));
    model.run(members[j], fc);
    for (double v : fc.raw()) fs << v << "\n";
  }
  ref::Ensemble ens(std::move(members));
  model.backend().runEnsemble(ens);
  for (int j = 0; j < k; ++j)
    for (double v : ens.GetMember(j).raw()) fe << v << "\n";
  // a bad scheme surfaces through Model's error wrapping
  ref::Config bad;
  bad.values = {{"scheme", ref::ConfigValue(std::string("leapfrog"))}};
  try {
    ref::Model<metada_amd::Lorenz96Model> m2(bad);
    std::printf("bad_scheme accepted\n");
  } catch (const std::exception& e) {
    std::printf("bad_scheme %s\n", e.what());
  }
  model.finalize();
  std::printf("initialized_after_finalize %d\n", (int)model.isInitialized());
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <input> <output prefix>\n", argv[0]);
    return 2;
  }
  try {
    std::ifstream in(argv[1]);
    std::string algo, form;
    in >> algo;
    if (algo == "l96") return run_l96(in, argv[2]);
    int nx, ny, nlev, k, p;
    double inflation, radius;
    in >> form >> nx >> ny >> nlev >> k >> p >> inflation >> radius;
    const size_t n = (size_t)nx * ny * nlev;
    ref::Geometry geom(nx, ny);
    std::vector<ref::State> members;
    for (int j = 0; j < k; ++j) {
      std::vector<double> v(n);
      for (auto& x : v) in >> x;
      members.emplace_back(&geom, nx, ny, std::move(v));
    }
    ref::Ensemble ens(std::move(members));
    ref::Observation obs;
    for (int o = 0; o < p; ++o) {
      int x, y, z;
      double val, err;
      in >> x >> y >> z >> val >> err;
      obs.pts.push_back({ref::Location(x, y, z), val, err, std::isfinite(err)});
    }
    std::vector<double> Z((size_t)p * k);
    if (algo == "enkf" || algo == "lwenkf")
      for (auto& z : Z) in >> z;
    if (!in) throw std::runtime_error("input file is short");
    const std::string out = argv[2];
    ref::ObsOperator op(nx, ny, nlev);
    ref::Config analysis;
    analysis.values = {{"inflation", ref::ConfigValue((float)inflation)},
                       {"localization_radius", ref::ConfigValue((float)radius)},
                       {"output_base_file", ref::ConfigValue(out + "_saved")},
                       {"format", ref::ConfigValue(std::string("txt"))},
                       {"inflation_method", ref::ConfigValue(std::string("relaxation"))}};
    metada_amd::Context ctx(0);
    if (algo == "letkf") {
      metada_amd::LETKF<ref::Tag> f(ens, obs, op, analysis,
                                    form == "refcompat"
                                        ? metada_amd::LETKF<ref::Tag>::Form::RefCompat
                                        : metada_amd::LETKF<ref::Tag>::Form::Standard,
                                    ctx);
      f.Analyse();
      f.saveEnsemble();
      std::printf("local_obs %lld\n", (long long)f.diagnostics().total_local_obs);
    } else if (algo == "etkf") {
      metada_amd::ETKF<ref::Tag> f(ens, obs, op, analysis, ctx);
      f.Analyse();
      f.saveEnsemble();
    } else if (algo == "lwenkf") {
      // form = <localization_function>:<weighting_scheme>[:standard]
      const size_t c1 = form.find(':');
      const size_t c2 = form.find(':', c1 + 1);
      analysis.values.push_back({"localization_function", ref::ConfigValue(form.substr(0, c1))});
      analysis.values.push_back(
          {"weighting_scheme", ref::ConfigValue(form.substr(c1 + 1, c2 - c1 - 1))});
      const bool standard_gc = c2 != std::string::npos && form.substr(c2 + 1) == "standard";
      ref::Config root;
      root.subsections = {{"analysis", analysis}};
      metada_amd::LWEnKF<ref::Tag> f(ens, obs, op, root, Z, standard_gc, ctx);
      f.Analyse();
      f.saveEnsemble();
      const metada_amd::LWAnalysisResults r = f.getAnalysisResults();
      std::printf("innovation_norm %.17g\nbackground_spread %.17g\nanalysis_spread %.17g\n"
                  "max_kalman_gain %.17g\nmin_kalman_gain %.17g\ncondition_number %.17g\n"
                  "max_weight %.17g\nmin_weight %.17g\nweight_variance %.17g\n"
                  "localization_function %s\nweighting_scheme %s\ninflation_method %s\n"
                  "localization_radius %.17g\nperturbations %zu\n",
                  r.innovation_norm, r.background_spread, r.analysis_spread, r.max_kalman_gain,
                  r.min_kalman_gain, r.condition_number, r.max_weight, r.min_weight,
                  r.weight_variance, r.localization_function.c_str(), r.weighting_scheme.c_str(),
                  r.inflation_method.c_str(), r.localization_radius, ens.perturbations());
    } else {
      ref::Config root;
      root.subsections = {{"analysis", analysis}};
      metada_amd::EnKF<ref::Tag> f(ens, obs, op, root, Z, ctx);
      f.Analyse();
      f.saveEnsemble();
      const metada_amd::AnalysisResults r = f.getAnalysisResults();
      std::printf("innovation_norm %.17g\nbackground_spread %.17g\nanalysis_spread %.17g\n"
                  "max_kalman_gain %.17g\nmin_kalman_gain %.17g\ncondition_number %.17g\n"
                  "ensemble_size %d\nobservation_count %d\ninflation_method %s\n"
                  "inflation_factor %.17g\nperturbations %zu\n",
                  r.innovation_norm, r.background_spread, r.analysis_spread,
                  r.max_kalman_gain, r.min_kalman_gain, r.condition_number, r.ensemble_size,
                  r.observation_count, r.inflation_method.c_str(), r.inflation_factor,
                  ens.perturbations());
    }
    std::ofstream fa(out + ".analysis"), fm(out + ".mean");
    fa << std::setprecision(17);
    fm << std::setprecision(17);
    for (int j = 0; j < k; ++j)
      for (double v : ens.GetMember(j).raw()) fa << v << "\n";
    for (double v : ens.Mean().raw()) fm << v << "\n";
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
