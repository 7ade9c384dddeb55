// Instantiates the C++ adapter (include/metada_amd.hpp) against minimal
// stand-ins with the reference adapter API's member names, so that a change
// to the adapter that breaks the interface it binds to fails to compile.
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "metada_amd.hpp"

namespace mock {
enum class CoordinateSystem { GRID, GEOGRAPHIC, CARTESIAN };
struct Location {
  CoordinateSystem system = CoordinateSystem::GRID;
  double x = 0, y = 0, z = 0;
  std::tuple<int, int, int> getGridCoords() const { return {int(x), int(y), int(z)}; }
  std::tuple<double, double, double> getGeographicCoords() const { return {x, y, z}; }
  std::tuple<double, double, double> getCartesianCoords() const { return {x, y, z}; }
};
struct Geometry {
  std::vector<Location> pts;
  auto begin() const { return pts.begin(); }
  auto end() const { return pts.end(); }
};
struct State {
  std::vector<double> v;
  const Geometry* g = nullptr;
  template <class T> T& getData() { return v; }
  const Geometry* geometry() const { return g; }
};
struct Ensemble {
  std::vector<State> m;
  size_t Size() const { return m.size(); }
  State& GetMember(size_t i) { return m.at(i); }
};
struct ObsPoint { Location location; double value, error; };
struct Observation {
  std::vector<ObsPoint> pts;
  std::vector<double> getObservationValues() const {
    std::vector<double> r; for (auto& p : pts) r.push_back(p.value); return r;
  }
  std::vector<double> getCovariance() const {
    std::vector<double> r; for (auto& p : pts) r.push_back(p.error * p.error); return r;
  }
  auto begin() const { return pts.begin(); }
  auto end() const { return pts.end(); }
};
struct ObsOperator {
  std::vector<double> apply(const State& s, const Observation& o) const {
    std::vector<double> r; for (auto& p : o.pts) r.push_back(s.v[size_t(p.location.x)]); return r;
  }
};
struct Value { double d; double asFloat() const { return double(float(d)); } };
struct Config {
  double infl = 1.05, radius = 2.0;
  Value Get(const std::string& k) const { return {k == "inflation" ? infl : radius}; }
};
struct RootConfig {
  Config c;
  const Config& Get(const std::string&) const { return c; }
};
struct Tag {};
}  // namespace mock

int run_all(mock::Ensemble& e, mock::Observation& o, const mock::ObsOperator& h,
            const mock::Config& c, const mock::RootConfig& rc) {
  metada_amd::Context ctx(0);
  metada_amd::LETKF<mock::Tag> letkf(e, o, h, c, metada_amd::LETKF<mock::Tag>::Form::Standard, ctx);
  letkf.Analyse();
  metada_amd::ETKF<mock::Tag> etkf(e, o, h, c, ctx);
  etkf.Analyse();
  std::vector<double> Z(o.pts.size() * e.Size(), 0.0);
  metada_amd::EnKF<mock::Tag> enkf(e, o, h, rc, Z, ctx);
  enkf.Analyse();
  return int(letkf.diagnostics().n_columns + enkf.getAnalysisResults().innovation_norm);
}
