// C entry points over metada_amd/csrc/mda_crmath.hpp for the CPU tests
// (ctypes); built with g++ by tests/test_crmath_cpu.py.
#include "mda_crmath.hpp"

extern "C" {
double cr_sin(double x) { return mda::cr::sin_cr(x); }
double cr_cos(double x) { return mda::cr::cos_cr(x); }
double cr_atan2(double y, double x) { return mda::cr::atan2_cr(y, x); }
double cr_haversine(double a, double b, double c, double d) {
  return mda::cr::haversine_km_cr(a, b, c, d);
}
}
