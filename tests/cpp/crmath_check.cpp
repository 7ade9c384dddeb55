// Checks the correctly rounded sin / cos / atan2 / haversine of
// metada_amd/csrc/mda_crmath.hpp against the C library on random arguments
// of the haversine's ranges; prints the number of results that differ in
// any bit.  Built with g++ by the CPU test suite.
#include <cmath>
#include <cstdio>
#include <random>

#include "mda_crmath.hpp"

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> ang(-3.14159265358979, 3.14159265358979);
  std::uniform_real_distribution<double> wide(-7.5, 7.5);
  std::uniform_real_distribution<double> unit(0.0, 1.0);
  std::uniform_real_distribution<double> lat(-90.0, 90.0), lon(-180.0, 360.0);
  const int n = 2000000;
  long bad_sin = 0, bad_cos = 0, bad_atan2 = 0, bad_hav = 0;
  for (int i = 0; i < n; ++i) {
    const double x = (i & 1) ? ang(rng) : wide(rng);
    bad_sin += mda::cr::sin_cr(x) != std::sin(x);
    bad_cos += mda::cr::cos_cr(x) != std::cos(x);
    const double a = unit(rng);
    const double y = std::sqrt(a), z = std::sqrt(1 - a);
    bad_atan2 += mda::cr::atan2_cr(y, z) != std::atan2(y, z);
    const double la1 = lat(rng), lo1 = lon(rng), la2 = lat(rng), lo2 = lon(rng);
    const double pi = 3.14159265358979323846;
    const double dlat = (la2 - la1) * pi / 180.0, dlon = (lo2 - lo1) * pi / 180.0;
    const double aa = std::sin(dlat / 2) * std::sin(dlat / 2) +
                      std::cos(la1 * pi / 180.0) * std::cos(la2 * pi / 180.0) *
                          std::sin(dlon / 2) * std::sin(dlon / 2);
    const double ref = 6371.0 * (2 * std::atan2(std::sqrt(aa), std::sqrt(1 - aa)));
    bad_hav += mda::cr::haversine_km_cr(la1, lo1, la2, lo2) != ref;
  }
  std::printf("n %d sin %ld cos %ld atan2 %ld haversine %ld\n", n, bad_sin, bad_cos, bad_atan2,
              bad_hav);
  return 0;
}
