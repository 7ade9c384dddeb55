// Correctly rounded double sin, cos and atan2 on the argument ranges of the
// haversine distance (Location.hpp:349-357): sin and cos of |x| <= 8,
// atan2(y, x) of y, x >= 0.  Each is evaluated in double-double arithmetic
// (about 2^-104 relative) and rounded once to double, so it returns the
// double nearest the exact value -- what a correctly rounded C library
// returns.  The device's own sin / cos / atan2 differ from that in the last
// bit for a few percent of arguments, enough to move a distance equal to the
// localisation radius across it; haversine_km_cr below therefore reproduces
// the reference's distance bit for bit wherever its C library rounds
// correctly.  Only boundary cases use it (it is far slower than the library
// functions).
//
// Header-only and host/device, so the CPU test suite compiles the same code
// with g++ and checks it against the C library.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define MDA_HD __host__ __device__
#else
#define MDA_HD
#endif

namespace mda {
namespace cr {

struct dd {
  double hi, lo;
};

MDA_HD inline dd two_sum(double a, double b) {
#pragma clang fp contract(off)
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

MDA_HD inline dd quick_two_sum(double a, double b) {  // |a| >= |b|
#pragma clang fp contract(off)
  const double s = a + b;
  return {s, b - (s - a)};
}

MDA_HD inline dd two_prod(double a, double b) {
#pragma clang fp contract(off)
  const double p = a * b;
  return {p, std::fma(a, b, -p)};
}

MDA_HD inline dd add(dd x, dd y) {
#pragma clang fp contract(off)
  dd s = two_sum(x.hi, y.hi);
  const dd t = two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}

MDA_HD inline dd mul(dd x, dd y) {
#pragma clang fp contract(off)
  dd p = two_prod(x.hi, y.hi);
  p.lo += x.hi * y.lo + x.lo * y.hi;
  return quick_two_sum(p.hi, p.lo);
}

MDA_HD inline dd mul_d(dd x, double d) {
#pragma clang fp contract(off)
  dd p = two_prod(x.hi, d);
  p.lo += x.lo * d;
  return quick_two_sum(p.hi, p.lo);
}

MDA_HD inline dd div_d(dd x, double d) {
#pragma clang fp contract(off)
  const double q1 = x.hi / d;
  const dd p = two_prod(q1, d);
  const double r = ((x.hi - p.hi) - p.lo) + x.lo;
  return quick_two_sum(q1, r / d);
}

MDA_HD inline dd neg(dd x) { return {-x.hi, -x.lo}; }

// the double nearest hi + lo
MDA_HD inline double round_dd(dd x) { return two_sum(x.hi, x.lo).hi; }

// sin and cos of r in double-double for |r| <= pi/4 (Taylor series to r^27)
MDA_HD inline void sincos_reduced(dd r, dd& s, dd& c) {
  const dd r2 = mul(r, r);
  dd t = r, sum_s = r;
  dd u = {1.0, 0.0}, sum_c = {1.0, 0.0};
  for (int i = 1; i <= 13; ++i) {
    t = div_d(mul(t, r2), (double)((2 * i) * (2 * i + 1)));
    u = div_d(mul(u, r2), (double)((2 * i - 1) * (2 * i)));
    sum_s = add(sum_s, (i & 1) ? neg(t) : t);
    sum_c = add(sum_c, (i & 1) ? neg(u) : u);
  }
  s = sum_s;
  c = sum_c;
}

// sin x and cos x in double-double for |x| <= 8: x = r + n pi/2, |r| <= pi/4,
// with pi/2 split into parts whose products by n (|n| <= 5) are exact
// (P1 33 bits, P2 and P3 with trailing zero bits, P4 the rest): the reduced
// argument carries an error below 1e-30, far inside the rounding of the
// result for every argument of this range
MDA_HD inline void sincos_dd(double x, dd& s, dd& c) {
#pragma clang fp contract(off)
  constexpr double P1 = 0x1.921fb544p+0;
  constexpr double P2 = 0x1.0b4611a6p-34;
  constexpr double P3 = 0x1.3198a2ep-69;
  constexpr double P4 = 0x1.b839a252049c1p-104;
  const double nf = std::nearbyint(x * 0.63661977236758134308);  // x / (pi/2)
  const int n = (int)nf;
  // x - n P1 is exact (|x| and |n P1| within a factor 2, or n = 0)
  dd r = two_sum(x - nf * P1, -nf * P2);
  r = add(r, dd{-nf * P3, -nf * P4});
  dd sr, cr;
  sincos_reduced(r, sr, cr);
  switch (n & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = neg(sr); break;
    case 2: s = neg(sr); c = neg(cr); break;
    default: s = neg(cr); c = sr; break;
  }
}

MDA_HD inline double sin_cr(double x) {
  dd s, c;
  sincos_dd(x, s, c);
  return round_dd(s);
}

MDA_HD inline double cos_cr(double x) {
  dd s, c;
  sincos_dd(x, s, c);
  return round_dd(c);
}

// atan2(y, x) for y, x >= 0: one Newton step in double-double on
// f(t) = y cos t - x sin t from the library's estimate t0 (about 1 ulp), so
// t = t0 + (y cos t0 - x sin t0) / (x cos t0 + y sin t0) to about 2^-104
MDA_HD inline double atan2_cr(double y, double x) {
  if (y == 0.0 && x == 0.0) return 0.0;
  const double t0 = std::atan2(y, x);
  dd s, c;
  sincos_dd(t0, s, c);
  const dd num = add(mul_d(c, y), neg(mul_d(s, x)));
  const dd den = add(mul_d(c, x), mul_d(s, y));
  const double corr = num.hi / den.hi;
  return round_dd(add(dd{t0, 0.0}, dd{corr, 0.0}));
}

// Location::haversine (Location.hpp:349-357): the same formula, constants
// and operation order, every product and sum rounded separately, with
// correctly rounded sin, cos and atan2.
MDA_HD inline double haversine_km_cr(double lat1, double lon1, double lat2, double lon2) {
#pragma clang fp contract(off)
  constexpr double PI = 3.14159265358979323846;
  const double dlat = (lat2 - lat1) * PI / 180.0;
  const double dlon = (lon2 - lon1) * PI / 180.0;
  const double s1 = sin_cr(dlat / 2), s2 = sin_cr(dlon / 2);
  const double a =
      s1 * s1 + cos_cr(lat1 * PI / 180.0) * cos_cr(lat2 * PI / 180.0) * s2 * s2;
  const double c = 2 * atan2_cr(std::sqrt(a), std::sqrt(1 - a));
  return 6371.0 * c;
}

}  // namespace cr
}  // namespace mda
