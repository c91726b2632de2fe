// Host-side helpers shared by the engine translation units.
#pragma once
#include <stdint.h>

namespace mpcx {
// records the message of the last failure (per calling thread) and returns code
int set_error(int code, const char* msg);
// x**n with an integer exponent, evaluated by repeated multiplication; exact
// for the integer and dyadic radii the simulators use
inline double ipow(double x, int n) {
  double r = x;
  for (int i = 1; i < n; i++) r = r * x;
  return r;
}
}  // namespace mpcx
