// Host-side helpers shared by the engine translation units.
#pragma once
#include <stdint.h>

struct mpcx_co;
struct mpcx_sk;

namespace mpcx {
// records the message of the last failure (per calling thread) and returns code
int set_error(int code, const char* msg);
// live handles of the C ABI, released together by mpcx_shutdown
enum HandleKind { H_COB = 0, H_CO = 1, H_SK = 2, H_MLP = 3, H_KINDS = 4 };
void handle_add(int kind, void* p);
void handle_del(int kind, void* p);
// x**n with an integer exponent, evaluated by repeated multiplication; exact
// for the integer and dyadic radii the simulators use
inline double ipow(double x, int n) {
  double r = x;
  for (int i = 1; i < n; i++) r = r * x;
  return r;
}

// Device arrays of one engine's system, for the diagnostics kernels.
struct DevView {
  const double *x, *y, *z, *vx, *vy, *vz;
  const int* flag;
  long long n;            // particle slots
  double box[3];
  double mass[4];         // particle mass by flag value 0..3 (0 for flags that are not particles)
  int species[4];         // species column of each flag value (-1: not counted)
  int periodic_z;         // collective_code has walls in z
  double dim[12][240];    // sphere fields (x y z vx vy vz fx fy fz ...) of up to 120 dimers, s = k + i*nd
  int nd;                 // dimers
  double msph[2];         // C and N sphere masses
  double tpotential;
  void* stream;           // hipStream_t of the engine
};
int co_view(struct ::mpcx_co* e, DevView* v);
int sk_view(struct ::mpcx_sk* e, DevView* v);
}  // namespace mpcx
