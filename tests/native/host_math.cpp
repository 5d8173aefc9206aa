// Host (g++) build of the kernels' DD headers, for CPU tests only: checks the
// device algorithms (multifloats_amd/csrc/dd_core.h, dd_math.h) against the
// oracle and the 320-bit fixture values before they run on a GPU.
#include "../../multifloats_amd/csrc/dd_math.h"

#define H1(name, fn)                                                   \
  extern "C" void name(const double *a, double *c, long n) {           \
    for (long i = 0; i < n; ++i) {                                     \
      ddv r = fn(mk(a[2 * i], a[2 * i + 1]));                          \
      c[2 * i] = r.hi;                                                 \
      c[2 * i + 1] = r.lo;                                             \
    }                                                                  \
  }
#define H2(name, fn)                                                          \
  extern "C" void name(const double *a, const double *b, double *c, long n) { \
    for (long i = 0; i < n; ++i) {                                            \
      ddv r = fn(mk(a[2 * i], a[2 * i + 1]), mk(b[2 * i], b[2 * i + 1]));     \
      c[2 * i] = r.hi;                                                        \
      c[2 * i + 1] = r.lo;                                                    \
    }                                                                         \
  }
H1(h_exp, dd_exp)
H1(h_log, dd_log)
H1(h_sin, dd_sin)
H1(h_erf, dd_erf)
H1(h_erfc, dd_erfc)
H1(h_erfcx, dd_erfcx)
H1(h_sqrt, dd_sqrt)
H2(h_add, dd_add)
H2(h_sub, dd_sub)
H2(h_mul, dd_mul)
H2(h_div, dd_div)
H1(h_atan, dd_atan)
H1(h_atanpi, dd_atanpi)
H2(h_atan2, dd_atan2)
H2(h_atan2pi, dd_atan2pi)
H1(h_asin, dd_asin)
H1(h_acos, dd_acos)
H1(h_asinpi, dd_asinpi)
H1(h_acospi, dd_acospi)
H2(h_fmod, dd_fmod)
H2(h_modulo, dd_modulo)
H2(h_nearest, dd_nearest)
