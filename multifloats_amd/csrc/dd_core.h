// Double-double (DD) core for the gfx950 kernels: error-free transformations,
// the compensated + - * / sqrt, and the GEMM multiply-accumulate.
//
// Every function keeps the exact operation order of the reference so that a
// kernel built from them (with -ffp-contract=off, which the build enforces and
// the pragma below repeats) returns the same bits as the reference CPU code:
//   two_sum / fast_two_sum / two_prod   include/multifloats/float64x2.h:83-107
//   dd_add / dd_sub                     float64x2.h:203-250
//   dd_mul                              float64x2.h:253-271
//   dd_div                              float64x2.h:274-311
//   dd_sqrt                             float64x2.h:2190-2213
//   mac / renorm                        src/float64x2/matmul.inc:6-39
// Non-finite short-circuits are kept because they define the lo limb of a
// non-finite result (NaN-in/NaN-out is the library's whole error model).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MF_HD __device__ __forceinline__
#define MF_HDM __device__ __forceinline__  // member functions
#else
#include <math.h>
#define MF_HD static inline
#define MF_HDM inline
#endif

#pragma clang fp contract(off)

struct ddv {
  double hi, lo;
};

MF_HD ddv mk(double h, double l) {
  ddv r;
  r.hi = h;
  r.lo = l;
  return r;
}

MF_HD bool mf_isfinite(double x) { return __builtin_isfinite(x); }

MF_HD void two_sum(double a, double b, double &s, double &e) {
  double t = a + b;
  double ap = t - b;
  double bp = t - ap;
  e = (a - ap) + (b - bp);
  s = t;
}

MF_HD void fast_two_sum(double a, double b, double &s, double &e) {
  double t = a + b;
  double bp = t - a;
  e = b - bp;
  s = t;
}

MF_HD void two_prod(double a, double b, double &p, double &e) {
  double q = a * b;
  e = __builtin_fma(a, b, -q);
  p = q;
}

MF_HD ddv dd_neg(ddv x) { return mk(-x.hi, -x.lo); }

MF_HD ddv dd_add(ddv x, ddv y) {
  double s = x.hi + y.hi;
  if (!mf_isfinite(s)) return mk(s, 0.0);
  if (x.hi == 0.0 && y.hi == 0.0) return mk(s, x.lo + y.lo);
  double a, b, c, d;
  two_sum(x.hi, y.hi, a, b);
  two_sum(x.lo, y.lo, c, d);
  fast_two_sum(a, c, a, c);
  b = b + d;
  b = b + c;
  ddv r;
  fast_two_sum(a, b, r.hi, r.lo);
  return r;
}

MF_HD ddv dd_sub(ddv x, ddv y) {
  double s = x.hi - y.hi;
  if (!mf_isfinite(s)) return mk(s, 0.0);
  if (x.hi == 0.0 && y.hi == 0.0) return mk(s, x.lo - y.lo);
  double a, b, c, d;
  two_sum(x.hi, -y.hi, a, b);
  two_sum(x.lo, -y.lo, c, d);
  fast_two_sum(a, c, a, c);
  b = b + d;
  b = b + c;
  ddv r;
  fast_two_sum(a, b, r.hi, r.lo);
  return r;
}

MF_HD ddv dd_mul(ddv x, ddv y) {
  double p00, e00;
  two_prod(x.hi, y.hi, p00, e00);
  if (!mf_isfinite(p00)) return mk(p00, 0.0);
  double p01 = x.hi * y.lo;
  double p10 = x.lo * y.hi;
  p01 = p01 + p10;
  e00 = e00 + p01;
  ddv r;
  fast_two_sum(p00, e00, r.hi, r.lo);
  return r;
}

MF_HD ddv dd_div(ddv x, ddv y) {
  double q1 = x.hi / y.hi;
  if (!mf_isfinite(q1)) return mk(q1, q1);
  if (!mf_isfinite(y.hi)) return mk(q1, 0.0);
  double p00, e00;
  two_prod(q1, y.hi, p00, e00);
  double p01 = q1 * y.lo;
  double qlo = e00 + p01;
  double r0, r0e;
  two_sum(x.hi, -p00, r0, r0e);
  double r1 = (x.lo - qlo) + r0e;
  double rh, rl;
  fast_two_sum(r0, r1, rh, rl);
  double q2 = rh / y.hi;
  ddv r;
  fast_two_sum(q1, q2, r.hi, r.lo);
  return r;
}

MF_HD ddv dd_sqrt(ddv x) {
  double c = __builtin_sqrt(x.hi);
  if (!(x.hi > 0.0) || !mf_isfinite(c)) return mk(c, 0.0);
  double u, uu;
  two_prod(c, c, u, uu);
  double cc = (((x.hi - u) - uu) + x.lo) / (2.0 * c);
  double h = c + cc;
  return mk(h, (c - h) + cc);
}

// One compensated multiply-accumulate of the reference GEMM: 2 MUL, 2 FMA,
// 9 ADD. s_lo absorbs the product error, the cross terms and the TwoSum error.
MF_HD void mac(double ah, double al, double bh, double bl, double &s_hi, double &s_lo) {
  double p = ah * bh;
  double e = __builtin_fma(ah, bh, -p);
  double cross = __builtin_fma(ah, bl, al * bh);
  double t = s_hi + p;
  double bp = t - s_hi;
  double ap = t - bp;
  double aerr = s_hi - ap;
  double berr = p - bp;
  s_lo = s_lo + ((e + cross) + (aerr + berr));
  s_hi = t;
}

// Full TwoSum renormalisation of an accumulator pair (also the finaliser).
MF_HD void renorm(double &s_hi, double &s_lo) {
  double t = s_hi + s_lo;
  double bp = t - s_hi;
  double ap = t - bp;
  double aerr = s_hi - ap;
  double berr = s_lo - bp;
  s_lo = aerr + berr;
  s_hi = t;
}
