// DD transcendentals for the elementwise kernels.
//
// exp, exp2, exp10, expm1, exp2m1, exp10m1, log, log2, log10, log1p,
// log2p1, log10p1, sin, cos, tan, sinpi, cospi, tanpi, sinh, cosh, tanh,
// asinh, acosh and atanh keep the reference's algorithms step for step (the same range
// reductions, table sizes, polynomial degrees, evaluation trees and special
// cases), so their results are bit-identical to the reference CPU functions:
//   dd_exp   src/float64x2/exp_log.inc:237-277 with exp2_from_reduced 26-72
//   dd_exp2  exp_log.inc:109-128;  dd_expm1 407-427
//   dd_exp10 exp_log.inc:364-397 with exp2_from_triple 86-106
//   dd_exp2m1 / dd_exp10m1 / dd_log2p1 / dd_log10p1  exp_log.inc:479-519
//   dd_log   exp_log.inc:329-341, log2 311-327, log2_kernel 280-308
//   dd_log10 exp_log.inc:343-353;  dd_log1p 429-469
//   dd_sin / dd_cos / dd_tan  src/float64x2/trig.inc:241-289,
//           reduce_pi_half 114-147, sin_eval/cos_eval/sincos_eval 165-215
//   dd_sinpi / dd_cospi / dd_tanpi  trig.inc:41-104
//   dd_pow   exp_log.inc:531-600 with powi (float64x2.h:2077-2102);
//   dd_hypot, dd_fdim, dd_fmin, dd_fmax, dd_copysign  float64x2.h:1299-1309,
//           1518-1520, 2130-2136, 2270-2303
//   dd_fmod, dd_modulo  float64x2.h:2010-2072 (a bounded form of the
//           reference's reduction loop: NaN where that loop never returns);
//           dd_nearest  fsrc/multifloats.fypp:1595-1607
//   dd_sinh / dd_cosh / dd_tanh  src/float64x2/hyp.inc:2-87
//   dd_asinh / dd_acosh / dd_atanh  src/float64x2/inv_trig.inc:226-291
//   estrin10/12/14/17/24  src/float64x2/poly.inc:33-217 (neval's trees)
// The constant tables are closed forms produced by gen/gen_tables.py.
//
// dd_atan (with atan2, atanpi, atan2pi around it), dd_erf, dd_erfc and
// dd_erfcx are this project's own design, not the
// reference's rational fits (src/float64x2/special.inc:18-188,
// inv_trig.inc:21-55): piecewise
// Taylor evaluations whose coefficient row is picked per lane. Their accuracy
// target is the reference's own error on the same inputs.
#pragma once
#include "dd_core.h"
#include "dd_tables.h"

MF_HD ddv cdd(const double *hi, const double *lo, int i) { return mk(hi[i], lo[i]); }

// c0 + c1 x + ... + c10 x^10, Estrin tree of the reference (degree-10 case).
MF_HD ddv estrin10(ddv x, const double *hi, const double *lo) {
  ddv x2 = dd_mul(x, x);
  ddv x4 = dd_mul(x2, x2);
  ddv x8 = dd_mul(x4, x4);
  ddv p01 = dd_add(cdd(hi, lo, 0), dd_mul(cdd(hi, lo, 1), x));
  ddv p23 = dd_add(cdd(hi, lo, 2), dd_mul(cdd(hi, lo, 3), x));
  ddv p45 = dd_add(cdd(hi, lo, 4), dd_mul(cdd(hi, lo, 5), x));
  ddv p67 = dd_add(cdd(hi, lo, 6), dd_mul(cdd(hi, lo, 7), x));
  ddv p89 = dd_add(cdd(hi, lo, 8), dd_mul(cdd(hi, lo, 9), x));
  ddv p03 = dd_add(p01, dd_mul(p23, x2));
  ddv p47 = dd_add(p45, dd_mul(p67, x2));
  ddv p810 = dd_add(p89, dd_mul(cdd(hi, lo, 10), x2));
  return dd_add(dd_add(p03, dd_mul(p47, x4)), dd_mul(p810, x8));
}

// Degree-12 case of the same tree.
MF_HD ddv estrin12(ddv x, const double *hi, const double *lo) {
  ddv x2 = dd_mul(x, x);
  ddv x4 = dd_mul(x2, x2);
  ddv x8 = dd_mul(x4, x4);
  ddv p01 = dd_add(cdd(hi, lo, 0), dd_mul(cdd(hi, lo, 1), x));
  ddv p23 = dd_add(cdd(hi, lo, 2), dd_mul(cdd(hi, lo, 3), x));
  ddv p45 = dd_add(cdd(hi, lo, 4), dd_mul(cdd(hi, lo, 5), x));
  ddv p67 = dd_add(cdd(hi, lo, 6), dd_mul(cdd(hi, lo, 7), x));
  ddv p89 = dd_add(cdd(hi, lo, 8), dd_mul(cdd(hi, lo, 9), x));
  ddv p1011 = dd_add(cdd(hi, lo, 10), dd_mul(cdd(hi, lo, 11), x));
  ddv p03 = dd_add(p01, dd_mul(p23, x2));
  ddv p47 = dd_add(p45, dd_mul(p67, x2));
  ddv p811 = dd_add(p89, dd_mul(p1011, x2));
  ddv top = dd_add(p811, dd_mul(cdd(hi, lo, 12), x4));
  return dd_add(dd_add(p03, dd_mul(p47, x4)), dd_mul(top, x8));
}

// Degree-14 case of the same tree (asinh, atanh).
MF_HD ddv estrin14(ddv x, const double *hi, const double *lo) {
  ddv x2 = dd_mul(x, x);
  ddv x4 = dd_mul(x2, x2);
  ddv x8 = dd_mul(x4, x4);
  ddv p01 = dd_add(cdd(hi, lo, 0), dd_mul(cdd(hi, lo, 1), x));
  ddv p23 = dd_add(cdd(hi, lo, 2), dd_mul(cdd(hi, lo, 3), x));
  ddv p45 = dd_add(cdd(hi, lo, 4), dd_mul(cdd(hi, lo, 5), x));
  ddv p67 = dd_add(cdd(hi, lo, 6), dd_mul(cdd(hi, lo, 7), x));
  ddv p89 = dd_add(cdd(hi, lo, 8), dd_mul(cdd(hi, lo, 9), x));
  ddv p1011 = dd_add(cdd(hi, lo, 10), dd_mul(cdd(hi, lo, 11), x));
  ddv p1213 = dd_add(cdd(hi, lo, 12), dd_mul(cdd(hi, lo, 13), x));
  ddv p03 = dd_add(p01, dd_mul(p23, x2));
  ddv p47 = dd_add(p45, dd_mul(p67, x2));
  ddv p811 = dd_add(p89, dd_mul(p1011, x2));
  ddv p1214 = dd_add(p1213, dd_mul(cdd(hi, lo, 14), x2));
  return dd_add(dd_add(p03, dd_mul(p47, x4)), dd_mul(dd_add(p811, dd_mul(p1214, x4)), x8));
}

// c0 + ... + c15 x^15 by the degree-15 tree (shared by the 17 and 24 cases).
MF_HD ddv estrin_0_15(ddv x, ddv x2, ddv x4, ddv x8, const double *hi, const double *lo) {
  ddv p01 = dd_add(cdd(hi, lo, 0), dd_mul(cdd(hi, lo, 1), x));
  ddv p23 = dd_add(cdd(hi, lo, 2), dd_mul(cdd(hi, lo, 3), x));
  ddv p45 = dd_add(cdd(hi, lo, 4), dd_mul(cdd(hi, lo, 5), x));
  ddv p67 = dd_add(cdd(hi, lo, 6), dd_mul(cdd(hi, lo, 7), x));
  ddv p89 = dd_add(cdd(hi, lo, 8), dd_mul(cdd(hi, lo, 9), x));
  ddv p1011 = dd_add(cdd(hi, lo, 10), dd_mul(cdd(hi, lo, 11), x));
  ddv p1213 = dd_add(cdd(hi, lo, 12), dd_mul(cdd(hi, lo, 13), x));
  ddv p1415 = dd_add(cdd(hi, lo, 14), dd_mul(cdd(hi, lo, 15), x));
  ddv p03 = dd_add(p01, dd_mul(p23, x2));
  ddv p47 = dd_add(p45, dd_mul(p67, x2));
  ddv p811 = dd_add(p89, dd_mul(p1011, x2));
  ddv p1215 = dd_add(p1213, dd_mul(p1415, x2));
  return dd_add(dd_add(p03, dd_mul(p47, x4)), dd_mul(dd_add(p811, dd_mul(p1215, x4)), x8));
}

// Degree 17: the degree-15 tree plus (c16 + c17 x) x^16 (poly.inc:174-191).
MF_HD ddv estrin17(ddv x, const double *hi, const double *lo) {
  ddv x2 = dd_mul(x, x);
  ddv x4 = dd_mul(x2, x2);
  ddv x8 = dd_mul(x4, x4);
  ddv x16 = dd_mul(x8, x8);
  ddv p = estrin_0_15(x, x2, x4, x8, hi, lo);
  ddv p1617 = dd_add(cdd(hi, lo, 16), dd_mul(cdd(hi, lo, 17), x));
  return dd_add(p, dd_mul(p1617, x16));
}

// Degree 24: the degree-15 tree plus a degree-8 tree times x^16 (poly.inc:192-217).
MF_HD ddv estrin24(ddv x, const double *hi, const double *lo) {
  ddv x2 = dd_mul(x, x);
  ddv x4 = dd_mul(x2, x2);
  ddv x8 = dd_mul(x4, x4);
  ddv x16 = dd_mul(x8, x8);
  ddv p = estrin_0_15(x, x2, x4, x8, hi, lo);
  ddv p1617 = dd_add(cdd(hi, lo, 16), dd_mul(cdd(hi, lo, 17), x));
  ddv p1819 = dd_add(cdd(hi, lo, 18), dd_mul(cdd(hi, lo, 19), x));
  ddv p2021 = dd_add(cdd(hi, lo, 20), dd_mul(cdd(hi, lo, 21), x));
  ddv p2223 = dd_add(cdd(hi, lo, 22), dd_mul(cdd(hi, lo, 23), x));
  ddv p1619 = dd_add(p1617, dd_mul(p1819, x2));
  ddv p2023 = dd_add(p2021, dd_mul(p2223, x2));
  ddv q = dd_add(dd_add(p1619, dd_mul(p2023, x4)), dd_mul(cdd(hi, lo, 24), x8));
  return dd_add(p, dd_mul(q, x16));
}

// ---------------------------------------------------------------- exp
MF_HD ddv exp2_reduced(double n_d, double rh, double rl) {
  if (n_d > 1100.0) return mk(__builtin_inf(), 0.0);
  if (n_d < -1100.0) return mk(0.0, 0.0);
  double k_d = __builtin_rint(rh * 256.0);
  int k = (int)k_d;
  k = k > 128 ? 128 : (k < -128 ? -128 : k);
  double sh = rh - (double)k * 0.00390625;
  double sl = rl;
  fast_two_sum(sh, sl, sh, sl);
  ddv s = mk(sh, sl);
  ddv q = cdd(mf_exp2c_hi, mf_exp2c_lo, 9);
  for (int j = 8; j >= 1; --j) {
    q = dd_mul(q, s);
    q = dd_add(q, cdd(mf_exp2c_hi, mf_exp2c_lo, j));
  }
  ddv pow2s = dd_add(mk(1.0, 0.0), dd_mul(s, q));
  ddv mant = dd_mul(cdd(mf_exp2t_hi, mf_exp2t_lo, k + 128), pow2s);
  int n = (int)n_d;
  return mk(__builtin_ldexp(mant.hi, n), __builtin_ldexp(mant.lo, n));
}

MF_HD ddv dd_exp(ddv x) {
  if (!mf_isfinite(x.hi)) {
    if (x.hi > 0.0) return mk(x.hi, 0.0);
    if (x.hi < 0.0) return mk(0.0, 0.0);
    return mk(x.hi, 0.0);
  }
  if (x.hi > mf_exp_max) return mk(__builtin_inf(), 0.0);
  if (x.hi < mf_exp_min) return mk(0.0, 0.0);
  double n_d = __builtin_rint(x.hi * mf_log2e_hi);
  double rh, rl, t;
  two_sum(x.hi, -(n_d * mf_ln2_cw1), rh, t);
  rl = t + x.lo;
  fast_two_sum(rh, rl, rh, rl);
  double p_hi = n_d * mf_ln2_cw2;
  double p_lo = __builtin_fma(n_d, mf_ln2_cw2, -p_hi);
  two_sum(rh, -p_hi, rh, t);
  rl = rl + (t - p_lo);
  fast_two_sum(rh, rl, rh, rl);
  two_sum(rh, -(n_d * mf_ln2_cw3), rh, t);
  rl = rl + t;
  ddv rr;
  fast_two_sum(rh, rl, rr.hi, rr.lo);
  ddv y = dd_mul(rr, mk(mf_log2e_hi, mf_log2e_lo));
  return exp2_reduced(n_d, y.hi, y.lo);
}

// exp_log.inc:119-128 (exp2) with exp2_kernel 109-116
MF_HD ddv dd_exp2(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(x.hi > 0.0 ? x.hi : (x.hi < 0.0 ? 0.0 : x.hi), 0.0);
  if (x.hi < mf_exp2_min) return mk(0.0, 0.0);
  if (x.hi > mf_exp2_max) return mk(__builtin_inf(), 0.0);
  double n_d = __builtin_rint(x.hi);
  double rh, rl, t;
  two_sum(x.hi, -n_d, rh, t);
  rl = t + x.lo;
  fast_two_sum(rh, rl, rh, rl);
  return exp2_reduced(n_d, rh, rl);
}

// exp_log.inc:407-427: exp(x) - 1 for |x| >= 1/2, else x * P(x), P of degree 24
MF_HD ddv dd_expm1(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi)) return mk(x.hi > 0.0 ? x.hi : (x.hi < 0.0 ? -1.0 : x.hi), 0.0);
  if (__builtin_fabs(x.hi) >= 0.5) return dd_sub(dd_exp(x), mk(1.0, 0.0));
  return dd_mul(x, estrin24(x, mf_expm1c_hi, mf_expm1c_lo));
}

// ---------------------------------------------------------------- log
MF_HD ddv log2_kernel(ddv x) {
  double hi = x.hi;
  if (hi > 15.0 / 16.0 && hi < 17.0 / 16.0) {
    ddv t = dd_div(dd_sub(x, mk(1.0, 0.0)), dd_add(x, mk(1.0, 0.0)));
    ddv p = estrin10(dd_mul(t, t), mf_log2c_hi, mf_log2c_lo);
    return dd_mul(t, p);
  }
  int fe;
  (void)__builtin_frexp(hi, &fe);
  int e = fe - 1;  // ilogb for finite nonzero input, subnormals included
  ddv m = mk(__builtin_ldexp(x.hi, -e), __builtin_ldexp(x.lo, -e));
  int idx = (int)((m.hi - 1.0) * 32.0);
  idx = idx < 0 ? 0 : (idx > 31 ? 31 : idx);
  ddv c = mk(mf_log2_centre[idx], 0.0);
  ddv v = cdd(mf_log2v_hi, mf_log2v_lo, idx);
  ddv t = dd_div(dd_sub(m, c), dd_add(m, c));
  ddv ts = dd_mul(t, t);
  ddv p = cdd(mf_log2c_hi, mf_log2c_lo, 6);
  for (int i = 5; i >= 0; --i) p = dd_add(dd_mul(p, ts), cdd(mf_log2c_hi, mf_log2c_lo, i));
  return dd_add(dd_add(mk((double)e, 0.0), v), dd_mul(t, p));
}

MF_HD ddv dd_log(ddv x) {
  if (!mf_isfinite(x.hi)) {
    // log(+inf) = +inf; log(-inf) and log(NaN) are NaN.
    return mk(x.hi > 0.0 ? x.hi : __builtin_nan(""), 0.0);
  }
  ddv l2;
  if (x.hi < 0.0) return mk(__builtin_nan(""), 0.0);
  if (x.hi == 0.0) return mk(-__builtin_inf(), 0.0);
  l2 = log2_kernel(x);
  if (!mf_isfinite(l2.hi)) return l2;
  return dd_mul(l2, mk(mf_ln2_hi, mf_ln2_lo));
}

// exp_log.inc:311-327
MF_HD ddv dd_log2(ddv x) {
  if (x.hi != x.hi || x.hi < 0.0) return mk(__builtin_nan(""), 0.0);
  if (x.hi == 0.0) return mk(-__builtin_inf(), 0.0);
  if (!mf_isfinite(x.hi)) return mk(__builtin_inf(), 0.0);
  return log2_kernel(x);
}

// exp_log.inc:343-353: log2(x) * log10(2)
MF_HD ddv dd_log10(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(x.hi > 0.0 ? x.hi : __builtin_nan(""), 0.0);
  ddv l2 = dd_log2(x);
  if (!mf_isfinite(l2.hi)) return l2;
  return dd_mul(l2, mk(mf_log10_2_hi, mf_log10_2_lo));
}

// exp_log.inc:429-469: log(1 + x) for |x| >= 1/4, else 2 atanh(x / (2 + x))
// as t * P(t^2), P of degree 17
MF_HD ddv dd_log1p(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi)) return mk(x.hi > 0.0 ? x.hi : __builtin_nan(""), 0.0);
  ddv arg = dd_add(mk(1.0, 0.0), x);
  if (arg.hi <= 0.0) {
    if (arg.hi == 0.0 && arg.lo == 0.0) return mk(-__builtin_inf(), 0.0);
    if (arg.hi < 0.0 || (arg.hi == 0.0 && arg.lo < 0.0)) return mk(__builtin_nan(""), 0.0);
    return dd_log(mk(arg.lo, 0.0));  // hi == 0 < lo: the value is the lo limb alone
  }
  if (__builtin_fabs(x.hi) >= 0.25) return dd_log(arg);
  ddv t = dd_div(x, dd_add(mk(2.0, 0.0), x));
  return dd_mul(t, estrin17(dd_mul(t, t), mf_log1pc_hi, mf_log1pc_lo));
}

// ---------------------------------------------------------------- sinh, cosh, tanh
// src/float64x2/hyp.inc:2-18: x * P(x^2) for |x| < 1/2, else (e^x - e^-x) / 2
MF_HD ddv dd_sinh(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi)) return mk(x.hi, 0.0);  // sinh(+-inf) = +-inf, NaN stays
  if (__builtin_fabs(x.hi) < 0.5) return dd_mul(x, estrin12(dd_mul(x, x), mf_sinhc_hi, mf_sinhc_lo));
  return dd_mul(dd_sub(dd_exp(x), dd_exp(dd_neg(x))), mk(0.5, 0.0));
}

// hyp.inc:20-29: (e^x + e^-x) / 2
MF_HD ddv dd_cosh(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(x.hi == x.hi ? __builtin_inf() : x.hi, 0.0);
  return dd_mul(dd_add(dd_exp(x), dd_exp(dd_neg(x))), mk(0.5, 0.0));
}

// hyp.inc:62-87: +-1 beyond |x| >= 38, sinh/cosh below 1/2, else
// (1 - e^-2|x|) / (1 + e^-2|x|) with the sign restored
MF_HD ddv dd_tanh(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi)) return mk(x.hi == x.hi ? (x.hi > 0.0 ? 1.0 : -1.0) : x.hi, 0.0);
  if (x.hi >= 38.0) return mk(1.0, 0.0);
  if (x.hi <= -38.0) return mk(-1.0, 0.0);
  if (__builtin_fabs(x.hi) < 0.5) return dd_div(dd_sinh(x), dd_cosh(x));
  const bool sign = x.hi < 0.0;
  ddv m2 = sign ? mk(2.0 * x.hi, 2.0 * x.lo) : mk(-2.0 * x.hi, -2.0 * x.lo);
  ddv em2 = dd_exp(m2);
  ddv r = dd_div(dd_sub(mk(1.0, 0.0), em2), dd_add(mk(1.0, 0.0), em2));
  return sign ? dd_neg(r) : r;
}

// ---------------------------------------------------------------- asinh, acosh, atanh
// src/float64x2/inv_trig.inc:226-254: x P(x^2) below |x| = 0.01; log(2|x|)
// beyond 1e150; else log1p(|x| + x^2 / (sqrt(x^2 + 1) + 1)) below 1 and
// log(|x| + sqrt(x^2 + 1)) from 1 on, with the sign restored.
MF_HD ddv dd_asinh(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi)) return mk(x.hi, 0.0);  // asinh(+-inf) = +-inf, NaN stays
  if (__builtin_fabs(x.hi) < 0.01) return dd_mul(x, estrin14(dd_mul(x, x), mf_asinhc_hi, mf_asinhc_lo));
  const bool sign = x.hi < 0.0;
  const ddv ax = sign ? dd_neg(x) : x;
  if (ax.hi > 1e150) {
    const ddv r = dd_add(dd_log(ax), mk(mf_ln2_hi, mf_ln2_lo));
    return sign ? dd_neg(r) : r;
  }
  const ddv axsq = dd_mul(ax, ax);
  const ddv root = dd_sqrt(dd_add(axsq, mk(1.0, 0.0)));
  ddv res;
  if (ax.hi < 1.0) res = dd_log1p(dd_add(ax, dd_div(axsq, dd_add(root, mk(1.0, 0.0)))));
  else res = dd_log(dd_add(ax, root));
  return sign ? dd_neg(res) : res;
}

// inv_trig.inc:256-268: NaN below 1, log(2x) beyond 1e150, else
// log(x + sqrt(x^2 - 1)).
MF_HD ddv dd_acosh(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(x.hi > 0.0 ? x.hi : __builtin_nan(""), 0.0);
  if (x.hi < 1.0) return mk(__builtin_nan(""), 0.0);
  if (x.hi > 1e150) return dd_add(dd_log(x), mk(mf_ln2_hi, mf_ln2_lo));
  return dd_log(dd_add(x, dd_sqrt(dd_sub(dd_mul(x, x), mk(1.0, 0.0)))));
}

// inv_trig.inc:270-291: x P(x^2) below |x| = 0.01, NaN beyond 1, else
// log1p(2x / (1 - x)) / 2.
MF_HD ddv dd_atanh(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi)) return mk(__builtin_nan(""), 0.0);  // atanh(+-inf) and NaN: NaN
  if (__builtin_fabs(x.hi) > 1.0) return mk(__builtin_nan(""), 0.0);
  if (__builtin_fabs(x.hi) < 0.01) return dd_mul(x, estrin14(dd_mul(x, x), mf_atanhc_hi, mf_atanhc_lo));
  const ddv arg = dd_div(dd_add(x, x), dd_sub(mk(1.0, 0.0), x));
  return dd_mul(mk(0.5, 0.0), dd_log1p(arg));
}

// ------------------------------------------------- exp10 and the C23 forms
// Three-limb running sum: T0 + T1 + T2 += x with two error-free additions
// (src/multifloats_td.hh:30-43, tsum3 and tsum3_finalize).
MF_HD void tsum3(double &T0, double &T1, double &T2, double x) {
  double e;
  two_sum(T0, x, T0, e);
  two_sum(T1, e, T1, e);
  T2 += e;
}

// 2^(h + l + e) for a triple with (h, l) normalised: split off the nearest
// integer, carry the remainder as three limbs, evaluate 2^(T0 + T1) and
// correct to first order for T2 (exp_log.inc:86-106).
MF_HD ddv exp2_triple(double h, double l, double e) {
  const double n_d = __builtin_rint(h);
  double rh, rt;
  two_sum(h, -n_d, rh, rt);
  double T0 = rh, T1 = 0.0, T2 = 0.0;
  tsum3(T0, T1, T2, rt);
  tsum3(T0, T1, T2, l);
  tsum3(T0, T1, T2, e);
  double u;
  two_sum(T1, T2, T1, u);
  T2 = u;
  two_sum(T0, T1, T0, u);
  T1 = u;
  const ddv res = exp2_reduced(n_d, T0, T1);
  if (!mf_isfinite(res.hi)) return res;
  const double corr = res.hi * (mf_ln2_hi * T2);
  return dd_add(res, mk(corr, 0.0));
}

// exp_log.inc:364-397: x log2(10) formed as a triple from the four leading
// exact products of x with the three-limb log2(10), then exp2_triple.
MF_HD ddv dd_exp10(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(x.hi > 0.0 ? x.hi : (x.hi < 0.0 ? 0.0 : x.hi), 0.0);
  constexpr double hi_lim = mf_exp2_max * mf_log10_2_hi, lo_lim = mf_exp2_min * mf_log10_2_hi;
  if (x.hi > hi_lim) return mk(__builtin_inf(), 0.0);
  if (x.hi < lo_lim) return mk(0.0, 0.0);
  double P1, P1e, P2, P2e, P3, P3e;
  two_prod(x.hi, mf_log2_10_hi, P1, P1e);
  two_prod(x.hi, mf_log2_10_lo, P2, P2e);
  two_prod(x.lo, mf_log2_10_hi, P3, P3e);
  const double PX = x.hi * mf_log2_10_lo2 + x.lo * mf_log2_10_lo;
  double h = P1, l = 0.0, c = 0.0, t, u;
  two_sum(h, P3, h, t);
  two_sum(l, t, l, u);
  c += u;
  const double rest[5] = {P1e, P2, P2e, P3e, PX};
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    two_sum(l, rest[i], l, u);
    c += u;
  }
  double e;
  two_sum(l, c, l, e);
  fast_two_sum(h, l, h, l);
  return exp2_triple(h, l, e);
}

// exp_log.inc:479-506: 2^x - 1 and 10^x - 1 by the direct form for
// |x| >= 1/2, else expm1 of x ln2 (x ln10); non-finite x through expm1.
MF_HD ddv dd_exp2m1(ddv x) {
  if (mf_isfinite(x.hi) && __builtin_fabs(x.hi) >= 0.5) return dd_sub(dd_exp2(x), mk(1.0, 0.0));
  return dd_expm1(dd_mul(x, mk(mf_ln2_hi, mf_ln2_lo)));
}
MF_HD ddv dd_exp10m1(ddv x) {
  if (mf_isfinite(x.hi) && __builtin_fabs(x.hi) >= 0.5) return dd_sub(dd_exp10(x), mk(1.0, 0.0));
  return dd_expm1(dd_mul(x, mk(mf_ln10_hi, mf_ln10_lo)));
}
// exp_log.inc:508-519: log1p(x) times log2(e) or 1/ln(10).
MF_HD ddv dd_log2p1(ddv x) { return dd_mul(dd_log1p(x), mk(mf_log2e_hi, mf_log2e_lo)); }
MF_HD ddv dd_log10p1(ddv x) { return dd_mul(dd_log1p(x), mk(mf_inv_ln10_hi, mf_inv_ln10_lo)); }

// ------------------------------------------------- two-argument functions
// Lexicographic DD order of the reference (include/multifloats/float64x2.h:
// 167-190): NaN unordered, +0 equal to -0.
MF_HD bool dd_lt(ddv a, ddv b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
// Sign of the first nonzero limb; of hi when both are zero (float64x2.h:1449-1455).
MF_HD bool dd_signbit(ddv x) {
  return __builtin_signbit(x.hi != 0.0 ? x.hi : (x.lo != 0.0 ? x.lo : x.hi));
}
// Unbiased exponent of a finite nonzero double.
MF_HD int mf_ilogb(double v) {
  int e;
  (void)__builtin_frexp(v, &e);
  return e - 1;
}
// float64x2.h:1299-1309: the non-NaN operand when exactly one is NaN
MF_HD ddv dd_fmin(ddv a, ddv b) { return (dd_lt(a, b) || b.hi != b.hi) ? a : b; }
MF_HD ddv dd_fmax(ddv a, ddv b) { return (dd_lt(a, b) || a.hi != a.hi) ? b : a; }
// The Fortran generics min, max and dim of the reference's module
// (fsrc/multifloats.fypp:1539-1552, 1573-1583): the order is its `<` (the same
// lexicographic rule as dd_lt), so a NaN in y gives y for min and x for max;
// dim gives x - y when either is NaN (C fdim gives (NaN, 0)).
MF_HD ddv dd_fortran_min(ddv x, ddv y) { return dd_lt(x, y) ? x : y; }
MF_HD ddv dd_fortran_max(ddv x, ddv y) { return dd_lt(x, y) ? y : x; }
MF_HD ddv dd_fortran_dim(ddv x, ddv y) {
  if (x.hi != x.hi || y.hi != y.hi) return dd_sub(x, y);
  return dd_lt(y, x) ? dd_sub(x, y) : mk(0.0, 0.0);
}
// float64x2.h:1518-1520
MF_HD ddv dd_copysign(ddv x, ddv y) { return dd_signbit(x) == dd_signbit(y) ? x : dd_neg(x); }
// float64x2.h:2130-2136: max(x - y, 0), NaN if either is NaN
MF_HD ddv dd_fdim(ddv x, ddv y) {
  if (x.hi != x.hi || y.hi != y.hi) return mk(__builtin_nan(""), 0.0);
  return dd_lt(y, x) ? dd_sub(x, y) : mk(0.0, 0.0);
}
// ------------------------------------------------- fmod, modulo, nearest
// float64x2.h:1540-1555 (floor) and 1574-1576 (trunc); zeros kept
MF_HD ddv dd_floor(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  const double f = __builtin_floor(x.hi);
  if (f != x.hi) return mk(f, 0.0);
  double h, l;
  fast_two_sum(f, __builtin_floor(x.lo), h, l);
  return mk(h, l);
}
MF_HD ddv dd_trunc(ddv x) { return __builtin_signbit(x.hi) ? dd_neg(dd_floor(dd_neg(x))) : dd_floor(x); }
// float64x2.h:186-190 (operator>=): an equal hi limb decides on lo; NaN false
MF_HD bool dd_ge(ddv a, ddv b) { return b.hi < a.hi || (a.hi == b.hi && a.lo >= b.lo); }
// std::ilogb as the reference's x86-64 build has it (glibc): INT_MIN for
// zero and NaN, INT_MAX for infinities, the true exponent of a subnormal.
MF_HD int ref_ilogb(double v) {
  if (v == 0.0 || v != v) return -2147483647 - 1;
  if (__builtin_isinf(v)) return 2147483647;
  return mf_ilogb(v);
}
// One step of the reference's fmod loop (float64x2.h:2028-2054) on the
// remainder r, ay = |y|: false at the loop's exit (r unchanged, its
// magnitude the result), true after a step that leaves r nonzero (a zero r
// also ends the loop). The exponent gap is the 32-bit difference with
// wrap-around, as the compiled loop computes it: with ilogb(0) = INT_MIN a
// divisor of -0 gives q = -inf and r - ay, the same r again.
MF_HD bool dd_fmod_step(ddv &r, ddv ay) {
  if (r.hi < 0.0) {
    const double r_abs = -r.hi;
    const int gap = (int)((unsigned)ref_ilogb(r_abs) - (unsigned)ref_ilogb(ay.hi));
    if (gap <= 0) {
      r = dd_add(r, ay);
    } else if (gap <= 53) {
      const double q = __builtin_trunc(r_abs / ay.hi) + 1.0;
      r = dd_add(r, dd_mul(ay, mk(q, 0.0)));
    } else {
      r = dd_add(r, dd_mul(dd_add(dd_trunc(dd_div(dd_neg(r), ay)), mk(1.0, 0.0)), ay));
    }
  } else if (dd_ge(r, ay)) {
    const int gap = (int)((unsigned)ref_ilogb(r.hi) - (unsigned)ref_ilogb(ay.hi));
    if (gap <= 53) {
      const double q = __builtin_trunc(r.hi / ay.hi);
      r = q <= 1.0 ? dd_sub(r, ay) : dd_sub(r, dd_mul(ay, mk(q, 0.0)));
    } else {
      r = dd_sub(r, dd_mul(dd_trunc(dd_div(r, ay)), ay));
    }
  } else {
    return false;
  }
  return !(r.hi == 0.0 && r.lo == 0.0);
}
// float64x2.h:2010-2059 (fmod) with a bounded loop. The reference's loop is a
// function of the remainder alone (ay is fixed), so it returns exactly when
// the orbit of r reaches an exit, and never when r comes back to a value it
// held (y = -0, subnormal or tiny divisors whose products vanish, quotients
// beyond the double range). Brent's search (r saved after 1, 2, 4, ... steps
// since the last save) sees such a return within twice the length of the
// orbit; the step cap bounds the search. Both give NaN (C fmod's value for a
// zero divisor). Every other call takes the reference's steps and returns its
// bits; those orbits end within a few steps (10 on the test sets).
constexpr int kFmodMaxSteps = 512;
MF_HD ddv dd_fmod(ddv x, ddv y) {
  const bool x_neg = x.hi < 0.0;
  const ddv ax = x_neg ? dd_neg(x) : x;
  const ddv ay = y.hi < 0.0 ? dd_neg(y) : y;
  if (dd_lt(ax, ay)) return x;
  ddv r = ax, saved = ax;
  int power = 1, since = 0;
  for (int n = 1;; ++n) {
    if (!dd_fmod_step(r, ay)) break;
    if ((r.hi == saved.hi && r.lo == saved.lo && __builtin_signbit(r.hi) == __builtin_signbit(saved.hi) &&
         __builtin_signbit(r.lo) == __builtin_signbit(saved.lo)) ||
        n >= kFmodMaxSteps)
      return mk(__builtin_nan(""), 0.0);
    if (++since == power) {
      saved = r;
      power *= 2;
      since = 0;
    }
  }
  return x_neg ? dd_neg(r) : r;
}
// float64x2.h:2068-2072 (modulo): the floored remainder, sign of y
MF_HD ddv dd_modulo(ddv x, ddv y) {
  const ddv r = dd_fmod(x, y);
  if (r.hi == 0.0 && r.lo == 0.0) return r;
  return dd_signbit(r) != dd_signbit(y) ? dd_add(r, y) : r;
}
// fsrc/multifloats.fypp:1595-1607 (nearest): the next double after x.hi
// toward +inf when y.hi > 0, else toward -inf (gfortran's NEAREST is
// nextafter toward the infinity of the direction's sign); low limb 0.
MF_HD ddv dd_nearest(ddv x, ddv y) {
  const double h = x.hi;
  double out;
  if (h != h) {
    out = h;
  } else if (h == 0.0) {
    out = y.hi > 0.0 ? 0x1p-1074 : -0x1p-1074;
  } else {
    const bool up = y.hi > 0.0;
    const long long b = __builtin_bit_cast(long long, h);
    // away from zero when the step's direction has h's sign; +-inf stays
    // put in its own direction (nextafter(inf, inf) = inf)
    const bool away = (h > 0.0) == up;
    out = (away && __builtin_isinf(h)) ? h : __builtin_bit_cast(double, away ? b + 1 : b - 1);
  }
  return mk(out, 0.0);
}

// float64x2.h:2270-2303: both scaled by 2^-e (e the exponent of the larger),
// big sqrt(1 + (small/big)^2), scaled back. Non-finite input: +inf if either
// is infinite, else NaN (C hypot).
MF_HD ddv dd_hypot(ddv x, ddv y) {
  if (!mf_isfinite(x.hi) || !mf_isfinite(y.hi)) {
    const bool inf = __builtin_isinf(x.hi) || __builtin_isinf(y.hi);
    return mk(inf ? __builtin_inf() : __builtin_nan(""), 0.0);
  }
  const ddv ax = dd_signbit(x) ? dd_neg(x) : x;
  const ddv ay = dd_signbit(y) ? dd_neg(y) : y;
  const bool gt = dd_lt(ay, ax);
  ddv big = gt ? ax : ay, small = gt ? ay : ax;
  if (big.hi == 0.0) return mk(0.0, 0.0);
  const int e = mf_ilogb(big.hi);
  const double down = __builtin_ldexp(1.0, -e), up = __builtin_ldexp(1.0, e);
  big = mk(big.hi * down, big.lo * down);
  small = mk(small.hi * down, small.lo * down);
  const ddv ratio = dd_div(small, big);
  const ddv root = dd_mul(big, dd_sqrt(dd_add(mk(1.0, 0.0), dd_mul(ratio, ratio))));
  const double rh = root.hi * up;
  return mk(rh, mf_isfinite(rh) ? root.lo * up : 0.0);
}
// float64x2.h:2077-2102 (powi): binary powering; 1 / x^|n| for n < 0
MF_HD ddv dd_powi(ddv base, int n) {
  if (n == 0) return mk(1.0, 0.0);
  if (base.hi == 0.0 && base.lo == 0.0) {
    const double sg = (n & 1) ? __builtin_copysign(1.0, base.hi) : 1.0;
    return mk(sg * (n > 0 ? 0.0 : __builtin_inf()), 0.0);
  }
  const bool neg = n < 0;
  unsigned long long u = neg ? (unsigned long long)(-(long long)n) : (unsigned long long)n;
  ddv result = mk(1.0, 0.0), b = base;
  for (;;) {
    if (u & 1ull) result = dd_mul(result, b);
    u >>= 1;
    if (u == 0ull) break;
    b = dd_mul(b, b);
  }
  return neg ? dd_div(mk(1.0, 0.0), result) : result;
}

// exp_log.inc:531-600: integer y (|y| <= 2^31-1) by dd_powi; non-finite or
// non-positive x by exp(y log x); else 2^(y log2 x) with log2 x = e + log2 m
// (m in [1, 2)) and y (e + log2 m) accumulated as a triple from exact
// products, then exp2_triple.
MF_HD ddv dd_pow(ddv x, ddv y) {
  if (x.hi == 0.0 && y.hi == 0.0) return mk(1.0, 0.0);
  if (x.hi == 1.0 && x.lo == 0.0) return mk(1.0, 0.0);
  if (x.hi == -1.0 && x.lo == 0.0 && __builtin_isinf(y.hi)) return mk(1.0, 0.0);
  if (y.lo == 0.0 && mf_isfinite(y.hi) && y.hi == __builtin_trunc(y.hi) && __builtin_fabs(y.hi) <= 2147483647.0)
    return dd_powi(x, (int)y.hi);
  if (!mf_isfinite(x.hi) || !mf_isfinite(y.hi) || x.hi <= 0.0) return dd_exp(dd_mul(y, dd_log(x)));
  const int e_x = mf_ilogb(x.hi);
  const ddv m = mk(__builtin_ldexp(x.hi, -e_x), __builtin_ldexp(x.lo, -e_x));
  const ddv Lf = log2_kernel(m);
  double Lx = 0.0;
  if (!(m.hi > 15.0 / 16.0 && m.hi < 17.0 / 16.0)) {
    int idx = (int)((m.hi - 1.0) * 32.0);
    idx = idx < 0 ? 0 : (idx > 31 ? 31 : idx);
    Lx = mf_log2v_x[idx];
  }
  const double ed = (double)e_x;
  double AH, AM, ALh, ALl, P1, P1e, P2, P2e, P3, P3e;
  two_prod(y.hi, ed, AH, AM);
  two_prod(y.lo, ed, ALh, ALl);
  two_prod(y.hi, Lf.hi, P1, P1e);
  two_prod(y.hi, Lf.lo, P2, P2e);
  two_prod(y.lo, Lf.hi, P3, P3e);
  const double PX = y.hi * Lx + y.lo * Lf.lo;
  double h = AH, l = 0.0, c = 0.0, t, u;
  const double lead[4] = {P1, AM, ALh, P3};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    two_sum(h, lead[i], h, t);
    two_sum(l, t, l, u);
    c += u;
  }
  const double rest[5] = {P1e, P2, P2e, P3e, PX};
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    two_sum(l, rest[i], l, u);
    c += u;
  }
  c += ALl;
  double e;
  two_sum(l, c, l, e);
  fast_two_sum(h, l, h, l);
  return exp2_triple(h, l, e);
}
// ---------------------------------------------------------------- sin
MF_HD void reduce_pio2(ddv x, ddv &r, int &q) {
  double n = __builtin_rint(x.hi * 0.6366197723675814);
  double rh = x.hi, rl = x.lo;
  const double cw[3] = {mf_pio2_cw1, mf_pio2_cw2, mf_pio2_cw3};
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    double ph = n * cw[s];
    double pl = s < 2 ? __builtin_fma(n, cw[s], -ph) : 0.0;
    double sh, sl, tl, tll;
    two_sum(rh, -ph, sh, sl);
    two_sum(rl, -pl, tl, tll);
    sl = sl + tl;
    fast_two_sum(sh, sl, rh, rl);
    rl = rl + tll;
    fast_two_sum(rh, rl, rh, rl);
  }
  r = mk(rh, rl);
  long long nn = (long long)__builtin_fabs(n) & 3LL;
  if (n < 0.0) nn = (4 - nn) & 3;
  q = (int)nn;
}

MF_HD ddv sin_kernel(ddv x) { return dd_mul(estrin12(dd_mul(x, x), mf_sinc_hi, mf_sinc_lo), x); }
MF_HD ddv cos_kernel(ddv x) { return estrin12(dd_mul(x, x), mf_cosc_hi, mf_cosc_lo); }

// want_cos = 0: sin(r); 1: cos(r), for |r| <= pi/4.
// Every lane evaluates the sin and the cos Taylor kernel exactly once, on r
// itself when |r| <= pi/8 and on r -+ pi/4 otherwise, and selects afterwards:
// the same operations per lane as the reference's branches (so the same
// bits), but a wavefront whose lanes take different branches runs two
// polynomial evaluations instead of up to four.
MF_HD ddv sincos_eval(ddv r, int want_cos) {
  const double pi8 = 0.392699081698724;
  const bool small = __builtin_fabs(r.hi) <= pi8;
  const bool pos = r.hi > 0.0;
  ddv pi4 = mk(mf_pi_quarter_hi, mf_pi_quarter_lo);
  ddv rp = small ? r : (pos ? dd_sub(r, pi4) : dd_add(r, pi4));
  ddv sk = sin_kernel(rp);
  ddv ck = cos_kernel(rp);
  if (small) return want_cos ? ck : sk;
  ddv is2 = mk(mf_inv_sqrt2_hi, mf_inv_sqrt2_lo);
  ddv comb;
  if (!want_cos) comb = pos ? dd_add(sk, ck) : dd_sub(sk, ck);
  else comb = pos ? dd_sub(ck, sk) : dd_add(ck, sk);
  return dd_mul(comb, is2);
}

MF_HD ddv dd_sin(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi) || !(__builtin_fabs(x.hi) < 0x1p55)) return mk(__builtin_nan(""), 0.0);
  ddv r;
  int q;
  reduce_pio2(x, r, q);
  ddv v = sincos_eval(r, q & 1);
  return (q & 2) ? dd_neg(v) : v;
}

// trig.inc:255-266
MF_HD ddv dd_cos(ddv x) {
  if (!mf_isfinite(x.hi) || !(__builtin_fabs(x.hi) < 0x1p55)) return mk(__builtin_nan(""), 0.0);
  ddv r;
  int q;
  reduce_pio2(x, r, q);
  ddv v = sincos_eval(r, !(q & 1));
  return (q == 1 || q == 2) ? dd_neg(v) : v;
}

// sin(r) and cos(r) together for |r| <= pi/4 (trig.inc:192-215)
MF_HD void sincos_both(ddv r, ddv &s, ddv &c) {
  const double pi8 = 0.392699081698724;
  const bool small = __builtin_fabs(r.hi) <= pi8;
  const bool pos = r.hi > 0.0;
  ddv pi4 = mk(mf_pi_quarter_hi, mf_pi_quarter_lo);
  ddv rp = small ? r : (pos ? dd_sub(r, pi4) : dd_add(r, pi4));
  ddv sk = sin_kernel(rp);
  ddv ck = cos_kernel(rp);
  s = sk;
  c = ck;
  if (!small) {
    ddv is2 = mk(mf_inv_sqrt2_hi, mf_inv_sqrt2_lo);
    s = dd_mul(pos ? dd_add(sk, ck) : dd_sub(sk, ck), is2);
    c = dd_mul(pos ? dd_sub(ck, sk) : dd_add(ck, sk), is2);
  }
}

// trig.inc:268-289: s / c, or -c / s in odd quadrants
MF_HD ddv dd_tan(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi) || !(__builtin_fabs(x.hi) < 0x1p55)) return mk(__builtin_nan(""), 0.0);
  ddv r, s, c;
  int q;
  reduce_pio2(x, r, q);
  sincos_both(r, s, c);
  return (q & 1) ? dd_div(dd_neg(c), s) : dd_div(s, c);
}

// ---------------------------------------------------------- sinpi, cospi, tanpi
// trig.inc:41-104: n = RN(2|x|), r = |x| - n/2 (|r| <= 1/4), evaluated at
// pi r by the sin/cos kernels above and placed by n mod 4. NaN for
// non-finite x and for |x| >= 2^52, where 2|x| has no fractional part left.
MF_HD void reduce_half(ddv ax, ddv &pr, int &nmod) {
  const double n = __builtin_rint(2.0 * ax.hi);
  pr = dd_mul(mk(mf_pi_hi, mf_pi_lo), dd_add(ax, mk(-0.5 * n, 0.0)));
  nmod = (int)((long long)n & 3LL);
}
MF_HD ddv dd_sinpi(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(__builtin_nan(""), 0.0);
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  const bool sign = x.hi < 0.0;
  const ddv ax = sign ? dd_neg(x) : x;
  if (!(ax.hi < 0x1p52)) return mk(__builtin_nan(""), 0.0);
  ddv pr;
  int q;
  reduce_half(ax, pr, q);
  const ddv v = sincos_eval(pr, q & 1);
  const ddv res = (q & 2) ? dd_neg(v) : v;
  return sign ? dd_neg(res) : res;
}
MF_HD ddv dd_cospi(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(__builtin_nan(""), 0.0);
  const ddv ax = x.hi < 0.0 ? dd_neg(x) : x;
  if (!(ax.hi < 0x1p52)) return mk(__builtin_nan(""), 0.0);
  ddv pr;
  int q;
  reduce_half(ax, pr, q);
  const ddv v = sincos_eval(pr, !(q & 1));
  return (q == 1 || q == 2) ? dd_neg(v) : v;
}
MF_HD ddv dd_tanpi(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(__builtin_nan(""), 0.0);
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  const bool sign = x.hi < 0.0;
  const ddv ax = sign ? dd_neg(x) : x;
  if (!(ax.hi < 0x1p52)) return mk(__builtin_nan(""), 0.0);
  ddv pr, s, c;
  int q;
  reduce_half(ax, pr, q);
  sincos_both(pr, s, c);
  const ddv res = (q & 1) ? dd_div(dd_neg(c), s) : dd_div(s, c);
  return sign ? dd_neg(res) : res;
}

// ---------------------------------------------------------------- erf
// base (three limbs) + corr, rounded once to a double-double. The leading
// low term w is first folded into s exactly (fast_two_sum), so the one
// rounding happens at the scale of the final low limb: when |w| exceeds
// ulp(s)/2, rounding it before the fold would cost up to 2^-105 of the
// result instead of 2^-106.
MF_HD ddv sum3_round(double b0, double b1, double b2, ddv corr) {
  double s, e1, u1, v1, w, z, t;
  two_sum(b0, corr.hi, s, e1);
  two_sum(b1, corr.lo, u1, v1);
  two_sum(e1, u1, w, z);
  fast_two_sum(s, w, s, t);
  double lo = t + (z + (v1 + b2));
  ddv r;
  fast_two_sum(s, lo, r.hi, r.lo);
  return r;
}

// Where erf reads its row tables: the constant tables here, or a copy that
// an array kernel stages in LDS (elementwise.hip), with the same values, so
// both give the same bits.
struct ErfTabConst {
  MF_HDM double centre(int row) const { return mf_erf_centre[row]; }
  MF_HDM double base(int row, int i) const { return mf_erf_base[3 * row + i]; }
  MF_HDM double coef_hi(int row, int j) const { return mf_erf_ddhi[row * MF_ERF_ND + j]; }
  MF_HDM double coef_lo(int row, int j) const { return mf_erf_ddlo[row * MF_ERF_ND + j]; }
  MF_HDM double tail(int row, int j) const { return mf_erf_d[row * (MF_ERF_NT - MF_ERF_ND) + j]; }
};

// P(u) of erf's row (ND double-double coefficients, the tail in double)
template <class Tab>
MF_HD ddv erf_poly_t(const Tab &tab, int row, ddv u) {
  double d = tab.tail(row, MF_ERF_NT - MF_ERF_ND - 1);
  for (int j = MF_ERF_NT - MF_ERF_ND - 2; j >= 0; --j) d = __builtin_fma(d, u.hi, tab.tail(row, j));
  ddv p = dd_add(mk(tab.coef_hi(row, MF_ERF_ND - 1), tab.coef_lo(row, MF_ERF_ND - 1)), dd_mul(u, mk(d, 0.0)));
  for (int j = MF_ERF_ND - 2; j >= 0; --j) p = dd_add(mk(tab.coef_hi(row, j), tab.coef_lo(row, j)), dd_mul(u, p));
  return p;
}
MF_HD ddv erf_poly(int row, ddv u) { return erf_poly_t(ErfTabConst{}, row, u); }

// Row 0 (|x| < 1/4):   erf(x) = x + x * Q(x^2)
// Row i (1/4 <= |x| < 9): erf(c_i + t) = a0 + t * P(t), |t| <= 1/32
// P and Q share one evaluator: the first MF_ERF_ND coefficients are
// double-doubles, the tail (contributing < 2^-56 of the result) is plain
// double Horner in u.hi. a0 is a three-limb constant and the final sum
// base + mult*P is rounded once, so the result error stays near one half-ulp
// of a double-double.
template <class Tab>
MF_HD ddv dd_erf_t(const Tab &tab, ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi)) {
    if (x.hi != x.hi) return mk(x.hi, 0.0);
    return mk(x.hi > 0.0 ? 1.0 : -1.0, 0.0);
  }
  bool neg = x.hi < 0.0;
  ddv ax = neg ? dd_neg(x) : x;
  if (ax.hi >= mf_erf_xmax) return mk(neg ? -1.0 : 1.0, 0.0);
  int row = 0;
  if (ax.hi >= mf_erf_x0) {
    row = 1 + (int)((ax.hi - mf_erf_x0) * mf_erf_inv_h);
    row = row > MF_ERF_ROWS - 1 ? MF_ERF_ROWS - 1 : row;
  }
  ddv u, mult;
  double b0, b1, b2;
  if (row == 0) {
    u = dd_mul(ax, ax);
    mult = ax;
    b0 = ax.hi;
    b1 = ax.lo;
    b2 = 0.0;
  } else {
    double th = ax.hi - tab.centre(row);  // exact (Sterbenz)
    double tl = ax.lo;
    fast_two_sum(th, tl, th, tl);
    u = mk(th, tl);
    mult = u;
    b0 = tab.base(row, 0);
    b1 = tab.base(row, 1);
    b2 = tab.base(row, 2);
  }
  ddv r = sum3_round(b0, b1, b2, dd_mul(mult, erf_poly_t(tab, row, u)));
  return neg ? dd_neg(r) : r;
}
MF_HD ddv dd_erf(ddv x) { return dd_erf_t(ErfTabConst{}, x); }

// ---------------------------------------------------------------- erfc, erfcx
// This project's own design (the reference uses libquadmath's rational fits,
// src/float64x2/special.inc:56-188), with the reference's regime split:
//   erfc(x), x >= 0:
//     x < 1/4       1 - erf(x) from erf's row 0, base 1 - x as three limbs;
//     1/4 <= x < 4  Taylor rows of erfc itself (all coefficients double-
//                   double, relative accuracy), base erfc(c) in three limbs;
//     4 <= x < 27.3 exp(-x^2) * erfcx(x), x^2 = s^2 + (x - s)(x + s) with s the
//                   leading 26 bits of x (s^2 exact), exp(-(x-s)(x+s)) by its
//                   series; beyond 27.3 erfc underflows to 0;
//   erfc(x), x < 0: 2 - erfc(-x) (1 + erf(-x) below 1/4);
//   erfcx(x), x >= -2: Taylor rows of erfcx (ODE-generated coefficients) up
//                   to 32, then the asymptotic series in 1/(2x^2);
//   erfcx(x), x < -2: 2 exp(x^2) - erfcx(-x), +inf past exp's overflow.
// Each row result is rounded once from a three-limb base plus a DD
// correction, as erf.

// 1/16-wide rows of erfcx on [0, 4), 32 rows per binade on [4, 32), and
// rows 160..191 for [-2, 0).
MF_HD int erfcx_row(double x) {
  if (x < 0.0) {
    int r = (int)(-x * 16.0);
    return 160 + (r > 31 ? 31 : r);
  }
  if (x < 4.0) {
    int r = (int)(x * 16.0);
    return r > 63 ? 63 : r;
  }
  int fe;
  const double m = __builtin_frexp(x, &fe);  // x = m 2^fe, m in [1/2, 1)
  int j = (int)((2.0 * m - 1.0) * 32.0);
  j = j > 31 ? 31 : j;
  return 64 + 32 * (fe - 3) + j;
}

// DD Horner over nt DD coefficients c1..c_nt: c1 + c2 t + ... (all DD)
MF_HD ddv horner_dd(ddv t, const double *hi, const double *lo, int nt) {
  ddv p = mk(hi[nt - 1], lo[nt - 1]);
  for (int j = nt - 2; j >= 0; --j) p = dd_add(mk(hi[j], lo[j]), dd_mul(t, p));
  return p;
}

// t = x - c as a double-double (x.hi - c taken exactly by TwoSum: near 0 the
// centre can exceed 2 x.hi)
MF_HD ddv row_offset(ddv x, double c) {
  double th, te;
  two_sum(x.hi, -c, th, te);
  double tl = te + x.lo;
  fast_two_sum(th, tl, th, tl);
  return mk(th, tl);
}

// exp(sg * x^2) for x >= 0, sg = +1 or -1: x^2 = s^2 + e with s the top 26
// significand bits of x.hi (s^2 exact) and e = (x - s)(x + s) in DD, so
// exp(sg x^2) = exp(sg s^2) * exp(sg e), the second by its degree-8 series.
MF_HD ddv exp_pm_sq(ddv x, double sg) {
  const double s = __builtin_bit_cast(double, __builtin_bit_cast(long long, x.hi) & ~((1LL << 27) - 1));
  double dh = x.hi - s, dl = x.lo;
  fast_two_sum(dh, dl, dh, dl);
  const ddv e = dd_mul(mk(dh, dl), dd_add(x, mk(s, 0.0)));
  const ddv y = mk(sg * e.hi, sg * e.lo);
  ddv q = cdd(mf_expm1c_hi, mf_expm1c_lo, 7);
  for (int k = 6; k >= 0; --k) q = dd_add(cdd(mf_expm1c_hi, mf_expm1c_lo, k), dd_mul(y, q));
  const ddv e1 = dd_exp(mk(sg * (s * s), 0.0));
  return dd_add(e1, dd_mul(e1, dd_mul(y, q)));  // e1 (1 + expm1(sg e)), one rounding at the end
}

// erfcx(x) for finite x >= -2
MF_HD ddv erfcx_pos(ddv x) {
  if (x.hi >= 32.0) {
    const ddv u = dd_div(mk(0.5, 0.0), dd_mul(x, x));  // 1/(2x^2)
    ddv sum = mk(1.0, 0.0);
    for (int k = MF_ERFCX_ASYM - 1; k >= 0; --k)
      sum = dd_sub(mk(1.0, 0.0), dd_mul(mk(2.0 * k + 1.0, 0.0), dd_mul(u, sum)));
    return dd_div(dd_mul(sum, mk(mf_inv_sqrtpi_hi, mf_inv_sqrtpi_lo)), x);
  }
  const int row = erfcx_row(x.hi);
  const ddv t = row_offset(x, mf_erfcx_centre[row]);
  const ddv p = horner_dd(t, mf_erfcx_hi + row * MF_ERFCX_NT, mf_erfcx_lo + row * MF_ERFCX_NT, MF_ERFCX_NT);
  const double *b = mf_erfcx_base + 3 * row;
  return sum3_round(b[0], b[1], b[2], dd_mul(t, p));
}

// erfc(x) for finite x >= 1/4
MF_HD ddv erfc_pos(ddv x) {
  if (x.hi >= 27.3) return mk(0.0, 0.0);
  if (x.hi >= 4.0) return dd_mul(exp_pm_sq(x, -1.0), erfcx_pos(x));
  int row = (int)((x.hi - mf_erf_x0) * mf_erf_inv_h);
  row = row > MF_ERFC_ROWS - 1 ? MF_ERFC_ROWS - 1 : row;
  const ddv t = row_offset(x, mf_erfc_centre[row]);
  const ddv p = horner_dd(t, mf_erfc_hi + row * MF_ERFC_NT, mf_erfc_lo + row * MF_ERFC_NT, MF_ERFC_NT);
  const double *b = mf_erfc_base + 3 * row;
  return sum3_round(b[0], b[1], b[2], dd_mul(t, p));
}

MF_HD ddv dd_erfc(ddv x) {
  if (!mf_isfinite(x.hi)) {
    if (x.hi != x.hi) return mk(x.hi, 0.0);
    return mk(x.hi > 0.0 ? 0.0 : 2.0, 0.0);
  }
  const bool neg = x.hi < 0.0;
  const ddv ax = neg ? dd_neg(x) : x;
  if (ax.hi < mf_erf_x0) {
    // 1 - erf(x) = (1 - x) - x Q(x^2): the base 1 - x kept exactly in three limbs
    const double sg = neg ? -1.0 : 1.0;
    double b0, e0, b1, b2;
    two_sum(1.0, -sg * ax.hi, b0, e0);
    two_sum(e0, -sg * ax.lo, b1, b2);
    const ddv c = dd_mul(ax, erf_poly(0, dd_mul(ax, ax)));
    return sum3_round(b0, b1, b2, mk(-sg * c.hi, -sg * c.lo));
  }
  const ddv r = erfc_pos(ax);
  return neg ? dd_sub(mk(2.0, 0.0), r) : r;
}

MF_HD ddv dd_erfcx(ddv x) {
  if (!mf_isfinite(x.hi)) {
    if (x.hi != x.hi) return mk(x.hi, 0.0);
    return mk(x.hi > 0.0 ? 0.0 : __builtin_inf(), 0.0);
  }
  if (x.hi >= -2.0) return erfcx_pos(x);
  const ddv ax = dd_neg(x);
  if (ax.hi > 26.7) return mk(__builtin_inf(), 0.0);  // exp(x^2) overflows
  const ddv ex2 = exp_pm_sq(ax, 1.0);
  if (!mf_isfinite(ex2.hi)) return mk(__builtin_inf(), 0.0);
  return dd_sub(dd_mul(mk(2.0, 0.0), ex2), erfcx_pos(ax));
}

// ------------------------------------------------------------------- atan
// Own design (the reference evaluates libquadmath's fitted rationals,
// src/float64x2/inv_trig.inc:21-55; its special cases and the atan2 / atanpi
// wrappers, 134-224, are kept). Every path ends in one rounding of a
// three-limb base plus a double-double correction that is at most 1/15 of
// the result:
//   |x| < 1/8:      x + x^3 Q(x^2)
//   1/8 <= |x| <= 8: atan(k/64) + atan(u), k = RN(64|x|),
//                   u = (|x| - k/64) / (1 + |x| k/64), |u| <= 1/128
//   |x| > 8:        pi/2 - atan(1/|x|)
// with Q the Maclaurin series of (atan(u) - u) / u^3 to 17 terms (|u| <= 1/8).
MF_HD ddv atan_tail(ddv u) {
  const ddv s = dd_mul(u, u);
  ddv q = cdd(mf_atanc_hi, mf_atanc_lo, 16);
  for (int j = 15; j >= 0; --j) q = dd_add(cdd(mf_atanc_hi, mf_atanc_lo, j), dd_mul(s, q));
  return dd_mul(dd_mul(s, u), q);
}
MF_HD ddv dd_atan(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(x.hi != x.hi ? x.hi : __builtin_copysign(mf_pio2_cw1, x.hi), 0.0);
  if (__builtin_fabs(x.hi) < 1.85e-16) return x;  // the cubic term is below half a DD ulp
  const bool neg = x.hi < 0.0;
  const ddv t = neg ? dd_neg(x) : x;
  ddv r;
  if (t.hi < 0.125) {
    r = sum3_round(t.hi, t.lo, 0.0, atan_tail(t));
  } else if (t.hi <= 8.0) {
    const int k = (int)__builtin_rint(t.hi * 64.0);
    const ddv c = mk((double)k * 0.015625, 0.0);
    const ddv u = dd_div(dd_sub(t, c), dd_add(mk(1.0, 0.0), dd_mul(t, c)));
    const double *b = mf_atant + 3 * (k - 8);
    r = sum3_round(b[0], b[1], b[2], dd_add(u, atan_tail(u)));
  } else {
    const ddv v = dd_div(mk(1.0, 0.0), t);
    r = sum3_round(mf_pio2_cw1, mf_pio2_cw2, mf_pio2_cw3, dd_neg(dd_add(v, atan_tail(v))));
  }
  return neg ? dd_neg(r) : r;
}
// atan2 for finite arguments, not both with a zero leading limb
// (inv_trig.inc:176-193): atan of the smaller ratio, moved by pi or pi/2.
MF_HD ddv atan2_main(ddv y, ddv x) {
  const ddv pi = mk(mf_pi_hi, mf_pi_lo), pio2 = mk(mf_pio2_cw1, mf_pio2_cw2);
  if (__builtin_fabs(x.hi) >= __builtin_fabs(y.hi)) {
    ddv res = dd_atan(dd_div(y, x));
    if (x.hi < 0.0) res = __builtin_signbit(y.hi) ? dd_sub(res, pi) : dd_add(res, pi);
    return res;
  }
  const ddv res = dd_atan(dd_div(x, y));
  return y.hi > 0.0 ? dd_sub(pio2, res) : dd_sub(dd_neg(pio2), res);
}
// inv_trig.inc:134-193: non-finite arguments give C atan2's double result
// (NaN; +-pi/4, +-3pi/4 or +-pi/2 for an infinite y; +-0 or +-pi for an
// infinite x); both leading limbs zero: the exact quadrant value for signed
// zeros, else atan2 of the low limbs rounded to a double.
MF_HD ddv dd_atan2(ddv y, ddv x) {
  if (!mf_isfinite(x.hi) || !mf_isfinite(y.hi)) {
    if (x.hi != x.hi || y.hi != y.hi) return mk(__builtin_nan(""), 0.0);
    const double pi = mf_pi_hi, sy = __builtin_copysign(1.0, y.hi);
    if (__builtin_isinf(y.hi)) {
      const double a = __builtin_isinf(x.hi) ? (x.hi > 0.0 ? 0x1.921fb54442d18p-1 : 0x1.2d97c7f3321d2p+1) : mf_pio2_cw1;
      return mk(sy * a, 0.0);
    }
    return mk(x.hi > 0.0 ? sy * 0.0 : sy * pi, 0.0);
  }
  if (x.hi == 0.0 && y.hi == 0.0) {
    if (y.lo != 0.0 || x.lo != 0.0) {
      const double yl = y.lo != 0.0 ? y.lo : __builtin_copysign(0.0, y.hi);
      const double xl = x.lo != 0.0 ? x.lo : __builtin_copysign(0.0, x.hi);
      if (yl == 0.0)  // C atan2(+-0, x): +-0 for x > 0, +-pi for x < 0
        return mk(__builtin_copysign(xl > 0.0 ? 0.0 : mf_pi_hi, yl), 0.0);
      if (xl == 0.0) return mk(__builtin_copysign(mf_pio2_cw1, yl), 0.0);
      return mk(atan2_main(mk(yl, 0.0), mk(xl, 0.0)).hi, 0.0);
    }
    if (!__builtin_signbit(x.hi)) return mk(__builtin_copysign(0.0, y.hi), 0.0);
    return __builtin_signbit(y.hi) ? mk(-mf_pi_hi, -mf_pi_lo) : mk(mf_pi_hi, mf_pi_lo);
  }
  return atan2_main(y, x);
}
// inv_trig.inc:207-223: atan / pi, with atanpi(+-inf) = +-1/2 exactly
MF_HD ddv dd_atanpi(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!mf_isfinite(x.hi)) return mk(x.hi != x.hi ? x.hi : (__builtin_signbit(x.hi) ? -0.5 : 0.5), 0.0);
  return dd_mul(dd_atan(x), mk(mf_inv_pi_hi, mf_inv_pi_lo));
}
MF_HD ddv dd_atan2pi(ddv y, ddv x) { return dd_mul(dd_atan2(y, x), mk(mf_inv_pi_hi, mf_inv_pi_lo)); }

// ------------------------------------------------------------- asin, acos
// Own design on the atan reduction above, with every quantity that can
// cancel carried in three limbs: s = sqrt(1 - x^2) from the exact expansion of
// x^2 and one Newton correction of the double-double root; then
// asin |x| = atan(|x| / s), acos |x| = atan(s / |x|), acos x = pi - acos |x|
// for x < 0 (the reference's special cases, inv_trig.inc:56-130, are kept).
struct td3 {
  double h, m, l;
};
// a - c b for three-limb a, b and a short double c, summed from exact products
MF_HD ddv td_sub_cmul(td3 a, double c, td3 b) {
  double p1, e1, p2, e2;
  two_prod(c, b.h, p1, e1);
  two_prod(c, b.m, p2, e2);
  double T0 = a.h, T1 = 0.0, T2 = 0.0;
  tsum3(T0, T1, T2, -p1);
  tsum3(T0, T1, T2, a.m);
  tsum3(T0, T1, T2, -p2);
  tsum3(T0, T1, T2, -e1);
  tsum3(T0, T1, T2, a.l - c * b.l - e2);
  double u;
  two_sum(T1, T2, T1, u);
  two_sum(T0, T1, T0, u);
  return mk(T0, u);
}
// atan(n / d) for n, d >= 0 given in three limbs, not both zero
MF_HD ddv atan_ratio(td3 n, td3 d) {
  if (d.h == 0.0) return mk(mf_pio2_cw1, mf_pio2_cw2);
  if (n.h == 0.0) return mk(0.0, 0.0);
  const ddv nd = mk(n.h, n.m), dd = mk(d.h, d.m);
  const ddv z = dd_div(nd, dd);
  if (z.hi < 0.125) {
    // z in three limbs: z + (n - z d) / d, the residual from exact products
    double p1, e1, p2, e2, p3, e3;
    two_prod(z.hi, d.h, p1, e1);
    two_prod(z.hi, d.m, p2, e2);
    two_prod(z.lo, d.h, p3, e3);
    // the partial sums are ~2^-53 of n and r only ~2^-106: add them exactly
    double R0 = n.h - p1, R1 = 0.0, R2 = 0.0;  // exact: p1 is within an ulp of n.h
    tsum3(R0, R1, R2, -e1);
    tsum3(R0, R1, R2, n.m);
    tsum3(R0, R1, R2, -p2);
    tsum3(R0, R1, R2, -p3);
    tsum3(R0, R1, R2, n.l - e2 - e3 - z.hi * d.l - z.lo * d.m);
    const double z2 = (R0 + (R1 + R2)) / d.h;
    double zh, zl, zc;
    fast_two_sum(z.hi, z.lo, zh, zl);
    two_sum(zl, z2, zl, zc);
    return sum3_round(zh, zl, zc, atan_tail(z));
  }
  if (z.hi <= 8.0) {
    const int k = (int)__builtin_rint(z.hi * 64.0);
    const double c = (double)k * 0.015625;
    const ddv num = td_sub_cmul(n, c, d);
    const ddv den = dd_add(dd, dd_mul(mk(c, 0.0), nd));
    const ddv u = dd_div(num, den);
    const double *b = mf_atant + 3 * (k - 8);
    return sum3_round(b[0], b[1], b[2], dd_add(u, atan_tail(u)));
  }
  const ddv w = dd_div(dd, nd);
  return sum3_round(mf_pio2_cw1, mf_pio2_cw2, mf_pio2_cw3, dd_neg(dd_add(w, atan_tail(w))));
}
// sqrt(1 - x^2) in three limbs for |x| <= 1 (NaN beyond, as sqrt of a negative)
MF_HD td3 sqrt_one_minus_sq(ddv x) {
  double p1, e1, p2, e2;
  two_prod(x.hi, x.hi, p1, e1);
  two_prod(2.0 * x.hi, x.lo, p2, e2);
  double T0 = 1.0, T1 = 0.0, T2 = 0.0;
  tsum3(T0, T1, T2, -p1);
  tsum3(T0, T1, T2, -p2);
  tsum3(T0, T1, T2, -e1);
  tsum3(T0, T1, T2, -e2 - x.lo * x.lo);
  double u;
  two_sum(T1, T2, T1, u);
  T2 = u;
  two_sum(T0, T1, T0, u);
  T1 = u;
  if (T0 == 0.0) return td3{0.0, 0.0, 0.0};
  const ddv sq = dd_sqrt(mk(T0, T1));
  if (!(sq.hi > 0.0)) return td3{sq.hi, 0.0, 0.0};  // NaN for 1 - x^2 < 0
  double q1, f1, q2, f2;
  two_prod(sq.hi, sq.hi, q1, f1);
  two_prod(2.0 * sq.hi, sq.lo, q2, f2);
  double P0 = T0 - q1, P1 = 0.0, P2 = 0.0;  // exact: q1 is within an ulp of T0
  tsum3(P0, P1, P2, T1);
  tsum3(P0, P1, P2, -q2);
  tsum3(P0, P1, P2, -f1);
  tsum3(P0, P1, P2, T2 - f2 - sq.lo * sq.lo);
  return td3{sq.hi, sq.lo, (P0 + (P1 + P2)) / (2.0 * sq.hi)};
}
MF_HD ddv dd_asin(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(__builtin_nan(""), 0.0);
  if (__builtin_fabs(x.hi) > 1.0) return mk(__builtin_nan(""), 0.0);
  if (__builtin_fabs(x.hi) < 1.85e-16) return x;  // the cubic term is below half a DD ulp
  const bool neg = x.hi < 0.0;
  const ddv ax = neg ? dd_neg(x) : x;
  const td3 s = sqrt_one_minus_sq(ax);
  if (s.h != s.h) return mk(__builtin_nan(""), 0.0);
  const ddv r = atan_ratio(td3{ax.hi, ax.lo, 0.0}, s);
  return neg ? dd_neg(r) : r;
}
MF_HD ddv dd_acos(ddv x) {
  if (!mf_isfinite(x.hi)) return mk(__builtin_nan(""), 0.0);
  if (__builtin_fabs(x.hi) > 1.0) return mk(__builtin_nan(""), 0.0);
  const bool neg = x.hi < 0.0;
  const ddv ax = neg ? dd_neg(x) : x;
  const td3 s = sqrt_one_minus_sq(ax);
  if (s.h != s.h) return mk(__builtin_nan(""), 0.0);
  const ddv r = atan_ratio(s, td3{ax.hi, ax.lo, 0.0});
  if (!neg) return r;
  return sum3_round(2.0 * mf_pio2_cw1, 2.0 * mf_pio2_cw2, 2.0 * mf_pio2_cw3, dd_neg(r));
}
// inv_trig.inc:197-205: asin / pi (odd, zeros kept) and acos / pi
MF_HD ddv dd_asinpi(ddv x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  return dd_mul(dd_asin(x), mk(mf_inv_pi_hi, mf_inv_pi_lo));
}
MF_HD ddv dd_acospi(ddv x) { return dd_mul(dd_acos(x), mk(mf_inv_pi_hi, mf_inv_pi_lo)); }
