/* Correctly rounded reference values (test infrastructure only) for the
 * functions this project designs itself rather than restating the
 * reference's code: erf, erfc, erfcx, atan, asin, acos (the reference fits
 * them with rationals, src/float64x2/special.inc:18-188 and
 * inv_trig.inc:21-130; its published accuracy bar for erf is max_rel
 * 1.55e-32 vs MPFR at 200 bits, doc/dev/benchmark/baseline/
 * benchmark-raptor-lake.json:1636). Every input is a double-double x = hi + lo
 * (exact when the limbs lie within 2^200 of each other, as in a canonical
 * pair); f(x) is evaluated with MPFR, correctly rounded to 256 bits, and
 * returned as three round-to-nearest limbs v0 + v1 + v2 (v1 = RN(f - v0),
 * v2 = RN(f - v0 - v1)), so errors of a double-double result can be measured
 * to about 2^-150 relative. erfcx(x) = exp(x^2) erfc(x) is formed directly for |x| <= 2^14
 * and from its asymptotic series sum_n (-1)^n (2n-1)!! / (2x^2)^n / (x sqrt(pi))
 * above (the first 24 terms shrink by at least 2^-23 a step there, so they
 * pass 256 bits), where erfc(x) itself would leave MPFR's exponent range. */
#include <stdint.h>
#include <mpfr.h>

#define PREC 256

enum { PIN_ERF = 0, PIN_ERFC = 1, PIN_ERFCX = 2, PIN_ATAN = 3, PIN_ASIN = 4, PIN_ACOS = 5 };

static void limbs3(mpfr_t v, mpfr_t t, double *out) {
  double v0 = mpfr_get_d(v, MPFR_RNDN);
  out[0] = v0;
  if (!(v0 - v0 == 0.0)) { /* Inf or NaN: no lower limbs */
    out[1] = out[2] = 0.0;
    return;
  }
  mpfr_sub_d(t, v, v0, MPFR_RNDN);
  double v1 = mpfr_get_d(t, MPFR_RNDN);
  out[1] = v1;
  mpfr_sub_d(t, t, v1, MPFR_RNDN);
  out[2] = mpfr_get_d(t, MPFR_RNDN);
}

static void erfcx_big(mpfr_t v, const mpfr_t x, mpfr_t t, mpfr_t u) {
  /* sum_{n=0}^{N} (-1)^n (2n-1)!! / (2 x^2)^n, then / (x sqrt(pi)) */
  mpfr_set_ui(v, 1, MPFR_RNDN);
  mpfr_set_ui(t, 1, MPFR_RNDN);
  mpfr_sqr(u, x, MPFR_RNDN);
  mpfr_mul_2ui(u, u, 1, MPFR_RNDN); /* 2 x^2 */
  for (int nn = 1; nn <= 24; ++nn) {
    mpfr_mul_si(t, t, -(2 * nn - 1), MPFR_RNDN);
    mpfr_div(t, t, u, MPFR_RNDN);
    mpfr_add(v, v, t, MPFR_RNDN);
  }
  mpfr_const_pi(t, MPFR_RNDN);
  mpfr_sqrt(t, t, MPFR_RNDN);
  mpfr_mul(t, t, x, MPFR_RNDN);
  mpfr_div(v, v, t, MPFR_RNDN);
}

/* out[3 i .. 3 i + 2] = limbs of f(x_i), x_i = x[2 i] + x[2 i + 1]. */
int pin_eval(int fn, const double *x, double *out, int64_t n) {
  mpfr_set_emin(mpfr_get_emin_min());
  mpfr_set_emax(mpfr_get_emax_max());
  mpfr_t a, v, t, u;
  mpfr_inits2(PREC, a, v, t, u, (mpfr_ptr)0);
  for (int64_t i = 0; i < n; ++i) {
    mpfr_set_d(a, x[2 * i], MPFR_RNDN);
    mpfr_add_d(a, a, x[2 * i + 1], MPFR_RNDN);
    switch (fn) {
      case PIN_ERF: mpfr_erf(v, a, MPFR_RNDN); break;
      case PIN_ERFC: mpfr_erfc(v, a, MPFR_RNDN); break;
      case PIN_ERFCX:
        if (mpfr_cmp_d(a, 16384.0) > 0) {
          erfcx_big(v, a, t, u);
        } else {
          mpfr_erfc(v, a, MPFR_RNDN);
          mpfr_sqr(t, a, MPFR_RNDN);
          mpfr_exp(t, t, MPFR_RNDN);
          mpfr_mul(v, v, t, MPFR_RNDN);
        }
        break;
      case PIN_ATAN: mpfr_atan(v, a, MPFR_RNDN); break;
      case PIN_ASIN: mpfr_asin(v, a, MPFR_RNDN); break;
      case PIN_ACOS: mpfr_acos(v, a, MPFR_RNDN); break;
      default: mpfr_clears(a, v, t, u, (mpfr_ptr)0); return -1;
    }
    limbs3(v, t, out + 3 * i);
  }
  mpfr_clears(a, v, t, u, (mpfr_ptr)0);
  return 0;
}
