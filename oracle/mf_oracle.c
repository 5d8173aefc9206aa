/* CPU oracle for the multifloats DD hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker or the timed CPU baseline. The product
 * (multifloats_amd/) never links or calls it.
 *
 * A plain-C restatement of the reference's CPU algorithms, each function citing
 * the reference file:line it follows (paths relative to the reference root).
 * Built with -O2 -ffp-contract=off (oracle/Makefile) so that no multiply-add is
 * fused behind our back; fma() is the correctly rounded libm call.
 *
 * Pinning (what checks each function, besides agreement with the GPU):
 *   or_add/sub/mul/div/sqrt, or_exp/log/sin, or_matmul_*: bit-exact golden
 *       vectors from the compiled reference (tests/golden), closed forms
 *       (A = I, small integers) and exact-accumulator bounds.
 *   or_sum/or_dot/or_norm2 (Fortran-only in the reference, which cannot be
 *       built here: no fypp, no gfortran): exact-accumulator bounds, invariants
 *       (integer sums exact, Pythagorean triples), and planted-mistake tests.
 *   or_exact_*: brute force against Python integers/fractions on small inputs.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../multifloats_amd/csrc/dd_tables.h"

typedef struct { double hi, lo; } dd;

static inline dd mkdd(double h, double l) { dd r; r.hi = h; r.lo = l; return r; }

/* include/multifloats/float64x2.h:83-107 */
static inline void tsum(double a, double b, double *s, double *e) {
  double t = a + b, ap = t - b, bp = t - ap;
  *e = (a - ap) + (b - bp);
  *s = t;
}
static inline void ftsum(double a, double b, double *s, double *e) {
  double t = a + b, bp = t - a;
  *e = b - bp;
  *s = t;
}
static inline void tprod(double a, double b, double *p, double *e) {
  double q = a * b;
  *e = fma(a, b, -q);
  *p = q;
}

/* float64x2.h:203-229 (operator+) */
static dd o_add(dd x, dd y) {
  double s = x.hi + y.hi, a, b, c, d;
  if (!isfinite(s)) return mkdd(s, 0.0);
  if (x.hi == 0.0 && y.hi == 0.0) return mkdd(s, x.lo + y.lo);
  tsum(x.hi, y.hi, &a, &b);
  tsum(x.lo, y.lo, &c, &d);
  ftsum(a, c, &a, &c);
  b += d;
  b += c;
  dd r;
  ftsum(a, b, &r.hi, &r.lo);
  return r;
}
/* float64x2.h:232-250 (operator-) */
static dd o_sub(dd x, dd y) {
  double s = x.hi - y.hi, a, b, c, d;
  if (!isfinite(s)) return mkdd(s, 0.0);
  if (x.hi == 0.0 && y.hi == 0.0) return mkdd(s, x.lo - y.lo);
  tsum(x.hi, -y.hi, &a, &b);
  tsum(x.lo, -y.lo, &c, &d);
  ftsum(a, c, &a, &c);
  b += d;
  b += c;
  dd r;
  ftsum(a, b, &r.hi, &r.lo);
  return r;
}
/* float64x2.h:253-271 (operator*) */
static dd o_mul(dd x, dd y) {
  double p, e;
  tprod(x.hi, y.hi, &p, &e);
  if (!isfinite(p)) return mkdd(p, 0.0);
  double p01 = x.hi * y.lo, p10 = x.lo * y.hi;
  p01 += p10;
  e += p01;
  dd r;
  ftsum(p, e, &r.hi, &r.lo);
  return r;
}
/* float64x2.h:274-311 (operator/) */
static dd o_div(dd x, dd y) {
  double q1 = x.hi / y.hi;
  if (!isfinite(q1)) return mkdd(q1, q1);
  if (!isfinite(y.hi)) return mkdd(q1, 0.0);
  double p, e, r0, r0e, rh, rl;
  tprod(q1, y.hi, &p, &e);
  double qlo = e + q1 * y.lo;
  tsum(x.hi, -p, &r0, &r0e);
  double r1 = (x.lo - qlo) + r0e;
  ftsum(r0, r1, &rh, &rl);
  double q2 = rh / y.hi;
  dd r;
  ftsum(q1, q2, &r.hi, &r.lo);
  return r;
}
/* float64x2.h:2190-2213 (sqrt) */
static dd o_sqrt(dd x) {
  double c = sqrt(x.hi);
  if (!(x.hi > 0.0) || !isfinite(c)) return mkdd(c, 0.0);
  double u, uu;
  tprod(c, c, &u, &uu);
  double cc = ((x.hi - u) - uu + x.lo) / (2.0 * c);
  double h = c + cc;
  return mkdd(h, (c - h) + cc);
}

#define ARR1(name, fn)                                              \
  void name(const dd *a, dd *c, int64_t n) {                        \
    for (int64_t i = 0; i < n; ++i) c[i] = fn(a[i]);                \
  }
#define ARR2(name, fn)                                              \
  void name(const dd *a, const dd *b, dd *c, int64_t n) {           \
    for (int64_t i = 0; i < n; ++i) c[i] = fn(a[i], b[i]);          \
  }
ARR2(or_add, o_add)
ARR2(or_sub, o_sub)
ARR2(or_mul, o_mul)
ARR2(or_div, o_div)
ARR1(or_sqrt, o_sqrt)

/* ---------------------------------------------------------------- GEMM */
/* src/float64x2/matmul.inc:6-19 (mac_inl) */
static inline void o_mac(double ah, double al, double bh, double bl, double *sh, double *sl) {
  double p = ah * bh;
  double e = fma(ah, bh, -p);
  double cross = fma(ah, bl, al * bh);
  double t = *sh + p, bp = t - *sh, ap = t - bp;
  double aerr = *sh - ap, berr = p - bp;
  *sl += (e + cross) + (aerr + berr);
  *sh = t;
}
/* matmul.inc:24-39 (renorm_inl / finalize_inl) */
static inline void o_renorm(double *sh, double *sl) {
  double t = *sh + *sl, bp = t - *sh, ap = t - bp;
  double aerr = *sh - ap, berr = *sl - bp;
  *sl = aerr + berr;
  *sh = t;
}
/* One output of matmul.inc's panels: sequential p, renorm after every full
 * chunk of ri MACs that is not the last (matmul.inc:55-87, 153-172). The
 * reference's MR x NR blocking never changes this per-output order. */
static dd o_chain(const dd *a, int64_t sa, const dd *b, int64_t sb, int64_t k, int64_t ri) {
  double sh = 0.0, sl = 0.0;
  int chunked = !(ri <= 0 || k <= ri);
  for (int64_t p = 0; p < k; ++p) {
    o_mac(a[p * sa].hi, a[p * sa].lo, b[p * sb].hi, b[p * sb].lo, &sh, &sl);
    if (chunked && (p + 1) % ri == 0 && p + 1 < k) o_renorm(&sh, &sl);
  }
  o_renorm(&sh, &sl);
  return mkdd(sh, sl);
}
/* matmul.inc:182-205 (mm_impl): C(m,n) = A(m,k) B(k,n), column-major */
void or_matmul_mm(const dd *a, const dd *b, dd *c, int64_t m, int64_t k, int64_t n, int64_t ri) {
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) c[i + j * m] = o_chain(a + i, m, b + j * k, 1, k, ri);
}
/* matmul.inc:207-211 (mv_impl) */
void or_matmul_mv(const dd *a, const dd *x, dd *y, int64_t m, int64_t k, int64_t ri) {
  for (int64_t i = 0; i < m; ++i) y[i] = o_chain(a + i, m, x, 1, k, ri);
}
/* matmul.inc:213-239 (vm_impl): x is the "a" operand of the MAC */
void or_matmul_vm(const dd *x, const dd *b, dd *y, int64_t k, int64_t n, int64_t ri) {
  for (int64_t j = 0; j < n; ++j) y[j] = o_chain(x, 1, b + j * k, 1, k, ri);
}
/* BLAS-style C = alpha op(A) op(B) + beta C with the argument rules and quick
 * returns of the reference's xGEMM shim (test/integration/blas/wgemm.f90:16-69),
 * op(A) op(B) formed as matmul.inc's per-output chains (o_chain) and combined
 * with the DD operators as o_add(o_mul(alpha, T), o_mul(beta, C)) (beta == 0:
 * o_mul(alpha, T), C not read). alpha == 0 or k == 0: C = beta C (zeros when
 * beta == 0). Returns the wgemm.f90 info code (0 = success, C untouched else). */
int or_gemm_blas(char ta, char tb, int64_t m, int64_t n, int64_t k, const dd *alpha, const dd *a, int64_t lda,
                 const dd *b, int64_t ldb, const dd *beta, dd *c, int64_t ldc, int64_t ri) {
  int nota = ta == 'N' || ta == 'n', notb = tb == 'N' || tb == 'n';
  int oka = nota || ta == 'T' || ta == 't' || ta == 'C' || ta == 'c';
  int okb = notb || tb == 'T' || tb == 't' || tb == 'C' || tb == 'c';
  int64_t nrowa = nota ? m : k, nrowb = notb ? k : n;
  if (!oka) return 1;
  if (!okb) return 2;
  if (m < 0) return 3;
  if (n < 0) return 4;
  if (k < 0) return 5;
  if (lda < (nrowa > 1 ? nrowa : 1)) return 8;
  if (ldb < (nrowb > 1 ? nrowb : 1)) return 10;
  if (ldc < (m > 1 ? m : 1)) return 13;
  int alpha0 = alpha->hi == 0.0 && alpha->lo == 0.0, beta0 = beta->hi == 0.0 && beta->lo == 0.0;
  int beta1 = beta->hi == 1.0 && beta->lo == 0.0;
  if (m == 0 || n == 0 || ((alpha0 || k == 0) && beta1)) return 0;
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      dd *cij = &c[i + j * ldc], o;
      if (alpha0 || k == 0) {
        o = beta0 ? mkdd(0.0, 0.0) : o_mul(*beta, *cij);
      } else {
        /* op(A)(i, p) and op(B)(p, j) as base pointer + stride over p */
        dd t = o_chain(nota ? a + i : a + i * lda, nota ? lda : 1, notb ? b + j * ldb : b + j, notb ? 1 : ldb,
                       k, ri);
        o = o_mul(*alpha, t);
        if (!beta0) o = o_add(o, o_mul(*beta, *cij));
      }
      *cij = o;
    }
  return 0;
}
/* DD comparisons of the Fortran module (fypp:2913-2940): both limbs equal. */
static inline int o_ne(dd x, dd y) { return x.hi != y.hi || x.lo != y.lo; }

/* Triangular solve op(A) X = alpha B (side 'L') or X op(A) = alpha B (side
 * 'R'), X overwriting B: the reference's xTRSM shim
 * test/integration/blas/wtrsm.f90, its argument checks (24-47), quick returns
 * (48-56) and all eight loop nests (57-206) restated with the DD operators in
 * the same order, including its skips (b(k,j) == 0, a(k,j) == 0). Returns the
 * info code (0 = success; B untouched otherwise). */
int or_trsm(char side, char uplo, char transa, char diag, int64_t m, int64_t n, const dd *alpha, const dd *a,
            int64_t lda, dd *b, int64_t ldb) {
#define A_(i, j) a[(i) + (j) * lda]
#define B_(i, j) b[(i) + (j) * ldb]
  int lside = side == 'L' || side == 'l', upper = uplo == 'U' || uplo == 'u';
  int notr = transa == 'N' || transa == 'n', nounit = diag == 'N' || diag == 'n';
  int64_t nrowa = lside ? m : n;
  const dd zero = mkdd(0.0, 0.0), one = mkdd(1.0, 0.0);
  if (!lside && !(side == 'R' || side == 'r')) return 1;
  if (!upper && !(uplo == 'L' || uplo == 'l')) return 2;
  if (!notr && !(transa == 'T' || transa == 't' || transa == 'C' || transa == 'c')) return 3;
  if (!(diag == 'U' || diag == 'u') && !nounit) return 4;
  if (m < 0) return 5;
  if (n < 0) return 6;
  if (lda < (nrowa > 1 ? nrowa : 1)) return 9;
  if (ldb < (m > 1 ? m : 1)) return 11;
  if (m == 0 || n == 0) return 0;
  if (!o_ne(*alpha, zero)) {
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < m; ++i) B_(i, j) = zero;
    return 0;
  }
  const int scale = o_ne(*alpha, one);
  if (lside && notr) { /* wtrsm.f90:58-101 */
    for (int64_t j = 0; j < n; ++j) {
      if (scale)
        for (int64_t i = 0; i < m; ++i) B_(i, j) = o_mul(*alpha, B_(i, j));
      for (int64_t t = 0; t < m; ++t) {
        int64_t k = upper ? m - 1 - t : t;
        if (!o_ne(B_(k, j), zero)) continue;
        if (nounit) B_(k, j) = o_div(B_(k, j), A_(k, k));
        if (upper)
          for (int64_t i = 0; i < k; ++i) B_(i, j) = o_sub(B_(i, j), o_mul(B_(k, j), A_(i, k)));
        else
          for (int64_t i = k + 1; i < m; ++i) B_(i, j) = o_sub(B_(i, j), o_mul(B_(k, j), A_(i, k)));
      }
    }
  } else if (lside) { /* wtrsm.f90:102-126 */
    for (int64_t j = 0; j < n; ++j)
      for (int64_t t = 0; t < m; ++t) {
        int64_t i = upper ? t : m - 1 - t;
        dd temp = o_mul(*alpha, B_(i, j));
        if (upper)
          for (int64_t k = 0; k < i; ++k) temp = o_sub(temp, o_mul(A_(k, i), B_(k, j)));
        else
          for (int64_t k = i + 1; k < m; ++k) temp = o_sub(temp, o_mul(A_(k, i), B_(k, j)));
        if (nounit) temp = o_div(temp, A_(i, i));
        B_(i, j) = temp;
      }
  } else if (notr) { /* wtrsm.f90:128-166 */
    for (int64_t t = 0; t < n; ++t) {
      int64_t j = upper ? t : n - 1 - t;
      if (scale)
        for (int64_t i = 0; i < m; ++i) B_(i, j) = o_mul(*alpha, B_(i, j));
      int64_t k0 = upper ? 0 : j + 1, k1 = upper ? j : n;
      for (int64_t k = k0; k < k1; ++k) {
        if (!o_ne(A_(k, j), zero)) continue;
        for (int64_t i = 0; i < m; ++i) B_(i, j) = o_sub(B_(i, j), o_mul(A_(k, j), B_(i, k)));
      }
      if (nounit) {
        dd temp = o_div(one, A_(j, j));
        for (int64_t i = 0; i < m; ++i) B_(i, j) = o_mul(temp, B_(i, j));
      }
    }
  } else { /* wtrsm.f90:167-205 */
    for (int64_t t = 0; t < n; ++t) {
      int64_t k = upper ? n - 1 - t : t;
      if (nounit) {
        dd temp = o_div(one, A_(k, k));
        for (int64_t i = 0; i < m; ++i) B_(i, k) = o_mul(temp, B_(i, k));
      }
      int64_t j0 = upper ? 0 : k + 1, j1 = upper ? k : n;
      for (int64_t j = j0; j < j1; ++j) {
        if (!o_ne(A_(j, k), zero)) continue;
        dd temp = A_(j, k);
        for (int64_t i = 0; i < m; ++i) B_(i, j) = o_sub(B_(i, j), o_mul(temp, B_(i, k)));
      }
      if (scale)
        for (int64_t i = 0; i < m; ++i) B_(i, k) = o_mul(*alpha, B_(i, k));
    }
  }
  return 0;
#undef A_
#undef B_
}

/* fsrc/multifloats.fypp:2609-2621 (cdd_matmul_mm): four real GEMMs, then the
 * native DD - and + (bit-identical to the C++ operators, fypp:2772-2790).
 * Complex arrays are interleaved {re.hi, re.lo, im.hi, im.lo}. */
void or_cmatmul_mm(const dd *a, const dd *b, dd *c, int64_t m, int64_t k, int64_t n, int64_t ri) {
  dd *ar = malloc(sizeof(dd) * m * k), *ai = malloc(sizeof(dd) * m * k);
  dd *br = malloc(sizeof(dd) * k * n), *bi = malloc(sizeof(dd) * k * n);
  dd *t1 = malloc(sizeof(dd) * m * n), *t2 = malloc(sizeof(dd) * m * n);
  for (int64_t i = 0; i < m * k; ++i) { ar[i] = a[2 * i]; ai[i] = a[2 * i + 1]; }
  for (int64_t i = 0; i < k * n; ++i) { br[i] = b[2 * i]; bi[i] = b[2 * i + 1]; }
  or_matmul_mm(ar, br, t1, m, k, n, ri);
  or_matmul_mm(ai, bi, t2, m, k, n, ri);
  for (int64_t i = 0; i < m * n; ++i) c[2 * i] = o_sub(t1[i], t2[i]);
  or_matmul_mm(ar, bi, t1, m, k, n, ri);
  or_matmul_mm(ai, br, t2, m, k, n, ri);
  for (int64_t i = 0; i < m * n; ++i) c[2 * i + 1] = o_add(t1[i], t2[i]);
  free(ar); free(ai); free(br); free(bi); free(t1); free(t2);
}

/* ----------------------------------------------------------- reductions */
/* fypp:2254-2308 (dd_dot_product): note cross = ah*bl + al*bh without fma,
 * s_lo += (((e + cross) + aerr) + berr), and a fast_two_sum after EVERY chunk
 * of ri elements including the last, then a final fast_two_sum. */
dd or_dot(const dd *a, const dd *b, int64_t n, int64_t ri) {
  double sh = 0.0, sl = 0.0;
  int64_t step = ri <= 0 ? (n > 0 ? n : 1) : ri;
  for (int64_t kk = 0; kk < n; kk += step) {
    int64_t ke = kk + step < n ? kk + step : n;
    for (int64_t k = kk; k < ke; ++k) {
      double ah = a[k].hi, al = a[k].lo, bh = b[k].hi, bl = b[k].lo;
      double p = ah * bh;
      double e = fma(ah, bh, -p);
      double cross = ah * bl + al * bh;
      double t = sh + p, bp = t - sh, ap = t - bp;
      double aerr = sh - ap, berr = p - bp;
      sl = sl + (((e + cross) + aerr) + berr);
      sh = t;
    }
    if (ri > 0) {
      double t = sh + sl, bp = t - sh;
      sh = t;
      sl = sl - bp;
    }
  }
  double t = sh + sl, bp = t - sh;
  return mkdd(t, sl - bp);
}
/* fypp:1980-2020 (dd_sum_1d, no mask) */
dd or_sum(const dd *a, int64_t n) {
  double sh = 0.0, sl = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double xh = a[i].hi, xl = a[i].lo;
    double t = sh + xh, bp = t - sh, ap = t - bp;
    double aerr = sh - ap, berr = xh - bp;
    sl = sl + ((xl + aerr) + berr);
    sh = t;
  }
  double t = sh + sl, bp = t - sh;
  return mkdd(t, sl - bp);
}
/* fypp:2338-2352 (dd_norm2_1d): s = s + x*x with the DD operators, then the
 * DD sqrt (sqrtdd, float64x2.h:2190-2213). No overflow scaling. */
dd or_norm2(const dd *a, int64_t n) {
  dd s = mkdd(0.0, 0.0);
  for (int64_t i = 0; i < n; ++i) s = o_add(s, o_mul(a[i], a[i]));
  return o_sqrt(s);
}

/* fypp:1980-2020 with mask present: masked-out elements are skipped
 * (`cycle`), not added. */
dd or_sum_mask(const dd *a, const int32_t *mask, int64_t n) {
  double sh = 0.0, sl = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    if (!mask[i]) continue;
    double xh = a[i].hi, xl = a[i].lo;
    double t = sh + xh, bp = t - sh, ap = t - bp;
    double aerr = sh - ap, berr = xh - bp;
    sl = sl + ((xl + aerr) + berr);
    sh = t;
  }
  double t = sh + sl, bp = t - sh;
  return mkdd(t, sl - bp);
}

/* fypp:2321-2331 (cdd_dot_product): conj(a).b as four real dot_products,
 * re = dot(ar,br) + dot(ai,bi), im = dot(ar,bi) - dot(ai,br). Complex
 * elements are {re, im} pairs of dd. */
void or_cdot(const dd *a, const dd *b, int64_t n, int64_t ri, dd *out) {
  size_t nn = n > 0 ? (size_t)n : 1;
  dd *ar = malloc(nn * sizeof(dd)), *ai = malloc(nn * sizeof(dd));
  dd *br = malloc(nn * sizeof(dd)), *bi = malloc(nn * sizeof(dd));
  for (int64_t i = 0; i < n; ++i) {
    ar[i] = a[2 * i]; ai[i] = a[2 * i + 1];
    br[i] = b[2 * i]; bi[i] = b[2 * i + 1];
  }
  out[0] = o_add(or_dot(ar, br, n, ri), or_dot(ai, bi, n, ri));
  out[1] = o_sub(or_dot(ar, bi, n, ri), or_dot(ai, br, n, ri));
  free(ar); free(ai); free(br); free(bi);
}

/* fypp:2035-2090 (dd_sum_Nd_dim) on an (inner, len, outer) column-major view:
 * out(i, o) is the Neumaier chain over j (masked elements skipped), closed by
 * the full TwoSum in the reference's order (t, bp, ap, aerr, berr). */
/* product / maxval / minval over a strided fibre: the reference's loops
 * (fypp:1786-1810 maxval/minval, 2005-2016 product; the _dim forms
 * 1825-1862, 2076-2087 run the same loop per output). kind 0 product
 * (res = res * x from 1), 1 maxval, 2 minval (res = first element, replaced
 * by x when x > res / x < res, lexicographic on (hi, lo), fypp:2960-2998;
 * 0 when nothing is selected). Masked-out elements are skipped. */
static int o_gt(dd x, dd y) { return y.hi < x.hi || (x.hi == y.hi && y.lo < x.lo); }
static dd o_fibre(int kind, const dd *a, const int32_t *mask, int64_t start, int64_t stride, int64_t len) {
  dd res = mkdd(kind == 0 ? 1.0 : 0.0, 0.0);
  int have = 0;
  for (int64_t j = 0; j < len; ++j) {
    int64_t e = start + stride * j;
    if (mask && !mask[e]) continue;
    if (kind == 0) {
      res = o_mul(res, a[e]);
    } else if (!have) {
      res = a[e];
      have = 1;
    } else if (kind == 1 ? o_gt(a[e], res) : o_gt(res, a[e])) {
      res = a[e];
    }
  }
  return res;
}
void or_fibre_reduce(int kind, const dd *a, const int32_t *mask, int64_t n, dd *out) {
  *out = o_fibre(kind, a, mask, 0, 1, n);
}
void or_fibre_reduce_dim(int kind, const dd *a, const int32_t *mask, int64_t inner, int64_t len, int64_t outer,
                         dd *out) {
  for (int64_t o = 0; o < outer; ++o)
    for (int64_t i = 0; i < inner; ++i) out[i + inner * o] = o_fibre(kind, a, mask, i + inner * len * o, inner, len);
}
void or_sum_dim(const dd *a, const int32_t *mask, int64_t inner, int64_t len, int64_t outer, dd *out) {
  for (int64_t o = 0; o < outer; ++o)
    for (int64_t i = 0; i < inner; ++i) {
      double sh = 0.0, sl = 0.0;
      for (int64_t j = 0; j < len; ++j) {
        int64_t e = i + inner * (j + len * o);
        if (mask && !mask[e]) continue;
        double xh = a[e].hi, xl = a[e].lo;
        double t = sh + xh, bp = t - sh, ap = t - bp;
        double aerr = sh - ap, berr = xh - bp;
        sl = sl + ((xl + aerr) + berr);
        sh = t;
      }
      double t = sh + sl, bp = t - sh, ap = t - bp;
      double aerr = sh - ap, berr = sl - bp;
      out[i + inner * o] = mkdd(t, aerr + berr);
    }
}

/* fypp:2356-2377 (dd_norm2_Nd_dim): cur = cur + x*x with the DD operators
 * along dimension `dim`, then DD sqrt. */
void or_norm2_dim(const dd *a, int64_t inner, int64_t len, int64_t outer, dd *out) {
  for (int64_t o = 0; o < outer; ++o)
    for (int64_t i = 0; i < inner; ++i) {
      dd s = mkdd(0.0, 0.0);
      for (int64_t j = 0; j < len; ++j) {
        dd x = a[i + inner * (j + len * o)];
        s = o_add(s, o_mul(x, x));
      }
      out[i + inner * o] = o_sqrt(s);
    }
}

/* ------------------------------------------------------ transcendentals */
static inline dd cdd(const double *hi, const double *lo, int i) { return mkdd(hi[i], lo[i]); }

/* src/float64x2/poly.inc:69-79 (neval, degree 10) */
static dd o_estrin10(dd x, const double *h, const double *l) {
  dd x2 = o_mul(x, x), x4 = o_mul(x2, x2), x8 = o_mul(x4, x4);
  dd p01 = o_add(cdd(h, l, 0), o_mul(cdd(h, l, 1), x));
  dd p23 = o_add(cdd(h, l, 2), o_mul(cdd(h, l, 3), x));
  dd p45 = o_add(cdd(h, l, 4), o_mul(cdd(h, l, 5), x));
  dd p67 = o_add(cdd(h, l, 6), o_mul(cdd(h, l, 7), x));
  dd p89 = o_add(cdd(h, l, 8), o_mul(cdd(h, l, 9), x));
  dd p03 = o_add(p01, o_mul(p23, x2));
  dd p47 = o_add(p45, o_mul(p67, x2));
  dd p810 = o_add(p89, o_mul(cdd(h, l, 10), x2));
  return o_add(o_add(p03, o_mul(p47, x4)), o_mul(p810, x8));
}
/* poly.inc:92-104 (neval, degree 12) */
static dd o_estrin12(dd x, const double *h, const double *l) {
  dd x2 = o_mul(x, x), x4 = o_mul(x2, x2), x8 = o_mul(x4, x4);
  dd p01 = o_add(cdd(h, l, 0), o_mul(cdd(h, l, 1), x));
  dd p23 = o_add(cdd(h, l, 2), o_mul(cdd(h, l, 3), x));
  dd p45 = o_add(cdd(h, l, 4), o_mul(cdd(h, l, 5), x));
  dd p67 = o_add(cdd(h, l, 6), o_mul(cdd(h, l, 7), x));
  dd p89 = o_add(cdd(h, l, 8), o_mul(cdd(h, l, 9), x));
  dd p1011 = o_add(cdd(h, l, 10), o_mul(cdd(h, l, 11), x));
  dd p03 = o_add(p01, o_mul(p23, x2));
  dd p47 = o_add(p45, o_mul(p67, x2));
  dd p811 = o_add(p89, o_mul(p1011, x2));
  dd top = o_add(p811, o_mul(cdd(h, l, 12), x4));
  return o_add(o_add(p03, o_mul(p47, x4)), o_mul(top, x8));
}
/* poly.inc:116-129 (neval, degree 14) */
static dd o_estrin14(dd x, const double *h, const double *l) {
  dd x2 = o_mul(x, x), x4 = o_mul(x2, x2), x8 = o_mul(x4, x4);
  dd p01 = o_add(cdd(h, l, 0), o_mul(cdd(h, l, 1), x));
  dd p23 = o_add(cdd(h, l, 2), o_mul(cdd(h, l, 3), x));
  dd p45 = o_add(cdd(h, l, 4), o_mul(cdd(h, l, 5), x));
  dd p67 = o_add(cdd(h, l, 6), o_mul(cdd(h, l, 7), x));
  dd p89 = o_add(cdd(h, l, 8), o_mul(cdd(h, l, 9), x));
  dd p1011 = o_add(cdd(h, l, 10), o_mul(cdd(h, l, 11), x));
  dd p1213 = o_add(cdd(h, l, 12), o_mul(cdd(h, l, 13), x));
  dd p03 = o_add(p01, o_mul(p23, x2));
  dd p47 = o_add(p45, o_mul(p67, x2));
  dd p811 = o_add(p89, o_mul(p1011, x2));
  dd p1214 = o_add(p1213, o_mul(cdd(h, l, 14), x2));
  return o_add(o_add(p03, o_mul(p47, x4)), o_mul(o_add(p811, o_mul(p1214, x4)), x8));
}

/* (int)x as the reference's static_cast<int> behaves on x86-64, without the
   undefined behaviour of the C cast: NaN and values outside int's range give
   INT_MIN (cvttsd2si's "integer indefinite"). Only NaN arguments reach that
   case here, and their results are NaN whatever the index. */
static int o_toint(double x) { return (x == x && fabs(x) < 2147483648.0) ? (int)x : INT_MIN; }

/* src/float64x2/exp_log.inc:26-72 (exp2_from_reduced) */
static dd o_exp2_reduced(double n_d, double rh, double rl) {
  if (n_d > 1100.0) return mkdd(INFINITY, 0.0);
  if (n_d < -1100.0) return mkdd(0.0, 0.0);
  int k = o_toint(nearbyint(rh * 256.0));
  if (k > 128) k = 128;
  if (k < -128) k = -128;
  double sh = rh - (double)k / 256.0, sl = rl;
  ftsum(sh, sl, &sh, &sl);
  dd s = mkdd(sh, sl);
  dd q = cdd(mf_exp2c_hi, mf_exp2c_lo, 9);
  for (int j = 8; j >= 1; --j) q = o_add(o_mul(q, s), cdd(mf_exp2c_hi, mf_exp2c_lo, j));
  dd pow2s = o_add(mkdd(1.0, 0.0), o_mul(s, q));
  dd mant = o_mul(cdd(mf_exp2t_hi, mf_exp2t_lo, k + 128), pow2s);
  int n = o_toint(n_d);
  return mkdd(ldexp(mant.hi, n), ldexp(mant.lo, n));
}
/* exp_log.inc:237-277 (exp) */
static dd o_exp(dd x) {
  if (!isfinite(x.hi)) return mkdd(x.hi > 0.0 ? x.hi : (x.hi < 0.0 ? 0.0 : x.hi), 0.0);
  if (x.hi > mf_exp_max) return mkdd(INFINITY, 0.0);
  if (x.hi < mf_exp_min) return mkdd(0.0, 0.0);
  double n_d = nearbyint(x.hi * mf_log2e_hi), rh, rl, t;
  tsum(x.hi, -(n_d * mf_ln2_cw1), &rh, &t);
  rl = t + x.lo;
  ftsum(rh, rl, &rh, &rl);
  double p_hi = n_d * mf_ln2_cw2, p_lo = fma(n_d, mf_ln2_cw2, -p_hi);
  tsum(rh, -p_hi, &rh, &t);
  rl += t - p_lo;
  ftsum(rh, rl, &rh, &rl);
  tsum(rh, -(n_d * mf_ln2_cw3), &rh, &t);
  rl += t;
  dd rr;
  ftsum(rh, rl, &rr.hi, &rr.lo);
  dd y = o_mul(rr, mkdd(mf_log2e_hi, mf_log2e_lo));
  return o_exp2_reduced(n_d, y.hi, y.lo);
}
/* exp_log.inc:280-308 (log2_kernel) */
static dd o_log2_kernel(dd x) {
  double hi = x.hi;
  if (hi > 15.0 / 16.0 && hi < 17.0 / 16.0) {
    dd t = o_div(o_sub(x, mkdd(1.0, 0.0)), o_add(x, mkdd(1.0, 0.0)));
    return o_mul(t, o_estrin10(o_mul(t, t), mf_log2c_hi, mf_log2c_lo));
  }
  int e = ilogb(hi);
  dd m = mkdd(ldexp(x.hi, -e), ldexp(x.lo, -e));
  int idx = (int)((m.hi - 1.0) * 32.0);
  if (idx < 0) idx = 0;
  if (idx > 31) idx = 31;
  dd c = mkdd(mf_log2_centre[idx], 0.0);
  dd v = cdd(mf_log2v_hi, mf_log2v_lo, idx);
  dd t = o_div(o_sub(m, c), o_add(m, c));
  dd ts = o_mul(t, t);
  dd p = cdd(mf_log2c_hi, mf_log2c_lo, 6);  /* horner, poly.inc:9-16 */
  for (int i = 5; i >= 0; --i) p = o_add(o_mul(p, ts), cdd(mf_log2c_hi, mf_log2c_lo, i));
  return o_add(o_add(mkdd((double)e, 0.0), v), o_mul(t, p));
}
/* exp_log.inc:311-341 (log2, log) */
static dd o_log(dd x) {
  if (!isfinite(x.hi)) return mkdd(log(x.hi), 0.0);
  if (isnan(x.hi) || x.hi < 0.0) return mkdd(NAN, 0.0);
  if (x.hi == 0.0) return mkdd(-INFINITY, 0.0);
  dd l2 = o_log2_kernel(x);
  if (!isfinite(l2.hi)) return l2;
  return o_mul(l2, mkdd(mf_ln2_hi, mf_ln2_lo));
}
/* src/float64x2/trig.inc:114-147 (reduce_pi_half) */
static void o_reduce(dd x, dd *r, int *q) {
  double n = nearbyint(x.hi * 0.6366197723675814), rh = x.hi, rl = x.lo;
  double cw[3] = {mf_pio2_cw1, mf_pio2_cw2, mf_pio2_cw3};
  for (int s = 0; s < 3; ++s) {
    double ph = n * cw[s], pl = s < 2 ? fma(n, cw[s], -ph) : 0.0, sh, sl, tl, tll;
    tsum(rh, -ph, &sh, &sl);
    tsum(rl, -pl, &tl, &tll);
    sl += tl;
    ftsum(sh, sl, &rh, &rl);
    rl += tll;
    ftsum(rh, rl, &rh, &rl);
  }
  *r = mkdd(rh, rl);
  long long nn = (long long)fabs(n) & 3LL;
  if (n < 0.0) nn = (4 - nn) & 3;
  *q = (int)nn;
}
/* trig.inc:151-187 (sin_kernel, cos_kernel, sin_eval, cos_eval) */
static dd o_sink(dd x) { return o_mul(o_estrin12(o_mul(x, x), mf_sinc_hi, mf_sinc_lo), x); }
static dd o_cosk(dd x) { return o_estrin12(o_mul(x, x), mf_cosc_hi, mf_cosc_lo); }
static dd o_eval(dd r, int want_cos) {
  if (fabs(r.hi) <= 0.392699081698724) return want_cos ? o_cosk(r) : o_sink(r);
  dd pi4 = mkdd(mf_pi_quarter_hi, mf_pi_quarter_lo), is2 = mkdd(mf_inv_sqrt2_hi, mf_inv_sqrt2_lo);
  int pos = r.hi > 0.0;
  dd rp = pos ? o_sub(r, pi4) : o_add(r, pi4);
  dd sk = o_sink(rp), ck = o_cosk(rp);
  if (!want_cos) return o_mul(pos ? o_add(sk, ck) : o_sub(sk, ck), is2);
  return o_mul(pos ? o_sub(ck, sk) : o_add(ck, sk), is2);
}
/* trig.inc:241-253 (sin) */
static dd o_sin(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!isfinite(x.hi) || !(fabs(x.hi) < 0x1p55)) return mkdd(NAN, 0.0);
  dd r;
  int q;
  o_reduce(x, &r, &q);
  dd v = o_eval(r, q & 1);
  return (q & 2) ? mkdd(-v.hi, -v.lo) : v;
}
/* poly.inc:33-217 (neval): degrees 17 and 24 extend the degree-15 tree by
 * the top coefficients times x^16 */
static dd o_estrin15(dd x, dd x2, dd x4, dd x8, const double *h, const double *l) {
  dd p[8];
  for (int i = 0; i < 8; ++i) p[i] = o_add(cdd(h, l, 2 * i), o_mul(cdd(h, l, 2 * i + 1), x));
  dd p03 = o_add(p[0], o_mul(p[1], x2)), p47 = o_add(p[2], o_mul(p[3], x2));
  dd p811 = o_add(p[4], o_mul(p[5], x2)), p1215 = o_add(p[6], o_mul(p[7], x2));
  return o_add(o_add(p03, o_mul(p47, x4)), o_mul(o_add(p811, o_mul(p1215, x4)), x8));
}
static dd o_estrin17(dd x, const double *h, const double *l) {
  dd x2 = o_mul(x, x), x4 = o_mul(x2, x2), x8 = o_mul(x4, x4), x16 = o_mul(x8, x8);
  dd p1617 = o_add(cdd(h, l, 16), o_mul(cdd(h, l, 17), x));
  return o_add(o_estrin15(x, x2, x4, x8, h, l), o_mul(p1617, x16));
}
static dd o_estrin24(dd x, const double *h, const double *l) {
  dd x2 = o_mul(x, x), x4 = o_mul(x2, x2), x8 = o_mul(x4, x4), x16 = o_mul(x8, x8);
  dd lo = o_estrin15(x, x2, x4, x8, h, l);
  dd q[4];
  for (int i = 0; i < 4; ++i) q[i] = o_add(cdd(h, l, 16 + 2 * i), o_mul(cdd(h, l, 17 + 2 * i), x));
  dd p1619 = o_add(q[0], o_mul(q[1], x2)), p2023 = o_add(q[2], o_mul(q[3], x2));
  dd hi = o_add(o_add(p1619, o_mul(p2023, x4)), o_mul(cdd(h, l, 24), x8));
  return o_add(lo, o_mul(hi, x16));
}
static dd o_neg(dd x) { return mkdd(-x.hi, -x.lo); }
/* trig.inc:255-266 (cos) */
static dd o_cos(dd x) {
  if (!isfinite(x.hi) || !(fabs(x.hi) < 0x1p55)) return mkdd(NAN, 0.0);
  dd r;
  int q;
  o_reduce(x, &r, &q);
  switch (q) {
    case 0: return o_eval(r, 1);
    case 1: return o_neg(o_eval(r, 0));
    case 2: return o_neg(o_eval(r, 1));
    default: return o_eval(r, 0);
  }
}
/* trig.inc:268-289 (tan) with sincos_eval 192-215 */
static dd o_tan(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!isfinite(x.hi) || !(fabs(x.hi) < 0x1p55)) return mkdd(NAN, 0.0);
  dd r, s, c;
  int q;
  o_reduce(x, &r, &q);
  if (fabs(r.hi) <= 0.392699081698724) {
    s = o_sink(r);
    c = o_cosk(r);
  } else {
    dd pi4 = mkdd(mf_pi_quarter_hi, mf_pi_quarter_lo), is2 = mkdd(mf_inv_sqrt2_hi, mf_inv_sqrt2_lo);
    int pos = r.hi > 0.0;
    dd rp = pos ? o_sub(r, pi4) : o_add(r, pi4);
    dd sk = o_sink(rp), ck = o_cosk(rp);
    s = o_mul(pos ? o_add(sk, ck) : o_sub(sk, ck), is2);
    c = o_mul(pos ? o_sub(ck, sk) : o_add(ck, sk), is2);
  }
  return (q == 0 || q == 2) ? o_div(s, c) : o_div(o_neg(c), s);
}
/* exp_log.inc:109-128 (exp2, exp2_kernel) */
static dd o_exp2(dd x) {
  if (!isfinite(x.hi)) return mkdd(x.hi > 0.0 ? x.hi : (x.hi < 0.0 ? 0.0 : x.hi), 0.0);
  if (x.hi < mf_exp2_min) return mkdd(0.0, 0.0);
  if (x.hi > mf_exp2_max) return mkdd(INFINITY, 0.0);
  double n_d = nearbyint(x.hi), rh, rl, t;
  tsum(x.hi, -n_d, &rh, &t);
  rl = t + x.lo;
  ftsum(rh, rl, &rh, &rl);
  return o_exp2_reduced(n_d, rh, rl);
}
/* exp_log.inc:407-427 (expm1) */
static dd o_expm1(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!isfinite(x.hi)) return mkdd(x.hi > 0.0 ? x.hi : (x.hi < 0.0 ? -1.0 : x.hi), 0.0);
  if (fabs(x.hi) >= 0.5) return o_sub(o_exp(x), mkdd(1.0, 0.0));
  return o_mul(x, o_estrin24(x, mf_expm1c_hi, mf_expm1c_lo));
}
/* exp_log.inc:311-327 (log2) */
static dd o_log2(dd x) {
  if (isnan(x.hi) || x.hi < 0.0) return mkdd(NAN, 0.0);
  if (x.hi == 0.0) return mkdd(-INFINITY, 0.0);
  if (!isfinite(x.hi)) return mkdd(INFINITY, 0.0);
  return o_log2_kernel(x);
}
/* exp_log.inc:343-353 (log10) */
static dd o_log10(dd x) {
  if (!isfinite(x.hi)) return mkdd(log10(x.hi), 0.0);
  dd l2 = o_log2(x);
  if (!isfinite(l2.hi)) return l2;
  return o_mul(l2, mkdd(mf_log10_2_hi, mf_log10_2_lo));
}
/* exp_log.inc:429-469 (log1p) */
static dd o_log1p(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!isfinite(x.hi)) return mkdd(log1p(x.hi), 0.0);
  dd arg = o_add(mkdd(1.0, 0.0), x);
  if (arg.hi <= 0.0) {
    if (arg.hi == 0.0 && arg.lo == 0.0) return mkdd(-INFINITY, 0.0);
    if (arg.hi < 0.0 || (arg.hi == 0.0 && arg.lo < 0.0)) return mkdd(NAN, 0.0);
    return o_log(mkdd(arg.lo, 0.0));
  }
  if (fabs(x.hi) >= 0.25) return o_log(arg);
  dd t = o_div(x, o_add(mkdd(2.0, 0.0), x));
  return o_mul(t, o_estrin17(o_mul(t, t), mf_log1pc_hi, mf_log1pc_lo));
}
/* src/float64x2/hyp.inc:2-18 (sinh) */
static dd o_sinh(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!isfinite(x.hi)) return mkdd(sinh(x.hi), 0.0);
  if (fabs(x.hi) < 0.5) return o_mul(x, o_estrin12(o_mul(x, x), mf_sinhc_hi, mf_sinhc_lo));
  return o_mul(o_sub(o_exp(x), o_exp(o_neg(x))), mkdd(0.5, 0.0));
}
/* hyp.inc:20-29 (cosh) */
static dd o_cosh(dd x) {
  if (!isfinite(x.hi)) return mkdd(cosh(x.hi), 0.0);
  return o_mul(o_add(o_exp(x), o_exp(o_neg(x))), mkdd(0.5, 0.0));
}
/* hyp.inc:62-87 (tanh) */
static dd o_tanh(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!isfinite(x.hi)) return mkdd(tanh(x.hi), 0.0);
  if (x.hi >= 38.0) return mkdd(1.0, 0.0);
  if (x.hi <= -38.0) return mkdd(-1.0, 0.0);
  if (fabs(x.hi) < 0.5) return o_div(o_sinh(x), o_cosh(x));
  int sign = x.hi < 0.0;
  dd m2 = sign ? mkdd(2.0 * x.hi, 2.0 * x.lo) : mkdd(-2.0 * x.hi, -2.0 * x.lo);
  dd em2 = o_exp(m2);
  dd res = o_div(o_sub(mkdd(1.0, 0.0), em2), o_add(mkdd(1.0, 0.0), em2));
  return sign ? o_neg(res) : res;
}
ARR1(or_exp, o_exp)
ARR1(or_log, o_log)
ARR1(or_sin, o_sin)
ARR1(or_cos, o_cos)
ARR1(or_tan, o_tan)
ARR1(or_exp2, o_exp2)
ARR1(or_expm1, o_expm1)
ARR1(or_log2, o_log2)
ARR1(or_log10, o_log10)
ARR1(or_log1p, o_log1p)
ARR1(or_sinh, o_sinh)
ARR1(or_cosh, o_cosh)
ARR1(or_tanh, o_tanh)

/* src/float64x2/inv_trig.inc:226-254 (asinh) */
static dd o_asinh(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!isfinite(x.hi)) return mkdd(asinh(x.hi), 0.0);
  if (fabs(x.hi) < 0.01) return o_mul(x, o_estrin14(o_mul(x, x), mf_asinhc_hi, mf_asinhc_lo));
  int sign = x.hi < 0.0;
  dd ax = sign ? o_neg(x) : x;
  if (ax.hi > 1e150) {
    dd r = o_add(o_log(ax), mkdd(mf_ln2_hi, mf_ln2_lo));
    return sign ? o_neg(r) : r;
  }
  dd axsq = o_mul(ax, ax);
  dd root = o_sqrt(o_add(axsq, mkdd(1.0, 0.0)));
  dd res = ax.hi < 1.0 ? o_log1p(o_add(ax, o_div(axsq, o_add(root, mkdd(1.0, 0.0)))))
                       : o_log(o_add(ax, root));
  return sign ? o_neg(res) : res;
}
/* inv_trig.inc:256-268 (acosh) */
static dd o_acosh(dd x) {
  if (!isfinite(x.hi)) return mkdd(acosh(x.hi), 0.0);
  if (x.hi < 1.0) return mkdd(NAN, 0.0);
  if (x.hi > 1e150) return o_add(o_log(x), mkdd(mf_ln2_hi, mf_ln2_lo));
  return o_log(o_add(x, o_sqrt(o_sub(o_mul(x, x), mkdd(1.0, 0.0)))));
}
/* inv_trig.inc:270-291 (atanh) */
static dd o_atanh(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  if (!isfinite(x.hi)) return mkdd(atanh(x.hi), 0.0);
  if (fabs(x.hi) > 1.0) return mkdd(NAN, 0.0);
  if (fabs(x.hi) < 0.01) return o_mul(x, o_estrin14(o_mul(x, x), mf_atanhc_hi, mf_atanhc_lo));
  dd arg = o_div(o_add(x, x), o_sub(mkdd(1.0, 0.0), x));
  return o_mul(mkdd(0.5, 0.0), o_log1p(arg));
}
ARR1(or_asinh, o_asinh)
ARR1(or_acosh, o_acosh)
ARR1(or_atanh, o_atanh)

/* src/multifloats_td.hh:30-35 (tsum3) */
static void o_tsum3(double *T0, double *T1, double *T2, double x) {
  double e;
  tsum(*T0, x, T0, &e);
  tsum(*T1, e, T1, &e);
  *T2 += e;
}
/* src/float64x2/exp_log.inc:86-106 (exp2_from_triple), with tsum3_finalize
   of multifloats_td.hh:39-43 */
static dd o_exp2_from_triple(double h, double l, double e) {
  double n_d = nearbyint(h), rh, rt, u;
  tsum(h, -n_d, &rh, &rt);
  double T0 = rh, T1 = 0.0, T2 = 0.0;
  o_tsum3(&T0, &T1, &T2, rt);
  o_tsum3(&T0, &T1, &T2, l);
  o_tsum3(&T0, &T1, &T2, e);
  tsum(T1, T2, &T1, &u);
  T2 = u;
  tsum(T0, T1, &T0, &u);
  T1 = u;
  dd res = o_exp2_reduced(n_d, T0, T1);
  if (!isfinite(res.hi)) return res;
  return o_add(res, mkdd(res.hi * (mf_ln2_hi * T2), 0.0));
}
/* exp_log.inc:364-397 (exp10) */
static dd o_exp10(dd x) {
  if (!isfinite(x.hi)) return mkdd(x.hi > 0.0 ? x.hi : (x.hi < 0.0 ? 0.0 : x.hi), 0.0);
  if (x.hi > mf_exp2_max * mf_log10_2_hi) return mkdd(INFINITY, 0.0);
  if (x.hi < mf_exp2_min * mf_log10_2_hi) return mkdd(0.0, 0.0);
  double P1, P1e, P2, P2e, P3, P3e;
  tprod(x.hi, mf_log2_10_hi, &P1, &P1e);
  tprod(x.hi, mf_log2_10_lo, &P2, &P2e);
  tprod(x.lo, mf_log2_10_hi, &P3, &P3e);
  double PX = x.hi * mf_log2_10_lo2 + x.lo * mf_log2_10_lo;
  double h = P1, l = 0.0, c = 0.0, t, u, e;
  tsum(h, P3, &h, &t);
  tsum(l, t, &l, &u); c += u;
  tsum(l, P1e, &l, &u); c += u;
  tsum(l, P2, &l, &u); c += u;
  tsum(l, P2e, &l, &u); c += u;
  tsum(l, P3e, &l, &u); c += u;
  tsum(l, PX, &l, &u); c += u;
  tsum(l, c, &l, &e);
  ftsum(h, l, &h, &l);
  return o_exp2_from_triple(h, l, e);
}
/* exp_log.inc:479-493 (exp2m1) */
static dd o_exp2m1(dd x) {
  if (!isfinite(x.hi)) return o_expm1(o_mul(x, mkdd(mf_ln2_hi, mf_ln2_lo)));
  if (fabs(x.hi) >= 0.5) return o_sub(o_exp2(x), mkdd(1.0, 0.0));
  return o_expm1(o_mul(x, mkdd(mf_ln2_hi, mf_ln2_lo)));
}
/* exp_log.inc:495-506 (exp10m1) */
static dd o_exp10m1(dd x) {
  if (!isfinite(x.hi)) return o_expm1(o_mul(x, mkdd(mf_ln10_hi, mf_ln10_lo)));
  if (fabs(x.hi) >= 0.5) return o_sub(o_exp10(x), mkdd(1.0, 0.0));
  return o_expm1(o_mul(x, mkdd(mf_ln10_hi, mf_ln10_lo)));
}
/* exp_log.inc:508-519 (log2p1, log10p1) */
static dd o_log2p1(dd x) { return o_mul(o_log1p(x), mkdd(mf_log2e_hi, mf_log2e_lo)); }
static dd o_log10p1(dd x) { return o_mul(o_log1p(x), mkdd(mf_inv_ln10_hi, mf_inv_ln10_lo)); }
ARR1(or_exp10, o_exp10)
ARR1(or_exp2m1, o_exp2m1)
ARR1(or_exp10m1, o_exp10m1)
ARR1(or_log2p1, o_log2p1)
ARR1(or_log10p1, o_log10p1)

/* trig.inc:192-215 (sincos_eval) */
static void o_sincos_eval(dd r, dd *s, dd *c) {
  if (fabs(r.hi) <= 0.392699081698724) {
    *s = o_sink(r);
    *c = o_cosk(r);
    return;
  }
  dd pi4 = mkdd(mf_pi_quarter_hi, mf_pi_quarter_lo), is2 = mkdd(mf_inv_sqrt2_hi, mf_inv_sqrt2_lo);
  int pos = r.hi > 0.0;
  dd rp = pos ? o_sub(r, pi4) : o_add(r, pi4);
  dd sk = o_sink(rp), ck = o_cosk(rp);
  *s = o_mul(pos ? o_add(sk, ck) : o_sub(sk, ck), is2);
  *c = o_mul(pos ? o_sub(ck, sk) : o_add(ck, sk), is2);
}
/* trig.inc:41-61 (sinpi) */
static dd o_sinpi(dd x) {
  if (!isfinite(x.hi)) return mkdd(NAN, 0.0);
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  int sign = x.hi < 0.0;
  dd ax = sign ? o_neg(x) : x;
  if (!(ax.hi < 0x1p52)) return mkdd(NAN, 0.0);
  double n = nearbyint(2.0 * ax.hi);
  dd pr = o_mul(mkdd(mf_pi_hi, mf_pi_lo), o_add(ax, mkdd(-0.5 * n, 0.0)));
  dd res;
  switch ((long long)n & 3LL) {
    case 0: res = o_eval(pr, 0); break;
    case 1: res = o_eval(pr, 1); break;
    case 2: res = o_neg(o_eval(pr, 0)); break;
    default: res = o_neg(o_eval(pr, 1)); break;
  }
  return sign ? o_neg(res) : res;
}
/* trig.inc:63-79 (cospi) */
static dd o_cospi(dd x) {
  if (!isfinite(x.hi)) return mkdd(NAN, 0.0);
  dd ax = x.hi < 0.0 ? o_neg(x) : x;
  if (!(ax.hi < 0x1p52)) return mkdd(NAN, 0.0);
  double n = nearbyint(2.0 * ax.hi);
  dd pr = o_mul(mkdd(mf_pi_hi, mf_pi_lo), o_add(ax, mkdd(-0.5 * n, 0.0)));
  switch ((long long)n & 3LL) {
    case 0: return o_eval(pr, 1);
    case 1: return o_neg(o_eval(pr, 0));
    case 2: return o_neg(o_eval(pr, 1));
    default: return o_eval(pr, 0);
  }
}
/* trig.inc:82-104 (tanpi) */
static dd o_tanpi(dd x) {
  if (!isfinite(x.hi)) return mkdd(NAN, 0.0);
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  int sign = x.hi < 0.0;
  dd ax = sign ? o_neg(x) : x;
  if (!(ax.hi < 0x1p52)) return mkdd(NAN, 0.0);
  double n = nearbyint(2.0 * ax.hi);
  dd pr = o_mul(mkdd(mf_pi_hi, mf_pi_lo), o_add(ax, mkdd(-0.5 * n, 0.0)));
  dd s, c;
  o_sincos_eval(pr, &s, &c);
  dd res = ((long long)n & 1LL) ? o_div(o_neg(c), s) : o_div(s, c);
  return sign ? o_neg(res) : res;
}
ARR1(or_sinpi, o_sinpi)
ARR1(or_cospi, o_cospi)
ARR1(or_tanpi, o_tanpi)

/* include/multifloats/float64x2.h:171-175 (operator<) */
static int o_lt(dd a, dd b) {
  if (a.hi < b.hi) return 1;
  if (a.hi == b.hi) return a.lo < b.lo;
  return 0;
}
/* float64x2.h:1449-1455 (signbit) with first_nonzero_limb_index 1276-1280 */
static int o_signbit(dd x) {
  int i = x.hi != 0.0 ? 0 : (x.lo != 0.0 ? 1 : 2);
  return signbit(i == 1 ? x.lo : x.hi) != 0;
}
/* float64x2.h:1299-1309 (fmin, fmax) */
static dd o_fmin(dd a, dd b) { return (o_lt(a, b) || b.hi != b.hi) ? a : b; }
static dd o_fmax(dd a, dd b) { return (o_lt(a, b) || a.hi != a.hi) ? b : a; }
/* float64x2.h:1518-1520 (copysign) */
static dd o_copysign(dd x, dd y) { return o_signbit(x) == o_signbit(y) ? x : o_neg(x); }
/* float64x2.h:2130-2136 (fdim) */
static dd o_fdim(dd x, dd y) {
  if (x.hi != x.hi || y.hi != y.hi) return mkdd(NAN, 0.0);
  return o_lt(y, x) ? o_sub(x, y) : mkdd(0.0, 0.0);
}
/* float64x2.h:2270-2303 (hypot) */
static dd o_hypot(dd x, dd y) {
  if (!isfinite(x.hi) || !isfinite(y.hi)) return mkdd(hypot(x.hi, y.hi), 0.0);
  dd ax = o_signbit(x) ? o_neg(x) : x, ay = o_signbit(y) ? o_neg(y) : y;
  dd big = o_lt(ay, ax) ? ax : ay, small = o_lt(ay, ax) ? ay : ax;
  if (big.hi == 0.0) return mkdd(0.0, 0.0);
  int e = ilogb(big.hi);
  double down = ldexp(1.0, -e), up = ldexp(1.0, e);
  big.hi *= down; big.lo *= down;
  small.hi *= down; small.lo *= down;
  dd ratio = o_div(small, big);
  dd root = o_mul(big, o_sqrt(o_add(mkdd(1.0, 0.0), o_mul(ratio, ratio))));
  dd r = mkdd(root.hi * up, root.lo * up);
  if (!isfinite(r.hi)) r.lo = 0.0;
  return r;
}
/* float64x2.h:2077-2102 (powi) */
static dd o_powi(dd base, int n) {
  if (n == 0) return mkdd(1.0, 0.0);
  if (base.hi == 0.0 && base.lo == 0.0) {
    double sg = (n & 1) ? copysign(1.0, base.hi) : 1.0;
    return mkdd(sg * (n > 0 ? 0.0 : INFINITY), 0.0);
  }
  int neg = n < 0;
  unsigned long long u = neg ? (unsigned long long)(-(long long)n) : (unsigned long long)n;
  dd result = mkdd(1.0, 0.0), b = base;
  for (;;) {
    if (u & 1ull) result = o_mul(result, b);
    u >>= 1;
    if (u == 0ull) break;
    b = o_mul(b, b);
  }
  return neg ? o_div(mkdd(1.0, 0.0), result) : result;
}
/* exp_log.inc:531-600 (pow) */
static dd o_pow(dd x, dd y) {
  if (x.hi == 0.0 && y.hi == 0.0) return mkdd(1.0, 0.0);
  if (x.hi == 1.0 && x.lo == 0.0) return mkdd(1.0, 0.0);
  if (x.hi == -1.0 && x.lo == 0.0 && isinf(y.hi)) return mkdd(1.0, 0.0);
  if (y.lo == 0.0 && isfinite(y.hi) && y.hi == trunc(y.hi) && fabs(y.hi) <= 2147483647.0)
    return o_powi(x, (int)y.hi);
  if (!isfinite(x.hi) || !isfinite(y.hi) || x.hi <= 0.0) return o_exp(o_mul(y, o_log(x)));
  int e_x = ilogb(x.hi);
  dd m = mkdd(ldexp(x.hi, -e_x), ldexp(x.lo, -e_x));
  dd Lf = o_log2_kernel(m);
  double Lx = 0.0, mh = m.hi;
  if (!(mh > 15.0 / 16.0 && mh < 17.0 / 16.0)) {
    int idx = (int)((mh - 1.0) * 32.0);
    if (idx < 0) idx = 0;
    if (idx > 31) idx = 31;
    Lx = mf_log2v_x[idx];
  }
  double ed = (double)e_x, AH, AM, ALh, ALl, P1, P1e, P2, P2e, P3, P3e;
  tprod(y.hi, ed, &AH, &AM);
  tprod(y.lo, ed, &ALh, &ALl);
  tprod(y.hi, Lf.hi, &P1, &P1e);
  tprod(y.hi, Lf.lo, &P2, &P2e);
  tprod(y.lo, Lf.hi, &P3, &P3e);
  double PX = y.hi * Lx + y.lo * Lf.lo;
  double h = AH, l = 0.0, c = 0.0, t, u, e;
  tsum(h, P1, &h, &t); tsum(l, t, &l, &u); c += u;
  tsum(h, AM, &h, &t); tsum(l, t, &l, &u); c += u;
  tsum(h, ALh, &h, &t); tsum(l, t, &l, &u); c += u;
  tsum(h, P3, &h, &t); tsum(l, t, &l, &u); c += u;
  tsum(l, P1e, &l, &u); c += u;
  tsum(l, P2, &l, &u); c += u;
  tsum(l, P2e, &l, &u); c += u;
  tsum(l, P3e, &l, &u); c += u;
  tsum(l, PX, &l, &u); c += u;
  c += ALl;
  tsum(l, c, &l, &e);
  ftsum(h, l, &h, &l);
  return o_exp2_from_triple(h, l, e);
}
/* float64x2.h:1540-1555 (floor) and 1574-1576 (trunc) */
static dd o_floor(dd x) {
  if (x.hi == 0.0 && x.lo == 0.0) return x;
  double f = floor(x.hi);
  if (f == x.hi) {
    double h, l;
    ftsum(f, floor(x.lo), &h, &l);
    return mkdd(h, l);
  }
  return mkdd(f, 0.0);
}
static dd o_trunc(dd x) { return signbit(x.hi) ? o_neg(o_floor(o_neg(x))) : o_floor(x); }
/* float64x2.h:186-190 (operator>=) */
static int o_ge(dd a, dd b) {
  if (b.hi < a.hi) return 1;
  if (a.hi == b.hi) return a.lo >= b.lo;
  return 0;
}
/* std::ilogb as glibc defines it on x86-64: FP_ILOGB0 = FP_ILOGBNAN =
 * INT_MIN, INT_MAX for infinities, the true exponent of a subnormal. */
static int ref_ilogb(double v) {
  if (v == 0.0 || v != v) return INT_MIN;
  if (isinf(v)) return INT_MAX;
  int e;
  (void)frexp(v, &e);
  return e - 1;
}
/* One step of the reference's fmod loop (float64x2.h:2028-2054) on the
 * remainder r (ay = |y|): returns 0 at the loop's exit (r then holds the
 * result's magnitude), 1 after a step. The exponent gap is the wrapping
 * 32-bit difference, as the compiled loop computes it (ilogb(0) = INT_MIN:
 * y = -0 gives the gap INT_MIN + ilogb(r)). */
static int o_fmod_step(dd *rp, dd ay) {
  dd r = *rp;
  if (r.hi < 0.0) {
    double r_abs = -r.hi;
    int gap = (int)((unsigned)ref_ilogb(r_abs) - (unsigned)ref_ilogb(ay.hi));
    if (gap <= 0) {
      r = o_add(r, ay);
    } else if (gap <= 53) {
      double q = trunc(r_abs / ay.hi) + 1.0;
      r = o_add(r, o_mul(ay, mkdd(q, 0.0)));
    } else {
      r = o_add(r, o_mul(o_add(o_trunc(o_div(o_neg(r), ay)), mkdd(1.0, 0.0)), ay));
    }
  } else if (o_ge(r, ay)) {
    int gap = (int)((unsigned)ref_ilogb(r.hi) - (unsigned)ref_ilogb(ay.hi));
    if (gap <= 53) {
      double q = trunc(r.hi / ay.hi);
      r = (q <= 1.0) ? o_sub(r, ay) : o_sub(r, o_mul(ay, mkdd(q, 0.0)));
    } else {
      r = o_sub(r, o_mul(o_trunc(o_div(r, ay)), ay));
    }
  } else {
    return 0;
  }
  *rp = r;
  return !(r.hi == 0.0 && r.lo == 0.0);
}
static int same_bits(dd a, dd b) { return memcmp(&a, &b, sizeof(dd)) == 0; }
/* float64x2.h:2010-2059 (fmod), bounded. The reference's loop is a function
 * of r alone (ay is fixed), so it returns iff its orbit reaches an exit; an
 * orbit that revisits a remainder never returns. Brent's cycle search
 * (remainder saved at step counts 1, 2, 4, ...) finds such a revisit within
 * twice the orbit's length, and FMOD_MAX_STEPS bounds the search: both give
 * NaN. Every returning call restates the reference's bits. *steps (if set)
 * receives the steps taken (-1: cycle, -2: cap). */
#define FMOD_MAX_STEPS 512
static dd o_fmod_n(dd x, dd y, int *steps) {
  int x_neg = x.hi < 0.0;
  dd ax = x_neg ? o_neg(x) : x;
  dd ay = (y.hi < 0.0) ? o_neg(y) : y;
  if (steps) *steps = 0;
  if (o_lt(ax, ay)) return x;
  dd r = ax, saved = ax;
  int n = 0, power = 1, since = 0;
  for (;;) {
    if (!o_fmod_step(&r, ay)) break;
    ++n;
    if (same_bits(r, saved)) {
      if (steps) *steps = -1;
      return mkdd(NAN, 0.0);
    }
    if (n >= FMOD_MAX_STEPS) {
      if (steps) *steps = -2;
      return mkdd(NAN, 0.0);
    }
    if (++since == power) {
      saved = r;
      power *= 2;
      since = 0;
    }
  }
  if (steps) *steps = n;
  return x_neg ? o_neg(r) : r;
}
static dd o_fmod(dd x, dd y) { return o_fmod_n(x, y, NULL); }
/* float64x2.h:2068-2072 (modulo) */
static dd o_modulo(dd x, dd y) {
  dd r = o_fmod(x, y);
  if (r.hi == 0.0 && r.lo == 0.0) return r;
  return (o_signbit(r) != o_signbit(y)) ? o_add(r, y) : r;
}
/* fsrc/multifloats.fypp:1595-1607 (nearest): the next double after x.hi in
 * the direction of +inf when y.hi > 0, else of -inf (gfortran's NEAREST is
 * nextafter toward the infinity of that sign); low limb 0. */
static dd o_nearest(dd x, dd y) { return mkdd(nextafter(x.hi, y.hi > 0.0 ? INFINITY : -INFINITY), 0.0); }
void or_fmod_steps(const dd *a, const dd *b, int *steps, int64_t n) {
  for (int64_t i = 0; i < n; ++i) (void)o_fmod_n(a[i], b[i], &steps[i]);
}
ARR2(or_fmod, o_fmod)
ARR2(or_modulo, o_modulo)
ARR2(or_nearest, o_nearest)
ARR2(or_pow, o_pow)
ARR2(or_hypot, o_hypot)
ARR2(or_fdim, o_fdim)
ARR2(or_fmin, o_fmin)
ARR2(or_fmax, o_fmax)
ARR2(or_copysign, o_copysign)

/* ------------------------------------------------- exact accumulator ---
 * A fixed-point superaccumulator over 2^-1216 .. 2^1120 in 32-bit digits held
 * in int64 (carry-save), enough for every product of two finite doubles and
 * for sums of up to 2^30 of them. Exact: adding a double splits its 53-bit
 * significand over at most three digits. Used only to measure errors. */
#define XA_DIGITS 74
#define XA_BIAS 1216 /* bit position of 2^0 */
typedef struct { int64_t d[XA_DIGITS]; int count; } xacc;

static void xa_norm(xacc *a) {
  for (int i = 0; i < XA_DIGITS - 1; ++i) {
    int64_t c = a->d[i] >> 32; /* arithmetic shift: floor division */
    a->d[i] -= c * ((int64_t)1 << 32);
    a->d[i + 1] += c;
  }
  a->count = 0;
}
static void xa_add(xacc *a, double x) {
  if (x == 0.0) return;
  int e;
  double m = frexp(x, &e); /* x = m 2^e, 0.5 <= |m| < 1 */
  int64_t im = (int64_t)ldexp(m, 53); /* exact, |im| < 2^53 */
  int pos = e - 53 + XA_BIAS;         /* bit position of im's lsb */
  if (pos < 0) { /* only reachable by subnormal inputs below 2^-1216: impossible for doubles */
    im >>= -pos; pos = 0;
  }
  int di = pos >> 5, sh = pos & 31;
  /* im * 2^sh spans < 85 bits: split into three 32-bit digits */
  int neg = im < 0;
  uint64_t u = neg ? (uint64_t)(-im) : (uint64_t)im;
  unsigned __int128 v = (unsigned __int128)u << sh;
  int64_t d0 = (int64_t)(uint32_t)(v), d1 = (int64_t)(uint32_t)(v >> 32), d2 = (int64_t)(uint32_t)(v >> 64);
  if (neg) { d0 = -d0; d1 = -d1; d2 = -d2; }
  a->d[di] += d0;
  a->d[di + 1] += d1;
  a->d[di + 2] += d2;
  if (++a->count >= (1 << 29)) xa_norm(a);
}
/* Round the accumulator to the nearest double (correct to within one ulp,
 * which is all an error measurement needs). */
static double xa_value(xacc *a) {
  xa_norm(a);
  /* after normalisation the top digit carries the sign: work on |value| so
   * that every digit is in [0, 2^32), then sum from the top down */
  int neg = a->d[XA_DIGITS - 1] < 0;
  if (neg) {
    for (int i = 0; i < XA_DIGITS; ++i) a->d[i] = -a->d[i];
    xa_norm(a);
  }
  double r = 0.0;
  int top = XA_DIGITS - 1;
  while (top > 0 && a->d[top] == 0) --top;
  for (int i = top; i >= 0 && i >= top - 3; --i) r += ldexp((double)a->d[i], 32 * i - XA_BIAS);
  return neg ? -r : r;
}
static void xa_add_dd_prod(xacc *a, dd x, dd y) {
  double p, e;
  tprod(x.hi, y.hi, &p, &e); xa_add(a, p); xa_add(a, e);
  tprod(x.hi, y.lo, &p, &e); xa_add(a, p); xa_add(a, e);
  tprod(x.lo, y.hi, &p, &e); xa_add(a, p); xa_add(a, e);
  tprod(x.lo, y.lo, &p, &e); xa_add(a, p); xa_add(a, e);
}

/* err = |sum_p a[p*sa] b[p*sb] - c| exactly (then rounded); mag = sum |a||b|. */
void or_exact_dot_err(const dd *a, int64_t sa, const dd *b, int64_t sb, int64_t k, const dd *c,
                      double *err, double *mag) {
  xacc acc;
  memset(&acc, 0, sizeof acc);
  double m = 0.0;
  for (int64_t p = 0; p < k; ++p) {
    xa_add_dd_prod(&acc, a[p * sa], b[p * sb]);
    m += fabs(a[p * sa].hi * b[p * sb].hi);
  }
  xa_add(&acc, -c->hi);
  xa_add(&acc, -c->lo);
  *err = fabs(xa_value(&acc));
  *mag = m;
}
/* Exact value of sum_p a_p b_p rounded to a double-double (hi = RN(sum)). */
void or_exact_dot(const dd *a, int64_t sa, const dd *b, int64_t sb, int64_t k, dd *out) {
  xacc acc;
  memset(&acc, 0, sizeof acc);
  for (int64_t p = 0; p < k; ++p) xa_add_dd_prod(&acc, a[p * sa], b[p * sb]);
  xacc acc2 = acc;
  double h = xa_value(&acc);
  xa_add(&acc2, -h);
  out->hi = h;
  out->lo = xa_value(&acc2);
}
/* err = |sum_i x_i - c|, mag = sum |x_i| */
void or_exact_sum_err(const dd *x, int64_t n, const dd *c, double *err, double *mag) {
  xacc acc;
  memset(&acc, 0, sizeof acc);
  double m = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    xa_add(&acc, x[i].hi);
    xa_add(&acc, x[i].lo);
    m += fabs(x[i].hi);
  }
  xa_add(&acc, -c->hi);
  xa_add(&acc, -c->lo);
  *err = fabs(xa_value(&acc));
  *mag = m;
}
/* GEMM error on a sampled set of outputs: for each listed (i,j), the exact
 * error of c(i,j) and sum_p |a_ip||b_pj|. Column-major, lda = m, ldb = k. */
void or_gemm_err(const dd *a, const dd *b, const dd *c, int64_t m, int64_t k, int64_t n,
                 const int64_t *rows, const int64_t *cols, int64_t npts, double *err, double *mag) {
  (void)n;
  for (int64_t t = 0; t < npts; ++t) {
    int64_t i = rows[t], j = cols[t];
    or_exact_dot_err(a + i, m, b + j * k, 1, k, &c[i + j * m], &err[t], &mag[t]);
  }
}
