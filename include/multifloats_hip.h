/* multifloats_hip -- MI355X (gfx950) kernels behind the multifloats C ABI.
 *
 * Drop-in boundary for the reference's data-parallel hot path. Paths in the
 * citations are relative to the reference (kyungminlee/multifloats 0.10.3).
 *
 *  1. Reference-compatible entry points (identical names, signatures and
 *     semantics): matmuldd_mm / _mv / _vm and their matmulf64x2_* spellings.
 *     Host pointers, synchronous, caller owns every buffer, C overwritten,
 *     column-major with leading dimension = first extent.
 *  2. Array entry points that replace the reference's per-element and
 *     Fortran-native loops (the reference has no array-level C ABI for these;
 *     its Fortran generics are the interface they stand in for). Host
 *     pointers, synchronous.
 *  3. A device-pointer API (mfhip_*) taking a HIP stream, returning a status,
 *     never allocating and never synchronising: for callers that keep data
 *     resident in HBM, chain operations, or capture them in a HIP graph.
 *
 * Error model. The reference is NaN-in/NaN-out with no status channel
 * (README.md:274-296). The numerical results here keep that model exactly.
 * A device-side failure (no GPU, launch or copy error) cannot be reported
 * through a void reference signature, so entry points of groups 1 and 2
 * print the HIP error to stderr, fill their outputs with NaN and record the
 * error for mfhip_last_error(); there is no silent CPU fallback.
 *
 * Thread safety: every host thread gets its own stream and staging buffers,
 * so the group-1/2 entry points may be called concurrently (the reference's
 * Fortran interfaces are `pure`, fsrc/multifloats.fypp:135-152).
 */
#ifndef MULTIFLOATS_HIP_H
#define MULTIFLOATS_HIP_H

#include <stddef.h>
#include <stdint.h>

#define MULTIFLOATS_HIP_VERSION 1

#ifndef MULTIFLOATS_ABI_VERSION
/* Same ABI version and POD layout as include/multifloats/float64x2.h:33,
 * 125-139, 348-351: two back-to-back IEEE binary64 limbs (hi, lo); complex
 * is re then im. When the reference header is included first, its types are
 * used unchanged. */
#define MULTIFLOATS_ABI_VERSION 2
#if defined(__GNUC__) || defined(__clang__)
#define MULTIFLOATS_API __attribute__((visibility("default")))
#else
#define MULTIFLOATS_API
#endif
#ifdef __cplusplus
namespace multifloats {
#endif
struct float64x2 {
  double limbs[2];
};
struct complex64x2 {
  struct float64x2 re;
  struct float64x2 im;
};
#ifndef __cplusplus
typedef struct float64x2 float64x2;
typedef struct complex64x2 complex64x2;
#endif
#ifdef __cplusplus
} /* namespace multifloats */
#endif
#endif /* MULTIFLOATS_ABI_VERSION */

#ifdef __cplusplus
namespace multifloats {
extern "C" {
#endif

/* ---- 1. reference-compatible GEMM entry points ------------------------- */

/* C(m,n) = A(m,k) * B(k,n). Replaces matmuldd_mm
 * (include/multifloats/float64x2.h:682-685, src/float64x2/abi.inc:133-137).
 * Bit-identical to the reference for every input: each output is the same
 * sequential compensated chain (src/float64x2/matmul.inc:6-39, 132-205).
 * renorm_interval: > 0 renormalises every that many MACs (Fortran default 8,
 * fsrc/multifloats.fypp:664); <= 0 only at the end. k = 0 writes zeros. */
MULTIFLOATS_API void matmuldd_mm(const float64x2 *a, const float64x2 *b, float64x2 *c,
                                 int64_t m, int64_t k, int64_t n, int64_t renorm_interval);
/* y(m) = A(m,k) * x(k). Replaces matmuldd_mv (float64x2.h:687-690). */
MULTIFLOATS_API void matmuldd_mv(const float64x2 *a, const float64x2 *x, float64x2 *y,
                                 int64_t m, int64_t k, int64_t renorm_interval);
/* y(n) = x(k) * B(k,n). Replaces matmuldd_vm (float64x2.h:692-695). */
MULTIFLOATS_API void matmuldd_vm(const float64x2 *x, const float64x2 *b, float64x2 *y,
                                 int64_t k, int64_t n, int64_t renorm_interval);
/* f64x2-tagged spellings (float64x2.h:1042-1046, src/float64x2/f64x2_aliases.inc:114-116). */
MULTIFLOATS_API void matmulf64x2_mm(const float64x2 *a, const float64x2 *b, float64x2 *c,
                                    int64_t m, int64_t k, int64_t n, int64_t renorm_interval);
MULTIFLOATS_API void matmulf64x2_mv(const float64x2 *a, const float64x2 *x, float64x2 *y,
                                    int64_t m, int64_t k, int64_t renorm_interval);
MULTIFLOATS_API void matmulf64x2_vm(const float64x2 *x, const float64x2 *b, float64x2 *y,
                                    int64_t k, int64_t n, int64_t renorm_interval);

/* Complex C(m,n) = A(m,k) * B(k,n) on complex64x2 (re then im), the Fortran
 * matmul on cmplx64x2 (cdd_matmul_mm, fsrc/multifloats.fypp:2609-2621, which
 * has no C symbol in the reference). Bit-identical to that composition:
 * re = mm(a_re,b_re) - mm(a_im,b_im), im = mm(a_re,b_im) + mm(a_im,b_re),
 * each mm a reference-order chain with the given renorm interval. */
MULTIFLOATS_API void cmatmuldd_mm(const complex64x2 *a, const complex64x2 *b, complex64x2 *c,
                                  int64_t m, int64_t k, int64_t n, int64_t renorm_interval);
/* cdd_matmul_mv / cdd_matmul_vm (fypp:2623-2649). */
MULTIFLOATS_API void cmatmuldd_mv(const complex64x2 *a, const complex64x2 *x, complex64x2 *y,
                                  int64_t m, int64_t k, int64_t renorm_interval);
MULTIFLOATS_API void cmatmuldd_vm(const complex64x2 *x, const complex64x2 *b, complex64x2 *y,
                                  int64_t k, int64_t n, int64_t renorm_interval);

/* ---- 2. array entry points -------------------------------------------- */

/* c[i] = a[i] op b[i], bit-identical to the reference operators
 * (float64x2.h:203-311; Fortran native fsrc/multifloats.fypp:2746-2837). */
MULTIFLOATS_API void adddd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void subdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void muldd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void divdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
/* c[i] = f(a[i], b[i]) for the reference's powdd, hypotdd, fdimdd, fmindd,
 * fmaxdd, copysigndd (float64x2.h:421-446; src/float64x2/exp_log.inc:531-600,
 * include/multifloats/float64x2.h:1299-1309, 1518-1520, 2077-2102,
 * 2130-2136, 2270-2303), bit-identical to them element by element. */
MULTIFLOATS_API void powdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void hypotdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void fdimdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void fmindd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void fmaxdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void copysigndd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
/* Array forms of atandd, atanpidd, atan2dd(a = y, b = x), atan2pidd
 * (float64x2.h:489-509; the Fortran atan, atanpi, atan2, atan2pi). Own design
 * for atan (the reference evaluates libquadmath's fitted rationals,
 * src/float64x2/inv_trig.inc:21-55): a table of atan(k/64) in three limbs and
 * a Maclaurin tail, one final rounding; the reference's special cases and
 * atan2 / atanpi wrappers (inv_trig.inc:134-223). No less accurate than the
 * reference on the same inputs. */
MULTIFLOATS_API void atandd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void atanpidd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void atan2dd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void atan2pidd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
/* Array forms of the Fortran generics min, max and dim on real64x2 (the
 * reference's module bodies, fsrc/multifloats.fypp:1539-1552, 1573-1583):
 * min = a < b ? a : b, max = a < b ? b : a (the module's lexicographic `<`),
 * dim = a - b when either is NaN, else a > b ? a - b : 0. Bit-identical to
 * them; they differ from fmindd / fmaxdd / fdimdd only for NaN operands. */
MULTIFLOATS_API void mindd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void maxdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void dimdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
/* Array forms of fmoddd and modulodd (float64x2.h:434-440, 2010-2072; the
 * Fortran mod and modulo, fsrc/multifloats.fypp:1562-1570): the reference's
 * reduction loop step for step, so every call on which that loop returns
 * gives its bits. The reference's loop never returns for some operands (a
 * divisor of -0, subnormal or tiny divisors of large dividends, quotients
 * beyond the double range: its remainder comes back to a value it held);
 * these give NaN here, found by a cycle search over the loop's states within
 * at most 512 steps (returning calls take a few). */
MULTIFLOATS_API void fmoddd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
MULTIFLOATS_API void modulodd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
/* Array form of the Fortran nearest(a, b) on real64x2 (the module's body,
 * fsrc/multifloats.fypp:1595-1607): the next double after a.hi toward +inf
 * when b.hi > 0, else toward -inf (gfortran's NEAREST, nextafter toward that
 * infinity), low limb 0. */
MULTIFLOATS_API void nearestdd_v(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n);
/* Array forms of asindd, acosdd, asinpidd, acospidd (float64x2.h:485-505;
 * the Fortran asin, acos, asinpi, acospi). Own design (the reference
 * evaluates libquadmath's fitted rationals, src/float64x2/inv_trig.inc:56-130):
 * sqrt(1 - x^2) in three limbs, then the atan reduction of x / s or s / x;
 * the reference's special cases and pi-scaled wrappers (inv_trig.inc:197-205).
 * No less accurate than the reference on the same inputs. */
MULTIFLOATS_API void asindd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void acosdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void asinpidd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void acospidd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* c[i] = sqrt(a[i]), bit-identical to sqrtdd (float64x2.h:2190-2213). */
MULTIFLOATS_API void sqrtdd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* Array forms of expdd / logdd / sindd (float64x2.h:453-483; src/float64x2/
 * exp_log.inc:237-341, trig.inc:241-253), bit-identical to them; they replace
 * the Fortran elemental loop over c_expdd/... (fypp:709-736, 1407-1412). */
MULTIFLOATS_API void expdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void logdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void sindd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* Array forms of cosdd, tandd, exp2dd, expm1dd, log2dd, log10dd, log1pdd,
 * sinhdd, coshdd, tanhdd (float64x2.h:455-517; src/float64x2/trig.inc:255-289,
 * exp_log.inc:109-128, 311-353, 407-469, hyp.inc:2-87), bit-identical to them
 * element by element (the Fortran generics cos, tan, expm1, log2, log10,
 * log1p, sinh, cosh, tanh are elemental loops over these, fypp:57). */
MULTIFLOATS_API void cosdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void tandd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void exp2dd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void expm1dd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void log2dd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void log10dd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void log1pdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void sinhdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void coshdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void tanhdd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* Array forms of asinhdd, acoshdd, atanhdd (float64x2.h:519-523; the Fortran
 * asinh, acosh, atanh): the reference's algorithms (src/float64x2/
 * inv_trig.inc:226-291), bit-identical to them element by element. */
MULTIFLOATS_API void asinhdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void acoshdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void atanhdd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* Array forms of exp10dd, exp2m1dd, exp10m1dd, log2p1dd, log10p1dd
 * (float64x2.h:457-475; src/float64x2/exp_log.inc:364-397, 479-519; the
 * Fortran exp10, exp2m1, exp10m1, log2p1, log10p1), bit-identical to them
 * element by element. */
MULTIFLOATS_API void exp10dd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void exp2m1dd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void exp10m1dd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void log2p1dd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void log10p1dd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* Array forms of sinpidd, cospidd, tanpidd (float64x2.h:497-501;
 * src/float64x2/trig.inc:41-104): sin, cos, tan of pi x, bit-identical to
 * them element by element (NaN for |x| >= 2^52, as the reference). */
MULTIFLOATS_API void sinpidd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void cospidd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void tanpidd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* Array forms of erfcdd and erfcxdd (float64x2.h:531-533; the Fortran erfc
 * and erfc_scaled, fypp:57). Own design, not the reference's rational fits
 * (src/float64x2/special.inc:56-188): piecewise Taylor rows of erfc and of
 * erfcx = exp(x^2) erfc(x), erfc = exp(-x^2) erfcx beyond 4. Accuracy
 * against 256-bit MPFR on 100k points per function covering every row edge
 * and switch point: max relative error 5.8e-32 (erfc) and 5.0e-32 (erfcx)
 * where the result is a normal double-double; the reference's erfcdd /
 * erfcxdd on the same points: 8.5e-32 and 8.0e-32. */
MULTIFLOATS_API void erfcdd_v(const float64x2 *a, float64x2 *c, int64_t n);
MULTIFLOATS_API void erfcxdd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* Array form of erfdd (float64x2.h:525; src/float64x2/special.inc:18-51): own
 * piecewise-Taylor algorithm held to the reference's published max relative
 * error 1.55e-32: 8.6e-33 against 256-bit MPFR on 100k points covering every
 * Taylor row, row edge and tiny / huge argument (the reference's erfdd on the
 * same points: 1.95e-32). */
MULTIFLOATS_API void erfdd_v(const float64x2 *a, float64x2 *c, int64_t n);
/* Fortran sum(a) on a rank-1 real64x2 array (fypp:1980-2020). */
MULTIFLOATS_API float64x2 sumdd(const float64x2 *a, int64_t n);
/* Fortran dot_product(a, b) (fypp:2254-2308); renorm_interval plays the role
 * of DD_FMA_RENORM_INTERVAL (fypp:664, 2268). */
MULTIFLOATS_API float64x2 dotdd(const float64x2 *a, const float64x2 *b, int64_t n,
                                int64_t renorm_interval);
/* Fortran norm2(a) (fypp:2338-2352); like the reference, no overflow scaling. */
MULTIFLOATS_API float64x2 norm2dd(const float64x2 *a, int64_t n);
/* Fortran sum(array, mask) (fypp:1980-2020): elements whose mask entry
 * (a default-kind Fortran LOGICAL, i.e. a 4-byte integer, nonzero = true) is
 * false are skipped; mask == NULL is sumdd. */
MULTIFLOATS_API float64x2 sumdd_mask(const float64x2 *a, const int32_t *mask, int64_t n);
/* Fortran dot_product(a, b) on cmplx64x2 = sum(conj(a) * b) (fypp:2321-2331:
 * four real dot chains combined with DD + and -). */
MULTIFLOATS_API complex64x2 cdotdd(const complex64x2 *a, const complex64x2 *b, int64_t n,
                                   int64_t renorm_interval);
/* Fortran sum(array, dim[, mask]) and norm2(array, dim) (fypp:2035-2090,
 * 2356-2377) for a contiguous column-major array viewed as (inner, len,
 * outer), with len the extent of dimension `dim`, inner the product of the
 * extents before it and outer the product of those after it. out receives
 * the inner x outer result in column-major order. Each output is the
 * reference's sequential chain over dimension dim with the DD operators, in
 * the loop order of the fypp source; the results are bit-identical to that
 * restated loop (the oracle's or_sum_dim / or_norm2_dim; the Fortran module
 * itself is not compiled by the tests). mask may be NULL. */
MULTIFLOATS_API void sumdd_dim(const float64x2 *a, const int32_t *mask, int64_t inner, int64_t len,
                               int64_t outer, float64x2 *out);
MULTIFLOATS_API void norm2dd_dim(const float64x2 *a, int64_t inner, int64_t len, int64_t outer,
                                 float64x2 *out);

/* Fortran product / maxval / minval on real64x2 (fypp:1783-1862, 1975-2090):
 *   <f>dd(a, n), <f>dd_mask(a, mask, n) (mask of int32, nonzero = included,
 *   NULL = all), <f>dd_dim(a, mask, inner, len, outer, out) as sumdd_dim.
 * product: rank 1 is a DD-multiply tree from 1 (not the sequential loop's
 * bits; relative error within 4 n 2^-104 for results clear of overflow and
 * underflow); dim = sequential chains, bit-identical to the restated loop
 * (the oracle's or_fibre_reduce_dim). maxval / minval: the
 * reference's scan (first element, replaced by any later x > res / x < res,
 * lexicographic on (hi, lo)), bit-identical in every form including NaN
 * limbs; 0 when no element is selected. */
MULTIFLOATS_API float64x2 productdd(const float64x2 *a, int64_t n);
MULTIFLOATS_API float64x2 productdd_mask(const float64x2 *a, const int32_t *mask, int64_t n);
MULTIFLOATS_API void productdd_dim(const float64x2 *a, const int32_t *mask, int64_t inner, int64_t len,
                                   int64_t outer, float64x2 *out);
MULTIFLOATS_API float64x2 maxvaldd(const float64x2 *a, int64_t n);
MULTIFLOATS_API float64x2 maxvaldd_mask(const float64x2 *a, const int32_t *mask, int64_t n);
MULTIFLOATS_API void maxvaldd_dim(const float64x2 *a, const int32_t *mask, int64_t inner, int64_t len,
                                  int64_t outer, float64x2 *out);
MULTIFLOATS_API float64x2 minvaldd(const float64x2 *a, int64_t n);
MULTIFLOATS_API float64x2 minvaldd_mask(const float64x2 *a, const int32_t *mask, int64_t n);
MULTIFLOATS_API void minvaldd_dim(const float64x2 *a, const int32_t *mask, int64_t inner, int64_t len,
                                  int64_t outer, float64x2 *out);

/* BLAS-style C = alpha op(A) op(B) + beta C on float64x2 (the interface of
 * the reference's xGEMM shim test/integration/blas/wgemm.f90:1-14, the GEMM
 * interface its architecture notes defer, doc/dev/architecture.md:307-310).
 * transa/transb: 'N', 'T' or 'C' (= 'T' for real);
 * column-major with leading dimensions lda, ldb, ldc; argument checks and
 * quick returns as wgemm.f90:16-69 (alpha == 0 or k == 0: C = beta C without
 * reading A, B; beta == 0: C is not read). op(A) op(B) is formed by the
 * reference-order chains of matmuldd_mm with the given renorm interval, then
 * alpha T + beta C with the DD operators. Invalid arguments write NaN to C
 * (as a failed call does) instead of calling xerbla. */
MULTIFLOATS_API void wgemmdd(char transa, char transb, int64_t m, int64_t n, int64_t k, float64x2 alpha,
                             const float64x2 *a, int64_t lda, const float64x2 *b, int64_t ldb,
                             float64x2 beta, float64x2 *c, int64_t ldc, int64_t renorm_interval);

/* Triangular solve op(A) X = alpha B (side 'L') or X op(A) = alpha B (side
 * 'R'), X overwriting B: the reference's xTRSM shim
 * test/integration/blas/wtrsm.f90:1-10, bit-identical to the restated loop
 * nests (the oracle's or_trsm restates its eight loop nests, 57-206, in the
 * same operation order, including the skips of zero pivots and zero
 * coefficients; the Fortran shim itself is not compiled by the tests).
 * uplo 'U'/'L', transa 'N'/'T'/'C', diag 'U' (unit, diagonal not read) or
 * 'N'. Argument checks and quick returns as wtrsm.f90:
 * 24-56 (alpha == 0: B = 0). Invalid arguments write NaN to B. */
MULTIFLOATS_API void wtrsmdd(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float64x2 alpha,
                             const float64x2 *a, int64_t lda, float64x2 *b, int64_t ldb);

/* Precision form of matmuldd_mm (no counterpart in the reference): the Ozaki
 * split-precision GEMM on the INT8 matrix cores. A and B are scaled per row /
 * column by a power of two and represented by `slices` base-256 digits (8
 * bits each). The digit integers are multiplied exactly modulo pairwise-
 * coprime moduli (one INT8 GEMM each) and C is rebuilt by the Chinese
 * remainder theorem (CRT form); with an explicit count and m or n below 512,
 * digit pairs whose levels sum to fewer than `slices` are multiplied instead
 * and recombined in DD (triangular form). Not bit-identical to matmuldd_mm.
 *
 * slices <= 0 (recommended): every output satisfies
 *     |C - AB|_ij <= 2^-100 (|A||B|)_ij
 * for any finite input whose products and results stay in the normal range
 * (2^-968 .. 2^1023), whatever the distribution of magnitudes. A certificate
 * (an FP8 MFMA product of the scaled magnitudes, bounding (|A||B|)_ij from
 * below, against a rigorous bound of the digit representation's error) gives
 * each output the digits it needs; the slice count is then chosen on the
 * device, and the outputs it does not cover (e.g. rows whose largest entries
 * meet small entries of the column) are recomputed by a compensated dot
 * accurate to 2^-102.6 (|A||B|)_ij. Nothing is synchronised with the host.
 * Explicit counts (7..16) skip the certificate: the error then depends on
 * the data (about 2^(-8 slices + 8) (|A||B|)_ij on well-spread inputs).
 * k <= 32768. Rows of A or columns of B holding Inf/NaN give NaN outputs. */
MULTIFLOATS_API void matmuldd_mm_ozaki(const float64x2 *a, const float64x2 *b, float64x2 *c,
                                       int64_t m, int64_t k, int64_t n, int slices);

/* ---- 3. device-pointer API ------------------------------------------- */
/* All pointers below are device pointers (16-byte aligned) and `stream` is a
 * hipStream_t (NULL = the default stream). Returns 0 on success, otherwise a
 * hipError_t value (see mfhip_error_string). Work is only enqueued. */

enum {
  MFHIP_SQRT = 0, MFHIP_EXP = 1, MFHIP_LOG = 2, MFHIP_SIN = 3, MFHIP_ERF = 4, MFHIP_NEG = 5,
  MFHIP_COS = 6, MFHIP_TAN = 7, MFHIP_EXP2 = 8, MFHIP_EXPM1 = 9, MFHIP_LOG2 = 10,
  MFHIP_LOG10 = 11, MFHIP_LOG1P = 12, MFHIP_SINH = 13, MFHIP_COSH = 14, MFHIP_TANH = 15,
  MFHIP_ERFC = 16, MFHIP_ERFCX = 17, MFHIP_ASINH = 18, MFHIP_ACOSH = 19, MFHIP_ATANH = 20,
  MFHIP_EXP10 = 21, MFHIP_EXP2M1 = 22, MFHIP_EXP10M1 = 23, MFHIP_LOG2P1 = 24, MFHIP_LOG10P1 = 25,
  MFHIP_SINPI = 26, MFHIP_COSPI = 27, MFHIP_TANPI = 28, MFHIP_ATAN = 29, MFHIP_ATANPI = 30,
  MFHIP_ASIN = 31, MFHIP_ACOS = 32, MFHIP_ASINPI = 33, MFHIP_ACOSPI = 34
};
enum {
  MFHIP_ADD = 0, MFHIP_SUB = 1, MFHIP_MUL = 2, MFHIP_DIV = 3, MFHIP_POW = 4, MFHIP_HYPOT = 5, MFHIP_FDIM = 6,
  MFHIP_FMIN = 7, MFHIP_FMAX = 8, MFHIP_COPYSIGN = 9, MFHIP_ATAN2 = 10, MFHIP_ATAN2PI = 11,
  MFHIP_MINF = 12, MFHIP_MAXF = 13, MFHIP_DIMF = 14, MFHIP_FMOD = 15, MFHIP_MODULO = 16, MFHIP_NEAREST = 17
};
enum { MFHIP_DOT = 0, MFHIP_SUM = 1, MFHIP_NORM2 = 2, MFHIP_PRODUCT = 3, MFHIP_MAXVAL = 4, MFHIP_MINVAL = 5 };

MULTIFLOATS_API int mfhip_matmul_mm(const float64x2 *a, const float64x2 *b, float64x2 *c,
                                    int64_t m, int64_t k, int64_t n, int64_t renorm_interval,
                                    void *stream);
/* Same, with an explicit tile-shape variant in [0, mfhip_matmul_variants()) --
 * the hook the autotuner drives; variant -1 is the built-in size rule. */
MULTIFLOATS_API int mfhip_matmul_variants(void);
MULTIFLOATS_API int mfhip_matmul_mm_variant(const float64x2 *a, const float64x2 *b, float64x2 *c,
                                            int64_t m, int64_t k, int64_t n,
                                            int64_t renorm_interval, int variant, void *stream);
MULTIFLOATS_API int mfhip_cmatmul_mm(const complex64x2 *a, const complex64x2 *b, complex64x2 *c,
                                     int64_t m, int64_t k, int64_t n, int64_t renorm_interval,
                                     void *stream);
MULTIFLOATS_API int mfhip_matmul_mv(const float64x2 *a, const float64x2 *x, float64x2 *y,
                                    int64_t m, int64_t k, int64_t renorm_interval, void *stream);
MULTIFLOATS_API int mfhip_matmul_vm(const float64x2 *x, const float64x2 *b, float64x2 *y,
                                    int64_t k, int64_t n, int64_t renorm_interval, void *stream);
MULTIFLOATS_API int mfhip_unary(int op, const float64x2 *a, float64x2 *c, int64_t n,
                                void *stream);
MULTIFLOATS_API int mfhip_binary(int op, const float64x2 *a, const float64x2 *b, float64x2 *c,
                                 int64_t n, void *stream);
/* Reductions write one float64x2 to *out (device). `work` is caller-owned
 * device scratch of at least mfhip_reduce_workspace_bytes(n) bytes. b is
 * only read for MFHIP_DOT; renorm_interval only matters for MFHIP_DOT. */
MULTIFLOATS_API size_t mfhip_reduce_workspace_bytes(int64_t n);
MULTIFLOATS_API int mfhip_reduce(int kind, const float64x2 *a, const float64x2 *b, int64_t n,
                                 int64_t renorm_interval, float64x2 *out, void *work,
                                 void *stream);
/* mfhip_reduce without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. The
 * partials go to a 16 MiB array that the library holds in each GPU's memory
 * until the process ends; the call itself allocates nothing. Same kernels
 * and results (bit for bit) as mfhip_reduce. n above MFHIP_NOALLOC_MAX_N is
 * rejected (hipErrorInvalidValue) before anything is enqueued. All calls on
 * one GPU share its array: calls that overlap on that GPU (on different
 * streams, or from different threads without ordering) race. */
#define MFHIP_NOALLOC_MAX_N ((int64_t)1 << 30)
MULTIFLOATS_API int mfhip_reduce_noalloc(int kind, const float64x2 *a, const float64x2 *b, int64_t n,
                                         int64_t renorm_interval, float64x2 *out, void *stream);

/* Device forms of sumdd_mask (work as for mfhip_reduce), cdotdd (work of
 * mfhip_cdot_workspace_bytes(n) bytes; *out is one complex64x2) and the
 * _dim reductions (kind MFHIP_SUM, MFHIP_NORM2, MFHIP_PRODUCT, MFHIP_MAXVAL
 * or MFHIP_MINVAL; no workspace; mask may be NULL and is refused with
 * MFHIP_NORM2). mfhip_reduce takes MFHIP_PRODUCT / MAXVAL / MINVAL too. */
MULTIFLOATS_API int mfhip_sum_masked(const float64x2 *a, const int32_t *mask, int64_t n,
                                     float64x2 *out, void *work, void *stream);
/* Masked rank-1 reduction of any kind but MFHIP_DOT / MFHIP_NORM2 (sum,
 * product, maxval, minval); work as for mfhip_reduce. */
MULTIFLOATS_API int mfhip_reduce_masked(int kind, const float64x2 *a, const int32_t *mask, int64_t n,
                                        float64x2 *out, void *work, void *stream);
MULTIFLOATS_API size_t mfhip_cdot_workspace_bytes(int64_t n);
MULTIFLOATS_API int mfhip_cdot(const complex64x2 *a, const complex64x2 *b, int64_t n,
                               int64_t renorm_interval, complex64x2 *out, void *work, void *stream);
MULTIFLOATS_API int mfhip_reduce_dim(int kind, const float64x2 *a, const int32_t *mask, int64_t inner,
                                     int64_t len, int64_t outer, float64x2 *out, void *stream);

/* Device form of wgemmdd. */
MULTIFLOATS_API int mfhip_gemm(char transa, char transb, int64_t m, int64_t n, int64_t k, float64x2 alpha,
                               const float64x2 *a, int64_t lda, const float64x2 *b, int64_t ldb,
                               float64x2 beta, float64x2 *c, int64_t ldc, int64_t renorm_interval,
                               void *stream);

/* Device form of wtrsmdd; work holds mfhip_trsm_workspace_bytes(side, m, n)
 * bytes (read and written; calls that share it must not overlap). */
MULTIFLOATS_API size_t mfhip_trsm_workspace_bytes(char side, int64_t m, int64_t n);
MULTIFLOATS_API int mfhip_trsm(char side, char uplo, char transa, char diag, int64_t m, int64_t n,
                               float64x2 alpha, const float64x2 *a, int64_t lda, float64x2 *b, int64_t ldb,
                               void *work, size_t work_bytes, void *stream);

/* Ozaki GEMM on device pointers; `work` is caller-owned device scratch of at
 * least mfhip_ozaki_workspace_bytes(m, k, n, slices) bytes (the size for
 * slices <= 0 also serves every explicit count). The call only enqueues work:
 * with slices <= 0 the slice count is chosen on the device, so the call may be
 * captured in a HIP graph. Calls sharing a workspace must not overlap. */
MULTIFLOATS_API size_t mfhip_ozaki_workspace_bytes(int64_t m, int64_t k, int64_t n, int slices);
MULTIFLOATS_API int mfhip_ozaki_matmul_mm(const float64x2 *a, const float64x2 *b, float64x2 *c,
                                          int64_t m, int64_t k, int64_t n, int slices, void *work,
                                          size_t work_bytes, void *stream);
/* What the last Ozaki call on `work` (same m, k, n, slices) did: the slice
 * count of its products and how many outputs the fix-up recomputed (0 for
 * explicit counts). Synchronises `stream`. No counterpart in the reference. */
MULTIFLOATS_API int mfhip_ozaki_info(const void *work, int64_t m, int64_t k, int64_t n, int slices,
                                     int *slices_used, int64_t *fixups, void *stream);
/* Moduli count of the residue GEMMs of the last Ozaki call on `work` (0 when
 * an explicit count ran the triangular form; automatic calls always run the
 * CRT form). An automatic call uses no more moduli than
 * mfhip_ozaki_moduli gives for its slice count, and fewer when the data
 * bound the product (the certificate's row and column sums bound every
 * |C_int| well below the worst case k 2^(16 slices - 2)). Synchronises
 * `stream`. No counterpart in the reference. */
MULTIFLOATS_API int mfhip_ozaki_moduli_used(const void *work, int64_t m, int64_t k, int64_t n, int slices,
                                            int *moduli, void *stream);
/* Slice count and fix-up count of the calling thread's last
 * matmuldd_mm_ozaki call; after mfhip_ozaki_matmul_mm the count is the
 * explicit one, or -1 for an automatic call (use mfhip_ozaki_info). */
MULTIFLOATS_API int mfhip_ozaki_slices_used(void);
MULTIFLOATS_API int64_t mfhip_ozaki_fixups(void);
/* The exponent-range rule with which an automatic call starts its search
 * (range_a, range_b: the spread of the row exponents of A and of the column
 * exponents of B). */
MULTIFLOATS_API int mfhip_ozaki_choose_slices(int64_t k, int range_a, int range_b);
/* Products form of an Ozaki call with these sizes and slice count (7..16):
 * the number N of pairwise-coprime moduli of the CRT form (N INT8 GEMMs with
 * residue operands, then Chinese-remainder reconstruction), which runs for
 * m, n >= 512; 0 when the triangular form runs (S(S+1)/2 digit-pair
 * products in one fused kernel). No counterpart in the reference. */
MULTIFLOATS_API int mfhip_ozaki_moduli(int64_t m, int64_t k, int64_t n, int slices);
/* mfhip_ozaki_matmul_mm in two steps with an explicit slice count (7..16):
 * prepare scales and slices A and B into
 * `work`; multiply forms C from `work` (same m, k, n, slices, work). A
 * prepared `work` may be multiplied again, e.g. to time the product alone. */
MULTIFLOATS_API int mfhip_ozaki_prepare(const float64x2 *a, const float64x2 *b, int64_t m,
                                        int64_t k, int64_t n, int slices, void *work,
                                        size_t work_bytes, void *stream);
MULTIFLOATS_API int mfhip_ozaki_multiply(float64x2 *c, int64_t m, int64_t k, int64_t n,
                                         int slices, void *work, size_t work_bytes,
                                         void *stream);
/* The products stage of mfhip_ozaki_multiply in the CRT form
 * (mfhip_ozaki_moduli(m, k, n, slices) > 0) alone: the residue INT8 GEMMs
 * from a prepared `work` into `work`, without the reconstruction of C (e.g.
 * to time or schedule the products separately). Returns an error for the
 * triangular form, whose products and recombination are one kernel. With
 * slices <= 0 it runs the residue GEMMs of the last automatic call on `work`
 * again, with the slice and moduli counts that call chose. No counterpart in
 * the reference. */
MULTIFLOATS_API int mfhip_ozaki_products(int64_t m, int64_t k, int64_t n, int slices, void *work,
                                         size_t work_bytes, void *stream);

/* Batching form (no counterpart in the reference): many independent DD GEMMs
 * (matmuldd_mm semantics, bit-identical per problem) and dot products
 * (dotdd semantics) in one persistent grouped launch. A plan is built once
 * (it allocates its device-side work list), then run any number of times on
 * the same buffers; run_graph captures the launch in a HIP graph on first use
 * and replays it afterwards. tile_variant[i] picks the GEMM tile for problem
 * i: 0 = 32x32, 1 = 64x64, 2 = 128x16, 3 = 16x128, 4 = 64x32, 5 = 32x64,
 * 6 = 16x16, 7 = 48x48, -1 = built-in rule: least padded work weighted by a
 * per-shape cost (NULL: rule for all). The problems must be independent: no
 * output may overlap another problem's input or output (the GEMM tile kinds
 * run in an order of the plan's choosing, and the dots of a plan with GEMMs
 * run on a second stream at the same time as the GEMMs). Runs of one plan
 * must not overlap one another. */
typedef struct {
  const float64x2 *a;
  const float64x2 *b;
  float64x2 *c;
  int64_t m, k, n, renorm_interval;
} mfhip_gemm_problem;
typedef struct {
  const float64x2 *a;
  const float64x2 *b;
  float64x2 *out;
  int64_t n, renorm_interval;
} mfhip_dot_problem;
MULTIFLOATS_API void *mfhip_grouped_create(const mfhip_gemm_problem *gemms, int ngemm,
                                           const int *tile_variant, const mfhip_dot_problem *dots,
                                           int ndot);
MULTIFLOATS_API int mfhip_grouped_items(void *plan);
MULTIFLOATS_API int mfhip_grouped_run(void *plan, void *stream);
MULTIFLOATS_API int mfhip_grouped_run_graph(void *plan, void *stream);
MULTIFLOATS_API void mfhip_grouped_destroy(void *plan);

/* Last HIP error recorded on the calling thread by any entry point (0 = none). */
MULTIFLOATS_API int mfhip_last_error(void);
MULTIFLOATS_API const char *mfhip_error_string(int err);

#ifdef __cplusplus
} /* extern "C" */
} /* namespace multifloats */
#endif

#endif /* MULTIFLOATS_HIP_H */
