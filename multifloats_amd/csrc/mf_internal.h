// Internal launcher interface shared by the kernel translation units and the
// C-ABI layer. Every launcher enqueues on the given stream and returns the
// launch status; none of them synchronises or allocates.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

typedef double v2d __attribute__((ext_vector_type(2)));

struct ddv;  // dd_core.h

namespace mfhip {

enum {
  OP_SQRT = 0, OP_EXP = 1, OP_LOG = 2, OP_SIN = 3, OP_ERF = 4, OP_NEG = 5,
  OP_COS = 6, OP_TAN = 7, OP_EXP2 = 8, OP_EXPM1 = 9, OP_LOG2 = 10,
  OP_LOG10 = 11, OP_LOG1P = 12, OP_SINH = 13, OP_COSH = 14, OP_TANH = 15,
  OP_ERFC = 16, OP_ERFCX = 17, OP_ASINH = 18, OP_ACOSH = 19, OP_ATANH = 20,
  OP_EXP10 = 21, OP_EXP2M1 = 22, OP_EXP10M1 = 23, OP_LOG2P1 = 24, OP_LOG10P1 = 25,
  OP_SINPI = 26, OP_COSPI = 27, OP_TANPI = 28, OP_ATAN = 29, OP_ATANPI = 30,
  OP_ASIN = 31, OP_ACOS = 32, OP_ASINPI = 33, OP_ACOSPI = 34
};
enum {
  OP_ADD = 0, OP_SUB = 1, OP_MUL = 2, OP_DIV = 3, OP_POW = 4, OP_HYPOT = 5, OP_FDIM = 6, OP_FMIN = 7,
  OP_FMAX = 8, OP_COPYSIGN = 9, OP_ATAN2 = 10, OP_ATAN2PI = 11,
  OP_MINF = 12, OP_MAXF = 13, OP_DIMF = 14, OP_FMOD = 15, OP_MODULO = 16, OP_NEAREST = 17
};
enum { RED_DOT = 0, RED_SUM = 1, RED_NORM2 = 2, RED_PROD = 3, RED_MAXVAL = 4, RED_MINVAL = 5 };

// Plain reference-order GEMM; variant < 0 picks by size, otherwise one of
// gemm_plain_variants() tile shapes (index order is stable for the tuner).
int gemm_plain_variants();
hipError_t gemm_plain(const double2 *A, const double2 *B, double2 *C, int64_t m, int64_t k,
                      int64_t n, int64_t ri, int variant, hipStream_t s);
// Complex GEMM (complex64x2 as double4), bit-identical to the reference's
// four-real-GEMM composition; also serves mv (n = 1) and vm (m = 1).
hipError_t gemm_cplx(const double4 *A, const double4 *B, double4 *C, int64_t m, int64_t k,
                     int64_t n, int64_t ri, hipStream_t s);
// C = beta C on an m x n block with leading dimension ldc (beta == 0: zeros,
// C not read).
hipError_t scale_matrix(double2 *C, int64_t ldc, int64_t m, int64_t n, ddv beta, hipStream_t s);
// xTRSM of wtrsm.f90 (side, uplo, transa, diag; X overwrites B); work holds
// trsm_workspace_bytes(side, m, n) bytes (pivot masks of side 'L').
size_t trsm_workspace_bytes(char side, int64_t m, int64_t n);
hipError_t trsm_check(char side, char uplo, char ta, char diag, int64_t m, int64_t n, int64_t lda, int64_t ldb);
hipError_t trsm(char side, char uplo, char ta, char diag, int64_t m, int64_t n, ddv alpha, const double2 *A,
                int64_t lda, double2 *B, int64_t ldb, void *work, size_t work_bytes, hipStream_t st);
// BLAS argument checks of wgemm.f90:16-51 (transposes, sizes, leading dimensions)
hipError_t gemm_blas_check(char ta, char tb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                           int64_t ldc);
// BLAS-style C = alpha op(A) op(B) + beta C (ta/tb 'N', 'T' or 'C'; leading
// dimensions lda/ldb/ldc); invalid arguments give hipErrorInvalidValue.
hipError_t gemm_blas(char ta, char tb, int64_t m, int64_t n, int64_t k, ddv alpha, const double2 *A,
                     int64_t lda, const double2 *B, int64_t ldb, ddv beta, double2 *C, int64_t ldc, int64_t ri,
                     hipStream_t s);
hipError_t gemv(const double2 *A, const double2 *x, double2 *y, int64_t m, int64_t k, int64_t ri,
                hipStream_t s);
hipError_t gevm(const double2 *x, const double2 *B, double2 *y, int64_t k, int64_t n, int64_t ri,
                hipStream_t s);

hipError_t unary(int op, const double2 *a, double2 *c, int64_t n, hipStream_t s);
hipError_t binary(int op, const double2 *a, const double2 *b, double2 *c, int64_t n, hipStream_t s);

int64_t reduce_groups(int64_t n);
size_t reduce_workspace_bytes(int64_t n);
// mask (sum only, may be null): elements with mask[i] == 0 are skipped.
hipError_t reduce(int kind, const double2 *a, const double2 *b, int64_t n, int64_t ri,
                  double2 *out, double2 *work, hipStream_t s, const int32_t *mask = nullptr);
// Complex conj(a).b; work holds cdot_workspace_bytes(n).
size_t cdot_workspace_bytes(int64_t n);
hipError_t cdot(const double4 *a, const double4 *b, int64_t n, int64_t ri, double4 *out, double4 *work,
                hipStream_t s);
// Bit-exact sum / norm2 along one dimension of an (inner, len, outer)
// column-major array into out(inner, outer); needs no workspace.
hipError_t reduce_dim(int kind, const double2 *a, const int32_t *mask, int64_t inner, int64_t len,
                      int64_t outer, double2 *out, hipStream_t s);

// Ozaki split-precision GEMM with S int8 slices per operand (7 <= S <= 16),
// or S <= 0: the CRT form with the slice count chosen on the device and every
// output certified to 2^-100 (|A||B|)_ij (uncertified outputs recomputed by
// a compensated dot). Enqueues only; never synchronises.
int ozaki_max_slices();
size_t ozaki_workspace_bytes(int64_t m, int64_t k, int64_t n, int S);
hipError_t ozaki_gemm(const double2 *A, const double2 *B, double2 *C, int64_t m, int64_t k,
                      int64_t n, int S, void *work, size_t work_bytes, hipStream_t s);
// The exponent-range rule that starts an automatic call's search.
int ozaki_choose_slices(int64_t k, int range_a, int range_b);
// Slice count, fix-up count and moduli count (CRT form; 0 otherwise) of the
// last call on `work` (synchronises s).
hipError_t ozaki_info(const void *work, int64_t m, int64_t k, int64_t n, int S, int *slices, int64_t *fixups,
                      int *moduli, hipStream_t s);
// Moduli of the CRT form that a call with these sizes and S slices runs
// (0: the triangular form runs).
int ozaki_moduli(int64_t m, int64_t k, int64_t n, int S);
// The two halves of ozaki_gemm: scaling + slicing into `work`, then the
// INT8 MFMA products and DD recombination from `work` into C.
hipError_t ozaki_prepare(const double2 *A, const double2 *B, int64_t m, int64_t k, int64_t n,
                         int S, void *work, size_t work_bytes, hipStream_t s);
hipError_t ozaki_multiply(double2 *C, int64_t m, int64_t k, int64_t n, int S, void *work,
                          size_t work_bytes, hipStream_t s);
// The products stage of the CRT form alone (ozaki_moduli(m, k, n, S) > 0):
// the residue GEMMs into `work`; invalid for the triangular form.
hipError_t ozaki_products(int64_t m, int64_t k, int64_t n, int S, void *work, size_t work_bytes,
                          hipStream_t s);

}  // namespace mfhip
