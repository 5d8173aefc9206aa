// Elementwise DD array kernels: + - * / sqrt (HBM-bound, bit-exact with the
// reference operators) and exp / log / sin / erf (FP64-issue-bound).
//
// The reference evaluates these one scalar C-ABI call per element from a
// compiler-generated Fortran elemental loop (fsrc/multifloats.fypp:1407-1412,
// 793-799); here one launch covers the whole array.
//
// HBM-bound ops (+ - * / sqrt): one element per lane and one lane per
// element (no grid-stride loop), 256-lane workgroups, nontemporal 16-byte
// loads and stores on the AoS {hi,lo} layout. Measured on MI355X at N = 2^26
// this streams at ~6.5 TB/s; grid-stride loops with several loads in flight
// per lane and a capped grid reached only ~4.8 TB/s.
//
// Transcendentals (FP64-issue-bound): grid-stride, one element per step.
#include "dd_math.h"
#include "mf_internal.h"

namespace {

constexpr int kThreads = 256;

template <int OP>
__device__ __forceinline__ ddv unary_op(ddv x) {
  if constexpr (OP == mfhip::OP_SQRT) return dd_sqrt(x);
  else if constexpr (OP == mfhip::OP_EXP) return dd_exp(x);
  else if constexpr (OP == mfhip::OP_LOG) return dd_log(x);
  else if constexpr (OP == mfhip::OP_SIN) return dd_sin(x);
  else if constexpr (OP == mfhip::OP_ERF) return dd_erf(x);
  else if constexpr (OP == mfhip::OP_NEG) return dd_neg(x);
  else if constexpr (OP == mfhip::OP_COS) return dd_cos(x);
  else if constexpr (OP == mfhip::OP_TAN) return dd_tan(x);
  else if constexpr (OP == mfhip::OP_EXP2) return dd_exp2(x);
  else if constexpr (OP == mfhip::OP_EXPM1) return dd_expm1(x);
  else if constexpr (OP == mfhip::OP_LOG2) return dd_log2(x);
  else if constexpr (OP == mfhip::OP_LOG10) return dd_log10(x);
  else if constexpr (OP == mfhip::OP_LOG1P) return dd_log1p(x);
  else if constexpr (OP == mfhip::OP_SINH) return dd_sinh(x);
  else if constexpr (OP == mfhip::OP_COSH) return dd_cosh(x);
  else if constexpr (OP == mfhip::OP_TANH) return dd_tanh(x);
  else if constexpr (OP == mfhip::OP_ERFC) return dd_erfc(x);
  else if constexpr (OP == mfhip::OP_ERFCX) return dd_erfcx(x);
  else if constexpr (OP == mfhip::OP_ASINH) return dd_asinh(x);
  else if constexpr (OP == mfhip::OP_ACOSH) return dd_acosh(x);
  else if constexpr (OP == mfhip::OP_ATANH) return dd_atanh(x);
  else if constexpr (OP == mfhip::OP_EXP10) return dd_exp10(x);
  else if constexpr (OP == mfhip::OP_EXP2M1) return dd_exp2m1(x);
  else if constexpr (OP == mfhip::OP_EXP10M1) return dd_exp10m1(x);
  else if constexpr (OP == mfhip::OP_LOG2P1) return dd_log2p1(x);
  else if constexpr (OP == mfhip::OP_LOG10P1) return dd_log10p1(x);
  else if constexpr (OP == mfhip::OP_SINPI) return dd_sinpi(x);
  else if constexpr (OP == mfhip::OP_COSPI) return dd_cospi(x);
  else if constexpr (OP == mfhip::OP_TANPI) return dd_tanpi(x);
  else if constexpr (OP == mfhip::OP_ATAN) return dd_atan(x);
  else if constexpr (OP == mfhip::OP_ATANPI) return dd_atanpi(x);
  else if constexpr (OP == mfhip::OP_ASIN) return dd_asin(x);
  else if constexpr (OP == mfhip::OP_ACOS) return dd_acos(x);
  else if constexpr (OP == mfhip::OP_ASINPI) return dd_asinpi(x);
  else if constexpr (OP == mfhip::OP_ACOSPI) return dd_acospi(x);
  else return x;
}

template <int OP>
__device__ __forceinline__ ddv binary_op(ddv x, ddv y) {
  if constexpr (OP == mfhip::OP_ADD) return dd_add(x, y);
  else if constexpr (OP == mfhip::OP_SUB) return dd_sub(x, y);
  else if constexpr (OP == mfhip::OP_MUL) return dd_mul(x, y);
  else if constexpr (OP == mfhip::OP_DIV) return dd_div(x, y);
  else if constexpr (OP == mfhip::OP_POW) return dd_pow(x, y);
  else if constexpr (OP == mfhip::OP_HYPOT) return dd_hypot(x, y);
  else if constexpr (OP == mfhip::OP_FDIM) return dd_fdim(x, y);
  else if constexpr (OP == mfhip::OP_FMIN) return dd_fmin(x, y);
  else if constexpr (OP == mfhip::OP_FMAX) return dd_fmax(x, y);
  else if constexpr (OP == mfhip::OP_COPYSIGN) return dd_copysign(x, y);
  else if constexpr (OP == mfhip::OP_ATAN2) return dd_atan2(x, y);
  else if constexpr (OP == mfhip::OP_ATAN2PI) return dd_atan2pi(x, y);
  else if constexpr (OP == mfhip::OP_MINF) return dd_fortran_min(x, y);
  else if constexpr (OP == mfhip::OP_MAXF) return dd_fortran_max(x, y);
  else if constexpr (OP == mfhip::OP_FMOD) return dd_fmod(x, y);
  else if constexpr (OP == mfhip::OP_MODULO) return dd_modulo(x, y);
  else if constexpr (OP == mfhip::OP_NEAREST) return dd_nearest(x, y);
  else return dd_fortran_dim(x, y);
}

template <int OP, int U>
__global__ __launch_bounds__(kThreads) void k_unary(const double2 *__restrict__ a,
                                                    double2 *__restrict__ c, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t base = (int64_t)blockIdx.x * kThreads + threadIdx.x; base < n; base += stride * U) {
    v2d v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int64_t i = base + u * stride;
      v[u] = i < n ? __builtin_nontemporal_load((const v2d *)a + i) : v2d{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int64_t i = base + u * stride;
      if (i < n) {
        ddv r = unary_op<OP>(mk(v[u].x, v[u].y));
        __builtin_nontemporal_store(v2d{r.hi, r.lo}, (v2d *)c + i);
      }
    }
  }
}

// One element per lane (HBM-bound ops).
template <int OP>
__global__ __launch_bounds__(kThreads) void k_unary1(const double2 *__restrict__ a, double2 *__restrict__ c,
                                                     int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const v2d v = __builtin_nontemporal_load((const v2d *)a + i);
  const ddv r = unary_op<OP>(mk(v.x, v.y));
  __builtin_nontemporal_store(v2d{r.hi, r.lo}, (v2d *)c + i);
}

template <int OP>
__global__ __launch_bounds__(kThreads) void k_binary1(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                      double2 *__restrict__ c, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const v2d va = __builtin_nontemporal_load((const v2d *)a + i);
  const v2d vb = __builtin_nontemporal_load((const v2d *)b + i);
  const ddv r = binary_op<OP>(mk(va.x, va.y), mk(vb.x, vb.y));
  __builtin_nontemporal_store(v2d{r.hi, r.lo}, (v2d *)c + i);
}

template <int OP, int U>
__global__ __launch_bounds__(kThreads) void k_binary(const double2 *__restrict__ a,
                                                     const double2 *__restrict__ b,
                                                     double2 *__restrict__ c, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t base = (int64_t)blockIdx.x * kThreads + threadIdx.x; base < n; base += stride * U) {
    v2d va[U], vb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int64_t i = base + u * stride;
      va[u] = i < n ? __builtin_nontemporal_load((const v2d *)a + i) : v2d{0.0, 0.0};
      vb[u] = i < n ? __builtin_nontemporal_load((const v2d *)b + i) : v2d{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int64_t i = base + u * stride;
      if (i < n) {
        ddv r = binary_op<OP>(mk(va[u].x, va[u].y), mk(vb[u].x, vb[u].y));
        __builtin_nontemporal_store(v2d{r.hi, r.lo}, (v2d *)c + i);
      }
    }
  }
}

// erf over an array with its row tables in LDS. The evaluation picks one of
// MF_ERF_ROWS Taylor rows per element, so the lanes of a wave read 28
// coefficients from as many different rows: from the constant tables every
// such load is a gather over up to 64 cache lines. Each workgroup copies the
// rows once into LDS as records of kErfRec doubles -- centre, three base
// limbs, the ND (hi, lo) coefficient pairs, the tail -- whose odd stride
// starts consecutive rows on different banks, then runs a grid-stride loop
// over the elements. The arithmetic is dd_erf's, so the bits are the same.
constexpr int kErfRec = 33;
static_assert(4 + 2 * MF_ERF_ND + (MF_ERF_NT - MF_ERF_ND) < kErfRec, "row record");
struct ErfTabLds {
  const double *t;
  __device__ __forceinline__ double centre(int row) const { return t[row * kErfRec]; }
  __device__ __forceinline__ double base(int row, int i) const { return t[row * kErfRec + 1 + i]; }
  __device__ __forceinline__ double coef_hi(int row, int j) const { return t[row * kErfRec + 4 + 2 * j]; }
  __device__ __forceinline__ double coef_lo(int row, int j) const { return t[row * kErfRec + 5 + 2 * j]; }
  __device__ __forceinline__ double tail(int row, int j) const { return t[row * kErfRec + 4 + 2 * MF_ERF_ND + j]; }
};
constexpr int kErfThreads = 512;  // 4 workgroups (37 KB of LDS each) and 8 waves per SIMD
__global__ __launch_bounds__(kErfThreads) void k_erf_lds(const double2 *__restrict__ a, double2 *__restrict__ c,
                                                         int64_t n) {
  __shared__ double tab[MF_ERF_ROWS * kErfRec];
  for (int i = threadIdx.x; i < MF_ERF_ROWS * kErfRec; i += kErfThreads) {
    const int row = i / kErfRec, f = i - row * kErfRec;
    double v = 0.0;
    if (f == 0) v = mf_erf_centre[row];
    else if (f < 4) v = mf_erf_base[3 * row + f - 1];
    else if (f < 4 + 2 * MF_ERF_ND) v = ((f - 4) & 1 ? mf_erf_ddlo : mf_erf_ddhi)[row * MF_ERF_ND + (f - 4) / 2];
    else if (f < 4 + 2 * MF_ERF_ND + (MF_ERF_NT - MF_ERF_ND))
      v = mf_erf_d[row * (MF_ERF_NT - MF_ERF_ND) + f - 4 - 2 * MF_ERF_ND];
    tab[i] = v;
  }
  __syncthreads();
  const ErfTabLds lt{tab};
  const int64_t stride = (int64_t)gridDim.x * kErfThreads;
  for (int64_t i = (int64_t)blockIdx.x * kErfThreads + threadIdx.x; i < n; i += stride) {
    const v2d v = __builtin_nontemporal_load((const v2d *)a + i);
    const ddv r = dd_erf_t(lt, mk(v.x, v.y));
    __builtin_nontemporal_store(v2d{r.hi, r.lo}, (v2d *)c + i);
  }
}

// Grid: enough workgroups to cover n once with U elements per lane, capped so
// that every CU holds several resident groups (256 CUs on MI355X).
unsigned grid_for(int64_t n, int U, int max_blocks) {
  int64_t per = (int64_t)kThreads * U;
  int64_t g = (n + per - 1) / per;
  if (g > max_blocks) g = max_blocks;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// Elements for one lane per element: a launch holds fewer than 2^32
// work-items in total; beyond that the grid-stride form takes over.
constexpr int64_t kMaxOneShot = (int64_t)(0xffffffffu / kThreads) * kThreads;

template <int OP>
hipError_t run_unary(const double2 *a, double2 *c, int64_t n, hipStream_t s) {
  if ((OP == mfhip::OP_SQRT || OP == mfhip::OP_NEG) && n <= kMaxOneShot) {
    hipLaunchKernelGGL((k_unary1<OP>), dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       a, c, n);
  } else if (OP == mfhip::OP_SQRT || OP == mfhip::OP_NEG) {
    hipLaunchKernelGGL((k_unary<OP, 4>), dim3(grid_for(n, 4, 8192)), dim3(kThreads), 0, s, a, c, n);
  } else if (OP == mfhip::OP_ERF) {
    // a few resident workgroups per CU, each staging the tables once
    hipLaunchKernelGGL(k_erf_lds, dim3(grid_for(n, 32, 1024)), dim3(kErfThreads), 0, s, a, c, n);
  } else {
    hipLaunchKernelGGL((k_unary<OP, 1>), dim3(grid_for(n, 1, 1 << 20)), dim3(kThreads), 0, s, a, c, n);
  }
  return hipGetLastError();
}

template <int OP>
hipError_t run_binary(const double2 *a, const double2 *b, double2 *c, int64_t n, hipStream_t s) {
  if (n <= kMaxOneShot)
    hipLaunchKernelGGL((k_binary1<OP>), dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       a, b, c, n);
  else
    hipLaunchKernelGGL((k_binary<OP, 4>), dim3(grid_for(n, 4, 8192)), dim3(kThreads), 0, s, a, b, c, n);
  return hipGetLastError();
}

}  // namespace

namespace mfhip {

hipError_t unary(int op, const double2 *a, double2 *c, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  switch (op) {
    case OP_SQRT: return run_unary<OP_SQRT>(a, c, n, s);
    case OP_EXP: return run_unary<OP_EXP>(a, c, n, s);
    case OP_LOG: return run_unary<OP_LOG>(a, c, n, s);
    case OP_SIN: return run_unary<OP_SIN>(a, c, n, s);
    case OP_ERF: return run_unary<OP_ERF>(a, c, n, s);
    case OP_NEG: return run_unary<OP_NEG>(a, c, n, s);
    case OP_COS: return run_unary<OP_COS>(a, c, n, s);
    case OP_TAN: return run_unary<OP_TAN>(a, c, n, s);
    case OP_EXP2: return run_unary<OP_EXP2>(a, c, n, s);
    case OP_EXPM1: return run_unary<OP_EXPM1>(a, c, n, s);
    case OP_LOG2: return run_unary<OP_LOG2>(a, c, n, s);
    case OP_LOG10: return run_unary<OP_LOG10>(a, c, n, s);
    case OP_LOG1P: return run_unary<OP_LOG1P>(a, c, n, s);
    case OP_SINH: return run_unary<OP_SINH>(a, c, n, s);
    case OP_COSH: return run_unary<OP_COSH>(a, c, n, s);
    case OP_TANH: return run_unary<OP_TANH>(a, c, n, s);
    case OP_ERFC: return run_unary<OP_ERFC>(a, c, n, s);
    case OP_ERFCX: return run_unary<OP_ERFCX>(a, c, n, s);
    case OP_ASINH: return run_unary<OP_ASINH>(a, c, n, s);
    case OP_ACOSH: return run_unary<OP_ACOSH>(a, c, n, s);
    case OP_ATANH: return run_unary<OP_ATANH>(a, c, n, s);
    case OP_EXP10: return run_unary<OP_EXP10>(a, c, n, s);
    case OP_EXP2M1: return run_unary<OP_EXP2M1>(a, c, n, s);
    case OP_EXP10M1: return run_unary<OP_EXP10M1>(a, c, n, s);
    case OP_LOG2P1: return run_unary<OP_LOG2P1>(a, c, n, s);
    case OP_LOG10P1: return run_unary<OP_LOG10P1>(a, c, n, s);
    case OP_SINPI: return run_unary<OP_SINPI>(a, c, n, s);
    case OP_COSPI: return run_unary<OP_COSPI>(a, c, n, s);
    case OP_TANPI: return run_unary<OP_TANPI>(a, c, n, s);
    case OP_ATAN: return run_unary<OP_ATAN>(a, c, n, s);
    case OP_ATANPI: return run_unary<OP_ATANPI>(a, c, n, s);
    case OP_ASIN: return run_unary<OP_ASIN>(a, c, n, s);
    case OP_ACOS: return run_unary<OP_ACOS>(a, c, n, s);
    case OP_ASINPI: return run_unary<OP_ASINPI>(a, c, n, s);
    case OP_ACOSPI: return run_unary<OP_ACOSPI>(a, c, n, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t binary(int op, const double2 *a, const double2 *b, double2 *c, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  switch (op) {
    case OP_ADD: return run_binary<OP_ADD>(a, b, c, n, s);
    case OP_SUB: return run_binary<OP_SUB>(a, b, c, n, s);
    case OP_MUL: return run_binary<OP_MUL>(a, b, c, n, s);
    case OP_DIV: return run_binary<OP_DIV>(a, b, c, n, s);
    case OP_POW: return run_binary<OP_POW>(a, b, c, n, s);
    case OP_HYPOT: return run_binary<OP_HYPOT>(a, b, c, n, s);
    case OP_FDIM: return run_binary<OP_FDIM>(a, b, c, n, s);
    case OP_FMIN: return run_binary<OP_FMIN>(a, b, c, n, s);
    case OP_FMAX: return run_binary<OP_FMAX>(a, b, c, n, s);
    case OP_COPYSIGN: return run_binary<OP_COPYSIGN>(a, b, c, n, s);
    case OP_ATAN2: return run_binary<OP_ATAN2>(a, b, c, n, s);
    case OP_ATAN2PI: return run_binary<OP_ATAN2PI>(a, b, c, n, s);
    case OP_MINF: return run_binary<OP_MINF>(a, b, c, n, s);
    case OP_MAXF: return run_binary<OP_MAXF>(a, b, c, n, s);
    case OP_DIMF: return run_binary<OP_DIMF>(a, b, c, n, s);
    case OP_FMOD: return run_binary<OP_FMOD>(a, b, c, n, s);
    case OP_MODULO: return run_binary<OP_MODULO>(a, b, c, n, s);
    case OP_NEAREST: return run_binary<OP_NEAREST>(a, b, c, n, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace mfhip
