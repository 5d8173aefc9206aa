// Triangular solve (xTRSM) on float64x2 for gfx950.
//
// Semantics: the reference's BLAS shim test/integration/blas/wtrsm.f90, bit
// for bit. Six of its eight loop nests (57-206) share one shape. B splits into
// independent vectors (columns of B for side 'L', rows for side 'R') of length
// s (m for 'L', n for 'R'). Each vector runs the same sequence of pivots p in
// a fixed order (ascending or descending): the pivot is finalised (x_p / a_pp,
// or (1 / a_pp) * x_p on side 'R'), then every later entry t takes
// x_t = x_t - coef(t, p) * x_p. Every entry therefore sees its updates in
// pivot order, which a right-looking blocked solve keeps:
//
//   for each block of 64 pivots (in pivot order):
//     k_trsm_diag    one wavefront per vector solves the block's own pivots
//                    (lane t holds x_t; pivots broadcast lane to lane);
//     k_trsm_update  a GEMM-shaped update of every later entry by the block's
//                    pivots, in pivot order, with x_t kept in registers.
//
// Two nests do not fit that shape: side 'L' transposed lower (107-125) and
// side 'R' untransposed lower (146-165) run their pivots in descending order
// but sum each entry's terms in ascending k, i.e. starting from the pivot
// finalised last. Entry t's chain cannot start before pivot t+1 is final, so
// the critical path is s^2/2 dependent DD steps per vector whatever the
// blocking; k_trsm_chain runs those two with one lane per vector.
//
// The shim's skips are kept: side 'L' without transpose skips a pivot whose
// value is zero before the division (wtrsm.f90:74, 89), recorded per vector in
// a 64-bit mask for the update kernel; side 'R' skips a zero coefficient
// (136, 157, 185, 196). alpha scales B up front where the shim scales a column
// before using it (N forms; always for side 'L' with transpose, 104, 116), and
// after the pivot's updates for side 'R' with transpose (199-203).
#include "dd_core.h"
#include "mf_internal.h"

namespace {

constexpr int kT = 64;         // pivots per block: one per lane of a wavefront
constexpr int kThreads = 256;  // 4 wavefronts
constexpr int kBK = 16;        // pivots per LDS stage in the update kernel (32 KiB of LDS)

enum Fin { FIN_NONE = 0, FIN_DIV = 1, FIN_RECIP = 2 };

__device__ __forceinline__ bool dd_nonzero(ddv x) { return x.hi != 0.0 || x.lo != 0.0; }

// x from a wavefront-uniform lane (v_readlane: no LDS round trip)
__device__ __forceinline__ double readlane_d(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ ddv shfl_dd(ddv x, int lane) { return mk(readlane_d(x.hi, lane), readlane_d(x.lo, lane)); }

// Actual index of block position q (pivot order) within [base, base + tb).
template <bool DESC>
__device__ __forceinline__ int64_t act(int64_t base, int tb, int q) {
  return DESC ? base + tb - 1 - q : base + q;
}

// One wavefront per vector: the block's tb pivots, in order. Side L: vector v
// is column v of B (x_t = B(t, v)); side R: row v (x_t = B(v, t)). coef(t, p)
// is a(p, t) when CT, else a(t, p).
template <bool R, bool CT, int FIN, bool ACTIVE_SKIP, bool COEF_SKIP, bool DESC>
__global__ __launch_bounds__(kThreads) void k_trsm_diag(const double2 *__restrict__ A, int64_t lda,
                                                        double2 *__restrict__ B, int64_t ldb, int64_t nvec,
                                                        int64_t base, int tb, uint64_t *__restrict__ mask) {
  __shared__ double2 coef[kT * kT];  // coef[q * kT + t], t > q
  __shared__ double2 diag[kT];
  const int tid = threadIdx.x;
  for (int e = tid; e < kT * kT; e += kThreads) {
    // threads walk the contiguous axis of A: t for a(t, p), q for a(p, t)
    const int q = CT ? e % kT : e / kT, t = CT ? e / kT : e % kT;
    if (q < tb && t < tb && t > q) {
      const int64_t pa = act<DESC>(base, tb, q), ta = act<DESC>(base, tb, t);
      coef[q * kT + t] = CT ? A[pa + ta * lda] : A[ta + pa * lda];
    }
  }
  if (FIN != FIN_NONE && tid < tb) {
    const int64_t pa = act<DESC>(base, tb, tid);
    diag[tid] = A[pa + pa * lda];
  }
  __syncthreads();
  const int lane = tid & 63;
  const int64_t v = (int64_t)blockIdx.x * (kThreads / 64) + (tid >> 6);
  if (v >= nvec) return;  // whole wavefront leaves together; no barrier follows
  double2 *bx = nullptr;
  if (lane < tb) {
    const int64_t ta = act<DESC>(base, tb, lane);
    bx = R ? B + v + ta * ldb : B + ta + v * ldb;
  }
  ddv x = mk(0.0, 0.0);
  if (bx) x = mk(bx->x, bx->y);
  uint64_t bits = 0;
  for (int q = 0; q < tb; ++q) {
    if (ACTIVE_SKIP) {
      if (!dd_nonzero(shfl_dd(x, q))) continue;  // uniform: the broadcast value
      bits |= 1ull << q;
    }
    if (FIN == FIN_DIV && lane == q) {
      const double2 d = diag[q];
      x = dd_div(x, mk(d.x, d.y));
    } else if (FIN == FIN_RECIP && lane == q) {
      const double2 d = diag[q];
      x = dd_mul(dd_div(mk(1.0, 0.0), mk(d.x, d.y)), x);
    }
    const ddv xq = shfl_dd(x, q);
    if (lane > q && lane < tb) {
      const double2 c = coef[q * kT + lane];
      const ddv cd = mk(c.x, c.y);
      if (!COEF_SKIP || dd_nonzero(cd)) x = dd_sub(x, dd_mul(cd, xq));
    }
  }
  if (bx) *bx = make_double2(x.hi, x.lo);
  if (ACTIVE_SKIP && lane == 0) mask[v] = bits;
}

// Later entries updated by one block's tb pivots, in pivot order:
// C(i, j) = C(i, j) - Aop(i, k) * Bop(k, j) for k = 0..tb-1 (k = pivot
// position), skipping where the shim skips. A 256-thread workgroup owns a
// 64 x 64 tile of C; thread (tr, tc) holds rows tr + 16a, columns tc + 16b.
//   side L: C = B(targets, vectors), Aop = coef(t, p), Bop = B(p, vector)
//   side R: C = B(vectors, targets), Aop = B(vector, p), Bop = coef(t, p)
template <bool R, bool CT, bool ACTIVE_SKIP, bool COEF_SKIP, bool DESC>
// Three waves per SIMD (stages of 16 pivots leave LDS for three workgroups per
// CU): 49.7-50.0 ms at m = n = 4096 against 54.3-54.7 ms at two.
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_trsm_update(
    const double2 *__restrict__ A, int64_t lda, double2 *__restrict__ B, int64_t ldb, int64_t r0, int64_t nrows,
    int64_t c0, int64_t ncols, int64_t base, int tb, int tiles_r, const uint64_t *__restrict__ mask) {
  __shared__ double2 As[kBK * 64];  // As[k * 64 + row]
  __shared__ double2 Bs[kBK * 64];  // Bs[k * 64 + col]
  const int tid = threadIdx.x, tr = tid & 15, tc = tid >> 4;
  const int64_t ti = (int64_t)(blockIdx.x % tiles_r) * 64, tj = (int64_t)(blockIdx.x / tiles_r) * 64;
  ddv acc[4][4];
  uint64_t mk_bits[4] = {~0ull, ~0ull, ~0ull, ~0ull};
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t i = ti + tr + 16 * a, j = tj + tc + 16 * b;
      acc[a][b] = mk(0.0, 0.0);
      if (i < nrows && j < ncols) {
        const double2 c = B[(r0 + i) + (c0 + j) * ldb];
        acc[a][b] = mk(c.x, c.y);
      }
    }
  if (ACTIVE_SKIP)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t j = tj + tc + 16 * b;
      mk_bits[b] = j < ncols ? mask[c0 + j] : 0ull;
    }
  for (int k0 = 0; k0 < tb; k0 += kBK) {
    __syncthreads();
    for (int e = tid; e < kBK * 64; e += kThreads) {
      // Aop: contiguous along rows except side L with transposed coef
      {
        const bool kmajor = !R && CT;
        const int k = kmajor ? e % kBK : e / 64, row = kmajor ? e / kBK : e % 64;
        const int64_t i = ti + row;
        double2 val = make_double2(0.0, 0.0);
        if (k0 + k < tb && i < nrows) {
          const int64_t p = act<DESC>(base, tb, k0 + k), ia = r0 + i;
          val = R ? B[ia + p * ldb] : (CT ? A[p + ia * lda] : A[ia + p * lda]);
        }
        As[k * 64 + row] = val;
      }
      // Bop: contiguous along k except side R with coef = a(t, p)
      {
        const bool cmajor = R && !CT;
        const int k = cmajor ? e / 64 : e % kBK, col = cmajor ? e % 64 : e / kBK;
        const int64_t j = tj + col;
        double2 val = make_double2(0.0, 0.0);
        if (k0 + k < tb && j < ncols) {
          const int64_t p = act<DESC>(base, tb, k0 + k), ja = c0 + j;
          val = R ? (CT ? A[p + ja * lda] : A[ja + p * lda]) : B[p + ja * ldb];
        }
        Bs[k * 64 + col] = val;
      }
    }
    __syncthreads();
    const int kn = tb - k0 < kBK ? tb - k0 : kBK;
    for (int k = 0; k < kn; ++k) {
      ddv av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double2 t = As[k * 64 + tr + 16 * a];
        av[a] = mk(t.x, t.y);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const double2 t = Bs[k * 64 + tc + 16 * b];
        bv[b] = mk(t.x, t.y);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        bool on = true;
        if (ACTIVE_SKIP) on = (mk_bits[b] >> (k0 + k)) & 1ull;
        if (COEF_SKIP) on = dd_nonzero(bv[b]);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const ddv u = dd_sub(acc[a][b], dd_mul(av[a], bv[b]));
          if (on) acc[a][b] = u;
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t i = ti + tr + 16 * a, j = tj + tc + 16 * b;
      if (i < nrows && j < ncols) B[(r0 + i) + (c0 + j) * ldb] = make_double2(acc[a][b].hi, acc[a][b].lo);
    }
}

// One lane per vector, the shim's loop as written, for the two nests whose
// per-entry sums run against the pivot order:
//   side L, 'T', lower (wtrsm.f90:107-125): x_i = (alpha x_i
//       - sum_{k=i+1..s-1} a(k,i) x_k) / a(i,i), i descending, k ascending;
//   side R, 'N', lower (146-165): x_j = alpha x_j (alpha != 1), then
//       x_j -= a(k,j) x_k for k = j+1..s-1 ascending with a(k,j) != 0, then
//       x_j = (1 / a(j,j)) x_j; j descending.
// Both read coefficients a(k, t), k > t (uniform across the wavefront).
// The products do not depend on the chain, so each batch of kChainU terms is
// loaded and multiplied ahead of its dependent subtractions.
constexpr int kChainThreads = 64;
constexpr int kChainU = 8;

template <bool R, bool NOUNIT>
__global__ __launch_bounds__(kChainThreads) void k_trsm_chain(const double2 *__restrict__ A, int64_t lda,
                                                              double2 *__restrict__ B, int64_t ldb, int64_t nvec,
                                                              int64_t s, ddv alpha, bool scale) {
  const int64_t v = (int64_t)blockIdx.x * kChainThreads + threadIdx.x;
  if (v >= nvec) return;
  auto xp = [&](int64_t t) -> double2 * { return R ? B + v + t * ldb : B + t + v * ldb; };
  for (int64_t t = s - 1; t >= 0; --t) {
    const double2 b0 = *xp(t);
    ddv x = mk(b0.x, b0.y);
    if (!R || scale) x = dd_mul(alpha, x);
    const double2 *col = A + t * lda;  // a(k, t)
    int64_t k = t + 1;
    for (; k + kChainU <= s; k += kChainU) {
      double2 c[kChainU], xv[kChainU];
#pragma unroll
      for (int u = 0; u < kChainU; ++u) {
        c[u] = col[k + u];
        xv[u] = *xp(k + u);
      }
      ddv pr[kChainU];
#pragma unroll
      for (int u = 0; u < kChainU; ++u) pr[u] = dd_mul(mk(c[u].x, c[u].y), mk(xv[u].x, xv[u].y));
#pragma unroll
      for (int u = 0; u < kChainU; ++u)
        if (!R || dd_nonzero(mk(c[u].x, c[u].y))) x = dd_sub(x, pr[u]);
    }
    for (; k < s; ++k) {
      const double2 c = col[k], xk = *xp(k);
      const ddv cd = mk(c.x, c.y);
      if (!R || dd_nonzero(cd)) x = dd_sub(x, dd_mul(cd, mk(xk.x, xk.y)));
    }
    if (NOUNIT) {
      const double2 d = col[t];
      x = R ? dd_mul(dd_div(mk(1.0, 0.0), mk(d.x, d.y)), x) : dd_div(x, mk(d.x, d.y));
    }
    *xp(t) = make_double2(x.hi, x.lo);
  }
}

bool is(char c, char u) { return c == u || c == u + ('a' - 'A'); }

template <bool R, bool CT, int FIN, bool ACTIVE_SKIP, bool COEF_SKIP, bool DESC>
hipError_t solve(const double2 *A, int64_t lda, double2 *B, int64_t ldb, int64_t m, int64_t n, ddv alpha,
                 bool post_scale, uint64_t *mask, hipStream_t st) {
  const int64_t s = R ? n : m, nvec = R ? m : n;
  for (int64_t q0 = 0; q0 < s; q0 += kT) {
    const int tb = (int)(s - q0 < kT ? s - q0 : kT);
    const int64_t base = DESC ? s - q0 - tb : q0;
    hipLaunchKernelGGL((k_trsm_diag<R, CT, FIN, ACTIVE_SKIP, COEF_SKIP, DESC>),
                       dim3((unsigned)((nvec + 3) / 4)), dim3(kThreads), 0, st, A, lda, B, ldb, nvec, base, tb,
                       mask);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    const int64_t rest = s - q0 - tb;  // later entries: [q0 + tb, s) or [0, rest)
    if (rest > 0) {
      const int64_t t0 = DESC ? 0 : q0 + tb;
      const int64_t r0 = R ? 0 : t0, nrows = R ? m : rest, c0 = R ? t0 : 0, ncols = R ? rest : n;
      const int64_t tiles_r = (nrows + 63) / 64, tiles_c = (ncols + 63) / 64;
      if (tiles_r * tiles_c > 0x7fffffff) return hipErrorInvalidValue;
      hipLaunchKernelGGL((k_trsm_update<R, CT, ACTIVE_SKIP, COEF_SKIP, DESC>), dim3((unsigned)(tiles_r * tiles_c)),
                         dim3(kThreads), 0, st, A, lda, B, ldb, r0, nrows, c0, ncols, base, tb, (int)tiles_r, mask);
      if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    }
    if (post_scale) {  // side R with transpose: x_p = alpha x_p after its updates
      if (hipError_t e = mfhip::scale_matrix(B + base * ldb, ldb, m, tb, alpha, st); e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

}  // namespace

namespace mfhip {

size_t trsm_workspace_bytes(char side, int64_t m, int64_t n) {
  (void)m;
  const int64_t nv = is(side, 'L') && n > 0 ? n : 1;
  return (size_t)((8 * nv + 255) / 256 * 256);
}

hipError_t trsm_check(char side, char uplo, char ta, char diag, int64_t m, int64_t n, int64_t lda, int64_t ldb) {
  const bool lside = is(side, 'L');
  const int64_t nrowa = lside ? m : n;
  if ((!lside && !is(side, 'R')) || (!is(uplo, 'U') && !is(uplo, 'L')) ||
      (!is(ta, 'N') && !is(ta, 'T') && !is(ta, 'C')) || (!is(diag, 'U') && !is(diag, 'N')) || m < 0 || n < 0 ||
      lda < (nrowa > 1 ? nrowa : 1) || ldb < (m > 1 ? m : 1))
    return hipErrorInvalidValue;
  return hipSuccess;
}

// Argument checks wtrsm.f90:24-47, quick returns 48-56, then the blocked solve.
hipError_t trsm(char side, char uplo, char ta, char diag, int64_t m, int64_t n, ddv alpha, const double2 *A,
                int64_t lda, double2 *B, int64_t ldb, void *work, size_t work_bytes, hipStream_t st) {
  if (hipError_t e = trsm_check(side, uplo, ta, diag, m, n, lda, ldb); e != hipSuccess) return e;
  if (m == 0 || n == 0) return hipSuccess;
  if (alpha.hi == 0.0 && alpha.lo == 0.0) return scale_matrix(B, ldb, m, n, alpha, st);  // zeros
  const bool R = is(side, 'R'), T = !is(ta, 'N'), upper = is(uplo, 'U'), unit = is(diag, 'U');
  const bool scale = alpha.hi != 1.0 || alpha.lo != 0.0;
  uint64_t *mask = (uint64_t *)work;
  if (!upper && (R != T)) {  // L/T/L and R/N/L: per-vector chains
    const int64_t nvec = R ? m : n, s = R ? n : m;
    const dim3 grid((unsigned)((nvec + kChainThreads - 1) / kChainThreads));
    auto kern = R ? (unit ? k_trsm_chain<true, false> : k_trsm_chain<true, true>)
                  : (unit ? k_trsm_chain<false, false> : k_trsm_chain<false, true>);
    hipLaunchKernelGGL(kern, grid, dim3(kChainThreads), 0, st, A, lda, B, ldb, nvec, s, alpha, scale);
    return hipGetLastError();
  }
  if (!R && !T && work_bytes < trsm_workspace_bytes(side, m, n)) return hipErrorInvalidValue;
  // pivot order: descending for L/N/U, L/T/L, R/N/L, R/T/U
  const bool desc = upper ^ T ^ R;
  if ((!R && !T && scale) || (!R && T) || (R && !T && scale))
    if (hipError_t e = scale_matrix(B, ldb, m, n, alpha, st); e != hipSuccess) return e;
#define MFHIP_TRSM(RR, CT, FIN, AS, CS)                                                                   \
  return desc ? solve<RR, CT, FIN, AS, CS, true>(A, lda, B, ldb, m, n, alpha, RR && T && scale, mask, st) \
              : solve<RR, CT, FIN, AS, CS, false>(A, lda, B, ldb, m, n, alpha, RR && T && scale, mask, st)
  if (!R && !T) {
    if (unit) MFHIP_TRSM(false, false, FIN_NONE, true, false);
    MFHIP_TRSM(false, false, FIN_DIV, true, false);
  }
  if (!R) {
    if (unit) MFHIP_TRSM(false, true, FIN_NONE, false, false);
    MFHIP_TRSM(false, true, FIN_DIV, false, false);
  }
  if (!T) {
    if (unit) MFHIP_TRSM(true, true, FIN_NONE, false, true);
    MFHIP_TRSM(true, true, FIN_RECIP, false, true);
  }
  if (unit) MFHIP_TRSM(true, false, FIN_NONE, false, true);
  MFHIP_TRSM(true, false, FIN_RECIP, false, true);
#undef MFHIP_TRSM
}

}  // namespace mfhip
