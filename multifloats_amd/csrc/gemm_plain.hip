// Plain (reference-order) double-double GEMM / GEMV kernels for gfx950.
//
// Semantics follow src/float64x2/matmul.inc:47-239 of the reference: every
// output C(i,j) is one sequential chain over p = 0..k-1 of mac() with A(i,p)
// as the "a" operand, a full TwoSum renormalisation after every complete chunk
// of `ri` MACs that is not the last (when ri > 0 and k > ri), then a final
// TwoSum. The reference's MR x NR blocking never changes that per-output
// order, so any kernel that keeps the chain sequential per output -- as these
// do -- returns the reference's bits.
//
// GEMM: a 256-thread workgroup owns a (16*TM) x (16*TN) tile of C; the tile
// body (LDS hi/lo staging, register-blocked MAC chains) is gemm_tile.h.
#include "dd_core.h"
#include "gemm_tile.h"
#include "mf_internal.h"

namespace {

constexpr int kThreads = mfhip_tile::kThreads;

template <int TM, int TN, int BK, int WPE, int R>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WPE))) void k_gemm_plain(
    const double2 *__restrict__ A, const double2 *__restrict__ B, double2 *__restrict__ C,
    int64_t m, int64_t k, int64_t n, int64_t ri, int tiles_m) {
  const int tile = blockIdx.x;
  const int64_t m0 = (int64_t)(tile % tiles_m) * (16 * TM), n0 = (int64_t)(tile / tiles_m) * (16 * TN);
  if constexpr (TM == 4 && TN == 4 && WPE == 2) {
    // 64 x 64 tiles at two waves per SIMD: operands copied straight into
    // two LDS stages (dynamic LDS, kDirectLds bytes)
    static_assert(BK == mfhip_tile::kDirectBK, "K block of the direct tile");
    extern __shared__ __attribute__((aligned(16))) char dyn_lds[];
    mfhip_tile::plain_tile_direct<R>(A, B, C, m, k, n, ri, m0, n0, dyn_lds);
  } else if constexpr (BK == mfhip_tile::kDirectBK) {
    // the other shapes with 16-k blocks: the general direct-copy tile
    extern __shared__ __attribute__((aligned(16))) char dyn_lds[];
    mfhip_tile::grouped_tile_direct<TM, TN, R>(A, m, B, k, C, m, m, k, n, ri, m0, n0, dyn_lds);
  } else {
    __shared__ double lds[2 * BK * (16 * TM + 16 * TN)];
    mfhip_tile::plain_tile<TM, TN, BK, R>(A, B, C, m, k, n, ri, m0, n0, lds);
  }
}

// BLAS-style C = alpha op(A) op(B) + beta C with leading dimensions
// (op = transpose when TA / TB); the chains are the matmuldd_mm chains of
// op(A) op(B), the epilogue applies alpha and beta with the DD operators.
template <int TM, int TN, int BK, int WPE, int R, bool TA, bool TB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WPE))) void k_gemm_blas(
    const double2 *__restrict__ A, int64_t lda, const double2 *__restrict__ B, int64_t ldb,
    double2 *__restrict__ C, int64_t ldc, int64_t m, int64_t k, int64_t n, int64_t ri, ddv alpha, ddv beta,
    int tiles_m) {
  const int tile = blockIdx.x;
  const int64_t m0 = (int64_t)(tile % tiles_m) * (16 * TM), n0 = (int64_t)(tile / tiles_m) * (16 * TN);
  if constexpr (BK == mfhip_tile::kDirectBK) {
    // the direct-copy tile for every transpose pair (dynamic LDS,
    // DirectShape<TM, TN, TA, TB>::kLds bytes)
    extern __shared__ __attribute__((aligned(16))) char dyn_lds[];
    mfhip_tile::grouped_tile_direct<TM, TN, R, true, TA, TB>(A, lda, B, ldb, C, ldc, m, k, n, ri, m0, n0, dyn_lds,
                                                             alpha, beta);
  } else {
    __shared__ double lds[2 * BK * (16 * TM + 16 * TN)];
    mfhip_tile::gemm_tile<TM, TN, BK, R, TA, TB, true>(A, lda, B, ldb, C, ldc, m, k, n, ri, m0, n0, lds, alpha,
                                                        beta);
  }
}

// C = beta C (beta == 0: zeros, C not read) for the BLAS quick paths.
__global__ __launch_bounds__(kThreads) void k_scale_c(double2 *__restrict__ C, int64_t ldc, int64_t m, int64_t n,
                                                      ddv beta) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= m * n) return;
  const int64_t i = t % m, j = t / m;
  double2 *c = C + i + j * ldc;
  if (beta.hi == 0.0 && beta.lo == 0.0) {
    *c = make_double2(0.0, 0.0);
  } else {
    const ddv r = dd_mul(beta, mk(c->x, c->y));
    *c = make_double2(r.hi, r.lo);
  }
}

// p[0..n) = 0 (the k = 0 products): a fill kernel rather than a memset node,
// so the call stays safe to record into a HIP graph.
__global__ __launch_bounds__(kThreads) void k_zero(double *__restrict__ p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

hipError_t zero_doubles(double *p, int64_t n, hipStream_t s) {
  const int64_t g = (n + kThreads - 1) / kThreads;
  if (g > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)g), dim3(kThreads), 0, s, p, n);
  return hipGetLastError();
}

// y(m) = A(m,k) x(k): one thread per row, A read down its columns (coalesced),
// x broadcast. Reference matmul.inc:207-211 via gaxpy_mv_panel (47-88).
__global__ __launch_bounds__(kThreads) void k_gemv(const double2 *__restrict__ A,
                                                   const double2 *__restrict__ x,
                                                   double2 *__restrict__ y, int64_t m, int64_t k,
                                                   int64_t ri) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= m) return;
  const int64_t rint = (ri > 0 && k > ri) ? ri : 0;
  double sh = 0.0, sl = 0.0;
  int64_t until = rint;
  for (int64_t p = 0; p < k; ++p) {
    double2 a = A[i + p * m];
    double2 b = x[p];
    mac(a.x, a.y, b.x, b.y, sh, sl);
    if (rint && --until == 0) {
      until = rint;
      if (p + 1 < k) renorm(sh, sl);
    }
  }
  renorm(sh, sl);
  y[i] = make_double2(sh, sl);
}

// y(n) = x(k) B(k,n): x is the "a" operand (reference matmul.inc:213-239).
// One thread per column; an 8-deep slab of B is transposed through LDS so the
// HBM reads stay coalesced along B's columns.
constexpr int kVmK = 8;
__global__ __launch_bounds__(kThreads) void k_gevm(const double2 *__restrict__ x,
                                                   const double2 *__restrict__ B,
                                                   double2 *__restrict__ y, int64_t k, int64_t n,
                                                   int64_t ri) {
  __shared__ double Bh[kVmK][kThreads], Bl[kVmK][kThreads];
  const int tid = threadIdx.x;
  const int64_t n0 = (int64_t)blockIdx.x * kThreads;
  const int64_t j = n0 + tid;
  const int64_t rint = (ri > 0 && k > ri) ? ri : 0;
  double sh = 0.0, sl = 0.0;
  int64_t until = rint, done = 0;
  for (int64_t p0 = 0; p0 < k; p0 += kVmK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kVmK; ++q) {
      int e = tid + q * kThreads;
      int r = e % kVmK, c = e / kVmK;
      int64_t gp = p0 + r, gj = n0 + c;
      double2 v = (gp < k && gj < n) ? B[gp + gj * k] : make_double2(0.0, 0.0);
      Bh[r][c] = v.x;
      Bl[r][c] = v.y;
    }
    __syncthreads();
    const int kmax = (k - p0) < kVmK ? (int)(k - p0) : kVmK;
    for (int r = 0; r < kmax; ++r) {
      double2 a = x[p0 + r];
      mac(a.x, a.y, Bh[r][tid], Bl[r][tid], sh, sl);
      ++done;
      if (rint && --until == 0) {
        until = rint;
        if (done < k) renorm(sh, sl);
      }
    }
  }
  if (j < n) {
    renorm(sh, sl);
    y[j] = make_double2(sh, sl);
  }
}

// Complex C(m,n) = A(m,k) B(k,n) on complex64x2 {re.hi, re.lo, im.hi, im.lo}.
// The reference forms it from four real GEMMs and the DD - and + operators
// (fsrc/multifloats.fypp:2609-2621):
//   re = mm(a_re, b_re) - mm(a_im, b_im),  im = mm(a_re, b_im) + mm(a_im, b_re).
// Here one pass runs those four sequential chains per output side by side
// (same MAC operand order, same renormalisation schedule), then applies the
// same subtraction and addition: identical bits, inputs read once, no
// temporaries. mv and vm are the n = 1 and m = 1 cases (fypp:2623-2649).
//
// A 32 x 32 tile per 256-lane workgroup, four chains per output and two
// outputs per lane in each direction (16 chains). Operands are copied
// straight from global memory into two LDS stages of 16 k (the plain 64 x 64
// kernel's scheme): an A column of 32 complex elements is one 1 KiB copy,
// two 16-k B columns are one, the next block's copies run during the current
// block's MACs, one barrier per block. Lane (tx, ty) owns rows tx + 16 i and
// columns 2 ty + j; its A reads take the real pair of rows 0-7 and the
// imaginary pair of rows 8-15 first (and the other halves second), so the
// 16 rows of a read fall on distinct banks. Copy addresses are clamped into
// the operands (edge rows and columns are not stored; k past the end is never
// taken). With ri = 8 (R) the renormalisation points are static within a
// block, as in the real kernels.
constexpr int kCplxBK = 16;
constexpr int kCplxStageA = kCplxBK * 1024;
constexpr int kCplxStageB = mfhip_tile::direct_b_base(15) + 1024;
constexpr int kCplxStage = kCplxStageA + kCplxStageB;
constexpr int kCplxLds = 2 * kCplxStage;
template <int R>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_cplx(
    const double4 *__restrict__ A, const double4 *__restrict__ B, double4 *__restrict__ C,
    int64_t m, int64_t k, int64_t n, int64_t ri, int tiles_m) {
  constexpr int BK = kCplxBK, TM = 2, TN = 2;
  extern __shared__ __attribute__((aligned(16))) char lds[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tx = tid & 15, ty = tid >> 4;
  const int tile = blockIdx.x;
  const int64_t m0 = (int64_t)(tile % tiles_m) * 32;
  const int64_t n0 = (int64_t)(tile / tiles_m) * 32;
  const int64_t nb = (k + BK - 1) / BK;
  // A copy: lane = (row lane / 2, half lane % 2); B copy q: (column 2q + lane / 32, k (lane / 2) % 16, half)
  const int64_t ar = m0 + (lane >> 1) < m ? m0 + (lane >> 1) : m - 1;
  const int bcol = lane >> 5, bkk = (lane >> 1) & 15, half = lane & 1;
  auto issue = [&](int64_t kb, int buf) {
    char *st = lds + buf * kCplxStage;
#pragma unroll
    for (int c = wave; c < BK; c += 4) {
      const int64_t p = kb * BK + c < k ? kb * BK + c : k - 1;
      __builtin_amdgcn_global_load_lds((const void *)((const char *)(A + ar + p * m) + 16 * (lane & 1)),
                                       (__attribute__((address_space(3))) void *)(st + c * 1024), 16, 0, 0);
    }
#pragma unroll
    for (int q = wave; q < 16; q += 4) {
      const int64_t col = n0 + 2 * q + bcol < n ? n0 + 2 * q + bcol : n - 1;
      const int64_t p = kb * BK + bkk < k ? kb * BK + bkk : k - 1;
      __builtin_amdgcn_global_load_lds((const void *)((const char *)(B + p + col * k) + 16 * half),
                                       (__attribute__((address_space(3))) void *)(st + kCplxStageA +
                                                                                   mfhip_tile::direct_b_base(q)),
                                       16, 0, 0);
    }
  };

  // chain 0: ar*br, 1: ai*bi, 2: ar*bi, 3: ai*br
  double sh[4][TM][TN], sl[4][TM][TN];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) sh[c][i][j] = sl[c][i][j] = 0.0;
  const int64_t rint = (ri > 0 && k > ri) ? ri : 0;
  int64_t until = rint, done = 0;
  const int hf = (tx >> 3) & 1;  // which half of a row this lane reads first
  auto macs = [&](const char *st, int kk) {
    double arh[TM], arl[TM], aih[TM], ail[TM], brh[TN], brl[TN], bih[TN], bil[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const char *e = st + kk * 1024 + (tx + 16 * i) * 32;
      const double2 v0 = *(const double2 *)(e + 16 * hf);
      const double2 v1 = *(const double2 *)(e + 16 * (1 - hf));
      const double2 re = hf ? v1 : v0, im = hf ? v0 : v1;
      arh[i] = re.x;
      arl[i] = re.y;
      aih[i] = im.x;
      ail[i] = im.y;
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      // column 2 ty + j: copy ty, column j of the copy
      const char *e = st + kCplxStageA + mfhip_tile::direct_b_base(ty) + j * 512 + kk * 32;
      const double2 re = *(const double2 *)e, im = *(const double2 *)(e + 16);
      brh[j] = re.x;
      brl[j] = re.y;
      bih[j] = im.x;
      bil[j] = im.y;
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        mac(arh[i], arl[i], brh[j], brl[j], sh[0][i][j], sl[0][i][j]);
        mac(aih[i], ail[i], bih[j], bil[j], sh[1][i][j], sl[1][i][j]);
        mac(arh[i], arl[i], bih[j], bil[j], sh[2][i][j], sl[2][i][j]);
        mac(aih[i], ail[i], brh[j], brl[j], sh[3][i][j], sl[3][i][j]);
      }
  };
  auto renorm_all = [&]() {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) renorm(sh[c][i][j], sl[c][i][j]);
  };

  if (nb > 0) issue(0, 0);
  for (int64_t kb = 0; kb < nb; ++kb) {
    // this wave's copies of block kb have landed, every wave's after the
    // barrier, and every wave is done with the other buffer
    __syncthreads();
    if (kb + 1 < nb) issue(kb + 1, (int)((kb + 1) & 1));
    const char *st = lds + (int)(kb & 1) * kCplxStage;
    const int64_t k0 = kb * BK;
    const int kmax = (int)((k - k0) < BK ? (k - k0) : BK);
    if (R > 0 && rint == R && k0 + BK < k) {
      constexpr int RR = R > 0 ? R : 1;
      static_assert(R == 0 || BK % RR == 0, "static renorm interval divides the K block");
#pragma unroll 1
      for (int c = 0; c < BK; c += RR) {
#pragma unroll
        for (int kk = 0; kk < RR; ++kk) macs(st, c + kk);
        renorm_all();
      }
      done += BK;
    } else {
#pragma unroll 1
      for (int kk = 0; kk < kmax; ++kk) {
        macs(st, kk);
        ++done;
        if (rint && --until == 0) {
          until = rint;
          if (done < k) renorm_all();
        }
      }
    }
  }

#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gj = n0 + 2 * ty + j;
    if (gj >= n) continue;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int64_t gi = m0 + tx + 16 * i;
      if (gi >= m) continue;
      ddv p[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double h = sh[c][i][j], l = sl[c][i][j];
        renorm(h, l);
        p[c] = mk(h, l);
      }
      const ddv re = dd_sub(p[0], p[1]);
      const ddv im = dd_add(p[2], p[3]);
      C[gi + gj * m] = make_double4(re.hi, re.lo, im.hi, im.lo);
    }
  }
}

template <int TM, int TN, int BK, int WPE>
hipError_t launch_gemm(const double2 *A, const double2 *B, double2 *C, int64_t m, int64_t k,
                       int64_t n, int64_t ri, hipStream_t s) {
  constexpr int BM = 16 * TM, BN = 16 * TN;
  const int64_t tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
  if (tm * tn > 0x7fffffff) return hipErrorInvalidValue;
  const bool r8 = ri == 8 && k > 8;
  constexpr bool direct = BK == mfhip_tile::kDirectBK;
  constexpr int dyn = !direct ? 0
                      : (TM == 4 && TN == 4 && WPE == 2) ? mfhip_tile::kDirectLds
                                                          : mfhip_tile::DirectShape<TM, TN>::kLds;
  static_assert(dyn * WPE <= 160 * 1024, "the workgroups per CU fit in LDS");
  if (direct && dyn > 65536) {
    // above the 64 KiB default: allowed on every call (any device current)
    const void *kern = r8 ? (const void *)k_gemm_plain<TM, TN, BK, WPE, 8> : (const void *)k_gemm_plain<TM, TN, BK, WPE, 0>;
    const hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    if (e != hipSuccess) return e;
  }
  if (r8)
    hipLaunchKernelGGL((k_gemm_plain<TM, TN, BK, WPE, 8>), dim3((unsigned)(tm * tn)), dim3(kThreads), dyn,
                       s, A, B, C, m, k, n, ri, (int)tm);
  else
    hipLaunchKernelGGL((k_gemm_plain<TM, TN, BK, WPE, 0>), dim3((unsigned)(tm * tn)), dim3(kThreads), dyn,
                       s, A, B, C, m, k, n, ri, (int)tm);
  return hipGetLastError();
}

}  // namespace

namespace mfhip {

int gemm_plain_variants() { return 8; }

// Tile rule. Every output is one sequential chain, so a call has exactly m*n
// chains of work to spread: 64x64 tiles (16 chains per lane, the most
// in-flight work per wave) once they give at least 1024 workgroups (4 per
// CU), else 64x32 tiles (8 chains, 3 waves per SIMD) once those give 512,
// else 32x32 tiles once those give 512, else 16x16 tiles (one chain per
// lane; 256^2 outputs then fill all 256 CUs instead of 16). Measured on
// MI355X (m = n = k): 64x64 best from 2048, 64x32 at 1024 (630 us against
// 699 us for 32x32), 16x16 up to 512.
int plain_variant_for(int64_t m, int64_t n) {
  const int64_t t64 = ((m + 63) / 64) * ((n + 63) / 64), t32 = ((m + 31) / 32) * ((n + 31) / 32);
  const int64_t t6432 = ((m + 63) / 64) * ((n + 31) / 32);
  if (t64 >= 1024) return 1;
  if (t6432 >= 512) return 2;
  if (t32 >= 512) return 0;
  return 7;
}

hipError_t gemm_plain(const double2 *A, const double2 *B, double2 *C, int64_t m, int64_t k,
                      int64_t n, int64_t ri, int variant, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (k <= 0) return zero_doubles((double *)C, 2 * m * n, s);
  if (variant < 0) variant = plain_variant_for(m, n);
  switch (variant) {
    case 0: return launch_gemm<2, 2, 16, 4>(A, B, C, m, k, n, ri, s);
    case 1: return launch_gemm<4, 4, 16, 2>(A, B, C, m, k, n, ri, s);
    case 2: return launch_gemm<4, 2, 16, 3>(A, B, C, m, k, n, ri, s);
    case 3: return launch_gemm<4, 4, 16, 1>(A, B, C, m, k, n, ri, s);
    case 4: return launch_gemm<4, 8, 16, 1>(A, B, C, m, k, n, ri, s);
    case 5: return launch_gemm<8, 4, 16, 1>(A, B, C, m, k, n, ri, s);
    case 6: return launch_gemm<2, 4, 16, 3>(A, B, C, m, k, n, ri, s);
    default: {
      // 16 x 16 tiles: with four or more workgroups per CU the direct copies
      // (16-k blocks) hide their latency behind the other workgroups; with
      // fewer, register staging with K blocks of 64 and the hand-pipelined
      // chain (one chain per lane) is faster (256^3: 17.0-17.4 us against
      // 32-33 us direct; 384^3: 55.0 against 57.1 us; 512^3: direct 2 %
      // faster). Same chains, same bits.
      const int64_t t16 = ((m + 15) / 16) * ((n + 15) / 16);
      if (t16 >= 1024) return launch_gemm<1, 1, 16, 4>(A, B, C, m, k, n, ri, s);
      return launch_gemm<1, 1, 64, 4>(A, B, C, m, k, n, ri, s);
    }
  }
}

namespace {
template <int TM, int TN, int WPE, bool TA, bool TB, int BKT = 16>
hipError_t launch_blas(const double2 *A, int64_t lda, const double2 *B, int64_t ldb, double2 *C, int64_t ldc,
                       int64_t m, int64_t k, int64_t n, int64_t ri, ddv alpha, ddv beta, hipStream_t s) {
  constexpr int BK = BKT;
  constexpr int BM = 16 * TM, BN = 16 * TN;
  const int64_t tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
  if (tm * tn > 0x7fffffff) return hipErrorInvalidValue;
  constexpr int dyn = BK == mfhip_tile::kDirectBK ? mfhip_tile::DirectShape<TM, TN, TA, TB>::kLds : 0;
  static_assert(dyn * WPE <= 160 * 1024, "the workgroups per CU fit in LDS");
  if (dyn > 65536) {  // above the 64 KiB default: allowed on every call (any device current)
    const void *kern = (ri == 8 && k > 8) ? (const void *)k_gemm_blas<TM, TN, BK, WPE, 8, TA, TB>
                                          : (const void *)k_gemm_blas<TM, TN, BK, WPE, 0, TA, TB>;
    const hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    if (e != hipSuccess) return e;
  }
  if (ri == 8 && k > 8)
    hipLaunchKernelGGL((k_gemm_blas<TM, TN, BK, WPE, 8, TA, TB>), dim3((unsigned)(tm * tn)), dim3(kThreads), dyn,
                       s, A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, (int)tm);
  else
    hipLaunchKernelGGL((k_gemm_blas<TM, TN, BK, WPE, 0, TA, TB>), dim3((unsigned)(tm * tn)), dim3(kThreads), dyn,
                       s, A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, (int)tm);
  return hipGetLastError();
}
template <bool TA, bool TB>
hipError_t blas_by_size(const double2 *A, int64_t lda, const double2 *B, int64_t ldb, double2 *C, int64_t ldc,
                        int64_t m, int64_t k, int64_t n, int64_t ri, ddv alpha, ddv beta, hipStream_t s) {
  switch (plain_variant_for(m, n)) {
    case 1: return launch_blas<4, 4, 2, TA, TB>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
    case 2: return launch_blas<4, 2, 3, TA, TB>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
    case 0: return launch_blas<2, 2, 4, TA, TB>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
    default:
      // 16 x 16 tiles below 512 of them: register staging with 64-k blocks
      // and the hand-pipelined chain (one workgroup per CU hides the direct
      // copies' latency poorly; 256^3: 13.7-19.8 us per transpose pair
      // against 14.8-20.9 with 32-k blocks; at 384^3 the direct copies stay
      // ahead for op(A) = A^T, op(B) = B)
      if (((m + 15) / 16) * ((n + 15) / 16) < 512)
        return launch_blas<1, 1, 4, TA, TB, 64>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
      return launch_blas<1, 1, 4, TA, TB>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
  }
}
bool trans_ok(char t) { return t == 'N' || t == 'n' || t == 'T' || t == 't' || t == 'C' || t == 'c'; }
}  // namespace

hipError_t scale_matrix(double2 *C, int64_t ldc, int64_t m, int64_t n, ddv beta, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_c, dim3((unsigned)((m * n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, C, ldc,
                     m, n, beta);
  return hipGetLastError();
}

// BLAS xGEMM argument rules (order of the reference's wgemm shim,
// test/integration/blas/wgemm.f90:16-51) and quick returns (52-69).
hipError_t gemm_blas_check(char ta, char tb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                           int64_t ldc) {
  const bool nota = ta == 'N' || ta == 'n', notb = tb == 'N' || tb == 'n';
  const int64_t nrowa = nota ? m : k, nrowb = notb ? k : n;
  if (!trans_ok(ta) || !trans_ok(tb) || m < 0 || n < 0 || k < 0 || lda < (nrowa > 1 ? nrowa : 1) ||
      ldb < (nrowb > 1 ? nrowb : 1) || ldc < (m > 1 ? m : 1))
    return hipErrorInvalidValue;
  return hipSuccess;
}

hipError_t gemm_blas(char ta, char tb, int64_t m, int64_t n, int64_t k, ddv alpha, const double2 *A,
                     int64_t lda, const double2 *B, int64_t ldb, ddv beta, double2 *C, int64_t ldc, int64_t ri,
                     hipStream_t s) {
  if (hipError_t e = gemm_blas_check(ta, tb, m, n, k, lda, ldb, ldc); e != hipSuccess) return e;
  const bool nota = ta == 'N' || ta == 'n', notb = tb == 'N' || tb == 'n';
  const bool alpha0 = alpha.hi == 0.0 && alpha.lo == 0.0;
  const bool beta1 = beta.hi == 1.0 && beta.lo == 0.0;
  if (m == 0 || n == 0 || ((alpha0 || k == 0) && beta1)) return hipSuccess;
  if (alpha0 || k == 0) return scale_matrix(C, ldc, m, n, beta, s);
  if (nota && notb) return blas_by_size<false, false>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
  if (nota) return blas_by_size<false, true>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
  if (notb) return blas_by_size<true, false>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
  return blas_by_size<true, true>(A, lda, B, ldb, C, ldc, m, k, n, ri, alpha, beta, s);
}

hipError_t gemm_cplx(const double4 *A, const double4 *B, double4 *C, int64_t m, int64_t k,
                     int64_t n, int64_t ri, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (k <= 0) return zero_doubles((double *)C, 4 * m * n, s);
  const int64_t tm = (m + 31) / 32, tn = (n + 31) / 32;
  if (tm * tn > 0x7fffffff) return hipErrorInvalidValue;
  // above the 64 KiB default: allowed on every call (any device current)
  const hipError_t e = hipFuncSetAttribute((ri == 8 && k > 8) ? (const void *)k_gemm_cplx<8> : (const void *)k_gemm_cplx<0>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kCplxLds);
  if (e != hipSuccess) return e;
  if (ri == 8 && k > 8)
    hipLaunchKernelGGL(k_gemm_cplx<8>, dim3((unsigned)(tm * tn)), dim3(kThreads), kCplxLds, s, A, B, C, m, k, n, ri,
                       (int)tm);
  else
    hipLaunchKernelGGL(k_gemm_cplx<0>, dim3((unsigned)(tm * tn)), dim3(kThreads), kCplxLds, s, A, B, C, m, k, n, ri,
                       (int)tm);
  return hipGetLastError();
}

hipError_t gemv(const double2 *A, const double2 *x, double2 *y, int64_t m, int64_t k, int64_t ri,
                hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gemv, dim3((unsigned)((m + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     s, A, x, y, m, k, ri);
  return hipGetLastError();
}

hipError_t gevm(const double2 *x, const double2 *B, double2 *y, int64_t k, int64_t n, int64_t ri,
                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gevm, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     s, x, B, y, k, n, ri);
  return hipGetLastError();
}

}  // namespace mfhip
