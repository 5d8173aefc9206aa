// Ozaki-style split-precision DD GEMM on the INT8 matrix cores of gfx950.
//
// The precision form of the reference's matmuldd_mm (src/float64x2/matmul.inc):
// same inputs and layout, results held to an error bound instead of the
// reference's bits.
//
// 1. Scaling. E_i = max_p ilogb(A(i,p).hi) per row of A, F_j likewise per
//    column of B. x = a * 2^(5 - E_i) then satisfies |x| <= 64.
// 2. Slicing. x = sum_s d_s 2^(-7 s) with balanced digits d_s = RN(x_s) in
//    [-64, 64], x_{s+1} = 128 (x_s - d_s). Every step is exact in DD
//    arithmetic, so the S digits are the leading 7S bits of x with round-to-
//    nearest digit boundaries. Digits are written as int8 in a tiled layout:
//    one contiguous (S x 64 x 32)-byte chunk per (64-row block, 32-k block).
// 3. Products. With G_L = sum_p sum_{s+t=L} d^A_s d^B_t (exact in int32:
//    |d d| <= 4096, so k * S * 4096 < 2^31 for k <= 32768, S <= 16),
//        C_ij = 2^(E_i + F_j - 10) * sum_{L < S} 2^(-7 L) G_L(i,j).
//    Levels L >= S are dropped (the triangular Ozaki truncation); each G_L is
//    accumulated by v_mfma_i32_32x32x32_i8 with exact INT32 accumulation.
// 4. Recombination. The S integers of an output are converted exactly to
//    double, scaled by exact powers of two and summed with accurate DD
//    additions (two_sum), then scaled by 2^(E_i + F_j - 10).
//
// Kernel shape. A workgroup owns a 64x64 tile of C and all S level
// accumulators of it: 4 waves, each holding up to 4 levels of the whole tile
// (4 x 4 x 16 = 256 accumulator registers per lane), levels assigned to waves
// by longest-processing-time so each wave issues the same number of MFMAs.
// Per 32-deep K block the S slices of both operand tiles (S x 4 KiB) are
// staged once in LDS (48-byte row pitch: conflict-free 16-byte fragment
// reads) and every slice pair of a wave's levels is multiplied out of LDS, so
// HBM/L2 see each slice tile once per output tile rather than once per pair.
#include "dd_core.h"
#include "mf_internal.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kT = 64;        // output tile edge (rows of A / cols of B per block)
constexpr int kK = 32;        // K block (one MFMA K step)
constexpr int kChunk = kT * kK;  // bytes of one slice tile (64 x 32)
constexpr int kPitch = 48;    // LDS bytes per tile row
constexpr int kLdsSlice = kT * kPitch;
constexpr int kThreads = 256;
constexpr int kEmpty = -2147483647 - 1;  // exponent of an all-zero row/column
constexpr int kNonFinite = 2147483647;   // a row/column holding Inf or NaN

__device__ __forceinline__ int dexp(double x) {
  int e;
  (void)__builtin_frexp(x, &e);
  return e - 1;
}

// Row exponents of A (column-major m x k): thread = row, block.y = k chunk.
__global__ __launch_bounds__(kThreads) void k_row_exp(const double2 *__restrict__ A, int64_t m,
                                                      int64_t k, int64_t kchunk,
                                                      int *__restrict__ E) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= m) return;
  const int64_t p0 = (int64_t)blockIdx.y * kchunk;
  const int64_t p1 = p0 + kchunk < k ? p0 + kchunk : k;
  int e = kEmpty;
  bool bad = false;
  for (int64_t p = p0; p < p1; ++p) {
    double h = A[i + p * m].x;
    if (!mf_isfinite(h)) bad = true;
    else if (h != 0.0) e = max(e, dexp(h));
  }
  if (bad) e = kNonFinite;
  if (e != kEmpty) atomicMax(&E[i], e);
}

// Column exponents of B (column-major k x n): block = column, threads walk p.
__global__ __launch_bounds__(kThreads) void k_col_exp(const double2 *__restrict__ B, int64_t k,
                                                      int *__restrict__ F) {
  const int64_t j = blockIdx.x;
  int e = kEmpty;
  bool bad = false;
  for (int64_t p = threadIdx.x; p < k; p += kThreads) {
    double h = B[p + j * k].x;
    if (!mf_isfinite(h)) bad = true;
    else if (h != 0.0) e = max(e, dexp(h));
  }
  if (bad) e = kNonFinite;
  __shared__ int red[kThreads];
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) F[j] = red[0];
}

__global__ void k_fill_int(int *__restrict__ p, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Balanced base-128 digits of x (|x| <= 64, exact DD), MAXS of them.
template <int MAXS>
__device__ __forceinline__ void digits(double h, double l, int S, signed char *d) {
#pragma unroll
  for (int s = 0; s < MAXS; ++s) {
    if (s < S) {
      double q = __builtin_rint(h);
      d[s] = (signed char)(int)q;
      h = h - q;  // exact
      fast_two_sum(h, l, h, l);
      h *= 128.0;
      l *= 128.0;
    }
  }
}

__device__ __forceinline__ int scale_of(int e) { return (e == kEmpty || e == kNonFinite) ? 0 : e; }

// Slice a 64-row x 32-k tile of A. Element (r, c) of the tile is A(i0+r, p0+c);
// tile bytes of slice s go to out[((ib*KB + kb)*S + s)*2048 + r*32 + c].
template <int MAXS>
__global__ __launch_bounds__(kThreads) void k_slice_a(const double2 *__restrict__ A, int64_t m,
                                                      int64_t k, const int *__restrict__ E,
                                                      int S, int64_t KB,
                                                      signed char *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) signed char tile[MAXS * kChunk];
  const int ib = blockIdx.x, kb = blockIdx.y;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kChunk / kThreads; ++q) {
    int e = tid + q * kThreads;
    int r = e % kT, c = e / kT;  // consecutive threads walk down a column of A
    int64_t i = (int64_t)ib * kT + r, p = (int64_t)kb * kK + c;
    double h = 0.0, l = 0.0;
    int ex = 0;
    if (i < m && p < k) {
      double2 v = A[i + p * m];
      int Ei = E[i];
      if (Ei != kEmpty && Ei != kNonFinite) {
        h = v.x;
        l = v.y;
        ex = 5 - Ei;
      }
    }
    h = __builtin_ldexp(h, ex);
    l = __builtin_ldexp(l, ex);
    signed char d[MAXS];
    digits<MAXS>(h, l, S, d);
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < S) tile[s * kChunk + r * kK + c] = d[s];
  }
  __syncthreads();
  const int nvec = S * kChunk / 16;
  v4i *dst = (v4i *)(out + ((int64_t)ib * KB + kb) * S * kChunk);
  const v4i *src = (const v4i *)tile;
  for (int v = tid; v < nvec; v += kThreads) dst[v] = src[v];
}

// Slice a 32-k x 64-col tile of B; bytes of slice t at [c*32 + p] (k fastest).
template <int MAXS>
__global__ __launch_bounds__(kThreads) void k_slice_b(const double2 *__restrict__ B, int64_t k,
                                                      int64_t n, const int *__restrict__ F,
                                                      int S, int64_t KB,
                                                      signed char *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) signed char tile[MAXS * kChunk];
  const int jb = blockIdx.x, kb = blockIdx.y;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kChunk / kThreads; ++q) {
    int e = tid + q * kThreads;
    int p = e % kK, c = e / kK;  // consecutive threads walk down a column of B
    int64_t gp = (int64_t)kb * kK + p, j = (int64_t)jb * kT + c;
    double h = 0.0, l = 0.0;
    int ex = 0;
    if (gp < k && j < n) {
      double2 v = B[gp + j * k];
      int Fj = F[j];
      if (Fj != kEmpty && Fj != kNonFinite) {
        h = v.x;
        l = v.y;
        ex = 5 - Fj;
      }
    }
    h = __builtin_ldexp(h, ex);
    l = __builtin_ldexp(l, ex);
    signed char d[MAXS];
    digits<MAXS>(h, l, S, d);
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < S) tile[s * kChunk + c * kK + p] = d[s];
  }
  __syncthreads();
  const int nvec = S * kChunk / 16;
  v4i *dst = (v4i *)(out + ((int64_t)jb * KB + kb) * S * kChunk);
  const v4i *src = (const v4i *)tile;
  for (int v = tid; v < nvec; v += kThreads) dst[v] = src[v];
}

// Level -> wave assignment: longest processing time first (level L has L+1
// slice pairs), at most 4 levels per wave.
template <int S>
struct Plan {
  int lv[4][4];
  int cnt[4];
  constexpr Plan() : lv{}, cnt{} {
    int load[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w)
      for (int j = 0; j < 4; ++j) lv[w][j] = -1;
    for (int L = S - 1; L >= 0; --L) {
      int best = -1;
      for (int w = 0; w < 4; ++w)
        if (cnt[w] < 4 && (best < 0 || load[w] < load[best])) best = w;
      lv[best][cnt[best]++] = L;
      load[best] += L + 1;
    }
  }
};

__device__ __forceinline__ v4i lds_frag(const signed char *base, int row, int half) {
  return *(const v4i *)(base + row * kPitch + half * 16);
}

template <int S, int W>
__device__ __forceinline__ void compute_block(const signed char *sa, const signed char *sb,
                                              v16i (&acc)[4][2][2], int lane) {
  constexpr Plan<S> P{};
  const int r = lane & 31, half = lane >> 5;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    bool any = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int L = P.lv[W][j];
      any = any || (L >= s && L - s < S);
    }
    if (!any) continue;
    const v4i a0 = lds_frag(sa + s * kLdsSlice, r, half);
    const v4i a1 = lds_frag(sa + s * kLdsSlice, r + 32, half);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int L = P.lv[W][j];
      const int t = L - s;
      if (L >= 0 && t >= 0 && t < S) {
        const v4i b0 = lds_frag(sb + t * kLdsSlice, r, half);
        const v4i b1 = lds_frag(sb + t * kLdsSlice, r + 32, half);
        acc[j][0][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b0, acc[j][0][0], 0, 0, 0);
        acc[j][0][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b1, acc[j][0][1], 0, 0, 0);
        acc[j][1][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b0, acc[j][1][0], 0, 0, 0);
        acc[j][1][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b1, acc[j][1][1], 0, 0, 0);
      }
    }
  }
}

// Sum of this wave's levels for one accumulator slot position, in DD,
// smallest-scale level first.
template <int S, int W>
__device__ __forceinline__ ddv wave_partial(const v16i (&acc)[4][2][2], int bi, int bj, int r) {
  constexpr Plan<S> P{};
  ddv sum = mk(0.0, 0.0);
#pragma unroll
  for (int j = 3; j >= 0; --j) {
    const int L = P.lv[W][j];
    if (L >= 0) {
      double v = __builtin_ldexp((double)acc[j][bi][bj][r], -7 * L);
      sum = dd_add(sum, mk(v, 0.0));
    }
  }
  return sum;
}

template <int S>
__global__ __launch_bounds__(kThreads) void k_ozaki(const signed char *__restrict__ As,
                                                    const signed char *__restrict__ Bs,
                                                    const int *__restrict__ E,
                                                    const int *__restrict__ F,
                                                    double2 *__restrict__ C, int64_t m, int64_t n,
                                                    int64_t KB, int tiles_m) {
  extern __shared__ __attribute__((aligned(16))) signed char lds[];
  signed char *sa = lds;
  signed char *sb = lds + S * kLdsSlice;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ib = blockIdx.x % tiles_m, jb = blockIdx.x / tiles_m;
  const v4i *ga = (const v4i *)(As + (int64_t)ib * KB * S * kChunk);
  const v4i *gb = (const v4i *)(Bs + (int64_t)jb * KB * S * kChunk);
  constexpr int V = S * kChunk / 16;  // 16-byte vectors per operand chunk
  constexpr int PER = (V + kThreads - 1) / kThreads;

  v16i acc[4][2][2];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[j][a][b] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

  v4i pa[PER], pb[PER];
  auto fetch = [&](int64_t kb) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      int v = tid + q * kThreads;
      if (v < V) {
        pa[q] = ga[kb * V + v];
        pb[q] = gb[kb * V + v];
      }
    }
  };
  // 16-byte vector v of a chunk: slice v/128, tile row (v%128)/2, half v%2.
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      int v = tid + q * kThreads;
      if (v < V) {
        int s = v >> 7, row = (v & 127) >> 1, hf = v & 1;
        int off = s * kLdsSlice + row * kPitch + hf * 16;
        *(v4i *)(sa + off) = pa[q];
        *(v4i *)(sb + off) = pb[q];
      }
    }
  };

  fetch(0);
  for (int64_t kb = 0; kb < KB; ++kb) {
    __syncthreads();
    stash();
    __syncthreads();
    if (kb + 1 < KB) fetch(kb + 1);
    switch (wave) {
      case 0: compute_block<S, 0>(sa, sb, acc, lane); break;
      case 1: compute_block<S, 1>(sa, sb, acc, lane); break;
      case 2: compute_block<S, 2>(sa, sb, acc, lane); break;
      default: compute_block<S, 3>(sa, sb, acc, lane); break;
    }
  }

  // Recombine: wave partials meet in LDS in wave order 0,1,2,3 (fixed, so
  // the result is deterministic); wave 3 scales and stores.
  __syncthreads();
  double *red = (double *)lds;  // 64 x 64 DD = 64 KiB
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            // 32x32 MFMA result layout: lane holds column (lane % 32) and
            // rows 8*(r/4) + 4*(lane/32) + r%4.
            const int row = bi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = bj * 32 + (lane & 31);
            ddv part;
            switch (w) {
              case 0: part = wave_partial<S, 0>(acc, bi, bj, r); break;
              case 1: part = wave_partial<S, 1>(acc, bi, bj, r); break;
              case 2: part = wave_partial<S, 2>(acc, bi, bj, r); break;
              default: part = wave_partial<S, 3>(acc, bi, bj, r); break;
            }
            double *slot = red + 2 * (row * kT + col);
            if (w > 0) part = dd_add(mk(slot[0], slot[1]), part);
            if (w < 3) {
              slot[0] = part.hi;
              slot[1] = part.lo;
            } else {
              const int64_t gi = (int64_t)ib * kT + row, gj = (int64_t)jb * kT + col;
              if (gi < m && gj < n) {
                const int ei = E[gi], fj = F[gj];
                double2 out;
                if (ei == kNonFinite || fj == kNonFinite) {
                  out = make_double2(__builtin_nan(""), 0.0);
                } else {
                  const int sc = scale_of(ei) + scale_of(fj) - 10;
                  out = make_double2(__builtin_ldexp(part.hi, sc), __builtin_ldexp(part.lo, sc));
                }
                C[gi + gj * m] = out;
              }
            }
          }
    }
    __syncthreads();
  }
}

template <int S>
hipError_t launch_main(const signed char *As, const signed char *Bs, const int *E, const int *F,
                       double2 *C, int64_t m, int64_t n, int64_t KB, int64_t tm, int64_t tn,
                       hipStream_t s) {
  // operand staging, reused by the 64 x 64 DD recombination buffer (64 KiB)
  const size_t lds = 2 * (size_t)S * kLdsSlice > 65536 ? 2 * (size_t)S * kLdsSlice : 65536;
  hipError_t e = hipFuncSetAttribute((const void *)k_ozaki<S>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ozaki<S>, dim3((unsigned)(tm * tn)), dim3(kThreads), lds, s, As, Bs, E, F,
                     C, m, n, KB, (int)tm);
  return hipGetLastError();
}

}  // namespace

namespace mfhip {

int ozaki_max_slices() { return 16; }

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t ozaki_workspace_bytes(int64_t m, int64_t k, int64_t n, int S) {
  const int64_t tm = (m + kT - 1) / kT, tn = (n + kT - 1) / kT, KB = (k + kK - 1) / kK;
  return align256((size_t)tm * KB * S * kChunk) + align256((size_t)tn * KB * S * kChunk) +
         align256(sizeof(int) * (size_t)m) + align256(sizeof(int) * (size_t)n);
}

hipError_t ozaki_gemm(const double2 *A, const double2 *B, double2 *C, int64_t m, int64_t k,
                      int64_t n, int S, void *work, size_t work_bytes, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (k <= 0) return hipMemsetAsync(C, 0, sizeof(double2) * m * n, s);
  // int32 level sums stay exact while k * S * 64 * 64 < 2^31
  if (S < 8 || S > ozaki_max_slices() || k > 32768) return hipErrorInvalidValue;
  if (work_bytes < ozaki_workspace_bytes(m, k, n, S)) return hipErrorInvalidValue;
  const int64_t tm = (m + kT - 1) / kT, tn = (n + kT - 1) / kT, KB = (k + kK - 1) / kK;
  char *w = (char *)work;
  signed char *As = (signed char *)w;
  w += align256((size_t)tm * KB * S * kChunk);
  signed char *Bs = (signed char *)w;
  w += align256((size_t)tn * KB * S * kChunk);
  int *E = (int *)w;
  w += align256(sizeof(int) * (size_t)m);
  int *F = (int *)w;

  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, E, m, kEmpty);
  const int64_t kchunk = 64;
  hipLaunchKernelGGL(k_row_exp, dim3((unsigned)((m + kThreads - 1) / kThreads), (unsigned)((k + kchunk - 1) / kchunk)),
                     dim3(kThreads), 0, s, A, m, k, kchunk, E);
  hipLaunchKernelGGL(k_col_exp, dim3((unsigned)n), dim3(kThreads), 0, s, B, k, F);
  hipLaunchKernelGGL(k_slice_a<16>, dim3((unsigned)tm, (unsigned)KB), dim3(kThreads), 0, s, A, m, k, E, S, KB, As);
  hipLaunchKernelGGL(k_slice_b<16>, dim3((unsigned)tn, (unsigned)KB), dim3(kThreads), 0, s, B, k, n, F, S, KB, Bs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  switch (S) {
    case 8: return launch_main<8>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
    case 9: return launch_main<9>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
    case 10: return launch_main<10>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
    case 11: return launch_main<11>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
    case 12: return launch_main<12>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
    case 13: return launch_main<13>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
    case 14: return launch_main<14>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
    case 15: return launch_main<15>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
    default: return launch_main<16>(As, Bs, E, F, C, m, n, KB, tm, tn, s);
  }
}

}  // namespace mfhip
