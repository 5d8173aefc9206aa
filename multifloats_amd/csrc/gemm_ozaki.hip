// Ozaki-style split-precision DD GEMM on the INT8 matrix cores of gfx950.
//
// The precision form of the reference's matmuldd_mm (src/float64x2/matmul.inc):
// same inputs and layout, results held to an error bound instead of the
// reference's bits.
//
// 1. Scaling. E_i = max_p ilogb(A(i,p).hi) per row of A, plus one when the
//    row's largest |hi| reaches 127/128 of 2^(E_i+1); F_j likewise per column
//    of B. x = a * 2^(6 - E_i) then satisfies |x| < 127.
// 2. Slicing. x = sum_s d_s 2^(-8 s): digits d_s = RN(x_s),
//    x_{s+1} = 256 (x_s - d_s), every step exact in DD arithmetic; a digit
//    that rounds to +128 becomes -128 with a carry of one into the digit
//    above (carries ripple upward; |x| < 127 keeps the top digit <= 127). So
//    the S digits are int8 in [-128, 127] and hold the leading 8S bits of x
//    with a round-to-nearest last digit. They are written in a tiled layout:
//    one contiguous (S x 64 x 32)-byte chunk per (64-row block, 32-k block).
// 3. Products. With G_L = sum_p sum_{s+t=L} d^A_s d^B_t,
//        C_ij = 2^(E_i + F_j - 12) * sum_{L < S} 2^(-8 L) G_L(i,j).
//    Levels L >= S are dropped (the triangular Ozaki truncation); each G_L is
//    accumulated by v_mfma_i32_32x32x32_i8 with exact INT32 accumulation,
//    which holds while k * S * 2^14 < 2^31: longer K ranges are cut into
//    chunks whose DD results are added into C.
// 4. Recombination, in the lane that owns the output (see k_ozaki).
// That is the triangular form, run while m or n is below 512. From m, n >= 512
// the CRT form runs instead (k_oz2_resid, k_oz2_gemm, k_oz2_crt below): the
// S-digit integers are multiplied exactly modulo N pairwise-coprime moduli
// (N plain INT8 GEMMs instead of S(S+1)/2) and C is rebuilt by the Chinese
// remainder theorem, with the digit representation as the only error.
#include "dd_core.h"
#include "dd_tables.h"
#include "mf_internal.h"
#include "ozaki_shared.h"

#include <math.h>

namespace {

using namespace mfhip_oz;

__device__ __forceinline__ int dexp(double x) {
  int e;
  (void)__builtin_frexp(x, &e);
  return e - 1;
}

// Row exponent maxima and minima of A (column-major m x k): thread = row of
// the 256-row block rb, over k chunk kc. Zeros are skipped; a non-finite
// entry marks the row.
// Mbits collects the largest |hi| of the row (the bit patterns of
// non-negative doubles order like the values).
__device__ __forceinline__ void row_exp(const double2 *__restrict__ A, int64_t m, int64_t k, int64_t kchunk,
                                        int64_t rb, int64_t kc, int *__restrict__ E, int *__restrict__ Emin,
                                        unsigned long long *__restrict__ Mbits) {
  const int64_t i = rb * kThreads + threadIdx.x;
  if (i >= m) return;
  const int64_t p0 = kc * kchunk;
  const int64_t p1 = p0 + kchunk < k ? p0 + kchunk : k;
  int e = kEmpty, lo = kNonFinite;
  double mx = 0.0;
  bool bad = false;
#pragma unroll 8
  for (int64_t p = p0; p < p1; ++p) {
    double h = A[i + p * m].x;
    if (!mf_isfinite(h)) bad = true;
    else if (h != 0.0) {
      const int x = dexp(h);
      e = max(e, x);
      lo = min(lo, x);
      mx = fmax(mx, fabs(h));
    }
  }
  if (bad) e = kNonFinite;
  if (e != kEmpty) atomicMax(&E[i], e);
  if (lo != kNonFinite) atomicMin(&Emin[i], lo);
  if (mx > 0.0) atomicMax(&Mbits[i], (unsigned long long)__double_as_longlong(mx));
}

// One more bit of headroom for a row or column whose largest |hi| reaches
// 127/128 of 2^(e+1), so that |x| < 127 after scaling by 2^(6-e).
__device__ __forceinline__ int headroom(int e, double mx) {
  if (e == kEmpty || e == kNonFinite) return e;
  return mx >= __builtin_ldexp(127.0, e - 6) ? e + 1 : e;
}

// Finish the row exponents (headroom) and fold the largest exponent range
// max - min over the finite, nonzero rows into *range.
__global__ void k_row_finish(int *__restrict__ E, const int *__restrict__ Emin,
                             const unsigned long long *__restrict__ Mbits, int64_t m, int *__restrict__ range) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int e = E[i];
  if (e != kEmpty && e != kNonFinite) {
    const int eh = headroom(e, __longlong_as_double((long long)Mbits[i]));
    E[i] = eh;
    atomicMax(range, eh - Emin[i]);
  }
}

// Column exponent maxima of B (column-major k x n) with the same headroom
// rule: one workgroup for column j, threads walk p (all of a thread's loads issued before
// its max/min chain); the block reduces within each wave by shuffles, then
// across the four waves through LDS. Also folds the column's exponent range
// into *range.
__device__ __forceinline__ void col_exp(const double2 *__restrict__ B, int64_t k, int64_t j,
                                        int *__restrict__ F, int *__restrict__ range) {
  int e = kEmpty, lo = kNonFinite;
  double mx = 0.0;
  bool bad = false;
  const double2 *col = B + j * k;
  for (int64_t p0 = 0; p0 < k; p0 += 8 * kThreads) {
    double h[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t p = p0 + threadIdx.x + (int64_t)u * kThreads;
      h[u] = p < k ? col[p].x : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (!mf_isfinite(h[u])) bad = true;
      else if (h[u] != 0.0) {
        const int x = dexp(h[u]);
        e = max(e, x);
        lo = min(lo, x);
        mx = fmax(mx, fabs(h[u]));
      }
    }
  }
  if (bad) e = kNonFinite;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e = max(e, __shfl_xor(e, off));
    lo = min(lo, __shfl_xor(lo, off));
    mx = fmax(mx, __shfl_xor(mx, off));
  }
  __shared__ int rmax[kThreads / 64], rmin[kThreads / 64];
  __shared__ double rmx[kThreads / 64];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    rmax[wv] = e;
    rmin[wv] = lo;
    rmx[wv] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) {
      e = max(e, rmax[w]);
      lo = min(lo, rmin[w]);
      mx = fmax(mx, rmx[w]);
    }
    const int eh = headroom(e, mx);
    F[j] = eh;
    if (eh != kEmpty && eh != kNonFinite) atomicMax(range, eh - lo);
  }
}

// The exponent scans of A's rows and B's columns in one launch, so that the
// two read streams overlap: workgroups below rblocks * kchunks take a
// (256-row block, k chunk) of A, the rest one column of B each. Both parts
// are uniform per workgroup (the column part synchronises its workgroup).
__global__ __launch_bounds__(kThreads) void k_scan_exp(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                       int64_t m, int64_t k, int64_t kchunk, int64_t kchunks,
                                                       int64_t rblocks, int *__restrict__ E,
                                                       int *__restrict__ Emin, unsigned long long *__restrict__ Mbits,
                                                       int *__restrict__ F, int *__restrict__ range) {
  const int64_t b = blockIdx.x, nrow = rblocks * kchunks;
  if (b < nrow)
    row_exp(A, m, k, kchunk, b / kchunks, b % kchunks, E, Emin, Mbits);
  else
    col_exp(B, k, b - nrow, F, range + 1);
}

// Initial state of the row scan: E = empty, Emin = non-finite, Mbits = 0,
// both ranges = 0, and the certificate's row and column sums = 0 (one launch
// over max(m, n) threads).
__global__ void k_scan_init(int *__restrict__ E, int *__restrict__ Emin, unsigned long long *__restrict__ Mbits,
                            int *__restrict__ range, unsigned long long *__restrict__ rsum,
                            unsigned long long *__restrict__ csum, int64_t m, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) {
    E[i] = kEmpty;
    Emin[i] = kNonFinite;
    Mbits[i] = 0ull;
    rsum[i] = 0ull;
  }
  if (i < n) csum[i] = 0ull;
  if (i < 2) range[i] = 0;
}

// Base-256 digits of x (|x| < 127, exact DD), S <= MAXS of them, in
// [-128, 127]: round to nearest digit by digit, then turn every +128 into
// -128 with a carry into the digit above (bottom-up, so carries ripple).
template <int MAXS>
__device__ __forceinline__ void digits(double h, double l, int S, int *d) {
#pragma unroll
  for (int s = 0; s < MAXS; ++s) {
    if (s < S) {
      double q = __builtin_rint(h);
      d[s] = (int)q;
      h = h - q;  // exact
      fast_two_sum(h, l, h, l);
      h *= 256.0;
      l *= 256.0;
    }
  }
#pragma unroll
  for (int s = MAXS - 1; s > 0; --s) {
    if (s < S && d[s] == 128) {
      d[s] = -128;
      d[s - 1] += 1;
    }
  }
}

__device__ __forceinline__ int scale_of(int e) { return (e == kEmpty || e == kNonFinite) ? 0 : e; }

// Slice a 64-row x 32-k tile of A. Element (r, c) of the tile is A(i0+r, p0+c);
// its digit s goes to out[((ib*KB + kb)*S + s)*2048 + tile_off(r, c)].
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
        ex = 6 - Ei;
      }
    }
    h = __builtin_ldexp(h, ex);
    l = __builtin_ldexp(l, ex);
    int d[MAXS];
    digits<MAXS>(h, l, S, d);
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < S) tile[s * kChunk + tile_off(r, c)] = (signed char)d[s];
  }
  __syncthreads();
  const int nvec = S * kChunk / 16;
  v4i *dst = (v4i *)(out + ((int64_t)ib * KB + kb) * S * kChunk);
  const v4i *src = (const v4i *)tile;
  for (int v = tid; v < nvec; v += kThreads) dst[v] = src[v];
}

// Slice a 32-k x 64-col tile of B; digit t of B(p0+p, j0+c) at tile_off(c, p).
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
        ex = 6 - Fj;
      }
    }
    h = __builtin_ldexp(h, ex);
    l = __builtin_ldexp(l, ex);
    int d[MAXS];
    digits<MAXS>(h, l, S, d);
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < S) tile[s * kChunk + tile_off(c, p)] = (signed char)d[s];
  }
  __syncthreads();
  const int nvec = S * kChunk / 16;
  v4i *dst = (v4i *)(out + ((int64_t)jb * KB + kb) * S * kChunk);
  const v4i *src = (const v4i *)tile;
  for (int v = tid; v < nvec; v += kThreads) dst[v] = src[v];
}

__device__ __forceinline__ v4i lds_frag(const signed char *base, int row, int half) {
  return *(const v4i *)(base + row * kK + ((half ^ swz(row)) << 4));
}

// Main kernel: a workgroup of 4 waves (one per SIMD) owns a 64x64 tile of C;
// wave w owns its 32x32 block (w & 1, w >> 1) with ALL S level accumulators
// (S x 16 registers per lane, held in the accumulation registers). Per K
// block the wave reads its S A-fragments once into registers and streams the
// S B-fragments: B fragment t meets A fragments s < S - t, so every fragment
// read from LDS feeds up to S MFMAs (LDS traffic 32 fragments per 136 MFMAs
// at S = 16). Consecutive MFMAs update distinct levels, so the matrix core
// never waits on its own previous result. The S level sums of an output are
// then all in the owning lane: recombination needs no data exchange.
//
// K range: blocks [kb0, kb0 + nkb) of the KB blocks in the workspace. With
// accumulate set, the result is added (DD) to what C holds (later chunks of a
// K range too long for exact INT32 accumulation).
template <int S>
__global__ __launch_bounds__(kMainThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_ozaki(
    const signed char *__restrict__ As, const signed char *__restrict__ Bs, const int *__restrict__ E,
    const int *__restrict__ F, double2 *__restrict__ C, int64_t m, int64_t n, int64_t KB, int tiles_m,
    int64_t kb0, int64_t nkb, int accumulate) {
  extern __shared__ __attribute__((aligned(16))) signed char lds[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar address math
  const int ib = blockIdx.x % tiles_m, jb = blockIdx.x / tiles_m;
  const int bi = wave & 1, bj = wave >> 1;
  const int r = lane & 31, half = lane >> 5;
  const signed char *ga = As + ((int64_t)ib * KB + kb0) * S * kChunk;
  const signed char *gb = Bs + ((int64_t)jb * KB + kb0) * S * kChunk;
  constexpr int kBuf = 2 * S * kChunk;                          // one stage: A chunk then B chunk
  constexpr int NV = S * kChunk / 1024;                         // 1 KiB wave-loads per operand chunk
  constexpr int NLOAD = 2 * ((NV + kWaves - 1) / kWaves);      // loads per wave per K block
  constexpr int kSpread = 4;                                    // one load per 4 MFMAs

  v16i acc[S];
#pragma unroll
  for (int L = 0; L < S; ++L) acc[L] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

  // Load j of a wave's share of the asynchronous HBM -> LDS copy of K block
  // kb into stage st: operand j & 1, 1 KiB piece wave + 4 (j >> 1) (16 bytes
  // per lane, consecutive in LDS).
  auto issue_one = [&](int64_t kb, int st, int j) {
    const int op = j & 1, v = wave + kWaves * (j >> 1);
    if (v < NV) {
      signed char *dst = lds + st * kBuf + op * S * kChunk + v * 1024;
      const signed char *src = (op ? gb : ga) + kb * S * kChunk + v * 1024 + lane * 16;
      __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)dst, 16,
                                       0, 0);
    }
  };

#pragma unroll
  for (int j = 0; j < NLOAD; ++j) issue_one(0, 0, j);
  for (int64_t kb = 0; kb < nkb; ++kb) {
    const int st = (int)(kb & 1);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of block kb have landed
    __syncthreads();                // ...and every wave's; stage st^1 is free
    const bool more = kb + 1 < nkb;
    const signed char *sa = lds + st * kBuf;
    const signed char *sb = sa + S * kChunk;
    v4i a[S];
    v4i b = lds_frag(sb, 32 * bj + r, half);
#pragma unroll
    for (int s = 0; s < S; ++s) a[s] = lds_frag(sa + s * kLdsSlice, 32 * bi + r, half);
    v4i bn = lds_frag(sb + kLdsSlice, 32 * bj + r, half);
    // The copy of block kb+1 is spread over the first MFMAs of this block
    // (one load per kSpread MFMAs; the instruction order is pinned by
    // scheduling barriers): a burst of loads at the block start stalls the
    // MFMA stream measurably.
    int slot = 0;
#pragma unroll
    for (int t = 0; t < S; ++t) {
      if (t > 0) {
        b = bn;
        if (t + 1 < S) bn = lds_frag(sb + (t + 1) * kLdsSlice, 32 * bj + r, half);
      }
#pragma unroll
      for (int s = 0; s + t < S; ++s) {
        acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[s], b, acc[s + t], 0, 0, 0);
        if (more && slot % kSpread == 0 && slot / kSpread < NLOAD) issue_one(kb + 1, st ^ 1, slot / kSpread);
        __builtin_amdgcn_sched_barrier(0);
        ++slot;
      }
    }
#pragma unroll
    for (int j = (S * (S + 1) / 2 + kSpread - 1) / kSpread; j < NLOAD; ++j)
      if (more) issue_one(kb + 1, st ^ 1, j);
  }
  __builtin_amdgcn_s_waitcnt(0);

  // Recombine in the owning lane. The 32x32 MFMA result layout puts column r
  // of the block in lane r + 32*h and row 8*(e/4) + 4*h + e%4 in element e.
  // Levels 3g..3g+2 are first combined exactly in 64-bit integers (|G_L| <
  // 2^31, so a group stays below 2^48 and converts to double exactly); the
  // group values are then summed smallest first with TwoSum, their rounding
  // errors collected, and closed by a Fast2Sum. The order is fixed.
  constexpr int NG = (S + 2) / 3;
  const int64_t gj = (int64_t)jb * kT + 32 * bj + r;
  const int fj = gj < n ? F[gj] : 0;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int64_t gi = (int64_t)ib * kT + 32 * bi + 8 * (e >> 2) + 4 * half + (e & 3);
    double hi = 0.0, lo = 0.0;
#pragma unroll
    for (int g = NG - 1; g >= 0; --g) {
      constexpr int kLast = S - 1;
      const int L1 = 3 * g + 2 < kLast ? 3 * g + 2 : kLast;  // last level of the group
      int64_t v = 0;
#pragma unroll
      for (int L = 3 * g; L <= L1; ++L) v = v * 256 + (int64_t)acc[L][e];
      const double x = __builtin_ldexp((double)v, -8 * L1);
      double t, err;
      two_sum(hi, x, t, err);
      hi = t;
      lo += err;
    }
    ddv part;
    fast_two_sum(hi, lo, part.hi, part.lo);
    if (gi < m && gj < n) {
      const int ei = E[gi];
      double2 out;
      if (ei == kNonFinite || fj == kNonFinite) {
        out = make_double2(__builtin_nan(""), 0.0);
      } else {
        const int sc = scale_of(ei) + scale_of(fj) - 12;
        out = make_double2(__builtin_ldexp(part.hi, sc), __builtin_ldexp(part.lo, sc));
        if (accumulate) {
          const double2 c0 = C[gi + gj * m];
          const ddv sum = dd_add(mk(c0.x, c0.y), mk(out.x, out.y));
          out = make_double2(sum.hi, sum.lo);
        }
      }
      C[gi + gj * m] = out;
    }
  }
}

// Longest K range (in 32-deep blocks) whose level sums stay exact in INT32:
// (L+1) k 2^14 < 2^31 for every level L < S.
int64_t max_exact_kblocks(int S) { return ((2147483647LL / ((int64_t)S << 14)) / kK); }

template <int S>
hipError_t launch_main(const signed char *As, const signed char *Bs, const int *E, const int *F,
                       double2 *C, int64_t m, int64_t n, int64_t KB, int64_t tm, int64_t tn,
                       hipStream_t s) {
  // two operand stages of 2 x S x 2 KiB each
  const size_t lds = 4 * (size_t)S * kChunk;
  hipError_t e = hipFuncSetAttribute((const void *)k_ozaki<S>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int64_t span = max_exact_kblocks(S);
  for (int64_t kb0 = 0; kb0 < KB; kb0 += span) {
    const int64_t nkb = KB - kb0 < span ? KB - kb0 : span;
    hipLaunchKernelGGL(k_ozaki<S>, dim3((unsigned)(tm * tn)), dim3(kMainThreads), lds, s, As, Bs, E, F, C, m,
                       n, KB, (int)tm, kb0, nkb, kb0 > 0 ? 1 : 0);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// CRT form (the modular "scheme II" of Ozaki, Uchino and Imamura, 2025, in a
// design of our own). The S digits of an element are one integer
// X = sum_s d_s 256^(S-1-s) (|X| < 2^(8S-1)), so C_int = A_int B_int is an
// exact integer product. Instead of S(S+1)/2 digit-pair products it is formed
// modulo N pairwise-coprime moduli m_t <= 256 (one plain INT8 GEMM each: the
// balanced residues fit int8 and k 2^14 < 2^31) and reassembled by the
// Chinese remainder theorem, with M = prod m_t > 2 k 2^(16S-2) > 2 |C_int|.
// The truncated levels of the triangular form are not dropped here, so the
// error is the digit representation's alone (no larger than the triangular
// form's with the same S).
// Moduli count of the CRT form for each slice count, computed once on the
// host for the call's k (the counts the workspace is sized for).
struct OzNtab {
  int n[kMaxSlices + 1];
};

// The exponent-range rule for the first slice count of an automatic call
// (host and device): an empirical model of how far (|A||B|)_ij falls below
// 2^(E_i + F_j) on iid inputs of the workloads (entries s*2^U(-e,e)*u),
// fitted on MI355X, in bits of digit precision b = 8S:
//  * error |err_ij| ~ 2^(E_i + F_j - b - 1) sqrt(k);
//  * (|A||B|)_ij / 2^(E_i + F_j) below 1 by at most W = (range_a + range_b)
//    f(k) binades, f the measured envelope (max over e = 10, 20, 30 of the
//    worst output among 512 x 512), interpolated in log2 k;
//  * 3 binades of margin.
// It only sets where the search starts: the certificate of the automatic
// call (k_oz2_gemm<kOzMag>) decides which outputs a count covers.
__host__ __device__ inline int oz_guess_slices(int64_t k, int range_a, int range_b) {
  if (k <= 0) return kMinSlices;
  const double lk[9] = {4, 5, 6, 7, 8.23, 9.97, 11, 12, 13};
  const double fr[9] = {0.42, 0.354, 0.28, 0.20, 0.12, 0.05, 0.0275, 0.004, 0.002};
  const double x = log2((double)k);
  double f = fr[0];
  if (x >= lk[8]) {
    f = fr[8];
  } else if (x > lk[0]) {
    for (int i = 0; i < 8; ++i)
      if (x <= lk[i + 1]) {
        const double t = (x - lk[i]) / (lk[i + 1] - lk[i]);
        f = fr[i] + t * (fr[i + 1] - fr[i]);
        break;
      }
  }
  const double need = 101.0 + 0.5 * x + (double)(range_a + range_b) * f + 3.0;
  const int S = (int)ceil(need / 8.0);
  return S < kMinSlices ? kMinSlices : S;
}

// Balanced int8 residue of p (0 <= p < m) as a byte.
__device__ __forceinline__ unsigned bal_byte(int p, int m) {
  const int bal = (m == 256 ? p >= 128 : p > (m - 1) / 2) ? p - m : p;
  return (unsigned)bal & 255u;
}

// Residue weights and reconstruction constants of S slices and N moduli into
// ctl: thread t (< 64) fills modulus t; byte j of w[t] is f_t 256^j mod m_t
// for j < S, zero above, except byte 15 = 64 when S <= 15 (k_oz2_resid's
// bias slot).
__device__ void oz_set_plan(OzCtl *ctl, int S, int N, int t) {
  if (t < MF_OZ2_MODULI) {
    unsigned wa[4] = {0, 0, 0, 0}, wb[4] = {0, 0, 0, 0};
    float inv = 0.0f;
    if (t < N) {
      const int mt = mf_oz2_m[t];
      int pa = 1 % mt, pb = pa;
      for (int j = 0; j < S; ++j) {
        wa[j / 4] |= bal_byte(pa, mt) << (8 * (j % 4));
        wb[j / 4] |= bal_byte(pb, mt) << (8 * (j % 4));
        pa = pa * 256 % mt;
        pb = pb * 256 % mt;
      }
      if (S <= 15) {  // slot 15: the residue pass's bias of 128 (digit 2 x weight 64)
        wa[3] |= 64u << 24;
        wb[3] |= 64u << 24;
      }
      inv = 1.0f / (float)mt;
    }
    for (int q = 0; q < 4; ++q) {
      ctl->wa.w[t][q] = (int)wa[q];
      ctl->wb.w[t][q] = (int)wb[q];
    }
    ctl->wa.inv[t] = ctl->wb.inv[t] = inv;
  }
  if (t == 0) {
    // M 2^-(16(S-1)) for the product M of the first N moduli (M rounded to a
    // double-double, so within 2^-106 M)
    ctl->cs.mp_hi = __builtin_ldexp(mf_oz2_mp[2 * (N - 1)], -16 * (S - 1));
    ctl->cs.mp_lo = __builtin_ldexp(mf_oz2_mp[2 * (N - 1) + 1], -16 * (S - 1));
  }
}

// First plan of a call (one wave): the explicit slice count, or for an
// automatic call (S_explicit <= 0) the exponent-range rule's count capped at
// kMaxSlices; counters cleared.
__global__ __launch_bounds__(64) void k_oz_plan_begin(OzCtl *__restrict__ ctl, const int *__restrict__ range,
                                                      OzNtab nt, int S_explicit, int64_t k) {
  int S = S_explicit;
  if (S <= 0) {
    S = oz_guess_slices(k, range[0], range[1]);
    if (S > kMaxSlices) S = kMaxSlices;
  }
  const int t = threadIdx.x;
  if (t == 0) {
    ctl->S = S;
    ctl->N = nt.n[S];
    ctl->redo = 0;
    ctl->S_first = S_explicit > 0 ? 0 : S;
    ctl->ncand = ctl->nfix = 0ull;
    ctl->rmax = ctl->cmax = 0ull;
  }
  if (t < kOzBins) ctl->hist[t] = 0ull;
  oz_set_plan(ctl, S, nt.n[S], t);
}

// Moduli count from the data (automatic calls): the products need
// M > 2 max |C_int|, and with X_ip = RN(a_ip 2^(8S-2-E_i)),
// |X_ip| <= |a_ip| 2^(8S-2-E_i) + 1/2 and every |Y_pj| < 2^(8S-1) (the
// headroom rule keeps |b| 2^(6-F) < 127),
//   |C_ij| <= sum_p |X_ip| |Y_pj| < 2^(8S-1) (2^(8S-2) R_A,i + k),
// and the same with the roles of A and B exchanged; R_A,i = sum_p |a_ip|
// 2^-E_i is bounded by the certificate's row sum (units of 2^-40, rounded up).
// rmax / cmax are the largest row / column sums of the live rows and columns,
// so C_max = 2^(8S-1) (2^(8S-2) min(rmax, cmax) 2^-40 + k) bounds every
// output (a dead row or column gives zero products). The count is the
// smallest N whose M (a double within 2^-53 of it) exceeds 2 C_max (1 + 2^-40),
// never more than n_worst, the count for the worst case k 2^(16S-2).
__device__ __forceinline__ int oz_data_moduli(int n_worst, int S, int64_t k, unsigned long long rmax,
                                              unsigned long long cmax) {
  const unsigned long long r = rmax < cmax ? rmax : cmax;
  const double cm = __builtin_ldexp(__builtin_ldexp((double)r * (1.0 + 0x1p-50), 8 * S - 42) + (double)k, 8 * S - 1);
  const double need = 2.0 * cm * (1.0 + 0x1p-40);
  for (int N = 1; N < n_worst; ++N)
    if (mf_oz2_mp[2 * (N - 1)] > need) return N;
  return n_worst;
}

// Final plan of an automatic call (one wave), after the certificate: the
// smallest S >= the first pass's count whose uncertified outputs (those that
// need more slices) number at most `budget` -- the fix-up recomputes those --
// or, when even kMaxSlices leaves more than that, the smallest S that leaves
// no more than kMaxSlices does. A raised S sets `redo`: the residues are then
// formed again with the new weights.
__global__ __launch_bounds__(64) void k_oz_plan_final(OzCtl *__restrict__ ctl, OzNtab nt,
                                                      unsigned long long budget, int64_t k) {
  const int S0 = ctl->S;
  unsigned long long over[kOzBins];  // over[s] = candidates needing more than s slices
  unsigned long long acc = 0;
  for (int s = kOverCap; s >= 0; --s) {
    over[s] = acc;
    acc += ctl->hist[s];
  }
  const unsigned long long target = budget > over[kMaxSlices] ? budget : over[kMaxSlices];
  int S = S0;
  while (S < kMaxSlices && over[S] > target) ++S;
  const int N = oz_data_moduli(nt.n[S], S, k, ctl->rmax, ctl->cmax);
  const int t = threadIdx.x;
  if (S != S0 || N != ctl->N) {
    __syncthreads();  // every lane has read ctl->S and ctl->N above
    if (t == 0) {
      ctl->S = S;
      ctl->N = N;
      ctl->redo = S != S0;
    }
    oz_set_plan(ctl, S, N, t);
  }
}

// One modular INT8 GEMM tile per workgroup: plane t = blockIdx.y, a 256 x 256
// output tile, 4 waves (one per SIMD) of 128 x 128 (4 x 4
// v_mfma_i32_32x32x32_i8 blocks, all 256 accumulators in AGPRs). Operands
// stream HBM -> LDS with asynchronous 16-byte copies, kO2KS K blocks per
// stage, in kO2Bufs stage buffers: the copies run kO2Bufs - 1 stages ahead
// of the MFMAs (a load latency of several microseconds under full load
// exceeds one stage of MFMAs), and each stage waits only for its own
// copies. The epilogue reduces each exact INT32 sum v (== c_t y_t, the CRT
// digit, mod m_t) to its balanced residue and writes the signed byte to plane
// t of R (column-major, like C).
constexpr int kO2Bufs = 4;
constexpr int kO2Stage = kO2KS * 8 * kChunk;              // 4 A + 4 B chunks per K block
constexpr int kO2NL = kO2Stage / 1024 / kWaves;            // 1 KiB copies per wave per stage
static_assert(kO2NL * 1024 * kWaves == kO2Stage, "stage copies split evenly over the waves");
static_assert(kO2Bufs * kO2Stage <= 160 * 1024, "stage buffers fit in LDS");
// Tile of workgroup x in a TM x TN tile grid. Workgroups are dispatched
// round-robin over the 8 XCDs by linear id, and each XCD has its own L2: when
// the grid splits evenly, the 8 residues of x get one compact (TM/2) x (TN/4)
// sub-grid each, so an operand panel is fetched by 4 (A) or 2 (B) XCDs rather
// than by all 8. (TM TN % 8 == 0 keeps the split the same in every plane.)
__device__ __forceinline__ void xcd_tile(int x, int TM, int TN, int &ib, int &jb) {
  if (TM % 2 == 0 && TN % 4 == 0) {
    const int xcd = x & 7, slot = x >> 3, bm = TM / 2, bn = TN / 4;
    ib = (xcd >> 2) * bm + slot % bm;
    jb = (xcd & 3) * bn + slot / bm;
  } else {
    ib = x % TM;
    jb = x / TM;
  }
}
// s_waitcnt immediate (gfx9 encoding) that waits for vmcnt <= n only
constexpr int waitcnt_vm(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }
// Workgroup barrier for LDS hand-offs only. __syncthreads() adds a fence that
// drains every outstanding global access (vmcnt(0)), which would also wait for
// the copies of the stages still in flight; the copies a stage needs are
// waited for explicitly before this barrier, and LDS needs no cache flush.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
// The two uses of the GEMM kernel below: the N residue products of the CRT
// form (INT8 MFMA, exact INT32 sums) and the certificate of an automatic call
// (FP8 MFMA on the magnitude operands, FP32 sums).
constexpr int kOzRes = 0;
constexpr int kOzMag = 1;
typedef float v16f __attribute__((ext_vector_type(16)));
template <int MODE>
struct O2Mma;
template <>
struct O2Mma<kOzRes> {
  typedef v16i acc_t;
  static __device__ __forceinline__ acc_t run(v4i a, v4i b, acc_t c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
  }
};
// The certificate's MFMA takes both K blocks of a stage at once: a lane's two
// 16-byte fragments (block 0, then block 1) form the 32 operand bytes of one
// v_mfma_scale_f32_32x32x64_f8f6f4 (format 0: OCP E4M3; unit E8M0 scales),
// the full-rate FP8 form. Both operands concatenate the same way, so every k
// meets its partner once (the sum does not depend on which k the instruction
// pairs in which step).
typedef int v8i __attribute__((ext_vector_type(8)));
template <>
struct O2Mma<kOzMag> {
  typedef v16f acc_t;
  static __device__ __forceinline__ acc_t run(v8i a, v8i b, acc_t c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, 127, 0, 127);
  }
};

struct Oz2GemmArgs {
  const signed char *As, *Bs;  // operand planes (residues, or FP8 magnitudes)
  unsigned char *R;            // digit planes (kOzRes)
  int64_t m, n, KB, TBA, TBB;  // KB: 32-byte K blocks per chunk row, padded to kO2KS
  int tiles_m, tiles_n;
  OzCtl *ctl;
  // certificate (kOzMag)
  const int *E, *F;
  const unsigned long long *rsum, *csum;  // sum_p (|hi| + |lo|) 2^-E in units of 2^-40, rounded up
  unsigned long long *cand;                // candidate list: (slices << 56) | (i + j m)
};

// Lower slice-count bound of one output (i, j) from its certificate sum G:
// with E = E_i, F = F_j, the CRT form's representation error is at most
// 2^(E + F + 1 - 8S) (R_A + R_B) (1 + 2^-20), R_A = sum_p |a_ip| 2^-E and
// R_B = sum_p |b_pj| 2^-F (each element within 2^(E + 1 - 8S) of its S-slice
// integer; the 2^-20 covers the second-order term and the reconstruction's
// limb truncation), and its reconstruction adds at most 2^-103 (|A||B|)_ij.
// (|A||B|)_ij >= 2^(E + F) G (1 - 2^-6): G sums products of FP8 values
// rounded down (the GEMM's sums of the 2^7-scaled magnitudes, times 2^-14),
// in FP32 with at most k roundings of non-negative terms (k <= 2^15). So
// 2^-100 (|A||B|)_ij holds once
// 8S >= 101.2 + log2((R_A + R_B) / (G (1 - 2^-6))); 0.1 more covers the FP32
// logarithms. Returns the smallest such S in kMinSlices..kMaxSlices, or
// kOverCap.
__device__ __forceinline__ int oz_cert_slices(float G, double ra, double rb) {
  const float L = G * (1.0f - 0x1p-6f);
  if (!(L > 0.0f)) return kOverCap;
  const float bits = 101.3f + __builtin_log2f((float)(ra + rb)) - __builtin_log2f(L);
  const int S = (int)__builtin_ceilf(bits * 0.125f);
  return S < kMinSlices ? kMinSlices : (S > kMaxSlices ? kOverCap : S);
}

// Exclusive prefix sum over the 64 lanes of a wave, and the total.
__device__ __forceinline__ unsigned wave_excl_scan(unsigned v, int lane, unsigned &total) {
  unsigned x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  total = __shfl(x, 63);
  return x - v;
}

template <int MODE>
__global__ __launch_bounds__(kMainThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_oz2_gemm(
    Oz2GemmArgs g) {
  typedef typename O2Mma<MODE>::acc_t acc_t;
  extern __shared__ __attribute__((aligned(16))) signed char lds[];
  const int t = blockIdx.y;
  const int Nc = MODE == kOzRes ? g.ctl->N : 0;
  if (MODE == kOzRes && t >= Nc) return;  // planes beyond the call's moduli count
  __shared__ unsigned hist_l[kOzBins];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (MODE == kOzMag && tid < kOzBins) hist_l[tid] = 0u;  // published by the first barrier below
  int ib, jb;
  xcd_tile(blockIdx.x, g.tiles_m, g.tiles_n, ib, jb);
  const int wi = wave & 1, wj = wave >> 1;
  const int r = lane & 31, half = lane >> 5;
  const int64_t KB = g.KB;
  const signed char *ga = g.As + ((int64_t)t * g.TBA + 4 * ib) * KB * kChunk;
  const signed char *gb = g.Bs + ((int64_t)t * g.TBB + 4 * jb) * KB * kChunk;
  // The output plane's address is formed (and its scalar argument load
  // completed) before the main loop: a scalar load still in flight there
  // would force every LDS wait in the loop to wait for all reads.
  unsigned char *gr = g.R + (int64_t)t * g.m * g.n;
  asm volatile("" : "+s"(gr));
  acc_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0;
  // copy v = wave + 4 j of stage kbs: K block q = v / 16, operand (v / 8) & 1,
  // chunk (v / 2) & 3, half KiB v & 1
  auto issue = [&](int64_t kbs, int j) {
    const int v = wave + kWaves * j;
    const int q = v >> 4, op = (v >> 3) & 1, c = (v >> 1) & 3, hh = v & 1;
    signed char *dst = lds + (int)(kbs % kO2Bufs) * kO2Stage + ((q * 2 + op) * 4 + c) * kChunk + hh * 1024;
    const signed char *src = (op ? gb : ga) + ((int64_t)c * KB + kbs * kO2KS + q) * kChunk + hh * 1024 + lane * 16;
    __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
  };
  const int64_t NS = KB / kO2KS;  // KB is padded to a multiple of kO2KS
  for (int p = 0; p < kO2Bufs - 1; ++p)
    if (p < NS)
#pragma unroll
      for (int j = 0; j < kO2NL; ++j) issue(p, j);
  // Wait until this wave's copies of stage x have landed; the copies of the
  // (at most kO2Bufs - 2) stages issued after it may stay in flight.
  auto wait_stage = [&](int64_t x) {
    if (x + 2 < NS) __builtin_amdgcn_s_waitcnt(waitcnt_vm(2 * kO2NL));
    else if (x + 1 < NS) __builtin_amdgcn_s_waitcnt(waitcnt_vm(kO2NL));
    else __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
  };
  // Stage s holds K blocks q = 0, 1. The fragments of block 0 of stage s + 1
  // are read from LDS while the MFMAs of block 1 of stage s run, so no stage
  // starts on an LDS read: the wait for stage s + 1's copies and the barrier
  // sit between the two blocks of stage s. The copies of stage s + 3, issued
  // during block 0 of stage s, overwrite the buffer of stage s - 1, whose
  // last reads completed before the barrier in the middle of stage s - 1.
  static_assert(kO2KS == 2 && 16 % kO2NL == 0, "two K blocks per stage; copies spread over one block");
  if constexpr (MODE == kOzRes) {
    v4i fa[2][4], fb[2][4];
    wait_stage(0);
    lds_barrier();
#pragma unroll
    for (int a = 0; a < 4; ++a) fa[0][a] = lds_frag(lds + (2 * wi + (a >> 1)) * kChunk, 32 * (a & 1) + r, half);
#pragma unroll
    for (int b = 0; b < 4; ++b) fb[0][b] = lds_frag(lds + (4 + 2 * wj + (b >> 1)) * kChunk, 32 * (b & 1) + r, half);
    for (int64_t s = 0; s < NS; ++s) {
      const bool more = s + kO2Bufs - 1 < NS;
      const signed char *sa1 = lds + (int)(s % kO2Bufs) * kO2Stage + 8 * kChunk;
#pragma unroll
      for (int a = 0; a < 4; ++a) fa[1][a] = lds_frag(sa1 + (2 * wi + (a >> 1)) * kChunk, 32 * (a & 1) + r, half);
#pragma unroll
      for (int b = 0; b < 4; ++b) fb[1][b] = lds_frag(sa1 + (4 + 2 * wj + (b >> 1)) * kChunk, 32 * (b & 1) + r, half);
      int slot = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          acc[a][b] = O2Mma<MODE>::run(fa[0][a], fb[0][b], acc[a][b]);
          constexpr int every = 16 / kO2NL;
          if (more && slot % every == 0) issue(s + kO2Bufs - 1, slot / every);
          ++slot;
        }
      // Unconditional, also after the last stage (the reads then fetch a stale
      // stage buffer inside the allocation and go unused): a branch here would
      // make the compiler wait for these reads before the MFMAs below.
      wait_stage(s + 1);
      lds_barrier();
      // One MFMA of block 1 goes first: the compiler's wait for block 1's
      // fragments then lands before the new reads instead of after them.
      acc[0][0] = O2Mma<MODE>::run(fa[1][0], fb[1][0], acc[0][0]);
      __builtin_amdgcn_sched_barrier(0);
      const signed char *sn = lds + (int)((s + 1) % kO2Bufs) * kO2Stage;
#pragma unroll
      for (int a = 0; a < 4; ++a) fa[0][a] = lds_frag(sn + (2 * wi + (a >> 1)) * kChunk, 32 * (a & 1) + r, half);
#pragma unroll
      for (int b = 0; b < 4; ++b) fb[0][b] = lds_frag(sn + (4 + 2 * wj + (b >> 1)) * kChunk, 32 * (b & 1) + r, half);
      __builtin_amdgcn_sched_barrier(0);  // the reads above stay ahead of these MFMAs
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if (a + b > 0) acc[a][b] = O2Mma<MODE>::run(fa[1][a], fb[1][b], acc[a][b]);
      // The next stage's fragments leave the loop through an empty register
      // asm: otherwise the compiler merges these reads with the ones before the
      // loop into one read at the top of the loop, behind the MFMAs again.
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int a = 0; a < 4; ++a) asm("" : "+v"(fa[0][a]));
#pragma unroll
      for (int b = 0; b < 4; ++b) asm("" : "+v"(fb[0][b]));
    }
  } else {
    // Certificate: one MFMA per block pair and stage (both K blocks), rows
    // of blocks in order. From the middle of stage s (after the wait for
    // stage s + 1's copies and the barrier) each fragment is reloaded with
    // stage s + 1's as soon as its last MFMA of stage s has issued: A
    // fragments 0 and 1 at once, A fragment 2 after its row, the B
    // fragments during the last row, A fragment 3 at the end. The stage's
    // copies go out during the first half.
    static_assert(kO2NL == 8, "one copy per MFMA of the first half");
    v8i xa[4], xb[4];
    auto rda = [&](const signed char *st, int a) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const v4i f = lds_frag(st + (q * 8 + 2 * wi + (a >> 1)) * kChunk, 32 * (a & 1) + r, half);
#pragma unroll
        for (int e = 0; e < 4; ++e) xa[a][4 * q + e] = f[e];
      }
    };
    auto rdb = [&](const signed char *st, int b) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const v4i f = lds_frag(st + (q * 8 + 4 + 2 * wj + (b >> 1)) * kChunk, 32 * (b & 1) + r, half);
#pragma unroll
        for (int e = 0; e < 4; ++e) xb[b][4 * q + e] = f[e];
      }
    };
    wait_stage(0);
    lds_barrier();
#pragma unroll
    for (int a = 0; a < 4; ++a) rda(lds, a);
#pragma unroll
    for (int b = 0; b < 4; ++b) rdb(lds, b);
    for (int64_t s = 0; s < NS; ++s) {
      const bool more = s + kO2Bufs - 1 < NS;
      // unconditional reloads, also after the last stage (a stale buffer
      // inside the allocation, unused)
      const signed char *sn = lds + (int)((s + 1) % kO2Bufs) * kO2Stage;
#pragma unroll
      for (int x = 0; x < 16; ++x) {
        const int a = x >> 2, b = x & 3;
        if (x == 8) {
          wait_stage(s + 1);
          lds_barrier();
          rda(sn, 0);
          rda(sn, 1);
          __builtin_amdgcn_sched_barrier(0);
        }
        acc[a][b] = O2Mma<MODE>::run(xa[a], xb[b], acc[a][b]);
        if (x < 8 && more) issue(s + kO2Bufs - 1, x);
        if (x == 11) rda(sn, 2);
        if (a == 3) rdb(sn, b);
        if (x == 15) rda(sn, 3);
        if (x >= 8) __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // Accumulator e of block (a, b) is row 8(e/4) + 4 half + e%4 and column r of
  // that 32 x 32 block.
  if constexpr (MODE == kOzRes) {
    // v == c_t (mod m_t); the CRT digit is u = v y_t mod m_t with the call's
    // final moduli count N. w = v y_t is exact in a double (|w| < 2^37), and
    // q = RN(w / m) is exact (w / m is an exact binary fraction for m = 256
    // and at least 1/(2m) from a half-integer otherwise, far beyond the
    // product's rounding), so the fma gives the balanced digit w - q m in
    // [-m/2, m/2] exactly; it is stored as a signed byte (+128 for m = 256
    // wraps to -128, the same residue).
    const int64_t m = g.m, n = g.n;
    const double md = (double)mf_oz2_m[t], inv = 1.0 / md;
    const double yt = (double)mf_oz2_y[(Nc - 1) * MF_OZ2_MODULI + t];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int64_t col = (int64_t)kO2Tile * jb + 128 * wj + 32 * b + r;
        if (col >= n) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t row0 = (int64_t)kO2Tile * ib + 128 * wi + 32 * a + 8 * q + 4 * half;
          unsigned word = 0;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const double v = (double)acc[a][b][4 * q + e] * yt;
            const double u = __builtin_fma(-__builtin_rint(v * inv), md, v);
            word |= ((unsigned)(int)u & 255u) << (8 * e);
          }
          if (row0 + 3 < m && (m & 3) == 0) {
            *(unsigned *)(gr + col * m + row0) = word;
          } else {
            for (int e = 0; e < 4; ++e)
              if (row0 + e < m) gr[col * m + row0 + e] = (unsigned char)(word >> (8 * e));
          }
        }
      }
  } else {
    // Certificate: the slices each output needs (oz_cert_slices). Outputs
    // that need more than the first residue pass's count are appended to the
    // candidate list (one atomic per wave and 32-column block) and counted by
    // slices; rows and columns that are empty (exact zero outputs) or hold
    // Inf/NaN (NaN outputs) need none.
    const int64_t m = g.m, n = g.n;
    const int Sf = g.ctl->S;
    unsigned long long rmx = 0ull, cmx = 0ull;  // largest row / column sums (live ones)
    // the lane's 4 columns, then per 32-row block its 16 rows: each row's and
    // column's exponent and sum are loaded once for the tile
    bool cl[4];
    double rb[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t col = (int64_t)kO2Tile * jb + 128 * wj + 32 * b + r;
      const int Fj = col < n ? g.F[col] : kEmpty;
      cl[b] = Fj != kEmpty && Fj != kNonFinite;
      const unsigned long long csj = cl[b] ? g.csum[col] : 0ull;
      cmx = csj > cmx ? csj : cmx;
      rb[b] = (double)csj * 0x1p-40;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      double ra[16];
      unsigned live = 0;  // bit e: row e of this block holds finite, nonzero entries
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t row = (int64_t)kO2Tile * ib + 128 * wi + 32 * a + 8 * (e >> 2) + 4 * half + (e & 3);
        const int Ei = row < m ? g.E[row] : kEmpty;
        const bool rl = Ei != kEmpty && Ei != kNonFinite;
        const unsigned long long rsi = rl ? g.rsum[row] : 0ull;
        rmx = rsi > rmx ? rsi : rmx;
        ra[e] = (double)rsi * 0x1p-40;
        live |= (unsigned)rl << e;
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        unsigned need = 0;  // outputs of this 32 x 32 block over the first count: 16 per lane
        unsigned long long packed[2] = {0ull, 0ull};
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int sl = cl[b] && ((live >> e) & 1u) ? oz_cert_slices(acc[a][b][e] * 0x1p-14f, ra[e], rb[b]) : 0;
          if (sl > Sf) {
            ++need;
            packed[e >> 3] |= (unsigned long long)sl << (8 * (e & 7));
          }
        }
        if (__ballot(need != 0)) {  // uniform: the candidates of this block
          const int64_t col = (int64_t)kO2Tile * jb + 128 * wj + 32 * b + r;
          unsigned total;
          const unsigned before = wave_excl_scan(need, lane, total);
          unsigned long long base = 0;
          if (lane == 0) base = atomicAdd(&g.ctl->ncand, (unsigned long long)total);
          base = __shfl(base, 0);
          unsigned long long slot = base + before;
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            const int sl = (int)((packed[e >> 3] >> (8 * (e & 7))) & 255ull);
            if (sl) {
              const int64_t row = (int64_t)kO2Tile * ib + 128 * wi + 32 * a + 8 * (e >> 2) + 4 * half + (e & 3);
              g.cand[slot++] = ((unsigned long long)sl << 56) | (unsigned long long)(row + col * m);
              atomicAdd(&hist_l[sl], 1u);
            }
          }
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long ro = __shfl_xor(rmx, off), co = __shfl_xor(cmx, off);
      rmx = ro > rmx ? ro : rmx;
      cmx = co > cmx ? co : cmx;
    }
    if (lane == 0) {
      if (rmx) atomicMax(&g.ctl->rmax, rmx);
      if (cmx) atomicMax(&g.ctl->cmax, cmx);
    }
    __syncthreads();
    if (tid < kOzBins && hist_l[tid]) atomicAdd(&g.ctl->hist[tid], (unsigned long long)hist_l[tid]);
  }
}

// CRT: x / M = frac(sum_t u_t / m_t) with the balanced digits
// u_t == c_t y_t (mod m_t), |u_t| <= 128, that the GEMM epilogue stored, summed
// exactly in five 36-bit fixed-point limbs of floor(2^180 / m_t) (each
// |u_t L_tj| < 2^43; N <= 40 terms stay below 2^49, so the double sums are
// exact), carried (floor, so negative limbs carry correctly), centred into
// [-1/2, 1/2), then
// scaled: C = (x / M) M 2^(E_i + F_j - 12 - 16(S - 1)).
// The truncated 1/m_t put an error below N 2^(8 - 180) on x / M, i.e. below
// 2^(16S + log2 k - 159) in units of x (M < 2^(16S + log2 k + 7.1)); the
// digit representation's own error in x is of order sqrt(k) 2^(8S - 2), which
// is larger by 2^21 or more for S <= 16 and k <= 32768.
// Each thread finishes VEC consecutive outputs (column-major, like C): one
// VEC-byte load per residue plane, eight planes in flight.
constexpr int kO2Limbs = 5;
template <int VEC>
__global__ __launch_bounds__(kThreads) void k_oz2_crt(const unsigned char *__restrict__ R,
                                                      const OzCtl *__restrict__ ctl, int64_t m, int64_t n,
                                                      const int *__restrict__ E, const int *__restrict__ F,
                                                      double2 *__restrict__ C) {
  const int N = ctl->N;
  const Oz2Consts cs = ctl->cs;
  const int64_t mn = m * n, groups = mn / VEC;  // VEC divides m n
  for (int64_t gi = (int64_t)blockIdx.x * kThreads + threadIdx.x; gi < groups;
       gi += (int64_t)gridDim.x * kThreads) {
    const int64_t o0 = gi * VEC;
    double L[VEC][kO2Limbs];
#pragma unroll
    for (int e = 0; e < VEC; ++e)
#pragma unroll
      for (int l = 0; l < kO2Limbs; ++l) L[e][l] = 0.0;
    for (int t0 = 0; t0 < N; t0 += 8) {
      unsigned w[8];
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        // planes past N re-read plane N - 1 (in bounds; their words are unused)
        const unsigned char *src = R + (int64_t)(t0 + g < N ? t0 + g : N - 1) * mn + o0;
        w[g] = VEC == 4 ? *(const unsigned *)src : (unsigned)*src;
      }
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        const int t = t0 + g;
        if (t < N) {
          const double *rl = mf_oz2_rlimb + kO2Limbs * t;
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const double ud = (double)(signed char)((w[g] >> (8 * e)) & 255u);
#pragma unroll
            for (int l = 0; l < kO2Limbs; ++l) L[e][l] = __builtin_fma(ud, rl[l], L[e][l]);
          }
        }
      }
    }
    int64_t j = o0 / m, i = o0 - j * m;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      double *Lr = L[e];
      const double T36 = 0x1p36, iT36 = 0x1p-36;
#pragma unroll
      for (int l = kO2Limbs - 1; l > 0; --l) {
        const double c = __builtin_floor(Lr[l] * iT36);
        Lr[l] -= c * T36;
        Lr[l - 1] += c;
      }
      Lr[0] -= __builtin_floor(Lr[0] * iT36) * T36;  // integer part dropped: L0 in [0, 2^36)
      // Centre x / M into [-1/2, 1/2): a fraction of 1/2 or more stands for
      // minus its 180-bit complement (limbwise 2^36 - 1 - L_l, plus one unit
      // at the bottom, carried), so every limb is non-negative.
      const bool neg = Lr[0] >= 0x1p35;
      if (neg) {
#pragma unroll
        for (int l = 0; l < kO2Limbs; ++l) Lr[l] = (T36 - 1.0) - Lr[l];
        Lr[kO2Limbs - 1] += 1.0;
#pragma unroll
        for (int l = kO2Limbs - 1; l > 0; --l) {
          const double c = __builtin_floor(Lr[l] * iT36);
          Lr[l] -= c * T36;
          Lr[l - 1] += c;
        }
      }
      // |x / M| = sum_l L_l 2^(-36 (l + 1)), non-overlapping non-negative
      // terms, summed from the bottom: a nonzero term is at least
      // 2^(-36 (l + 1)), no less than the rounded sum of those below it, so
      // every fast_two_sum is exact; the errors, decreasing by 2^36 a level,
      // form the low part.
      static_assert(kO2Limbs == 5, "five limbs below");
      double sm = Lr[4] * 0x1p-180, e3, e2, e1, e0;
      fast_two_sum(Lr[3] * 0x1p-144, sm, sm, e3);
      fast_two_sum(Lr[2] * 0x1p-108, sm, sm, e2);
      fast_two_sum(Lr[1] * 0x1p-72, sm, sm, e1);
      fast_two_sum(Lr[0] * 0x1p-36, sm, sm, e0);
      ddv f;
      fast_two_sum(sm, e0 + (e1 + (e2 + e3)), f.hi, f.lo);
      if (neg) f = mk(-f.hi, -f.lo);
      const int ei = E[i], fj = F[j];
      double2 out;
      if (ei == kNonFinite || fj == kNonFinite) {
        out = make_double2(__builtin_nan(""), 0.0);
      } else {
        const ddv v = dd_mul(f, mk(cs.mp_hi, cs.mp_lo));
        const int sc = scale_of(ei) + scale_of(fj) - 12;
        out = make_double2(__builtin_ldexp(v.hi, sc), __builtin_ldexp(v.lo, sc));
      }
      C[o0 + e] = out;
      if (++i == m) {
        i = 0;
        ++j;
      }
    }
  }
}

// Adds x to the value s1 + s2 + s3 kept as a cascade: (s1, e) = TwoSum(s1, x),
// (s2, e2) = TwoSum(s2, e), s3 += e2. Only the last addition rounds; after n
// additions of terms with sum |x| = T, |s2| <= n u T and the rounding errors
// of s3 total at most (n u)^3 T (u = 2^-53).
__device__ __forceinline__ void casc3(double &s1, double &s2, double &s3, double x) {
  double e, e2;
  two_sum(s1, x, s1, e);
  two_sum(s2, e, s2, e2);
  s3 += e2;
}

// 64-bit value of lane 0, uniform for the compiler (two readfirstlanes).
__device__ __forceinline__ unsigned long long rfl64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Outputs of an automatic call that the slice count in force does not
// certify (candidates needing more slices than ctl->S), recomputed one per
// wave: lane l takes p = l, l + 64, ...; each product a b of double-doubles is
// ph + q with ph = RN(ah bh), q = RN(fma(ah, bh, -ph) + fma(ah, bl, RN(al bh)))
// (al bl dropped), within 7 2^-106 |a||b| of the product; the ph go through a
// three-term cascade and the q through a two-term one, the lanes' cascades
// are merged by a butterfly, and s1 + s2 + s3 is rounded to a double-double.
// Error: 7 2^-106 (|A||B|)_ij from the products, (8192 u)^3 from the
// cascades, 2^-104.9 from the final rounding: below 2^-102.6 (|A||B|)_ij,
// with no dependence on how the row's and column's magnitudes are spread.
// Wave w of the grid takes list entries w, w + W, w + 2W, ... (W waves): the
// loop index, the bound and each entry are wave-uniform scalars, so the loop
// keeps one level of uniform control flow (a shared work counter whose ticket
// is broadcast from one lane makes the loop test divergent for the compiler,
// which then splits it into nested loops in which lanes can run on a stale
// ticket and never leave). Entries cost the same k-long dot, so the static
// order balances.
__global__ __launch_bounds__(kThreads) void k_oz_fixup(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                       double2 *__restrict__ C, int64_t m, int64_t k,
                                                       const unsigned long long *__restrict__ cand,
                                                       OzCtl *__restrict__ ctl) {
  const unsigned long long nc = rfl64(ctl->ncand);
  const int S = __builtin_amdgcn_readfirstlane(ctl->S);
  const int lane = threadIdx.x & 63;
  const unsigned long long nw = (unsigned long long)gridDim.x * (kThreads / 64);
  const unsigned long long w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6));
  for (unsigned long long idx = w0; idx < nc; idx += nw) {
    const unsigned long long ent = rfl64(cand[idx]);
    if ((int)(ent >> 56) <= S) continue;  // certified by the slice count in force
    const int64_t o = (int64_t)(ent & ((1ull << 56) - 1));
    const int64_t j = o / m, i = o - j * m;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, t1 = 0.0, t2 = 0.0;
    const double2 *arow = A + i, *bcol = B + j * k;
    for (int64_t p = lane; p < k; p += 64) {
      const double2 a = arow[p * m], b = bcol[p];
      const double ph = a.x * b.x;
      const double pe = __builtin_fma(a.x, b.x, -ph);
      const double q = pe + __builtin_fma(a.x, b.y, a.y * b.x);
      casc3(s1, s2, s3, ph);
      double e;
      two_sum(t1, q, t1, e);
      t2 += e;
    }
    casc3(s1, s2, s3, t1);
    casc3(s1, s2, s3, t2);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o1 = __shfl_xor(s1, off), o2 = __shfl_xor(s2, off), o3 = __shfl_xor(s3, off);
      casc3(s1, s2, s3, o1);
      casc3(s1, s2, s3, o2);
      casc3(s1, s2, s3, o3);
    }
    if (lane == 0) {
      double h, l1, l2;
      two_sum(s1, s2, h, l1);
      two_sum(h, s3, h, l2);
      double hi, lo;
      two_sum(h, l1 + l2, hi, lo);  // h may have cancelled below l1 + l2
      C[o] = make_double2(hi, lo);
      atomicAdd(&ctl->nfix, 1ull);
    }
  }
}

// C = 0 (m n elements) by a fill kernel, which graph capture records safely.
__global__ void k_zero2(double2 *__restrict__ C, int64_t count) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < count; x += (int64_t)gridDim.x * blockDim.x)
    C[x] = make_double2(0.0, 0.0);
}

}  // namespace

namespace mfhip {

int ozaki_max_slices() { return kMaxSlices; }

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

namespace {
// CRT form: moduli needed for S digits and inner dimension k, i.e. the
// smallest N with M = prod_{t<N} m_t > 2 k 2^(16S-2) (with margin); 0 if the
// 40 moduli do not reach it (they reach it for every S <= 16, k <= 32768).
int oz2_count(int S, int64_t k) {
  const double need = 16.0 * S - 1.0 + log2((double)(k > 0 ? k : 1)) + 0.05;
  double have = 0.0;
  for (int t = 0; t < MF_OZ2_MODULI; ++t) {
    have += log2((double)mf_oz2_m_host[t]);
    if (have > need) return t + 1;
  }
  return 0;
}
OzNtab oz2_ntab(int64_t k) {
  OzNtab t{};
  for (int S = kMinSlices; S <= kMaxSlices; ++S) t.n[S] = oz2_count(S, k);
  return t;
}
// With an explicit slice count the CRT form runs once the 256 x 256 tiles
// are mostly full and the triangular form below that; an automatic call
// always runs the CRT form (its slice count is chosen on the device, and the
// triangular kernel is specialised on it).
bool use_oz2(int64_t m, int64_t n, int64_t k, int S) {
  return m >= 512 && n >= 512 && oz2_count(S, k) > 0;
}
struct Oz2Dims {
  int N;
  int64_t TBA, TBB, KBp;  // 64-row blocks of A / B padded to whole tiles; K blocks padded to kO2KS
};
Oz2Dims oz2_dims(int64_t m, int64_t k, int64_t n, int S) {
  Oz2Dims d;
  d.N = oz2_count(S, k);
  d.TBA = (m + kO2Tile - 1) / kO2Tile * 4;
  d.TBB = (n + kO2Tile - 1) / kO2Tile * 4;
  const int64_t KB = (k + kK - 1) / kK;
  d.KBp = (KB + kO2KS - 1) / kO2KS * kO2KS;
  return d;
}
// residue planes of A and B, digit planes; with `cert` (automatic calls) also
// the FP8 magnitude operands (the residues' chunk layout: KBp K blocks) and the
// candidate list (one 8-byte entry per output at most)
size_t oz2_bytes(int64_t m, int64_t k, int64_t n, int S, bool cert) {
  const Oz2Dims d = oz2_dims(m, k, n, S);
  size_t b = align256((size_t)d.N * d.TBA * d.KBp * kChunk) + align256((size_t)d.N * d.TBB * d.KBp * kChunk) +
             align256((size_t)d.N * m * n);
  if (cert)
    b += align256((size_t)d.TBA * d.KBp * kChunk) + align256((size_t)d.TBB * d.KBp * kChunk) +
         align256(sizeof(unsigned long long) * (size_t)m * n);
  return b;
}
// Bytes of the parts that do not depend on the form: exponent arrays,
// ranges, control block, certificate sums.
size_t oz_head_bytes(int64_t m, int64_t n) {
  return 2 * align256(sizeof(int) * (size_t)m) + align256(sizeof(int) * (size_t)n) + 256 +
         align256(sizeof(unsigned long long) * (size_t)m) + align256(sizeof(OzCtl)) +
         align256(sizeof(unsigned long long) * (size_t)m) + align256(sizeof(unsigned long long) * (size_t)n);
}
}  // namespace

int ozaki_moduli(int64_t m, int64_t k, int64_t n, int S) {
  return (S >= kMinSlices && S <= kMaxSlices && use_oz2(m, n, k, S)) ? oz2_count(S, k) : 0;
}

size_t ozaki_workspace_bytes(int64_t m, int64_t k, int64_t n, int S) {
  if (m < 0) m = 0;
  if (n < 0) n = 0;
  // automatic calls: the CRT form at kMaxSlices with the certificate; also
  // enough for any explicit count, so one workspace serves both
  const bool aut = S <= 0;
  if (aut) S = kMaxSlices;
  const int64_t tm = (m + kT - 1) / kT, tn = (n + kT - 1) / kT, KB = (k + kK - 1) / kK;
  const size_t one = align256((size_t)tm * KB * S * kChunk) + align256((size_t)tn * KB * S * kChunk);
  const size_t two = oz2_count(S, k) ? oz2_bytes(m, k, n, S, aut) : 0;
  return oz_head_bytes(m, n) + (one > two ? one : two);
}

namespace {
struct OzWork {
  signed char *As, *Bs;
  int *E, *Emin, *F, *range;  // range[0] rows of A, range[1] columns of B
  unsigned long long *Mbits;  // largest |hi| per row of A (bit pattern)
  OzCtl *ctl;
  unsigned long long *rsum, *csum;  // certificate: sum |a| 2^-E per row of A / column of B
  int64_t tm, tn, KB;
};
OzWork carve(void *work, int64_t m, int64_t k, int64_t n, int S) {
  OzWork w;
  w.tm = (m + kT - 1) / kT;
  w.tn = (n + kT - 1) / kT;
  w.KB = (k + kK - 1) / kK;
  // the parts every form uses first, at offsets independent of S
  char *p = (char *)work;
  w.E = (int *)p;
  p += align256(sizeof(int) * (size_t)m);
  w.Emin = (int *)p;
  p += align256(sizeof(int) * (size_t)m);
  w.F = (int *)p;
  p += align256(sizeof(int) * (size_t)n);
  w.range = (int *)p;
  p += 256;
  w.Mbits = (unsigned long long *)p;
  p += align256(sizeof(unsigned long long) * (size_t)m);
  w.ctl = (OzCtl *)p;
  p += align256(sizeof(OzCtl));
  w.rsum = (unsigned long long *)p;
  p += align256(sizeof(unsigned long long) * (size_t)m);
  w.csum = (unsigned long long *)p;
  p += align256(sizeof(unsigned long long) * (size_t)n);
  w.As = (signed char *)p;
  p += align256((size_t)w.tm * w.KB * (S > 0 ? S : kMaxSlices) * kChunk);
  w.Bs = (signed char *)p;
  return w;
}
// CRT-form residue planes of A and B, digit planes and (automatic calls, with
// planes for kMaxSlices) the certificate's operands and candidate list, in
// place of the slices
struct Oz2Work {
  signed char *As, *Bs;
  unsigned char *R;
  unsigned char *magA, *magB;
  unsigned long long *cand;
  Oz2Dims d;
};
Oz2Work carve2(const OzWork &w, int64_t m, int64_t k, int64_t n, int S) {
  Oz2Work v;
  v.d = oz2_dims(m, k, n, S > 0 ? S : kMaxSlices);
  char *p = (char *)w.As;
  v.As = (signed char *)p;
  p += align256((size_t)v.d.N * v.d.TBA * v.d.KBp * kChunk);
  v.Bs = (signed char *)p;
  p += align256((size_t)v.d.N * v.d.TBB * v.d.KBp * kChunk);
  v.R = (unsigned char *)p;
  p += align256((size_t)v.d.N * m * n);
  v.magA = (unsigned char *)p;
  p += align256((size_t)v.d.TBA * v.d.KBp * kChunk);
  v.magB = (unsigned char *)p;
  p += align256((size_t)v.d.TBB * v.d.KBp * kChunk);
  v.cand = (unsigned long long *)p;
  if (S > 0) v.magA = v.magB = nullptr, v.cand = nullptr;  // explicit calls have no certificate
  return v;
}
bool oz_args_ok(int64_t k, int S) { return S >= kMinSlices && S <= kMaxSlices && k <= 32768; }

// Exponent scan of A's rows and B's columns (every form).
void oz_scan(const double2 *A, const double2 *B, int64_t m, int64_t k, int64_t n, const OzWork &w, hipStream_t s) {
  const int64_t mx = m > n ? m : n;
  hipLaunchKernelGGL(k_scan_init, dim3((unsigned)((mx + 255) / 256)), dim3(256), 0, s, w.E, w.Emin, w.Mbits, w.range,
                     w.rsum, w.csum, m, n);
  const int64_t kchunk = 64, kchunks = (k + kchunk - 1) / kchunk, rblocks = (m + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_scan_exp, dim3((unsigned)(rblocks * kchunks + n)), dim3(kThreads), 0, s, A, B, m, k, kchunk,
                     kchunks, rblocks, w.E, w.Emin, w.Mbits, w.F, w.range);
  hipLaunchKernelGGL(k_row_finish, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, w.E, w.Emin, w.Mbits, m,
                     w.range);
}

// Residue pass of the CRT form over both operands (pass 0: S from the first
// plan, with the certificate operands when v.magA is set; pass 1: again if
// the final plan raised S).
void oz2_resid(const double2 *A, const double2 *B, int64_t m, int64_t k, int64_t n, const OzWork &w,
               const Oz2Work &v, int pass, hipStream_t s) {
  const bool mag = pass == 0 && v.magA != nullptr;
  oz2_resid_launch(false, s, A, m, k, w.E, v.d.KBp, v.d.TBA, w.ctl, pass, v.As, mag ? v.magA : nullptr, w.rsum);
  oz2_resid_launch(true, s, B, n, k, w.F, v.d.KBp, v.d.TBB, w.ctl, pass, v.Bs, mag ? v.magB : nullptr, w.csum);
}

Oz2GemmArgs oz2_args(const OzWork &w, const Oz2Work &v, int64_t m, int64_t n) {
  Oz2GemmArgs g{};
  g.m = m;
  g.n = n;
  g.TBA = v.d.TBA;
  g.TBB = v.d.TBB;
  g.tiles_m = (int)(v.d.TBA / 4);
  g.tiles_n = (int)(v.d.TBB / 4);
  g.ctl = w.ctl;
  g.E = w.E;
  g.F = w.F;
  g.rsum = w.rsum;
  g.csum = w.csum;
  return g;
}

template <int MODE>
hipError_t oz2_launch_gemm(const Oz2GemmArgs &g, int planes, hipStream_t s) {
  const size_t lds = (size_t)kO2Bufs * kO2Stage;
  hipError_t e =
      hipFuncSetAttribute((const void *)k_oz2_gemm<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_oz2_gemm<MODE>, dim3((unsigned)(g.tiles_m * g.tiles_n), (unsigned)planes),
                     dim3(kMainThreads), lds, s, g);
  return hipGetLastError();
}

// The residue GEMMs of the CRT form (planes beyond the call's moduli count
// leave at once), from the residue planes into R.
hipError_t oz2_products(const OzWork &w, const Oz2Work &v, int64_t m, int64_t n, hipStream_t s) {
  Oz2GemmArgs g = oz2_args(w, v, m, n);
  g.As = v.As;
  g.Bs = v.Bs;
  g.R = v.R;
  g.KB = v.d.KBp;
  return oz2_launch_gemm<kOzRes>(g, v.d.N, s);
}

// The Chinese-remainder reconstruction of C from the digit planes.
void oz2_reconstruct(const OzWork &w, const Oz2Work &v, double2 *C, int64_t m, int64_t n, hipStream_t s) {
  if ((m * n) % 4 == 0) {
    const int64_t g = (m * n / 4 + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_oz2_crt<4>, dim3((unsigned)(g < 65536 ? g : 65536)), dim3(kThreads), 0, s, v.R, w.ctl, m,
                       n, w.E, w.F, C);
  } else {
    const int64_t g = (m * n + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_oz2_crt<1>, dim3((unsigned)(g < 65536 ? g : 65536)), dim3(kThreads), 0, s, v.R, w.ctl, m,
                       n, w.E, w.F, C);
  }
}

// Fix-up budget of an automatic call: how many uncertified outputs the fix-up
// kernel takes before the plan raises S instead (each output costs one
// wave's k-long compensated dot; a higher S costs a second residue pass and
// more residue GEMMs).
unsigned long long oz_fix_budget(int64_t m, int64_t n) {
  const unsigned long long b = (unsigned long long)m * (unsigned long long)n / 1024ull;
  return b > 4096ull ? b : 4096ull;
}
constexpr int kFixupGroups = 512;
}  // namespace

hipError_t ozaki_scan(const double2 *A, const double2 *B, int64_t m, int64_t k, int64_t n,
                      int S, void *work, size_t work_bytes, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
  if (!oz_args_ok(k, S) || work_bytes < ozaki_workspace_bytes(m, k, n, S)) return hipErrorInvalidValue;
  oz_scan(A, B, m, k, n, carve(work, m, k, n, S), s);
  return hipGetLastError();
}

int ozaki_choose_slices(int64_t k, int range_a, int range_b) { return oz_guess_slices(k, range_a, range_b); }

hipError_t ozaki_slice(const double2 *A, const double2 *B, int64_t m, int64_t k, int64_t n, int S,
                       void *work, size_t work_bytes, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
  if (!oz_args_ok(k, S) || work_bytes < ozaki_workspace_bytes(m, k, n, S)) return hipErrorInvalidValue;
  OzWork w = carve(work, m, k, n, S);
  if (use_oz2(m, n, k, S)) {
    const Oz2Work v = carve2(w, m, k, n, S);
    hipLaunchKernelGGL(k_oz_plan_begin, dim3(1), dim3(64), 0, s, w.ctl, w.range, oz2_ntab(k), S, k);
    oz2_resid(A, B, m, k, n, w, v, 0, s);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_slice_a<16>, dim3((unsigned)w.tm, (unsigned)w.KB), dim3(kThreads), 0, s, A, m, k, w.E, S, w.KB, w.As);
  hipLaunchKernelGGL(k_slice_b<16>, dim3((unsigned)w.tn, (unsigned)w.KB), dim3(kThreads), 0, s, B, k, n, w.F, S, w.KB, w.Bs);
  return hipGetLastError();
}

hipError_t ozaki_prepare(const double2 *A, const double2 *B, int64_t m, int64_t k, int64_t n,
                         int S, void *work, size_t work_bytes, hipStream_t s) {
  hipError_t e = ozaki_scan(A, B, m, k, n, S, work, work_bytes, s);
  if (e != hipSuccess) return e;
  return ozaki_slice(A, B, m, k, n, S, work, work_bytes, s);
}

hipError_t ozaki_products(int64_t m, int64_t k, int64_t n, int S, void *work, size_t work_bytes, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
  if (S <= 0) {
    // the residue GEMMs of the last automatic call on `work`, with the slice
    // and moduli counts its plan chose (read from the control block)
    if (k > 32768 || work_bytes < ozaki_workspace_bytes(m, k, n, 0) || oz2_count(kMaxSlices, k) <= 0)
      return hipErrorInvalidValue;
  } else if (!oz_args_ok(k, S) || work_bytes < ozaki_workspace_bytes(m, k, n, S) || !use_oz2(m, n, k, S)) {
    return hipErrorInvalidValue;
  }
  const OzWork w = carve(work, m, k, n, S);
  return oz2_products(w, carve2(w, m, k, n, S), m, n, s);
}

hipError_t ozaki_multiply(double2 *C, int64_t m, int64_t k, int64_t n, int S, void *work,
                          size_t work_bytes, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (k <= 0) {
    hipLaunchKernelGGL(k_zero2, dim3(1024), dim3(256), 0, s, C, m * n);
    return hipGetLastError();
  }
  if (!oz_args_ok(k, S) || work_bytes < ozaki_workspace_bytes(m, k, n, S)) return hipErrorInvalidValue;
  OzWork w = carve(work, m, k, n, S);
  if (use_oz2(m, n, k, S)) {
    const Oz2Work v = carve2(w, m, k, n, S);
    hipError_t e = oz2_products(w, v, m, n, s);
    if (e != hipSuccess) return e;
    oz2_reconstruct(w, v, C, m, n, s);
    return hipGetLastError();
  }
  switch (S) {
    case 7: return launch_main<7>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    case 8: return launch_main<8>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    case 9: return launch_main<9>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    case 10: return launch_main<10>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    case 11: return launch_main<11>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    case 12: return launch_main<12>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    case 13: return launch_main<13>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    case 14: return launch_main<14>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    case 15: return launch_main<15>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
    default: return launch_main<16>(w.As, w.Bs, w.E, w.F, C, m, n, w.KB, w.tm, w.tn, s);
  }
}

// An automatic call: every stage enqueued, the slice count chosen on the
// device (no synchronisation, so the call can be captured in a graph).
//   scan -> first plan (exponent-range rule) -> residues + certificate
//   operands -> certificate GEMM -> final plan -> residues again only if S
//   was raised -> residue GEMMs -> reconstruction -> fix-up of the outputs
//   the final S does not certify.
hipError_t ozaki_auto(const double2 *A, const double2 *B, double2 *C, int64_t m, int64_t k, int64_t n, void *work,
                      size_t work_bytes, hipStream_t s) {
  if (k > 32768 || work_bytes < ozaki_workspace_bytes(m, k, n, 0)) return hipErrorInvalidValue;
  const OzWork w = carve(work, m, k, n, 0);
  const Oz2Work v = carve2(w, m, k, n, 0);
  const OzNtab nt = oz2_ntab(k);
  oz_scan(A, B, m, k, n, w, s);
  hipLaunchKernelGGL(k_oz_plan_begin, dim3(1), dim3(64), 0, s, w.ctl, w.range, nt, 0, k);
  oz2_resid(A, B, m, k, n, w, v, 0, s);
  Oz2GemmArgs g = oz2_args(w, v, m, n);
  g.As = (const signed char *)v.magA;
  g.Bs = (const signed char *)v.magB;
  g.R = nullptr;
  g.KB = v.d.KBp;
  g.cand = v.cand;
  hipError_t e = oz2_launch_gemm<kOzMag>(g, 1, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_oz_plan_final, dim3(1), dim3(64), 0, s, w.ctl, nt, oz_fix_budget(m, n), k);
  oz2_resid(A, B, m, k, n, w, v, 1, s);
  e = oz2_products(w, v, m, n, s);
  if (e != hipSuccess) return e;
  oz2_reconstruct(w, v, C, m, n, s);
  hipLaunchKernelGGL(k_oz_fixup, dim3(kFixupGroups), dim3(kThreads), 0, s, A, B, C, m, k, v.cand, w.ctl);
  return hipGetLastError();
}

hipError_t ozaki_gemm(const double2 *A, const double2 *B, double2 *C, int64_t m, int64_t k,
                      int64_t n, int S, void *work, size_t work_bytes, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (k <= 0) {
    hipLaunchKernelGGL(k_zero2, dim3(1024), dim3(256), 0, s, C, m * n);
    return hipGetLastError();
  }
  if (S <= 0) return ozaki_auto(A, B, C, m, k, n, work, work_bytes, s);
  hipError_t e = ozaki_prepare(A, B, m, k, n, S, work, work_bytes, s);
  if (e != hipSuccess) return e;
  return ozaki_multiply(C, m, k, n, S, work, work_bytes, s);
}

hipError_t ozaki_info(const void *work, int64_t m, int64_t k, int64_t n, int S, int *slices, int64_t *fixups,
                      int *moduli, hipStream_t s) {
  if (moduli) *moduli = 0;
  if (m <= 0 || n <= 0 || k <= 0) {
    if (slices) *slices = 0;
    if (fixups) *fixups = 0;
    return hipSuccess;
  }
  const OzWork w = carve(const_cast<void *>(work), m, k, n, S);
  if (!use_oz2(m, n, k, S) && S > 0) {  // triangular form: no control block
    if (slices) *slices = S;
    if (fixups) *fixups = 0;
    return hipSuccess;
  }
  int hS = 0, hN = 0;
  unsigned long long hf = 0;
  hipError_t e = hipMemcpyAsync(&hS, &w.ctl->S, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&hN, &w.ctl->N, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&hf, &w.ctl->nfix, sizeof(hf), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  if (slices) *slices = hS;
  if (fixups) *fixups = (int64_t)hf;
  // automatic calls always run the CRT form; an explicit count runs it when
  // use_oz2 says so (the triangular form has no moduli)
  if (moduli) *moduli = (S <= 0 || use_oz2(m, n, k, hS)) ? hN : 0;
  return hipSuccess;
}

}  // namespace mfhip
