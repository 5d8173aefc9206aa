// DD reductions: dot_product, sum, norm2, product, maxval, minval (rank 1,
// and along one dimension).
//
// The reference computes these as single sequential Fortran loops
// (fsrc/multifloats.fypp: dd_dot_product 2254-2308, dd_sum_1d 1980-2020,
// dd_norm2_1d 2338-2352), a latency-bound chain through s_hi.
//
// Rank 1: pass 1 gives every workgroup a contiguous span of 256*E elements;
// lane l runs the reference's own per-element update over elements
// l, l+256, ... of the span (E of them, all loaded up front with
// nontemporal 16-byte loads, so each lane's partial carries the same
// compensation the reference's accumulator does); the lanes meet in a fixed
// LDS tree of accurate DD additions. Fold passes then combine 2048 partials
// per workgroup the same way until one remains. Every launch shape depends
// only on n, so results are reproducible run to run. The reassociation means
// results are not the sequential loop's bits; they are held to the
// c*n*2^-104*sum|term| bound against an exact accumulator instead.
// (Measured on MI355X, N = 2^26 dot: ~6.7 TB/s; the earlier grid-stride form
// with a 2048-workgroup cap reached ~4.7 TB/s.)
#include "dd_core.h"
#include "mf_internal.h"

namespace {

constexpr int kThreads = 256;

enum { R_DOT = 0, R_SUM = 1, R_NORM2 = 2, R_SUM_MASK = 3, R_PROD = 4, R_PROD_MASK = 5, R_MAXV = 6, R_MINV = 7 };

// Fortran dot update (fypp:2286-2302): cross = ah*bl + al*bh without an FMA,
// s_lo += (((e + cross) + aerr) + berr).
__device__ __forceinline__ void dot_step(double2 a, double2 b, double &sh, double &sl) {
  double p = a.x * b.x;
  double e = __builtin_fma(a.x, b.x, -p);
  double cross = a.x * b.y + a.y * b.x;
  double t = sh + p;
  double bp = t - sh;
  double ap = t - bp;
  double aerr = sh - ap;
  double berr = p - bp;
  sl = sl + (((e + cross) + aerr) + berr);
  sh = t;
}

// Fortran sum update (fypp:1998-2004): s_lo += ((x_lo + aerr) + berr).
__device__ __forceinline__ void sum_step(double2 x, double &sh, double &sl) {
  double t = sh + x.x;
  double bp = t - sh;
  double ap = t - bp;
  double aerr = sh - ap;
  double berr = x.x - bp;
  sl = sl + ((x.y + aerr) + berr);
  sh = t;
}

// Block tree: 256 normalised pairs -> one, in LDS, fixed shape; DD + (sums),
// or DD * for products (PROD).
template <bool PROD = false>
__device__ ddv block_tree(ddv v) {
  __shared__ double th[kThreads], tl[kThreads];
  const int tid = threadIdx.x;
  th[tid] = v.hi;
  tl[tid] = v.lo;
  __syncthreads();
#pragma unroll
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const ddv x = mk(th[tid], tl[tid]), y = mk(th[tid + s], tl[tid + s]);
      const ddv r = PROD ? dd_mul(x, y) : dd_add(x, y);
      th[tid] = r.hi;
      tl[tid] = r.lo;
    }
    __syncthreads();
  }
  return mk(th[0], tl[0]);
}

// Pass 1: lane-local reference accumulation over a strided slice, then the
// workgroup tree; one normalised DD partial per workgroup.
// Elements per lane in pass 1 (16-byte loads in flight per lane: dot 2E,
// the single-operand reductions E).
template <int KIND>
constexpr int elems_per_lane() {
  return KIND == R_DOT ? 8 : 16;
}
template <int KIND>
constexpr bool is_prod() {
  return KIND == R_PROD || KIND == R_PROD_MASK;
}
constexpr int kCdotE = 4;   // complex dot: 2 x 32 bytes per element
constexpr int kFold = 8;    // partials per lane in a fold pass
constexpr int kFoldSpan = kThreads * kFold;

// Pass 1. R_SUM_MASK: the sum update runs only where mask[i] != 0 (Fortran
// sum(array, mask), fypp:1993-1995: masked elements are skipped, not added).
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_reduce_partial(const double2 *__restrict__ a,
                                                             const double2 *__restrict__ b,
                                                             int64_t n, int64_t ri,
                                                             double2 *__restrict__ partial,
                                                             const int32_t *__restrict__ mask) {
  constexpr int E = elems_per_lane<KIND>();
  const int64_t base = (int64_t)blockIdx.x * kThreads * E + threadIdx.x;
  v2d va[E], vb[E];
  int32_t vm[E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int64_t i = base + (int64_t)kThreads * j;
    const bool ok = i < n;
    va[j] = ok ? __builtin_nontemporal_load((const v2d *)a + i) : v2d{0.0, 0.0};
    if constexpr (KIND == R_DOT) vb[j] = ok ? __builtin_nontemporal_load((const v2d *)b + i) : v2d{0.0, 0.0};
    if constexpr (KIND == R_SUM_MASK || KIND == R_PROD_MASK) vm[j] = ok ? __builtin_nontemporal_load(mask + i) : 0;
  }
  ddv acc;
  if constexpr (is_prod<KIND>()) {
    // product (fypp:2005-2016): res = res * x from 1, masked elements skipped
    acc = mk(1.0, 0.0);
#pragma unroll
    for (int j = 0; j < E; ++j)
      if (base + (int64_t)kThreads * j < n && (KIND == R_PROD || vm[j])) acc = dd_mul(acc, mk(va[j].x, va[j].y));
  } else if constexpr (KIND == R_NORM2) {
    acc = mk(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < E; ++j)
      if (base + (int64_t)kThreads * j < n) {
        const ddv x = mk(va[j].x, va[j].y);
        acc = dd_add(acc, dd_mul(x, x));
      }
  } else {
    double sh = 0.0, sl = 0.0;
    int64_t until = ri;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if (base + (int64_t)kThreads * j >= n) break;
      if constexpr (KIND == R_DOT) {
        dot_step(make_double2(va[j].x, va[j].y), make_double2(vb[j].x, vb[j].y), sh, sl);
        if (ri > 0 && --until == 0) {
          until = ri;
          fast_two_sum(sh, sl, sh, sl);
        }
      } else if constexpr (KIND == R_SUM_MASK) {
        if (vm[j]) sum_step(make_double2(va[j].x, va[j].y), sh, sl);
      } else {
        sum_step(make_double2(va[j].x, va[j].y), sh, sl);
      }
    }
    renorm(sh, sl);
    acc = mk(sh, sl);
  }
  const ddv r = block_tree<is_prod<KIND>()>(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(r.hi, r.lo);
}

// Fold pass: workgroup w combines partials [w*2048, (w+1)*2048) of NC-component
// values (NC = 1 real, 2 complex) through the same tree; the last pass writes
// the result (norm2 takes the DD square root there).
template <int KIND, int NC>
__global__ __launch_bounds__(kThreads) void k_fold(const double2 *__restrict__ in, int64_t g,
                                                   double2 *__restrict__ out, int last) {
  const int64_t base = (int64_t)blockIdx.x * kFoldSpan + threadIdx.x;
  constexpr bool P = is_prod<KIND>();
  ddv acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = P ? mk(1.0, 0.0) : mk(0.0, 0.0);
#pragma unroll
  for (int j = 0; j < kFold; ++j) {
    const int64_t i = base + (int64_t)kThreads * j;
    if (i < g) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const ddv x = mk(in[i * NC + c].x, in[i * NC + c].y);
        acc[c] = P ? dd_mul(acc[c], x) : dd_add(acc[c], x);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c) __syncthreads();
    ddv r = block_tree<P>(acc[c]);
    if (threadIdx.x == 0) {
      if (KIND == R_NORM2 && last) r = dd_sqrt(r);
      out[(int64_t)blockIdx.x * NC + c] = make_double2(r.hi, r.lo);
    }
  }
}

// Complex dot_product(a, b) = sum(conj(a) * b) (fypp:2321-2331): the reference
// runs four real dot chains, re = dot(ar, br) + dot(ai, bi) and
// im = dot(ar, bi) - dot(ai, br), and combines them with DD + and -. Here
// each lane keeps the four chains for its elements in registers (one pass
// over both inputs), forms its (re, im) with the same DD + / -, and the
// lanes' complex partials meet in the fixed tree.
__global__ __launch_bounds__(kThreads) void k_cdot_partial(const double4 *__restrict__ a,
                                                           const double4 *__restrict__ b, int64_t n,
                                                           int64_t ri, double2 *__restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * kThreads * kCdotE + threadIdx.x;
  v2d x[kCdotE][2], y[kCdotE][2];
#pragma unroll
  for (int j = 0; j < kCdotE; ++j) {
    const int64_t i = base + (int64_t)kThreads * j;
    const bool ok = i < n;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      x[j][h] = ok ? __builtin_nontemporal_load((const v2d *)a + 2 * i + h) : v2d{0.0, 0.0};
      y[j][h] = ok ? __builtin_nontemporal_load((const v2d *)b + 2 * i + h) : v2d{0.0, 0.0};
    }
  }
  double h[4] = {0.0, 0.0, 0.0, 0.0}, l[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t until = ri;
#pragma unroll
  for (int j = 0; j < kCdotE; ++j) {
    if (base + (int64_t)kThreads * j >= n) break;
    const double2 ar = make_double2(x[j][0].x, x[j][0].y), ai = make_double2(x[j][1].x, x[j][1].y);
    const double2 br = make_double2(y[j][0].x, y[j][0].y), bi = make_double2(y[j][1].x, y[j][1].y);
    dot_step(ar, br, h[0], l[0]);
    dot_step(ai, bi, h[1], l[1]);
    dot_step(ar, bi, h[2], l[2]);
    dot_step(ai, br, h[3], l[3]);
    if (ri > 0 && --until == 0) {
      until = ri;
#pragma unroll
      for (int q = 0; q < 4; ++q) fast_two_sum(h[q], l[q], h[q], l[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) renorm(h[q], l[q]);
  const ddv re = dd_add(mk(h[0], l[0]), mk(h[1], l[1]));
  const ddv im = dd_sub(mk(h[2], l[2]), mk(h[3], l[3]));
  const ddv r = block_tree(re);
  __syncthreads();
  const ddv q = block_tree(im);
  if (threadIdx.x == 0) {
    partial[2 * (int64_t)blockIdx.x] = make_double2(r.hi, r.lo);
    partial[2 * (int64_t)blockIdx.x + 1] = make_double2(q.hi, q.lo);
  }
}

// Reductions along one dimension, bit-identical to the reference: the array
// is viewed as (inner, len, outer) in Fortran order and every output
// out[i + inner*o] is ONE sequential chain over j = 0..len-1 of the
// reference's own update, exactly as the Fortran loop runs it (masked
// elements skipped where the reference skips them):
//   sum  (dd_sum_Nd_dim, fypp:2052-2075): the Neumaier update, then a full
//        TwoSum written in the reference's order (t, bp, ap, aerr, berr;
//        lo = aerr + berr);
//   norm2 (dd_norm2_Nd_dim, fypp:2356-2377): cur = cur + x*x with the DD
//        operators, then DD sqrt;
//   product (fypp:2076-2087): cur = cur * x from 1;
//   maxval / minval (dd_maxval_Nd_dim, fypp:1825-1862): the first element
//        taken, then replaced by any later x with x > cur (x < cur), the DD
//        comparison being lexicographic on (hi, lo) (fypp:2960-2998); 0 when
//        nothing is selected.
struct DimAcc {
  double sh, sl;
  bool have;
};

template <int KIND>
__device__ __forceinline__ DimAcc dim_init() {
  return DimAcc{KIND == R_PROD ? 1.0 : 0.0, 0.0, false};
}

// x > c (MAX) or x < c (MIN) as the reference's real64x2 operators
template <bool MAX>
__device__ __forceinline__ bool dd_better(double xh, double xl, double ch, double cl) {
  return MAX ? (ch < xh || (xh == ch && cl < xl)) : (xh < ch || (xh == ch && xl < cl));
}

template <int KIND>
__device__ __forceinline__ void dim_step(double2 x, DimAcc &s) {
  if constexpr (KIND == R_NORM2) {
    ddv xx = mk(x.x, x.y);
    ddv r = dd_add(mk(s.sh, s.sl), dd_mul(xx, xx));
    s.sh = r.hi;
    s.sl = r.lo;
  } else if constexpr (KIND == R_PROD) {
    ddv r = dd_mul(mk(s.sh, s.sl), mk(x.x, x.y));
    s.sh = r.hi;
    s.sl = r.lo;
  } else if constexpr (KIND == R_MAXV || KIND == R_MINV) {
    if (!s.have || dd_better<KIND == R_MAXV>(x.x, x.y, s.sh, s.sl)) {
      s.sh = x.x;
      s.sl = x.y;
      s.have = true;
    }
  } else {
    sum_step(x, s.sh, s.sl);
  }
}

template <int KIND>
__device__ __forceinline__ double2 dim_finish(DimAcc s) {
  if constexpr (KIND == R_NORM2) {
    ddv r = dd_sqrt(mk(s.sh, s.sl));
    return make_double2(r.hi, r.lo);
  } else if constexpr (KIND == R_SUM) {
    double t = s.sh + s.sl;
    double bp = t - s.sh;
    double ap = t - bp;
    double aerr = s.sh - ap;
    double berr = s.sl - bp;
    return make_double2(t, aerr + berr);
  } else {
    return make_double2(s.sh, s.sl);
  }
}

// inner > 1: one lane per output; lanes with consecutive i read consecutive
// addresses at every step j.
template <int KIND, bool MASKED>
__global__ __launch_bounds__(kThreads) void k_dim_strided(const double2 *__restrict__ a,
                                                          const int32_t *__restrict__ mask, int64_t inner,
                                                          int64_t len, int64_t outer, double2 *__restrict__ out) {
  const int64_t tot = inner * outer;
  for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < tot; f += (int64_t)gridDim.x * kThreads) {
    const int64_t o = f / inner, i = f - o * inner;
    const int64_t base = i + inner * len * o;
    DimAcc st = dim_init<KIND>();
    for (int64_t j = 0; j < len; ++j) {
      const int64_t e = base + inner * j;
      if (!MASKED || mask[e]) dim_step<KIND>(a[e], st);
    }
    out[f] = dim_finish<KIND>(st);
  }
}

// inner == 1 (reduction along the contiguous first dimension): a workgroup
// owns 256 consecutive outputs; each step stages a kDimJ-deep slab of their
// columns through LDS (8 lanes read one column's kDimJ contiguous elements),
// then every lane advances its own chain from LDS.
constexpr int kDimJ = 8;
template <int KIND, bool MASKED>
__global__ __launch_bounds__(kThreads) void k_dim_cols(const double2 *__restrict__ a,
                                                       const int32_t *__restrict__ mask, int64_t len,
                                                       int64_t outer, double2 *__restrict__ out) {
  __shared__ double th[kDimJ][kThreads + 1], tl[kDimJ][kThreads + 1];
  __shared__ int32_t tm[kDimJ][kThreads + 1];
  const int tid = threadIdx.x;
  const int64_t o0 = (int64_t)blockIdx.x * kThreads;
  DimAcc st = dim_init<KIND>();
  for (int64_t j0 = 0; j0 < len; j0 += kDimJ) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kDimJ; ++q) {
      const int e = q * kThreads + tid;
      const int oc = e / kDimJ, jj = e % kDimJ;
      const int64_t o = o0 + oc, j = j0 + jj;
      double2 v = make_double2(0.0, 0.0);
      int32_t mv = 1;
      if (o < outer && j < len) {
        v = a[o * len + j];
        if constexpr (MASKED) mv = mask[o * len + j];
      }
      th[jj][oc] = v.x;
      tl[jj][oc] = v.y;
      if constexpr (MASKED) tm[jj][oc] = mv;
    }
    __syncthreads();
    const int jn = (len - j0) < kDimJ ? (int)(len - j0) : kDimJ;
    for (int jj = 0; jj < jn; ++jj)
      if (!MASKED || tm[jj][tid]) dim_step<KIND>(make_double2(th[jj][tid], tl[jj][tid]), st);
  }
  if (o0 + tid < outer) out[o0 + tid] = dim_finish<KIND>(st);
}

// Rank-1 maxval / minval, bit-identical to the reference's sequential scan
// (dd_maxval_1d, fypp:1786-1810): res = first unmasked element, replaced by
// every later x with x > res. With the lexicographic comparison that scan
// returns
//   - the first element itself if its hi limb is NaN (nothing compares
//     greater than it);
//   - otherwise, with M the largest non-NaN hi limb and s0 the first element
//     whose hi == M: s0 if its lo limb is NaN, else the FIRST element with
//     hi == M and lo == L, L the largest non-NaN lo among hi == M.
// Those three indices are an associative, order-independent reduction of
// per-element tuples (ties go to the smaller index), so lanes and
// workgroups can combine in any tree and still return the scan's bits.
constexpr int64_t kNone = INT64_MAX;
struct Ext {
  int64_t first, f2, f3;  // first unmasked; first with hi == M; first with hi == M, lo == L
  double M, L;            // valid when f2 / f3 != kNone
};

__device__ __forceinline__ Ext ext_empty() { return Ext{kNone, kNone, kNone, 0.0, 0.0}; }

template <bool MAX>
__device__ __forceinline__ bool better(double x, double c) {
  return MAX ? x > c : x < c;
}

template <bool MAX>
__device__ __forceinline__ void ext_add(Ext &t, int64_t i, double2 x) {
  if (t.first == kNone) t.first = i;  // indices arrive in increasing order
  if (x.x != x.x) return;
  if (t.f2 == kNone || better<MAX>(x.x, t.M)) {
    t.M = x.x;
    t.f2 = i;
    t.f3 = kNone;
    if (x.y == x.y) {
      t.L = x.y;
      t.f3 = i;
    }
  } else if (x.x == t.M && x.y == x.y && (t.f3 == kNone || better<MAX>(x.y, t.L))) {
    t.L = x.y;
    t.f3 = i;
  }
}

template <bool MAX>
__device__ __forceinline__ Ext ext_combine(Ext a, Ext b) {
  Ext r;
  r.first = a.first < b.first ? a.first : b.first;
  if (b.f2 == kNone || (a.f2 != kNone && better<MAX>(a.M, b.M))) {
    r.M = a.M, r.f2 = a.f2, r.L = a.L, r.f3 = a.f3;
  } else if (a.f2 == kNone || better<MAX>(b.M, a.M)) {
    r.M = b.M, r.f2 = b.f2, r.L = b.L, r.f3 = b.f3;
  } else {  // equal M
    r.M = a.M;
    r.f2 = a.f2 < b.f2 ? a.f2 : b.f2;
    if (b.f3 == kNone || (a.f3 != kNone && better<MAX>(a.L, b.L))) {
      r.L = a.L, r.f3 = a.f3;
    } else if (a.f3 == kNone || better<MAX>(b.L, a.L)) {
      r.L = b.L, r.f3 = b.f3;
    } else {
      r.L = a.L;
      r.f3 = a.f3 < b.f3 ? a.f3 : b.f3;
    }
  }
  return r;
}

template <bool MAX>
__device__ Ext ext_block(Ext t) {
  __shared__ int64_t s_first[kThreads], s_f2[kThreads], s_f3[kThreads];
  __shared__ double s_M[kThreads], s_L[kThreads];
  const int tid = threadIdx.x;
  s_first[tid] = t.first, s_f2[tid] = t.f2, s_f3[tid] = t.f3, s_M[tid] = t.M, s_L[tid] = t.L;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const Ext a{s_first[tid], s_f2[tid], s_f3[tid], s_M[tid], s_L[tid]};
      const Ext b{s_first[tid + s], s_f2[tid + s], s_f3[tid + s], s_M[tid + s], s_L[tid + s]};
      const Ext r = ext_combine<MAX>(a, b);
      s_first[tid] = r.first, s_f2[tid] = r.f2, s_f3[tid] = r.f3, s_M[tid] = r.M, s_L[tid] = r.L;
    }
    __syncthreads();
  }
  return Ext{s_first[0], s_f2[0], s_f3[0], s_M[0], s_L[0]};
}

constexpr int kExtE = 16;  // elements per lane in the extrema pass

template <bool MAX, bool MASKED>
__global__ __launch_bounds__(kThreads) void k_ext_partial(const double2 *__restrict__ a,
                                                          const int32_t *__restrict__ mask, int64_t n,
                                                          Ext *__restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * kThreads * kExtE + threadIdx.x;
  v2d va[kExtE];
  int32_t vm[kExtE];
#pragma unroll
  for (int j = 0; j < kExtE; ++j) {
    const int64_t i = base + (int64_t)kThreads * j;
    const bool ok = i < n;
    va[j] = ok ? __builtin_nontemporal_load((const v2d *)a + i) : v2d{0.0, 0.0};
    vm[j] = ok ? (MASKED ? __builtin_nontemporal_load(mask + i) : 1) : 0;
  }
  Ext t = ext_empty();
#pragma unroll
  for (int j = 0; j < kExtE; ++j)
    if (vm[j]) ext_add<MAX>(t, base + (int64_t)kThreads * j, make_double2(va[j].x, va[j].y));
  const Ext r = ext_block<MAX>(t);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// One workgroup folds the g partials and writes the selected element.
template <bool MAX>
__global__ __launch_bounds__(kThreads) void k_ext_final(const double2 *__restrict__ a, const Ext *__restrict__ partial,
                                                        int64_t g, double2 *__restrict__ out) {
  Ext t = ext_empty();
  for (int64_t i = threadIdx.x; i < g; i += kThreads) t = ext_combine<MAX>(t, partial[i]);
  const Ext r = ext_block<MAX>(t);
  if (threadIdx.x == 0) {
    double2 v = make_double2(0.0, 0.0);  // empty or fully masked: 0 (fypp:1794)
    if (r.first != kNone) {
      v = a[r.first];
      if (v.x == v.x) {
        v = a[r.f2];
        if (v.y == v.y) v = a[r.f3];
      }
    }
    *out = v;
  }
}

}  // namespace

namespace mfhip {

namespace {
int64_t groups_for(int64_t n, int per_lane) {
  const int64_t span = (int64_t)kThreads * per_lane;
  const int64_t g = (n + span - 1) / span;
  return g < 1 ? 1 : g;
}
// Partials held in the workspace: every level of the fold chain except the
// last, which goes straight to the output.
int64_t chain_partials(int64_t g) {
  int64_t tot = 0;
  while (true) {
    tot += g;
    const int64_t g2 = (g + kFoldSpan - 1) / kFoldSpan;
    if (g2 <= 1) break;
    g = g2;
  }
  return tot;
}
template <int KIND, int NC>
hipError_t fold_chain(double2 *work, int64_t g, double2 *out, hipStream_t s) {
  double2 *src = work, *dst = work + g * NC;
  while (true) {
    const int64_t g2 = (g + kFoldSpan - 1) / kFoldSpan;
    const int last = g2 <= 1;
    hipLaunchKernelGGL((k_fold<KIND, NC>), dim3((unsigned)g2), dim3(kThreads), 0, s, src, g, last ? out : dst,
                       last);
    if (last) break;
    src = dst;
    dst += g2 * NC;
    g = g2;
  }
  return hipGetLastError();
}
template <int KIND>
hipError_t reduce_kind(const double2 *a, const double2 *b, int64_t n, int64_t ri, double2 *out, double2 *work,
                       hipStream_t s, const int32_t *mask) {
  const int64_t g = groups_for(n, elems_per_lane<KIND>());
  hipLaunchKernelGGL(k_reduce_partial<KIND>, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, work, mask);
  return fold_chain<KIND == R_SUM_MASK ? R_SUM : (KIND == R_PROD_MASK ? R_PROD : KIND), 1>(work, g, out, s);
}
template <bool MAX>
hipError_t extrema(const double2 *a, const int32_t *mask, int64_t n, double2 *out, double2 *work, hipStream_t s) {
  const int64_t g = groups_for(n, kExtE);
  Ext *part = reinterpret_cast<Ext *>(work);
  if (mask)
    hipLaunchKernelGGL((k_ext_partial<MAX, true>), dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, n, part);
  else
    hipLaunchKernelGGL((k_ext_partial<MAX, false>), dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, n, part);
  hipLaunchKernelGGL(k_ext_final<MAX>, dim3(1), dim3(kThreads), 0, s, a, part, g, out);
  return hipGetLastError();
}
}  // namespace

int64_t reduce_groups(int64_t n) { return groups_for(n, elems_per_lane<R_DOT>()); }

// Sized for the kind that needs the most: the dot's fold chain (8 elements
// per lane) or the extrema's tuples (one per 256 * kExtE elements).
size_t reduce_workspace_bytes(int64_t n) {
  if (n < 0) n = 0;
  const size_t dot = sizeof(double2) * (size_t)chain_partials(groups_for(n, elems_per_lane<R_DOT>()));
  const size_t ext = sizeof(Ext) * (size_t)groups_for(n, kExtE);
  return dot > ext ? dot : ext;
}

size_t cdot_workspace_bytes(int64_t n) {
  return 2 * sizeof(double2) * (size_t)chain_partials(groups_for(n < 0 ? 0 : n, kCdotE));
}

hipError_t reduce(int kind, const double2 *a, const double2 *b, int64_t n, int64_t ri,
                  double2 *out, double2 *work, hipStream_t s, const int32_t *mask) {
  if (n < 0) n = 0;
  switch (kind) {
    case RED_DOT: return reduce_kind<R_DOT>(a, b, n, ri, out, work, s, nullptr);
    case RED_SUM:
      return mask ? reduce_kind<R_SUM_MASK>(a, b, n, ri, out, work, s, mask)
                  : reduce_kind<R_SUM>(a, b, n, ri, out, work, s, nullptr);
    case RED_NORM2: return reduce_kind<R_NORM2>(a, b, n, ri, out, work, s, nullptr);
    case RED_PROD:
      return mask ? reduce_kind<R_PROD_MASK>(a, b, n, ri, out, work, s, mask)
                  : reduce_kind<R_PROD>(a, b, n, ri, out, work, s, nullptr);
    case RED_MAXVAL: return extrema<true>(a, mask, n, out, work, s);
    case RED_MINVAL: return extrema<false>(a, mask, n, out, work, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t cdot(const double4 *a, const double4 *b, int64_t n, int64_t ri, double4 *out, double4 *work,
                hipStream_t s) {
  if (n < 0) n = 0;
  const int64_t g = groups_for(n, kCdotE);
  double2 *w = reinterpret_cast<double2 *>(work);
  hipLaunchKernelGGL(k_cdot_partial, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, w);
  return fold_chain<R_DOT, 2>(w, g, reinterpret_cast<double2 *>(out), s);
}

namespace {
template <int KIND, bool MASKED>
void launch_dim(const double2 *a, const int32_t *mask, int64_t inner, int64_t len, int64_t outer, double2 *out,
                hipStream_t s) {
  if (inner == 1) {
    const int64_t g = (outer + kThreads - 1) / kThreads;
    hipLaunchKernelGGL((k_dim_cols<KIND, MASKED>), dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, len, outer, out);
  } else {
    const int64_t tot = inner * outer;
    int64_t g = (tot + kThreads - 1) / kThreads;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL((k_dim_strided<KIND, MASKED>), dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, inner, len,
                       outer, out);
  }
}
template <int KIND>
void launch_dim_m(const double2 *a, const int32_t *mask, int64_t inner, int64_t len, int64_t outer, double2 *out,
                  hipStream_t s) {
  if (mask) launch_dim<KIND, true>(a, mask, inner, len, outer, out, s);
  else launch_dim<KIND, false>(a, mask, inner, len, outer, out, s);
}
}  // namespace

hipError_t reduce_dim(int kind, const double2 *a, const int32_t *mask, int64_t inner, int64_t len,
                      int64_t outer, double2 *out, hipStream_t s) {
  if (inner <= 0 || outer <= 0) return hipSuccess;
  if (len < 0) len = 0;
  switch (kind) {
    case RED_SUM: launch_dim_m<R_SUM>(a, mask, inner, len, outer, out, s); break;
    case RED_NORM2:
      if (mask) return hipErrorInvalidValue;  // Fortran norm2 has no mask argument
      launch_dim<R_NORM2, false>(a, mask, inner, len, outer, out, s);
      break;
    case RED_PROD: launch_dim_m<R_PROD>(a, mask, inner, len, outer, out, s); break;
    case RED_MAXVAL: launch_dim_m<R_MAXV>(a, mask, inner, len, outer, out, s); break;
    case RED_MINVAL: launch_dim_m<R_MINV>(a, mask, inner, len, outer, out, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mfhip
