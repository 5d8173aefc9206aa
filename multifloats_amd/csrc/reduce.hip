// DD reductions: dot_product, sum, norm2.
//
// The reference computes these as single sequential Fortran loops
// (fsrc/multifloats.fypp: dd_dot_product 2254-2308, dd_sum_1d 1980-2020,
// dd_norm2_1d 2338-2352), a latency-bound chain through s_hi. Here every lane
// runs the reference's own per-element update on a strided slice of the
// input (so each partial carries the same compensation the reference's
// accumulator does), then the partials are combined by a fixed binary tree of
// accurate DD additions -- first across the workgroup through LDS, then across
// workgroups in a second single-workgroup pass. The tree shape depends only on
// n, so results are reproducible run to run. The reassociation means results
// are not the sequential loop's bits; they are held to the c*n*2^-104*sum|term|
// bound against an exact accumulator instead.
#include "dd_core.h"
#include "mf_internal.h"

namespace {

constexpr int kThreads = 256;

enum { R_DOT = 0, R_SUM = 1, R_NORM2 = 2, R_SUM_MASK = 3 };

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

// Block tree: 256 normalised pairs -> one, in LDS, fixed shape.
__device__ ddv block_tree(ddv v) {
  __shared__ double th[kThreads], tl[kThreads];
  const int tid = threadIdx.x;
  th[tid] = v.hi;
  tl[tid] = v.lo;
  __syncthreads();
#pragma unroll
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      ddv r = dd_add(mk(th[tid], tl[tid]), mk(th[tid + s], tl[tid + s]));
      th[tid] = r.hi;
      tl[tid] = r.lo;
    }
    __syncthreads();
  }
  return mk(th[0], tl[0]);
}

// Pass 1: lane-local reference accumulation over a strided slice, then the
// workgroup tree; one normalised DD partial per workgroup.
// R_SUM_MASK: the sum update runs only where mask[i] != 0 (Fortran
// sum(array, mask), fypp:1993-1995: masked elements are skipped, not added).
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_reduce_partial(const double2 *__restrict__ a,
                                                             const double2 *__restrict__ b,
                                                             int64_t n, int64_t ri,
                                                             double2 *__restrict__ partial,
                                                             const int32_t *__restrict__ mask) {
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  ddv acc;
  if constexpr (KIND == R_NORM2) {
    acc = mk(0.0, 0.0);
    for (; i < n; i += stride) {
      double2 x = a[i];
      ddv xx = mk(x.x, x.y);
      acc = dd_add(acc, dd_mul(xx, xx));
    }
  } else {
    double sh = 0.0, sl = 0.0;
    int64_t until = ri;
    constexpr int U = 4;
    // U independent loads in flight per lane; the update order per lane is
    // still strictly increasing in i.
    for (; i + (U - 1) * stride < n; i += U * stride) {
      double2 va[U], vb[U];
      int32_t vm[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        va[u] = a[i + u * stride];
        if constexpr (KIND == R_DOT) vb[u] = b[i + u * stride];
        if constexpr (KIND == R_SUM_MASK) vm[u] = mask[i + u * stride];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if constexpr (KIND == R_DOT) {
          dot_step(va[u], vb[u], sh, sl);
          if (ri > 0 && --until == 0) {
            until = ri;
            fast_two_sum(sh, sl, sh, sl);
          }
        } else if constexpr (KIND == R_SUM_MASK) {
          if (vm[u]) sum_step(va[u], sh, sl);
        } else {
          sum_step(va[u], sh, sl);
        }
      }
    }
    for (; i < n; i += stride) {
      if constexpr (KIND == R_DOT) {
        dot_step(a[i], b[i], sh, sl);
        if (ri > 0 && --until == 0) {
          until = ri;
          fast_two_sum(sh, sl, sh, sl);
        }
      } else if constexpr (KIND == R_SUM_MASK) {
        if (mask[i]) sum_step(a[i], sh, sl);
      } else {
        sum_step(a[i], sh, sl);
      }
    }
    renorm(sh, sl);
    acc = mk(sh, sl);
  }
  ddv r = block_tree(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(r.hi, r.lo);
}

// Pass 2: one workgroup folds the partials (strided, then the same tree).
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_reduce_final(const double2 *__restrict__ partial,
                                                           int64_t g, double2 *__restrict__ out) {
  ddv acc = mk(0.0, 0.0);
  for (int64_t i = threadIdx.x; i < g; i += kThreads) acc = dd_add(acc, mk(partial[i].x, partial[i].y));
  ddv r = block_tree(acc);
  if (threadIdx.x == 0) {
    if constexpr (KIND == R_NORM2) r = dd_sqrt(r);
    out[0] = make_double2(r.hi, r.lo);
  }
}


// Complex dot_product(a, b) = sum(conj(a) * b) (fypp:2321-2331): the reference
// runs four real dot chains, re = dot(ar, br) + dot(ai, bi) and
// im = dot(ar, bi) - dot(ai, br), and combines them with DD + and -. Here
// each lane keeps the four chains for its slice in registers (one pass over
// both inputs), forms its (re, im) with the same DD + / -, and the lanes'
// complex partials meet in the fixed tree.
__device__ __forceinline__ void cdot_lane(const double4 *__restrict__ a, const double4 *__restrict__ b,
                                          int64_t n, int64_t ri, ddv &re, ddv &im) {
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  double h[4] = {0.0, 0.0, 0.0, 0.0}, l[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t until = ri;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const double4 x = a[i], y = b[i];
    const double2 ar = make_double2(x.x, x.y), ai = make_double2(x.z, x.w);
    const double2 br = make_double2(y.x, y.y), bi = make_double2(y.z, y.w);
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
  re = dd_add(mk(h[0], l[0]), mk(h[1], l[1]));
  im = dd_sub(mk(h[2], l[2]), mk(h[3], l[3]));
}

__global__ __launch_bounds__(kThreads) void k_cdot_partial(const double4 *__restrict__ a,
                                                           const double4 *__restrict__ b, int64_t n,
                                                           int64_t ri, double4 *__restrict__ partial) {
  ddv re, im;
  cdot_lane(a, b, n, ri, re, im);
  ddv r = block_tree(re);
  __syncthreads();
  ddv q = block_tree(im);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double4(r.hi, r.lo, q.hi, q.lo);
}

__global__ __launch_bounds__(kThreads) void k_cdot_final(const double4 *__restrict__ partial, int64_t g,
                                                         double4 *__restrict__ out) {
  ddv re = mk(0.0, 0.0), im = mk(0.0, 0.0);
  for (int64_t i = threadIdx.x; i < g; i += kThreads) {
    const double4 p = partial[i];
    re = dd_add(re, mk(p.x, p.y));
    im = dd_add(im, mk(p.z, p.w));
  }
  ddv r = block_tree(re);
  __syncthreads();
  ddv q = block_tree(im);
  if (threadIdx.x == 0) out[0] = make_double4(r.hi, r.lo, q.hi, q.lo);
}

// Reductions along one dimension, bit-identical to the reference: the array
// is viewed as (inner, len, outer) in column-major order and every output
// out[i + inner*o] is ONE sequential chain over j = 0..len-1 of the
// reference's own update, exactly as the Fortran loop runs it:
//   sum  (dd_sum_Nd_dim, fypp:2052-2075): the Neumaier update, masked
//        elements skipped, then a full TwoSum written in the reference's
//        order (t, bp, ap, aerr, berr; lo = aerr + berr);
//   norm2 (dd_norm2_Nd_dim, fypp:2356-2377): cur = cur + x*x with the DD
//        operators, then DD sqrt.
template <int KIND>
__device__ __forceinline__ void dim_step(double2 x, int32_t mk_, double &sh, double &sl) {
  if constexpr (KIND == R_NORM2) {
    ddv xx = mk(x.x, x.y);
    ddv r = dd_add(mk(sh, sl), dd_mul(xx, xx));
    sh = r.hi;
    sl = r.lo;
  } else {
    if (KIND == R_SUM_MASK && !mk_) return;
    sum_step(x, sh, sl);
  }
}

template <int KIND>
__device__ __forceinline__ double2 dim_finish(double sh, double sl) {
  if constexpr (KIND == R_NORM2) {
    ddv r = dd_sqrt(mk(sh, sl));
    return make_double2(r.hi, r.lo);
  } else {
    double t = sh + sl;
    double bp = t - sh;
    double ap = t - bp;
    double aerr = sh - ap;
    double berr = sl - bp;
    return make_double2(t, aerr + berr);
  }
}

// inner > 1: one lane per output; lanes with consecutive i read consecutive
// addresses at every step j.
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_dim_strided(const double2 *__restrict__ a,
                                                          const int32_t *__restrict__ mask, int64_t inner,
                                                          int64_t len, int64_t outer, double2 *__restrict__ out) {
  const int64_t tot = inner * outer;
  for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < tot; f += (int64_t)gridDim.x * kThreads) {
    const int64_t o = f / inner, i = f - o * inner;
    const int64_t base = i + inner * len * o;
    double sh = 0.0, sl = 0.0;
    for (int64_t j = 0; j < len; ++j) {
      const int64_t e = base + inner * j;
      dim_step<KIND>(a[e], KIND == R_SUM_MASK ? mask[e] : 1, sh, sl);
    }
    out[f] = dim_finish<KIND>(sh, sl);
  }
}

// inner == 1 (reduction along the contiguous first dimension): a workgroup
// owns 256 consecutive outputs; each step stages a kDimJ-deep slab of their
// columns through LDS (8 lanes read one column's kDimJ contiguous elements),
// then every lane advances its own chain from LDS.
constexpr int kDimJ = 8;
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_dim_cols(const double2 *__restrict__ a,
                                                       const int32_t *__restrict__ mask, int64_t len,
                                                       int64_t outer, double2 *__restrict__ out) {
  __shared__ double th[kDimJ][kThreads + 1], tl[kDimJ][kThreads + 1];
  __shared__ int32_t tm[kDimJ][kThreads + 1];
  const int tid = threadIdx.x;
  const int64_t o0 = (int64_t)blockIdx.x * kThreads;
  double sh = 0.0, sl = 0.0;
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
        if constexpr (KIND == R_SUM_MASK) mv = mask[o * len + j];
      }
      th[jj][oc] = v.x;
      tl[jj][oc] = v.y;
      if constexpr (KIND == R_SUM_MASK) tm[jj][oc] = mv;
    }
    __syncthreads();
    const int jn = (len - j0) < kDimJ ? (int)(len - j0) : kDimJ;
    for (int jj = 0; jj < jn; ++jj)
      dim_step<KIND>(make_double2(th[jj][tid], tl[jj][tid]), KIND == R_SUM_MASK ? tm[jj][tid] : 1, sh, sl);
  }
  if (o0 + tid < outer) out[o0 + tid] = dim_finish<KIND>(sh, sl);
}

}  // namespace

namespace mfhip {

// Number of pass-1 workgroups for n elements: a function of n only (the tree
// shape, hence the result bits, must not depend on the device), capped at 2048
// (8 per CU on a 256-CU part), at least 2048 elements per workgroup.
int64_t reduce_groups(int64_t n) {
  int64_t g = (n + 2047) / 2048;
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  return g;
}

size_t reduce_workspace_bytes(int64_t n) { return sizeof(double2) * (size_t)reduce_groups(n); }

hipError_t reduce(int kind, const double2 *a, const double2 *b, int64_t n, int64_t ri,
                  double2 *out, double2 *work, hipStream_t s, const int32_t *mask) {
  if (n < 0) n = 0;
  const int64_t g = reduce_groups(n);
  switch (kind) {
    case RED_DOT:
      hipLaunchKernelGGL(k_reduce_partial<R_DOT>, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, work,
                         nullptr);
      hipLaunchKernelGGL(k_reduce_final<R_DOT>, dim3(1), dim3(kThreads), 0, s, work, g, out);
      break;
    case RED_SUM:
      if (mask) {
        hipLaunchKernelGGL(k_reduce_partial<R_SUM_MASK>, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri,
                           work, mask);
      } else {
        hipLaunchKernelGGL(k_reduce_partial<R_SUM>, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, work,
                           nullptr);
      }
      hipLaunchKernelGGL(k_reduce_final<R_SUM>, dim3(1), dim3(kThreads), 0, s, work, g, out);
      break;
    case RED_NORM2:
      hipLaunchKernelGGL(k_reduce_partial<R_NORM2>, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, work,
                         nullptr);
      hipLaunchKernelGGL(k_reduce_final<R_NORM2>, dim3(1), dim3(kThreads), 0, s, work, g, out);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t cdot(const double4 *a, const double4 *b, int64_t n, int64_t ri, double4 *out, double4 *work,
                hipStream_t s) {
  if (n < 0) n = 0;
  const int64_t g = reduce_groups(n);
  hipLaunchKernelGGL(k_cdot_partial, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, work);
  hipLaunchKernelGGL(k_cdot_final, dim3(1), dim3(kThreads), 0, s, work, g, out);
  return hipGetLastError();
}

hipError_t reduce_dim(int kind, const double2 *a, const int32_t *mask, int64_t inner, int64_t len,
                      int64_t outer, double2 *out, hipStream_t s) {
  if (inner <= 0 || outer <= 0) return hipSuccess;
  if (len < 0) len = 0;
  if (kind != RED_SUM && kind != RED_NORM2) return hipErrorInvalidValue;
  if (kind == RED_NORM2 && mask) return hipErrorInvalidValue;
  if (inner == 1) {
    const int64_t g = (outer + kThreads - 1) / kThreads;
    if (kind == RED_NORM2)
      hipLaunchKernelGGL(k_dim_cols<R_NORM2>, dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, len, outer, out);
    else if (mask)
      hipLaunchKernelGGL(k_dim_cols<R_SUM_MASK>, dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, len, outer, out);
    else
      hipLaunchKernelGGL(k_dim_cols<R_SUM>, dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, len, outer, out);
  } else {
    const int64_t tot = inner * outer;
    int64_t g = (tot + kThreads - 1) / kThreads;
    if (g > 65536) g = 65536;
    if (kind == RED_NORM2)
      hipLaunchKernelGGL(k_dim_strided<R_NORM2>, dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, inner, len,
                         outer, out);
    else if (mask)
      hipLaunchKernelGGL(k_dim_strided<R_SUM_MASK>, dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, inner,
                         len, outer, out);
    else
      hipLaunchKernelGGL(k_dim_strided<R_SUM>, dim3((unsigned)g), dim3(kThreads), 0, s, a, mask, inner, len,
                         outer, out);
  }
  return hipGetLastError();
}

}  // namespace mfhip
