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

enum { R_DOT = 0, R_SUM = 1, R_NORM2 = 2 };

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
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_reduce_partial(const double2 *__restrict__ a,
                                                             const double2 *__restrict__ b,
                                                             int64_t n, int64_t ri,
                                                             double2 *__restrict__ partial) {
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
#pragma unroll
      for (int u = 0; u < U; ++u) {
        va[u] = a[i + u * stride];
        if constexpr (KIND == R_DOT) vb[u] = b[i + u * stride];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if constexpr (KIND == R_DOT) {
          dot_step(va[u], vb[u], sh, sl);
          if (ri > 0 && --until == 0) {
            until = ri;
            fast_two_sum(sh, sl, sh, sl);
          }
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
                  double2 *out, double2 *work, hipStream_t s) {
  if (n < 0) n = 0;
  const int64_t g = reduce_groups(n);
  switch (kind) {
    case RED_DOT:
      hipLaunchKernelGGL(k_reduce_partial<R_DOT>, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, work);
      hipLaunchKernelGGL(k_reduce_final<R_DOT>, dim3(1), dim3(kThreads), 0, s, work, g, out);
      break;
    case RED_SUM:
      hipLaunchKernelGGL(k_reduce_partial<R_SUM>, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, work);
      hipLaunchKernelGGL(k_reduce_final<R_SUM>, dim3(1), dim3(kThreads), 0, s, work, g, out);
      break;
    case RED_NORM2:
      hipLaunchKernelGGL(k_reduce_partial<R_NORM2>, dim3((unsigned)g), dim3(kThreads), 0, s, a, b, n, ri, work);
      hipLaunchKernelGGL(k_reduce_final<R_NORM2>, dim3(1), dim3(kThreads), 0, s, work, g, out);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mfhip
