// Batching form: many independent, uneven DD GEMMs and dot products in one
// persistent grouped launch (optionally captured in a HIP graph).
//
// Semantics per problem are the reference's: each GEMM is matmuldd_mm
// (src/float64x2/matmul.inc, reference-order chains, so results are
// bit-identical to a standalone matmuldd_mm), each dot is the Fortran
// dot_product (fsrc/multifloats.fypp:2254-2308) evaluated like dotdd: the
// reference's per-element update per lane, then a fixed tree of accurate DD
// additions.
//
// Plan: the host cuts every GEMM into output tiles of its tile variant (one
// of eight shapes, chosen by the caller's tuned table or by least padded
// work) and every dot into chunks of kDotChunk (2^20) elements, then groups
// the items by kind (tile shape, or dot). Within a kind, the items are dealt
// to the launch's workgroups by longest processing time first: largest
// estimated cost first, each to the workgroup with the least cost so far.
// One persistent launch per kind; workgroup g takes its own list, in the
// order dealt (a static schedule: no queue, every loop bounded by its list),
// so each grid drains. A dot of one chunk is a single item that
// writes its result directly. A longer dot's chunks leave DD partials; the
// chunk that completes the dot last (per-dot atomic counter) folds the
// partials in index order, so results do not depend on scheduling. That
// hand-off needs device-scope release/acquire fences, which on this part
// write back / invalidate an XCD's whole L2: with 16 Ki-element chunks the
// fences of ~64 Ki dot items held 50 000 dots to ~1 TB/s, hence the large
// chunk. The counters are cleared by a fill kernel at the start of every run
// (graph-capture safe: no memset nodes).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <queue>
#include <utility>
#include <vector>

#include "../../include/multifloats_hip.h"
#include "dd_core.h"
#include "gemm_tile.h"
#include "mf_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kBK = 16;
constexpr int64_t kDotChunk = 1 << 20;  // elements per work item of one dot

// GEMM tile variants (kind of a GEMM work item): (TM, TN) outputs per lane,
// (16 TM) x (16 TN) tiles. 0: 32x32, 1: 64x64, 2: 128x16, 3: 16x128,
// 4: 64x32, 5: 32x64, 6: 16x16, 7: 48x48. The skinny shapes keep the padding
// of tall or wide problems small, 16x16 that of tiny ones, 48x48 that of
// edges between the powers of two.
constexpr int kNumTileVariants = 8;
constexpr int kTileTM[kNumTileVariants] = {2, 4, 8, 1, 4, 2, 1, 3};
constexpr int kTileTN[kNumTileVariants] = {2, 4, 1, 8, 2, 4, 1, 3};
// Relative cost per padded MAC of each variant (fewer chains per lane or
// more LDS reads per MAC cost more); used by the built-in choice.
constexpr double kTileCost[kNumTileVariants] = {1.3, 1.05, 1.25, 1.3, 1.0, 1.06, 2.0, 1.03};
enum { ITEM_DOT = kNumTileVariants };

struct Item {
  int kind;
  int prob;  // GEMM or dot index
  int a, b;  // GEMM: tile row, tile col; dot: chunk index, unused
};

struct GemmD {
  const double2 *A, *B;
  double2 *C;
  int64_t m, k, n, ri;
};

struct DotD {
  const double2 *a, *b;
  double2 *out;
  int64_t n, ri;
  int64_t first;  // index of the dot's first partial
  int nchunks;
  int id;  // index of the dot's completion counter
};

// One output tile of a reference-order GEMM: the direct-copy tile body
// (gemm_tile.h: operands copied straight into two LDS stages, the same
// chains as the standalone GEMM), with the static renormalisation schedule
// when the problem's interval is 8 (the Fortran default).
template <int TM, int TN>
__device__ __forceinline__ void gemm_tile(const GemmD &g, int tr, int tc, char *lds) {
  static_assert(mfhip_tile::kDirectBK == kBK, "K block");
  const int64_t m0 = (int64_t)tr * (16 * TM), n0 = (int64_t)tc * (16 * TN);
  if (g.ri == 8)
    mfhip_tile::grouped_tile_direct<TM, TN, 8>(g.A, g.m, g.B, g.k, g.C, g.m, g.m, g.k, g.n, g.ri, m0, n0, lds);
  else
    mfhip_tile::grouped_tile_direct<TM, TN, 0>(g.A, g.m, g.B, g.k, g.C, g.m, g.m, g.k, g.n, g.ri, m0, n0, lds);
}

// Block tree over 256 normalised pairs, fixed shape (the same as reduce.hip).
__device__ ddv tree256(ddv v, double *th, double *tl) {
  const int tid = threadIdx.x;
  __syncthreads();
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

template <int KE>
__device__ void dot_chunk(const DotD &d, int c, double2 *partials, int *dot_done, double *lds) {
  const int tid = threadIdx.x;
  const int64_t p0 = (int64_t)c * kDotChunk;
  const int64_t p1 = p0 + kDotChunk < d.n ? p0 + kDotChunk : d.n;
  double sh = 0.0, sl = 0.0;
  int64_t until = d.ri;
  // Lane tid takes elements p0 + tid + 256 j in order. They are loaded KE
  // at a time with the next KE already in flight (nontemporal 16-byte loads,
  // two register buffers), so the latency-bound accumulation chain does not
  // starve the memory system. KE changes the load batching, not the order of
  // the chain: the result's bits do not depend on it.
  constexpr int kE = KE;
  constexpr int64_t kRound = (int64_t)kThreads * kE;
  v2d xa[kE], xb[kE], ya[kE], yb[kE];  // two register buffers, indexed statically
  auto fetch = [&](v2d (&va)[kE], v2d (&vb)[kE], int64_t i0) {
#pragma unroll
    for (int j = 0; j < kE; ++j) {
      const int64_t i = i0 + (int64_t)kThreads * j;
      const bool ok = i < p1;
      va[j] = ok ? __builtin_nontemporal_load((const v2d *)d.a + i) : v2d{0.0, 0.0};
      vb[j] = ok ? __builtin_nontemporal_load((const v2d *)d.b + i) : v2d{0.0, 0.0};
    }
  };
  auto consume = [&](const v2d (&va)[kE], const v2d (&vb)[kE], int64_t i0) {
#pragma unroll
    for (int j = 0; j < kE; ++j) {
      if (i0 + (int64_t)kThreads * j >= p1) break;
      const v2d a = va[j], b = vb[j];
      // Fortran dot update (fypp:2286-2302)
      const double p = a.x * b.x;
      const double e = __builtin_fma(a.x, b.x, -p);
      const double cross = a.x * b.y + a.y * b.x;
      const double t = sh + p, bp = t - sh, ap = t - bp;
      sl = sl + (((e + cross) + (sh - ap)) + (p - bp));
      sh = t;
      if (d.ri > 0 && --until == 0) {
        until = d.ri;
        fast_two_sum(sh, sl, sh, sl);
      }
    }
  };
  fetch(xa, xb, p0 + tid);
  for (int64_t i0 = p0 + tid; i0 < p1; i0 += 2 * kRound) {
    if (i0 + kRound < p1) fetch(ya, yb, i0 + kRound);
    consume(xa, xb, i0);
    if (i0 + kRound >= p1) break;
    if (i0 + 2 * kRound < p1) fetch(xa, xb, i0 + 2 * kRound);
    consume(ya, yb, i0 + kRound);
  }
  renorm(sh, sl);
  double *th = lds, *tl = lds + kThreads;
  const ddv r = tree256(mk(sh, sl), th, tl);
  if (d.nchunks == 1) {  // the whole dot: no partials, no cross-workgroup hand-off
    if (tid == 0) *d.out = make_double2(r.hi, r.lo);
    return;
  }
  __shared__ int last;
  if (tid == 0) {
    partials[d.first + c] = make_double2(r.hi, r.lo);
    __threadfence();  // release: the partial is visible before the count
    last = atomicAdd(&dot_done[d.id], 1) == d.nchunks - 1;
  }
  __syncthreads();
  // broadcast through LDS, then made wave-uniform so the branch around the
  // barriers of tree256 is a scalar branch
  if (__builtin_amdgcn_readfirstlane(last)) {
    __threadfence();  // acquire: see every other chunk's partial
    ddv acc = mk(0.0, 0.0);
    for (int q = tid; q < d.nchunks; q += kThreads) {
      const double2 v = partials[d.first + q];
      acc = dd_add(acc, mk(v.x, v.y));
    }
    const ddv s = tree256(acc, th, tl);
    if (tid == 0) *d.out = make_double2(s.hi, s.lo);
  }
}

// Static longest-processing-time schedule: workgroup g takes items
// [off[g], off[g + 1]) of its kind's slice. The item index derives from
// blockIdx only, so every branch and loop around the barriers inside the
// item bodies is uniform. Each kind has a kernel of its
// own: one kernel holding every tile body needs more registers than any body
// alone and spilled (the 64 x 64 body alone already does at 256 VGPRs); split,
// each body gets its own register budget and occupancy.
template <int TM, int TN, int WPE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WPE))) void k_grouped_gemm(
    const Item *__restrict__ items, const int *__restrict__ off, const GemmD *__restrict__ gemms) {
  extern __shared__ __attribute__((aligned(16))) char lds[];  // DirectShape<TM, TN>::kLds bytes
  const int end = off[blockIdx.x + 1];
  for (int it = off[blockIdx.x]; it < end; ++it) {
    const Item w = items[it];
    gemm_tile<TM, TN>(gemms[w.prob], w.a, w.b, lds);
  }
}

// The dots. Alone (a plan without GEMMs): eight elements per register buffer,
// two waves per SIMD. Beside the GEMM kinds (the
// dots' launch on a second stream, running while the GEMM kinds run): two
// elements per buffer and at most 64 VGPRs (the eight-per-SIMD budget), so
// that two dot workgroups per CU fit in the VGPRs a GEMM kind leaves free
// (64 x 32 tiles: three workgroups of 128 VGPRs per lane, 128 left of 512).
template <int KE, int WPE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WPE))) void k_grouped_dot(
    const Item *__restrict__ items, const int *__restrict__ off, const DotD *__restrict__ dots,
    double2 *__restrict__ partials, int *__restrict__ dot_done) {
  __shared__ __attribute__((aligned(16))) double lds[2 * kThreads];
  const int end = off[blockIdx.x + 1];
  for (int it = off[blockIdx.x]; it < end; ++it) {
    const Item w = items[it];
    dot_chunk<KE>(dots[w.prob], w.a, partials, dot_done, lds);
  }
}

__global__ void k_zero_ints(int *__restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

// ---------------------------------------------------------------- host side
namespace {

struct Plan {
  int device = -1;
  Item *items = nullptr;  // grouped by kind, each kind's slice sorted by cost
  GemmD *gemms = nullptr;
  DotD *dots = nullptr;
  double2 *partials = nullptr;
  int *counters = nullptr;  // per-dot completion counts
  int nitems = 0, ndot = 0, cus = 1;
  int first[ITEM_DOT + 2] = {};  // items of kind q: [first[q], first[q + 1])
  int *offs = nullptr;           // per kind: its workgroups' item ranges, relative to first[q]
  int ofirst[ITEM_DOT + 1] = {};  // kind q's ranges: offs[ofirst[q] .. ofirst[q] + groups[q]]
  int groups[ITEM_DOT + 1] = {};  // kind q's launch size
  int64_t ncounters = 0;
  hipStream_t gstream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // dots beside the GEMM kinds: their stream and the fork / join events
  hipStream_t dstream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int order[ITEM_DOT] = {};  // GEMM kinds in launch order: largest estimated cost first
};

// Workgroups per CU of each kind's launch (its waves per SIMD: 4-wave
// workgroups), in kind order; the dots run two, alone or beside a GEMM kind.
constexpr int kKindWPE[ITEM_DOT + 1] = {4, 2, 2, 2, 3, 3, 5, 3, 2};

// One persistent launch per kind with items: grid = the kind's resident
// workgroups over all CUs (its occupancy), or fewer when it has fewer items.
template <int TM, int TN, int WPE, int Q>
void launch_kind(const Plan *p, hipStream_t s) {
  static_assert(kKindWPE[Q] == WPE, "the plan deals kind Q's items to kKindWPE[Q] workgroups per CU");
  constexpr int q = Q;
  if (p->groups[q] <= 0) return;
  constexpr int lds = mfhip_tile::DirectShape<TM, TN>::kLds;
  static_assert(lds * WPE <= 160 * 1024, "the kind's workgroups per CU fit in LDS");
  hipLaunchKernelGGL((k_grouped_gemm<TM, TN, WPE>), dim3((unsigned)p->groups[q]), dim3(kThreads), lds, s,
                     p->items + p->first[q], p->offs + p->ofirst[q], p->gemms);
}

// Dynamic LDS above the 64 KiB default, allowed once per plan on its device
// (at creation: nothing but launches happens while a run is captured).
template <int TM, int TN, int WPE>
bool allow_kind_lds() {
  constexpr int lds = mfhip_tile::DirectShape<TM, TN>::kLds;
  return lds <= 65536 || hipFuncSetAttribute((const void *)k_grouped_gemm<TM, TN, WPE>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, lds) == hipSuccess;
}
bool allow_lds() {
  return allow_kind_lds<2, 2, 4>() && allow_kind_lds<4, 4, 2>() && allow_kind_lds<8, 1, 2>() &&
         allow_kind_lds<1, 8, 2>() && allow_kind_lds<4, 2, 3>() && allow_kind_lds<2, 4, 3>() &&
         allow_kind_lds<1, 1, 5>() && allow_kind_lds<3, 3, 3>();
}

void launch_gemm_kind(const Plan *p, int q, hipStream_t s) {
  switch (q) {
    case 0: launch_kind<2, 2, 4, 0>(p, s); break;
    case 1: launch_kind<4, 4, 2, 1>(p, s); break;
    case 2: launch_kind<8, 1, 2, 2>(p, s); break;
    case 3: launch_kind<1, 8, 2, 3>(p, s); break;
    case 4: launch_kind<4, 2, 3, 4>(p, s); break;
    case 5: launch_kind<2, 4, 3, 5>(p, s); break;
    case 6: launch_kind<1, 1, 5, 6>(p, s); break;
    default: launch_kind<3, 3, 3, 7>(p, s); break;
  }
}

void launch_dots(const Plan *p, bool beside, hipStream_t s) {
  const int q = ITEM_DOT;
  if (beside)
    hipLaunchKernelGGL((k_grouped_dot<2, 8>), dim3((unsigned)p->groups[q]), dim3(kThreads), 0, s, p->items + p->first[q],
                       p->offs + p->ofirst[q], p->dots, p->partials, p->counters);
  else
    hipLaunchKernelGGL((k_grouped_dot<8, 2>), dim3((unsigned)p->groups[q]), dim3(kThreads), 0, s, p->items + p->first[q],
                       p->offs + p->ofirst[q], p->dots, p->partials, p->counters);
}

// The counters are cleared, then the GEMM kinds run one after another on s,
// largest estimated cost first. With GEMMs in the plan, the dots run on the
// plan's second stream from the start of the first GEMM kind (forked after the
// clear, joined back into s at the end): the dots are HBM-bound, the GEMM
// kinds VALU-bound, and the dot workgroups fit in the registers and LDS the
// first kind leaves free. Without GEMMs the dots run alone on s.
hipError_t run_plan(Plan *p, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)((p->ncounters + 255) / 256)), dim3(256), 0, s,
                     p->counters, p->ncounters);
  const bool dots = p->groups[ITEM_DOT] > 0, beside = dots && p->dstream;
  hipError_t e = hipSuccess;
  if (beside) {
    e = hipEventRecord(p->ev_fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(p->dstream, p->ev_fork, 0);
    if (e != hipSuccess) return e;
  }
  for (int i = 0; i < ITEM_DOT; ++i) {
    launch_gemm_kind(p, p->order[i], s);
    if (i == 0 && beside) {
      launch_dots(p, true, p->dstream);
      e = hipEventRecord(p->ev_join, p->dstream);
      if (e != hipSuccess) return e;
    }
  }
  if (beside)
    e = hipStreamWaitEvent(s, p->ev_join, 0);
  else if (dots)
    launch_dots(p, false, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

void free_plan(Plan *p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->ev_in) (void)hipEventDestroy(p->ev_in);
  if (p->ev_out) (void)hipEventDestroy(p->ev_out);
  if (p->gstream) (void)hipStreamDestroy(p->gstream);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->dstream) (void)hipStreamDestroy(p->dstream);
  (void)hipFree(p->items);
  (void)hipFree(p->offs);
  (void)hipFree(p->gemms);
  (void)hipFree(p->dots);
  (void)hipFree(p->partials);
  (void)hipFree(p->counters);
  delete p;
}

}  // namespace

namespace multifloats {
extern "C" {

void *mfhip_grouped_create(const mfhip_gemm_problem *gemms, int ngemm, const int *tile_variant,
                           const mfhip_dot_problem *dots, int ndot) {
  if (ngemm < 0 || ndot < 0 || (ngemm > 0 && !gemms) || (ndot > 0 && !dots)) return nullptr;
  std::vector<GemmD> hg((size_t)ngemm);
  std::vector<DotD> hd((size_t)ndot);
  std::vector<Item> items;
  std::vector<double> cost;
  for (int i = 0; i < ngemm; ++i) {
    const mfhip_gemm_problem &g = gemms[i];
    hg[i] = GemmD{(const double2 *)g.a, (const double2 *)g.b, (double2 *)g.c, g.m, g.k, g.n,
                  g.renorm_interval};
    if (g.m <= 0 || g.n <= 0) continue;
    int v = tile_variant ? tile_variant[i] : -1;
    if (v >= kNumTileVariants) v = -1;
    if (v < 0) {  // least padded work, weighted by the variant's relative cost
      double best = 0.0;
      for (int q = 0; q < kNumTileVariants; ++q) {
        const int64_t bm = 16 * kTileTM[q], bn = 16 * kTileTN[q];
        const double c = (double)((g.m + bm - 1) / bm * bm) * (double)((g.n + bn - 1) / bn * bn) * kTileCost[q];
        if (v < 0 || c < best) {
          v = q;
          best = c;
        }
      }
    }
    const int64_t bm = 16 * kTileTM[v], bn = 16 * kTileTN[v];
    const int64_t tm = (g.m + bm - 1) / bm, tn = (g.n + bn - 1) / bn;
    for (int64_t a = 0; a < tm; ++a)
      for (int64_t b = 0; b < tn; ++b) {
        items.push_back(Item{v, i, (int)a, (int)b});
        // MAC count of the padded tile (k = 0 still writes zeros: count 1)
        cost.push_back((double)(bm * bn) * (double)(g.k > 0 ? g.k : 1) * kTileCost[v]);
      }
  }
  int64_t nparts = 0;
  for (int i = 0; i < ndot; ++i) {
    const mfhip_dot_problem &d = dots[i];
    const int64_t n = d.n > 0 ? d.n : 0;
    const int nch = (int)((n + kDotChunk - 1) / kDotChunk);
    const int chunks = nch > 0 ? nch : 1;
    hd[i] = DotD{(const double2 *)d.a, (const double2 *)d.b, (double2 *)d.out, n,
                 d.renorm_interval, nparts, chunks, i};
    for (int c = 0; c < chunks; ++c) {
      items.push_back(Item{ITEM_DOT, i, c, 0});
      const int64_t len = std::min<int64_t>(kDotChunk, n - (int64_t)c * kDotChunk);
      // one element streams 32 bytes: about 14 DD-MACs of VALU time
      cost.push_back(14.0 * (double)(len > 0 ? len : 1));
    }
    nparts += chunks;
  }
  // grouped by kind; longest processing time first within a kind (stable
  // for determinism of the plan)
  std::vector<size_t> order(items.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    if (items[x].kind != items[y].kind) return items[x].kind < items[y].kind;
    return cost[x] > cost[y];
  });
  std::vector<Item> sorted(items.size());
  std::vector<double> scost(items.size());
  for (size_t i = 0; i < order.size(); ++i) {
    sorted[i] = items[order[i]];
    scost[i] = cost[order[i]];
  }

  Plan *p = new Plan();
  for (const Item &it : sorted) ++p->first[it.kind + 1];
  for (int q = 0; q <= ITEM_DOT; ++q) p->first[q + 1] += p->first[q];
  bool ok = hipGetDevice(&p->device) == hipSuccess;
  int cus = 0;
  ok = ok && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device) == hipSuccess;
  ok = ok && allow_lds();
  p->cus = cus > 0 ? cus : 1;
  // Deal each kind's items (largest first) to the workgroup with the least
  // estimated cost so far; ties go to the lower index (the plan is
  // deterministic). Workgroup g's list is offs[g] .. offs[g + 1] - 1 of the
  // kind's slice, in the order dealt.
  std::vector<Item> dealt(sorted.size());
  std::vector<int> offs;
  double kcost[ITEM_DOT + 1] = {};
  for (int q = 0; q <= ITEM_DOT; ++q) {
    const int b = p->first[q], n = p->first[q + 1] - b;
    for (int i = 0; i < n; ++i) kcost[q] += scost[(size_t)(b + i)];
    const int resident = p->cus * kKindWPE[q];
    const int G = n < resident ? n : resident;
    p->groups[q] = G;
    p->ofirst[q] = (int)offs.size();
    if (G <= 0) continue;
    std::vector<std::vector<int>> lists((size_t)G);
    typedef std::pair<double, int> Load;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (int g = 0; g < G; ++g) heap.push(Load(0.0, g));
    for (int i = 0; i < n; ++i) {
      const Load l = heap.top();
      heap.pop();
      lists[(size_t)l.second].push_back(b + i);
      heap.push(Load(l.first + scost[(size_t)(b + i)], l.second));
    }
    int pos = b;
    for (int g = 0; g < G; ++g) {
      offs.push_back(pos - b);
      for (int i : lists[(size_t)g]) dealt[(size_t)pos++] = sorted[(size_t)i];
    }
    offs.push_back(pos - b);
  }
  for (int q = 0; q < ITEM_DOT; ++q) p->order[q] = q;
  std::stable_sort(p->order, p->order + ITEM_DOT, [&](int x, int y) { return kcost[x] > kcost[y]; });
  if (p->groups[ITEM_DOT] > 0 && p->first[ITEM_DOT] > 0) {  // dots and GEMMs: the dots run beside
    ok = ok && hipStreamCreateWithFlags(&p->dstream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) == hipSuccess;
  }
  p->nitems = (int)sorted.size();
  p->ndot = ndot;
  p->ncounters = ndot > 0 ? (int64_t)ndot : 1;
  auto upload = [&](void **dst, const void *src, size_t bytes) {
    if (!ok) return;
    ok = hipMalloc(dst, bytes ? bytes : 16) == hipSuccess;
    if (ok && bytes) ok = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  upload((void **)&p->items, dealt.data(), sizeof(Item) * dealt.size());
  upload((void **)&p->offs, offs.data(), sizeof(int) * offs.size());
  upload((void **)&p->gemms, hg.data(), sizeof(GemmD) * hg.size());
  upload((void **)&p->dots, hd.data(), sizeof(DotD) * hd.size());
  ok = ok && hipMalloc((void **)&p->partials, sizeof(double2) * (size_t)(nparts > 0 ? nparts : 1)) == hipSuccess;
  ok = ok && hipMalloc((void **)&p->counters, sizeof(int) * (size_t)p->ncounters) == hipSuccess;
  if (!ok) {
    free_plan(p);
    return nullptr;
  }
  return p;
}

int mfhip_grouped_items(void *plan) { return plan ? ((Plan *)plan)->nitems : 0; }

int mfhip_grouped_run(void *plan, void *stream) {
  if (!plan) return (int)hipErrorInvalidValue;
  return (int)run_plan((Plan *)plan, (hipStream_t)stream);
}

int mfhip_grouped_run_graph(void *plan, void *stream) {
  if (!plan) return (int)hipErrorInvalidValue;
  Plan *p = (Plan *)plan;
  hipError_t e = hipSuccess;
  if (!p->exec) {
    // capture on the plan's own stream (capturing the legacy default stream
    // is not allowed); ordered against the caller's stream with events
    e = hipStreamCreateWithFlags(&p->gstream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamBeginCapture(p->gstream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t le = run_plan(p, p->gstream);
      e = hipStreamEndCapture(p->gstream, &p->graph);
      if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return (int)e;
  }
  e = hipEventRecord(p->ev_in, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(p->gstream, p->ev_in, 0);
  if (e == hipSuccess) e = hipGraphLaunch(p->exec, p->gstream);
  if (e == hipSuccess) e = hipEventRecord(p->ev_out, p->gstream);
  if (e == hipSuccess) e = hipStreamWaitEvent((hipStream_t)stream, p->ev_out, 0);
  return (int)e;
}

void mfhip_grouped_destroy(void *plan) { free_plan((Plan *)plan); }

}  // extern "C"
}  // namespace multifloats
