// C-ABI layer: the reference-compatible host-pointer entry points, the array
// entry points and the device-pointer API declared in include/multifloats_hip.h.
//
// Host-pointer calls stage through per-thread device buffers and a per-thread
// stream (created once, default priority), copy in, launch, copy out and
// synchronise -- so they are synchronous and re-entrant like the reference's.
// Nothing here computes on the CPU: on any HIP failure the outputs are set to
// NaN, the error is printed and kept for mfhip_last_error().
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>


#include "../../include/multifloats_hip.h"
#include "dd_core.h"
#include "mf_internal.h"

using multifloats::complex64x2;
using multifloats::float64x2;

// Device scratch of mfhip_reduce_noalloc: the fold partials of one reduction,
// held by each GPU the library runs on until the process ends. Its name is a
// host-side stand-in: host code uses only the address that
// hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Per-device address of `scratch`, looked up once per device (with that
// device current) on the first mfhip_reduce_noalloc call there.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

hipError_t noalloc_scratch(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

// Per-thread device buffers and stream of the host-pointer entry points. They
// are released when the thread exits and when the thread moves to another
// device, with their own device made current for the frees.
struct Staging {
  int device = -1;
  hipStream_t stream = nullptr;
  void *buf[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t cap[4] = {0, 0, 0, 0};

  Staging() = default;
  Staging(const Staging &) = delete;
  Staging &operator=(const Staging &) = delete;
  ~Staging() { release(); }

  void release() {
    if (device < 0) return;
    int cur = -1;
    const bool switched = hipGetDevice(&cur) == hipSuccess && cur != device && hipSetDevice(device) == hipSuccess;
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    for (int i = 0; i < 4; ++i)
      if (buf[i]) (void)hipFree(buf[i]);
    if (switched) (void)hipSetDevice(cur);
    device = -1;
    stream = nullptr;
    for (int i = 0; i < 4; ++i) {
      buf[i] = nullptr;
      cap[i] = 0;
    }
  }
};

thread_local Staging tl_stage;
thread_local int tl_err = 0;

hipError_t stage_init() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (tl_stage.stream && tl_stage.device == dev) return hipSuccess;
  tl_stage.release();  // a device switch: free the other device's set first
  tl_stage.device = dev;
  return hipStreamCreateWithFlags(&tl_stage.stream, hipStreamNonBlocking);
}

hipError_t stage_buf(int slot, size_t bytes, void **out) {
  if (bytes == 0) bytes = 16;
  if (tl_stage.cap[slot] < bytes) {
    if (tl_stage.buf[slot]) {
      hipError_t e = hipFree(tl_stage.buf[slot]);
      tl_stage.buf[slot] = nullptr;
      tl_stage.cap[slot] = 0;
      if (e != hipSuccess) return e;
    }
    size_t want = bytes + bytes / 4;  // grow with headroom
    hipError_t e = hipMalloc(&tl_stage.buf[slot], want);
    if (e != hipSuccess) return e;
    tl_stage.cap[slot] = want;
  }
  *out = tl_stage.buf[slot];
  return hipSuccess;
}

void nan_fill(float64x2 *out, int64_t n) {
  if (!out) return;
  for (int64_t i = 0; i < n; ++i) out[i].limbs[0] = out[i].limbs[1] = NAN;
}

void report(const char *fn, hipError_t e) {
  tl_err = (int)e;
  fprintf(stderr, "multifloats_hip: %s failed: %s\n", fn, hipGetErrorString(e));
}

#define TRY(x)                     \
  do {                             \
    hipError_t e_ = (x);           \
    if (e_ != hipSuccess) return e_; \
  } while (0)

hipError_t h2d(void *dst, const void *src, size_t bytes) {
  if (!bytes) return hipSuccess;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, tl_stage.stream);
}
hipError_t d2h(void *dst, const void *src, size_t bytes) {
  if (!bytes) return hipSuccess;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, tl_stage.stream);
}

// Column-major block of `rows` x `cols` elements of `esz` bytes with leading
// dimension ld: only the first `rows` entries of each column are copied, so
// the copy stays inside the ld (cols - 1) + rows elements a BLAS caller owns
// and leaves the rows below `rows` untouched (other threads may be updating
// them). Device and host blocks use the same leading dimension.
hipError_t h2d_block(void *dst, const void *src, int64_t rows, int64_t cols, int64_t ld, size_t esz) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  return hipMemcpy2DAsync(dst, esz * (size_t)ld, src, esz * (size_t)ld, esz * (size_t)rows, (size_t)cols,
                          hipMemcpyHostToDevice, tl_stage.stream);
}
hipError_t d2h_block(void *dst, const void *src, int64_t rows, int64_t cols, int64_t ld, size_t esz) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  return hipMemcpy2DAsync(dst, esz * (size_t)ld, src, esz * (size_t)ld, esz * (size_t)rows, (size_t)cols,
                          hipMemcpyDeviceToHost, tl_stage.stream);
}

const double2 *D(const void *p) { return static_cast<const double2 *>(p); }
double2 *W(void *p) { return static_cast<double2 *>(p); }

hipError_t host_mm(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t m, int64_t k,
                   int64_t n, int64_t ri) {
  tl_err = 0;
  if (m <= 0 || n <= 0) return hipSuccess;
  TRY(stage_init());
  const size_t kk = k > 0 ? (size_t)k : 0;
  void *da, *db, *dc;
  TRY(stage_buf(0, 16 * (size_t)m * kk, &da));
  TRY(stage_buf(1, 16 * kk * (size_t)n, &db));
  TRY(stage_buf(2, 16 * (size_t)m * n, &dc));
  TRY(h2d(da, a, 16 * (size_t)m * kk));
  TRY(h2d(db, b, 16 * kk * (size_t)n));
  TRY(mfhip::gemm_plain(D(da), D(db), W(dc), m, k, n, ri, -1, tl_stage.stream));
  TRY(d2h(c, dc, 16 * (size_t)m * n));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_mv(const float64x2 *a, const float64x2 *x, float64x2 *y, int64_t m, int64_t k,
                   int64_t ri) {
  tl_err = 0;
  if (m <= 0) return hipSuccess;
  TRY(stage_init());
  const size_t kk = k > 0 ? (size_t)k : 0;
  void *da, *dx, *dy;
  TRY(stage_buf(0, 16 * (size_t)m * kk, &da));
  TRY(stage_buf(1, 16 * kk, &dx));
  TRY(stage_buf(2, 16 * (size_t)m, &dy));
  TRY(h2d(da, a, 16 * (size_t)m * kk));
  TRY(h2d(dx, x, 16 * kk));
  TRY(mfhip::gemv(D(da), D(dx), W(dy), m, k, ri, tl_stage.stream));
  TRY(d2h(y, dy, 16 * (size_t)m));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_vm(const float64x2 *x, const float64x2 *b, float64x2 *y, int64_t k, int64_t n,
                   int64_t ri) {
  tl_err = 0;
  if (n <= 0) return hipSuccess;
  TRY(stage_init());
  const size_t kk = k > 0 ? (size_t)k : 0;
  void *dx, *db, *dy;
  TRY(stage_buf(0, 16 * kk, &dx));
  TRY(stage_buf(1, 16 * kk * (size_t)n, &db));
  TRY(stage_buf(2, 16 * (size_t)n, &dy));
  TRY(h2d(dx, x, 16 * kk));
  TRY(h2d(db, b, 16 * kk * (size_t)n));
  TRY(mfhip::gevm(D(dx), D(db), W(dy), k, n, ri, tl_stage.stream));
  TRY(d2h(y, dy, 16 * (size_t)n));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_unary(int op, const float64x2 *a, float64x2 *c, int64_t n) {
  tl_err = 0;
  if (n <= 0) return hipSuccess;
  TRY(stage_init());
  void *da, *dc;
  TRY(stage_buf(0, 16 * (size_t)n, &da));
  TRY(stage_buf(2, 16 * (size_t)n, &dc));
  TRY(h2d(da, a, 16 * (size_t)n));
  TRY(mfhip::unary(op, D(da), W(dc), n, tl_stage.stream));
  TRY(d2h(c, dc, 16 * (size_t)n));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_binary(int op, const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n) {
  tl_err = 0;
  if (n <= 0) return hipSuccess;
  TRY(stage_init());
  void *da, *db, *dc;
  TRY(stage_buf(0, 16 * (size_t)n, &da));
  TRY(stage_buf(1, 16 * (size_t)n, &db));
  TRY(stage_buf(2, 16 * (size_t)n, &dc));
  TRY(h2d(da, a, 16 * (size_t)n));
  TRY(h2d(db, b, 16 * (size_t)n));
  TRY(mfhip::binary(op, D(da), D(db), W(dc), n, tl_stage.stream));
  TRY(d2h(c, dc, 16 * (size_t)n));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_reduce(int kind, const float64x2 *a, const float64x2 *b, int64_t n, int64_t ri,
                       float64x2 *out) {
  tl_err = 0;
  TRY(stage_init());
  const size_t nn = n > 0 ? (size_t)n : 0;
  void *da, *db = nullptr, *dw;
  TRY(stage_buf(0, 16 * nn + 16, &da));
  if (kind == mfhip::RED_DOT) TRY(stage_buf(1, 16 * nn, &db));
  TRY(stage_buf(3, mfhip::reduce_workspace_bytes(n) + 16, &dw));
  TRY(h2d(da, a, 16 * nn));
  if (kind == mfhip::RED_DOT) TRY(h2d(db, b, 16 * nn));
  double2 *dout = W(dw);  // result slot first, partials after it
  TRY(mfhip::reduce(kind, D(da), D(db), n, ri, dout, dout + 1, tl_stage.stream));
  TRY(d2h(out, dout, 16));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_reduce_mask(int kind, const float64x2 *a, const int32_t *mask, int64_t n, float64x2 *out) {
  tl_err = 0;
  TRY(stage_init());
  const size_t nn = n > 0 ? (size_t)n : 0;
  void *da, *dm, *dw;
  TRY(stage_buf(0, 16 * nn + 16, &da));
  TRY(stage_buf(1, 4 * nn + 16, &dm));
  TRY(stage_buf(3, mfhip::reduce_workspace_bytes(n) + 16, &dw));
  TRY(h2d(da, a, 16 * nn));
  TRY(h2d(dm, mask, 4 * nn));
  double2 *dout = W(dw);
  TRY(mfhip::reduce(kind, D(da), nullptr, n, 0, dout, dout + 1, tl_stage.stream, static_cast<const int32_t *>(dm)));
  TRY(d2h(out, dout, 16));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_cdot(const complex64x2 *a, const complex64x2 *b, int64_t n, int64_t ri,
                     complex64x2 *out) {
  tl_err = 0;
  TRY(stage_init());
  const size_t nn = n > 0 ? (size_t)n : 0;
  void *da, *db, *dw;
  TRY(stage_buf(0, 32 * nn + 32, &da));
  TRY(stage_buf(1, 32 * nn + 32, &db));
  TRY(stage_buf(3, mfhip::cdot_workspace_bytes(n) + 32, &dw));
  TRY(h2d(da, a, 32 * nn));
  TRY(h2d(db, b, 32 * nn));
  double4 *dout = static_cast<double4 *>(dw);
  TRY(mfhip::cdot((const double4 *)da, (const double4 *)db, n, ri, dout, dout + 1, tl_stage.stream));
  TRY(d2h(out, dout, 32));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_dim(int kind, const float64x2 *a, const int32_t *mask, int64_t inner, int64_t len,
                    int64_t outer, float64x2 *out) {
  tl_err = 0;
  if (inner <= 0 || outer <= 0) return hipSuccess;
  if (len < 0) len = 0;
  TRY(stage_init());
  const size_t tot = (size_t)inner * (size_t)len * (size_t)outer;
  const size_t nout = (size_t)inner * (size_t)outer;
  void *da, *dm = nullptr, *dc;
  TRY(stage_buf(0, 16 * tot, &da));
  if (mask) TRY(stage_buf(1, 4 * tot, &dm));
  TRY(stage_buf(2, 16 * nout, &dc));
  TRY(h2d(da, a, 16 * tot));
  if (mask) TRY(h2d(dm, mask, 4 * tot));
  TRY(mfhip::reduce_dim(kind, D(da), static_cast<const int32_t *>(dm), inner, len, outer, W(dc),
                        tl_stage.stream));
  TRY(d2h(out, dc, 16 * nout));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_gemm_blas(char ta, char tb, int64_t m, int64_t n, int64_t k, float64x2 alpha,
                          const float64x2 *a, int64_t lda, const float64x2 *b, int64_t ldb, float64x2 beta,
                          float64x2 *c, int64_t ldc, int64_t ri) {
  tl_err = 0;
  TRY(mfhip::gemm_blas_check(ta, tb, m, n, k, lda, ldb, ldc));
  if (m == 0 || n == 0) return hipSuccess;
  TRY(stage_init());
  const bool nota = ta == 'N' || ta == 'n', notb = tb == 'N' || tb == 'n';
  // the blocks of A, B, C the call reads (column-major with leading dims);
  // device copies keep the caller's leading dimensions
  const int64_t ar = nota ? m : k, ac = nota ? k : m, br = notb ? k : n, bc = notb ? n : k;
  void *da, *db, *dc;
  TRY(stage_buf(0, 16 * (size_t)lda * (size_t)(ac > 0 ? ac : 0), &da));
  TRY(stage_buf(1, 16 * (size_t)ldb * (size_t)(bc > 0 ? bc : 0), &db));
  TRY(stage_buf(2, 16 * (size_t)ldc * (size_t)n, &dc));
  TRY(h2d_block(da, a, ar, ac, lda, 16));
  TRY(h2d_block(db, b, br, bc, ldb, 16));
  TRY(h2d_block(dc, c, m, n, ldc, 16));
  TRY(mfhip::gemm_blas(ta, tb, m, n, k, ddv{alpha.limbs[0], alpha.limbs[1]}, D(da), lda, D(db), ldb,
                       ddv{beta.limbs[0], beta.limbs[1]}, W(dc), ldc, ri, tl_stage.stream));
  TRY(d2h_block(c, dc, m, n, ldc, 16));
  return hipStreamSynchronize(tl_stage.stream);
}

hipError_t host_trsm(char side, char uplo, char ta, char diag, int64_t m, int64_t n, float64x2 alpha,
                     const float64x2 *a, int64_t lda, float64x2 *b, int64_t ldb) {
  tl_err = 0;
  TRY(mfhip::trsm_check(side, uplo, ta, diag, m, n, lda, ldb));
  if (m == 0 || n == 0) return hipSuccess;
  TRY(stage_init());
  const int64_t s = (side == 'L' || side == 'l') ? m : n;
  const size_t wb = mfhip::trsm_workspace_bytes(side, m, n);
  void *da, *db, *dw;
  TRY(stage_buf(0, 16 * (size_t)lda * (size_t)s, &da));
  TRY(stage_buf(1, 16 * (size_t)ldb * (size_t)n, &db));
  TRY(stage_buf(3, wb, &dw));
  TRY(h2d_block(da, a, s, s, lda, 16));  // the s x s triangle's block
  TRY(h2d_block(db, b, m, n, ldb, 16));
  TRY(mfhip::trsm(side, uplo, ta, diag, m, n, ddv{alpha.limbs[0], alpha.limbs[1]}, D(da), lda, W(db), ldb, dw, wb,
                  tl_stage.stream));
  TRY(d2h_block(b, db, m, n, ldb, 16));
  return hipStreamSynchronize(tl_stage.stream);
}

thread_local int tl_oz_used = 0;
thread_local int64_t tl_oz_fix = 0;

hipError_t host_cmm(const complex64x2 *a, const complex64x2 *b, complex64x2 *c, int64_t m,
                    int64_t k, int64_t n, int64_t ri) {
  tl_err = 0;
  if (m <= 0 || n <= 0) return hipSuccess;
  TRY(stage_init());
  const size_t kk = k > 0 ? (size_t)k : 0;
  void *da, *db, *dc;
  TRY(stage_buf(0, 32 * (size_t)m * kk, &da));
  TRY(stage_buf(1, 32 * kk * (size_t)n, &db));
  TRY(stage_buf(2, 32 * (size_t)m * n, &dc));
  TRY(h2d(da, a, 32 * (size_t)m * kk));
  TRY(h2d(db, b, 32 * kk * (size_t)n));
  TRY(mfhip::gemm_cplx((const double4 *)da, (const double4 *)db, (double4 *)dc, m, k, n, ri,
                       tl_stage.stream));
  TRY(d2h(c, dc, 32 * (size_t)m * n));
  return hipStreamSynchronize(tl_stage.stream);
}

void nan_fill_c(complex64x2 *out, int64_t n) {
  if (!out) return;
  for (int64_t i = 0; i < n; ++i)
    out[i].re.limbs[0] = out[i].re.limbs[1] = out[i].im.limbs[0] = out[i].im.limbs[1] = NAN;
}

hipError_t host_mm_ozaki(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t m,
                         int64_t k, int64_t n, int slices) {
  tl_err = 0;
  if (m <= 0 || n <= 0) return hipSuccess;
  if (slices > mfhip::ozaki_max_slices() || (slices > 0 && slices < 7)) return hipErrorInvalidValue;
  TRY(stage_init());
  const size_t kk = k > 0 ? (size_t)k : 0;
  void *da, *db, *dc, *dw;
  const size_t wb = mfhip::ozaki_workspace_bytes(m, k, n, slices);
  TRY(stage_buf(0, 16 * (size_t)m * kk, &da));
  TRY(stage_buf(1, 16 * kk * (size_t)n, &db));
  TRY(stage_buf(2, 16 * (size_t)m * n, &dc));
  TRY(stage_buf(3, wb, &dw));
  TRY(h2d(da, a, 16 * (size_t)m * kk));
  TRY(h2d(db, b, 16 * kk * (size_t)n));
  TRY(mfhip::ozaki_gemm(D(da), D(db), W(dc), m, k, n, slices, dw, wb, tl_stage.stream));
  TRY(d2h(c, dc, 16 * (size_t)m * n));
  // the call is synchronous anyway: read back what the device chose
  return mfhip::ozaki_info(dw, m, k, n, slices, &tl_oz_used, &tl_oz_fix, nullptr, tl_stage.stream);
}

float64x2 nan_dd() {
  float64x2 r;
  r.limbs[0] = r.limbs[1] = NAN;
  return r;
}

}  // namespace

namespace multifloats {
extern "C" {

void matmuldd_mm(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t m, int64_t k,
                 int64_t n, int64_t ri) {
  hipError_t e = host_mm(a, b, c, m, k, n, ri);
  if (e != hipSuccess) {
    report("matmuldd_mm", e);
    nan_fill(c, (m > 0 && n > 0) ? m * n : 0);
  }
}
void matmuldd_mv(const float64x2 *a, const float64x2 *x, float64x2 *y, int64_t m, int64_t k,
                 int64_t ri) {
  hipError_t e = host_mv(a, x, y, m, k, ri);
  if (e != hipSuccess) {
    report("matmuldd_mv", e);
    nan_fill(y, m > 0 ? m : 0);
  }
}
void matmuldd_vm(const float64x2 *x, const float64x2 *b, float64x2 *y, int64_t k, int64_t n,
                 int64_t ri) {
  hipError_t e = host_vm(x, b, y, k, n, ri);
  if (e != hipSuccess) {
    report("matmuldd_vm", e);
    nan_fill(y, n > 0 ? n : 0);
  }
}
void matmuldd_mm_ozaki(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t m,
                       int64_t k, int64_t n, int slices) {
  hipError_t e = host_mm_ozaki(a, b, c, m, k, n, slices);
  if (e != hipSuccess) {
    report("matmuldd_mm_ozaki", e);
    nan_fill(c, (m > 0 && n > 0) ? m * n : 0);
  }
}
void cmatmuldd_mm(const complex64x2 *a, const complex64x2 *b, complex64x2 *c, int64_t m,
                  int64_t k, int64_t n, int64_t ri) {
  hipError_t e = host_cmm(a, b, c, m, k, n, ri);
  if (e != hipSuccess) {
    report("cmatmuldd_mm", e);
    nan_fill_c(c, (m > 0 && n > 0) ? m * n : 0);
  }
}
void cmatmuldd_mv(const complex64x2 *a, const complex64x2 *x, complex64x2 *y, int64_t m,
                  int64_t k, int64_t ri) {
  hipError_t e = host_cmm(a, x, y, m, k, 1, ri);
  if (e != hipSuccess) {
    report("cmatmuldd_mv", e);
    nan_fill_c(y, m > 0 ? m : 0);
  }
}
void cmatmuldd_vm(const complex64x2 *x, const complex64x2 *b, complex64x2 *y, int64_t k,
                  int64_t n, int64_t ri) {
  hipError_t e = host_cmm(x, b, y, 1, k, n, ri);
  if (e != hipSuccess) {
    report("cmatmuldd_vm", e);
    nan_fill_c(y, n > 0 ? n : 0);
  }
}
void matmulf64x2_mm(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t m, int64_t k,
                    int64_t n, int64_t ri) {
  matmuldd_mm(a, b, c, m, k, n, ri);
}
void matmulf64x2_mv(const float64x2 *a, const float64x2 *x, float64x2 *y, int64_t m, int64_t k,
                    int64_t ri) {
  matmuldd_mv(a, x, y, m, k, ri);
}
void matmulf64x2_vm(const float64x2 *x, const float64x2 *b, float64x2 *y, int64_t k, int64_t n,
                    int64_t ri) {
  matmuldd_vm(x, b, y, k, n, ri);
}

#define MF_BINARY(name, op)                                                        \
  void name(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n) {      \
    hipError_t e = host_binary(op, a, b, c, n);                                    \
    if (e != hipSuccess) {                                                          \
      report(#name, e);                                                             \
      nan_fill(c, n > 0 ? n : 0);                                                   \
    }                                                                               \
  }
#define MF_UNARY(name, op)                                                         \
  void name(const float64x2 *a, float64x2 *c, int64_t n) {                          \
    hipError_t e = host_unary(op, a, c, n);                                        \
    if (e != hipSuccess) {                                                          \
      report(#name, e);                                                             \
      nan_fill(c, n > 0 ? n : 0);                                                   \
    }                                                                               \
  }
MF_BINARY(adddd_v, mfhip::OP_ADD)
MF_BINARY(subdd_v, mfhip::OP_SUB)
MF_BINARY(muldd_v, mfhip::OP_MUL)
MF_BINARY(divdd_v, mfhip::OP_DIV)
MF_BINARY(powdd_v, mfhip::OP_POW)
MF_BINARY(hypotdd_v, mfhip::OP_HYPOT)
MF_BINARY(fdimdd_v, mfhip::OP_FDIM)
MF_BINARY(fmindd_v, mfhip::OP_FMIN)
MF_BINARY(fmaxdd_v, mfhip::OP_FMAX)
MF_BINARY(copysigndd_v, mfhip::OP_COPYSIGN)
MF_BINARY(atan2dd_v, mfhip::OP_ATAN2)
MF_BINARY(atan2pidd_v, mfhip::OP_ATAN2PI)
MF_BINARY(mindd_v, mfhip::OP_MINF)
MF_BINARY(maxdd_v, mfhip::OP_MAXF)
MF_BINARY(dimdd_v, mfhip::OP_DIMF)
MF_BINARY(fmoddd_v, mfhip::OP_FMOD)
MF_BINARY(modulodd_v, mfhip::OP_MODULO)
MF_BINARY(nearestdd_v, mfhip::OP_NEAREST)
MF_UNARY(sqrtdd_v, mfhip::OP_SQRT)
MF_UNARY(expdd_v, mfhip::OP_EXP)
MF_UNARY(logdd_v, mfhip::OP_LOG)
MF_UNARY(sindd_v, mfhip::OP_SIN)
MF_UNARY(erfdd_v, mfhip::OP_ERF)
MF_UNARY(cosdd_v, mfhip::OP_COS)
MF_UNARY(tandd_v, mfhip::OP_TAN)
MF_UNARY(exp2dd_v, mfhip::OP_EXP2)
MF_UNARY(expm1dd_v, mfhip::OP_EXPM1)
MF_UNARY(log2dd_v, mfhip::OP_LOG2)
MF_UNARY(log10dd_v, mfhip::OP_LOG10)
MF_UNARY(log1pdd_v, mfhip::OP_LOG1P)
MF_UNARY(sinhdd_v, mfhip::OP_SINH)
MF_UNARY(coshdd_v, mfhip::OP_COSH)
MF_UNARY(tanhdd_v, mfhip::OP_TANH)
MF_UNARY(erfcdd_v, mfhip::OP_ERFC)
MF_UNARY(erfcxdd_v, mfhip::OP_ERFCX)
MF_UNARY(asinhdd_v, mfhip::OP_ASINH)
MF_UNARY(acoshdd_v, mfhip::OP_ACOSH)
MF_UNARY(atanhdd_v, mfhip::OP_ATANH)
MF_UNARY(exp10dd_v, mfhip::OP_EXP10)
MF_UNARY(exp2m1dd_v, mfhip::OP_EXP2M1)
MF_UNARY(exp10m1dd_v, mfhip::OP_EXP10M1)
MF_UNARY(log2p1dd_v, mfhip::OP_LOG2P1)
MF_UNARY(log10p1dd_v, mfhip::OP_LOG10P1)
MF_UNARY(sinpidd_v, mfhip::OP_SINPI)
MF_UNARY(cospidd_v, mfhip::OP_COSPI)
MF_UNARY(tanpidd_v, mfhip::OP_TANPI)
MF_UNARY(atandd_v, mfhip::OP_ATAN)
MF_UNARY(atanpidd_v, mfhip::OP_ATANPI)
MF_UNARY(asindd_v, mfhip::OP_ASIN)
MF_UNARY(acosdd_v, mfhip::OP_ACOS)
MF_UNARY(asinpidd_v, mfhip::OP_ASINPI)
MF_UNARY(acospidd_v, mfhip::OP_ACOSPI)

float64x2 sumdd(const float64x2 *a, int64_t n) {
  float64x2 r;
  hipError_t e = host_reduce(mfhip::RED_SUM, a, nullptr, n, 0, &r);
  if (e != hipSuccess) {
    report("sumdd", e);
    return nan_dd();
  }
  return r;
}
float64x2 dotdd(const float64x2 *a, const float64x2 *b, int64_t n, int64_t ri) {
  float64x2 r;
  hipError_t e = host_reduce(mfhip::RED_DOT, a, b, n, ri, &r);
  if (e != hipSuccess) {
    report("dotdd", e);
    return nan_dd();
  }
  return r;
}
float64x2 norm2dd(const float64x2 *a, int64_t n) {
  float64x2 r;
  hipError_t e = host_reduce(mfhip::RED_NORM2, a, nullptr, n, 0, &r);
  if (e != hipSuccess) {
    report("norm2dd", e);
    return nan_dd();
  }
  return r;
}

float64x2 sumdd_mask(const float64x2 *a, const int32_t *mask, int64_t n) {
  float64x2 r;
  hipError_t e = mask ? host_reduce_mask(mfhip::RED_SUM, a, mask, n, &r)
                      : host_reduce(mfhip::RED_SUM, a, nullptr, n, 0, &r);
  if (e != hipSuccess) {
    report("sumdd_mask", e);
    return nan_dd();
  }
  return r;
}
complex64x2 cdotdd(const complex64x2 *a, const complex64x2 *b, int64_t n, int64_t ri) {
  complex64x2 r;
  hipError_t e = host_cdot(a, b, n, ri, &r);
  if (e != hipSuccess) {
    report("cdotdd", e);
    nan_fill_c(&r, 1);
  }
  return r;
}
void sumdd_dim(const float64x2 *a, const int32_t *mask, int64_t inner, int64_t len, int64_t outer,
               float64x2 *out) {
  hipError_t e = host_dim(mfhip::RED_SUM, a, mask, inner, len, outer, out);
  if (e != hipSuccess) {
    report("sumdd_dim", e);
    nan_fill(out, (inner > 0 && outer > 0) ? inner * outer : 0);
  }
}
#define MF_REDUCE_FAMILY(name, kind)                                                              \
  float64x2 name##dd(const float64x2 *a, int64_t n) {                                             \
    float64x2 r;                                                                                  \
    hipError_t e = host_reduce(kind, a, nullptr, n, 0, &r);                                       \
    if (e != hipSuccess) {                                                                        \
      report(#name "dd", e);                                                                      \
      return nan_dd();                                                                            \
    }                                                                                             \
    return r;                                                                                     \
  }                                                                                               \
  float64x2 name##dd_mask(const float64x2 *a, const int32_t *mask, int64_t n) {                  \
    float64x2 r;                                                                                  \
    hipError_t e = mask ? host_reduce_mask(kind, a, mask, n, &r) : host_reduce(kind, a, nullptr, n, 0, &r); \
    if (e != hipSuccess) {                                                                        \
      report(#name "dd_mask", e);                                                                 \
      return nan_dd();                                                                            \
    }                                                                                             \
    return r;                                                                                     \
  }                                                                                               \
  void name##dd_dim(const float64x2 *a, const int32_t *mask, int64_t inner, int64_t len, int64_t outer, \
                    float64x2 *out) {                                                             \
    hipError_t e = host_dim(kind, a, mask, inner, len, outer, out);                               \
    if (e != hipSuccess) {                                                                        \
      report(#name "dd_dim", e);                                                                  \
      nan_fill(out, (inner > 0 && outer > 0) ? inner * outer : 0);                                \
    }                                                                                             \
  }
MF_REDUCE_FAMILY(product, mfhip::RED_PROD)
MF_REDUCE_FAMILY(maxval, mfhip::RED_MAXVAL)
MF_REDUCE_FAMILY(minval, mfhip::RED_MINVAL)

void norm2dd_dim(const float64x2 *a, int64_t inner, int64_t len, int64_t outer, float64x2 *out) {
  hipError_t e = host_dim(mfhip::RED_NORM2, a, nullptr, inner, len, outer, out);
  if (e != hipSuccess) {
    report("norm2dd_dim", e);
    nan_fill(out, (inner > 0 && outer > 0) ? inner * outer : 0);
  }
}

void wgemmdd(char transa, char transb, int64_t m, int64_t n, int64_t k, float64x2 alpha, const float64x2 *a,
             int64_t lda, const float64x2 *b, int64_t ldb, float64x2 beta, float64x2 *c, int64_t ldc,
             int64_t renorm_interval) {
  hipError_t e = host_gemm_blas(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, renorm_interval);
  if (e != hipSuccess) {
    report("wgemmdd", e);
    if (c && m > 0 && n > 0 && ldc >= m)
      for (int64_t j = 0; j < n; ++j) nan_fill(c + j * ldc, m);
  }
}

void wtrsmdd(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float64x2 alpha, const float64x2 *a,
             int64_t lda, float64x2 *b, int64_t ldb) {
  hipError_t e = host_trsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
  if (e != hipSuccess) {
    report("wtrsmdd", e);
    if (b && m > 0 && n > 0 && ldb >= m)
      for (int64_t j = 0; j < n; ++j) nan_fill(b + j * ldb, m);
  }
}

// ---- device-pointer API ----

static int status(hipError_t e) {
  if (e != hipSuccess) tl_err = (int)e;
  return (int)e;
}
static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }
#define CHECK_ALIGN(...)                                                   \
  do {                                                                     \
    const void *ps_[] = {__VA_ARGS__};                                     \
    for (const void *p_ : ps_)                                             \
      if (p_ && !aligned16(p_)) return status(hipErrorInvalidValue);       \
  } while (0)

int mfhip_matmul_variants(void) { return mfhip::gemm_plain_variants(); }

int mfhip_matmul_mm_variant(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t m,
                            int64_t k, int64_t n, int64_t ri, int variant, void *stream) {
  CHECK_ALIGN(a, b, c);
  if (variant >= mfhip::gemm_plain_variants()) return status(hipErrorInvalidValue);
  return status(mfhip::gemm_plain(D(a), D(b), W(c), m, k, n, ri, variant, (hipStream_t)stream));
}
int mfhip_matmul_mm(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t m, int64_t k,
                    int64_t n, int64_t ri, void *stream) {
  return mfhip_matmul_mm_variant(a, b, c, m, k, n, ri, -1, stream);
}
int mfhip_cmatmul_mm(const complex64x2 *a, const complex64x2 *b, complex64x2 *c, int64_t m,
                     int64_t k, int64_t n, int64_t ri, void *stream) {
  CHECK_ALIGN(a, b, c);
  return status(mfhip::gemm_cplx((const double4 *)a, (const double4 *)b, (double4 *)c, m, k, n, ri,
                                 (hipStream_t)stream));
}
int mfhip_matmul_mv(const float64x2 *a, const float64x2 *x, float64x2 *y, int64_t m, int64_t k,
                    int64_t ri, void *stream) {
  CHECK_ALIGN(a, x, y);
  return status(mfhip::gemv(D(a), D(x), W(y), m, k, ri, (hipStream_t)stream));
}
int mfhip_matmul_vm(const float64x2 *x, const float64x2 *b, float64x2 *y, int64_t k, int64_t n,
                    int64_t ri, void *stream) {
  CHECK_ALIGN(x, b, y);
  return status(mfhip::gevm(D(x), D(b), W(y), k, n, ri, (hipStream_t)stream));
}
int mfhip_unary(int op, const float64x2 *a, float64x2 *c, int64_t n, void *stream) {
  CHECK_ALIGN(a, c);
  return status(mfhip::unary(op, D(a), W(c), n, (hipStream_t)stream));
}
int mfhip_binary(int op, const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t n,
                 void *stream) {
  CHECK_ALIGN(a, b, c);
  return status(mfhip::binary(op, D(a), D(b), W(c), n, (hipStream_t)stream));
}
size_t mfhip_reduce_workspace_bytes(int64_t n) { return mfhip::reduce_workspace_bytes(n); }
int mfhip_reduce(int kind, const float64x2 *a, const float64x2 *b, int64_t n, int64_t ri,
                 float64x2 *out, void *work, void *stream) {
  CHECK_ALIGN(a, b, out, work);
  if (!work || !out) return status(hipErrorInvalidValue);
  return status(mfhip::reduce(kind, D(a), D(b), n, ri, W(out), W(work), (hipStream_t)stream));
}

int mfhip_reduce_noalloc(int kind, const float64x2 *a, const float64x2 *b, int64_t n, int64_t ri,
                         float64x2 *out, void *stream) {
  CHECK_ALIGN(a, b, out);
  if (!out) return status(hipErrorInvalidValue);
  if (n > MFHIP_NOALLOC_MAX_N || mfhip::reduce_workspace_bytes(n) > kNoAllocScratchBytes)
    return status(hipErrorInvalidValue);
  void *work = nullptr;
  hipError_t e = noalloc_scratch(&work);
  if (e != hipSuccess) return status(e);
  return status(mfhip::reduce(kind, D(a), D(b), n, ri, W(out), W(work), (hipStream_t)stream));
}

int mfhip_sum_masked(const float64x2 *a, const int32_t *mask, int64_t n, float64x2 *out, void *work,
                     void *stream) {
  CHECK_ALIGN(a, out, work);
  if (!work || !out || !mask) return status(hipErrorInvalidValue);
  return status(mfhip::reduce(mfhip::RED_SUM, D(a), nullptr, n, 0, W(out), W(work), (hipStream_t)stream,
                              mask));
}
int mfhip_reduce_masked(int kind, const float64x2 *a, const int32_t *mask, int64_t n, float64x2 *out, void *work,
                        void *stream) {
  CHECK_ALIGN(a, out, work);
  if (!work || !out || !mask || kind == mfhip::RED_DOT || kind == mfhip::RED_NORM2)
    return status(hipErrorInvalidValue);
  return status(mfhip::reduce(kind, D(a), nullptr, n, 0, W(out), W(work), (hipStream_t)stream, mask));
}
size_t mfhip_cdot_workspace_bytes(int64_t n) { return mfhip::cdot_workspace_bytes(n); }
int mfhip_cdot(const complex64x2 *a, const complex64x2 *b, int64_t n, int64_t ri, complex64x2 *out,
               void *work, void *stream) {
  CHECK_ALIGN(a, b, out, work);
  if (!work || !out) return status(hipErrorInvalidValue);
  return status(mfhip::cdot((const double4 *)a, (const double4 *)b, n, ri, (double4 *)out,
                            (double4 *)work, (hipStream_t)stream));
}
int mfhip_reduce_dim(int kind, const float64x2 *a, const int32_t *mask, int64_t inner, int64_t len,
                     int64_t outer, float64x2 *out, void *stream) {
  CHECK_ALIGN(a, out);
  return status(mfhip::reduce_dim(kind, D(a), mask, inner, len, outer, W(out), (hipStream_t)stream));
}

int mfhip_gemm(char transa, char transb, int64_t m, int64_t n, int64_t k, float64x2 alpha, const float64x2 *a,
               int64_t lda, const float64x2 *b, int64_t ldb, float64x2 beta, float64x2 *c, int64_t ldc,
               int64_t renorm_interval, void *stream) {
  CHECK_ALIGN(a, b, c);
  return status(mfhip::gemm_blas(transa, transb, m, n, k, ddv{alpha.limbs[0], alpha.limbs[1]}, D(a), lda, D(b), ldb,
                                 ddv{beta.limbs[0], beta.limbs[1]}, W(c), ldc, renorm_interval,
                                 (hipStream_t)stream));
}

size_t mfhip_trsm_workspace_bytes(char side, int64_t m, int64_t n) { return mfhip::trsm_workspace_bytes(side, m, n); }

int mfhip_trsm(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float64x2 alpha, const float64x2 *a,
               int64_t lda, float64x2 *b, int64_t ldb, void *work, size_t work_bytes, void *stream) {
  CHECK_ALIGN(a, b, b);
  return status(mfhip::trsm(side, uplo, transa, diag, m, n, ddv{alpha.limbs[0], alpha.limbs[1]}, D(a), lda, W(b),
                            ldb, work, work_bytes, (hipStream_t)stream));
}

size_t mfhip_ozaki_workspace_bytes(int64_t m, int64_t k, int64_t n, int slices) {
  return mfhip::ozaki_workspace_bytes(m, k, n, slices);
}
int mfhip_ozaki_matmul_mm(const float64x2 *a, const float64x2 *b, float64x2 *c, int64_t m,
                          int64_t k, int64_t n, int slices, void *work, size_t work_bytes,
                          void *stream) {
  CHECK_ALIGN(a, b, c, work);
  if (slices > mfhip::ozaki_max_slices() || (slices > 0 && slices < 7))
    return status(hipErrorInvalidValue);
  tl_oz_used = slices > 0 ? slices : -1;  // automatic: chosen on the device (mfhip_ozaki_info)
  tl_oz_fix = 0;
  return status(mfhip::ozaki_gemm(D(a), D(b), W(c), m, k, n, slices, work, work_bytes, (hipStream_t)stream));
}
int mfhip_ozaki_info(const void *work, int64_t m, int64_t k, int64_t n, int slices, int *slices_used,
                     int64_t *fixups, void *stream) {
  CHECK_ALIGN(work);
  if (!work) return status(hipErrorInvalidValue);
  return status(mfhip::ozaki_info(work, m, k, n, slices, slices_used, fixups, nullptr, (hipStream_t)stream));
}
int mfhip_ozaki_moduli_used(const void *work, int64_t m, int64_t k, int64_t n, int slices, int *moduli,
                            void *stream) {
  CHECK_ALIGN(work);
  if (!work || !moduli) return status(hipErrorInvalidValue);
  return status(mfhip::ozaki_info(work, m, k, n, slices, nullptr, nullptr, moduli, (hipStream_t)stream));
}
int mfhip_ozaki_slices_used(void) { return tl_oz_used; }
int64_t mfhip_ozaki_fixups(void) { return tl_oz_fix; }
int mfhip_ozaki_moduli(int64_t m, int64_t k, int64_t n, int slices) {
  return mfhip::ozaki_moduli(m, k, n, slices);
}
int mfhip_ozaki_choose_slices(int64_t k, int range_a, int range_b) {
  return mfhip::ozaki_choose_slices(k, range_a, range_b);
}
int mfhip_ozaki_prepare(const float64x2 *a, const float64x2 *b, int64_t m, int64_t k, int64_t n,
                        int slices, void *work, size_t work_bytes, void *stream) {
  CHECK_ALIGN(a, b, work);
  return status(mfhip::ozaki_prepare(D(a), D(b), m, k, n, slices, work, work_bytes,
                                     (hipStream_t)stream));
}
int mfhip_ozaki_multiply(float64x2 *c, int64_t m, int64_t k, int64_t n, int slices, void *work,
                         size_t work_bytes, void *stream) {
  CHECK_ALIGN(c, work);
  return status(mfhip::ozaki_multiply(W(c), m, k, n, slices, work, work_bytes,
                                      (hipStream_t)stream));
}
int mfhip_ozaki_products(int64_t m, int64_t k, int64_t n, int slices, void *work, size_t work_bytes,
                         void *stream) {
  CHECK_ALIGN(work);
  return status(mfhip::ozaki_products(m, k, n, slices, work, work_bytes, (hipStream_t)stream));
}

int mfhip_last_error(void) { return tl_err; }
const char *mfhip_error_string(int err) { return hipGetErrorString((hipError_t)err); }

}  // extern "C"
}  // namespace multifloats
