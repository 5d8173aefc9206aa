// One (16*TM) x (16*TN) output tile of the reference-order DD GEMM, run by a
// 256-lane workgroup; shared by the standalone GEMM kernel and the grouped
// (batched) kernel.
//
// Semantics follow src/float64x2/matmul.inc:47-239 of the reference: every
// output C(i,j) is one sequential chain over p = 0..k-1 of mac() with A(i,p)
// as the "a" operand, a full TwoSum renormalisation after every complete chunk
// of `ri` MACs that is not the last (when ri > 0 and k > ri), then a final
// TwoSum. The reference's MR x NR blocking never changes that per-output
// order, so any kernel that keeps the chain sequential per output -- as this
// one does -- returns the reference's bits.
//
// Per K-block of BK, A and B tiles are de-interleaved from the AoS {hi,lo}
// HBM layout into separate hi and lo LDS arrays (SoA, carved from `lds`,
// 2*BK*(16*TM + 16*TN) doubles), so each lane reads its TM A-values and TN
// B-values with 16-byte LDS loads and runs TM*TN independent MAC chains out
// of registers: the FP64 VALU, not memory, is the bound (13 FP64 instructions
// per MAC). The next K-block is prefetched into registers while the current
// one is consumed.
//
// R > 0: when the renorm interval ri equals R and R divides BK, in every K
// block that is not the last one the renormalisation points sit at the
// static steps (kk+1) % R == 0 and the block is unrolled without branches
// (the per-step branch otherwise costs a register copy per accumulator at
// each join). R == 0 (or ri != R): ri is checked per step.

#pragma once
#include "dd_core.h"

namespace mfhip_tile {

constexpr int kThreads = 256;

// General form: C(m,n) = op(A) op(B) with BLAS-style leading dimensions.
// TA: op(A) = A^T (element (i,p) at A[p + i*lda]), else A[i + p*lda]; TB
// likewise for B ((p,j) at B[j + p*ldb] or B[p + j*ldb]); C(i,j) at
// C[i + j*ldc]. AB: the result T of the chains is written as
// alpha*T + beta*C (DD operators; C is not read when beta == 0).
// Threads walk the contiguous dimension of each operand, so every layout
// loads coalesced.
// One K block of a single reference-order chain per lane (1 x 1 outputs per
// lane, renormalisation every 8 MACs at fixed points), software-pipelined by
// hand. One such chain per lane is all the parallelism a small product has
// (m n chains, one wave per SIMD), and a MAC's 13 FP64 operations are 8
// deep, so the compiler's order (each MAC's operations one after another)
// leaves most operations waiting on the one just before. Here the 13
// operations of step j are
//   1 p = ah bh, 2 e = fma(ah, bh, -p), 3 q = al bh, 4 cr = fma(ah, bl, q),
//   5 t = s + p, 6 bp = t - s, 7 ap = t - bp, 8 ae = s - ap, 9 be = p - bp,
//   10 x1 = e + cr, 11 x2 = ae + be, 12 x3 = x1 + x2, 13 sl = sl + x3
// (s = the running hi part before the step, t after it), and each issue slot
// of the steady state takes the tail (10-13) of step j, the middle (5-9) of
// step j + 1 and the head (1-4) of step j + 2, in the order
//   5(j+1) 10(j) 6(j+1) 1(j+2) 11(j) 7(j+1) 3(j+2) 12(j) 9(j+1) 8(j+1) 13(j) 2(j+2) 4(j+2)
// so that every operand was produced at least two operations earlier. A group
// of 8 steps starts with the head of its first two steps and the middle of
// the first, and ends with the tail of its last step and the TwoSum
// renormalisation (skipped after the chain's final MAC, as in gemm_tile).
// Every operation, its operands and each chain's order are those of mac() and
// renorm(): the bits are the same. Scheduling barriers pin the order; the LDS
// reads of step j + 4 are issued in slot j, and the first four steps of the
// block are read at its start.
//   Ah, Al: this lane's A entries of the block's first k (hi, lo), step sa
//   apart; Bh, Bl likewise with step sb.
template <int BK>
__device__ __forceinline__ void chain_block8(const double *Ah, const double *Al, const double *Bh, const double *Bl,
                                             int sa, int sb, double &sh, double &sl, bool renorm_end) {
  static_assert(BK % 8 == 0, "groups of 8 steps");
  constexpr int W = 4;  // steps in flight: ring slots
  constexpr int D = 4;  // LDS read distance in steps
  double oh[2 * D], ol[2 * D], ph[2 * D], pl[2 * D];
  double p[W], e[W], q[W], cr[W], t[W], s0[W], bp[W], ap[W], ae[W], be[W], x1[W], x2[W];
  auto pin = [] { __builtin_amdgcn_sched_barrier(0); };
  auto rd = [&](int j) {
    if (j >= BK) return;
    const int r = j % (2 * D);
    oh[r] = Ah[j * sa];
    ol[r] = Al[j * sa];
    ph[r] = Bh[j * sb];
    pl[r] = Bl[j * sb];
  };
  auto op1 = [&](int j) { p[j % W] = oh[j % (2 * D)] * ph[j % (2 * D)]; pin(); };
  auto op2 = [&](int j) { e[j % W] = __builtin_fma(oh[j % (2 * D)], ph[j % (2 * D)], -p[j % W]); pin(); };
  auto op3 = [&](int j) { q[j % W] = ol[j % (2 * D)] * ph[j % (2 * D)]; pin(); };
  auto op4 = [&](int j) { cr[j % W] = __builtin_fma(oh[j % (2 * D)], pl[j % (2 * D)], q[j % W]); pin(); };
  auto op5 = [&](int j) { s0[j % W] = sh; t[j % W] = sh + p[j % W]; sh = t[j % W]; pin(); };
  auto op6 = [&](int j) { bp[j % W] = t[j % W] - s0[j % W]; pin(); };
  auto op7 = [&](int j) { ap[j % W] = t[j % W] - bp[j % W]; pin(); };
  auto op8 = [&](int j) { ae[j % W] = s0[j % W] - ap[j % W]; pin(); };
  auto op9 = [&](int j) { be[j % W] = p[j % W] - bp[j % W]; pin(); };
  auto op10 = [&](int j) { x1[j % W] = e[j % W] + cr[j % W]; pin(); };
  auto op11 = [&](int j) { x2[j % W] = ae[j % W] + be[j % W]; pin(); };
  auto op13 = [&](int j) {  // ops 12 and 13
    const double x3 = x1[j % W] + x2[j % W];
    pin();
    sl = sl + x3;
    pin();
  };
#pragma unroll
  for (int j = 0; j < D; ++j) rd(j);
#pragma unroll
  for (int g = 0; g < BK; g += 8) {
    op1(g); op3(g); op1(g + 1); op3(g + 1); op2(g); op4(g);
    op5(g); op2(g + 1); op4(g + 1); op6(g); op7(g); op9(g); op8(g);
#pragma unroll
    for (int j = g; j < g + 7; ++j) {
      rd(j + D);
      const bool head = j + 2 < g + 8;
      op5(j + 1); op10(j); op6(j + 1);
      if (head) op1(j + 2);
      op11(j); op7(j + 1);
      if (head) op3(j + 2);
      op13(j); op9(j + 1); op8(j + 1);
      if (head) { op2(j + 2); op4(j + 2); }
    }
    rd(g + 7 + D);
    op10(g + 7); op11(g + 7); op13(g + 7);
    if (g + 8 < BK || renorm_end) {
      renorm(sh, sl);
      pin();
    }
  }
}

template <int TM, int TN, int BK, int R, bool TA = false, bool TB = false, bool AB = false>
__device__ __forceinline__ void gemm_tile(const double2 *__restrict__ A, int64_t lda,
                                          const double2 *__restrict__ B, int64_t ldb,
                                          double2 *__restrict__ C, int64_t ldc, int64_t m, int64_t k,
                                          int64_t n, int64_t ri, int64_t m0, int64_t n0, double *lds,
                                          ddv alpha = ddv{1.0, 0.0}, ddv beta = ddv{0.0, 0.0}) {
  static_assert(R == 0 || BK % R == 0, "static renorm interval must divide the K block");
  constexpr int BM = 16 * TM, BN = 16 * TN;
  constexpr int A_PER = BM * BK / kThreads;
  constexpr int B_PER = BN * BK / kThreads;
  static_assert(A_PER * kThreads == BM * BK && B_PER * kThreads == BN * BK, "tile");
  double(*As_h)[BM] = (double(*)[BM])lds;
  double(*As_l)[BM] = (double(*)[BM])(lds + BK * BM);
  double(*Bs_h)[BN] = (double(*)[BN])(lds + 2 * BK * BM);
  double(*Bs_l)[BN] = (double(*)[BN])(lds + 2 * BK * BM + BK * BN);

  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;

  // Global -> register staging coordinates: element e of the A tile
  // (BM rows x BK cols) is (ar(e), ac(e)), of the B tile (BK x BN) (br, bc);
  // consecutive threads walk the operand's contiguous dimension.
  auto ar = [](int e) { return TA ? e / BK : e % BM; };
  auto ac = [](int e) { return TA ? e % BK : e / BM; };
  auto br = [](int e) { return TB ? e / BN : e % BK; };
  auto bc = [](int e) { return TB ? e % BN : e / BK; };
  double2 ra[A_PER], rb[B_PER];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = tid + q * kThreads;
      const int64_t gi = m0 + ar(e), gp = k0 + ac(e);
      ra[q] = (gi < m && gp < k) ? A[TA ? gp + gi * lda : gi + gp * lda] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * kThreads;
      const int64_t gp = k0 + br(e), gj = n0 + bc(e);
      rb[q] = (gp < k && gj < n) ? B[TB ? gj + gp * ldb : gp + gj * ldb] : make_double2(0.0, 0.0);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int e = tid + q * kThreads;
      As_h[ac(e)][ar(e)] = ra[q].x;
      As_l[ac(e)][ar(e)] = ra[q].y;
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int e = tid + q * kThreads;
      Bs_h[br(e)][bc(e)] = rb[q].x;
      Bs_l[br(e)][bc(e)] = rb[q].y;
    }
  };

  double sh[TM][TN], sl[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) sh[i][j] = sl[i][j] = 0.0;

  const int64_t rint = (ri > 0 && k > ri) ? ri : 0;
  int64_t until = rint;  // MACs left before the next renormalisation point
  int64_t done = 0;

  auto macs = [&](int kk) {
    double ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      ah[i] = As_h[kk][tx * TM + i];
      al[i] = As_l[kk][tx * TM + i];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      bh[j] = Bs_h[kk][ty * TN + j];
      bl[j] = Bs_l[kk][ty * TN + j];
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) mac(ah[i], al[i], bh[j], bl[j], sh[i][j], sl[i][j]);
  };
  auto renorm_all = [&]() {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) renorm(sh[i][j], sl[i][j]);
  };
  auto step = [&](int kk) {
    macs(kk);
    ++done;
    if (rint && --until == 0) {
      until = rint;
      if (done < k) renorm_all();
    }
  };

  load(0);
  for (int64_t k0 = 0; k0 < k; k0 += BK) {
    __syncthreads();
    stash();
    __syncthreads();
    if (k0 + BK < k) load(k0 + BK);
    const int64_t kmax = (k - k0) < BK ? (k - k0) : BK;
    if (R > 0 && rint == R && kmax == BK) {
      // static renormalisation points; `until` stays in phase since R | BK.
      // A full last block (k0 + BK == k) skips the point after its final
      // MAC, as the counted path does (done == k there). One chain per lane
      // runs the hand-pipelined block above; with more chains per lane one
      // group of R steps already holds enough independent work.
      constexpr int RR = R > 0 ? R : 1;
      const bool last = k0 + BK >= k;
      if constexpr (TM == 1 && TN == 1 && R == 8) {
        chain_block8<BK>(&As_h[0][tx], &As_l[0][tx], &Bs_h[0][ty], &Bs_l[0][ty], BM, BN, sh[0][0], sl[0][0], !last);
      } else {
#pragma unroll 1
        for (int c = 0; c < BK; c += RR) {
#pragma unroll
          for (int kk = 0; kk < RR; ++kk) macs(c + kk);
          if (!last || c + RR < BK) renorm_all();
        }
      }
      done += BK;
    } else if (kmax == BK) {
#pragma unroll
      for (int kk = 0; kk < BK; ++kk) step(kk);
    } else {
      for (int kk = 0; kk < (int)kmax; ++kk) step(kk);
    }
  }

#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gj = n0 + ty * TN + j;
    if (gj >= n) continue;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int64_t gi = m0 + tx * TM + i;
      if (gi >= m) continue;
      double h = sh[i][j], l = sl[i][j];
      renorm(h, l);
      if (AB) {
        ddv o = dd_mul(alpha, mk(h, l));
        if (beta.hi != 0.0 || beta.lo != 0.0) {
          const double2 c0 = C[gi + gj * ldc];
          o = dd_add(o, dd_mul(beta, mk(c0.x, c0.y)));
        }
        h = o.hi;
        l = o.lo;
      }
      C[gi + gj * ldc] = make_double2(h, l);
    }
  }
}

// 64 x 64 tile of the standalone matmuldd_mm (contiguous column-major
// operands, 4 x 4 chains per lane) with operands copied straight from global
// memory into LDS: no staging registers, no de-interleave pass, one barrier
// per K block. Two LDS stages of BK k: the copies of block b + 1 run during
// the MACs of block b.
//   A stage: BK columns of 64 rows, {hi, lo} pairs, column kk at kk * 1 KiB
//     (one 1 KiB copy per column: 64 lanes x 16 bytes);
//   B stage: 64 columns of BK k, 1 KiB copy q holding columns 4q .. 4q + 3
//     (BK = 16) at q * 1 KiB + 16 (q % 4) + 64 (q / 4): the four copies read
//     by one wave's four lane groups sit on different LDS banks.
// Lane (tx, ty) owns rows tx + 16 i and columns 4 ty + j (i, j < 4): the A
// reads of a step are 16 consecutive pairs (all banks once), broadcast over
// the four lane groups. Copies clamp their rows, columns and k into the
// operand, so an edge tile never reads outside it: the rows and columns past
// m and n are computed from repeated data and not stored, and the k past the
// operand's end are never taken (the last block runs kmax steps).
// In the static renormalisation path a step's 16 MACs run as four 2 x 2
// blocks (rows 0-1 x cols 0-1, rows 0-1 x cols 2-3, rows 2-3 x cols 0-1,
// rows 2-3 x cols 2-3), and each operand pair is reloaded with the next
// step's values as soon as its last MAC of the step has issued (A rows 0-1
// after the second block, B cols 0-1 after the third, the rest after the
// fourth; scheduling barriers keep each reload ahead of the next block, which
// the scheduler would otherwise sink to its use). Each output's MACs,
// renormalisation points and final TwoSum are those of gemm_tile, in the same
// order, so the bits are the same.
constexpr int kDirectBK = 16;
constexpr int kDirectStageA = kDirectBK * 1024;
constexpr int kDirectStageB = 16 * 1024 + 3 * 16 + 3 * 64 + 64;  // copy 15 ends at 15 KiB + 48 + 192 + 1 KiB
constexpr int kDirectStage = kDirectStageA + kDirectStageB;
constexpr int kDirectLds = 2 * kDirectStage;
constexpr int direct_b_base(int q) { return q * 1024 + 16 * (q & 3) + 64 * (q >> 2); }

template <int R>
__device__ __forceinline__ void plain_tile_direct(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                  double2 *__restrict__ C, int64_t m, int64_t k, int64_t n,
                                                  int64_t ri, int64_t m0, int64_t n0, char *lds) {
  constexpr int BK = kDirectBK;
  static_assert(BK == 16, "B copies of four 16-k columns");
  static_assert(direct_b_base(15) + 1024 <= kDirectStageB, "B stage size");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tx = tid & 15, ty = tid >> 4;
  const int64_t nb = (k + BK - 1) / BK;
  // copies of K block kb into stage buffer buf: this wave's A columns
  // wave, wave + 4, ... and B copies wave, wave + 4, ...
  const int64_t ra = m0 + lane < m ? m0 + lane : m - 1;   // A row of this lane
  const int bc = lane >> 4, bk = lane & 15;                 // B copy: column in the copy, k
  auto issue = [&](int64_t kb, int buf) {
    char *st = lds + buf * kDirectStage;
#pragma unroll
    for (int c = wave; c < BK; c += 4) {
      const int64_t p = kb * BK + c < k ? kb * BK + c : k - 1;
      __builtin_amdgcn_global_load_lds((const void *)(A + ra + p * m),
                                       (__attribute__((address_space(3))) void *)(st + c * 1024), 16, 0, 0);
    }
#pragma unroll
    for (int q = wave; q < 16; q += 4) {
      const int64_t col = n0 + 4 * q + bc < n ? n0 + 4 * q + bc : n - 1;
      const int64_t p = kb * BK + bk < k ? kb * BK + bk : k - 1;
      __builtin_amdgcn_global_load_lds((const void *)(B + p + col * k),
                                       (__attribute__((address_space(3))) void *)(st + kDirectStageA + direct_b_base(q)),
                                       16, 0, 0);
    }
  };

  double sh[4][4], sl[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) sh[i][j] = sl[i][j] = 0.0;
  const int64_t rint = (ri > 0 && k > ri) ? ri : 0;
  int64_t until = rint;
  int64_t done = 0;
  double ah[4], al[4], bh[4], bl[4];
  auto ld_a = [&](const char *st, int kk, int i0) {
#pragma unroll
    for (int i = i0; i < i0 + 2; ++i) {
      const double2 v = *(const double2 *)(st + kk * 1024 + (tx + 16 * i) * 16);
      ah[i] = v.x;
      al[i] = v.y;
    }
  };
  auto ld_b = [&](const char *st, int kk, int j0) {
#pragma unroll
    for (int j = j0; j < j0 + 2; ++j) {
      const double2 v = *(const double2 *)(st + kDirectStageA + direct_b_base(ty) + j * BK * 16 + kk * 16);
      bh[j] = v.x;
      bl[j] = v.y;
    }
  };
  auto blk = [&](int i0, int j0) {
#pragma unroll
    for (int i = i0; i < i0 + 2; ++i)
#pragma unroll
      for (int j = j0; j < j0 + 2; ++j) mac(ah[i], al[i], bh[j], bl[j], sh[i][j], sl[i][j]);
  };
  auto renorm_all = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) renorm(sh[i][j], sl[i][j]);
  };

  if (nb > 0) issue(0, 0);  // (the launcher never passes k = 0)
  for (int64_t kb = 0; kb < nb; ++kb) {
    // this wave's copies of block kb have landed (the only ones in flight),
    // every wave's too after the barrier, and every wave is done with the
    // other buffer (block kb - 1)
    __syncthreads();
    if (kb + 1 < nb) issue(kb + 1, (int)((kb + 1) & 1));
    const char *st = lds + (int)(kb & 1) * kDirectStage;
    const int64_t k0 = kb * BK;
    const int kmax = (int)((k - k0) < BK ? (k - k0) : BK);
    if (R > 0 && rint == R && k0 + BK < k) {
      constexpr int RR = R > 0 ? R : 1;
      ld_a(st, 0, 0);
      ld_a(st, 0, 2);
      ld_b(st, 0, 0);
      ld_b(st, 0, 2);
#pragma unroll 1
      for (int c = 0; c < BK; c += RR) {
#pragma unroll
        for (int kk = 0; kk < RR; ++kk) {
          // the block's last step reloads step 0 of this block (unused): no branch
          const int nk = (c + kk + 1) & (BK - 1);
          static_assert((BK & (BK - 1)) == 0, "BK a power of 2");
          blk(0, 0);
          blk(0, 2);
          ld_a(st, nk, 0);
          __builtin_amdgcn_sched_barrier(0);
          blk(2, 0);
          ld_b(st, nk, 0);
          __builtin_amdgcn_sched_barrier(0);
          blk(2, 2);
          ld_a(st, nk, 2);
          ld_b(st, nk, 2);
          __builtin_amdgcn_sched_barrier(0);
        }
        renorm_all();
      }
      done += BK;
    } else {
      auto step = [&](int kk) {
        ld_a(st, kk, 0);
        ld_a(st, kk, 2);
        ld_b(st, kk, 0);
        ld_b(st, kk, 2);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) mac(ah[i], al[i], bh[j], bl[j], sh[i][j], sl[i][j]);
        ++done;
        if (rint && --until == 0) {
          until = rint;
          if (done < k) renorm_all();
        }
      };
      if (kmax == BK) {
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) step(kk);
      } else {
        for (int kk = 0; kk < kmax; ++kk) step(kk);
      }
    }
  }

#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t gj = n0 + 4 * ty + j;
    if (gj >= n) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t gi = m0 + tx + 16 * i;
      if (gi >= m) continue;
      double h = sh[i][j], l = sl[i][j];
      renorm(h, l);
      C[gi + gj * m] = make_double2(h, l);
    }
  }
}

// The same direct copies for a (16 TM) x (16 TN) tile of any of the grouped
// shapes (TM, TN in 1..8; one MAC per chain and step, in the order of
// gemm_tile's steps), for op(A) = A or A^T and op(B) = B or B^T. Layout of
// a stage:
//   A (op(A) = A): column kk at kk * BM * 16 bytes, its BM rows contiguous
//     (ceil(BM / 64) copies of up to 64 rows; lanes past the tile's rows copy
//     nothing); lane tx owns rows tx + 16 i, and a read takes 16 consecutive
//     rows;
//   A (TA: rows contiguous along k): copy q holds rows 4q .. 4q + 3 by 16 k
//     at direct_b_base(q); lane tx owns rows TM tx + i, so for TM = 4 the 16
//     rows of a read sit in 16 copies at distinct bank offsets (two-way
//     conflicts for TM = 2, none for TM = 1);
//   B (op(B) = B): 4 TN copies of four 16-k columns, copy q at
//     direct_b_base(q). Lane group ty = 4 w + u (w = ty / 4, u = ty % 4)
//     owns, for j < TN, the column 4 q + c of copy q = u + 4 t, with
//     t = (TN w + j) % TN and c = (TN w + j) / TN: a bijection onto the
//     tile's columns in which the four lane groups of a wave always read four
//     copies with different bank offsets;
//   B (TB: columns contiguous for a given k): row kk at kk * BN * 16 bytes
//     (ceil(BN / 64) copies); lane group ty owns columns ty + 16 j, so the
//     four lane groups of a wave read four consecutive columns.
template <int TM, int TN, bool TA = false, bool TB = false>
struct DirectShape {
  static constexpr int BK = kDirectBK;
  static constexpr int BM = 16 * TM, BN = 16 * TN;
  static constexpr int kACopies = TA ? BM / 4 : (BM + 63) / 64;  // TA: in total; else per A column
  static constexpr int kBCopies = TB ? (BN + 63) / 64 : 4 * TN;   // TB: per B row; else in total
  static constexpr int kStageA = TA ? ((direct_b_base(kACopies - 1) + 1024) + 15) / 16 * 16 : BK * BM * 16;
  static constexpr int kStageB = TB ? BK * BN * 16 : ((direct_b_base(kBCopies - 1) + 1024) + 15) / 16 * 16;
  static constexpr int kStage = kStageA + kStageB;
  static constexpr int kLds = 2 * kStage;
};

// Leading dimensions lda, ldb, ldc as in gemm_tile (op(A)(i, p) at
// A[i + p lda], or A[p + i lda] with TA; op(B)(p, j) at B[p + j ldb], or
// B[j + p ldb] with TB; C(i, j) at C[i + j ldc]); AB: the result T of the
// chains is written as alpha T + beta C.
template <int TM, int TN, int R, bool AB = false, bool TA = false, bool TB = false>
__device__ __forceinline__ void grouped_tile_direct(const double2 *__restrict__ A, int64_t lda,
                                                    const double2 *__restrict__ B, int64_t ldb,
                                                    double2 *__restrict__ C, int64_t ldc, int64_t m, int64_t k,
                                                    int64_t n, int64_t ri, int64_t m0, int64_t n0, char *lds,
                                                    ddv alpha = ddv{1.0, 0.0}, ddv beta = ddv{0.0, 0.0}) {
  typedef DirectShape<TM, TN, TA, TB> S;
  constexpr int BK = S::BK, BM = S::BM, BN = S::BN;
  static_assert(BK == 16, "copies of 16-k rows and columns");
  static_assert((TA ? S::kACopies : BK * S::kACopies) % 4 == 0 && (TB ? BK * S::kBCopies : S::kBCopies) % 4 == 0,
                "copies split over 4 waves");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tx = tid & 15, ty = tid >> 4;
  const int tw = ty >> 2, tu = ty & 3;
  const int64_t nb = (k + BK - 1) / BK;
  const int bc = lane >> 4, bk = lane & 15;  // a 4 x 16 copy: row or column in the copy, k
  auto issue = [&](int64_t kb, int buf) {
    char *st = lds + buf * S::kStage;
    if constexpr (TA) {
#pragma unroll
      for (int q = wave; q < S::kACopies; q += 4) {
        const int64_t r = m0 + 4 * q + bc < m ? m0 + 4 * q + bc : m - 1;
        const int64_t p = kb * BK + bk < k ? kb * BK + bk : k - 1;
        __builtin_amdgcn_global_load_lds((const void *)(A + p + r * lda),
                                         (__attribute__((address_space(3))) void *)(st + direct_b_base(q)), 16, 0, 0);
      }
    } else {
#pragma unroll
      for (int a = wave; a < BK * S::kACopies; a += 4) {
        const int kk = a / S::kACopies, cc = a % S::kACopies;
        const int rows = BM - 64 * cc < 64 ? BM - 64 * cc : 64;  // rows of this copy
        if (lane < rows) {
          const int64_t r = m0 + 64 * cc + lane < m ? m0 + 64 * cc + lane : m - 1;
          const int64_t p = kb * BK + kk < k ? kb * BK + kk : k - 1;
          __builtin_amdgcn_global_load_lds(
              (const void *)(A + r + p * lda),
              (__attribute__((address_space(3))) void *)(st + kk * BM * 16 + 64 * cc * 16), 16, 0, 0);
        }
      }
    }
    if constexpr (TB) {
#pragma unroll
      for (int a = wave; a < BK * S::kBCopies; a += 4) {
        const int kk = a / S::kBCopies, cc = a % S::kBCopies;
        const int cols = BN - 64 * cc < 64 ? BN - 64 * cc : 64;  // columns of this copy
        if (lane < cols) {
          const int64_t col = n0 + 64 * cc + lane < n ? n0 + 64 * cc + lane : n - 1;
          const int64_t p = kb * BK + kk < k ? kb * BK + kk : k - 1;
          __builtin_amdgcn_global_load_lds(
              (const void *)(B + col + p * ldb),
              (__attribute__((address_space(3))) void *)(st + S::kStageA + kk * BN * 16 + 64 * cc * 16), 16, 0, 0);
        }
      }
    } else {
#pragma unroll
      for (int q = wave; q < S::kBCopies; q += 4) {
        const int64_t col = n0 + 4 * q + bc < n ? n0 + 4 * q + bc : n - 1;
        const int64_t p = kb * BK + bk < k ? kb * BK + bk : k - 1;
        __builtin_amdgcn_global_load_lds(
            (const void *)(B + p + col * ldb),
            (__attribute__((address_space(3))) void *)(st + S::kStageA + direct_b_base(q)), 16, 0, 0);
      }
    }
  };
  // this lane's rows and columns, and where a step's operands sit in a stage
  auto arow = [&](int i) { return TA ? TM * tx + i : tx + 16 * i; };
  int bq[TN], bcol[TN];  // op(B) = B: copy and column in the copy
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int idx = TN * tw + j;
    bq[j] = tu + 4 * (idx % TN);
    bcol[j] = idx / TN;
  }
  auto bcolumn = [&](int j) { return TB ? ty + 16 * j : 4 * bq[j] + bcol[j]; };
  auto a_at = [&](const char *st, int kk, int i) -> const double2 * {
    const int r = arow(i);
    return (const double2 *)(TA ? st + direct_b_base(r >> 2) + (r & 3) * 256 + kk * 16 : st + kk * BM * 16 + r * 16);
  };
  auto b_at = [&](const char *st, int kk, int j) -> const double2 * {
    return (const double2 *)(TB ? st + S::kStageA + kk * BN * 16 + bcolumn(j) * 16
                                : st + S::kStageA + direct_b_base(bq[j]) + bcol[j] * BK * 16 + kk * 16);
  };

  double sh[TM][TN], sl[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) sh[i][j] = sl[i][j] = 0.0;
  const int64_t rint = (ri > 0 && k > ri) ? ri : 0;
  int64_t until = rint;
  int64_t done = 0;
  auto macs = [&](const char *st, int kk) {
    double ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const double2 v = *a_at(st, kk, i);
      ah[i] = v.x;
      al[i] = v.y;
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const double2 v = *b_at(st, kk, j);
      bh[j] = v.x;
      bl[j] = v.y;
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) mac(ah[i], al[i], bh[j], bl[j], sh[i][j], sl[i][j]);
  };
  auto renorm_all = [&]() {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) renorm(sh[i][j], sl[i][j]);
  };

  if (nb > 0) issue(0, 0);  // k = 0: no copies (the outputs are zeros)
  for (int64_t kb = 0; kb < nb; ++kb) {
    // this wave's copies of block kb have landed, every wave's after the
    // barrier, and every wave is done with the other buffer
    __syncthreads();
    if (kb + 1 < nb) issue(kb + 1, (int)((kb + 1) & 1));
    const char *st = lds + (int)(kb & 1) * S::kStage;
    const int64_t k0 = kb * BK;
    const int kmax = (int)((k - k0) < BK ? (k - k0) : BK);
    if (R > 0 && rint == R && k0 + BK < k) {
      constexpr int RR = R > 0 ? R : 1;
#pragma unroll 1
      for (int c = 0; c < BK; c += RR) {
#pragma unroll
        for (int kk = 0; kk < RR; ++kk) macs(st, c + kk);
        renorm_all();
      }
      done += BK;
    } else {
      auto step = [&](int kk) {
        macs(st, kk);
        ++done;
        if (rint && --until == 0) {
          until = rint;
          if (done < k) renorm_all();
        }
      };
      if (kmax == BK) {
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) step(kk);
      } else {
        for (int kk = 0; kk < kmax; ++kk) step(kk);
      }
    }
  }
  // the next item's copies reuse both buffers: every wave done reading them
  __syncthreads();

#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gj = n0 + bcolumn(j);
    if (gj >= n) continue;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int64_t gi = m0 + arow(i);
      if (gi >= m) continue;
      double h = sh[i][j], l = sl[i][j];
      renorm(h, l);
      if (AB) {
        ddv o = dd_mul(alpha, mk(h, l));
        if (beta.hi != 0.0 || beta.lo != 0.0) {
          const double2 c0 = C[gi + gj * ldc];
          o = dd_add(o, dd_mul(beta, mk(c0.x, c0.y)));
        }
        h = o.hi;
        l = o.lo;
      }
      C[gi + gj * ldc] = make_double2(h, l);
    }
  }
}

// The standalone matmuldd_mm tile: contiguous column-major operands.
template <int TM, int TN, int BK, int R>
__device__ __forceinline__ void plain_tile(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                           double2 *__restrict__ C, int64_t m, int64_t k, int64_t n,
                                           int64_t ri, int64_t m0, int64_t n0, double *lds) {
  gemm_tile<TM, TN, BK, R>(A, m, B, k, C, m, m, k, n, ri, m0, n0, lds);
}

}  // namespace mfhip_tile
