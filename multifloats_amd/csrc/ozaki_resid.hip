// The CRT form's residue pass (k_oz2_resid): the residues of every operand
// element modulo the call's moduli, formed on the matrix cores. Built with
// -mllvm -amdgpu-mfma-vgpr-form=1 (its Makefile rule): the MFMA results then
// land in VGPRs, where the reduction reads them, and the accumulator input
// stays put.
#include <type_traits>

#include "ozaki_shared.h"

namespace {

using namespace mfhip_oz;

// Two's complement 128-bit value (lo, hi) of an integer-valued double D,
// |D| < 2^127: |D| = mant 2^sh with -52 <= sh <= 74, and the bits that a right
// shift drops are zero.
__device__ __forceinline__ void int_double_to_i128(double D, uint64_t &lo, uint64_t &hi) {
  const uint64_t bits = (uint64_t)__double_as_longlong(D);
  const int e = (int)((bits >> 52) & 0x7FF);
  const uint64_t mant = (bits & 0xFFFFFFFFFFFFFull) | (1ull << 52);
  const int sh = e - 1075;
  uint64_t l = 0, h = 0;
  if (e != 0) {
    if (sh <= 0) {
      l = mant >> (-sh);
    } else if (sh < 64) {
      l = mant << sh;
      h = mant >> (64 - sh);
    } else {
      h = mant << (sh - 64);
    }
  }
  if (bits >> 63) {
    l = ~l + 1;
    h = ~h + (l == 0 ? 1 : 0);
  }
  lo = l;
  hi = h;
}

// Magnitude of a certificate operand: x = |h| 2^(7-e) (0 <= x < 256 here) as
// an OCP FP8 E4M3 byte rounded toward zero, so it never exceeds the
// element's scaled magnitude: normal values keep the top three mantissa bits
// of the double (truncation), x < 2^-6 becomes the subnormal m 2^-9 with
// m = floor(x 2^9) (0 below 2^-9).
__device__ __forceinline__ unsigned fp8_floor(double x) {
  const uint64_t bits = (uint64_t)__double_as_longlong(x);
  const int ex = (int)((bits >> 52) & 0x7FF) - 1023;
  if (ex >= -6) return (unsigned)(((ex + 7) << 3) | (int)((bits >> 49) & 7));
  return (unsigned)(int)__builtin_ldexp(x, 9);
}

// Residue planes of A (or of B with TRANS): plane t holds the 64 x 32 chunks
// of tile row block ib (TB of them, padded to whole 256-row tiles) and K
// block kb at ((t TB + ib) KB + kb) 2048 in the slices' swizzled layout.
// The chunk (rows of A, or columns of B, by 32 k) is first staged through LDS
// with coalesced reads (lanes walk the operand's contiguous dimension).
// Each element's integer X = RN(x 256^(S-1)) (x = a 2^(6-E), |x| < 127) is
// formed exactly in 128 bits from the integer part of H = h 2^P and the rounded
// rest RN(frac(H) + l 2^P), P = 6 - E + 8(S-1). Its balanced base-256 digits
// e_j in [-128, 127] (X = sum_j e_j 256^j, j < S; |X| <= 127 256^(S-1) + 1/2
// is within their range) are the bytes of (X + K) xor K, K = 128 sum_{j<S}
// 256^j: X + K = sum_j (e_j + 128) 256^j is the plain base-256 expansion.
// These are the digits of the triangular form's slicing (the same integer,
// up to the rounding of an exact tie).
//
// The residues come from the matrix cores: V_t = sum_j e_j w_tj with
// w_tj = f_t 256^j mod m_t (balanced bytes) is an [elements x 16 digits] x
// [16 digits x moduli] INT8 product. One v_mfma_i32_32x32x32_i8 takes two
// matrix rows R0, R1 of a wave's 16 at once, in block form: lane L < 32
// supplies the 16 digits of element (R0, kappa(L)) as K bytes 0..15, lane
// L + 32 those of (R1, kappa(L)) as K bytes 16..31; B column c < 16 holds the
// weights of modulus 16 q + c in K 0..15, column c >= 16 those of modulus
// 16 q + c - 16 in K 16..31 (zeros elsewhere). With kappa(8 g + 4 h + i) =
// 16 h + 4 g + i, output lane (c, h) then holds V for row R0 + (c >= 16),
// modulus 16 q + (c & 15) and k = 16 h + j in accumulator j: one 16-byte half
// of a chunk row, stored with one write. A bias of 128 enters the sum (for
// S <= 15 as digit 2 against weight 64 in the free slot 15, for S = 16 as the
// accumulators' start value), so the lane reduces V' = V + 128 to r' = V' - m floor(V' / m) in [0, m]
// (a float quotient that lands on an integer from below gives m instead of 0:
// the same residue, still a byte), packs r' as an unsigned byte and flips the
// top bit: r' - 128 == V (mod m) as a signed byte. |V'| <= 2^18 + 128, so the
// float conversion and the fma are exact, and a quotient V'/m that is not an
// integer lies at least 1/m from one, far beyond the product's rounding
// (2^-23 |V'/m| < 2^-11).
// S, N and the weights come from the call's control block. pass 1 is the
// second residue pass of an automatic call: it runs only when the plan raised
// S (ctl->redo). With mag set (first pass of an automatic call) the kernel
// also writes the certificate operand -- |hi| 2^(7-E) as FP8 E4M3 rounded
// down, in the residues' chunk layout (32 k-elements per 32-byte row) -- and
// adds an upper bound of sum_p (|hi| + |lo|) 2^-E over its chunk,
// in units of 2^-40 rounded up, to sums[row] (integer atomics: the total does
// not depend on their order).
template <bool TRANS>
__global__ __launch_bounds__(kThreads) void k_oz2_resid(const double2 *__restrict__ X, int64_t rows,
                                                        int64_t k, const int *__restrict__ Ex, int64_t KB,
                                                        int64_t TB, const OzCtl *__restrict__ ctl, int pass,
                                                        signed char *__restrict__ out,
                                                        unsigned char *__restrict__ mag,
                                                        unsigned long long *__restrict__ sums) {
  if (pass == 1 && !ctl->redo) return;  // uniform: one control word for the grid
  constexpr int kPad = kK + 1;
  __shared__ double2 stage[kT * kPad];
  __shared__ int Es[kT];
  __shared__ unsigned long long rsum[kThreads / 64][kT];
  const int S = ctl->S, N = ctl->N;
  const Oz2Weights *__restrict__ wt = TRANS ? &ctl->wb : &ctl->wa;
  const int ib = blockIdx.x, kb = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rho = lane & 31, hl = lane >> 5, c = lane & 31;
  // This lane's B fragment and modulus in each group of 16 moduli (N <= 48),
  // loaded first so that their latency hides behind the staging below.
  constexpr int kGroups = 3;
  static_assert(16 * kGroups >= MF_OZ2_MODULI, "modulus groups");
  v4i bw[kGroups];
  float mf[kGroups], inv[kGroups];
#pragma unroll
  for (int g = 0; g < kGroups; ++g) {
    const int t = 16 * g + (c & 15);
    const bool tl = t < N;
    bw[g] = tl && (c >> 4) == hl ? v4i{wt->w[t][0], wt->w[t][1], wt->w[t][2], wt->w[t][3]} : v4i{0, 0, 0, 0};
    mf[g] = tl ? (float)mf_oz2_m[t] : 1.0f;
    inv[g] = tl ? wt->inv[t] : 1.0f;
  }
#pragma unroll
  for (int q = 0; q < kChunk / kThreads; ++q) {
    const int e = tid + q * kThreads;
    const int rr = TRANS ? e / kK : e % kT, pp = TRANS ? e % kK : e / kT;
    const int64_t gi = (int64_t)ib * kT + rr, gp = (int64_t)kb * kK + pp;
    stage[rr * kPad + pp] = (gi < rows && gp < k) ? X[TRANS ? gp + gi * k : gi + gp * rows] : make_double2(0.0, 0.0);
  }
  if (tid < kT) {
    const int64_t i = (int64_t)ib * kT + tid;
    Es[tid] = i < rows ? Ex[i] : kEmpty;
  }
  __syncthreads();
  if (mag) {
    // lane = row r, k = c0 .. c0 + 7
    const int r = lane, c0 = 8 * wave;
    const int Ei = Es[r];
    const bool live = Ei != kEmpty && Ei != kNonFinite;
    unsigned mg[2] = {0, 0};
    unsigned long long part = 0;
    if (live) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double2 v = stage[r * kPad + c0 + q];
        const double ah = fabs(v.x);
        mg[q >> 2] |= fp8_floor(__builtin_ldexp(ah, 7 - Ei)) << (8 * (q & 3));
        // (|hi| + |lo|) 2^(40-E) < 2^41; + 1 covers the rounding of the sum
        part += (unsigned long long)__builtin_ceil(__builtin_ldexp(ah + fabs(v.y), 40 - Ei)) + (ah != 0.0);
      }
    }
    // 8 FP8 bytes: k = c0 .. c0 + 7 of row r in the chunk of K block kb
    unsigned char *dm = mag + ((int64_t)ib * KB + kb) * kChunk + tile_off(r, c0);
    *(uint2 *)dm = make_uint2(mg[0], mg[1]);
    rsum[wave][r] = part;
    __syncthreads();
    if (tid < kT) {
      const unsigned long long tot = rsum[0][tid] + rsum[1][tid] + rsum[2][tid] + rsum[3][tid];
      const int64_t i = (int64_t)ib * kT + tid;
      if (tot && i < rows) atomicAdd(&sums[i], tot);
    }
  }
  // digits of this lane's 8 elements: pair p = rows 16 wave + 2 p + (lane >= 32), k = kappa(lane & 31)
  const int kap = 16 * ((rho >> 2) & 1) + 4 * (rho >> 3) + (rho & 3);
  const uint64_t k80 = 0x8080808080808080ull;
  const uint64_t K0 = S >= 8 ? k80 : k80 & ((1ull << (8 * S)) - 1);
  const uint64_t K1 = S <= 8 ? 0ull : S >= 16 ? k80 : k80 & ((1ull << (8 * (S - 8))) - 1);
  const int slot15 = S <= 15 ? 0x02000000 : 0;  // bias digit 2 (weight 64) in the free top slot
  v4i dp[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const int R = 16 * wave + 2 * p + hl;
    const int Ei = Es[R];
    const bool live = Ei != kEmpty && Ei != kNonFinite;
    const int P = live ? 6 - Ei + 8 * (S - 1) : 0;
    const double2 v = stage[R * kPad + kap];
    const double H = live ? __builtin_ldexp(v.x, P) : 0.0;
    const double L = live ? __builtin_ldexp(v.y, P) : 0.0;
    const double Hi = __builtin_trunc(H);
    const double Lr = __builtin_rint((H - Hi) + L);
    uint64_t a0, a1, b0, b1;
    int_double_to_i128(Hi, a0, a1);
    int_double_to_i128(Lr, b0, b1);
    uint64_t lo = a0 + b0, hi = a1 + b1 + (lo < a0 ? 1 : 0);
    const uint64_t lk = lo + K0;
    hi = (hi + K1 + (lk < lo ? 1 : 0)) ^ K1;
    lo = lk ^ K0;
    dp[p] = v4i{(int)(uint32_t)lo, (int)(uint32_t)(lo >> 32), (int)(uint32_t)hi, (int)(uint32_t)(hi >> 32) | slot15};
  }
  const int64_t plane = TB * KB * kChunk;
  signed char *base = out + ((int64_t)ib * KB + kb) * kChunk;
  const int G = (N + 15) >> 4;
  // CB: the bias rides in the accumulator input (S = 16 leaves no free digit
  // slot); otherwise the MFMA starts from zero.
  auto run = [&](auto cb) {
    constexpr bool CB = decltype(cb)::value;
    v16i cin;
#pragma unroll
    for (int e = 0; e < 16; ++e) cin[e] = CB ? 128 : 0;
    // opaque to the register allocator: kept in registers for every MFMA
    // rather than rebuilt each time the result lands on its registers
    asm volatile("" : "+v"(cin));
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
      if (g >= G) break;  // uniform
      const int t = 16 * g + (c & 15);
      signed char *dt = base + (int64_t)(t < N ? t : 0) * plane;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        const v16i V = __builtin_amdgcn_mfma_i32_32x32x32_i8(dp[p], bw[g], cin, 0, 0, 0);
        unsigned wd[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const float vf = (float)V[j];
          const float rr = __builtin_fmaf(-__builtin_floorf(vf * inv[g]), mf[g], vf);
          wd[j >> 2] = __builtin_amdgcn_cvt_pk_u8_f32(rr, j & 3, wd[j >> 2]);
        }
        const int Ro = 16 * wave + 2 * p + (c >> 4);
        if (t < N)
          *(v4i *)(dt + Ro * kK + ((hl ^ swz(Ro)) << 4)) =
              v4i{(int)(wd[0] ^ 0x80808080u), (int)(wd[1] ^ 0x80808080u), (int)(wd[2] ^ 0x80808080u),
                  (int)(wd[3] ^ 0x80808080u)};
      }
    }
  };
  if (S <= 15)
    run(std::integral_constant<bool, false>{});
  else
    run(std::integral_constant<bool, true>{});
}

}  // namespace

namespace mfhip_oz {

void oz2_resid_launch(bool trans, hipStream_t s, const double2 *X, int64_t rows, int64_t k, const int *Ex,
                      int64_t KB, int64_t TB, const OzCtl *ctl, int pass, signed char *out, unsigned char *mag,
                      unsigned long long *sums) {
  if (trans)
    hipLaunchKernelGGL(k_oz2_resid<true>, dim3((unsigned)TB, (unsigned)KB), dim3(kThreads), 0, s, X, rows, k, Ex, KB,
                       TB, ctl, pass, out, mag, sums);
  else
    hipLaunchKernelGGL(k_oz2_resid<false>, dim3((unsigned)TB, (unsigned)KB), dim3(kThreads), 0, s, X, rows, k, Ex,
                       KB, TB, ctl, pass, out, mag, sums);
}

}  // namespace mfhip_oz
