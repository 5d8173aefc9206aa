// Definitions shared by the Ozaki translation units (gemm_ozaki.hip: scan,
// slices, products, reconstruction; ozaki_resid.hip: the CRT form's residue
// pass, built with MFMA results in VGPRs).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dd_tables.h"

namespace mfhip_oz {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kT = 64;        // output tile edge (rows of A / cols of B per block)
constexpr int kK = 32;        // K block (one MFMA K step)
constexpr int kChunk = kT * kK;  // bytes of one slice tile (64 x 32)
constexpr int kLdsSlice = kChunk;  // LDS image of a slice tile = its global image
constexpr int kThreads = 256;  // pre-pass kernels
constexpr int kWaves = 4;       // main kernel: 4 waves, 1 per SIMD
constexpr int kMainThreads = 64 * kWaves;
constexpr int kEmpty = -2147483647 - 1;  // exponent of an all-zero row/column
constexpr int kNonFinite = 2147483647;   // a row/column holding Inf or NaN

// Byte offset of element (r, c) in a 64 x 32 slice tile. Row r is 32 bytes;
// its two 16-byte halves are swapped when bit 2 xor bit 3 of r is set, so the
// 16-byte fragment reads of 8 or 16 consecutive lanes (rows r..r+7 or
// r..r+15, same half) land on distinct LDS banks.
__device__ __forceinline__ int swz(int r) { return ((r >> 2) ^ (r >> 3)) & 1; }
__device__ __forceinline__ int tile_off(int r, int c) {
  return r * kK + ((((c >> 4) ^ swz(r))) << 4) + (c & 15);
}

constexpr int kO2Tile = 256;  // modular GEMM output tile edge (4 x 4 blocks of 64)
constexpr int kO2KS = 2;      // K blocks per stage

struct Oz2Consts {
  double mp_hi, mp_lo;  // M 2^-(16(S-1)) as a double-double
};

// Weights of the residue pass: byte j of w[t] (16 bytes) is 256^j mod m_t,
// balanced into int8, and zero for S <= j < 15, so that X mod m_t ==
// sum_j e_j w_tj for X = sum_j e_j 256^j; byte 15 is 64 when S <= 15 (the
// residue pass's bias slot). The residues do not depend on how many moduli
// the call ends up using: the CRT factor y_t = (M / m_t)^-1 mod m_t enters in
// the residue GEMM's epilogue, once the moduli count is final.
struct Oz2Weights {
  int w[MF_OZ2_MODULI][4];
  float inv[MF_OZ2_MODULI];  // RN(1 / m_t)
};

// Slice counts the certificate may ask for: kMinSlices..kMaxSlices, and
// kOverCap for an output that no count up to kMaxSlices certifies.
constexpr int kMinSlices = 7;
constexpr int kMaxSlices = 16;
constexpr int kOverCap = kMaxSlices + 1;
constexpr int kOzBins = kOverCap + 1;

// Control block of one call of the CRT form, in the workspace. The plan
// kernels write it and every later kernel reads it, so the slice count S and
// the moduli count N can be chosen on the device from the data (no host
// round trip; the call can be captured in a graph).
struct OzCtl {
  int S;        // slices of the residues in the workspace (and of the products)
  int N;        // moduli of the products and the reconstruction
  int redo;     // 1: the certificate raised S, the residues are formed again
  int S_first;  // slices of the first residue pass (0: explicit slice count)
  unsigned long long hist[kOzBins];  // candidate outputs by the slices they need
  unsigned long long ncand;          // outputs in the candidate list
  unsigned long long rmax, cmax;     // largest certificate sum over live rows of A / columns of B
  unsigned long long nfix;           // outputs the fix-up recomputed
  Oz2Consts cs;
  Oz2Weights wa, wb;  // residue weights of A (f_t = 1) and of B (f_t = y_t)
};

// Residue pass of the CRT form over one operand (ozaki_resid.hip): grid
// (TB, KB) workgroups of kThreads; TRANS for B.
void oz2_resid_launch(bool trans, hipStream_t s, const double2 *X, int64_t rows, int64_t k, const int *Ex,
                      int64_t KB, int64_t TB, const OzCtl *ctl, int pass, signed char *out, unsigned char *mag,
                      unsigned long long *sums);

}  // namespace mfhip_oz
