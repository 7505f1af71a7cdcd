// Fused-part path: dense 2^k x 2^k unitary applied on the gfx950 matrix cores.
//
// The state viewed through a part with qubit set Q (k = |Q|) is a batch of
// 2^(n-k) vectors of length D = 2^k; the fused unitary U maps each of them, so
// the pass is the complex GEMM  Y = U X  with X = D x 2^(n-k). A workgroup
// owns a D x C tile of X whose C columns run along the lowest slots outside Q
// (contiguous in HBM), stages it in LDS, and each of its D/16 waves computes
// 16 rows of Y with 16x16x4 MFMAs, its slice of U held in registers for the
// whole (persistent) launch.
//
// MFMA 16x16x4 operand layouts (checked on gfx950 with a lane probe): lane l
// holds A[l%16][l/16] and B[l/16][l%16]; result item i of lane l is
// D[4*(l/16) + i][l%16] for the f32 instruction but D[l/16 + 4*i][l%16] for
// the f64 one.
//
//  * HSV_FUSED_F64: v_mfma_f64_16x16x4_f64, FP64 products and accumulation.
//  * HSV_FUSED_F32SPLIT: v_mfma_f32_16x16x4_f32 on hi/lo FP32 splits of U and
//    X (x = hi + lo, hi = fp32(x), lo = fp32(x - hi)); each real product is
//    hi*hi + hi*lo + lo*hi, accumulated in FP32; the result is stored back in
//    the state's precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/hsv.h"
#include "fused.hpp"
#include "kernels.hpp"
#include "partition.hpp"

namespace hsv {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

struct FusedArgs {
  uint64_t ntiles;
  uint64_t outer_mask;
  int C;       // tile columns = 16 * NC
  int logC;
  int NC;
};

constexpr int kFusedMaxC = 128;

template <typename R, int K, bool SPLIT>
__global__ __launch_bounds__(4 * (1 << K)) void k_fused(typename VecT<R>::type* __restrict__ a,
                                                        const double2* __restrict__ U,
                                                        const uint64_t* __restrict__ qoff,
                                                        const uint64_t* __restrict__ toff, FusedArgs f) {
  constexpr int D = 1 << K;
  constexpr int NT = 4 * D;  // D/16 waves of 64 lanes
  constexpr int KS = D / 4;  // k-steps of the 16x16x4 MFMA
  using V = typename VecT<R>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* X = reinterpret_cast<double2*>(smem_raw);  // D x C tile, row-major
  __shared__ uint64_t sq[D];
  __shared__ uint64_t st[kFusedMaxC];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int C = f.C;
  for (int i = threadIdx.x; i < D; i += NT) sq[i] = qoff[i];
  for (int i = threadIdx.x; i < C; i += NT) st[i] = toff[i];

  // A-operand fragments: lane holds U[16*wave + lane%16][4*ks + lane/16].
  const int arow = 16 * wave + (lane & 15);
  const int akk = lane >> 4;
  double ar[SPLIT ? 1 : KS], ai[SPLIT ? 1 : KS];
  float arh[SPLIT ? KS : 1], arl[SPLIT ? KS : 1], aih[SPLIT ? KS : 1], ail[SPLIT ? KS : 1];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const double2 u = U[arow * D + 4 * ks + akk];
    if constexpr (SPLIT) {
      arh[ks] = (float)u.x; arl[ks] = (float)(u.x - (double)arh[ks]);
      aih[ks] = (float)u.y; ail[ks] = (float)(u.y - (double)aih[ks]);
    } else {
      ar[ks] = u.x; ai[ks] = u.y;
    }
  }
  const int ccol = lane & 15;  // B column / D column within a 16-wide block
  const int krow = lane >> 4;  // B row within a k-step; D row group
  for (uint64_t t = blockIdx.x; t < f.ntiles; t += gridDim.x) {
    const uint64_t base = pdep64(t, f.outer_mask);
    __syncthreads();  // the previous tile's LDS reads are done
    for (int e = threadIdx.x; e < D * C; e += NT) {
      const int r = e >> f.logC, c = e & (C - 1);
      const V v = a[base + sq[r] + st[c]];
      X[e] = make_double2((double)v.x, (double)v.y);
    }
    __syncthreads();
    for (int cb = 0; cb < f.NC; ++cb) {
      if constexpr (SPLIT) {
        v4f accR = {0, 0, 0, 0}, accT = {0, 0, 0, 0}, accI = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const double2 x = X[(4 * ks + krow) * C + 16 * cb + ccol];
          const float xrh = (float)x.x, xrl = (float)(x.x - (double)xrh);
          const float xih = (float)x.y, xil = (float)(x.y - (double)xih);
          // Re: Ur*Xr - Ui*Xi ; Im: Ur*Xi + Ui*Xr (three split products each)
          accR = __builtin_amdgcn_mfma_f32_16x16x4f32(arh[ks], xrh, accR, 0, 0, 0);
          accR = __builtin_amdgcn_mfma_f32_16x16x4f32(arh[ks], xrl, accR, 0, 0, 0);
          accR = __builtin_amdgcn_mfma_f32_16x16x4f32(arl[ks], xrh, accR, 0, 0, 0);
          accT = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[ks], xih, accT, 0, 0, 0);
          accT = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[ks], xil, accT, 0, 0, 0);
          accT = __builtin_amdgcn_mfma_f32_16x16x4f32(ail[ks], xih, accT, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f32_16x16x4f32(arh[ks], xih, accI, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f32_16x16x4f32(arh[ks], xil, accI, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f32_16x16x4f32(arl[ks], xih, accI, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[ks], xrh, accI, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f32_16x16x4f32(aih[ks], xrl, accI, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f32_16x16x4f32(ail[ks], xrh, accI, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * wave + 4 * krow + i;
          const int col = 16 * cb + ccol;
          V o;
          o.x = (R)((double)accR[i] - (double)accT[i]);
          o.y = (R)accI[i];
          a[base + sq[row] + st[col]] = o;
        }
      } else {
        v4d accR = {0, 0, 0, 0}, accT = {0, 0, 0, 0}, accI = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const double2 x = X[(4 * ks + krow) * C + 16 * cb + ccol];
          accR = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[ks], x.x, accR, 0, 0, 0);
          accT = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[ks], x.y, accT, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[ks], x.y, accI, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[ks], x.x, accI, 0, 0, 0);
        }
        // f64 result layout differs from f32: item i of lane l is row l/16 + 4i
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * wave + krow + 4 * i;
          const int col = 16 * cb + ccol;
          V o;
          o.x = (R)(accR[i] - accT[i]);
          o.y = (R)accI[i];
          a[base + sq[row] + st[col]] = o;
        }
      }
    }
  }
}

// ---- BF16-split fused path ------------------------------------------------------
// gfx950's dense BF16 rate is 16x its FP32 rate, so here U and X are split
// into BF16 pairs (v = hi + lo, hi = bf16(v), lo = bf16(v - hi): 16
// significant bits) and every real product is hi*hi + hi*lo + lo*hi on
// v_mfma_f32_16x16x32_bf16, accumulated in FP32. U's fragments live in LDS
// for the whole (persistent) launch; each wave owns column blocks of 16
// columns (16 amplitudes along the lowest slots outside the part) and
// streams their D rows from HBM straight into B operands, 32 rows per
// k-step, the next k-step's loads in flight while the current one runs.
//
// 16x16x32 operand layouts (checked on gfx950 with a lane probe): lane l holds
// A[l%16][8*(l/16) + e] and B[8*(l/16) + e][l%16] for e = 0..7 (any k order
// shared by A and B gives the same product), and result item i of lane l is
// D[4*(l/16) + i][l%16].
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));

// Two floats to their BF16 hi parts (round to nearest even, packed in one
// dword as element pair 2j, 2j+1 of a bf8 operand) and the BF16 of the
// remainders: a = hi_a + lo_a to 16 significant bits. One v_cvt_pk_bf16_f32
// per pair for each of hi and lo; hi back to FP32 is a shift or a mask.
__device__ __forceinline__ void split_pair(float a, float b, uint32_t* hi, uint32_t* lo) {
  const f2v ab = {a, b};
  const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(ab, bf2v));
  const f2v rest = {a - __builtin_bit_cast(float, h << 16), b - __builtin_bit_cast(float, h & 0xffff0000u)};
  *hi = h;
  *lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(rest, bf2v));
}

// The three-piece split (Henry, Tang and Heinecke, "Leveraging the bfloat16
// artificial intelligence datatype for higher-precision computations",
// ARITH 2019): a = hi + mid + lo with each piece the BF16 of what the
// previous ones leave (both remainders are exact in FP32), 24 significant
// bits in all; the residual is below 2^-26 |a|.
__device__ __forceinline__ void split_pair3(float a, float b, uint32_t* hi, uint32_t* mid, uint32_t* lo) {
  const f2v ab = {a, b};
  const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(ab, bf2v));
  const f2v r1 = {a - __builtin_bit_cast(float, h << 16), b - __builtin_bit_cast(float, h & 0xffff0000u)};
  const uint32_t m = __builtin_bit_cast(uint32_t, __builtin_convertvector(r1, bf2v));
  const f2v r2 = {r1[0] - __builtin_bit_cast(float, m << 16), r1[1] - __builtin_bit_cast(float, m & 0xffff0000u)};
  *hi = h;
  *mid = m;
  *lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r2, bf2v));
}

// 8 complex values (rows of a B operand) to the four BF16 operands: real hi,
// real lo, imaginary hi, imaginary lo.
__device__ __forceinline__ void split_bf8(const float2 (&v)[8], bf8* rh, bf8* rl, bf8* ih, bf8* il) {
  u4v a, b, c, d;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint32_t x0, x1, y0, y1;
    split_pair(v[2 * j].x, v[2 * j + 1].x, &x0, &x1);
    split_pair(v[2 * j].y, v[2 * j + 1].y, &y0, &y1);
    a[j] = x0;
    b[j] = x1;
    c[j] = y0;
    d[j] = y1;
  }
  *rh = __builtin_bit_cast(bf8, a);
  *rl = __builtin_bit_cast(bf8, b);
  *ih = __builtin_bit_cast(bf8, c);
  *il = __builtin_bit_cast(bf8, d);
}

// The same with three pieces: real hi, mid, lo, imaginary hi, mid, lo.
__device__ __forceinline__ void split_bf8_3(const float2 (&v)[8], bf8* rh, bf8* rm, bf8* rl, bf8* ih, bf8* im,
                                            bf8* il) {
  u4v a, b, c, d, e, f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint32_t x0, x1, x2, y0, y1, y2;
    split_pair3(v[2 * j].x, v[2 * j + 1].x, &x0, &x1, &x2);
    split_pair3(v[2 * j].y, v[2 * j + 1].y, &y0, &y1, &y2);
    a[j] = x0;
    b[j] = x1;
    c[j] = x2;
    d[j] = y0;
    e[j] = y1;
    f[j] = y2;
  }
  *rh = __builtin_bit_cast(bf8, a);
  *rm = __builtin_bit_cast(bf8, b);
  *rl = __builtin_bit_cast(bf8, c);
  *ih = __builtin_bit_cast(bf8, d);
  *im = __builtin_bit_cast(bf8, e);
  *il = __builtin_bit_cast(bf8, f);
}

// Column signs of the three-piece split. column_hash: murmur3's 32-bit
// finaliser of (tile << 8 | column) (columns < 256), once per tile;
// column_sign: 0 or the FP32 sign bit for one step, the top bit of
// (hash ^ step * 0x85EBCA77) * 0x9E3779B1.
__device__ __forceinline__ uint32_t column_hash(uint64_t tile, int col) {
  uint32_t k = uint32_t(tile) << 8 | uint32_t(col);
  k ^= k >> 16;
  k *= 0x85EBCA6Bu;
  k ^= k >> 13;
  k *= 0xC2B2AE35u;
  k ^= k >> 16;
  return k;
}
__device__ __forceinline__ uint32_t column_sign(uint32_t hash, int step) {
  return ((hash ^ (uint32_t(step) * 0x85EBCA77u)) * 0x9E3779B1u) & 0x80000000u;
}

struct BfArgs {
  uint64_t nblocks;     // column blocks of 16 columns
  uint64_t outer_mask;  // outer bits above the 4 column bits
};

constexpr int kBfThreads = 512;

// SWZ: row order of the B operands and of the results. false: lane l, element
// e reads row 8*(l/16) + e of its k-step and stores result item i to row
// 4*(l/16) + i. true: rows (l/16) + 4*e and 4*i + (l/16), so the 4 lanes of
// a column cover 4 consecutive rows in each load and store instruction; the
// host permutes U's columns and rows to match (any k order shared by A and
// B, and any row order of A, leave the product the same up to that order).
template <typename R, int K, bool SWZ>
__global__ __launch_bounds__(kBfThreads) void k_fused_bf(typename VecT<R>::type* __restrict__ a,
                                                       const bf8* __restrict__ ufrag,
                                                       const uint64_t* __restrict__ qoff,
                                                       const uint64_t* __restrict__ toff, BfArgs f) {
  constexpr int D = 1 << K;
  constexpr int RB = D / 16;  // row blocks of the 16x16 result
  constexpr int KS = D / 32;  // k-steps of 32
  using V = typename VecT<R>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  bf8* su = reinterpret_cast<bf8*>(smem_raw);  // [rb][ks][set][lane], set = Ur_h, Ur_l, Ui_h, Ui_l
  __shared__ uint64_t sq[D];
  for (int i = threadIdx.x; i < RB * KS * 4 * 64; i += kBfThreads) su[i] = ufrag[i];
  for (int i = threadIdx.x; i < D; i += kBfThreads) sq[i] = qoff[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kgrp = lane >> 4;
  const uint64_t coff = toff[lane & 15];
  const uint64_t nw = uint64_t(gridDim.x) * (kBfThreads / 64);
  typedef float v4f_t __attribute__((ext_vector_type(4)));
  // raw: this lane's 8 amplitudes of each k-step (rows 32 ks + 8 kgrp + e,
  // column lane % 16); the next block's are loaded while this one's MFMAs run
  V raw[KS][8];
  uint64_t cb = uint64_t(blockIdx.x) * (kBfThreads / 64) + wave;
  if (cb < f.nblocks) {
    const uint64_t base = pdep64(cb, f.outer_mask) + coff;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int e = 0; e < 8; ++e) raw[ks][e] = a[base + sq[32 * ks + (SWZ ? kgrp + 4 * e : 8 * kgrp + e)]];
  }
  for (; cb < f.nblocks; cb += nw) {
    const uint64_t base = pdep64(cb, f.outer_mask) + coff;
    bf8 xrh[KS], xrl[KS], xih[KS], xil[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float xr = (float)raw[ks][e].x, xi = (float)raw[ks][e].y;
        const __bf16 hr = (__bf16)xr, hi = (__bf16)xi;
        xrh[ks][e] = hr;
        xrl[ks][e] = (__bf16)(xr - (float)hr);
        xih[ks][e] = hi;
        xil[ks][e] = (__bf16)(xi - (float)hi);
      }
    const uint64_t cn = cb + nw;
    if (cn < f.nblocks) {
      const uint64_t nbase = pdep64(cn, f.outer_mask) + coff;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int e = 0; e < 8; ++e) raw[ks][e] = a[nbase + sq[32 * ks + (SWZ ? kgrp + 4 * e : 8 * kgrp + e)]];
    }
    // every row of this column block is in registers now: results can be
    // written as each row block finishes
#pragma unroll 1
    for (int rb = 0; rb < RB; ++rb) {
      v4f_t accR = {0, 0, 0, 0}, accT = {0, 0, 0, 0}, accI = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const bf8* A = su + ((rb * KS + ks) * 4) * 64 + lane;
        const bf8 urh = A[0], url = A[64], uih = A[128], uil = A[192];
        // Re = Ur Xr - Ui Xi ; Im = Ur Xi + Ui Xr, three split products each
        accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xrh[ks], accR, 0, 0, 0);
        accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xrl[ks], accR, 0, 0, 0);
        accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(url, xrh[ks], accR, 0, 0, 0);
        accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xih[ks], accT, 0, 0, 0);
        accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xil[ks], accT, 0, 0, 0);
        accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uil, xih[ks], accT, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xih[ks], accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xil[ks], accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(url, xih[ks], accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xrh[ks], accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xrl[ks], accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uil, xrh[ks], accI, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        V o;
        o.x = (R)(accR[i] - accT[i]);
        o.y = (R)accI[i];
        a[base + sq[16 * rb + (SWZ ? 4 * i + kgrp : 4 * kgrp + i)]] = o;
      }
    }
  }
}

// ---- BF16-split fused path, U-stationary tile form ---------------------------------
// A workgroup (8 waves) owns a tile of 4096 amplitudes: the part's D rows by
// C = 4096 / D columns along the lowest slots outside the part. Tile element
// t (12 bits) maps to the tile's 12 slots in ascending order, so consecutive
// threads load and store consecutive amplitudes. Loaded values are split to
// BF16 and scattered into B-operand fragments in LDS; wave w keeps row block
// w % RB of U in registers (its A fragments) and computes that row block for
// its share of the columns; results go through an FP32 staging tile in LDS
// and are stored back coalesced. LDS: 32 KB of fragments + 32 KB of staging.
struct BfTileTables {
  uint64_t alo[64], ahi[64];    // address offset of tile index bits 0-5 / 6-11
  uint16_t rlo[64], rhi[64];    // row index contribution of those bits
  uint16_t clo[64], chi[64];    // column index contribution
  uint16_t tr[128], tc[128];    // tile index of row r / column c (OR-ed)
};

constexpr int kBfTileThreads = 512;

template <typename R, int K>
__global__ __launch_bounds__(kBfTileThreads) void k_fused_bft(typename VecT<R>::type* __restrict__ a,
                                                            const bf8* __restrict__ ufrag,
                                                            const BfTileTables* __restrict__ tabs,
                                                            uint64_t ntiles, uint64_t outer_mask, uint64_t top,
                                                            int tiles_per_wg) {
  constexpr int D = 1 << K;
  constexpr int RB = D / 16, KS = D / 32;
  constexpr int C = 4096 / D, NCB = C / 16;
  constexpr int W = kBfTileThreads / 64;
  constexpr int CG = W / RB;        // waves sharing one row block
  constexpr int CBW = NCB / CG;     // column blocks per wave
  constexpr int PER = 4096 / kBfTileThreads;
  using V = typename VecT<R>::type;
  __shared__ bf8 xs[KS * NCB * 4 * 64];   // [ks][cb][set][lane], set = Xr_h, Xr_l, Xi_h, Xi_l
  __shared__ float2 ys[4096];             // results in tile order
  __shared__ BfTileTables tb;
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(tabs);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&tb);
    for (int i = threadIdx.x; i < (int)(sizeof(BfTileTables) / 4); i += kBfTileThreads) dst[i] = src[i];
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rb = wave % RB, cg = wave / RB;
  bf8 A[KS][4];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int s = 0; s < 4; ++s) A[ks][s] = ufrag[((rb * KS + ks) * 4 + s) * 64 + lane];
  __syncthreads();
  __bf16* xb = reinterpret_cast<__bf16*>(xs);
  typedef float v4f_t __attribute__((ext_vector_type(4)));
  // consecutive tiles per workgroup, no prefetch (the other workgroups on
  // the CU overlap this one's compute); the U fragments above are loaded
  // once per workgroup
  for (int it = 0; it < tiles_per_wg; ++it) {
    const uint64_t tile = uint64_t(blockIdx.x) * tiles_per_wg + it;
    if (tile >= ntiles) break;  // uniform across the workgroup
    const uint64_t base = tile_base(tile, outer_mask, top);
    V raw[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int t = threadIdx.x + kBfTileThreads * j;
      raw[j] = ld_stream<R>(&a[base + tb.alo[t & 63] + tb.ahi[t >> 6]]);
    }
    __syncthreads();  // the previous tile's stores have read ys (and the xs reads are done)
    // split and scatter this thread's elements into the B fragments:
    // row r = 32 ks + g + 4 e goes to lane 16 g + (c % 16), element e
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int t = threadIdx.x + kBfTileThreads * j;
      const int r = tb.rlo[t & 63] + tb.rhi[t >> 6];
      const int c = tb.clo[t & 63] + tb.chi[t >> 6];
      const int ks = r >> 5, g = r & 3, e = (r >> 2) & 7;
      const int frag = (ks * NCB + (c >> 4)) * 4 * 64 + 16 * g + (c & 15);
      const float xr = (float)raw[j].x, xi = (float)raw[j].y;
      const __bf16 hr = (__bf16)xr, hi = (__bf16)xi;
      xb[(frag + 0 * 64) * 8 + e] = hr;
      xb[(frag + 1 * 64) * 8 + e] = (__bf16)(xr - (float)hr);
      xb[(frag + 2 * 64) * 8 + e] = hi;
      xb[(frag + 3 * 64) * 8 + e] = (__bf16)(xi - (float)hi);
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < CBW; ++w) {
      const int cb = cg * CBW + w;
      v4f_t accR = {0, 0, 0, 0}, accT = {0, 0, 0, 0}, accI = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const bf8* B = xs + (ks * NCB + cb) * 4 * 64 + lane;
        const bf8 xrh = B[0], xrl = B[64], xih = B[128], xil = B[192];
        accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][0], xrh, accR, 0, 0, 0);
        accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][0], xrl, accR, 0, 0, 0);
        accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][1], xrh, accR, 0, 0, 0);
        accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][2], xih, accT, 0, 0, 0);
        accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][2], xil, accT, 0, 0, 0);
        accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][3], xih, accT, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][0], xih, accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][0], xil, accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][1], xih, accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][2], xrh, accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][2], xrl, accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][3], xrh, accI, 0, 0, 0);
      }
      // item i of lane l: row 16 rb + 4 i + l/16 (the swizzled A rows), column 16 cb + l % 16
      const int col = 16 * cb + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * rb + 4 * i + (lane >> 4);
        ys[tb.tr[row] | tb.tc[col]] = make_float2(accR[i] - accT[i], accI[i]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int t = threadIdx.x + kBfTileThreads * j;
      const float2 y = ys[t];
      V o;
      o.x = (R)y.x;
      o.y = (R)y.y;
      st_stream<R>(&a[base + tb.alo[t & 63] + tb.ahi[t >> 6]], o);
    }
  }
}

// ---- BF16-split wide parts: 12- to 14-qubit tiles, 5-qubit fused steps ------------
// For a part of 8..14 qubits the host splits the gates into steps of at most
// 5 qubits (the qubit-set closure method), fuses each step into a 32 x 32
// FP64 unitary and sends its BF16 A fragments. A workgroup owns one tile of
// 2^KT amplitudes (the part's qubits, padded with the lowest free ones, in
// ascending order: tile index bit j <-> the j-th of them): it loads the tile
// coalesced into LDS as FP32, applies the steps there one after another
// (rows = the step's 5 tile bits, columns = the other KT - 5; each wave owns
// column blocks of 16 and updates them in place; a barrier between steps),
// and stores the tile back coalesced. One HBM pass per part, the gate work
// on the matrix cores.
//
// LDS bank swizzle: tile element t lives at ys[t ^ h(t >> 4)], h linear over
// GF(2) (swz[] tabulates it; tile bit 4 feeds only bits 0-3, so the map is
// invertible), chosen on the host per part so that the 16
// lanes of a quarter-wave (16 columns of one row) and the two rows of a
// half-wave fall in distinct banks for every step. Because the map is
// linear, a step's row and column offsets are tabulated already swizzled and
// combined with XOR. Without it, a step on the low tile bits reads its
// columns at a stride of 32+ elements: 16-way bank conflicts.
struct PartBfStep {
  uint16_t rt[32];              // swizzled LDS offset of step row r
  uint16_t ctl[16], cth[32];    // of step column c: ctl[c & 15] ^ cth[c >> 4]
};
struct PartBfTables {
  uint64_t alo[64], ahi[256];   // address offset of tile index bits 0-5 / 6-13
  // A layout change moves qubits between the tile's slots: the element
  // stored at tile position u is the one loaded at position plo[u & 63] |
  // phi[u >> 6] (the identity without a change), so the stores stay as
  // coalesced as the loads.
  uint16_t plo[64], phi[256];
  uint8_t swz[1024];            // h(t >> 4): XOR into the low 5 bits of t
};
// 512 threads: 8 amplitudes per thread for 12-qubit tiles, 16 for 13; 1024
// threads (16 amplitudes each) for 14. 12-qubit tiles keep every step's U
// fragments in LDS; 13-qubit tiles (64 KB of FP32 in LDS) either stream them
// (STREAM_U) through one 8 KB buffer, each of 512 threads holding its
// 16-byte share of the next step in a register, so two workgroups fit a CU
// and one computes while the other loads or stores; or keep them resident
// with 1024 threads (8 amplitudes each) and one workgroup per CU, which
// then loads the next tile into registers during the steps, as 12-qubit
// tiles do. A 14-qubit tile (128 KB) fills the CU alone and streams, so its
// compute does not overlap its I/O (66 ms against 54 ms per pass at 33
// qubits): it saves passes only where parts get much wider.
constexpr int part_bf_threads(int kt, bool stream_u) { return kt >= 14 || (kt == 13 && !stream_u) ? 1024 : 512; }
constexpr int kPartBfMaxSteps = 12;
// steps per launch: 8 for the streamed three-piece 13-qubit form (its 12 KB
// fragment buffer leaves room for 8 step tables, two workgroups per CU)
constexpr int part_bf_max_steps(int kt, int np) { return np == 3 && kt == 13 ? 8 : kPartBfMaxSteps; }

// NP: BF16 pieces per value, 2 (hi*hi + hi*lo + lo*hi, 16 significant bits)
// or 3 (the three-piece split: hi*hi, hi*mid, mid*hi, mid*mid, hi*lo, lo*hi,
// 24 bits, FP32-class products; 12-qubit tiles with resident fragments, or
// 13-qubit tiles streaming 12 KB of fragments per step, with at most 8 steps
// per launch so that two workgroups still fit a CU).
template <typename R, int KT, bool STREAM_U, bool U_DIRECT = false, int NP = 2>
// 16 waves per CU (two workgroups of 8, or one of 16): at most 128 VGPRs
__global__ __launch_bounds__(part_bf_threads(KT, STREAM_U)) __attribute__((amdgpu_waves_per_eu(4))) void k_part_bf(typename VecT<R>::type* __restrict__ a,
                                                          const PartBfTables* __restrict__ tabs,
                                                          const PartBfStep* __restrict__ steps_g,
                                                          const bf8* __restrict__ ufrag, int nsteps,
                                                          uint64_t ntiles, uint64_t outer_mask, uint64_t top,
                                                          int tiles_per_wg) {
  constexpr int T = 1 << KT;
  constexpr int NCB = (1 << (KT - 5)) / 16;   // column blocks of a step
  constexpr int NTH = part_bf_threads(KT, STREAM_U);
  constexpr int W = NTH / 64;
  constexpr int CBW = NCB / W;                // column blocks per wave
  constexpr int PER = T / NTH;
  static_assert(KT == 13 || STREAM_U == (KT >= 14), "12-qubit tiles keep U resident, 14-qubit tiles stream it");
  constexpr int STEP_FRAGS = 2 * 2 * NP * 64;  // bf8 per step: [rb][set][lane]
  static_assert(NP == 2 || (NP == 3 && ((KT == 12 && !STREAM_U) || (KT == 13 && STREAM_U && !U_DIRECT))),
                "three pieces: 12-qubit tiles with resident fragments or streamed 13-qubit tiles");
  constexpr int FRPT = (STEP_FRAGS + NTH - 1) / NTH;  // streamed fragments held per thread
  constexpr int MAX_STEPS = part_bf_max_steps(KT, NP);
  static_assert(CBW >= 1 && NCB % W == 0 && PER * NTH == T, "tile shape");
  static_assert(!STREAM_U || STEP_FRAGS <= 2 * NTH, "at most two fragments per thread per step");
  static_assert(!U_DIRECT || STREAM_U, "direct fragment loads replace the streaming buffer");
  constexpr bool SU_BUF = STREAM_U && !U_DIRECT;
  using V = typename VecT<R>::type;
  typedef float v4f_t __attribute__((ext_vector_type(4)));
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ bf8 su_one[SU_BUF ? STEP_FRAGS : 1];
  bf8* su = STREAM_U ? su_one : reinterpret_cast<bf8*>(smem_raw);     // [step][rb][set][lane]
  __shared__ float2 ys[T];
  __shared__ PartBfStep st[MAX_STEPS];
  __shared__ uint64_t alo[64], ahi[T >> 6];
  __shared__ uint16_t plo[64], phi[T >> 6];
  __shared__ uint8_t swz[T >> 4];
  if constexpr (!STREAM_U)
    for (int i = threadIdx.x; i < nsteps * STEP_FRAGS; i += NTH) su[i] = ufrag[i];
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(steps_g);
    uint32_t* dst = reinterpret_cast<uint32_t*>(st);
    for (int i = threadIdx.x; i < nsteps * (int)(sizeof(PartBfStep) / 4); i += NTH) dst[i] = src[i];
  }
  for (int i = threadIdx.x; i < 64; i += NTH) alo[i] = tabs->alo[i];
  for (int i = threadIdx.x; i < (T >> 6); i += NTH) ahi[i] = tabs->ahi[i];
  for (int i = threadIdx.x; i < 64; i += NTH) plo[i] = tabs->plo[i];
  for (int i = threadIdx.x; i < (T >> 6); i += NTH) phi[i] = tabs->phi[i];
  for (int i = threadIdx.x; i < (T >> 4); i += NTH) swz[i] = tabs->swz[i];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, cc = lane & 15;
  // A few consecutive tiles per workgroup (tiles_per_wg; the grid covers
  // the rest). With the fragments resident, the next tile's loads are issued
  // before this tile's steps and overlap them. (With streamed fragments the
  // per-step fragment loads would wait behind such a prefetch, so there the
  // other workgroup on the CU provides the overlap.)
  const uint64_t t0 = uint64_t(blockIdx.x) * tiles_per_wg;
  if (t0 >= ntiles) return;  // uniform across the workgroup
  const int nt = (int)min(uint64_t(tiles_per_wg), ntiles - t0);
  bf8 fr[FRPT];  // streamed fragments: this thread's share of the next step
  if constexpr (SU_BUF) {
#pragma unroll
    for (int i = 0; i < FRPT; ++i)
      if (threadIdx.x + NTH * i < STEP_FRAGS) fr[i] = ufrag[threadIdx.x + NTH * i];
  }
  __syncthreads();  // tables ready
  // Thread x handles tile positions t = x + NTH j: t & 63 is its lane and
  // t >> 6 = wave + W j, so the per-j table entries sit at fixed offsets
  // from one base (no per-j address registers).
  const uint64_t a_lo = alo[lane];
  const int p_lo = plo[lane];
  V raw[PER];
  if constexpr (!STREAM_U) {
    const uint64_t base = tile_base(t0, outer_mask, top);
#pragma unroll
    for (int j = 0; j < PER; ++j) raw[j] = ld_stream<R>(&a[base + a_lo + ahi[wave + W * j]]);
  }
  for (int it = 0; it < nt; ++it) {
    const uint64_t base = tile_base(t0 + it, outer_mask, top);
    if constexpr (STREAM_U) {
#pragma unroll
      for (int j = 0; j < PER; ++j) raw[j] = ld_stream<R>(&a[base + a_lo + ahi[wave + W * j]]);
    }
    if (it > 0) __syncthreads();  // the previous tile's stores have read ys
    // t ^ swz[t >> 4] with swz < 32 and NTH j a multiple of 512. The swizzle
    // is linear and t >> 4 = (x >> 4) | (NTH / 16) j with disjoint bits, so
    // swz[t >> 4] = swz[x >> 4] ^ swz[(NTH / 16) j]: one per-thread table
    // read, the rest wave-uniform (scalar loads of the global copy, so no
    // per-j address registers stay live across the tile loop)
    {
      const uint32_t sw_x = swz[threadIdx.x >> 4];
      // swz[(NTH / 16) 2^b], b < log2(PER), read as aligned words
      const uint32_t* sw_words = reinterpret_cast<const uint32_t*>(tabs->swz);
      uint32_t hb[4];
#pragma unroll
      for (int b = 0; b < 4; ++b)
        hb[b] = (1 << b) < PER ? __builtin_amdgcn_readfirstlane(sw_words[((NTH >> 4) << b) >> 2] & 0xffu) : 0u;
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        uint32_t hj = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if (j >> b & 1) hj ^= hb[b];
        ys[(threadIdx.x ^ sw_x ^ hj) + NTH * j] = make_float2((float)raw[j].x, (float)raw[j].y);
      }
    }
    if (!STREAM_U && it + 1 < nt) {
      const uint64_t nbase = tile_base(t0 + it + 1, outer_mask, top);
#pragma unroll
      for (int j = 0; j < PER; ++j) raw[j] = ld_stream<R>(&a[nbase + a_lo + ahi[wave + W * j]]);
    }
    // three pieces: each of this lane's columns keeps its hash for the tile
    uint32_t colh[NP == 3 ? CBW : 1];
    if constexpr (NP == 3) {
#pragma unroll
      for (int w = 0; w < CBW; ++w) colh[w] = column_hash(t0 + it, 16 * (wave * CBW + w) + cc);
    }
    for (int s = 0; s < nsteps; ++s) {
      const bf8* As;
      bf8 Af[2][4];
      if constexpr (U_DIRECT) {
        // each wave reads the step's fragments straight from L2 into
        // registers, issued before the barrier so that their latency
        // overlaps it
        As = ufrag + s * STEP_FRAGS + lane;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
          for (int q = 0; q < 4; ++q) Af[rb][q] = As[(rb * 4 + q) * 64];
      }
      __syncthreads();  // the tile (or the previous step's results) is in LDS
      const PartBfStep& S = st[s];
      if constexpr (U_DIRECT) {
      } else if constexpr (STREAM_U) {
#pragma unroll
        for (int i = 0; i < FRPT; ++i) {
          const int f = threadIdx.x + NTH * i;
          if (f < STEP_FRAGS) {
            su_one[f] = fr[i];  // every wave is past the previous step's reads
            fr[i] = ufrag[(s + 1 < nsteps ? s + 1 : 0) * STEP_FRAGS + f];
          }
        }
        __syncthreads();
        As = su_one + lane;
      } else {
        As = su + s * STEP_FRAGS + lane;
      }
      // This lane reads rows g + 4e (e = 0..7) of its columns and writes
      // result item i of row block rb to row 16 rb + 4 i + g: the same 8
      // rows, so one set of row offsets serves both. All of the wave's column
      // blocks are read and split before any MFMA and written after them
      // (each wave owns its columns for the whole step), so the blocks'
      // product chains interleave instead of running one block at a time.
      // (two 16-bit offsets per register)
      uint32_t roff2[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) roff2[e] = uint32_t(S.rt[g + 8 * e]) | (uint32_t(S.rt[g + 8 * e + 4]) << 16);
      auto roff = [&](int e) -> uint32_t { return e & 1 ? roff2[e >> 1] >> 16 : roff2[e >> 1] & 0xffffu; };
      uint32_t tc[CBW];
#pragma unroll
      for (int w = 0; w < CBW; ++w) {
        const int col = 16 * (wave * CBW + w) + cc;
        tc[w] = S.ctl[col & 15] ^ S.cth[col >> 4];
      }
      if constexpr (NP == 3) {
        // three pieces: the six products of each real product, smallest
        // first into the FP32 accumulator; one column block at a time (each
        // wave owns its columns, so a block is read, multiplied and written
        // before the next without a hazard), which keeps the eighteen BF16
        // operands of a block and its accumulators within 128 VGPRs.
        // The matrix cores' FP32 accumulation rounds with a small bias of
        // one direction: a pass's error has a component that is the same in
        // every tile and for every input, which later Hadamards (a QFT's)
        // add up onto a few amplitudes. Each column's 32 values are
        // negated before the products and the results negated back, by a
        // sign hashed from (tile, column, step): exact, and bit-identical
        // under unbiased rounding, while the bias of each column then
        // enters with a pseudo-random sign and no longer adds up.
#pragma unroll
        for (int w = 0; w < CBW; ++w) {
          const uint32_t fl = column_sign(colh[w], s);
          bf8 xrh, xrm, xrl, xih, xim, xil;
          {
            float2 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              const float2 y = ys[roff(e) ^ tc[w]];
              v[e] = make_float2(flip_sign(y.x, fl), flip_sign(y.y, fl));
            }
            split_bf8_3(v, &xrh, &xrm, &xrl, &xih, &xim, &xil);
          }
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            const bf8* Ar = As + rb * 6 * 64;
            const bf8 urh = Ar[0], urm = Ar[64], url = Ar[128], uih = Ar[192], uim = Ar[256], uil = Ar[320];
            const v4f_t z = {0, 0, 0, 0};
            v4f_t accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xrl, z, 0, 0, 0);
            v4f_t accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xil, z, 0, 0, 0);
            v4f_t accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xil, z, 0, 0, 0);
            v4f_t accJ = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xrl, z, 0, 0, 0);
            accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(url, xrh, accR, 0, 0, 0);
            accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uil, xih, accT, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(url, xih, accI, 0, 0, 0);
            accJ = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uil, xrh, accJ, 0, 0, 0);
            accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urm, xrm, accR, 0, 0, 0);
            accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uim, xim, accT, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urm, xim, accI, 0, 0, 0);
            accJ = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uim, xrm, accJ, 0, 0, 0);
            accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xrm, accR, 0, 0, 0);
            accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xim, accT, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xim, accI, 0, 0, 0);
            accJ = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xrm, accJ, 0, 0, 0);
            accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urm, xrh, accR, 0, 0, 0);
            accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uim, xih, accT, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urm, xih, accI, 0, 0, 0);
            accJ = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uim, xrh, accJ, 0, 0, 0);
            accR = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xrh, accR, 0, 0, 0);
            accT = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xih, accT, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xih, accI, 0, 0, 0);
            accJ = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xrh, accJ, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i)
              ys[roff(4 * rb + i) ^ tc[w]] =
                  make_float2(flip_sign(accR[i] - accT[i], fl), flip_sign(accI[i] + accJ[i], fl));
          }
        }
        continue;
      }
      bf8 xrh[CBW], xrl[CBW], xih[CBW], xil[CBW];
#pragma unroll
      for (int w = 0; w < CBW; ++w) {
        float2 v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = ys[roff(e) ^ tc[w]];
        split_bf8(v, &xrh[w], &xrl[w], &xih[w], &xil[w]);
      }
      // a row block's U fragments are read (once for all the wave's column
      // blocks) just before its products; the second block's reads are
      // issued before the first block's results are written
      bf8 urh, url, uih, uil;
      if constexpr (U_DIRECT) {
        urh = Af[0][0], url = Af[0][1], uih = Af[0][2], uil = Af[0][3];
      } else {
        urh = As[0], url = As[64], uih = As[128], uil = As[192];
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        // Re = Ur Xr - Ui Xi, Im = Ur Xi + Ui Xr, each real product as
        // hi*hi + hi*lo + lo*hi: four independent chains of three per block
        v4f_t accR[CBW], accT[CBW], accI[CBW], accJ[CBW];
#pragma unroll
        for (int w = 0; w < CBW; ++w) {
          const v4f_t z = {0, 0, 0, 0};
          accR[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xrh[w], z, 0, 0, 0);
          accT[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xih[w], z, 0, 0, 0);
          accI[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xih[w], z, 0, 0, 0);
          accJ[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xrh[w], z, 0, 0, 0);
        }
#pragma unroll
        for (int w = 0; w < CBW; ++w) {
          accR[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xrl[w], accR[w], 0, 0, 0);
          accT[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xil[w], accT[w], 0, 0, 0);
          accI[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(urh, xil[w], accI[w], 0, 0, 0);
          accJ[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uih, xrl[w], accJ[w], 0, 0, 0);
        }
#pragma unroll
        for (int w = 0; w < CBW; ++w) {
          accR[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(url, xrh[w], accR[w], 0, 0, 0);
          accT[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uil, xih[w], accT[w], 0, 0, 0);
          accI[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(url, xih[w], accI[w], 0, 0, 0);
          accJ[w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uil, xrh[w], accJ[w], 0, 0, 0);
        }
        if (rb == 0) {
          if constexpr (U_DIRECT) {
            urh = Af[1][0], url = Af[1][1], uih = Af[1][2], uil = Af[1][3];
          } else {
            urh = As[256], url = As[320], uih = As[384], uil = As[448];
          }
        }
        // every wave has read its rows of this step above: write in place
#pragma unroll
        for (int w = 0; w < CBW; ++w)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            ys[roff(4 * rb + i) ^ tc[w]] = make_float2(accR[w][i] - accT[w][i], accI[w][i] + accJ[w][i]);
      }
    }
    __syncthreads();
#pragma unroll 8
    for (int j = 0; j < PER; ++j) {
      const int t = p_lo | phi[wave + W * j];
      const float2 y = ys[t ^ swz[t >> 4]];
      V o;
      o.x = (R)y.x;
      o.y = (R)y.y;
      st_stream<R>(&a[base + a_lo + ahi[wave + W * j]], o);
    }
  }
}

namespace {

// the highest `count` slots of an outer mask (the tile kernels' fast slots)
uint64_t top_slots(uint64_t outer, int count) {
  uint64_t top = 0;
  for (int i = 0; i < count && (outer & ~top); ++i) top |= 1ull << (63 - __builtin_clzll(outer & ~top));
  return top;
}

uint64_t pdep_host(uint64_t x, uint64_t mask) {
  uint64_t r = 0;
  for (uint64_t bit = 1; mask; bit <<= 1) {
    const uint64_t low = mask & (~mask + 1);
    if (x & bit) r |= low;
    mask &= mask - 1;
  }
  return r;
}

template <typename R, int K, bool SPLIT>
void launch(void* amp, const void* scratch, int D, int C, int grid, hipStream_t st, const FusedArgs& f) {
  const double2* U = static_cast<const double2*>(scratch);
  const uint64_t* qoff = reinterpret_cast<const uint64_t*>(U + size_t(D) * D);
  const uint64_t* toff = qoff + D;
  const size_t lds = size_t(D) * C * sizeof(double2);
  hipLaunchKernelGGL((k_fused<R, K, SPLIT>), grid, 4 * (1 << K), lds, st,
                     static_cast<typename VecT<R>::type*>(amp), U, qoff, toff, f);
}

template <typename R, bool SPLIT>
void launch_k(int K, void* amp, const void* scratch, int D, int C, int grid, hipStream_t st, const FusedArgs& f) {
  switch (K) {
    case 4: launch<R, 4, SPLIT>(amp, scratch, D, C, grid, st, f); break;
    case 5: launch<R, 5, SPLIT>(amp, scratch, D, C, grid, st, f); break;
    case 6: launch<R, 6, SPLIT>(amp, scratch, D, C, grid, st, f); break;
    default: launch<R, 7, SPLIT>(amp, scratch, D, C, grid, st, f); break;
  }
}

}  // namespace

int fused_unitary(int n, const std::vector<Lowered>& gates, std::vector<int>* qubits, std::vector<cd>* U,
                  std::string* err, int min_qubits) {
  uint64_t mask = 0;
  for (const Lowered& L : gates) {
    for (int i = 0; i < L.nt; ++i) mask |= 1ull << L.t[i];
    for (int i = 0; i < L.nc; ++i) mask |= 1ull << L.c[i];
  }
  if (__builtin_popcountll(mask) > 7) {
    *err = "fused parts take at most 7 qubits";
    return HSV_ERR_LIMIT;
  }
  if (n < min_qubits) {
    *err = "fused parts need at least " + std::to_string(min_qubits) + " qubits in the state";
    return HSV_ERR_LIMIT;
  }
  for (int b = 0; b < n && __builtin_popcountll(mask) < min_qubits; ++b) mask |= 1ull << b;  // pad: U (x) I
  qubits->clear();
  for (int b = 0; b < 64; ++b)
    if (mask >> b & 1) qubits->push_back(b);
  build_fused(gates, *qubits, U);
  return HSV_OK;
}

namespace {

// Round-to-nearest-even FP32 -> BF16 (finite inputs; NaN stays NaN).
uint16_t bf16_rne(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
double bf16_value(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

int ensure_scratch(DeviceScratch* scratch, size_t bytes, hipStream_t stream, std::string* err) {
  if (scratch->bytes >= bytes) return HSV_OK;
  if (scratch->fixed) {
    *err = "this part needs " + std::to_string(bytes) + " bytes of device scratch; the no-allocation entry point "
           "has " + std::to_string(scratch->bytes);
    return HSV_ERR_LIMIT;
  }
  if (hipStreamSynchronize(stream) != hipSuccess) { *err = "stream sync failed"; return HSV_ERR_DEVICE; }
  if (scratch->ptr) (void)hipFree(scratch->ptr);
  scratch->ptr = nullptr;
  scratch->bytes = 0;
  const size_t cap = std::max<size_t>(bytes, 1 << 20);
  if (hipMalloc(&scratch->ptr, cap) != hipSuccess) { *err = "fused scratch allocation failed"; return HSV_ERR_ALLOC; }
  scratch->bytes = cap;
  return HSV_OK;
}

template <typename R, int K, bool SWZ>
void launch_bf(void* amp, const void* scratch, int grid, hipStream_t st, const BfArgs& f, int* rc) {
  constexpr int D = 1 << K;
  const size_t frag_bytes = size_t(8) * D * D;
  const auto* U = static_cast<const bf8*>(scratch);
  const uint64_t* qoff = reinterpret_cast<const uint64_t*>(static_cast<const unsigned char*>(scratch) + frag_bytes);
  const uint64_t* toff = qoff + D;
  auto kern = &k_fused_bf<R, K, SWZ>;
  if (frag_bytes > 65536 &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)frag_bytes) != hipSuccess) {
    *rc = HSV_ERR_DEVICE;
    return;
  }
  hipLaunchKernelGGL(kern, grid, kBfThreads, frag_bytes, st, static_cast<typename VecT<R>::type*>(amp), U, qoff,
                     toff, f);
}

// The BF16-split path: U as BF16 hi/lo fragments in the MFMA A-operand order,
// columns in blocks of 16 along the lowest 4 slots outside the part.
template <typename R, int K>
void launch_bft(void* amp, const void* scratch, int grid, hipStream_t st, uint64_t ntiles, uint64_t outer, uint64_t top,
                int tpw) {
  const size_t frag_bytes = size_t(8) << (2 * K);
  const auto* U = static_cast<const bf8*>(scratch);
  const auto* T = reinterpret_cast<const BfTileTables*>(static_cast<const unsigned char*>(scratch) + frag_bytes);
  hipLaunchKernelGGL((k_fused_bft<R, K>), grid, kBfTileThreads, 0, st,
                     static_cast<typename VecT<R>::type*>(amp), U, T, ntiles, outer, top, tpw);
}

// Tables of the tile form: tile index bit j <-> the j-th lowest of the tile's
// 12 slots (the part's qubits plus the 12 - K lowest slots outside it).
void build_tile_tables(uint64_t qmask, uint64_t cmask, BfTileTables* T) {
  int slot[12], rowpos[12], colpos[12];
  int j = 0, nr = 0, nc = 0;
  for (int b = 0; b < 64 && j < 12; ++b) {
    if (qmask >> b & 1) { slot[j] = b; rowpos[j] = nr++; colpos[j] = -1; ++j; }
    else if (cmask >> b & 1) { slot[j] = b; rowpos[j] = -1; colpos[j] = nc++; ++j; }
  }
  std::memset(T, 0, sizeof(*T));
  for (int v = 0; v < 64; ++v)
    for (int bit = 0; bit < 6; ++bit) {
      if (!(v >> bit & 1)) continue;
      T->alo[v] += 1ull << slot[bit];
      T->ahi[v] += 1ull << slot[6 + bit];
      if (rowpos[bit] >= 0) T->rlo[v] += (uint16_t)(1u << rowpos[bit]);
      if (rowpos[6 + bit] >= 0) T->rhi[v] += (uint16_t)(1u << rowpos[6 + bit]);
      if (colpos[bit] >= 0) T->clo[v] += (uint16_t)(1u << colpos[bit]);
      if (colpos[6 + bit] >= 0) T->chi[v] += (uint16_t)(1u << colpos[6 + bit]);
    }
  for (int q = 0; q < 12; ++q) {
    for (int r = 0; r < (1 << nr); ++r)
      if (rowpos[q] >= 0 && (r >> rowpos[q] & 1)) T->tr[r] |= (uint16_t)(1u << q);
    for (int c = 0; c < (1 << nc); ++c)
      if (colpos[q] >= 0 && (c >> colpos[q] & 1)) T->tc[c] |= (uint16_t)(1u << q);
  }
}

// LDS swizzle of k_part_bf. nimg[c] = h applied to tile bit c: 0 for
// c < 4, a mask of bits 0-3 for c = 4, of bits 0-4 for c >= 5. A step is
// conflict-free when the low-4-bit images of its 4 lowest column bits are
// independent (the 16 lanes of a quarter-wave) and, with its lowest row bit,
// the low-5-bit images are too (the 32 lanes of a half-wave).
// rank over GF(2) of n <= 5 vectors of 5 bits
int rank_gf2(uint32_t* v, int n) {
  int r = 0;
  for (int bit = 4; bit >= 0; --bit) {
    int piv = -1;
    for (int i = r; i < n; ++i)
      if (v[i] >> bit & 1) { piv = i; break; }
    if (piv < 0) continue;
    std::swap(v[r], v[piv]);
    for (int i = 0; i < n; ++i)
      if (i != r && (v[i] >> bit & 1)) v[i] ^= v[r];
    ++r;
  }
  return r;
}

struct PartSwizzleKey { int cbit[4]; int rbit0; };

void choose_part_swizzle(int KT, const std::vector<PartSwizzleKey>& keys, uint32_t nimg[14]) {
  auto score = [&](const uint32_t* m) {
    int sc = 0;
    for (const PartSwizzleKey& k : keys) {
      auto low5 = [&](int c) { return ((c < 5) ? (1u << c) : 0u) ^ m[c]; };
      uint32_t a[4], b[5];
      for (int j = 0; j < 4; ++j) { b[j] = low5(k.cbit[j]); a[j] = b[j] & 15u; }
      b[4] = low5(k.rbit0);
      sc += 2 * (rank_gf2(a, 4) == 4) + (rank_gf2(b, 5) == 5);
    }
    return sc;
  };
  const int best_possible = 3 * (int)keys.size();
  uint32_t cand[14] = {0};
  std::memcpy(nimg, cand, sizeof(cand));
  int best = score(cand);
  uint64_t rng = 0x9E3779B97F4A7C15ull;
  for (int it = 0; it < 2000 && best < best_possible; ++it) {
    for (int c = 0; c < 14; ++c) {
      rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
      cand[c] = c < 4 ? 0u : (c == 4 ? uint32_t(rng & 15u) : (c < KT ? uint32_t(rng & 31u) : 0u));
    }
    const int sc = score(cand);
    if (sc > best) { best = sc; std::memcpy(nimg, cand, sizeof(cand)); }
  }
}

template <typename R, int KT, bool STREAM_U, bool U_DIRECT = false, int NP = 2>
void launch_part_bf(void* amp, const unsigned char* dev, size_t tabs_off, size_t steps_off, size_t frag_off,
                    int nsteps, int grid, hipStream_t st, uint64_t ntiles, uint64_t outer, uint64_t top, int tpw,
                    int* rc) {
  const size_t lds = STREAM_U ? 0 : size_t(nsteps) * 2 * 2 * NP * 64 * sizeof(bf8);
  auto kern = &k_part_bf<R, KT, STREAM_U, U_DIRECT, NP>;
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)std::max<size_t>(lds, 1)) != hipSuccess) {
    *rc = HSV_ERR_DEVICE;
    return;
  }
  hipLaunchKernelGGL(kern, grid, part_bf_threads(KT, STREAM_U), lds, st, static_cast<typename VecT<R>::type*>(amp),
                     reinterpret_cast<const PartBfTables*>(dev + tabs_off),
                     reinterpret_cast<const PartBfStep*>(dev + steps_off),
                     reinterpret_cast<const bf8*>(dev + frag_off), nsteps, ntiles, outer, top, tpw);
}

// A part of 8..14 qubits on the BF16-split path (see k_part_bf).
// tile_qubits: the (logical) qubits the tile must hold, at least those of
// the gates. Logical qubit q sits in slot phys[q] (q when phys is null) and
// the store leaves it in slot phys_after[q] (unchanged when null): a layout
// change that must keep every qubit outside the tile where it is.
// pieces: BF16 pieces per value (2, or 3 for FP32-class products; 3 takes
// 12- and 13-qubit tiles).
int part_apply_bf(void* amp, int n, int dtype, hipStream_t stream, DeviceScratch* scratch,
                  const std::vector<Lowered>& gates, uint64_t tile_qubits, const int* phys,
                  const int* phys_after, uint64_t live_slots, const BfTune& bf, std::string* err, int pieces) {
  int tpw = bf.part_tpw;
  int ph[64], pa[64], inv[64];
  for (int q = 0; q < n; ++q) ph[q] = phys ? phys[q] : q;
  for (int q = 0; q < n; ++q) {
    pa[q] = phys_after ? phys_after[q] : ph[q];
    inv[ph[q]] = q;
  }
  uint64_t pmask = 0;
  for (int q = 0; q < n; ++q)
    if (tile_qubits >> q & 1) pmask |= 1ull << ph[q];
  const int K = __builtin_popcountll(pmask);
  // The tile is the part's qubits padded with the lowest free slots. It
  // should hold slots 0-4, so that each wave's loads and stores cover
  // 512-byte runs: a part that fits 12 qubits only without them takes a
  // 13-qubit tile.
  const uint64_t low5 = (1ull << std::min(n, 5)) - 1;
  const int KT = std::max(12, std::min(14, __builtin_popcountll(pmask | low5)));
  if (K > 14) {
    *err = "the bf16 fused path takes parts of at most 14 qubits";
    return HSV_ERR_LIMIT;
  }
  if (n < KT) {
    *err = "a wide bf16 fused part needs a state of at least " + std::to_string(KT) + " qubits";
    return HSV_ERR_LIMIT;
  }
  if (pieces == 3 && KT > 13) {
    *err = "the three-piece bf16 split takes parts whose qubits and slots 0-4 fit 13 qubits";
    return HSV_ERR_LIMIT;
  }
  uint64_t tmask = pmask;
  for (int b = 0; b < n && __builtin_popcountll(tmask) < KT; ++b) tmask |= 1ull << b;
  int slot[14], sslot[14];
  uint64_t smask = 0;
  for (int b = 0, j = 0; b < 64 && j < KT; ++b)
    if (tmask >> b & 1) {
      slot[j] = b;
      sslot[j] = pa[inv[b]];
      smask |= 1ull << sslot[j];
      ++j;
    }
  for (int q = 0; q < n; ++q)
    if (!(tmask >> ph[q] & 1) && pa[q] != ph[q]) smask = 0;
  if (smask != tmask) {
    *err = "the layout change moves a qubit into or out of the part's tile";
    return HSV_ERR_ARG;
  }
  auto inner = [&](int q) { return __builtin_popcountll(tmask & ((1ull << ph[q]) - 1)); };
  for (const Lowered& L : gates) {
    if (L.kind == K_IDENT) continue;  // no-ops: their qubits need no tile slot
    for (int i = 0; i < L.nt + L.nc; ++i)
      if (!(tmask >> ph[i < L.nt ? L.t[i] : L.c[i - L.nt]] & 1)) {
        *err = "a gate acts on a qubit outside the part's tile";
        return HSV_ERR_ARG;
      }
  }
  std::vector<uint64_t> qm;
  std::vector<const Lowered*> gl;
  for (const Lowered& L : gates) {
    if (L.kind == K_IDENT) continue;
    uint64_t m = 0;
    for (int i = 0; i < L.nt; ++i) m |= 1ull << inner(L.t[i]);
    for (int i = 0; i < L.nc; ++i) m |= 1ull << inner(L.c[i]);
    qm.push_back(m);
    gl.push_back(&L);
  }
  bool moves = false;
  for (int j = 0; j < KT; ++j) moves = moves || sslot[j] != slot[j];
  if (qm.empty() && !moves) return HSV_OK;  // identities only
  // the steps: groups of <= 5 qubits, as few as a beam of 16 finds
  std::vector<std::vector<int>> groups = qm.empty() ? std::vector<std::vector<int>>() : step_groups(qm, 5, 16);
  if (groups.empty() && !qm.empty()) {
    *err = "a gate acts on more than 5 qubits: the wide bf16 fused path cannot take it";
    return HSV_ERR_LIMIT;
  }
  const int ns = (int)groups.size();
  std::vector<PartBfStep> steps(ns);
  std::vector<PartSwizzleKey> keys(ns);
  const int nsets = 2 * pieces;   // real pieces, then imaginary pieces
  const size_t step_frag_bytes = size_t(2) * nsets * 64 * 16;
  std::vector<uint16_t> frags(size_t(ns) * 2 * nsets * 64 * 8);
  for (int s = 0; s < ns; ++s) {
    uint64_t gm = 0;
    std::vector<Lowered> sg;
    for (int i : groups[s]) { gm |= qm[i]; sg.push_back(*gl[i]); }
    for (int b = 0; b < KT && __builtin_popcountll(gm) < 5; ++b) gm |= 1ull << b;  // pad: U (x) I
    int rbit[5], cbit[9];
    std::vector<int> qubits;
    for (int b = 0, r = 0, c = 0; b < KT; ++b) {
      if (gm >> b & 1) { rbit[r++] = b; qubits.push_back(inv[slot[b]]); }
      else cbit[c++] = b;
    }
    std::vector<cd> U;
    build_fused(sg, qubits, &U);
    const double uerr = unitarity_error(U, 32);
    if (!(uerr <= 1e-10)) {
      *err = "fused step matrix is not unitary (max |U^H U - I| = " + std::to_string(uerr) + ")";
      return HSV_ERR_UNITARY;
    }
    PartBfStep& S = steps[s];
    std::memset(&S, 0, sizeof(S));
    keys[s] = PartSwizzleKey{{cbit[0], cbit[1], cbit[2], cbit[3]}, rbit[0]};
    for (int r = 0; r < 32; ++r)
      for (int b = 0; b < 5; ++b)
        if (r >> b & 1) S.rt[r] |= (uint16_t)(1u << rbit[b]);
    const int ncb = KT - 5;
    for (int v = 0; v < 32; ++v)
      for (int j = 0; j < 5; ++j) {
        if (!(v >> j & 1)) continue;
        if (j < 4 && j < ncb && v < 16) S.ctl[v] |= (uint16_t)(1u << cbit[j]);
        if (4 + j < ncb) S.cth[v] |= (uint16_t)(1u << cbit[4 + j]);
      }
    for (int rb = 0; rb < 2; ++rb)
      for (int lane = 0; lane < 64; ++lane)
        for (int e = 0; e < 8; ++e) {
          const int mm = lane % 16;
          const int row = 16 * rb + 4 * (mm % 4) + mm / 4;
          const int col = lane / 16 + 4 * e;
          const cd u = U[size_t(row) * 32 + col];
          const size_t f = size_t((s * 2 + rb) * nsets) * 64 + lane;
          // each piece the BF16 of what the previous ones leave of the
          // FP64 entry
          double re = u.real(), im = u.imag();
          for (int k = 0; k < pieces; ++k) {
            const uint16_t pr = bf16_rne((float)re), pi = bf16_rne((float)im);
            frags[(f + size_t(k) * 64) * 8 + e] = pr;
            frags[(f + size_t(pieces + k) * 64) * 8 + e] = pi;
            re -= bf16_value(pr);
            im -= bf16_value(pi);
          }
        }
  }
  // tile bit j's qubit moves to tile bit sig[j] (the tile position of its new
  // slot); store position u reads load position sig^-1(u)
  int sig_inv[14];
  for (int j = 0; j < KT; ++j)
    for (int k = 0; k < KT; ++k)
      if (slot[k] == sslot[j]) sig_inv[k] = j;
  // the store gather of a layout change reads a quarter-wave's 16 positions
  // along load bits sig^-1(0..3): its banks count like a step's columns
  if (moves) keys.push_back(PartSwizzleKey{{sig_inv[0], sig_inv[1], sig_inv[2], sig_inv[3]}, sig_inv[4]});
  // swizzle the steps' LDS offsets (h is linear: XOR of the bits' images)
  uint32_t nimg[14];
  choose_part_swizzle(KT, keys, nimg);
  auto hmap = [&](uint32_t t) {
    uint32_t h = 0;
    for (int c = 4; c < KT; ++c)
      if (t >> c & 1) h ^= nimg[c];
    return h;
  };
  for (PartBfStep& S : steps) {
    for (auto& x : S.rt) x = (uint16_t)(x ^ hmap(x));
    for (auto& x : S.ctl) x = (uint16_t)(x ^ hmap(x));
    for (auto& x : S.cth) x = (uint16_t)(x ^ hmap(x));
  }
  // two table sets: T0 stores where it loads (passes before the last, when
  // the steps need several), T applies the layout change
  PartBfTables T0, T;
  std::memset(&T0, 0, sizeof(T0));
  for (int i = 0; i < (1 << (KT - 4)); ++i) T0.swz[i] = (uint8_t)hmap(uint32_t(i) << 4);
  for (int v = 0; v < 256; ++v)
    for (int j = 0; j < 8; ++j) {
      if (!(v >> j & 1)) continue;
      if (v < 64 && j < 6) T0.alo[v] += 1ull << slot[j];
      if (6 + j < KT) T0.ahi[v] += 1ull << slot[6 + j];
    }
  for (int v = 0; v < 64; ++v) T0.plo[v] = (uint16_t)v;
  for (int v = 0; v < 256; ++v) T0.phi[v] = (uint16_t)(v << 6);
  T = T0;
  std::memset(T.plo, 0, sizeof(T.plo));
  std::memset(T.phi, 0, sizeof(T.phi));
  for (int v = 0; v < 256; ++v)
    for (int k = 0; k < 8; ++k) {
      if (!(v >> k & 1)) continue;
      if (v < 64 && k < 6) T.plo[v] |= (uint16_t)(1u << sig_inv[k]);
      if (6 + k < KT) T.phi[v] |= (uint16_t)(1u << sig_inv[6 + k]);
    }
  // LDS: the tile (FP32) + step tables + tile tables statically, 8 KB per step
  const size_t static_lds =
      (size_t(1) << KT) * 8 + sizeof(PartBfStep) * kPartBfMaxSteps + (64 + (size_t(1) << (KT - 6))) * 10 +
      (size_t(1) << (KT - 4));
  // 13-qubit tiles: resident fragments (1024 threads) when the part's steps
  // fit LDS beside the tile, else streamed
  // (the three-piece split streams them at 13 qubits)
  const bool stream_u = KT >= 14 || (KT == 13 && (pieces == 3 || bf.part13 != 1 ||
                                                  (160 * 1024 - static_lds) / step_frag_bytes < size_t(ns)));
  const bool u_direct = KT == 13 && stream_u && bf.part13 == 2 && pieces == 2;
  const int max_steps = stream_u ? part_bf_max_steps(KT, pieces)
                                 : (int)std::min<size_t>(kPartBfMaxSteps, (160 * 1024 - static_lds) / step_frag_bytes);
  const size_t steps_off = 2 * sizeof(PartBfTables);
  const size_t frag_off = (steps_off + sizeof(PartBfStep) * ns + 255) & ~size_t(255);
  const size_t bytes = frag_off + frags.size() * 2;
  std::vector<unsigned char> host(bytes);
  std::memcpy(host.data(), &T0, sizeof(T0));
  std::memcpy(host.data() + sizeof(T0), &T, sizeof(T));
  std::memcpy(host.data() + steps_off, steps.data(), sizeof(PartBfStep) * ns);
  std::memcpy(host.data() + frag_off, frags.data(), frags.size() * 2);
  if (int rc = ensure_scratch(scratch, bytes, stream, err)) return rc;
  if (hipMemcpyAsync(scratch->ptr, host.data(), bytes, hipMemcpyHostToDevice, stream) != hipSuccess) {
    *err = "fused upload failed";
    return HSV_ERR_DEVICE;
  }
  const uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1);
  // tiles with a set bit in a slot outside live_slots hold only zeros (and
  // stay zero): they are skipped
  const uint64_t outer = all & ~tmask & live_slots;
  const uint64_t ntiles = 1ull << __builtin_popcountll(outer);
  if (tpw <= 0) tpw = stream_u ? 1 : 8;
  const int grid = (int)((ntiles + tpw - 1) / tpw);
  const uint64_t top = top_slots(outer, bf.tile_order);
  const auto* dev = static_cast<const unsigned char*>(scratch->ptr);
  const bool c128 = dtype == HSV_C128;
  int rc = HSV_OK;
  // more steps than fit in LDS at once: one pass per chunk of steps, the
  // last one with the layout change (a layout change alone: one pass, no steps)
  for (int s0 = 0; (s0 < ns || (ns == 0 && s0 == 0)) && rc == HSV_OK; s0 += std::max(max_steps, 1)) {
    const int cnt = std::min(max_steps, ns - s0);
    const size_t to = s0 + cnt >= ns ? sizeof(PartBfTables) : 0;
    const size_t so = steps_off + sizeof(PartBfStep) * s0;
    const size_t fo = frag_off + size_t(s0) * step_frag_bytes;
    if (KT == 12 && pieces == 3)
      c128 ? launch_part_bf<double, 12, false, false, 3>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc)
           : launch_part_bf<float, 12, false, false, 3>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc);
    else if (KT == 13 && pieces == 3)
      c128 ? launch_part_bf<double, 13, true, false, 3>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc)
           : launch_part_bf<float, 13, true, false, 3>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc);
    else if (KT == 12) c128 ? launch_part_bf<double, 12, false>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc)
                       : launch_part_bf<float, 12, false>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc);
    else if (KT == 13 && u_direct)
      c128 ? launch_part_bf<double, 13, true, true>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc)
           : launch_part_bf<float, 13, true, true>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc);
    else if (KT == 13 && stream_u)
      c128 ? launch_part_bf<double, 13, true>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc)
           : launch_part_bf<float, 13, true>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc);
    else if (KT == 13)
      c128 ? launch_part_bf<double, 13, false>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc)
           : launch_part_bf<float, 13, false>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc);
    else c128 ? launch_part_bf<double, 14, true>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc)
              : launch_part_bf<float, 14, true>(amp, dev, to, so, fo, cnt, grid, stream, ntiles, outer, top, tpw, &rc);
  }
  if (rc) { *err = "wide bf16 fused kernel attribute setup failed"; return rc; }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("fused launch: ") + hipGetErrorString(e);
    return HSV_ERR_DEVICE;
  }
  return HSV_OK;
}

// variant: 0 / 1 = column-block kernel (1: swizzled row order), 2 = tile form
// (needs n >= 12; falls back to 1 below that).
int fused_apply_bf(void* amp, int n, int dtype, hipStream_t stream, int num_cu, DeviceScratch* scratch,
                   const std::vector<int>& Q, const std::vector<cd>& Um, int variant, int tpw, int tile_order,
                   std::string* err) {
  const int K = (int)Q.size();
  const int D = 1 << K;
  if (K < 5 || K > 7) {
    *err = "the bf16 fused path takes parts of 5..7 qubits";
    return HSV_ERR_LIMIT;
  }
  const bool tile = variant == 2 && n >= 12;
  const int swizzle = variant >= 1 ? 1 : 0;
  const int RB = D / 16, KS = D / 32;
  const size_t frag_bytes = size_t(8) * D * D;
  std::vector<unsigned char> host(frag_bytes + std::max(size_t(D) * 8 + 16 * 8, sizeof(BfTileTables)));
  uint16_t* hf = reinterpret_cast<uint16_t*>(host.data());
  for (int rb = 0; rb < RB; ++rb)
    for (int ks = 0; ks < KS; ++ks)
      for (int lane = 0; lane < 64; ++lane)
        for (int e = 0; e < 8; ++e) {
          const int mm = lane % 16, g = lane / 16;
          const int row = 16 * rb + (swizzle ? 4 * (mm % 4) + mm / 4 : mm);
          const int col = 32 * ks + (swizzle ? g + 4 * e : 8 * g + e);
          const cd u = Um[size_t(row) * D + col];
          const uint16_t rh = bf16_rne((float)u.real()), ih = bf16_rne((float)u.imag());
          const uint16_t rl = bf16_rne((float)(u.real() - bf16_value(rh)));
          const uint16_t il = bf16_rne((float)(u.imag() - bf16_value(ih)));
          const size_t frag = size_t((rb * KS + ks) * 4) * 64;  // first of the 4 sets, in 8-element units
          hf[(frag + 0 * 64 + lane) * 8 + e] = rh;
          hf[(frag + 1 * 64 + lane) * 8 + e] = rl;
          hf[(frag + 2 * 64 + lane) * 8 + e] = ih;
          hf[(frag + 3 * 64 + lane) * 8 + e] = il;
        }
  uint64_t qmask = 0;
  for (int q : Q) qmask |= 1ull << q;
  const uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1);
  if (tile) {
    uint64_t cmask = 0;
    for (int b = 0, got = 0; b < n && got < 12 - K; ++b)
      if (!(qmask >> b & 1)) { cmask |= 1ull << b; ++got; }
    build_tile_tables(qmask, cmask, reinterpret_cast<BfTileTables*>(host.data() + frag_bytes));
    const size_t bytes = frag_bytes + sizeof(BfTileTables);
    if (int rc = ensure_scratch(scratch, bytes, stream, err)) return rc;
    if (hipMemcpyAsync(scratch->ptr, host.data(), bytes, hipMemcpyHostToDevice, stream) != hipSuccess) {
      *err = "fused upload failed";
      return HSV_ERR_DEVICE;
    }
    const uint64_t ntiles = 1ull << (n - 12);
    const uint64_t outer = all & ~qmask & ~cmask;
    const int grid = (int)((ntiles + tpw - 1) / tpw);
    const uint64_t top = top_slots(outer, tile_order);
    const bool c128 = dtype == HSV_C128;
    if (K == 5) c128 ? launch_bft<double, 5>(amp, scratch->ptr, grid, stream, ntiles, outer, top, tpw)
                     : launch_bft<float, 5>(amp, scratch->ptr, grid, stream, ntiles, outer, top, tpw);
    else if (K == 6) c128 ? launch_bft<double, 6>(amp, scratch->ptr, grid, stream, ntiles, outer, top, tpw)
                          : launch_bft<float, 6>(amp, scratch->ptr, grid, stream, ntiles, outer, top, tpw);
    else c128 ? launch_bft<double, 7>(amp, scratch->ptr, grid, stream, ntiles, outer, top, tpw)
              : launch_bft<float, 7>(amp, scratch->ptr, grid, stream, ntiles, outer, top, tpw);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      *err = std::string("fused launch: ") + hipGetErrorString(e);
      return HSV_ERR_DEVICE;
    }
    return HSV_OK;
  }
  uint64_t tmask = 0;
  for (int b = 0, got = 0; b < n && got < 4; ++b)
    if (!(qmask >> b & 1)) { tmask |= 1ull << b; ++got; }
  uint64_t* hq = reinterpret_cast<uint64_t*>(host.data() + frag_bytes);
  for (int r = 0; r < D; ++r) hq[r] = pdep_host((uint64_t)r, qmask);
  for (int c = 0; c < 16; ++c) hq[D + c] = pdep_host((uint64_t)c, tmask);
  const size_t bytes = frag_bytes + size_t(D) * 8 + 16 * 8;
  if (int rc = ensure_scratch(scratch, bytes, stream, err)) return rc;
  if (hipMemcpyAsync(scratch->ptr, host.data(), bytes, hipMemcpyHostToDevice, stream) != hipSuccess) {
    *err = "fused upload failed";
    return HSV_ERR_DEVICE;
  }
  BfArgs f{};
  f.outer_mask = all & ~qmask & ~tmask;
  f.nblocks = 1ull << (n - K - 4);
  const uint64_t waves = kBfThreads / 64;
  const int grid = (int)std::min<uint64_t>((f.nblocks + waves - 1) / waves, uint64_t(num_cu) * 2);
  int rc = HSV_OK;
  const bool c128 = dtype == HSV_C128;
#define HSV_BF_LAUNCH(K_, S_)                                                                  \
  (c128 ? launch_bf<double, K_, S_>(amp, scratch->ptr, grid, stream, f, &rc)                  \
        : launch_bf<float, K_, S_>(amp, scratch->ptr, grid, stream, f, &rc))
  if (swizzle) {
    if (K == 5) HSV_BF_LAUNCH(5, true); else if (K == 6) HSV_BF_LAUNCH(6, true); else HSV_BF_LAUNCH(7, true);
  } else {
    if (K == 5) HSV_BF_LAUNCH(5, false); else if (K == 6) HSV_BF_LAUNCH(6, false); else HSV_BF_LAUNCH(7, false);
  }
#undef HSV_BF_LAUNCH
  if (rc) { *err = "bf16 fused kernel attribute setup failed"; return rc; }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("fused launch: ") + hipGetErrorString(e);
    return HSV_ERR_DEVICE;
  }
  return HSV_OK;
}

}  // namespace

int fused_part_bf_layout(void* amp, int n, int dtype, hipStream_t stream, DeviceScratch* scratch,
                         const std::vector<Lowered>& gates, uint64_t tile_qubits, const int* phys,
                         const int* phys_after, uint64_t live_slots, const BfTune& bf, std::string* err,
                         int pieces) {
  return part_apply_bf(amp, n, dtype, stream, scratch, gates, tile_qubits, phys, phys_after, live_slots, bf, err,
                       pieces);
}

int fused_apply(void* amp, int n, int dtype, hipStream_t stream, int num_cu, DeviceScratch* scratch,
                const std::vector<Lowered>& gates, int mode, std::string* err, const BfTune& bf) {
  if (mode != HSV_FUSED_F64 && mode != HSV_FUSED_F32SPLIT && mode != HSV_FUSED_BF16SPLIT &&
      mode != HSV_FUSED_BF16X3) {
    *err = "unknown fused mode";
    return HSV_ERR_ARG;
  }
  if (mode == HSV_FUSED_BF16X3) {
    // the three-piece split runs every part on the wide-part kernel, in a
    // 12- or 13-qubit tile with slots 0-4
    uint64_t pmask = 0;
    for (const Lowered& L : gates) {
      if (L.kind == K_IDENT) continue;
      for (int i = 0; i < L.nt; ++i) pmask |= 1ull << L.t[i];
      for (int i = 0; i < L.nc; ++i) pmask |= 1ull << L.c[i];
    }
    return part_apply_bf(amp, n, dtype, stream, scratch, gates, pmask, nullptr, nullptr, ~0ull, bf, err, 3);
  }
  if (mode == HSV_FUSED_BF16SPLIT) {
    uint64_t pmask = 0;
    for (const Lowered& L : gates) {
      if (L.kind == K_IDENT) continue;
      for (int i = 0; i < L.nt; ++i) pmask |= 1ull << L.t[i];
      for (int i = 0; i < L.nc; ++i) pmask |= 1ull << L.c[i];
    }
    if (__builtin_popcountll(pmask) > 7)
      return part_apply_bf(amp, n, dtype, stream, scratch, gates, pmask, nullptr, nullptr, ~0ull, bf, err, 2);
  }
  std::vector<int> Q;
  std::vector<cd> Um;
  if (int rc = fused_unitary(n, gates, &Q, &Um, err, mode == HSV_FUSED_BF16SPLIT ? 5 : 4)) return rc;
  const int K = (int)Q.size();
  const int D = 1 << K;
  const double uerr = unitarity_error(Um, D);
  if (!(uerr <= 1e-10)) {
    *err = "fused matrix is not unitary (max |U^H U - I| = " + std::to_string(uerr) + ")";
    return HSV_ERR_UNITARY;
  }
  const int free_bits = n - K;
  if (free_bits < 4) {
    *err = "fused parts need at least 4 state qubits outside the part";
    return HSV_ERR_LIMIT;
  }
  if (mode == HSV_FUSED_BF16SPLIT)
    return fused_apply_bf(amp, n, dtype, stream, num_cu, scratch, Q, Um, bf.variant, std::max(bf.tile_tpw, 1),
                          bf.tile_order, err);
  int NC = 2048 / (16 * D);
  while (NC > 1 && 4 + __builtin_ctz(NC) > free_bits) NC >>= 1;
  const int C = 16 * NC;
  const int logC = __builtin_ctz(C);
  uint64_t qmask = 0;
  for (int q : Q) qmask |= 1ull << q;
  uint64_t tmask = 0;
  for (int b = 0, got = 0; b < n && got < logC; ++b)
    if (!(qmask >> b & 1)) { tmask |= 1ull << b; ++got; }
  const uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1);
  FusedArgs f{};
  f.outer_mask = all & ~qmask & ~tmask;
  f.ntiles = 1ull << (n - K - logC);
  f.C = C;
  f.logC = logC;
  f.NC = NC;
  // scratch layout: U (D*D double2) | qoff (D u64) | toff (C u64)
  std::vector<unsigned char> host(size_t(D) * D * 16 + size_t(D) * 8 + size_t(C) * 8);
  double* hu = reinterpret_cast<double*>(host.data());
  for (size_t i = 0; i < size_t(D) * D; ++i) { hu[2 * i] = Um[i].real(); hu[2 * i + 1] = Um[i].imag(); }
  uint64_t* hq = reinterpret_cast<uint64_t*>(host.data() + size_t(D) * D * 16);
  for (int r = 0; r < D; ++r) hq[r] = pdep_host((uint64_t)r, qmask);
  uint64_t* ht = hq + D;
  for (int c = 0; c < C; ++c) ht[c] = pdep_host((uint64_t)c, tmask);
  if (int rc = ensure_scratch(scratch, host.size(), stream, err)) return rc;
  if (hipMemcpyAsync(scratch->ptr, host.data(), host.size(), hipMemcpyHostToDevice, stream) != hipSuccess) {
    *err = "fused upload failed";
    return HSV_ERR_DEVICE;
  }
  const int per_cu = K >= 7 ? 2 : 4;
  const int grid = (int)std::min<uint64_t>(f.ntiles, uint64_t(num_cu) * per_cu);
  const bool split = mode == HSV_FUSED_F32SPLIT;
  if (dtype == HSV_C128) {
    if (split) launch_k<double, true>(K, amp, scratch->ptr, D, C, grid, stream, f);
    else launch_k<double, false>(K, amp, scratch->ptr, D, C, grid, stream, f);
  } else {
    if (split) launch_k<float, true>(K, amp, scratch->ptr, D, C, grid, stream, f);
    else launch_k<float, false>(K, amp, scratch->ptr, D, C, grid, stream, f);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("fused launch: ") + hipGetErrorString(e);
    return HSV_ERR_DEVICE;
  }
  return HSV_OK;
}

}  // namespace hsv
