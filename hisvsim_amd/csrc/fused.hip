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

namespace {

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
int fused_apply_bf(void* amp, int n, int dtype, hipStream_t stream, int num_cu, DeviceScratch* scratch,
                   const std::vector<int>& Q, const std::vector<cd>& Um, int swizzle, std::string* err) {
  const int K = (int)Q.size();
  const int D = 1 << K;
  if (K < 5 || K > 7) {
    *err = "the bf16 fused path takes parts of 5..7 qubits";
    return HSV_ERR_LIMIT;
  }
  const int RB = D / 16, KS = D / 32;
  const size_t frag_bytes = size_t(8) * D * D;
  std::vector<unsigned char> host(frag_bytes + size_t(D) * 8 + 16 * 8);
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
  uint64_t tmask = 0;
  for (int b = 0, got = 0; b < n && got < 4; ++b)
    if (!(qmask >> b & 1)) { tmask |= 1ull << b; ++got; }
  uint64_t* hq = reinterpret_cast<uint64_t*>(host.data() + frag_bytes);
  for (int r = 0; r < D; ++r) hq[r] = pdep_host((uint64_t)r, qmask);
  for (int c = 0; c < 16; ++c) hq[D + c] = pdep_host((uint64_t)c, tmask);
  if (int rc = ensure_scratch(scratch, host.size(), stream, err)) return rc;
  if (hipMemcpyAsync(scratch->ptr, host.data(), host.size(), hipMemcpyHostToDevice, stream) != hipSuccess) {
    *err = "fused upload failed";
    return HSV_ERR_DEVICE;
  }
  const uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1);
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

int fused_apply(void* amp, int n, int dtype, hipStream_t stream, int num_cu, DeviceScratch* scratch,
                const std::vector<Lowered>& gates, int mode, std::string* err, int bf_swizzle) {
  if (mode != HSV_FUSED_F64 && mode != HSV_FUSED_F32SPLIT && mode != HSV_FUSED_BF16SPLIT) {
    *err = "unknown fused mode";
    return HSV_ERR_ARG;
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
    return fused_apply_bf(amp, n, dtype, stream, num_cu, scratch, Q, Um, bf_swizzle, err);
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
