// Device kernels of libhsv (gfx950). Included by hsv.hip only.
//
// Layout: the state is one array of 2^n interleaved complex amplitudes in HBM
// (double2 for complex128, float2 for complex64); amplitude index bit q is
// qubit q. Every kernel addresses amplitudes through "insert zero bits at the
// fixed positions" index maps, so consecutive work items touch consecutive
// addresses along every free bit and loads/stores stay coalesced.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsv {

template <typename R> struct VecT;
template <> struct VecT<double> { using type = double2; };
template <> struct VecT<float> { using type = float2; };

// ---- small complex helpers (R = double or float) --------------------------
template <typename V, typename R>
__device__ __forceinline__ V cmk(R x, R y) { V v; v.x = x; v.y = y; return v; }

template <typename V, typename R>
__device__ __forceinline__ V cmul(V a, R bx, R by) {
  return cmk<V, R>(a.x * bx - a.y * by, a.x * by + a.y * bx);
}
// acc + a*b as two FMA chains (4 FMAs; `acc += a*b - c*d` written naively
// compiles to mul + fma + add per component).
template <typename V, typename R>
__device__ __forceinline__ V cmac(V acc, V a, R bx, R by) {
  acc.x = fma(a.x, bx, fma(-a.y, by, acc.x));  // HIP's fma overloads float and double
  acc.y = fma(a.x, by, fma(a.y, bx, acc.y));
  return acc;
}

// Streaming loads/stores; NT = non-temporal (the state is touched once per
// pass, so keeping it out of the caches leaves them to the other stream).
template <bool NT, typename V>
__device__ __forceinline__ V ldg(const V* p) {
  if constexpr (NT) {
    V v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    return v;
  } else {
    return *p;
  }
}
template <bool NT, typename V>
__device__ __forceinline__ void stg(V* p, V v) {
  if constexpr (NT) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
  } else {
    *p = v;
  }
}

// Insert a zero bit at each position in `fixed` (ascending), then set `set`.
template <typename I>
__device__ __forceinline__ I expand_index(I w, int nfixed, const int* fixed, I set) {
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (i < nfixed) {
      const int p = fixed[i];
      const I lo = w & ((I(1) << p) - 1);
      w = ((w ^ lo) << 1) | lo;
    }
  }
  return w | set;
}

__device__ __forceinline__ uint64_t pdep64(uint64_t x, uint64_t mask) {
  uint64_t r = 0;
  for (uint64_t bit = 1; mask; bit <<= 1) {
    const uint64_t low = mask & (~mask + 1);
    if (x & bit) r |= low;
    mask &= mask - 1;
  }
  return r;
}

// ---- per-gate sweep ---------------------------------------------------------
struct GateArgs {
  uint64_t nwork;    // number of work items (2^(n - #fixed bits))
  uint64_t setmask;  // bits forced to 1 (controls; the target of a phase gate)
  uint64_t off1;     // 1 << targets[0]
  uint64_t off2;     // 1 << targets[1] (two-target gates)
  int nfixed;
  int fixed[7];      // ascending fixed bit positions (targets and controls)
  double m[32];      // matrix entries (re, im), row-major
};

enum { GK_PHASE1 = 1, GK_DIAG1 = 2, GK_XSWAP1 = 3, GK_DENSE1 = 4, GK_DIAG2 = 5, GK_SWAP2 = 6, GK_DENSE2 = 7 };

constexpr int kGateBlock = 256;

template <typename R, int KIND, int ITEMS, bool NT>
__global__ __launch_bounds__(kGateBlock) void k_gate(typename VecT<R>::type* __restrict__ a, GateArgs g) {
  using V = typename VecT<R>::type;
  constexpr uint64_t per_block = uint64_t(kGateBlock) * ITEMS;
  R m[32];
  constexpr int NM = (KIND == GK_DENSE2) ? 32 : (KIND == GK_DIAG2 || KIND == GK_DENSE1) ? 8 : 4;
#pragma unroll
  for (int i = 0; i < NM; ++i) m[i] = (R)g.m[i];
  for (uint64_t blk = blockIdx.x; blk * per_block < g.nwork; blk += gridDim.x) {
    uint64_t idx[ITEMS];
    bool ok[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint64_t w = blk * per_block + uint64_t(j) * kGateBlock + threadIdx.x;
      ok[j] = w < g.nwork;
      idx[j] = expand_index<uint64_t>(ok[j] ? w : 0, g.nfixed, g.fixed, g.setmask);
    }
    if constexpr (KIND == GK_PHASE1) {
      V v[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) v[j] = ldg<NT>(&a[idx[j]]);
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) stg<NT>(&a[idx[j]], cmul<V, R>(v[j], m[0], m[1]));
    } else if constexpr (KIND == GK_DIAG1 || KIND == GK_XSWAP1 || KIND == GK_DENSE1) {
      V v0[ITEMS], v1[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) { v0[j] = ldg<NT>(&a[idx[j]]); v1[j] = ldg<NT>(&a[idx[j] + g.off1]); }
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        if (!ok[j]) continue;
        V o0, o1;
        if constexpr (KIND == GK_DIAG1) {
          o0 = cmul<V, R>(v0[j], m[0], m[1]);
          o1 = cmul<V, R>(v1[j], m[2], m[3]);
        } else if constexpr (KIND == GK_XSWAP1) {
          o0 = v1[j];
          o1 = v0[j];
        } else {
          o0 = cmac<V, R>(cmul<V, R>(v0[j], m[0], m[1]), v1[j], m[2], m[3]);
          o1 = cmac<V, R>(cmul<V, R>(v0[j], m[4], m[5]), v1[j], m[6], m[7]);
        }
        stg<NT>(&a[idx[j]], o0);
        stg<NT>(&a[idx[j] + g.off1], o1);
      }
    } else if constexpr (KIND == GK_SWAP2) {
      V v0[ITEMS], v1[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) { v0[j] = ldg<NT>(&a[idx[j] + g.off1]); v1[j] = ldg<NT>(&a[idx[j] + g.off2]); }
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) { stg<NT>(&a[idx[j] + g.off1], v1[j]); stg<NT>(&a[idx[j] + g.off2], v0[j]); }
    } else {  // GK_DIAG2, GK_DENSE2: v[2*b(t0) + b(t1)]
      V v[ITEMS][4];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        if (!ok[j]) continue;
        v[j][0] = ldg<NT>(&a[idx[j]]);
        v[j][1] = ldg<NT>(&a[idx[j] + g.off2]);
        v[j][2] = ldg<NT>(&a[idx[j] + g.off1]);
        v[j][3] = ldg<NT>(&a[idx[j] + g.off1 + g.off2]);
      }
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        if (!ok[j]) continue;
        V o[4];
        if constexpr (KIND == GK_DIAG2) {
#pragma unroll
          for (int r = 0; r < 4; ++r) o[r] = cmul<V, R>(v[j][r], m[2 * r], m[2 * r + 1]);
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            V acc = cmul<V, R>(v[j][0], m[8 * r], m[8 * r + 1]);
            acc = cmac<V, R>(acc, v[j][1], m[8 * r + 2], m[8 * r + 3]);
            acc = cmac<V, R>(acc, v[j][2], m[8 * r + 4], m[8 * r + 5]);
            o[r] = cmac<V, R>(acc, v[j][3], m[8 * r + 6], m[8 * r + 7]);
          }
        }
        stg<NT>(&a[idx[j]], o[0]);
        stg<NT>(&a[idx[j] + g.off2], o[1]);
        stg<NT>(&a[idx[j] + g.off1], o[2]);
        stg<NT>(&a[idx[j] + g.off1 + g.off2], o[3]);
      }
    }
  }
}

// ---- LDS part kernel --------------------------------------------------------
// One workgroup owns one outer base at a time: it gathers the part's 2^k
// amplitudes into LDS, applies every gate of the part there, and scatters
// them back -- one HBM pass per part instead of one per gate.
struct PartGate {
  int kind;
  int nfixed;
  uint32_t setmask;
  uint32_t off1;
  uint32_t off2;
  uint32_t nwork;
  int fixed[7];
  int pad;
  double m[32];
};

constexpr int kMaxPartQubits = 14;

template <typename R, int BS>
__device__ __forceinline__ void part_gate_lds(typename VecT<R>::type* lds, const PartGate& G) {
  using V = typename VecT<R>::type;
  const int kind = G.kind;
  const uint32_t nwork = G.nwork;
  const uint32_t off1 = G.off1, off2 = G.off2;
  if (kind == GK_PHASE1) {
    const R dx = (R)G.m[0], dy = (R)G.m[1];
    for (uint32_t w = threadIdx.x; w < nwork; w += BS) {
      const uint32_t i = expand_index<uint32_t>(w, G.nfixed, G.fixed, G.setmask);
      lds[i] = cmul<V, R>(lds[i], dx, dy);
    }
  } else if (kind == GK_DIAG1 || kind == GK_XSWAP1 || kind == GK_DENSE1) {
    const R m0 = (R)G.m[0], m1 = (R)G.m[1], m2 = (R)G.m[2], m3 = (R)G.m[3];
    const R m4 = (R)G.m[4], m5 = (R)G.m[5], m6 = (R)G.m[6], m7 = (R)G.m[7];
    for (uint32_t w = threadIdx.x; w < nwork; w += BS) {
      const uint32_t i = expand_index<uint32_t>(w, G.nfixed, G.fixed, G.setmask);
      const V v0 = lds[i], v1 = lds[i + off1];
      V o0, o1;
      if (kind == GK_DIAG1) {
        o0 = cmul<V, R>(v0, m0, m1);
        o1 = cmul<V, R>(v1, m2, m3);
      } else if (kind == GK_XSWAP1) {
        o0 = v1;
        o1 = v0;
      } else {
        o0 = cmac<V, R>(cmul<V, R>(v0, m0, m1), v1, m2, m3);
        o1 = cmac<V, R>(cmul<V, R>(v0, m4, m5), v1, m6, m7);
      }
      lds[i] = o0;
      lds[i + off1] = o1;
    }
  } else if (kind == GK_SWAP2) {
    for (uint32_t w = threadIdx.x; w < nwork; w += BS) {
      const uint32_t i = expand_index<uint32_t>(w, G.nfixed, G.fixed, G.setmask);
      const V v0 = lds[i + off1], v1 = lds[i + off2];
      lds[i + off1] = v1;
      lds[i + off2] = v0;
    }
  } else {  // GK_DIAG2 / GK_DENSE2
    for (uint32_t w = threadIdx.x; w < nwork; w += BS) {
      const uint32_t i = expand_index<uint32_t>(w, G.nfixed, G.fixed, G.setmask);
      V v[4] = {lds[i], lds[i + off2], lds[i + off1], lds[i + off1 + off2]};
      V o[4];
      if (kind == GK_DIAG2) {
#pragma unroll
        for (int r = 0; r < 4; ++r) o[r] = cmul<V, R>(v[r], (R)G.m[2 * r], (R)G.m[2 * r + 1]);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          V acc = cmul<V, R>(v[0], (R)G.m[8 * r], (R)G.m[8 * r + 1]);
          acc = cmac<V, R>(acc, v[1], (R)G.m[8 * r + 2], (R)G.m[8 * r + 3]);
          acc = cmac<V, R>(acc, v[2], (R)G.m[8 * r + 4], (R)G.m[8 * r + 5]);
          o[r] = cmac<V, R>(acc, v[3], (R)G.m[8 * r + 6], (R)G.m[8 * r + 7]);
        }
      }
      lds[i] = o[0];
      lds[i + off2] = o[1];
      lds[i + off1] = o[2];
      lds[i + off1 + off2] = o[3];
    }
  }
}

template <typename R, int BS>
__global__ __launch_bounds__(BS) void k_part(typename VecT<R>::type* __restrict__ a, uint64_t nbases,
                                             uint64_t outer_mask, uint64_t part_mask, int k,
                                             const PartGate* __restrict__ gates, int ngates) {
  using V = typename VecT<R>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  __shared__ uint64_t tlo[64];
  __shared__ uint64_t thi[1 << (kMaxPartQubits - 6)];
  const int klo = k < 6 ? k : 6;
  for (int j = threadIdx.x; j < 64; j += BS) tlo[j] = j < (1 << klo) ? pdep64((uint64_t)j, part_mask) : 0;
  for (int j = threadIdx.x; j < (1 << (kMaxPartQubits - 6)); j += BS)
    thi[j] = j < (1 << (k - klo)) ? pdep64((uint64_t)j << klo, part_mask) : 0;
  __syncthreads();
  const uint32_t nin = 1u << k;
  for (uint64_t w = blockIdx.x; w < nbases; w += gridDim.x) {
    const uint64_t base = pdep64(w, outer_mask);
    for (uint32_t j = threadIdx.x; j < nin; j += BS) lds[j] = a[base + tlo[j & 63] + thi[j >> 6]];
    __syncthreads();
    for (int gi = 0; gi < ngates; ++gi) {
      part_gate_lds<R, BS>(lds, gates[gi]);
      __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j < nin; j += BS) a[base + tlo[j & 63] + thi[j >> 6]] = lds[j];
  }
}

// ---- reductions (FP64 accumulation, fixed grid -> deterministic) -----------
enum { RD_NORM2 = 0, RD_INNER = 1, RD_MAXDIFF = 2, RD_PROB0 = 3, RD_CHUNK = 4 };
constexpr int kRedBlocks = 1024;
constexpr int kRedBlock = 256;

template <typename R, int MODE>
__global__ __launch_bounds__(kRedBlock) void k_reduce(const typename VecT<R>::type* __restrict__ a,
                                                      const typename VecT<R>::type* __restrict__ b,
                                                      uint64_t len, int qubit, double* __restrict__ partial) {
  double s0 = 0.0, s1 = 0.0;
  if constexpr (MODE == RD_CHUNK) {
    const uint64_t chunk = len / gridDim.x;
    const uint64_t lo = blockIdx.x * chunk;
    for (uint64_t i = lo + threadIdx.x; i < lo + chunk; i += kRedBlock) {
      const double x = a[i].x, y = a[i].y;
      s0 += x * x + y * y;
    }
  } else {
    const uint64_t stride = uint64_t(gridDim.x) * kRedBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kRedBlock + threadIdx.x; i < len; i += stride) {
      const double x = a[i].x, y = a[i].y;
      if constexpr (MODE == RD_NORM2) {
        s0 += x * x + y * y;
      } else if constexpr (MODE == RD_PROB0) {
        if (!((i >> qubit) & 1)) s0 += x * x + y * y;
      } else if constexpr (MODE == RD_INNER) {
        const double u = b[i].x, v = b[i].y;
        s0 += x * u + y * v;   // Re(conj(a) b)
        s1 += x * v - y * u;   // Im(conj(a) b)
      } else {                 // RD_MAXDIFF; max |a_i| without b
        const double dx = b ? x - (double)b[i].x : x, dy = b ? y - (double)b[i].y : y;
        s0 = fmax(s0, sqrt(dx * dx + dy * dy));
      }
    }
  }
  __shared__ double r0[kRedBlock], r1[kRedBlock];
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int h = kRedBlock / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      if constexpr (MODE == RD_MAXDIFF) {
        r0[threadIdx.x] = fmax(r0[threadIdx.x], r0[threadIdx.x + h]);
      } else {
        r0[threadIdx.x] += r0[threadIdx.x + h];
        r1[threadIdx.x] += r1[threadIdx.x + h];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r0[0];
    partial[2 * blockIdx.x + 1] = r1[0];
  }
}

// <psi|P|psi> for a Pauli string: P|j> = i^ny (-1)^popc(j & z) |j ^ x>, so the
// sum is i^ny * sum_j conj(a[j ^ x]) a[j] (-1)^popc(j & z); the host applies
// i^ny. Same fixed grid and partials layout as k_reduce.
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_pauli(const typename VecT<R>::type* __restrict__ a, uint64_t len,
                                                     uint64_t xmask, uint64_t zmask, double* __restrict__ partial) {
  double s0 = 0.0, s1 = 0.0;
  const uint64_t stride = uint64_t(gridDim.x) * kRedBlock;
  for (uint64_t j = uint64_t(blockIdx.x) * kRedBlock + threadIdx.x; j < len; j += stride) {
    const double x = a[j].x, y = a[j].y;
    const uint64_t k = j ^ xmask;
    const double u = a[k].x, v = a[k].y;
    const double sg = (__popcll(j & zmask) & 1) ? -1.0 : 1.0;
    s0 += sg * (u * x + v * y);   // Re(conj(a_k) a_j)
    s1 += sg * (u * y - v * x);   // Im(conj(a_k) a_j)
  }
  __shared__ double r0[kRedBlock], r1[kRedBlock];
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int h = kRedBlock / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      r0[threadIdx.x] += r0[threadIdx.x + h];
      r1[threadIdx.x] += r1[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r0[0];
    partial[2 * blockIdx.x + 1] = r1[0];
  }
}

// Probability mass of the amplitudes with (i & mask) == value.
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_mass(const typename VecT<R>::type* __restrict__ a, uint64_t len,
                                                    uint64_t mask, uint64_t value, double* __restrict__ partial) {
  double s0 = 0.0;
  const uint64_t stride = uint64_t(gridDim.x) * kRedBlock;
  for (uint64_t i = uint64_t(blockIdx.x) * kRedBlock + threadIdx.x; i < len; i += stride) {
    if ((i & mask) != value) continue;
    const double x = a[i].x, y = a[i].y;
    s0 += x * x + y * y;
  }
  __shared__ double r0[kRedBlock];
  r0[threadIdx.x] = s0;
  __syncthreads();
  for (int h = kRedBlock / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) r0[threadIdx.x] += r0[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r0[0];
    partial[2 * blockIdx.x + 1] = 0.0;
  }
}

// Projection onto (i & mask) == value with renormalisation by `scale`.
template <typename R>
__global__ __launch_bounds__(256) void k_project(typename VecT<R>::type* __restrict__ a, uint64_t len, uint64_t mask,
                                                 uint64_t value, double scale) {
  using V = typename VecT<R>::type;
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < len; i += stride) {
    const V v = a[i];
    a[i] = (i & mask) == value ? cmk<V, R>(R(v.x * scale), R(v.y * scale)) : cmk<V, R>(R(0), R(0));
  }
}

// ---- fill |0...0> -----------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_zero_state(typename VecT<R>::type* __restrict__ a, uint64_t len) {
  using V = typename VecT<R>::type;
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < len; i += stride)
    a[i] = cmk<V, R>(i == 0 ? R(1) : R(0), R(0));
}

// ---- in-place qubit involution ---------------------------------------------
// The qubits of each slot pair trade slots: amplitudes a[i] and a[sigma(i)]
// swap, sigma exchanging the two bit values of every pair. Any qubit
// permutation is two such involutions. Tiled so that every access covers
// 512-byte runs: a tile is the 2^nb amplitudes over the slot set B = slots
// 0-4 plus their partners (nb <= 10); sigma maps the tile at outer address o
// onto the tile at sigma_hi(o) (the pairs outside B act on o) with its
// positions permuted by sigma_lo (the pairs inside B). A workgroup takes the
// pair of tiles (o, sigma_hi(o)) with o <= sigma_hi(o): both are loaded
// coalesced into LDS, and each position f of one is stored from position
// sigma_lo(f) of the other. Exact (FP64 stays FP64); every amplitude is read
// and written once.
struct SwapTiles {
  uint64_t outer_mask;           // slots outside B
  uint64_t lo[32], hi[32];       // address of tile position bits 0-4 / 5-9
  uint16_t slo[32], shi[32];     // sigma_lo of a tile position: slo[f & 31] | shi[f >> 5]
  int nb;                        // |B|
  int npairs;                    // pairs outside B
  uint8_t pa[32], pb[32];
};
constexpr int kSwapThreads = 512;
constexpr int kSwapPer = 1024 / kSwapThreads;  // positions per thread and tile (nb <= 10)
// Optionally fused with a renormalisation: norm_out (one FP64 partial of
// sum |a|^2 per workgroup, in a fixed order) in one pass, scale (a device
// value, multiplied into every stored amplitude) in the next. The next pair
// of tiles is loaded into registers while the current one is stored, so
// every workgroup keeps loads in flight through its store phase.
template <typename R>
__global__ __launch_bounds__(kSwapThreads) void k_swap_tiles(typename VecT<R>::type* __restrict__ a, uint64_t nouter,
                                                           SwapTiles st, double* __restrict__ norm_out,
                                                           const double* __restrict__ scale) {
  using V = typename VecT<R>::type;
  // 40 KB of LDS in all (FP64): four workgroups per CU
  __shared__ V ta[1024], tb[1024];
  double acc = 0.0;
  const R sc = scale ? R(*scale) : R(1);
  // outer address of tile w and of its partner, per 7-bit chunk of w
  __shared__ uint64_t ot[4][128], ot2[4][128];
  for (int i = threadIdx.x; i < 512; i += kSwapThreads) {
    const int c = i >> 7, v = i & 127;
    const uint64_t o = pdep64(uint64_t(v) << (7 * c), st.outer_mask);
    uint64_t o2 = o;
    for (int p = 0; p < st.npairs; ++p) {
      const uint64_t d = ((o >> st.pa[p]) ^ (o >> st.pb[p])) & 1;
      o2 ^= (d << st.pa[p]) | (d << st.pb[p]);
    }
    ot[c][v] = o;
    ot2[c][v] = o2;
  }
  const int T = 1 << st.nb;
  __syncthreads();  // tables built
  // the workgroup's tiles w = blockIdx.x + j gridDim.x whose partner is not
  // below them (the partner tile's workgroup does the others); uniform
  uint64_t w = blockIdx.x, o = 0, o2 = 0;
  auto advance = [&]() {
    for (; w < nouter; w += gridDim.x) {
      const int w0 = w & 127, w1 = (w >> 7) & 127, w2 = (w >> 14) & 127, w3 = (w >> 21) & 127;
      o = ot[0][w0] | ot[1][w1] | ot[2][w2] | ot[3][w3];
      o2 = ot2[0][w0] | ot2[1][w1] | ot2[2][w2] | ot2[3][w3];
      if (o2 >= o) return;
    }
  };
  V xa[kSwapPer], xb[kSwapPer];
  auto load = [&]() {
#pragma unroll
    for (int i = 0; i < kSwapPer; ++i) {
      const int f = threadIdx.x + kSwapThreads * i;
      if (f < T) {
        const uint64_t off = st.lo[f & 31] + st.hi[f >> 5];
        xa[i] = ldg<true>(&a[o + off]);
        if (o2 != o) xb[i] = ldg<true>(&a[o2 + off]);
      }
    }
  };
  advance();
  if (w < nouter) load();
  while (w < nouter) {
    const uint64_t co = o, co2 = o2;
    const bool pair = co2 != co;
#pragma unroll
    for (int i = 0; i < kSwapPer; ++i) {
      const int f = threadIdx.x + kSwapThreads * i;
      if (f < T) {
        ta[f] = xa[i];
        if (norm_out) acc += double(xa[i].x) * double(xa[i].x) + double(xa[i].y) * double(xa[i].y);
        if (pair) {
          tb[f] = xb[i];
          if (norm_out) acc += double(xb[i].x) * double(xb[i].x) + double(xb[i].y) * double(xb[i].y);
        }
      }
    }
    __syncthreads();  // the tiles are in LDS
    w += gridDim.x;
    advance();
    if (w < nouter) load();  // in flight during the stores below
#pragma unroll
    for (int i = 0; i < kSwapPer; ++i) {
      const int f = threadIdx.x + kSwapThreads * i;
      if (f < T) {
        const uint64_t off = st.lo[f & 31] + st.hi[f >> 5];
        const int g = st.slo[f & 31] | st.shi[f >> 5];
        V x = ta[g];
        x.x *= sc;
        x.y *= sc;
        if (pair) {
          V y = tb[g];
          y.x *= sc;
          y.y *= sc;
          stg<true>(&a[co2 + off], x);
          stg<true>(&a[co + off], y);
        } else {
          stg<true>(&a[co + off], x);
        }
      }
    }
    __syncthreads();  // LDS reads done before the next tile's writes
  }
  if (norm_out) {
    double* red = reinterpret_cast<double*>(ta);  // free: the loop ended on a barrier
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < kSwapThreads / 64; ++i) t += red[i];
      norm_out[blockIdx.x] = t;
    }
  }
}

// sum of the partials (in order) and the renormalisation factor: out[0] =
// norm^2, out[1] = 1 / norm (1 for a zero state, which stays unchanged)
static __global__ __launch_bounds__(64) void k_norm_scale(const double* __restrict__ partials, int count,
                                                         double* __restrict__ out) {
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < count; ++i) t += partials[i];
    out[0] = t;
    out[1] = t > 0.0 ? 1.0 / sqrt(t) : 1.0;
  }
}

// ---- qubit permutation (out of place) --------------------------------------
// out[sigma(i)] = in[i]; sigma moves bit s to bit new_slot[s]. sigma(i) is
// looked up per 9-bit chunk of i (4 chunks: n <= 36) from tables built per
// workgroup in LDS (16 KB).
constexpr int kPermChunk = 9;
template <typename R>
__global__ __launch_bounds__(256) void k_permute(const typename VecT<R>::type* __restrict__ in,
                                                 typename VecT<R>::type* __restrict__ out, uint64_t len,
                                                 int n, const int* __restrict__ new_slot_dev) {
  __shared__ uint64_t tab[4][1 << kPermChunk];
  __shared__ int ns[64];
  if (threadIdx.x < 64) ns[threadIdx.x] = threadIdx.x < n ? new_slot_dev[threadIdx.x] : 0;
  __syncthreads();
  for (int c = 0; c < 4; ++c) {
    for (int v = threadIdx.x; v < (1 << kPermChunk); v += 256) {
      uint64_t r = 0;
      for (int b = 0; b < kPermChunk; ++b) {
        const int s = c * kPermChunk + b;
        if (s < n && ((v >> b) & 1)) r |= uint64_t(1) << ns[s];
      }
      tab[c][v] = r;
    }
  }
  __syncthreads();
  const uint64_t mask = (uint64_t(1) << kPermChunk) - 1;
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < len; i += stride) {
    const uint64_t j = tab[0][i & mask] | tab[1][(i >> kPermChunk) & mask] |
                       tab[2][(i >> (2 * kPermChunk)) & mask] | tab[3][(i >> (3 * kPermChunk)) & mask];
    out[j] = in[i];
  }
}

// Outer address of tile w: the tile index deposited into the slots outside
// the tile. top holds the highest m outer slots; they vary fastest (the
// lowest bits of w go to them, highest slot first), so the workgroups
// running at one time spread over 2^m parts of the index space: measured on
// MI355X at 33 qubits, passes whose concurrent tiles sit in one half of the
// state run slower.
__device__ __forceinline__ uint64_t tile_base(uint64_t w, uint64_t outer_mask, uint64_t top) {
  if (!top) return pdep64(w, outer_mask);
  const int nt = __popcll(top);
  uint64_t hi = 0, t = top;
  for (int i = 0; i < nt; ++i) {  // bit i of w -> the (i+1)-th highest slot of top
    const uint64_t b = 1ull << (63 - __clzll(t));
    if (w >> i & 1) hi |= b;
    t &= ~b;
  }
  return pdep64(w >> nt, outer_mask & ~top) | hi;
}

// Non-temporal 2-element loads and stores of an amplitude (one dwordx4 or
// dwordx2 access with the streaming hint): the tile kernels touch each
// amplitude once per pass.
template <typename R> struct Pair2 { typedef R type __attribute__((ext_vector_type(2))); };
template <typename R>
__device__ __forceinline__ typename VecT<R>::type ld_stream(const typename VecT<R>::type* p) {
  using E = typename Pair2<R>::type;
  const E v = __builtin_nontemporal_load(reinterpret_cast<const E*>(p));
  typename VecT<R>::type o;
  o.x = v.x;
  o.y = v.y;
  return o;
}
template <typename R>
__device__ __forceinline__ void st_stream(typename VecT<R>::type* p, typename VecT<R>::type x) {
  using E = typename Pair2<R>::type;
  E v;
  v.x = x.x;
  v.y = x.y;
  __builtin_nontemporal_store(v, reinterpret_cast<E*>(p));
}

}  // namespace hsv

// ---- register-blocked part kernel ---------------------------------------------
// Each thread keeps NR = 2^RB of the part's 2^k amplitudes in registers; the
// k inner bits are split into RB "register bits" (which of a thread's values)
// and TB = k - RB "thread bits" (which thread). A layout is the set of inner
// bits that are register bits (`regmask`); register slot r holds the inner
// index jt | jr[r], jt = pdep(thread, ~regmask), jr[r] = pdep(r, regmask).
// A non-diagonal gate needs its targets among the register bits and then
// runs entirely in registers; diagonal gates and controls work in any
// layout. The host plan splits the part into segments, each a layout plus
// the gates it can apply; between segments the values go through LDS once to
// change layout. Global loads and stores use the first and last layouts.
namespace hsv {

struct RegSeg {
  uint32_t regmask;
  int first;   // first gate of the segment
  int count;
  int fast;    // tile kernel: first/count index rounds (TileRound), not gates
};

// 96 bytes, so the next gate's record can be prefetched into scalar
// registers while the current one runs; dense two-qubit matrices (32
// doubles) live in a side array indexed by m2.
struct RegGate {
  int kind;
  int p0, p1;       // register slots of targets[0] / targets[1] (non-diagonal kinds)
  uint32_t t0bit;   // inner bit of targets[0] (diagonal kinds)
  uint32_t t1bit;   // inner bit of targets[1]
  uint32_t cmask;   // inner bits that must be 1 (controls; the target of a phase gate)
  const double* m2;  // GK_DENSE2: device address of its 4x4 matrix (32 doubles)
  double m[8];       // 2x2 matrix or up to 4 diagonal entries, (re, im)
};

// Component-wise select and copy: whole-struct `c ? a : b` and struct copies
// make clang select between addresses and memcpy, which keeps the register
// array in scratch memory; per-component forms keep it in VGPRs.
template <typename V>
__device__ __forceinline__ void vsel(V& dst, bool on, const V& a, const V& b) {
  dst.x = on ? a.x : b.x;
  dst.y = on ? a.y : b.y;
}
template <typename V, typename W>
__device__ __forceinline__ void vcopy(V& dst, const W& src) {
  dst.x = src.x;
  dst.y = src.y;
}

// CTRL = false: the gate has no controls, so no per-value select is needed.
template <typename V, typename R, int NR, int P, bool CTRL>
__device__ __forceinline__ void reg_dense1_impl(V* v, const RegGate& G, uint32_t jt, const uint32_t* jr,
                                                bool swap_only) {
  const R m0 = (R)G.m[0], m1 = (R)G.m[1], m2 = (R)G.m[2], m3 = (R)G.m[3];
  const R m4 = (R)G.m[4], m5 = (R)G.m[5], m6 = (R)G.m[6], m7 = (R)G.m[7];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & (1 << P)) continue;
    constexpr int S = 1 << P;
    const bool on = CTRL ? ((jt | jr[r]) & G.cmask) == G.cmask : true;
    const V a0 = v[r], a1 = v[r + S];
    V o0, o1;
    if (swap_only) {
      o0 = a1;
      o1 = a0;
    } else {
      o0 = cmac<V, R>(cmul<V, R>(a0, m0, m1), a1, m2, m3);
      o1 = cmac<V, R>(cmul<V, R>(a0, m4, m5), a1, m6, m7);
    }
    vsel(v[r], on, o0, a0);
    vsel(v[r + S], on, o1, a1);
  }
}

template <typename V, typename R, int NR, int P>
__device__ __forceinline__ void reg_dense1(V* v, const RegGate& G, uint32_t jt, const uint32_t* jr, bool swap_only) {
  if (G.cmask) reg_dense1_impl<V, R, NR, P, true>(v, G, jt, jr, swap_only);
  else reg_dense1_impl<V, R, NR, P, false>(v, G, jt, jr, swap_only);
}

template <typename V, typename R, int NR, int P0, int P1>
__device__ __forceinline__ void reg_dense2(V* v, const RegGate& G, uint32_t jt, const uint32_t* jr, bool swap_only) {
  constexpr int S0 = 1 << P0, S1 = 1 << P1;
  const bool ctrl = G.cmask != 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if ((r & S0) || (r & S1)) continue;
    const bool on = !ctrl || ((jt | jr[r]) & G.cmask) == G.cmask;
    const V x0 = v[r], x1 = v[r + S1], x2 = v[r + S0], x3 = v[r + S0 + S1];
    V o[4];
    if (swap_only) {
      o[0] = x0; o[1] = x2; o[2] = x1; o[3] = x3;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double* M = G.m2;
        V acc = cmul<V, R>(x0, (R)M[8 * q], (R)M[8 * q + 1]);
        acc = cmac<V, R>(acc, x1, (R)M[8 * q + 2], (R)M[8 * q + 3]);
        acc = cmac<V, R>(acc, x2, (R)M[8 * q + 4], (R)M[8 * q + 5]);
        o[q] = cmac<V, R>(acc, x3, (R)M[8 * q + 6], (R)M[8 * q + 7]);
      }
    }
    vsel(v[r], on, o[0], x0);
    vsel(v[r + S1], on, o[1], x1);
    vsel(v[r + S0], on, o[2], x2);
    vsel(v[r + S0 + S1], on, o[3], x3);
  }
}

template <typename V, typename R, int NR, int RB, int P0>
__device__ __forceinline__ void reg_dense2_p1(V* v, const RegGate& G, uint32_t jt, const uint32_t* jr, bool sw) {
  switch (G.p1) {
    case 0: if constexpr (P0 != 0) reg_dense2<V, R, NR, P0, 0>(v, G, jt, jr, sw); break;
    case 1: if constexpr (P0 != 1 && RB > 1) reg_dense2<V, R, NR, P0, 1>(v, G, jt, jr, sw); break;
    case 2: if constexpr (P0 != 2 && RB > 2) reg_dense2<V, R, NR, P0, 2>(v, G, jt, jr, sw); break;
    case 3: if constexpr (P0 != 3 && RB > 3) reg_dense2<V, R, NR, P0, 3>(v, G, jt, jr, sw); break;
    case 4: if constexpr (P0 != 4 && RB > 4) reg_dense2<V, R, NR, P0, 4>(v, G, jt, jr, sw); break;
    case 5: if constexpr (P0 != 5 && RB > 5) reg_dense2<V, R, NR, P0, 5>(v, G, jt, jr, sw); break;
    default: break;
  }
}

template <typename V, typename R, int NR, int RB>
__device__ __forceinline__ void reg_apply(V* v, const RegGate& G, uint32_t jt, const uint32_t* jr) {
  const int kind = G.kind;
  if (kind == GK_PHASE1 && G.m[0] == -1.0 && G.m[1] == 0.0) {
    // Z / CZ / multi-controlled Z: a sign flip, no multiply
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const bool on = ((jt | jr[r]) & G.cmask) == G.cmask;
      v[r].x = on ? -v[r].x : v[r].x;
      v[r].y = on ? -v[r].y : v[r].y;
    }
  } else if (kind == GK_PHASE1 || kind == GK_DIAG1 || kind == GK_DIAG2) {
    const R d0x = (R)G.m[0], d0y = (R)G.m[1], d1x = (R)G.m[2], d1y = (R)G.m[3];
    const R d2x = (R)G.m[4], d2y = (R)G.m[5], d3x = (R)G.m[6], d3y = (R)G.m[7];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const uint32_t j = jt | jr[r];
      const bool on = (j & G.cmask) == G.cmask;
      R dx, dy;
      if (kind == GK_PHASE1) {
        dx = d0x; dy = d0y;
      } else if (kind == GK_DIAG1) {
        const bool b = j & G.t0bit;
        dx = b ? d1x : d0x; dy = b ? d1y : d0y;
      } else {
        const bool b0 = j & G.t0bit, b1 = j & G.t1bit;
        dx = b0 ? (b1 ? d3x : d2x) : (b1 ? d1x : d0x);
        dy = b0 ? (b1 ? d3y : d2y) : (b1 ? d1y : d0y);
      }
      const V o = cmul<V, R>(v[r], dx, dy);
      vsel(v[r], on, o, v[r]);
    }
  } else if (kind == GK_DENSE1 || kind == GK_XSWAP1) {
    const bool sw = kind == GK_XSWAP1;
    switch (G.p0) {
      case 0: reg_dense1<V, R, NR, 0>(v, G, jt, jr, sw); break;
      case 1: if constexpr (RB > 1) reg_dense1<V, R, NR, 1>(v, G, jt, jr, sw); break;
      case 2: if constexpr (RB > 2) reg_dense1<V, R, NR, 2>(v, G, jt, jr, sw); break;
      case 3: if constexpr (RB > 3) reg_dense1<V, R, NR, 3>(v, G, jt, jr, sw); break;
      case 4: if constexpr (RB > 4) reg_dense1<V, R, NR, 4>(v, G, jt, jr, sw); break;
      case 5: if constexpr (RB > 5) reg_dense1<V, R, NR, 5>(v, G, jt, jr, sw); break;
      default: break;
    }
  } else {  // GK_DENSE2 / GK_SWAP2
    const bool sw = kind == GK_SWAP2;
    switch (G.p0) {
      case 0: reg_dense2_p1<V, R, NR, RB, 0>(v, G, jt, jr, sw); break;
      case 1: if constexpr (RB > 1) reg_dense2_p1<V, R, NR, RB, 1>(v, G, jt, jr, sw); break;
      case 2: if constexpr (RB > 2) reg_dense2_p1<V, R, NR, RB, 2>(v, G, jt, jr, sw); break;
      case 3: if constexpr (RB > 3) reg_dense2_p1<V, R, NR, RB, 3>(v, G, jt, jr, sw); break;
      case 4: if constexpr (RB > 4) reg_dense2_p1<V, R, NR, RB, 4>(v, G, jt, jr, sw); break;
      case 5: if constexpr (RB > 5) reg_dense2_p1<V, R, NR, RB, 5>(v, G, jt, jr, sw); break;
      default: break;
    }
  }
}

__device__ __forceinline__ uint32_t pdep32(uint32_t x, uint32_t mask) {
  uint32_t r = 0;
  for (uint32_t bit = 1; mask; bit <<= 1) {
    const uint32_t low = mask & (~mask + 1);
    if (x & bit) r |= low;
    mask &= mask - 1;
  }
  return r;
}

template <int NR>
__device__ __forceinline__ void reg_layout(uint32_t regmask, uint32_t inner, uint32_t* jt, uint32_t* jr) {
  *jt = pdep32(threadIdx.x, inner & ~regmask);
  uint32_t bits[6] = {0, 0, 0, 0, 0, 0};
  uint32_t m = regmask;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    bits[i] = m & (~m + 1);
    m &= m - 1;
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
      if (r & (1 << i)) x |= bits[i];
    jr[r] = x;
  }
}

// Plain complex pair for register arrays: HIP's union-based double2 keeps
// the compiler from splitting a 32-entry array into registers (it lands in
// scratch); this struct has the same size and alignment and does split.
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R x, y;
};

template <typename R, int TB, int RB, bool NT>
__global__ __launch_bounds__(1 << TB) void k_part_reg(typename VecT<R>::type* __restrict__ a_vec, uint64_t nbases,
                                                      uint64_t outer_mask, uint64_t part_mask,
                                                      const RegSeg* __restrict__ segs, int nseg,
                                                      const RegGate* __restrict__ gates) {
  using V = cplx<R>;
  V* __restrict__ a = reinterpret_cast<V*>(a_vec);
  constexpr int NR = 1 << RB;
  constexpr int K = TB + RB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  __shared__ uint64_t tlo[64];
  __shared__ uint64_t thi[1 << (K > 6 ? K - 6 : 0)];
  constexpr int KLO = K < 6 ? K : 6;
  for (int j = threadIdx.x; j < 64; j += (1 << TB)) tlo[j] = j < (1 << KLO) ? pdep64((uint64_t)j, part_mask) : 0;
  for (int j = threadIdx.x; j < (1 << (K - KLO)); j += (1 << TB)) thi[j] = pdep64((uint64_t)j << KLO, part_mask);
  __syncthreads();
  const uint32_t inner = (1u << K) - 1;
  // The first layout is fixed, so the next base's values can be fetched into
  // `nxt` while the current base is processed: its loads are in flight
  // during this base's gates and stores.
  uint32_t jt0, jr0[NR];
  reg_layout<NR>(segs[0].regmask, inner, &jt0, jr0);
  V nxt[NR];
  if (blockIdx.x < nbases) {
    const uint64_t base = pdep64(blockIdx.x, outer_mask);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const uint32_t j = jt0 | jr0[r];
      vcopy(nxt[r], ldg<NT>(&a[base + tlo[j & 63] + thi[j >> 6]]));
    }
  }
  for (uint64_t w = blockIdx.x; w < nbases; w += gridDim.x) {
    const uint64_t base = pdep64(w, outer_mask);
    V v[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) vcopy(v[r], nxt[r]);
    const uint64_t wn = w + gridDim.x;
    if (wn < nbases) {
      const uint64_t nbase = pdep64(wn, outer_mask);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const uint32_t j = jt0 | jr0[r];
        vcopy(nxt[r], ldg<NT>(&a[nbase + tlo[j & 63] + thi[j >> 6]]));
      }
    }
    uint32_t jt = jt0, jr[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) jr[r] = jr0[r];
    for (int s = 0; s < nseg; ++s) {
      if (s > 0) {
        __syncthreads();  // everyone is done reading the previous exchange
#pragma unroll
        for (int r = 0; r < NR; ++r) vcopy(lds[jt | jr[r]], v[r]);
        __syncthreads();
        reg_layout<NR>(segs[s].regmask, inner, &jt, jr);
#pragma unroll
        for (int r = 0; r < NR; ++r) vcopy(v[r], lds[jt | jr[r]]);
      }
      const int g1 = segs[s].first + segs[s].count;
      // the next gate's record is fetched while the current one is applied
      RegGate cur;
      if (segs[s].first < g1) cur = gates[segs[s].first];
      for (int gi = segs[s].first; gi < g1; ++gi) {
        RegGate nxtg;
        if (gi + 1 < g1) nxtg = gates[gi + 1];
        reg_apply<V, R, NR, RB>(v, cur, jt, jr);
        cur = nxtg;
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const uint32_t j = jt | jr[r];
      stg<NT>(&a[base + tlo[j & 63] + thi[j >> 6]], v[r]);
    }
  }
}

// Sign bit flip of a real component: XOR of its top word with s (0 or the
// sign bit), one integer op on the high half of a double.
__device__ __forceinline__ double flip_sign(double x, uint32_t s) {
  return __hiloint2double(__double2hiint(x) ^ (int)s, __double2loint(x));
}
__device__ __forceinline__ float flip_sign(float x, uint32_t s) { return __int_as_float(__float_as_int(x) ^ (int)s); }

// One-qubit dense gate without controls on register slot P. Each pair's
// products that read a0 run first, so both outputs can land in the
// registers their inputs held (no copies where the gate loop's paths meet).
template <typename V, typename R, int NR, int P>
__device__ __forceinline__ void tile_dense1(V* v, const RegGate& G) {
  constexpr int S = 1 << P;
  const R m0 = (R)G.m[0], m1 = (R)G.m[1], m2 = (R)G.m[2], m3 = (R)G.m[3];
  const R m4 = (R)G.m[4], m5 = (R)G.m[5], m6 = (R)G.m[6], m7 = (R)G.m[7];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & S) continue;
    const R a0x = v[r].x, a0y = v[r].y, a1x = v[r + S].x, a1y = v[r + S].y;
    // every read of a0 precedes the writes of the first output, every read
    // of a1 but the last precedes the writes of the second
    R t1x = fma(-a0y, m5, a0x * m4), t1y = fma(a0y, m4, a0x * m5);
    R t0x = fma(-a0y, m1, a0x * m0), t0y = fma(a0y, m0, a0x * m1);
    t0x = fma(-a1y, m3, fma(a1x, m2, t0x));
    t0y = fma(a1y, m2, fma(a1x, m3, t0y));
    t1x = fma(-a1y, m7, t1x);
    t1y = fma(a1x, m7, t1y);
    t1x = fma(a1x, m6, t1x);
    t1y = fma(a1y, m6, t1y);
    v[r].x = t0x;
    v[r].y = t0y;
    v[r + S].x = t1x;
    v[r + S].y = t1y;
  }
}

// A fast segment of the tile kernel is a list of rounds. A round applies
// uncontrolled one-qubit dense gates on distinct register slots (they
// commute, so in slot order), then sign flips (Z, CZ, multi-controlled Z;
// diagonal, so they commute with each other): gates [sfirst, sfirst + ns).
// For a sign flip the host stores in t0bit the thread-bit part of its
// control and target bits and in t1bit the register slots r whose bits hold
// the register part; all of a round's flips make one per-lane mask over r.
struct TileRound {
  int g[5];   // gate on register slot P, or -1
  int sfirst;
  int ns;
  int pad;
};

template <typename V, typename R, int NR, int RB>
__device__ __forceinline__ void tile_round(V* v, const TileRound* __restrict__ q, const RegGate* __restrict__ gates,
                                           uint32_t jt) {
  if (q->g[0] >= 0) tile_dense1<V, R, NR, 0>(v, gates[q->g[0]]);
  if constexpr (RB > 1) if (q->g[1] >= 0) tile_dense1<V, R, NR, 1>(v, gates[q->g[1]]);
  if constexpr (RB > 2) if (q->g[2] >= 0) tile_dense1<V, R, NR, 2>(v, gates[q->g[2]]);
  if constexpr (RB > 3) if (q->g[3] >= 0) tile_dense1<V, R, NR, 3>(v, gates[q->g[3]]);
  if constexpr (RB > 4) if (q->g[4] >= 0) tile_dense1<V, R, NR, 4>(v, gates[q->g[4]]);
  const int s0 = q->sfirst, s1 = q->sfirst + q->ns;
  if (s0 < s1) {
    uint32_t mask = 0;
    for (int i = s0; i < s1; ++i) {
      const uint32_t tm = gates[i].t0bit, rsel = gates[i].t1bit;
      mask ^= (jt & tm) == tm ? rsel : 0u;
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const uint32_t sr = (mask << (31 - r)) & 0x80000000u;
      v[r].x = flip_sign(v[r].x, sr);
      v[r].y = flip_sign(v[r].y, sr);
    }
  }
}

// pdep of the thread index into the tile bits outside regmask (RB bits):
// a zero bit inserted at each register position, lowest first.
template <int RB>
__device__ __forceinline__ uint32_t tile_thread_bits(uint32_t regmask) {
  uint32_t x = threadIdx.x, m = regmask;
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const uint32_t lo = (m & (~m + 1)) - 1;
    m &= m - 1;
    x = (x & lo) | ((x & ~lo) << 1);
  }
  return x;
}

// The segments of a part on an LDS-resident tile of 2^KT amplitudes held by
// NTH threads: per segment, each thread reads the 2^RB values of its
// register layout (RB = KT - log2 NTH), applies the segment's gates (rounds
// or gate records) and writes them back in place.
__device__ __forceinline__ uint32_t tile_sw(uint32_t t) { return t ^ (((t >> 4) ^ (t >> 8)) & 15u); }
constexpr int tile_log2(int x) { return x <= 1 ? 0 : 1 + tile_log2(x >> 1); }

template <typename R, int KT, int NTH>
__device__ __forceinline__ void tile_segments(cplx<R>* ys, const RegSeg* __restrict__ segs, int nseg,
                                              const TileRound* __restrict__ rounds,
                                              const RegGate* __restrict__ gates) {
  using V = cplx<R>;
  constexpr int RB = KT - tile_log2(NTH), NR = 1 << RB;
  auto sw = tile_sw;
  for (int s = 0; s < nseg; ++s) {
    __syncthreads();  // the tile, or the previous segment's results, is in LDS
    const RegSeg S = segs[s];
    // jt: this thread's tile bits, jr[r]: register slot r's (uniform). The
    // swizzle is linear over XOR and jt, jr[r] have disjoint bits, so
    // sw(jt | jr[r]) = sw(jt) ^ sw(jr[r]): one XOR per LDS access.
    const uint32_t jt = tile_thread_bits<RB>(S.regmask);
    uint32_t jr[NR];
    {
      uint32_t bit[RB], m = S.regmask;
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        bit[i] = m & (~m + 1);
        m &= m - 1;
      }
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        uint32_t x = 0;
#pragma unroll
        for (int i = 0; i < RB; ++i)
          if (r >> i & 1) x |= bit[i];
        jr[r] = x;
      }
    }
    const uint32_t swt = sw(jt);
    V v[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = ys[swt ^ sw(jr[r])];
    const int g1 = S.first + S.count;
    if (S.fast) {
      // rounds [first, first + count): straight-line gate code joined by
      // plain if-then branches, so each gate's results overwrite its inputs
      for (int ri = S.first; ri < g1; ++ri) tile_round<V, R, NR, RB>(v, rounds + ri, gates, jt);
    } else {
      // the next gate's record is fetched while the current one is applied
      RegGate cur;
      if (S.first < g1) cur = gates[S.first];
      for (int gi = S.first; gi < g1; ++gi) {
        RegGate nxtg;
        if (gi + 1 < g1) nxtg = gates[gi + 1];
        reg_apply<V, R, NR, RB>(v, cur, jt, jr);
        cur = nxtg;
      }
    }
    // this thread read exactly these positions: write them back in place
#pragma unroll
    for (int r = 0; r < NR; ++r) ys[swt ^ sw(jr[r])] = v[r];
  }
}

// ---- FP64 tile part kernel ----------------------------------------------------
// The part path's exact arithmetic (one FP64 sweep per gate, as the
// reference's execute_on) on the tile layout of the BF16-split kernels: a
// 2^KT-amplitude tile (the part's qubits, slots 0-4 and the lowest free
// slots) is loaded with 512-byte runs per wave into LDS, the part's gates
// run on it segment by segment in registers (each of the NTH threads holds
// the 2^RB amplitudes of the segment's RB = KT - log2 NTH register bits, as
// in the register kernel: reg_apply), and the tile is stored the same way.
// One workgroup per tile: two per CU at KT = 12, NTH = 512 in complex128
// (64 KB of LDS each), four at KT = 11, NTH = 256.
struct TileTabs {
  uint64_t alo[64], ahi[256];   // address offset of tile bits 0-5 / 6-13
};

template <typename R, int KT, int NTH>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(4))) void k_part_tile(
    typename VecT<R>::type* __restrict__ a_vec, const TileTabs* __restrict__ tabs, const RegSeg* __restrict__ segs,
    int nseg, const TileRound* __restrict__ rounds, const RegGate* __restrict__ gates, uint64_t ntiles,
    uint64_t outer_mask, uint64_t top) {
  using V = cplx<R>;
  using G = typename VecT<R>::type;
  constexpr int W = NTH / 64, T = 1 << KT, PER = T / NTH;
  constexpr int RB = KT - tile_log2(NTH);
  static_assert(RB >= 1 && RB <= 5 && (T >> 6) <= 256, "tile shape");
  __shared__ V ys[T];
  __shared__ uint64_t alo[64], ahi[T >> 6];
  // LDS bank swizzle (tile_sw): tile element t at t ^ (((t >> 4) ^ (t >> 8))
  // & 15), so that the 16-byte elements a wave reads (its lanes spanning 6
  // tile bits) spread over the 16 four-bank groups for most choices of
  // register bits
  auto sw = tile_sw;
  for (int i = threadIdx.x; i < 64; i += NTH) alo[i] = tabs->alo[i];
  for (int i = threadIdx.x; i < (T >> 6); i += NTH) ahi[i] = tabs->ahi[i];
  __syncthreads();
  const uint64_t t0 = blockIdx.x;
  if (t0 >= ntiles) return;  // uniform across the workgroup
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = tile_base(t0, outer_mask, top) + alo[lane];
  {
    // tile position t = x + NTH j: its lane bits select alo, the rest ahi
    G raw[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) raw[j] = ld_stream<R>(&a_vec[base + ahi[wave + W * j]]);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      V v;
      v.x = raw[j].x;
      v.y = raw[j].y;
      ys[sw(threadIdx.x + NTH * j)] = v;
    }
  }
  tile_segments<R, KT, NTH>(ys, segs, nseg, rounds, gates);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const V v = ys[sw(threadIdx.x + NTH * j)];
    G o;
    o.x = v.x;
    o.y = v.y;
    st_stream<R>(&a_vec[base + ahi[wave + W * j]], o);
  }
}

}  // namespace hsv
