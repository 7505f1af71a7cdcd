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
template <typename V, typename R>
__device__ __forceinline__ V cmac(V acc, V a, R bx, R by) {
  acc.x += a.x * bx - a.y * by;
  acc.y += a.x * by + a.y * bx;
  return acc;
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

template <typename R, int KIND, int ITEMS>
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
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) v[j] = a[idx[j]];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) a[idx[j]] = cmul<V, R>(v[j], m[0], m[1]);
    } else if constexpr (KIND == GK_DIAG1 || KIND == GK_XSWAP1 || KIND == GK_DENSE1) {
      V v0[ITEMS], v1[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) { v0[j] = a[idx[j]]; v1[j] = a[idx[j] + g.off1]; }
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
        a[idx[j]] = o0;
        a[idx[j] + g.off1] = o1;
      }
    } else if constexpr (KIND == GK_SWAP2) {
      V v0[ITEMS], v1[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) { v0[j] = a[idx[j] + g.off1]; v1[j] = a[idx[j] + g.off2]; }
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) if (ok[j]) { a[idx[j] + g.off1] = v1[j]; a[idx[j] + g.off2] = v0[j]; }
    } else {  // GK_DIAG2, GK_DENSE2: v[2*b(t0) + b(t1)]
      V v[ITEMS][4];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        if (!ok[j]) continue;
        v[j][0] = a[idx[j]];
        v[j][1] = a[idx[j] + g.off2];
        v[j][2] = a[idx[j] + g.off1];
        v[j][3] = a[idx[j] + g.off1 + g.off2];
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
        a[idx[j]] = o[0];
        a[idx[j] + g.off2] = o[1];
        a[idx[j] + g.off1] = o[2];
        a[idx[j] + g.off1 + g.off2] = o[3];
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
      } else {                 // RD_MAXDIFF
        const double dx = x - (double)b[i].x, dy = y - (double)b[i].y;
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

// ---- fill |0...0> -----------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_zero_state(typename VecT<R>::type* __restrict__ a, uint64_t len) {
  using V = typename VecT<R>::type;
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < len; i += stride)
    a[i] = cmk<V, R>(i == 0 ? R(1) : R(0), R(0));
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

}  // namespace hsv
