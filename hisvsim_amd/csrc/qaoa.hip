// Batched QAOA MaxCut on one MI355X: many instances of uneven qubit counts,
// complex64 states, FP64 expectation values, one hipGraph per batch.
//
// Circuit per instance (n qubits, graph E, p layers): H on every qubit, then
// per layer ZZ(gamma) on every edge -- exp(-i gamma/2 Z_u Z_v), the reference's
// RZZ (basic_gates.hpp:1249-1252) -- and RX(2 beta) on every qubit; output
// <C> with C = sum_edges (1 - Z_u Z_v) / 2.
//
// Schedule (all instances advance together, one kernel launch per stage):
//  * init: every amplitude = 2^(-n/2) (the state after the H layer);
//  * stage A, per layer: contiguous chunks of 2^L amplitudes (the L low
//    qubits) in LDS: multiply by the layer's diagonal
//    prod_edges ZZ = exp(-i gamma/2 (|E| - 2 cut(x))), looked up by cut(x)
//    (cut(x) itself from a per-instance table over the low qubits and two
//    per-chunk tables for the edges that reach the chunk's fixed high bits),
//    then RX(2 beta) on each low qubit, three qubits per LDS sweep in registers;
//  * stage B, per layer, instances with n > L: chunks over the high qubits
//    plus the 3 to 6 lowest (contiguous 64 to 512 B runs, as many as a
//    2^13-amplitude chunk leaves room for), RX(2 beta) on each high qubit;
//  * reduce: sum_x |a_x|^2 cut(x) per stage-A chunk in FP64, then per instance.
// The stage kernels are persistent: a workgroup walks the flattened list of
// (instance, chunk) items of its stage. The whole sequence (plus the copy of
// the results to pinned host memory) is captured once into a hipGraph and
// replayed for each evaluation; angles are read from device memory so new
// parameters need no re-capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hsv.h"

// Shared with hsv.hip so that hsv_last_error() reports QAOA failures too.
void hsv_set_last_error(const std::string& msg);

namespace {

int qfail(int code, const std::string& m) {
  hsv_set_last_error(m);
  return code;
}
#define QCHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return qfail(HSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kMaxEdges = 64;
constexpr int kMaxQaoaQubits = 26;
// Stage B chunks hold an instance's hb = n - L high qubits and its lowest
// low_contig(hb) qubits (contiguous runs of 2^that amplitudes), at most
// kStageBBits qubits in all: from kLowContig (64-byte runs) up to
// kMaxLowContig (512-byte runs) for the narrower instances.
constexpr int kLowContig = 3;
constexpr int kMaxLowContig = 6;
constexpr int kStageBBits = 13;
__host__ __device__ inline int low_contig(int hb) {
  return kStageBBits - hb < kMaxLowContig ? kStageBBits - hb : kMaxLowContig;
}

struct Inst {
  uint64_t off;      // first amplitude in the batch buffer
  int n;             // qubits
  int ne;            // edges
  int edge_off;      // first edge in the edge arrays
  int p;             // layers
  int angle_off;     // first angle in gamma/beta
  int pad;
};

// One work list per stage: item prefix per instance (cum[i] = first item of
// instance i) over the instances that take part.
struct Stage {
  int num;           // instances in the list
  int first;         // offset into the device index arrays
  uint64_t items;    // total items
};

struct cf { float x, y; };

// Cut sizes of a chunk x = h + l (h: the fixed bits from B up, l < 2^B, B <= 13)
// from tables: cut(x) = low[l] + T1[l & 63] + T2[l >> 6], where low[] (host,
// per instance) counts the cut edges with both ends below B, and the chunk's
// tables hold the rest: an edge with both ends at or above B adds the
// constant h_u ^ h_v; an edge from a low qubit a to a high qubit c adds
// l_a ^ h_c = h_c + l_a (1 - 2 h_c), a constant plus a weight d_a on bit a.
// T1 sums the weights of bits 0-5 of its index, T2 those of bits 6-12 plus
// the constant. All threads of the workgroup take part (two barriers).
template <int NT>
__device__ __forceinline__ void chunk_cut_tables(uint64_t h, int B, const unsigned char* su, const unsigned char* sv,
                                                 int ne, int* sd, short* T1, short* T2) {
  for (int t = threadIdx.x; t <= B; t += NT) {
    // t < B: weight d_t of low bit t; t == B: the constant
    int acc = 0;
    for (int e = 0; e < ne; ++e) {
      int a = su[e], c = sv[e];
      if (a >= B) { const int x = a; a = c; c = x; }   // a: the lower end if any end is low
      const int hc = (int)((h >> c) & 1u);
      if (a < B && c >= B) {
        if (t == a) acc += 1 - 2 * hc;
        else if (t == B) acc += hc;
      } else if (a >= B && t == B) {
        acc += (int)(((h >> a) ^ (h >> c)) & 1u);
      }
    }
    sd[t] = acc;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 64 + 128; t += NT) {
    int acc = 0;
    if (t < 64) {
      for (int u = 0; u < 6 && u < B; ++u)
        if (t >> u & 1) acc += sd[u];
      T1[t] = (short)acc;
    } else {
      const int i = t - 64;
      acc = sd[B];
      for (int u = 6; u < B; ++u)
        if (i >> (u - 6) & 1) acc += sd[u];
      T2[i] = (short)acc;
    }
  }
  __syncthreads();
}

// LDS position of chunk element i: bits 0-2 XOR bits 5-7 and bits 3-4 XOR
// bits 6-7, so that the 64 lanes of a wave reading 8-byte elements spread
// over all bank pairs, for consecutive elements as for the three-bit
// register groups of rx_block (a bijection on [0, 2^bits) for any bits).
__device__ __forceinline__ uint32_t qsw(uint32_t i) { return i ^ ((i >> 5) & 7u) ^ (((i >> 6) & 3u) << 3); }

// RX(2 beta) = [[cos b, -i sin b], [-i sin b, cos b]] on a pair.
__device__ __forceinline__ void rx_pair(cf& a, cf& b, float c, float sn) {
  const cf x = a, y = b;
  a.x = c * x.x + sn * y.y;   // c a - i s b
  a.y = c * x.y - sn * y.x;
  b.x = sn * x.y + c * y.x;   // -i s a + c b
  b.y = -sn * x.x + c * y.y;
}

// RX on LDS bit q of a 2^bits block, one pair per thread and step.
__device__ __forceinline__ void rx_lds(cf* s, int bits, int q, float c, float sn, int nt) {
  const uint32_t half = 1u << (bits - 1);
  for (uint32_t w = threadIdx.x; w < half; w += nt) {
    const uint32_t lo = w & ((1u << q) - 1);
    const uint32_t i0 = ((w ^ lo) << 1) | lo;
    const uint32_t i1 = i0 | (1u << q);
    cf a = s[qsw(i0)], b = s[qsw(i1)];
    rx_pair(a, b, c, sn);
    s[qsw(i0)] = a;
    s[qsw(i1)] = b;
  }
}

template <int S>
__device__ __forceinline__ void rx_regs(cf* v, float c, float sn) {
#pragma unroll
  for (int r = 0; r < 8; ++r)
    if (!(r & S)) rx_pair(v[r], v[r + S], c, sn);
}

// RX on bits [from, bits) of a 2^bits LDS block, three bits per sweep: each
// thread takes the 8 elements of three consecutive bits q0..q0+2 into
// registers, applies RX on those of them not done yet and writes them back;
// one barrier per sweep (the block must be in LDS, after a barrier, on entry).
template <int NT>
__device__ __forceinline__ void rx_block(cf* s, int bits, int from, float c, float sn) {
  if (bits < 3) {
    for (int q = from; q < bits; ++q) {
      rx_lds(s, bits, q, c, sn, NT);
      __syncthreads();
    }
    return;
  }
  for (int done = from; done < bits;) {
    const int q0 = done < bits - 3 ? done : bits - 3;
    const int first = done - q0;  // 0..2: register bits below it are done
    const uint32_t nw = 1u << (bits - 3);
    for (uint32_t w = threadIdx.x; w < nw; w += NT) {
      const uint32_t lo = w & ((1u << q0) - 1);
      const uint32_t base = ((w ^ lo) << 3) | lo;
      cf v[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = s[qsw(base + ((uint32_t)r << q0))];
      if (first == 0) rx_regs<1>(v, c, sn);
      if (first <= 1) rx_regs<2>(v, c, sn);
      rx_regs<4>(v, c, sn);
#pragma unroll
      for (int r = 0; r < 8; ++r) s[qsw(base + ((uint32_t)r << q0))] = v[r];
    }
    __syncthreads();
    done = q0 + 3;
  }
}

__device__ __forceinline__ int find_inst(const uint64_t* cum, int num, uint64_t item) {
  int lo = 0, hi = num - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cum[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

constexpr int kInitItem = 4096;  // amplitudes per item of k_qaoa_init

__global__ __launch_bounds__(256) void k_qaoa_init(cf* __restrict__ st, const Inst* __restrict__ inst,
                                                   const int* __restrict__ ids, const uint64_t* __restrict__ cum,
                                                   int num, uint64_t items) {
  // item = kInitItem consecutive amplitudes (fewer for instances below 12 qubits)
  for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int k = find_inst(cum, num, it);
    const Inst I = inst[ids[k]];
    const uint64_t n_amp = 1ull << I.n;
    const uint64_t lo = (it - cum[k]) * kInitItem;
    cf v;
    v.x = (float)(1.0 / std::sqrt((double)n_amp));
    v.y = 0.0f;
    for (uint64_t j = lo + threadIdx.x; j < lo + kInitItem && j < n_amp; j += 256) st[I.off + j] = v;
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_qaoa_low(cf* __restrict__ st, const Inst* __restrict__ inst,
                                                 const int* __restrict__ ids, const uint64_t* __restrict__ cum,
                                                 int num, uint64_t items, int L, int layer,
                                                 const unsigned char* __restrict__ eu,
                                                 const unsigned char* __restrict__ ev,
                                                 const unsigned char* __restrict__ cut_low,
                                                 const uint64_t* __restrict__ cut_off,
                                                 const double* __restrict__ gamma, const double* __restrict__ beta) {
  extern __shared__ __align__(16) unsigned char smem[];
  cf* s = reinterpret_cast<cf*>(smem);
  __shared__ float ph_x[kMaxEdges + 1], ph_y[kMaxEdges + 1];
  __shared__ unsigned char su[kMaxEdges], sv[kMaxEdges];
  __shared__ int sd[14];
  __shared__ short T1[64], T2[128];
  int cur = -1;
  for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int k = find_inst(cum, num, it);
    const int id = ids[k];
    const Inst I = inst[id];
    const int bits = I.n < L ? I.n : L;
    const uint64_t base = I.off + ((it - cum[k]) << bits);
    const uint64_t gbase = (it - cum[k]) << bits;  // index of the chunk's first amplitude in the instance
    __syncthreads();
    if (id != cur) {  // per-instance tables: the layer's phase by cut size, the edges
      cur = id;
      const double g = gamma[I.angle_off + layer];
      for (int c = threadIdx.x; c <= I.ne; c += NT) {
        const double a = -0.5 * g * (double)(I.ne - 2 * c);
        ph_x[c] = (float)std::cos(a);
        ph_y[c] = (float)std::sin(a);
      }
      for (int e = threadIdx.x; e < I.ne; e += NT) {
        su[e] = eu[I.edge_off + e];
        sv[e] = ev[I.edge_off + e];
      }
      __syncthreads();
    }
    chunk_cut_tables<NT>(gbase, bits, su, sv, I.ne, sd, T1, T2);
    const unsigned char* low = cut_low + cut_off[id];
    const uint32_t sz = 1u << bits;
    for (uint32_t j = threadIdx.x; j < sz; j += NT) {
      const cf a = st[base + j];
      const int c = (int)low[j] + T1[j & 63] + T2[j >> 6];
      cf o;
      o.x = a.x * ph_x[c] - a.y * ph_y[c];
      o.y = a.x * ph_y[c] + a.y * ph_x[c];
      s[qsw(j)] = o;
    }
    __syncthreads();
    const double b = beta[I.angle_off + layer];
    const float cb = (float)std::cos(b), sb = (float)std::sin(b);
    rx_block<NT>(s, bits, 0, cb, sb);
    for (uint32_t j = threadIdx.x; j < sz; j += NT) st[base + j] = s[qsw(j)];
  }
}

__device__ __forceinline__ uint64_t pdep_q(uint64_t x, uint64_t mask) {
  uint64_t r = 0;
  for (uint64_t bit = 1; mask; bit <<= 1) {
    const uint64_t low = mask & (~mask + 1);
    if (x & bit) r |= low;
    mask &= mask - 1;
  }
  return r;
}

template <int NT>
__global__ __launch_bounds__(NT) void k_qaoa_high(cf* __restrict__ st, const Inst* __restrict__ inst,
                                                  const int* __restrict__ ids, const uint64_t* __restrict__ cum,
                                                  int num, uint64_t items, int L, int layer,
                                                  const double* __restrict__ beta) {
  extern __shared__ __align__(16) unsigned char smem[];
  cf* s = reinterpret_cast<cf*>(smem);
  for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int k = find_inst(cum, num, it);
    const Inst I = inst[ids[k]];
    // chunk bits: the lowest lc qubits + the high qubits L..n-1
    const int hb = I.n - L;
    const int lc = low_contig(hb);
    const int bits = lc + hb;
    const uint64_t cmask = ((1ull << lc) - 1) | (((1ull << hb) - 1) << L);
    const uint64_t omask = ((1ull << I.n) - 1) & ~cmask;
    const uint64_t base = pdep_q(it - cum[k], omask);
    const uint32_t sz = 1u << bits;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < sz; j += NT) {
      const uint64_t idx = base + (j & ((1u << lc) - 1)) + ((uint64_t)(j >> lc) << L);
      s[qsw(j)] = st[I.off + idx];
    }
    __syncthreads();
    const double b = beta[I.angle_off + layer];
    const float cb = (float)std::cos(b), sb = (float)std::sin(b);
    rx_block<NT>(s, bits, lc, cb, sb);
    for (uint32_t j = threadIdx.x; j < sz; j += NT) {
      const uint64_t idx = base + (j & ((1u << lc) - 1)) + ((uint64_t)(j >> lc) << L);
      st[I.off + idx] = s[qsw(j)];
    }
  }
}

// sum |a|^2 cut(x) over the chunks of stage A (2^min(n, L) amplitudes), FP64,
// one partial per item
__global__ __launch_bounds__(256) void k_qaoa_cost(const cf* __restrict__ st, const Inst* __restrict__ inst,
                                                   const int* __restrict__ ids, const uint64_t* __restrict__ cum,
                                                   int num, uint64_t items, int L, const unsigned char* __restrict__ eu,
                                                   const unsigned char* __restrict__ ev,
                                                   const unsigned char* __restrict__ cut_low,
                                                   const uint64_t* __restrict__ cut_off, double* __restrict__ part) {
  __shared__ double red[256];
  __shared__ unsigned char su[kMaxEdges], sv[kMaxEdges];
  __shared__ int sd[14];
  __shared__ short T1[64], T2[128];
  int cur = -1;
  for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int k = find_inst(cum, num, it);
    const int id = ids[k];
    const Inst I = inst[id];
    __syncthreads();
    if (id != cur) {
      cur = id;
      for (int e = threadIdx.x; e < I.ne; e += 256) {
        su[e] = eu[I.edge_off + e];
        sv[e] = ev[I.edge_off + e];
      }
      __syncthreads();
    }
    const int bits = I.n < L ? I.n : L;
    const uint64_t lo = (it - cum[k]) << bits;
    chunk_cut_tables<256>(lo, bits, su, sv, I.ne, sd, T1, T2);
    const unsigned char* low = cut_low + cut_off[id];
    double acc = 0.0;
    for (uint32_t j = threadIdx.x; j < (1u << bits); j += 256) {
      const cf a = st[I.off + lo + j];
      acc += ((double)a.x * a.x + (double)a.y * a.y) * (int)(low[j] + T1[j & 63] + T2[j >> 6]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[it] = red[0];
  }
}

__global__ void k_qaoa_sum(const double* __restrict__ part, const uint64_t* __restrict__ cum, int num,
                           uint64_t items, double* __restrict__ out) {
  // one thread per instance, fixed order -> deterministic
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num) return;
  const uint64_t a = cum[i], b = (i + 1 < num) ? cum[i + 1] : items;
  double s = 0.0;
  for (uint64_t t = a; t < b; ++t) s += part[t];
  out[i] = s;
}

}  // namespace

struct hsv_qaoa_batch {
  int num = 0;
  std::vector<Inst> inst;
  std::vector<unsigned char> eu, ev;
  std::vector<int> layers;
  int pmax = 0;
  int total_angles = 0;
  uint64_t total_amps = 0;
  int L = 12;          // stage-A chunk bits (tuned)
  int nt = 256;        // threads per workgroup (tuned)
  int grid_per_cu = 4; // persistent workgroups per CU (tuned)
  int num_cu = 256;
  hipStream_t stream = nullptr;
  cf* state = nullptr;
  Inst* d_inst = nullptr;
  unsigned char *d_eu = nullptr, *d_ev = nullptr;
  double *d_gamma = nullptr, *d_beta = nullptr;
  // stage work lists (device): ids and cum for init, low(l), high(l), cost
  void* d_lists = nullptr;
  unsigned char* d_cut = nullptr;    // per instance: cut sizes of the edges among its low min(n, L) qubits
  uint64_t* d_cut_off = nullptr;     // per instance: offset of its table in d_cut
  double* d_part = nullptr;
  double* d_out = nullptr;
  double* h_out = nullptr;  // pinned
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float last_ms = 0;
};

namespace {

struct HostList {
  std::vector<int> ids;
  std::vector<uint64_t> cum;
  uint64_t items = 0;
};

template <class F>
HostList make_list(const hsv_qaoa_batch* b, F items_of) {
  HostList h;
  for (int i = 0; i < b->num; ++i) {
    const uint64_t c = items_of(b->inst[i]);
    if (!c) continue;
    h.ids.push_back(i);
    h.cum.push_back(h.items);
    h.items += c;
  }
  return h;
}

void destroy_graph(hsv_qaoa_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (b->d_lists) (void)hipFree(b->d_lists);
  if (b->d_part) (void)hipFree(b->d_part);
  if (b->d_cut) (void)hipFree(b->d_cut);
  if (b->d_cut_off) (void)hipFree(b->d_cut_off);
  b->d_lists = nullptr;
  b->d_part = nullptr;
  b->d_cut = nullptr;
  b->d_cut_off = nullptr;
}

template <int NT>
void launch_low(hsv_qaoa_batch* b, int grid, const int* ids, const uint64_t* cum, int num, uint64_t items, int layer) {
  const size_t lds = (size_t(1) << b->L) * sizeof(cf);
  hipLaunchKernelGGL((k_qaoa_low<NT>), grid, NT, lds, b->stream, b->state, b->d_inst, ids, cum, num, items, b->L,
                     layer, b->d_eu, b->d_ev, b->d_cut, b->d_cut_off, b->d_gamma, b->d_beta);
}
template <int NT>
void launch_high(hsv_qaoa_batch* b, int grid, const int* ids, const uint64_t* cum, int num, uint64_t items, int layer,
                 int maxbits) {
  const size_t lds = (size_t(1) << maxbits) * sizeof(cf);
  hipLaunchKernelGGL((k_qaoa_high<NT>), grid, NT, lds, b->stream, b->state, b->d_inst, ids, cum, num, items, b->L,
                     layer, b->d_beta);
}

// Builds the work lists and captures the whole batch into a hipGraph.
int build_graph(hsv_qaoa_batch* b) {
  destroy_graph(b);
  const int L = b->L;
  std::vector<HostList> lists;
  lists.push_back(make_list(b, [](const Inst& I) { return ((1ull << I.n) + kInitItem - 1) / kInitItem; }));  // init
  for (int l = 0; l < b->pmax; ++l) {
    lists.push_back(make_list(b, [&](const Inst& I) -> uint64_t {
      return I.p > l ? (1ull << I.n) >> std::min(I.n, L) : 0;
    }));
    lists.push_back(make_list(b, [&](const Inst& I) -> uint64_t {
      return (I.p > l && I.n > L) ? (1ull << I.n) >> (low_contig(I.n - L) + I.n - L) : 0;
    }));
  }
  lists.push_back(make_list(b, [&](const Inst& I) { return (1ull << I.n) >> std::min(I.n, L); }));  // cost (all)
  {
    // cut sizes over each instance's low min(n, L) qubits, from the edges
    // with both ends there (the kernels add the other edges per chunk)
    std::vector<uint64_t> coff(b->num);
    uint64_t tot = 0;
    for (int i = 0; i < b->num; ++i) {
      coff[i] = tot;
      tot += 1ull << std::min(b->inst[i].n, L);
    }
    std::vector<unsigned char> cut(tot, 0);
    for (int i = 0; i < b->num; ++i) {
      const Inst& I = b->inst[i];
      const int B = std::min(I.n, L);
      for (int e = 0; e < I.ne; ++e) {
        const int u = b->eu[I.edge_off + e], v = b->ev[I.edge_off + e];
        if (u >= B || v >= B) continue;
        for (uint64_t x = 0; x < (1ull << B); ++x) cut[coff[i] + x] += (unsigned char)(((x >> u) ^ (x >> v)) & 1u);
      }
    }
    QCHK(hipMalloc(reinterpret_cast<void**>(&b->d_cut), std::max<uint64_t>(tot, 1)));
    QCHK(hipMalloc(reinterpret_cast<void**>(&b->d_cut_off), sizeof(uint64_t) * b->num));
    QCHK(hipMemcpy(b->d_cut, cut.data(), tot, hipMemcpyHostToDevice));
    QCHK(hipMemcpy(b->d_cut_off, coff.data(), sizeof(uint64_t) * b->num, hipMemcpyHostToDevice));
  }
  // device layout per list: ids (int, padded to 8 B) | cum (u64)
  std::vector<size_t> offs;
  size_t bytes = 0;
  for (auto& h : lists) {
    offs.push_back(bytes);
    bytes += ((h.ids.size() * 4 + 7) / 8) * 8 + h.cum.size() * 8 + 8;
  }
  std::vector<unsigned char> host(bytes, 0);
  for (size_t i = 0; i < lists.size(); ++i) {
    std::memcpy(host.data() + offs[i], lists[i].ids.data(), lists[i].ids.size() * 4);
    std::memcpy(host.data() + offs[i] + ((lists[i].ids.size() * 4 + 7) / 8) * 8, lists[i].cum.data(),
                lists[i].cum.size() * 8);
  }
  QCHK(hipMalloc(&b->d_lists, bytes));
  QCHK(hipMemcpy(b->d_lists, host.data(), bytes, hipMemcpyHostToDevice));
  const HostList& cost = lists.back();
  QCHK(hipMalloc(reinterpret_cast<void**>(&b->d_part), std::max<uint64_t>(cost.items, 1) * sizeof(double)));
  auto ids_of = [&](size_t i) {
    return reinterpret_cast<const int*>(static_cast<unsigned char*>(b->d_lists) + offs[i]);
  };
  auto cum_of = [&](size_t i) {
    return reinterpret_cast<const uint64_t*>(static_cast<unsigned char*>(b->d_lists) + offs[i] +
                                             ((lists[i].ids.size() * 4 + 7) / 8) * 8);
  };
  int maxhigh = 0;
  for (const Inst& I : b->inst)
    if (I.n > L) maxhigh = std::max(maxhigh, low_contig(I.n - L) + I.n - L);
  const int grid = b->num_cu * b->grid_per_cu;

  QCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  {
    const HostList& h = lists[0];
    if (h.items)
      hipLaunchKernelGGL(k_qaoa_init, (int)std::min<uint64_t>(h.items, (uint64_t)b->num_cu * 8), 256, 0, b->stream,
                         b->state, b->d_inst, ids_of(0), cum_of(0), (int)h.ids.size(), h.items);
    for (int l = 0; l < b->pmax; ++l) {
      const size_t ia = 1 + 2 * l, ib = 2 + 2 * l;
      const HostList& A = lists[ia];
      const HostList& B = lists[ib];
      const int ga = (int)std::min<uint64_t>(A.items, grid);
      if (A.items) {
        if (b->nt == 128) launch_low<128>(b, ga, ids_of(ia), cum_of(ia), (int)A.ids.size(), A.items, l);
        else if (b->nt == 512) launch_low<512>(b, ga, ids_of(ia), cum_of(ia), (int)A.ids.size(), A.items, l);
        else launch_low<256>(b, ga, ids_of(ia), cum_of(ia), (int)A.ids.size(), A.items, l);
      }
      const int gb = (int)std::min<uint64_t>(B.items, grid);
      if (B.items) {
        if (b->nt == 128) launch_high<128>(b, gb, ids_of(ib), cum_of(ib), (int)B.ids.size(), B.items, l, maxhigh);
        else if (b->nt == 512) launch_high<512>(b, gb, ids_of(ib), cum_of(ib), (int)B.ids.size(), B.items, l, maxhigh);
        else launch_high<256>(b, gb, ids_of(ib), cum_of(ib), (int)B.ids.size(), B.items, l, maxhigh);
      }
    }
    const size_t ic = lists.size() - 1;
    const HostList& C = lists[ic];
    if (C.items) {
      hipLaunchKernelGGL(k_qaoa_cost, (int)std::min<uint64_t>(C.items, grid), 256, 0, b->stream, b->state, b->d_inst,
                         ids_of(ic), cum_of(ic), (int)C.ids.size(), C.items, L, b->d_eu, b->d_ev, b->d_cut,
                         b->d_cut_off, b->d_part);
      hipLaunchKernelGGL(k_qaoa_sum, (b->num + 255) / 256, 256, 0, b->stream, b->d_part, cum_of(ic),
                         (int)C.ids.size(), C.items, b->d_out);
    }
    (void)hipMemcpyAsync(b->h_out, b->d_out, sizeof(double) * b->num, hipMemcpyDeviceToHost, b->stream);
  }
  hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
  if (ce != hipSuccess) return qfail(HSV_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
  QCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return HSV_OK;
}

}  // namespace

extern "C" {

int hsv_qaoa_create(int num_instances, const int* num_qubits, const int* edge_offsets, const int* edges,
                    const int* layers, hsv_qaoa_batch** out) {
  if (!out || num_instances <= 0 || !num_qubits || !edge_offsets || !edges || !layers)
    return qfail(HSV_ERR_ARG, "bad arguments");
  *out = nullptr;
  auto* b = new hsv_qaoa_batch();
  b->num = num_instances;
  uint64_t off = 0;
  int aoff = 0;
  for (int i = 0; i < num_instances; ++i) {
    const int n = num_qubits[i];
    const int e0 = edge_offsets[i], e1 = edge_offsets[i + 1];
    if (n < 1 || n > kMaxQaoaQubits || e1 < e0 || e1 - e0 > kMaxEdges || layers[i] < 0) {
      delete b;
      return qfail(HSV_ERR_ARG, "instance " + std::to_string(i) + ": qubits, edges or layers out of range");
    }
    for (int e = e0; e < e1; ++e) {
      const int u = edges[2 * e], v = edges[2 * e + 1];
      if (u < 0 || v < 0 || u >= n || v >= n || u == v) {
        delete b;
        return qfail(HSV_ERR_ARG, "instance " + std::to_string(i) + ": bad edge");
      }
    }
    Inst I{off, n, e1 - e0, (int)b->eu.size(), layers[i], aoff, 0};
    for (int e = e0; e < e1; ++e) {
      b->eu.push_back((unsigned char)edges[2 * e]);
      b->ev.push_back((unsigned char)edges[2 * e + 1]);
    }
    b->inst.push_back(I);
    off += 1ull << n;
    aoff += layers[i];
    b->pmax = std::max(b->pmax, layers[i]);
  }
  b->total_amps = off;
  b->total_angles = aoff;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    delete b;
    return qfail(HSV_ERR_DEVICE, "no HIP device");
  }
  b->num_cu = prop.multiProcessorCount;
  if (hipStreamCreate(&b->stream) != hipSuccess) { delete b; return qfail(HSV_ERR_DEVICE, "stream"); }
  if (hipMalloc(reinterpret_cast<void**>(&b->state), off * sizeof(cf)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_inst), b->inst.size() * sizeof(Inst)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_eu), std::max<size_t>(b->eu.size(), 1)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_ev), std::max<size_t>(b->ev.size(), 1)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_gamma), std::max(aoff, 1) * sizeof(double)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_beta), std::max(aoff, 1) * sizeof(double)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_out), b->num * sizeof(double)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&b->h_out), b->num * sizeof(double), 0) != hipSuccess) {
    hsv_qaoa_destroy(b);
    return qfail(HSV_ERR_ALLOC, "cannot allocate the batch");
  }
  QCHK(hipMemcpy(b->d_inst, b->inst.data(), b->inst.size() * sizeof(Inst), hipMemcpyHostToDevice));
  if (!b->eu.empty()) {
    QCHK(hipMemcpy(b->d_eu, b->eu.data(), b->eu.size(), hipMemcpyHostToDevice));
    QCHK(hipMemcpy(b->d_ev, b->ev.data(), b->ev.size(), hipMemcpyHostToDevice));
  }
  QCHK(hipEventCreate(&b->e0));
  QCHK(hipEventCreate(&b->e1));
  *out = b;
  return HSV_OK;
}

int hsv_qaoa_destroy(hsv_qaoa_batch* b) {
  if (!b) return HSV_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  destroy_graph(b);
  (void)hipFree(b->state);
  (void)hipFree(b->d_inst);
  (void)hipFree(b->d_eu);
  (void)hipFree(b->d_ev);
  (void)hipFree(b->d_gamma);
  (void)hipFree(b->d_beta);
  (void)hipFree(b->d_out);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return HSV_OK;
}

int hsv_qaoa_tune(hsv_qaoa_batch* b, const char* key, int value) {
  if (!b || !key) return qfail(HSV_ERR_ARG, "null argument");
  const std::string k = key;
  if (k == "chunk_bits" && value >= 8 && value <= 13) b->L = value;
  else if (k == "threads" && (value == 128 || value == 256 || value == 512)) b->nt = value;
  else if (k == "grid_per_cu" && value >= 1 && value <= 64) b->grid_per_cu = value;
  else return qfail(HSV_ERR_ARG, "unknown QAOA tuning key or value: " + k);
  // stage B holds the high qubits plus the low ones; keep it within 64 KB of LDS
  int maxn = 0;
  for (const Inst& I : b->inst) maxn = std::max(maxn, I.n);
  if (kLowContig + maxn - b->L > kStageBBits) return qfail(HSV_ERR_LIMIT, "chunk_bits too small for the widest instance");
  destroy_graph(b);
  return HSV_OK;
}

int hsv_qaoa_set_angles(hsv_qaoa_batch* b, const double* gammas, const double* betas) {
  if (!b || (b->total_angles && (!gammas || !betas))) return qfail(HSV_ERR_ARG, "null argument");
  if (!b->total_angles) return HSV_OK;
  QCHK(hipMemcpyAsync(b->d_gamma, gammas, sizeof(double) * b->total_angles, hipMemcpyHostToDevice, b->stream));
  QCHK(hipMemcpyAsync(b->d_beta, betas, sizeof(double) * b->total_angles, hipMemcpyHostToDevice, b->stream));
  QCHK(hipStreamSynchronize(b->stream));
  return HSV_OK;
}

int hsv_qaoa_run(hsv_qaoa_batch* b, double* expectations) {
  if (!b || !expectations) return qfail(HSV_ERR_ARG, "null argument");
  int maxn = 0;
  for (const Inst& I : b->inst) maxn = std::max(maxn, I.n);
  if (kLowContig + maxn - b->L > kStageBBits) return qfail(HSV_ERR_LIMIT, "chunk_bits too small for the widest instance");
  if (!b->exec)
    if (int rc = build_graph(b)) return rc;
  QCHK(hipEventRecord(b->e0, b->stream));
  QCHK(hipGraphLaunch(b->exec, b->stream));
  QCHK(hipEventRecord(b->e1, b->stream));
  QCHK(hipStreamSynchronize(b->stream));
  QCHK(hipEventElapsedTime(&b->last_ms, b->e0, b->e1));
  std::memcpy(expectations, b->h_out, sizeof(double) * b->num);
  return HSV_OK;
}

int hsv_qaoa_last_ms(hsv_qaoa_batch* b, double* ms) {
  if (!b || !ms) return qfail(HSV_ERR_ARG, "null argument");
  *ms = b->last_ms;
  return HSV_OK;
}

int hsv_qaoa_read_state(hsv_qaoa_batch* b, int instance, double* amps) {
  if (!b || !amps || instance < 0 || instance >= b->num) return qfail(HSV_ERR_ARG, "bad instance");
  const Inst& I = b->inst[instance];
  std::vector<cf> tmp(1ull << I.n);
  QCHK(hipMemcpyAsync(tmp.data(), b->state + I.off, tmp.size() * sizeof(cf), hipMemcpyDeviceToHost, b->stream));
  QCHK(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < tmp.size(); ++i) {
    amps[2 * i] = tmp[i].x;
    amps[2 * i + 1] = tmp[i].y;
  }
  return HSV_OK;
}

int hsv_qaoa_total_amplitudes(hsv_qaoa_batch* b, uint64_t* out) {
  if (!b || !out) return qfail(HSV_ERR_ARG, "null argument");
  *out = b->total_amps;
  return HSV_OK;
}

}  // extern "C"
