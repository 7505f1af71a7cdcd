// Batched small circuits: many circuits (or parameter sets of one circuit) of
// uneven qubit counts, all states in one HBM buffer, run as one hipGraph.
//
// The reference runs one circuit per call (qasm_assembler_standalone.py:
// 1377-1617, per-gate; 1768-2030, per part). Here each instance's gate list
// is cut on the host into segments whose qubits fit a chunk of 2^L
// amplitudes (L = chunk_bits, tuned): an instance of at most L qubits is one
// chunk and one segment, its whole circuit applied in LDS in one pass; a
// wider instance gets the DAG-aware greedy segments of libhsv's part path
// (a skipped gate blocks its qubits for the rest of the scan, so the segment
// order is a topological order of the gate DAG). Stage s applies the s-th
// segment of every instance that has one: a persistent kernel walks the
// flattened (instance, chunk) list of the stage, gathers each chunk (the
// segment's qubits plus the instance's lowest other qubits) into LDS,
// applies the segment's gates there and scatters it back. A chunk of 8-12
// qubits (13 in complex64) runs the FP64 tile kernel's plan (gates in
// registers, one LDS exchange per register layout) when the thread count
// leaves it 2..5 register bits; other chunks sweep LDS once per gate. The stage launches (after a fill kernel that sets
// every instance to |0...0>) are captured once into a hipGraph; gate
// matrices live in device memory, so new parameters for the same gate
// structure replay the same graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hsv.h"
#include "gates.hpp"
#include "kernels.hpp"
#include "tileplan.hpp"

void hsv_set_last_error(const std::string& msg);

using namespace hsv;

namespace {

int bfail(int code, const std::string& m) {
  hsv_set_last_error(m);
  return code;
}
#define BCHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return bfail(HSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kBatchMaxQubits = 30;
constexpr int kBatchMaxChunk = 13;

// One segment of one instance.
struct BSeg {
  uint64_t amp_off;     // the instance's first amplitude in the batch buffer
  uint64_t chunk_mask;  // instance qubits held in a chunk
  uint64_t outer_mask;  // the others (one chunk per assignment of them)
  int bits;             // popcount(chunk_mask)
  int gate_first;       // first PartGate record
  int gate_count;
  int tseg_first;       // chunks with a tile plan: its first register segment
  int tseg_count;       // -1: the gates sweep the chunk in LDS one by one
  int pad;
};

// A chunk of KT qubits runs the FP64 tile kernel's register segments
// (tile_segments: each thread holds 2^RB amplitudes, RB = KT - log2 threads,
// one LDS exchange per segment instead of one LDS sweep per gate) when RB is
// 2..5 (a two-target gate needs both targets in registers); other chunks
// sweep LDS once per gate.
template <typename R, int KT, int BS>
constexpr bool tile_chunk_ok() {
  return KT - tile_log2(BS) >= 2 && KT - tile_log2(BS) <= 5 && (sizeof(R) == 4 || KT <= 12);
}

// Stage work list: the segments taking part (seg_ids) and the first
// flattened item of each (cum).
struct StageList {
  std::vector<int> seg_ids;
  std::vector<uint64_t> cum;
  uint64_t items = 0;
};

__device__ __forceinline__ int find_seg(const uint64_t* cum, int num, uint64_t item) {
  int lo = 0, hi = num - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cum[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// |0...0> for every instance: zero fill (a kernel, not a memset node, in the
// captured graph), then amplitude 0 of each instance set to 1
template <typename R>
__global__ void k_batch_zero(typename VecT<R>::type* __restrict__ st, uint64_t total) {
  using V = typename VecT<R>::type;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += uint64_t(gridDim.x) * blockDim.x) {
    V v;
    v.x = 0;
    v.y = 0;
    st[i] = v;
  }
}

template <typename R>
__global__ void k_batch_one(typename VecT<R>::type* __restrict__ st, const uint64_t* __restrict__ starts, int num) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < num) {
    typename VecT<R>::type v;
    v.x = 1;
    v.y = 0;
    st[starts[i]] = v;
  }
}

template <typename R, int BS>
__global__ __launch_bounds__(BS) void k_batch_stage(typename VecT<R>::type* __restrict__ st,
                                                    const BSeg* __restrict__ segs, const int* __restrict__ ids,
                                                    const uint64_t* __restrict__ cum, int num, uint64_t items,
                                                    const PartGate* __restrict__ gates,
                                                    const RegSeg* __restrict__ tsegs,
                                                    const TileRound* __restrict__ trounds,
                                                    const RegGate* __restrict__ tgates) {
  using V = typename VecT<R>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  __shared__ uint64_t tlo[64];
  __shared__ uint64_t thi[1 << (kBatchMaxChunk - 6)];
  int cur = -1;
  for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int k = find_seg(cum, num, it);
    const int sid = ids[k];
    const BSeg S = segs[sid];
    __syncthreads();  // the previous item's scatter has read lds and the tables
    if (sid != cur) {
      cur = sid;
      const int klo = S.bits < 6 ? S.bits : 6;
      for (int j = threadIdx.x; j < 64; j += BS) tlo[j] = j < (1 << klo) ? pdep64((uint64_t)j, S.chunk_mask) : 0;
      for (int j = threadIdx.x; j < (1 << (S.bits - klo)); j += BS)
        thi[j] = pdep64((uint64_t)j << klo, S.chunk_mask);
      __syncthreads();
    }
    const uint64_t base = S.amp_off + pdep64(it - cum[k], S.outer_mask);
    const uint32_t nin = 1u << S.bits;
    if constexpr (BS >= 128) {
      if (S.tseg_count >= 0) {  // uniform: a chunk with a tile plan
        cplx<R>* ys = reinterpret_cast<cplx<R>*>(smem_raw);
        for (uint32_t j = threadIdx.x; j < nin; j += BS) {
          const V a = st[base + tlo[j & 63] + thi[j >> 6]];
          cplx<R> c;
          c.x = a.x;
          c.y = a.y;
          ys[tile_sw(j)] = c;
        }
        const RegSeg* ts = tsegs + S.tseg_first;
        const int ns = S.tseg_count;
        switch (S.bits) {  // the host plans only the widths tile_chunk_ok accepts
          case 8: if constexpr (tile_chunk_ok<R, 8, BS>()) tile_segments<R, 8, BS>(ys, ts, ns, trounds, tgates); break;
          case 9: if constexpr (tile_chunk_ok<R, 9, BS>()) tile_segments<R, 9, BS>(ys, ts, ns, trounds, tgates); break;
          case 10: if constexpr (tile_chunk_ok<R, 10, BS>()) tile_segments<R, 10, BS>(ys, ts, ns, trounds, tgates); break;
          case 11: if constexpr (tile_chunk_ok<R, 11, BS>()) tile_segments<R, 11, BS>(ys, ts, ns, trounds, tgates); break;
          case 12: if constexpr (tile_chunk_ok<R, 12, BS>()) tile_segments<R, 12, BS>(ys, ts, ns, trounds, tgates); break;
          case 13: if constexpr (tile_chunk_ok<R, 13, BS>()) tile_segments<R, 13, BS>(ys, ts, ns, trounds, tgates); break;
          default: break;
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < nin; j += BS) {
          const cplx<R> c = ys[tile_sw(j)];
          V o;
          o.x = c.x;
          o.y = c.y;
          st[base + tlo[j & 63] + thi[j >> 6]] = o;
        }
        continue;
      }
    }
    for (uint32_t j = threadIdx.x; j < nin; j += BS) lds[j] = st[base + tlo[j & 63] + thi[j >> 6]];
    __syncthreads();
    for (int g = 0; g < S.gate_count; ++g) {
      part_gate_lds<R, BS>(lds, gates[S.gate_first + g]);
      __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j < nin; j += BS) st[base + tlo[j & 63] + thi[j >> 6]] = lds[j];
  }
}

// <Z_mask> per instance: sum_x |a_x|^2 (-1)^popcount(x & mask) in FP64 over
// chunks of 4096 amplitudes (one partial per chunk), then per instance in a
// fixed order.
template <typename R>
__global__ __launch_bounds__(256) void k_batch_z(const typename VecT<R>::type* __restrict__ st,
                                                 const uint64_t* __restrict__ starts, const int* __restrict__ nq,
                                                 const uint64_t* __restrict__ masks, const uint64_t* __restrict__ cum,
                                                 int num, uint64_t items, double* __restrict__ part) {
  __shared__ double red[256];
  for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int k = find_seg(cum, num, it);
    const uint64_t len = 1ull << nq[k];
    const uint64_t lo = (it - cum[k]) * 4096;
    const uint64_t m = masks[k];
    double acc = 0.0;
    for (uint64_t j = lo + threadIdx.x; j < lo + 4096 && j < len; j += 256) {
      const typename VecT<R>::type a = st[starts[k] + j];
      const double p = (double)a.x * a.x + (double)a.y * a.y;
      acc += (__popcll(j & m) & 1) ? -p : p;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[it] = red[0];
    __syncthreads();
  }
}

__global__ void k_batch_zsum(const double* __restrict__ part, const uint64_t* __restrict__ cum, int num,
                             uint64_t items, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num) return;
  const uint64_t a = cum[i], b = (i + 1 < num) ? cum[i + 1] : items;
  double s = 0.0;
  for (uint64_t t = a; t < b; ++t) s += part[t];
  out[i] = s;
}

uint64_t gate_mask(const Lowered& L) {
  uint64_t m = 0;
  if (L.kind == K_IDENT) return 0;
  for (int i = 0; i < L.nt; ++i) m |= 1ull << L.t[i];
  for (int i = 0; i < L.nc; ++i) m |= 1ull << L.c[i];
  return m;
}

// A gate's sweep over a chunk (the LDS part kernel's record): fixed bit
// positions, forced-one mask, target offsets and work items, with qubit q at
// chunk bit inner[q].
PartGate part_gate_of(const Lowered& L, int bits, const int* inner) {
  PartGate G{};
  uint64_t fixedmask = 0, set = 0;
  for (int i = 0; i < L.nc; ++i) {
    fixedmask |= 1ull << inner[L.c[i]];
    set |= 1ull << inner[L.c[i]];
  }
  for (int i = 0; i < L.nt; ++i) fixedmask |= 1ull << inner[L.t[i]];
  if (L.kind == K_PHASE1) set |= 1ull << inner[L.t[0]];
  int nf = 0;
  for (int b = 0; b < 64; ++b)
    if (fixedmask >> b & 1) G.fixed[nf++] = b;
  G.nfixed = nf;
  G.kind = L.kind;
  G.setmask = (uint32_t)set;
  G.off1 = 1u << inner[L.t[0]];
  G.off2 = L.nt == 2 ? 1u << inner[L.t[1]] : 0;
  G.nwork = 1u << (bits - nf);
  for (int i = 0; i < 16; ++i) {
    G.m[2 * i] = L.m[i].real();
    G.m[2 * i + 1] = L.m[i].imag();
  }
  return G;
}

}  // namespace

struct hsv_batch {
  int num = 0;
  int dtype = HSV_C128;
  std::vector<int> nq;
  std::vector<int> gate_off;          // num + 1
  std::vector<hsv_gate> gates;        // as given
  std::vector<uint64_t> starts;       // first amplitude per instance
  uint64_t total = 0;
  int L = 12;                         // chunk bits (tuned)
  int nt = 256;                       // threads per workgroup (tuned)
  int grid_per_cu = 8;                // persistent workgroups per CU (tuned)
  int num_cu = 256;
  hipStream_t stream = nullptr;
  void* state = nullptr;
  // plan (rebuilt when L changes): segments, their gate records, stage lists
  std::vector<BSeg> segs;
  std::vector<int> rec_gate;          // source gate of each PartGate record
  std::vector<std::vector<int>> rec_inner;  // qubit -> chunk bit of each record's segment
  std::vector<StageList> stages;
  BSeg* d_segs = nullptr;
  PartGate* d_gates = nullptr;
  // tile plans of the segments that have one, one device buffer:
  // register segments | rounds | gate records | dense two-qubit matrices
  void* d_tile = nullptr;
  size_t tile_bytes = 0;
  size_t tile_nseg = 0, tile_nround = 0, tile_ngate = 0, tile_nmat = 0;
  void* d_lists = nullptr;
  std::vector<size_t> list_off;
  uint64_t* d_starts = nullptr;
  int* d_nq = nullptr;
  // Z expectations
  uint64_t* d_masks = nullptr;
  uint64_t* d_zcum = nullptr;
  uint64_t z_items = 0;
  double* d_part = nullptr;
  double* d_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float last_ms = 0;
  size_t elem() const { return dtype == HSV_C128 ? 16 : 8; }
};

namespace {

void drop_plan(hsv_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  (void)hipFree(b->d_segs);
  (void)hipFree(b->d_gates);
  (void)hipFree(b->d_lists);
  (void)hipFree(b->d_tile);
  b->d_segs = nullptr;
  b->d_gates = nullptr;
  b->d_lists = nullptr;
  b->d_tile = nullptr;
  b->tile_bytes = 0;
  b->segs.clear();
  b->rec_gate.clear();
  b->rec_inner.clear();
  b->stages.clear();
}

// Lowers every gate, cuts each instance into segments and builds the gate
// records; fails on a bad gate before anything is uploaded.
int plan(hsv_batch* b, std::vector<PartGate>* recs) {
  std::vector<std::vector<int>> inst_segs(b->num);
  for (int i = 0; i < b->num; ++i) {
    const int n = b->nq[i];
    const int L = std::min(n, b->L);
    const int g0 = b->gate_off[i], g1 = b->gate_off[i + 1];
    std::vector<Lowered> Ls(g1 - g0);
    std::vector<uint64_t> qm(g1 - g0);
    for (int g = g0; g < g1; ++g) {
      std::string err;
      if (int rc = lower(b->gates[g], &Ls[g - g0], &err))
        return bfail(rc, "instance " + std::to_string(i) + ", gate " + std::to_string(g - g0) + ": " + err);
      const Lowered& Lw = Ls[g - g0];
      for (int j = 0; j < Lw.nt; ++j)
        if (Lw.t[j] >= n) return bfail(HSV_ERR_ARG, "instance " + std::to_string(i) + ": target out of range");
      for (int j = 0; j < Lw.nc; ++j)
        if (Lw.c[j] >= n) return bfail(HSV_ERR_ARG, "instance " + std::to_string(i) + ": control out of range");
      qm[g - g0] = gate_mask(Lw);
      if (__builtin_popcountll(qm[g - g0]) > L)
        return bfail(HSV_ERR_LIMIT, "instance " + std::to_string(i) + ": a gate spans more than chunk_bits qubits");
    }
    // DAG-aware greedy segments (identities take no qubits and ride along)
    std::vector<int> rem(g1 - g0);
    for (int g = 0; g < g1 - g0; ++g) rem[g] = g;
    const uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1);
    do {
      uint64_t Q = 0, blocked = 0;
      std::vector<int> taken, keep;
      for (int g : rem) {
        const uint64_t m = qm[g];
        if (m & blocked) {
          blocked |= m;
          keep.push_back(g);
        } else if (__builtin_popcountll(Q | m) <= L) {
          Q |= m;
          taken.push_back(g);
        } else {
          blocked |= m;
          keep.push_back(g);
        }
      }
      // the chunk: the segment's qubits plus the lowest other qubits up to L
      uint64_t chunk = Q;
      for (int q = 0; q < n && __builtin_popcountll(chunk) < L; ++q) chunk |= 1ull << q;
      int inner[64];
      for (int q = 0, r = 0; q < 64; ++q) inner[q] = (chunk >> q & 1) ? r++ : -1;
      BSeg S{};
      S.amp_off = b->starts[i];
      S.chunk_mask = chunk;
      S.outer_mask = all & ~chunk;
      S.bits = L;
      S.gate_first = (int)recs->size();
      for (int g : taken) {
        if (Ls[g].kind == K_IDENT) continue;
        recs->push_back(part_gate_of(Ls[g], L, inner));
        b->rec_gate.push_back(g0 + g);
        b->rec_inner.emplace_back(inner, inner + 64);
      }
      S.gate_count = (int)recs->size() - S.gate_first;
      inst_segs[i].push_back((int)b->segs.size());
      b->segs.push_back(S);
      rem = keep;
    } while (!rem.empty());
  }
  size_t nst = 0;
  for (auto& v : inst_segs) nst = std::max(nst, v.size());
  b->stages.assign(nst, StageList{});
  for (size_t s = 0; s < nst; ++s)
    for (int i = 0; i < b->num; ++i) {
      if (s >= inst_segs[i].size()) continue;
      const BSeg& S = b->segs[inst_segs[i][s]];
      if (S.gate_count == 0) continue;   // nothing to apply (an empty or identity-only circuit)
      StageList& st = b->stages[s];
      st.seg_ids.push_back(inst_segs[i][s]);
      st.cum.push_back(st.items);
      st.items += 1ull << (b->nq[i] - S.bits);
    }
  return HSV_OK;
}

// The tile plans of every segment whose chunk width k_batch_stage has a
// tile form for (8..12 qubits, 13 in complex64, 2..5 register bits at the
// batch's thread count, at least 128 threads), as one set of arrays with
// absolute indices; sets each segment's tseg_first/count (-1: LDS sweeps).
// all[g]: gate g of the batch, lowered.
struct TileArrays {
  std::vector<RegSeg> segs;
  std::vector<TileRound> rounds;
  std::vector<RegGate> gates;
  std::vector<int> m2idx;
  std::vector<double> mat2;
};

void build_tiles(hsv_batch* b, const std::vector<Lowered>& all, std::vector<BSeg>* segs, TileArrays* A) {
  int tb = 0;
  while ((1 << tb) < b->nt) ++tb;
  for (BSeg& S : *segs) {
    S.tseg_first = 0;
    S.tseg_count = -1;
    const int RB = S.bits - tb;
    // the widths k_batch_stage instantiates (tile_chunk_ok)
    const bool ok = S.bits >= 8 && S.bits <= (b->dtype == HSV_C128 ? 12 : 13) && RB >= 2 && RB <= 5 && b->nt >= 128;
    if (!ok || S.gate_count == 0) continue;
    std::vector<Lowered> Ls;
    for (int r = S.gate_first; r < S.gate_first + S.gate_count; ++r) Ls.push_back(all[b->rec_gate[r]]);
    TilePlan TP;
    if (build_tile_plan(Ls, b->rec_inner[S.gate_first].data(), S.bits, RB, &TP) != HSV_OK) continue;
    const int rb = (int)A->rounds.size(), gb = (int)A->gates.size(), mb = (int)(A->mat2.size() / 32);
    S.tseg_first = (int)A->segs.size();
    S.tseg_count = (int)TP.segs.size();
    for (RegSeg G : TP.segs) {
      G.first += G.fast ? rb : gb;
      A->segs.push_back(G);
    }
    for (TileRound Q : TP.rounds) {
      for (int& g : Q.g)
        if (g >= 0) g += gb;
      Q.sfirst += gb;
      A->rounds.push_back(Q);
    }
    for (size_t i = 0; i < TP.gates.size(); ++i) {
      A->gates.push_back(TP.gates[i]);
      A->m2idx.push_back(TP.m2idx[i] >= 0 ? TP.m2idx[i] + mb : -1);
    }
    A->mat2.insert(A->mat2.end(), TP.mat2.begin(), TP.mat2.end());
  }
}

// Writes the arrays into b->d_tile (allocated here when absent), the gate
// records' matrix pointers into the same buffer.
int upload_tiles(hsv_batch* b, const TileArrays& A) {
  const size_t seg_off = 0;
  const size_t round_off = seg_off + A.segs.size() * sizeof(RegSeg);
  const size_t gate_off = round_off + A.rounds.size() * sizeof(TileRound);
  const size_t mat_off = gate_off + A.gates.size() * sizeof(RegGate);
  const size_t bytes = mat_off + A.mat2.size() * sizeof(double);
  if (!b->d_tile) {
    BCHK(hipMalloc(&b->d_tile, std::max<size_t>(bytes, 16)));
    b->tile_bytes = bytes;
    b->tile_nseg = A.segs.size();
    b->tile_nround = A.rounds.size();
    b->tile_ngate = A.gates.size();
    b->tile_nmat = A.mat2.size();
  }
  std::vector<unsigned char> host(std::max<size_t>(bytes, 16), 0);
  auto* dev = static_cast<unsigned char*>(b->d_tile);
  if (!A.segs.empty()) std::memcpy(host.data() + seg_off, A.segs.data(), A.segs.size() * sizeof(RegSeg));
  if (!A.rounds.empty()) std::memcpy(host.data() + round_off, A.rounds.data(), A.rounds.size() * sizeof(TileRound));
  RegGate* hg = reinterpret_cast<RegGate*>(host.data() + gate_off);
  for (size_t i = 0; i < A.gates.size(); ++i) {
    hg[i] = A.gates[i];
    hg[i].m2 = A.m2idx[i] >= 0 ? reinterpret_cast<const double*>(dev + mat_off) + 32 * size_t(A.m2idx[i]) : nullptr;
  }
  if (!A.mat2.empty()) std::memcpy(host.data() + mat_off, A.mat2.data(), A.mat2.size() * sizeof(double));
  BCHK(hipMemcpyAsync(b->d_tile, host.data(), bytes, hipMemcpyHostToDevice, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return HSV_OK;
}

template <typename T>
const T* tile_part(const hsv_batch* b, size_t off) {
  return reinterpret_cast<const T*>(static_cast<const unsigned char*>(b->d_tile) + off);
}

template <typename R, int BS>
int stage_attr(hsv_batch* b) {
  const size_t lds = (size_t(1) << b->L) * sizeof(typename VecT<R>::type);
  BCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_stage<R, BS>),
                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return HSV_OK;
}

template <typename R>
int stage_attr_nt(hsv_batch* b) {
  if (b->nt == 64) return stage_attr<R, 64>(b);
  if (b->nt == 128) return stage_attr<R, 128>(b);
  if (b->nt == 512) return stage_attr<R, 512>(b);
  return stage_attr<R, 256>(b);
}

template <typename R, int BS>
void launch_stage(hsv_batch* b, const StageList& st, const int* ids, const uint64_t* cum) {
  const size_t lds = (size_t(1) << b->L) * sizeof(typename VecT<R>::type);
  const int grid = (int)std::min<uint64_t>(st.items, uint64_t(b->num_cu) * b->grid_per_cu);
  const size_t round_off = b->tile_nseg * sizeof(RegSeg);
  const size_t gate_off = round_off + b->tile_nround * sizeof(TileRound);
  hipLaunchKernelGGL((k_batch_stage<R, BS>), grid, BS, lds, b->stream,
                     static_cast<typename VecT<R>::type*>(b->state), b->d_segs, ids, cum, (int)st.seg_ids.size(),
                     st.items, b->d_gates, tile_part<RegSeg>(b, 0), tile_part<TileRound>(b, round_off),
                     tile_part<RegGate>(b, gate_off));
}

template <typename R>
void launch_stage_nt(hsv_batch* b, const StageList& st, const int* ids, const uint64_t* cum) {
  if (b->nt == 64) launch_stage<R, 64>(b, st, ids, cum);
  else if (b->nt == 128) launch_stage<R, 128>(b, st, ids, cum);
  else if (b->nt == 512) launch_stage<R, 512>(b, st, ids, cum);
  else launch_stage<R, 256>(b, st, ids, cum);
}

int build(hsv_batch* b) {
  drop_plan(b);
  std::vector<PartGate> recs;
  if (int rc = plan(b, &recs)) {
    drop_plan(b);
    return rc;
  }
  {
    std::vector<Lowered> all(b->gates.size());
    for (size_t g = 0; g < b->gates.size(); ++g) {
      std::string err;
      if (int rc = lower(b->gates[g], &all[g], &err)) return bfail(rc, "gate " + std::to_string(g) + ": " + err);
    }
    TileArrays A;
    build_tiles(b, all, &b->segs, &A);
    if (int rc = upload_tiles(b, A)) return rc;
  }
  BCHK(hipMalloc(reinterpret_cast<void**>(&b->d_segs), std::max<size_t>(b->segs.size(), 1) * sizeof(BSeg)));
  BCHK(hipMalloc(reinterpret_cast<void**>(&b->d_gates), std::max<size_t>(recs.size(), 1) * sizeof(PartGate)));
  BCHK(hipMemcpy(b->d_segs, b->segs.data(), b->segs.size() * sizeof(BSeg), hipMemcpyHostToDevice));
  if (!recs.empty())
    BCHK(hipMemcpy(b->d_gates, recs.data(), recs.size() * sizeof(PartGate), hipMemcpyHostToDevice));
  size_t bytes = 0;
  b->list_off.clear();
  for (const StageList& st : b->stages) {
    b->list_off.push_back(bytes);
    bytes += ((st.seg_ids.size() * 4 + 7) / 8) * 8 + st.cum.size() * 8 + 8;
  }
  std::vector<unsigned char> host(std::max<size_t>(bytes, 8), 0);
  for (size_t s = 0; s < b->stages.size(); ++s) {
    const StageList& st = b->stages[s];
    std::memcpy(host.data() + b->list_off[s], st.seg_ids.data(), st.seg_ids.size() * 4);
    std::memcpy(host.data() + b->list_off[s] + ((st.seg_ids.size() * 4 + 7) / 8) * 8, st.cum.data(),
                st.cum.size() * 8);
  }
  BCHK(hipMalloc(&b->d_lists, host.size()));
  BCHK(hipMemcpy(b->d_lists, host.data(), host.size(), hipMemcpyHostToDevice));
  if (int rc = b->dtype == HSV_C128 ? stage_attr_nt<double>(b) : stage_attr_nt<float>(b)) return rc;
  // capture: |0...0> for every instance, then the stages in order
  BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int zgrid = (int)std::min<uint64_t>((b->total + 255) / 256, uint64_t(b->num_cu) * 16);
  if (b->dtype == HSV_C128) {
    hipLaunchKernelGGL(k_batch_zero<double>, zgrid, 256, 0, b->stream, static_cast<double2*>(b->state), b->total);
    hipLaunchKernelGGL(k_batch_one<double>, (b->num + 255) / 256, 256, 0, b->stream, static_cast<double2*>(b->state),
                       b->d_starts, b->num);
  } else {
    hipLaunchKernelGGL(k_batch_zero<float>, zgrid, 256, 0, b->stream, static_cast<float2*>(b->state), b->total);
    hipLaunchKernelGGL(k_batch_one<float>, (b->num + 255) / 256, 256, 0, b->stream, static_cast<float2*>(b->state),
                       b->d_starts, b->num);
  }
  for (size_t s = 0; s < b->stages.size(); ++s) {
    const StageList& st = b->stages[s];
    if (!st.items) continue;
    const auto* base = static_cast<const unsigned char*>(b->d_lists) + b->list_off[s];
    const int* ids = reinterpret_cast<const int*>(base);
    const uint64_t* cum = reinterpret_cast<const uint64_t*>(base + ((st.seg_ids.size() * 4 + 7) / 8) * 8);
    if (b->dtype == HSV_C128) launch_stage_nt<double>(b, st, ids, cum);
    else launch_stage_nt<float>(b, st, ids, cum);
  }
  const hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
  if (ce != hipSuccess) return bfail(HSV_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
  BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return HSV_OK;
}

bool same_structure(const hsv_gate& a, const hsv_gate& b) {
  if (a.type != b.type || a.num_controls != b.num_controls || a.num_targets != b.num_targets) return false;
  for (int i = 0; i < a.num_controls; ++i)
    if (a.controls[i] != b.controls[i]) return false;
  for (int i = 0; i < a.num_targets; ++i)
    if (a.targets[i] != b.targets[i]) return false;
  return true;
}

}  // namespace

extern "C" {

int hsv_batch_create(int num_instances, const int* num_qubits, const int* gate_offsets, const hsv_gate* gates,
                     int dtype, hsv_batch** out) {
  if (!out || num_instances <= 0 || !num_qubits || !gate_offsets || (gate_offsets[num_instances] > 0 && !gates))
    return bfail(HSV_ERR_ARG, "bad arguments");
  if (dtype != HSV_C128 && dtype != HSV_C64) return bfail(HSV_ERR_ARG, "dtype must be HSV_C128 or HSV_C64");
  *out = nullptr;
  auto* b = new hsv_batch();
  b->num = num_instances;
  b->dtype = dtype;
  if (gate_offsets[0] != 0) {
    delete b;
    return bfail(HSV_ERR_ARG, "gate_offsets[0] must be 0");
  }
  uint64_t off = 0;
  for (int i = 0; i < num_instances; ++i) {
    const int n = num_qubits[i];
    if (n < 1 || n > kBatchMaxQubits || gate_offsets[i + 1] < gate_offsets[i]) {
      delete b;
      return bfail(HSV_ERR_ARG, "instance " + std::to_string(i) + ": qubits or gate offsets out of range");
    }
    b->nq.push_back(n);
    b->starts.push_back(off);
    off += 1ull << n;
  }
  b->total = off;
  b->gate_off.assign(gate_offsets, gate_offsets + num_instances + 1);
  b->gates.assign(gates, gates + gate_offsets[num_instances]);
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    delete b;
    return bfail(HSV_ERR_DEVICE, "no HIP device");
  }
  b->num_cu = prop.multiProcessorCount;
  if (hipStreamCreate(&b->stream) != hipSuccess) {
    delete b;
    return bfail(HSV_ERR_DEVICE, "hipStreamCreate failed");
  }
  if (hipMalloc(&b->state, off * b->elem()) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_starts), sizeof(uint64_t) * num_instances) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_nq), sizeof(int) * num_instances) != hipSuccess) {
    hsv_batch_destroy(b);
    return bfail(HSV_ERR_ALLOC, "cannot allocate the batch");
  }
  BCHK(hipMemcpy(b->d_starts, b->starts.data(), sizeof(uint64_t) * num_instances, hipMemcpyHostToDevice));
  BCHK(hipMemcpy(b->d_nq, b->nq.data(), sizeof(int) * num_instances, hipMemcpyHostToDevice));
  BCHK(hipEventCreate(&b->e0));
  BCHK(hipEventCreate(&b->e1));
  std::vector<PartGate> recs;
  if (int rc = plan(b, &recs)) {   // report bad gates at creation
    drop_plan(b);
    hsv_batch_destroy(b);
    return rc;
  }
  drop_plan(b);
  *out = b;
  return HSV_OK;
}

int hsv_batch_destroy(hsv_batch* b) {
  if (!b) return HSV_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_plan(b);
  (void)hipFree(b->state);
  (void)hipFree(b->d_starts);
  (void)hipFree(b->d_nq);
  (void)hipFree(b->d_masks);
  (void)hipFree(b->d_zcum);
  (void)hipFree(b->d_part);
  (void)hipFree(b->d_out);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return HSV_OK;
}

int hsv_batch_set_gates(hsv_batch* b, const hsv_gate* gates) {
  if (!b || (!gates && !b->gates.empty())) return bfail(HSV_ERR_ARG, "null argument");
  for (size_t g = 0; g < b->gates.size(); ++g)
    if (!same_structure(b->gates[g], gates[g]))
      return bfail(HSV_ERR_ARG, "gate " + std::to_string(g) + " differs in type or qubits from the batch's");
  std::vector<Lowered> Ls(b->gates.size());
  for (size_t g = 0; g < b->gates.size(); ++g) {
    std::string err;
    if (int rc = lower(gates[g], &Ls[g], &err)) return bfail(rc, "gate " + std::to_string(g) + ": " + err);
  }
  std::vector<Lowered> old(b->gates.size());
  bool same_kinds = true;
  for (size_t g = 0; g < b->gates.size(); ++g) {
    std::string err;
    same_kinds = same_kinds && lower(b->gates[g], &old[g], &err) == HSV_OK && old[g].kind == Ls[g].kind;
  }
  std::copy(gates, gates + b->gates.size(), b->gates.begin());
  if (!same_kinds) drop_plan(b);  // a caller-supplied matrix changed class: plan again
  if (!b->exec) return HSV_OK;    // the next run plans with the new gates
  // same geometry, new matrices: rewrite the records in place, keep the graph
  std::vector<PartGate> recs(b->rec_gate.size());
  for (size_t r = 0; r < recs.size(); ++r) {
    const int g = b->rec_gate[r];
    int inst = 0;
    while (b->gate_off[inst + 1] <= g) ++inst;
    recs[r] = part_gate_of(Ls[g], std::min(b->nq[inst], b->L), b->rec_inner[r].data());
  }
  // the tile plans: the same layout of segments and rounds keeps the graph
  // (new matrices in place); a different one (a phase gate that became or
  // stopped being a sign flip) plans again at the next run
  std::vector<BSeg> segs = b->segs;
  TileArrays A;
  build_tiles(b, Ls, &segs, &A);
  bool same_tiles = A.segs.size() == b->tile_nseg && A.rounds.size() == b->tile_nround &&
                    A.gates.size() == b->tile_ngate && A.mat2.size() == b->tile_nmat;
  for (size_t i = 0; same_tiles && i < segs.size(); ++i)
    same_tiles = segs[i].tseg_first == b->segs[i].tseg_first && segs[i].tseg_count == b->segs[i].tseg_count;
  if (!same_tiles) {
    drop_plan(b);
    return HSV_OK;
  }
  if (int rc = upload_tiles(b, A)) return rc;
  if (!recs.empty())
    BCHK(hipMemcpyAsync(b->d_gates, recs.data(), recs.size() * sizeof(PartGate), hipMemcpyHostToDevice, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return HSV_OK;
}

int hsv_batch_run(hsv_batch* b) {
  if (!b) return bfail(HSV_ERR_ARG, "null handle");
  if (!b->exec)
    if (int rc = build(b)) return rc;
  BCHK(hipEventRecord(b->e0, b->stream));
  BCHK(hipGraphLaunch(b->exec, b->stream));
  BCHK(hipEventRecord(b->e1, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  BCHK(hipEventElapsedTime(&b->last_ms, b->e0, b->e1));
  return HSV_OK;
}

int hsv_batch_expect_z(hsv_batch* b, const uint64_t* z_masks, double* out) {
  if (!b || !z_masks || !out) return bfail(HSV_ERR_ARG, "null argument");
  if (!b->d_out) {
    std::vector<uint64_t> cum(b->num);
    uint64_t items = 0;
    for (int i = 0; i < b->num; ++i) {
      cum[i] = items;
      items += ((1ull << b->nq[i]) + 4095) / 4096;
    }
    b->z_items = items;
    if (hipMalloc(reinterpret_cast<void**>(&b->d_masks), sizeof(uint64_t) * b->num) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->d_zcum), sizeof(uint64_t) * b->num) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->d_part), sizeof(double) * items) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&b->d_out), sizeof(double) * b->num) != hipSuccess)
      return bfail(HSV_ERR_ALLOC, "cannot allocate the expectation buffers");
    BCHK(hipMemcpy(b->d_zcum, cum.data(), sizeof(uint64_t) * b->num, hipMemcpyHostToDevice));
  }
  BCHK(hipMemcpyAsync(b->d_masks, z_masks, sizeof(uint64_t) * b->num, hipMemcpyHostToDevice, b->stream));
  const int grid = (int)std::min<uint64_t>(b->z_items, uint64_t(b->num_cu) * 8);
  if (b->dtype == HSV_C128)
    hipLaunchKernelGGL(k_batch_z<double>, grid, 256, 0, b->stream, static_cast<const double2*>(b->state), b->d_starts,
                       b->d_nq, b->d_masks, b->d_zcum, b->num, b->z_items, b->d_part);
  else
    hipLaunchKernelGGL(k_batch_z<float>, grid, 256, 0, b->stream, static_cast<const float2*>(b->state), b->d_starts,
                       b->d_nq, b->d_masks, b->d_zcum, b->num, b->z_items, b->d_part);
  hipLaunchKernelGGL(k_batch_zsum, (b->num + 255) / 256, 256, 0, b->stream, b->d_part, b->d_zcum, b->num,
                     b->z_items, b->d_out);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(out, b->d_out, sizeof(double) * b->num, hipMemcpyDeviceToHost, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return HSV_OK;
}

int hsv_batch_read(hsv_batch* b, int instance, double* amps) {
  if (!b || !amps || instance < 0 || instance >= b->num) return bfail(HSV_ERR_ARG, "bad instance");
  const uint64_t len = 1ull << b->nq[instance];
  const auto* src = static_cast<const unsigned char*>(b->state) + b->starts[instance] * b->elem();
  if (b->dtype == HSV_C128) {
    BCHK(hipMemcpyAsync(amps, src, len * 16, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
  } else {
    std::vector<float> tmp(2 * len);
    BCHK(hipMemcpyAsync(tmp.data(), src, len * 8, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    for (uint64_t i = 0; i < 2 * len; ++i) amps[i] = tmp[i];
  }
  return HSV_OK;
}

int hsv_batch_tune(hsv_batch* b, const char* key, int value) {
  if (!b || !key) return bfail(HSV_ERR_ARG, "null argument");
  const std::string k = key;
  const int max_bits = b->dtype == HSV_C128 ? 12 : kBatchMaxChunk;   // 64 KB of LDS per workgroup
  if (k == "chunk_bits" && value >= 6 && value <= max_bits) b->L = value;
  else if (k == "threads" && (value == 64 || value == 128 || value == 256 || value == 512)) b->nt = value;
  else if (k == "grid_per_cu" && value >= 1 && value <= 64) b->grid_per_cu = value;
  else return bfail(HSV_ERR_ARG, "unknown batch tuning key or value out of range: " + k);
  drop_plan(b);
  return HSV_OK;
}

int hsv_batch_info(hsv_batch* b, double* last_ms, int* stages, uint64_t* pass_amplitudes) {
  if (!b) return bfail(HSV_ERR_ARG, "null handle");
  if (last_ms) *last_ms = b->last_ms;
  if (stages) *stages = (int)b->stages.size();
  if (pass_amplitudes) {
    // amplitudes the stages move: each segment with gates reads and writes
    // its whole instance once
    uint64_t t = 0;
    for (const BSeg& S : b->segs)
      if (S.gate_count) t += 1ull << (S.bits + __builtin_popcountll(S.outer_mask));
    *pass_amplitudes = t;
  }
  return HSV_OK;
}

}  // extern "C"
