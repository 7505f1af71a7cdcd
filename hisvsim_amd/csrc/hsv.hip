// libhsv: MI355X (gfx950) hierarchical state-vector library -- C ABI.
//
// Host C++ that replaces the reference's state_vector.hpp / loop.hpp /
// basic_gates.hpp / execute.hpp for one GPU: the state lives in HBM, each gate
// is lowered on the host (gates.cpp) to a small matrix and applied by one
// coalesced sweep kernel, and each part runs as one LDS-staged pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hsv.h"
#include "fused.hpp"
#include "gates.hpp"
#include "kernels.hpp"
#include "tileplan.hpp"

using namespace hsv;

namespace {

thread_local std::string g_err;
std::atomic<int> g_opt_slots{0};
std::atomic<long long> g_gate_count{0}, g_part_count{0}, g_part_us{0}, g_apply_us{0};

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(HSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

void hsv_set_last_error(const std::string& msg) { g_err = msg; }

struct hsv_state {
  int device = 0;                  // the HIP device current at hsv_create
  int n = 0;
  int dtype = HSV_C128;
  uint64_t len = 0;
  void* amp = nullptr;
  hipStream_t stream = nullptr;
  void* gate_buf = nullptr;        // device copy of the current part's gates
  size_t gate_buf_bytes = 0;
  double* partial = nullptr;       // device reduction partials
  std::vector<double> host_partial;
  // Event pairs around each timed launch (a pool reused across collections).
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  size_t events_used = 0;
  int timing = 0;
  int max_lds = 65536;
  int num_cu = 256;
  // tuning knobs (hsv_tune)
  int reg_waves = 4;                 // persistent-grid factor of the register part kernel
  int part_kernel = 2;               // 2: FP64 tile kernel for parts that fit a tile, else 1; 1: register-blocked kernel for k >= 8; 0: LDS kernel only
  int reg_bits = 0;                  // preferred register bits (0: per dtype)
  int tile_bits = 0;                 // tile kernel width: 0 = 12 qubits (complex128) / 13 (complex64); one less = 256-thread tiles
  // Slots 0..low_slots-1 join every LDS-staged part when they fit: gathered
  // runs of 2^low_slots amplitudes (512 B for complex128 at 5) sustain
  // 5.0 TB/s on MI355X against 3.9 TB/s for 128 B runs.
  int low_slots = 5;
  // Low inner bits the register kernel keeps on the lanes in its first and
  // last layouts (coalesced global loads and stores); 0 turns this off. Off
  // by default: at 31 qubits it measured no gain (the extra layout change
  // costs what the coalescing saves).
  int io_bits = 0;
  // hsv_permute_qubits: 0 = in place when a second state does not fit or
  // from 32 qubits on, 1 = always in place, 2 = always out of place
  int permute_in_place = 0;
  BfTune bf;                         // BF16-split fused kernel form and launch shape (see hsv_tune)
  // The state is |0...0> (set by create and reset, cleared by every call
  // that may change it): the chain then skips the tiles that are still zero.
  bool basis0 = false;
  DeviceScratch fused_scratch;       // fused unitary + offset tables
  size_t elem() const { return dtype == HSV_C128 ? 16 : 8; }
};

namespace {

template <typename R>
typename VecT<R>::type* amp_as(hsv_state* s) {
  return static_cast<typename VecT<R>::type*>(s->amp);
}

int begin_timed(hsv_state* s) {
  if (!s->timing) return HSV_OK;
  if (s->events_used == s->events.size()) {
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    s->events.emplace_back(a, b);
  }
  HIPCHK(hipEventRecord(s->events[s->events_used].first, s->stream));
  return HSV_OK;
}
int end_timed(hsv_state* s) {
  if (!s->timing) return HSV_OK;
  HIPCHK(hipEventRecord(s->events[s->events_used].second, s->stream));
  ++s->events_used;
  return HSV_OK;
}

int check_gate_qubits(const hsv_state* s, const Lowered& L) {
  for (int i = 0; i < L.nt; ++i)
    if (L.t[i] >= s->n) return fail(HSV_ERR_ARG, "target qubit " + std::to_string(L.t[i]) + " out of range");
  for (int i = 0; i < L.nc; ++i)
    if (L.c[i] >= s->n) return fail(HSV_ERR_ARG, "control qubit " + std::to_string(L.c[i]) + " out of range");
  return HSV_OK;
}

// Sweep geometry shared by the global and LDS kernels: fixed bit positions
// (ascending), forced-one mask, target offsets and the work-item count over
// an index space of `nbits` bits. `map` renames qubits to slots.
template <typename Map>
void sweep_geometry(const Lowered& L, int nbits, Map map, int* nfixed, int* fixed, uint64_t* setmask,
                    uint64_t* off1, uint64_t* off2, uint64_t* nwork) {
  uint64_t fixedmask = 0, set = 0;
  for (int i = 0; i < L.nc; ++i) { fixedmask |= 1ull << map(L.c[i]); set |= 1ull << map(L.c[i]); }
  for (int i = 0; i < L.nt; ++i) fixedmask |= 1ull << map(L.t[i]);
  if (L.kind == K_PHASE1) set |= 1ull << map(L.t[0]);
  int nf = 0;
  for (int b = 0; b < 64; ++b)
    if (fixedmask >> b & 1) fixed[nf++] = b;
  for (int i = nf; i < 7; ++i) fixed[i] = 0;
  *nfixed = nf;
  *setmask = set;
  *off1 = 1ull << map(L.t[0]);
  *off2 = L.nt == 2 ? 1ull << map(L.t[1]) : 0;
  *nwork = 1ull << (nbits - nf);
}

void pack_matrix(const Lowered& L, double* m) {
  for (int i = 0; i < 16; ++i) { m[2 * i] = L.m[i].real(); m[2 * i + 1] = L.m[i].imag(); }
}

int grid_for(uint64_t items, uint64_t per_block, uint64_t cap) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (int)std::min<uint64_t>(g, cap);
}

template <typename R, int ITEMS, bool NT>
int launch_gate_k(hsv_state* s, const Lowered& L, const GateArgs& g) {
  const int grid = grid_for(g.nwork, uint64_t(kGateBlock) * ITEMS, 1u << 20);
  auto* a = amp_as<R>(s);
  switch (L.kind) {
    case K_PHASE1: hipLaunchKernelGGL((k_gate<R, GK_PHASE1, ITEMS, NT>), grid, kGateBlock, 0, s->stream, a, g); break;
    case K_DIAG1: hipLaunchKernelGGL((k_gate<R, GK_DIAG1, ITEMS, NT>), grid, kGateBlock, 0, s->stream, a, g); break;
    case K_XSWAP1: hipLaunchKernelGGL((k_gate<R, GK_XSWAP1, ITEMS, NT>), grid, kGateBlock, 0, s->stream, a, g); break;
    case K_DENSE1: hipLaunchKernelGGL((k_gate<R, GK_DENSE1, ITEMS, NT>), grid, kGateBlock, 0, s->stream, a, g); break;
    case K_DIAG2: hipLaunchKernelGGL((k_gate<R, GK_DIAG2, ITEMS, NT>), grid, kGateBlock, 0, s->stream, a, g); break;
    case K_SWAP2: hipLaunchKernelGGL((k_gate<R, GK_SWAP2, ITEMS, NT>), grid, kGateBlock, 0, s->stream, a, g); break;
    case K_DENSE2: {
      constexpr int IT2 = ITEMS > 4 ? 4 : ITEMS;
      hipLaunchKernelGGL((k_gate<R, GK_DENSE2, IT2, NT>), grid_for(g.nwork, uint64_t(kGateBlock) * IT2, 1u << 20),
                         kGateBlock, 0, s->stream, a, g);
      break;
    }
    default: return HSV_OK;
  }
  HIPCHK(hipGetLastError());
  return HSV_OK;
}

template <typename R>
int launch_gate_t(hsv_state* s, const Lowered& L) {
  GateArgs g{};
  sweep_geometry(L, s->n, [](int q) { return q; }, &g.nfixed, g.fixed, &g.setmask, &g.off1, &g.off2, &g.nwork);
  pack_matrix(L, g.m);
  // Measured on MI355X (2^31 amplitudes): when the pair stride is >= 8
  // amplitudes, 2 items per thread with non-temporal access sustain the most
  // bandwidth; for the lowest targets, 8 items per thread with cached access.
  const uint64_t min_off = L.nt == 2 ? std::min(g.off1, g.off2) : g.off1;
  if (min_off >= 8) return launch_gate_k<R, 2, true>(s, L, g);
  return launch_gate_k<R, 8, false>(s, L, g);
}

int launch_gate(hsv_state* s, const Lowered& L) {
  if (L.kind == K_IDENT) return HSV_OK;
  if (int rc = check_gate_qubits(s, L)) return rc;
  return s->dtype == HSV_C128 ? launch_gate_t<double>(s, L) : launch_gate_t<float>(s, L);
}

int max_part_qubits(const hsv_state* s) {
  int k = 0;
  while (k < kMaxPartQubits && (size_t(1) << (k + 1)) * s->elem() + 4096 <= (size_t)s->max_lds) ++k;
  return k;
}

int ensure_gate_buf(hsv_state* s, size_t bytes) {
  if (bytes <= s->gate_buf_bytes) return HSV_OK;
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->gate_buf) HIPCHK(hipFree(s->gate_buf));
  s->gate_buf = nullptr;
  s->gate_buf_bytes = 0;
  size_t cap = std::max<size_t>(bytes, 1 << 16);
  if (hipMalloc(&s->gate_buf, cap) != hipSuccess) return fail(HSV_ERR_ALLOC, "gate buffer allocation failed");
  s->gate_buf_bytes = cap;
  return HSV_OK;
}

template <typename R>
int launch_part_t(hsv_state* s, const std::vector<Lowered>& Ls, uint64_t part_mask) {
  const int k = __builtin_popcountll(part_mask);
  int inner_of[64];
  for (int b = 0, r = 0; b < 64; ++b) inner_of[b] = (part_mask >> b & 1) ? r++ : -1;
  std::vector<PartGate> pg;
  pg.reserve(Ls.size());
  for (const Lowered& L : Ls) {
    if (L.kind == K_IDENT) continue;
    PartGate G{};
    uint64_t set, o1, o2, nw;
    sweep_geometry(L, k, [&](int q) { return inner_of[q]; }, &G.nfixed, G.fixed, &set, &o1, &o2, &nw);
    G.kind = L.kind;
    G.setmask = (uint32_t)set;
    G.off1 = (uint32_t)o1;
    G.off2 = (uint32_t)o2;
    G.nwork = (uint32_t)nw;
    pack_matrix(L, G.m);
    pg.push_back(G);
  }
  const size_t bytes = pg.size() * sizeof(PartGate);
  if (int rc = ensure_gate_buf(s, bytes)) return rc;
  if (bytes) HIPCHK(hipMemcpyAsync(s->gate_buf, pg.data(), bytes, hipMemcpyHostToDevice, s->stream));
  const uint64_t all = s->n == 64 ? ~0ull : ((1ull << s->n) - 1);
  const uint64_t outer = all & ~part_mask;
  const uint64_t nbases = 1ull << (s->n - k);
  const size_t lds = (size_t(1) << k) * sizeof(typename VecT<R>::type);
  auto* a = amp_as<R>(s);
  auto* gates = static_cast<const PartGate*>(s->gate_buf);
  const int ng = (int)pg.size();
  if (k >= 9) {
    const int grid = (int)std::min<uint64_t>(nbases, uint64_t(s->num_cu) * 16);
    if (lds > 65536) HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_part<R, 256>),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_part<R, 256>), grid, 256, lds, s->stream, a, nbases, outer, part_mask, k, gates, ng);
  } else {
    const int grid = (int)std::min<uint64_t>(nbases, uint64_t(s->num_cu) * 32);
    hipLaunchKernelGGL((k_part<R, 64>), grid, 64, lds, s->stream, a, nbases, outer, part_mask, k, gates, ng);
  }
  HIPCHK(hipGetLastError());
  return HSV_OK;
}

// ---- register-blocked part plan ---------------------------------------------
bool is_diagonal(int kind) { return kind == K_PHASE1 || kind == K_DIAG1 || kind == K_DIAG2; }

struct RegPlan {
  std::vector<RegSeg> segs;
  std::vector<RegGate> gates;
  std::vector<int> order;   // index into the part's gate list, in plan order
  std::vector<double> mat2; // 32 doubles per dense two-qubit gate
  std::vector<int> m2idx;   // per gate: its matrix in mat2, or -1
};

// Picks the next layout: the targets of the earliest pending non-diagonal
// gates, topped up with bits of the current layout, then with high bits
// (low inner bits stay thread bits when possible, for coalesced access).
// `forbid`: inner bits that must stay thread bits (the first layout keeps
// the low inner bits on the lanes so the global loads are coalesced); gates
// needing them are skipped over, not waited for.
uint32_t choose_layout(const std::vector<int>& pending, const std::vector<uint32_t>& need_bits,
                       const std::vector<std::array<int, 2>>& need_list, uint32_t cur, int k, int RB,
                       uint32_t forbid = 0) {
  uint32_t want = 0;
  int cnt = 0;
  for (int i : pending) {
    if (!need_bits[i]) continue;
    if (need_bits[i] & forbid) continue;
    if (cnt + __builtin_popcount(need_bits[i] & ~want) > RB) break;
    for (int t : need_list[i])
      if (t >= 0 && !(want >> t & 1)) { want |= 1u << t; ++cnt; }
    if (cnt == RB) break;
  }
  for (int b = k - 1; b >= 0 && cnt < RB; --b)
    if ((cur >> b & 1) && !(want >> b & 1) && !(forbid >> b & 1)) { want |= 1u << b; ++cnt; }
  for (int b = k - 1; b >= 0 && cnt < RB; --b)
    if (!(want >> b & 1) && !(forbid >> b & 1)) { want |= 1u << b; ++cnt; }
  return want;
}

// Splits a part (gates in inner-bit coordinates) into layout segments. Gates
// may move ahead of later gates only past gates on disjoint qubits, so the
// result applies the same operator as the part in its given order.
// io_forbid: inner bits the first and the last layout keep as thread bits,
// so that the global loads and stores of consecutive lanes hit consecutive
// amplitudes; a layout-only segment is appended when the last layout has
// them as register bits.
void plan_register_part(const std::vector<Lowered>& Ls, const int* inner_of, int k, int RB, RegPlan* P,
                        uint32_t io_forbid) {
  const int ng = (int)Ls.size();
  std::vector<uint32_t> qmask(ng, 0), need_bits(ng, 0);
  std::vector<std::array<int, 2>> need_list(ng, {-1, -1});
  std::vector<int> pending;
  for (int i = 0; i < ng; ++i) {
    const Lowered& L = Ls[i];
    if (L.kind == K_IDENT) continue;
    for (int t = 0; t < L.nt; ++t) qmask[i] |= 1u << inner_of[L.t[t]];
    for (int c = 0; c < L.nc; ++c) qmask[i] |= 1u << inner_of[L.c[c]];
    if (!is_diagonal(L.kind))
      for (int t = 0; t < L.nt; ++t) {
        need_bits[i] |= 1u << inner_of[L.t[t]];
        need_list[i][t] = inner_of[L.t[t]];
      }
    pending.push_back(i);
  }
  uint32_t cur = choose_layout(pending, need_bits, need_list, 0, k, RB, io_forbid);
  P->segs.clear();
  P->gates.clear();
  P->order.clear();
  P->mat2.clear();
  P->m2idx.clear();
  do {
    uint32_t blocked = 0;
    std::vector<int> applied, rest;
    for (int i : pending) {
      if (qmask[i] & blocked) { blocked |= qmask[i]; rest.push_back(i); continue; }
      if ((need_bits[i] & ~cur) == 0) applied.push_back(i);
      else { blocked |= qmask[i]; rest.push_back(i); }
    }
    RegSeg seg{cur, (int)P->gates.size(), (int)applied.size(), 0};
    P->segs.push_back(seg);
    for (int i : applied) {
      const Lowered& L = Ls[i];
      RegGate G{};
      G.kind = L.kind;
      G.t0bit = 1u << inner_of[L.t[0]];
      G.t1bit = L.nt == 2 ? 1u << inner_of[L.t[1]] : 0;
      G.p0 = __builtin_popcount(cur & (G.t0bit - 1));
      G.p1 = L.nt == 2 ? __builtin_popcount(cur & (G.t1bit - 1)) : 0;
      for (int c = 0; c < L.nc; ++c) G.cmask |= 1u << inner_of[L.c[c]];
      if (L.kind == K_PHASE1) G.cmask |= G.t0bit;
      double m[32];
      pack_matrix(L, m);
      std::memcpy(G.m, m, sizeof(G.m));
      int m2 = -1;
      if (L.kind == K_DENSE2) {
        m2 = (int)(P->mat2.size() / 32);
        P->mat2.insert(P->mat2.end(), m, m + 32);
      }
      P->gates.push_back(G);
      P->order.push_back(i);
      P->m2idx.push_back(m2);
    }
    pending.swap(rest);
    if (!pending.empty()) cur = choose_layout(pending, need_bits, need_list, cur, k, RB);
  } while (!pending.empty());
  if (cur & io_forbid) {
    const uint32_t io = choose_layout({}, need_bits, need_list, cur, k, RB, io_forbid);
    P->segs.push_back(RegSeg{io, (int)P->gates.size(), 0, 0});
  }
}

// Register bits and thread bits for a k-qubit part (k >= 8). With complex128,
// 16 values per thread (RB = 4) fit in 159 VGPRs (3 waves/SIMD) while 32 take
// all 512 registers; complex64 affords RB = 5 at 174 VGPRs.
// Instantiated (TB, RB) shapes; the nearest one to the preferred RB is used.
// With the prefetch buffer a complex128 thread holds 2 x 16 values in 252
// VGPRs; 2 x 32 would spill, so complex128 stops at RB = 4.
bool reg_shape_available(int tb, int rb, int elem) {
  if (elem == 16 && rb > 4) return false;
  switch (tb) {
    case 6: return rb >= 2 && rb <= 5;
    case 7: case 8: return rb >= 3 && rb <= 5;
    case 9: return rb == 3 || rb == 4;
    default: return false;
  }
}
bool reg_shape(int k, int rb_pref, int elem, int* TB, int* RB) {
  for (int d = 0; d <= 3; ++d)
    for (int rb : {rb_pref + d, rb_pref - d})
      if (reg_shape_available(k - rb, rb, elem)) {
        *TB = k - rb;
        *RB = rb;
        return true;
      }
  return false;
}

template <typename R, int TB, int RB>
int launch_reg(hsv_state* s, const RegPlan& P, uint64_t nbases, uint64_t outer, uint64_t part_mask) {
  // device layout: segments | gate records | 4x4 matrices
  const size_t seg_bytes = P.segs.size() * sizeof(RegSeg);
  const size_t gate_bytes = P.gates.size() * sizeof(RegGate);
  const size_t bytes = seg_bytes + gate_bytes + P.mat2.size() * sizeof(double);
  if (int rc = ensure_gate_buf(s, bytes)) return rc;
  std::vector<unsigned char> host(bytes);
  std::memcpy(host.data(), P.segs.data(), seg_bytes);
  const double* dev_mat2 =
      reinterpret_cast<const double*>(static_cast<const unsigned char*>(s->gate_buf) + seg_bytes + gate_bytes);
  RegGate* hg = reinterpret_cast<RegGate*>(host.data() + seg_bytes);
  for (size_t i = 0; i < P.gates.size(); ++i) {
    hg[i] = P.gates[i];
    hg[i].m2 = P.m2idx[i] >= 0 ? dev_mat2 + 32 * size_t(P.m2idx[i]) : nullptr;
  }
  if (!P.mat2.empty()) std::memcpy(host.data() + seg_bytes + gate_bytes, P.mat2.data(), P.mat2.size() * sizeof(double));
  HIPCHK(hipMemcpyAsync(s->gate_buf, host.data(), bytes, hipMemcpyHostToDevice, s->stream));
  const auto* segs = static_cast<const RegSeg*>(s->gate_buf);
  const auto* gates = reinterpret_cast<const RegGate*>(static_cast<const unsigned char*>(s->gate_buf) + seg_bytes);
  constexpr int K = TB + RB;
  const size_t lds = P.segs.size() > 1 ? (size_t(1) << K) * sizeof(typename VecT<R>::type) : 0;
  auto kern = &k_part_reg<R, TB, RB, true>;
  if (lds > 65536)
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  const int per_cu = lds > 65536 ? 1 : lds > 32768 ? 2 : 4;
  const int grid = (int)std::min<uint64_t>(nbases, uint64_t(s->num_cu) * per_cu * s->reg_waves);
  hipLaunchKernelGGL(kern, grid, 1 << TB, lds, s->stream, amp_as<R>(s), nbases, outer, part_mask, segs,
                     (int)P.segs.size(), gates);
  HIPCHK(hipGetLastError());
  return HSV_OK;
}

template <typename R>
int launch_reg_part_t(hsv_state* s, const std::vector<Lowered>& Ls, uint64_t part_mask) {
  const int k = __builtin_popcountll(part_mask);
  int TB, RB;
  const int rb_pref = s->reg_bits > 0 ? s->reg_bits : (sizeof(R) == 8 ? 4 : 5);
  if (!reg_shape(k, rb_pref, (int)sizeof(typename VecT<R>::type), &TB, &RB))
    return fail(HSV_ERR_LIMIT, "no register-kernel shape for a " + std::to_string(k) + "-qubit part");
  int inner_of[64];
  for (int b = 0, r = 0; b < 64; ++b) inner_of[b] = (part_mask >> b & 1) ? r++ : -1;
  RegPlan P;
  plan_register_part(Ls, inner_of, k, RB, &P, (1u << std::min(s->io_bits, TB)) - 1);
  const uint64_t all = s->n == 64 ? ~0ull : ((1ull << s->n) - 1);
  const uint64_t outer = all & ~part_mask;
  const uint64_t nbases = 1ull << (s->n - k);
#define HSV_REG_CASE(T_, R_) \
  if (TB == T_ && RB == R_) return launch_reg<R, T_, R_>(s, P, nbases, outer, part_mask);
  HSV_REG_CASE(6, 2) HSV_REG_CASE(6, 3) HSV_REG_CASE(6, 4)
  HSV_REG_CASE(7, 3) HSV_REG_CASE(8, 3) HSV_REG_CASE(9, 3)
  HSV_REG_CASE(7, 4) HSV_REG_CASE(8, 4) HSV_REG_CASE(9, 4)
  if constexpr (sizeof(R) == 4) {
    HSV_REG_CASE(6, 5) HSV_REG_CASE(7, 5) HSV_REG_CASE(8, 5)
  }
#undef HSV_REG_CASE
  return fail(HSV_ERR_LIMIT, "no register-kernel shape for a " + std::to_string(k) + "-qubit part");
}

// Tile width of the FP64 tile part kernel: by default 64 KB of LDS per
// workgroup of 512 threads (two workgroups per CU); tile_bits one less
// gives 32 KB tiles of 256 threads (four per CU).
int tile_part_default(const hsv_state* s) { return s->dtype == HSV_C128 ? 12 : 13; }
int tile_part_bits(const hsv_state* s) {
  const int def = tile_part_default(s);
  return s->tile_bits == def - 1 ? def - 1 : def;
}

}  // namespace

// Runs of uncontrolled one-qubit gates on one qubit (no other gate on that
// qubit in between) multiplied into one gate, later matrix times earlier:
// the gates between them act on other qubits and commute with them. The
// product is dense unless both are diagonal. Identities are dropped.
std::vector<Lowered> hsv::fuse_one_qubit_runs(const std::vector<Lowered>& Ls) {
  std::vector<Lowered> out;
  int last[64];
  for (int& x : last) x = -1;
  for (const Lowered& L : Ls) {
    if (L.kind == K_IDENT) continue;
    if (L.nt == 1 && L.nc == 0) {
      const int q = L.t[0];
      if (last[q] < 0) {
        last[q] = (int)out.size();
        out.push_back(L);
        continue;
      }
      Lowered& A = out[last[q]];
      cd a[16], b[16];
      dense_of(A, a);
      dense_of(L, b);
      const cd p0 = b[0] * a[0] + b[1] * a[2], p1 = b[0] * a[1] + b[1] * a[3];
      const cd p2 = b[2] * a[0] + b[3] * a[2], p3 = b[2] * a[1] + b[3] * a[3];
      const bool diag = (A.kind == K_PHASE1 || A.kind == K_DIAG1) && (L.kind == K_PHASE1 || L.kind == K_DIAG1);
      if (diag) {
        A.kind = K_DIAG1;
        A.m[0] = p0;
        A.m[1] = p3;
      } else {
        A.kind = K_DENSE1;
        A.m[0] = p0;
        A.m[1] = p1;
        A.m[2] = p2;
        A.m[3] = p3;
      }
      continue;
    }
    for (int i = 0; i < L.nt; ++i) last[L.t[i]] = -1;
    for (int i = 0; i < L.nc; ++i) last[L.c[i]] = -1;
    out.push_back(L);
  }
  return out;
}

// The FP64 tile kernel's plan of one part (shared with the circuit batch):
// the part's one-qubit runs fused (fuse_one_qubit_runs), then the register
// kernel's layout segments over the tile's KT bits with RB register bits;
// fast segments (only uncontrolled one-qubit dense gates and sign flips)
// become rounds. Indices are relative to the plan.
int hsv::build_tile_plan(const std::vector<Lowered>& Ls_in, const int* inner_of, int KT, int RB, TilePlan* T) {
  if (RB < 2 || RB > 5 || RB > KT) return HSV_ERR_LIMIT;  // the planner places two-target gates in registers
  const std::vector<Lowered> Ls = fuse_one_qubit_runs(Ls_in);
  RegPlan P;
  plan_register_part(Ls, inner_of, KT, RB, &P, 0);
  // fast segments (only uncontrolled one-qubit dense gates and sign flips)
  // become rounds: the dense gates on distinct register slots, then a run of
  // sign flips; a dense gate after a flip, or on a slot the round already
  // has, starts the next round (so no gate moves past one it does not
  // commute with)
  std::vector<TileRound>& rounds = T->rounds;
  for (RegSeg& S : P.segs) {
    if (__builtin_popcount(S.regmask) != RB) return HSV_ERR_LIMIT;
    S.fast = 1;
    for (int i = S.first; i < S.first + S.count; ++i) {
      const RegGate& G = P.gates[i];
      const bool dense1 = G.kind == GK_DENSE1 && G.cmask == 0;
      const bool sign = G.kind == GK_PHASE1 && G.m[0] == -1.0 && G.m[1] == 0.0;
      if (!dense1 && !sign) S.fast = 0;
    }
    if (!S.fast) continue;
    const int r0 = (int)rounds.size();
    TileRound Q{{-1, -1, -1, -1, -1}, 0, 0, 0};
    bool open = false;
    for (int i = S.first; i < S.first + S.count; ++i) {
      RegGate& G = P.gates[i];
      if (G.kind == GK_DENSE1) {
        if (Q.ns > 0 || Q.g[G.p0] >= 0) {
          rounds.push_back(Q);
          Q = TileRound{{-1, -1, -1, -1, -1}, 0, 0, 0};
        }
        Q.g[G.p0] = i;
      } else {
        // the thread part of the flip's bits, and the register slots r whose
        // bits (r's bits on the set bits of regmask, lowest first) hold the rest
        const uint32_t rm = G.cmask & S.regmask;
        uint32_t rsel = 0;
        for (uint32_t r = 0; r < (1u << RB); ++r) {
          uint32_t jr = 0, m = S.regmask;
          for (int b = 0; m; ++b, m &= m - 1)
            if (r >> b & 1) jr |= m & (~m + 1);
          if ((jr & rm) == rm) rsel |= 1u << r;
        }
        G.t0bit = G.cmask & ~S.regmask;
        G.t1bit = rsel;
        if (Q.ns == 0) Q.sfirst = i;
        ++Q.ns;
      }
      open = true;
    }
    if (open) rounds.push_back(Q);
    S.first = r0;
    S.count = (int)rounds.size() - r0;
  }
  T->segs = std::move(P.segs);
  T->gates = std::move(P.gates);
  T->m2idx = std::move(P.m2idx);
  T->mat2 = std::move(P.mat2);
  return HSV_OK;
}

namespace {

// A part on the FP64 tile kernel (k_part_tile): the tile holds the part's
// qubits padded with the lowest free slots (so slots 0-4 first) to 2^KT
// amplitudes; the register plan is the register kernel's, over the tile's
// bits, with KT - log2(NTH) register bits per thread.
template <typename R, int KT, int NTH>
int launch_tile_part_t(hsv_state* s, const std::vector<Lowered>& Ls, uint64_t part_mask) {
  constexpr int RB = KT - tile_log2(NTH);
  const int n = s->n;
  uint64_t tmask = part_mask;
  for (int b = 0; b < n && __builtin_popcountll(tmask) < KT; ++b) tmask |= 1ull << b;
  if (__builtin_popcountll(tmask) != KT) return fail(HSV_ERR_LIMIT, "part does not fit the tile kernel");
  int inner_of[64], slot[KT];
  for (int b = 0, r = 0; b < 64; ++b) {
    inner_of[b] = (tmask >> b & 1) ? r : -1;
    if (tmask >> b & 1) slot[r++] = b;
  }
  TilePlan TP;
  if (build_tile_plan(Ls, inner_of, KT, RB, &TP) != HSV_OK)
    return fail(HSV_ERR_LIMIT, "tile layout without its register bits");
  const std::vector<TileRound>& rounds = TP.rounds;
  RegPlan P;
  P.segs = std::move(TP.segs);
  P.gates = std::move(TP.gates);
  P.m2idx = std::move(TP.m2idx);
  P.mat2 = std::move(TP.mat2);
  TileTabs T{};
  for (int v = 0; v < 256; ++v)
    for (int j = 0; j < 8; ++j) {
      if (!(v >> j & 1)) continue;
      if (v < 64 && j < 6) T.alo[v] += 1ull << slot[j];
      if (6 + j < KT) T.ahi[v] += 1ull << slot[6 + j];
    }
  // device layout: tables | segments | rounds | gate records | 4x4 matrices
  const size_t seg_off = sizeof(TileTabs);
  const size_t round_off = seg_off + P.segs.size() * sizeof(RegSeg);
  const size_t gate_off = round_off + rounds.size() * sizeof(TileRound);
  const size_t mat_off = gate_off + P.gates.size() * sizeof(RegGate);
  const size_t bytes = mat_off + P.mat2.size() * sizeof(double);
  if (int rc = ensure_gate_buf(s, bytes)) return rc;
  std::vector<unsigned char> host(bytes);
  auto* dev = static_cast<unsigned char*>(s->gate_buf);
  std::memcpy(host.data(), &T, sizeof(T));
  std::memcpy(host.data() + seg_off, P.segs.data(), P.segs.size() * sizeof(RegSeg));
  if (!rounds.empty()) std::memcpy(host.data() + round_off, rounds.data(), rounds.size() * sizeof(TileRound));
  RegGate* hg = reinterpret_cast<RegGate*>(host.data() + gate_off);
  for (size_t i = 0; i < P.gates.size(); ++i) {
    hg[i] = P.gates[i];
    hg[i].m2 = P.m2idx[i] >= 0 ? reinterpret_cast<const double*>(dev + mat_off) + 32 * size_t(P.m2idx[i]) : nullptr;
  }
  if (!P.mat2.empty()) std::memcpy(host.data() + mat_off, P.mat2.data(), P.mat2.size() * sizeof(double));
  HIPCHK(hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, s->stream));
  const uint64_t all = n == 64 ? ~0ull : ((1ull << n) - 1);
  const uint64_t outer = all & ~tmask;
  const uint64_t ntiles = 1ull << (n - KT);
  uint64_t top = 0;  // the highest tile_order outer slots vary fastest across workgroups
  for (int i = 0; i < s->bf.tile_order && (outer & ~top); ++i) top |= 1ull << (63 - __builtin_clzll(outer & ~top));
  hipLaunchKernelGGL((k_part_tile<R, KT, NTH>), (unsigned)ntiles, NTH, 0, s->stream, amp_as<R>(s),
                     reinterpret_cast<const TileTabs*>(dev), reinterpret_cast<const RegSeg*>(dev + seg_off),
                     (int)P.segs.size(), reinterpret_cast<const TileRound*>(dev + round_off),
                     reinterpret_cast<const RegGate*>(dev + gate_off), ntiles, outer, top);
  HIPCHK(hipGetLastError());
  return HSV_OK;
}

template <typename R>
int launch_tile_part(hsv_state* s, const std::vector<Lowered>& Ls, uint64_t part_mask) {
  constexpr int KT = sizeof(R) == 8 ? 12 : 13;
  return tile_part_bits(s) == KT ? launch_tile_part_t<R, KT, 512>(s, Ls, part_mask)
                                 : launch_tile_part_t<R, KT - 1, 256>(s, Ls, part_mask);
}

template <typename R, int MODE>
int reduce_t(hsv_state* a, hsv_state* b, int qubit, double* out0, double* out1) {
  auto* pa = amp_as<R>(a);
  auto* pb = b ? amp_as<R>(b) : nullptr;
  int blocks = kRedBlocks;
  if (MODE == RD_CHUNK) blocks = (int)std::min<uint64_t>(kRedBlocks, a->len);
  hipLaunchKernelGGL((k_reduce<R, MODE>), blocks, kRedBlock, 0, a->stream, pa, pb, a->len, qubit, a->partial);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(a->host_partial.data(), a->partial, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost,
                        a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  if (MODE == RD_CHUNK) return blocks;
  double s0 = 0, s1 = 0;
  for (int i = 0; i < blocks; ++i) {
    if (MODE == RD_MAXDIFF) s0 = std::max(s0, a->host_partial[2 * i]);
    else { s0 += a->host_partial[2 * i]; s1 += a->host_partial[2 * i + 1]; }
  }
  if (out0) *out0 = s0;
  if (out1) *out1 = s1;
  return HSV_OK;
}

template <int MODE>
int reduce(hsv_state* a, hsv_state* b, int qubit, double* o0, double* o1) {
  return a->dtype == HSV_C128 ? reduce_t<double, MODE>(a, b, qubit, o0, o1) : reduce_t<float, MODE>(a, b, qubit, o0, o1);
}

int lower_all(const hsv_gate* gates, int num_gates, std::vector<Lowered>* out, const hsv_state* s) {
  out->resize(num_gates);
  for (int i = 0; i < num_gates; ++i) {
    std::string err;
    if (int rc = lower(gates[i], &(*out)[i], &err)) return fail(rc, "gate " + std::to_string(i) + ": " + err);
    if (int rc = check_gate_qubits(s, (*out)[i])) return rc;
  }
  return HSV_OK;
}

}  // namespace

// In-place qubit permutation (amplitude bit s moves to bit new_slot[s]) as
// two involutions: on each cycle c_0 -> c_1 -> ... -> c_{m-1} of the slot
// map, the reflections s1: c_k -> c_{-k} and s2: c_k -> c_{1-k} (indices
// mod m) compose to s2(s1(c_k)) = c_{k+1}. Two tiled swap passes, no second
// buffer. Needs n >= 10 (a tile of up to 10 slots).
// With renorm, the first pass also sums |a|^2 and the second scales by
// 1 / norm (both passes must move data; *renormalized tells whether it was done).
static int permute_in_place(hsv_state* s, const int* new_slot, bool renorm = false, bool* renormalized = nullptr) {
  const int n = s->n;
  if (renormalized) *renormalized = false;
  if (n < 10) return fail(HSV_ERR_LIMIT, "the in-place permutation needs at least 10 qubits");
  std::vector<int> s1(n), s2(n);
  for (int q = 0; q < n; ++q) s1[q] = s2[q] = q;
  std::vector<char> done(n, 0);
  for (int q0 = 0; q0 < n; ++q0) {
    if (done[q0]) continue;
    std::vector<int> cyc;
    for (int q = q0; !done[q]; q = new_slot[q]) {
      done[q] = 1;
      cyc.push_back(q);
    }
    const int m = (int)cyc.size();
    for (int k = 0; k < m; ++k) {
      s1[cyc[k]] = cyc[(m - k) % m];
      s2[cyc[k]] = cyc[(1 - k + m) % m];
    }
  }
  bool moves1 = false, moves2 = false;
  for (int q = 0; q < n; ++q) {
    moves1 = moves1 || s1[q] != q;
    moves2 = moves2 || s2[q] != q;
  }
  renorm = renorm && moves1 && moves2;
  if (int rc = begin_timed(s)) return rc;
  for (const std::vector<int>* sp : {&s1, &s2}) {
    const std::vector<int>& sig = *sp;
    bool any = false;
    for (int q = 0; q < n; ++q) any = any || sig[q] != q;
    if (!any) continue;
    const bool first = sp == &s1;
    // the tile's slot set B: slots 0-4 and their partners, then more low
    // slots (unpaired, or whole pairs) up to 9..10 slots
    uint64_t B = 0;
    for (int q = 0; q < 5; ++q) B |= (1ull << q) | (1ull << sig[q]);
    for (int q = 5; q < n && __builtin_popcountll(B) < 10; ++q) {
      if (B >> q & 1) continue;
      const uint64_t add = (1ull << q) | (1ull << sig[q]);
      if (__builtin_popcountll(B | add) <= 10) B |= add;
    }
    SwapTiles st{};
    st.nb = __builtin_popcountll(B);
    int bslot[10], pos[64];
    for (int q = 0, j = 0; q < n; ++q)
      if (B >> q & 1) {
        bslot[j] = q;
        pos[q] = j++;
      }
    for (int v = 0; v < 32; ++v)
      for (int j = 0; j < 5; ++j) {
        if (!(v >> j & 1)) continue;
        st.lo[v] += 1ull << bslot[j];
        st.slo[v] |= (uint16_t)(1u << pos[sig[bslot[j]]]);
        if (5 + j < st.nb) {
          st.hi[v] += 1ull << bslot[5 + j];
          st.shi[v] |= (uint16_t)(1u << pos[sig[bslot[5 + j]]]);
        }
      }
    const uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1);
    st.outer_mask = all & ~B;
    for (int q = 0; q < n; ++q)
      if (!(B >> q & 1) && sig[q] > q) {
        st.pa[st.npairs] = (uint8_t)q;
        st.pb[st.npairs] = (uint8_t)sig[q];
        ++st.npairs;
      }
    const uint64_t nouter = 1ull << (n - st.nb);
    if (n - st.nb > 28) return fail(HSV_ERR_LIMIT, "in-place permutation: more than 2^28 tiles");
    // partials: s->partial[0 .. grid), then norm^2 and the factor
    const int grid = (int)std::min<uint64_t>(nouter, std::min<uint64_t>(uint64_t(s->num_cu) * 8, 2 * kRedBlocks - 2));
    double* norm_out = renorm && first ? s->partial : nullptr;
    const double* scale = renorm && !first ? s->partial + 2 * kRedBlocks - 1 : nullptr;
    if (s->dtype == HSV_C128)
      hipLaunchKernelGGL((k_swap_tiles<double>), grid, kSwapThreads, 0, s->stream, amp_as<double>(s), nouter, st,
                         norm_out, scale);
    else
      hipLaunchKernelGGL((k_swap_tiles<float>), grid, kSwapThreads, 0, s->stream, amp_as<float>(s), nouter, st,
                         norm_out, scale);
    HIPCHK(hipGetLastError());
    if (norm_out) {
      hipLaunchKernelGGL(k_norm_scale, 1, 64, 0, s->stream, s->partial, grid, s->partial + 2 * kRedBlocks - 2);
      HIPCHK(hipGetLastError());
    }
  }
  if (renormalized) *renormalized = renorm;
  return end_timed(s);
}

// Qubit permutation: amplitude bit s moves to bit new_slot[s]. Out of place
// (one k_permute pass into a second buffer, which then replaces the state)
// when a second state fits in free HBM next to a 1 GiB margin; otherwise, or
// from 32 qubits on (a 33-qubit complex128 state is 137 GB: two of them leave
// no room for anything else), or when permute_in_place is 1, in place as two
// k_swap_tiles involution passes (twice the traffic, no allocation).
static int permute_impl(hsv_state* s, const int* new_slot) {
  uint64_t seen = 0;
  bool identity = true;
  for (int q = 0; q < s->n; ++q) {
    if (new_slot[q] < 0 || new_slot[q] >= s->n || (seen >> new_slot[q] & 1))
      return fail(HSV_ERR_ARG, "new_slot is not a permutation of 0..n-1");
    seen |= 1ull << new_slot[q];
    identity = identity && new_slot[q] == q;
  }
  if (identity) return HSV_OK;
  const size_t bytes = s->len * s->elem();
  bool in_place = s->permute_in_place == 1 || (s->permute_in_place == 0 && s->n >= 32);
  if (s->permute_in_place == 0 && !in_place) {
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    in_place = free_b < bytes + (size_t(1) << 30);
  }
  if (in_place && s->n >= 10) return permute_in_place(s, new_slot);
  if (s->permute_in_place == 1) return fail(HSV_ERR_LIMIT, "the in-place permutation needs at least 10 qubits");
  void* out = nullptr;
  if (hipMalloc(&out, bytes) != hipSuccess)
    return fail(HSV_ERR_ALLOC, "no room for the permutation's second buffer");
  int* dslot = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&dslot), sizeof(int) * 64) != hipSuccess) {
    hipFree(out);
    return fail(HSV_ERR_ALLOC, "permutation table allocation failed");
  }
  HIPCHK(hipMemcpyAsync(dslot, new_slot, sizeof(int) * s->n, hipMemcpyHostToDevice, s->stream));
  if (int rc = begin_timed(s)) return rc;
  const int grid = grid_for(s->len, 256, uint64_t(s->num_cu) * 8);
  if (s->dtype == HSV_C128)
    hipLaunchKernelGGL((k_permute<double>), grid, 256, 0, s->stream, amp_as<double>(s),
                       static_cast<double2*>(out), s->len, s->n, dslot);
  else
    hipLaunchKernelGGL((k_permute<float>), grid, 256, 0, s->stream, amp_as<float>(s),
                       static_cast<float2*>(out), s->len, s->n, dslot);
  HIPCHK(hipGetLastError());
  if (int rc = end_timed(s)) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipFree(s->amp));
  HIPCHK(hipFree(dslot));
  s->amp = out;
  return HSV_OK;
}

// Device scratch of hsv_apply_fused_chain_noalloc: one static array per GPU
// (each device holds its own copy until the process ends). The first
// kNoAllocPartialBytes hold the FP64 reduction partials of the
// renormalisation (2 * kRedBlocks doubles, the size of a state's own
// partials); the rest the fused-step tables and BF16 fragments of one part
// at a time. Host code never passes `scratch` itself to a kernel: that name
// is a host-side stand-in. noalloc_scratch_ptr() looks up the device address
// with hipGetSymbolAddress, once per device with that device current.
constexpr size_t kNoAllocScratchBytes = 256 * 1024;
constexpr size_t kNoAllocPartialBytes = sizeof(double) * 2 * kRedBlocks;
static_assert(kNoAllocPartialBytes % 256 == 0 && kNoAllocPartialBytes < kNoAllocScratchBytes, "scratch layout");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch_ptr(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(HSV_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(HSV_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return HSV_OK;
}
}  // namespace

// ============================================================================
extern "C" {

const char* hsv_last_error(void) { return g_err.c_str(); }

int hsv_create(int num_qubits, int dtype, hsv_state** out) {
  if (!out) return fail(HSV_ERR_ARG, "null output handle");
  *out = nullptr;
  if (num_qubits < 1 || num_qubits > 36) return fail(HSV_ERR_ARG, "num_qubits must be in [1, 36]");
  if (dtype != HSV_C128 && dtype != HSV_C64) return fail(HSV_ERR_ARG, "dtype must be HSV_C128 or HSV_C64");
  hsv_state* s = new hsv_state();
  s->n = num_qubits;
  s->dtype = dtype;
  s->len = 1ull << num_qubits;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  s->device = dev;
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  s->max_lds = (int)prop.sharedMemPerBlock;
  s->num_cu = prop.multiProcessorCount;
  if (hipStreamCreate(&s->stream) != hipSuccess) { delete s; return fail(HSV_ERR_DEVICE, "hipStreamCreate failed"); }
  if (hipMalloc(&s->amp, s->len * s->elem()) != hipSuccess) {
    hipStreamDestroy(s->stream);
    delete s;
    return fail(HSV_ERR_ALLOC, "cannot allocate " + std::to_string(s->len * (dtype == HSV_C128 ? 16 : 8)) +
                                   " bytes of device memory for the state");
  }
  if (hipMalloc(reinterpret_cast<void**>(&s->partial), sizeof(double) * 2 * kRedBlocks) != hipSuccess) {
    hipFree(s->amp);
    hipStreamDestroy(s->stream);
    delete s;
    return fail(HSV_ERR_ALLOC, "cannot allocate reduction scratch");
  }
  s->host_partial.resize(2 * kRedBlocks);
  *out = s;
  return hsv_reset(s);
}

int hsv_destroy(hsv_state* s) {
  if (!s) return HSV_OK;
  hipStreamSynchronize(s->stream);
  hipFree(s->amp);
  hipFree(s->partial);
  if (s->gate_buf) hipFree(s->gate_buf);
  if (s->fused_scratch.ptr) hipFree(s->fused_scratch.ptr);
  for (auto& e : s->events) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  hipStreamDestroy(s->stream);
  delete s;
  return HSV_OK;
}

int hsv_num_qubits(const hsv_state* s) { return s ? s->n : HSV_ERR_ARG; }
int hsv_dtype(const hsv_state* s) { return s ? s->dtype : HSV_ERR_ARG; }

int hsv_reset(hsv_state* s) {
  if (!s) return fail(HSV_ERR_ARG, "null handle");
  const int grid = grid_for(s->len, 256, 1u << 16);
  if (s->dtype == HSV_C128)
    hipLaunchKernelGGL((k_zero_state<double>), grid, 256, 0, s->stream, amp_as<double>(s), s->len);
  else
    hipLaunchKernelGGL((k_zero_state<float>), grid, 256, 0, s->stream, amp_as<float>(s), s->len);
  HIPCHK(hipGetLastError());
  s->basis0 = true;
  return HSV_OK;
}

int hsv_apply_gate(hsv_state* s, const hsv_gate* g) {
  if (!s || !g) return fail(HSV_ERR_ARG, "null argument");
  Lowered L;
  std::string err;
  if (int rc = lower(*g, &L, &err)) return fail(rc, err);
  s->basis0 = false;
  if (int rc = begin_timed(s)) return rc;
  if (int rc = launch_gate(s, L)) return rc;
  g_gate_count++;
  return end_timed(s);
}

int hsv_apply_part(hsv_state* s, const hsv_gate* gates, int num_gates, int opt_slots) {
  if (!s || (num_gates > 0 && !gates) || num_gates < 0) return fail(HSV_ERR_ARG, "bad arguments");
  std::vector<Lowered> Ls;
  if (int rc = lower_all(gates, num_gates, &Ls, s)) return rc;
  uint64_t part = 0;
  for (const Lowered& L : Ls) {
    for (int i = 0; i < L.nt; ++i) part |= 1ull << L.t[i];
    for (int i = 0; i < L.nc; ++i) part |= 1ull << L.c[i];
  }
  if (opt_slots) part |= (s->n >= 4 ? 0xFull : ((1ull << s->n) - 1));  // execute.hpp:184-199
  s->basis0 = false;
  // part_kernel 2: every part whose qubits and slots 0-4 fit a tile takes
  // the FP64 tile kernel (it pads the tile with the lowest free slots itself)
  const int kt = tile_part_bits(s);
  const uint64_t low5 = (1ull << std::min(s->n, 5)) - 1;
  const bool tile = s->part_kernel == 2 && s->n >= kt && __builtin_popcountll(part | low5) <= kt;
  const int kmax = max_part_qubits(s);
  // Slots 0..2 make every gathered run at least 8 consecutive amplitudes; they
  // join the part whenever the LDS budget allows and, for the register
  // kernel, up to its widest shape of 13 qubits (results are unchanged).
  const int kpad = s->part_kernel >= 1 ? std::min(kmax, 13) : kmax;
  for (int b = 0; !tile && b < s->low_slots && b < s->n; ++b)
    if (!(part >> b & 1) && __builtin_popcountll(part) < kpad) part |= 1ull << b;
  if (int rc = begin_timed(s)) return rc;
  if (__builtin_popcountll(part) > kmax) {
    // Too wide for one workgroup's LDS: apply the part gate by gate.
    for (const Lowered& L : Ls)
      if (int rc = launch_gate(s, L)) return rc;
  } else if (num_gates > 0) {
    const int k = __builtin_popcountll(part);
    const bool reg = !tile && s->part_kernel >= 1 && k >= 8 && k <= 13;
    int rc = tile  ? (s->dtype == HSV_C128 ? launch_tile_part<double>(s, Ls, part) : launch_tile_part<float>(s, Ls, part))
             : reg ? (s->dtype == HSV_C128 ? launch_reg_part_t<double>(s, Ls, part) : launch_reg_part_t<float>(s, Ls, part))
                   : (s->dtype == HSV_C128 ? launch_part_t<double>(s, Ls, part) : launch_part_t<float>(s, Ls, part));
    if (rc) return rc;
  }
  g_part_count++;
  g_gate_count += num_gates;
  return end_timed(s);
}

int hsv_apply_fused(hsv_state* s, const hsv_gate* gates, int num_gates, int mode) {
  if (!s || (num_gates > 0 && !gates) || num_gates < 0) return fail(HSV_ERR_ARG, "bad arguments");
  std::vector<Lowered> Ls;
  if (int rc = lower_all(gates, num_gates, &Ls, s)) return rc;
  s->basis0 = false;
  if (int rc = begin_timed(s)) return rc;
  std::string err;
  int rc = fused_apply(s->amp, s->n, s->dtype, s->stream, s->num_cu, &s->fused_scratch, Ls, mode, &err,
                       s->bf);
  if (rc) return fail(rc, err);
  g_part_count++;
  g_gate_count += num_gates;
  return end_timed(s);
}

int hsv_apply_fused_chain(hsv_state* s, const hsv_gate* gates, int num_gates, const int* part_sizes,
                          int num_parts, const int* lows, int mode) {
  if (!s || num_gates < 0 || num_parts < 0 || (num_gates > 0 && !gates) ||
      (num_parts > 0 && (!part_sizes || !lows)))
    return fail(HSV_ERR_ARG, "bad arguments");
  const bool renorm = (mode & HSV_CHAIN_RENORMALIZE) != 0;
  mode &= ~HSV_CHAIN_RENORMALIZE;
  if (mode != HSV_FUSED_BF16SPLIT && mode != HSV_FUSED_BF16X3)
    return fail(HSV_ERR_ARG, "the part chain runs on the BF16-split paths only");
  const int pieces = mode == HSV_FUSED_BF16X3 ? 3 : 2;
  const int n = s->n;
  if (n < 12) return fail(HSV_ERR_LIMIT, "the part chain needs a state of at least 12 qubits");
  long total = 0;
  for (int p = 0; p < num_parts; ++p) {
    if (part_sizes[p] < 0) return fail(HSV_ERR_ARG, "negative part size");
    total += part_sizes[p];
    uint64_t seen = 0;
    for (int i = 0; i < 5; ++i) {
      const int q = lows[5 * p + i];
      if (q < 0 || q >= n || (seen >> q & 1)) return fail(HSV_ERR_ARG, "lows: 5 distinct qubits per part");
      seen |= 1ull << q;
    }
  }
  if (total != num_gates) return fail(HSV_ERR_ARG, "part sizes do not add up to num_gates");
  std::vector<Lowered> Ls;
  if (int rc = lower_all(gates, num_gates, &Ls, s)) return rc;
  // the layouts before and after each part, worked out and checked first
  std::vector<std::vector<int>> layout(num_parts + 1, std::vector<int>(n));
  std::vector<uint64_t> tile(num_parts, 0);
  for (int q = 0; q < n; ++q) layout[0][q] = q;
  size_t off = 0;
  for (int p = 0; p < num_parts; ++p) {
    const std::vector<int>& phys = layout[p];
    uint64_t& tq = tile[p];
    for (int i = 0; i < 5; ++i) {
      const int q = lows[5 * p + i];
      if (phys[q] >= 5) return fail(HSV_ERR_ARG, "lows of part " + std::to_string(p) + " are not the qubits in slots 0-4");
      tq |= 1ull << q;
    }
    for (size_t g = off; g < off + part_sizes[p]; ++g) {
      const Lowered& L = Ls[g];
      if (L.kind == K_IDENT) continue;
      for (int i = 0; i < L.nt; ++i) tq |= 1ull << L.t[i];
      for (int i = 0; i < L.nc; ++i) tq |= 1ull << L.c[i];
    }
    off += part_sizes[p];
    std::vector<int>& after = layout[p + 1];
    after = phys;
    if (p + 1 < num_parts) {
      // bring the next part's lows into slots 0-4, each swapping slots with
      // a current occupant the next part does not list
      const int* nl = lows + 5 * (p + 1);
      uint64_t want = 0;
      for (int i = 0; i < 5; ++i) {
        if (!(tq >> nl[i] & 1))
          return fail(HSV_ERR_ARG, "lows of part " + std::to_string(p + 1) + " must be qubits of part " +
                                       std::to_string(p) + " or of its lows");
        want |= 1ull << nl[i];
      }
      std::vector<int> spare;
      for (int i = 0; i < 5; ++i)
        if (!(want >> lows[5 * p + i] & 1)) spare.push_back(lows[5 * p + i]);
      for (int i = 0; i < 5; ++i) {
        const int q = nl[i];
        if (after[q] < 5) continue;
        const int o = spare.back();
        spare.pop_back();
        std::swap(after[q], after[o]);
      }
    } else {
      // the last part sends each of its qubits whose own slot is among its
      // slots home; the others fill the remaining slots in order
      std::vector<int> slots, rest, free_slots;
      for (int q = 0; q < n; ++q)
        if (tq >> q & 1) slots.push_back(phys[q]);
      std::sort(slots.begin(), slots.end());
      uint64_t home = 0;  // slots whose own qubit is in the part
      for (int sl : slots)
        if (tq >> sl & 1) home |= 1ull << sl;
      for (int sl : slots)
        if (!(home >> sl & 1)) free_slots.push_back(sl);
      for (int sl : slots) {
        int q = 0;
        while (phys[q] != sl) ++q;
        if (home >> q & 1) after[q] = q;
        else rest.push_back(q);
      }
      for (size_t i = 0; i < rest.size(); ++i) after[rest[i]] = free_slots[i];
    }
  }
  // back to the standard layout: the bit in slot phys[q] moves to slot q;
  // one tile pass without gates when the qubits still away from home and
  // those in slots 0-4 fit a tile, else two in-place swap passes
  auto restore = [&](const std::vector<int>& phys, bool with_renorm) -> int {
    std::vector<int> new_slot(n), ident(n);
    uint64_t away = 0, lowq = 0;
    for (int q = 0; q < n; ++q) {
      new_slot[phys[q]] = q;
      ident[q] = q;
      if (phys[q] != q) away |= 1ull << q;
      if (phys[q] < 5) lowq |= 1ull << q;
    }
    bool done = false;
    if (!away) {
    } else if (__builtin_popcountll(away | lowq) <= 14) {
      if (int rc = begin_timed(s)) return rc;
      std::string err;
      if (int rc = fused_part_bf_layout(s->amp, n, s->dtype, s->stream, &s->fused_scratch, {}, away | lowq,
                                        phys.data(), ident.data(), ~0ull, s->bf, &err))
        return fail(rc, "layout restore: " + err);
      if (int rc = end_timed(s)) return rc;
    } else if (int rc = permute_in_place(s, new_slot.data(), with_renorm, &done)) {
      return rc;
    }
    // the renormalisation, when the swap passes could not carry it
    if (with_renorm && !done) {
      const int rc = hsv_renormalize(s, nullptr);
      if (rc == HSV_ERR_ARG) return HSV_OK;   // a zero state stays as it is
      return rc;
    }
    return HSV_OK;
  };
  // starting from |0...0>, a qubit no gate has touched yet is |0> everywhere:
  // the passes skip the tiles with a 1 in its slot (the first parts run on a
  // small corner of the state)
  const uint64_t all_q = (n == 64) ? ~0ull : ((1ull << n) - 1);
  uint64_t support = s->basis0 ? 0 : all_q;
  s->basis0 = false;
  off = 0;
  for (int p = 0; p < num_parts; ++p) {
    const std::vector<Lowered> part(Ls.begin() + off, Ls.begin() + off + part_sizes[p]);
    off += part_sizes[p];
    uint64_t live = 0;
    for (int q = 0; q < n; ++q)
      if (support >> q & 1) live |= 1ull << layout[p][q];
    if (int rc = begin_timed(s)) return rc;
    std::string err;
    if (int rc = fused_part_bf_layout(s->amp, n, s->dtype, s->stream, &s->fused_scratch, part, tile[p],
                                      layout[p].data(), layout[p + 1].data(), live, s->bf, &err, pieces)) {
      // close this part's timing pair, then bring the parts applied so far
      // back to the standard layout; a failed restore is reported too
      std::string msg = "part " + std::to_string(p) + ": " + err;
      if (end_timed(s)) msg += "; " + g_err;
      if (restore(layout[p], false)) msg += "; the layout restore failed as well (" + g_err + ")";
      return fail(rc, msg);
    }
    for (const Lowered& L : part) {
      if (L.kind == K_IDENT) continue;
      for (int i = 0; i < L.nt; ++i) support |= 1ull << L.t[i];
      for (int i = 0; i < L.nc; ++i) support |= 1ull << L.c[i];
    }
    if (int rc = end_timed(s)) return rc;
    g_part_count++;
    g_gate_count += part_sizes[p];
  }
  return restore(layout[num_parts], renorm);
}

int hsv_apply_fused_chain_noalloc(hsv_state* s, const hsv_gate* gates, int num_gates, const int* part_sizes,
                                  int num_parts, const int* lows, int mode) {
  if (!s) return fail(HSV_ERR_ARG, "null handle");
  void* p = nullptr;
  if (int rc = noalloc_scratch_ptr(s->device, &p)) return rc;
  // the same chain, with the static array in place of the state's own
  // partials and fused scratch for the duration of the call
  const DeviceScratch own = s->fused_scratch;
  double* const own_partial = s->partial;
  s->partial = static_cast<double*>(p);
  s->fused_scratch = DeviceScratch{static_cast<unsigned char*>(p) + kNoAllocPartialBytes,
                                   kNoAllocScratchBytes - kNoAllocPartialBytes, true};
  const int rc = hsv_apply_fused_chain(s, gates, num_gates, part_sizes, num_parts, lows, mode);
  s->fused_scratch = own;
  s->partial = own_partial;
  return rc;
}

int hsv_permute_qubits(hsv_state* s, const int* new_slot) {
  if (!s || !new_slot) return fail(HSV_ERR_ARG, "null argument");
  s->basis0 = false;
  return permute_impl(s, new_slot);
}


int hsv_read(hsv_state* s, uint64_t offset, uint64_t count, double* host_out) {
  if (!s || (count && !host_out) || offset > s->len || count > s->len - offset)
    return fail(HSV_ERR_ARG, "read range out of bounds");
  if (!count) return HSV_OK;
  if (s->dtype == HSV_C128) {
    HIPCHK(hipMemcpyAsync(host_out, static_cast<char*>(s->amp) + offset * 16, count * 16, hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  } else {
    std::vector<float> tmp(2 * count);
    HIPCHK(hipMemcpyAsync(tmp.data(), static_cast<char*>(s->amp) + offset * 8, count * 8, hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (uint64_t i = 0; i < 2 * count; ++i) host_out[i] = tmp[i];
  }
  return HSV_OK;
}

int hsv_write(hsv_state* s, uint64_t offset, uint64_t count, const double* host_in) {
  if (!s || (count && !host_in) || offset > s->len || count > s->len - offset)
    return fail(HSV_ERR_ARG, "write range out of bounds");
  if (!count) return HSV_OK;
  s->basis0 = false;
  if (s->dtype == HSV_C128) {
    HIPCHK(hipMemcpyAsync(static_cast<char*>(s->amp) + offset * 16, host_in, count * 16, hipMemcpyHostToDevice,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  } else {
    std::vector<float> tmp(2 * count);
    for (uint64_t i = 0; i < 2 * count; ++i) tmp[i] = (float)host_in[i];
    HIPCHK(hipMemcpyAsync(static_cast<char*>(s->amp) + offset * 8, tmp.data(), count * 8, hipMemcpyHostToDevice,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  return HSV_OK;
}

int hsv_norm2(hsv_state* s, double* out) {
  if (!s || !out) return fail(HSV_ERR_ARG, "null argument");
  return reduce<RD_NORM2>(s, nullptr, 0, out, nullptr);
}

int hsv_inner(hsv_state* a, hsv_state* b, double* out2) {
  if (!a || !b || !out2) return fail(HSV_ERR_ARG, "null argument");
  if (a->n != b->n || a->dtype != b->dtype) return fail(HSV_ERR_ARG, "states differ in size or dtype");
  HIPCHK(hipStreamSynchronize(b->stream));
  return reduce<RD_INNER>(a, b, 0, &out2[0], &out2[1]);
}

int hsv_max_abs_diff(hsv_state* a, hsv_state* b, double* out) {
  if (!a || !out) return fail(HSV_ERR_ARG, "null argument");
  if (b && (a->n != b->n || a->dtype != b->dtype)) return fail(HSV_ERR_ARG, "states differ in size or dtype");
  if (b) HIPCHK(hipStreamSynchronize(b->stream));
  return reduce<RD_MAXDIFF>(a, b, 0, out, nullptr);
}

int hsv_probability(hsv_state* s, int qubit, int value, double* out) {
  if (!s || !out || qubit < 0 || qubit >= s->n || (value != 0 && value != 1))
    return fail(HSV_ERR_ARG, "bad qubit or value");
  double p0 = 0;
  if (int rc = reduce<RD_PROB0>(s, nullptr, qubit, &p0, nullptr)) return rc;
  *out = value == 0 ? p0 : 1.0 - p0;
  return HSV_OK;
}

int hsv_renormalize(hsv_state* s, double* norm2_before) {
  if (!s) return fail(HSV_ERR_ARG, "null handle");
  s->basis0 = false;
  double n2 = 0;
  if (int rc = reduce<RD_NORM2>(s, nullptr, 0, &n2, nullptr)) return rc;
  if (norm2_before) *norm2_before = n2;
  if (!(n2 > 0.0)) return fail(HSV_ERR_ARG, "cannot renormalise a zero state");
  const double scale = 1.0 / std::sqrt(n2);
  const int grid = grid_for(s->len, 256, 1u << 16);
  // k_project with an empty mask keeps every amplitude and scales it
  if (s->dtype == HSV_C128)
    hipLaunchKernelGGL((k_project<double>), grid, 256, 0, s->stream, amp_as<double>(s), s->len, 0ull, 0ull, scale);
  else
    hipLaunchKernelGGL((k_project<float>), grid, 256, 0, s->stream, amp_as<float>(s), s->len, 0ull, 0ull, scale);
  HIPCHK(hipGetLastError());
  return HSV_OK;
}

int hsv_project(hsv_state* s, uint64_t qubit_mask, uint64_t values, double* p_out) {
  if (!s) return fail(HSV_ERR_ARG, "null handle");
  if (qubit_mask >> s->n) return fail(HSV_ERR_ARG, "qubit mask names a qubit beyond the state");
  if (values & ~qubit_mask) return fail(HSV_ERR_ARG, "values has bits outside qubit_mask");
  s->basis0 = false;
  const int blocks = kRedBlocks;
  if (s->dtype == HSV_C128)
    hipLaunchKernelGGL((k_mass<double>), blocks, kRedBlock, 0, s->stream, amp_as<double>(s), s->len, qubit_mask,
                       values, s->partial);
  else
    hipLaunchKernelGGL((k_mass<float>), blocks, kRedBlock, 0, s->stream, amp_as<float>(s), s->len, qubit_mask,
                       values, s->partial);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(s->host_partial.data(), s->partial, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  double p = 0;
  for (int i = 0; i < blocks; ++i) p += s->host_partial[2 * i];
  if (p_out) *p_out = p;
  if (!(p > 0.0)) return fail(HSV_ERR_ARG, "the outcome has probability 0");
  const double scale = 1.0 / std::sqrt(p);
  const int grid = grid_for(s->len, 256, 1u << 16);
  if (s->dtype == HSV_C128)
    hipLaunchKernelGGL((k_project<double>), grid, 256, 0, s->stream, amp_as<double>(s), s->len, qubit_mask, values,
                       scale);
  else
    hipLaunchKernelGGL((k_project<float>), grid, 256, 0, s->stream, amp_as<float>(s), s->len, qubit_mask, values,
                       scale);
  HIPCHK(hipGetLastError());
  return HSV_OK;
}

int hsv_expect_pauli(hsv_state* s, uint64_t x_mask, uint64_t z_mask, double* out2) {
  if (!s || !out2) return fail(HSV_ERR_ARG, "null argument");
  if ((x_mask | z_mask) >> s->n) return fail(HSV_ERR_ARG, "Pauli mask names a qubit beyond the state");
  const int blocks = kRedBlocks;
  if (s->dtype == HSV_C128)
    hipLaunchKernelGGL((k_pauli<double>), blocks, kRedBlock, 0, s->stream, amp_as<double>(s), s->len, x_mask,
                       z_mask, s->partial);
  else
    hipLaunchKernelGGL((k_pauli<float>), blocks, kRedBlock, 0, s->stream, amp_as<float>(s), s->len, x_mask,
                       z_mask, s->partial);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(s->host_partial.data(), s->partial, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  double re = 0, im = 0;
  for (int i = 0; i < blocks; ++i) {
    re += s->host_partial[2 * i];
    im += s->host_partial[2 * i + 1];
  }
  // multiply by i^ny, ny = number of Y factors
  switch (__builtin_popcountll(x_mask & z_mask) & 3) {
    case 0: out2[0] = re; out2[1] = im; break;
    case 1: out2[0] = -im; out2[1] = re; break;
    case 2: out2[0] = -re; out2[1] = -im; break;
    default: out2[0] = im; out2[1] = -re; break;
  }
  return HSV_OK;
}

int hsv_measure(hsv_state* s, const int* qubits, int num_qubits, const double* uniforms, int num_samples,
                int32_t* out) {
  if (!s || num_qubits < 0 || num_samples < 0 || (num_qubits && !qubits) || (num_samples && (!uniforms || !out)))
    return fail(HSV_ERR_ARG, "bad arguments");
  for (int j = 0; j < num_qubits; ++j)
    if (qubits[j] < 0 || qubits[j] >= s->n) return fail(HSV_ERR_ARG, "measured qubit out of range");
  if (!num_samples) return HSV_OK;
  std::vector<double> u(uniforms, uniforms + num_samples);
  std::sort(u.begin(), u.end());
  // Probability mass per contiguous chunk, then the reference's sequential
  // walk (state_vector.hpp:946-972) inside the chunks the samples land in.
  int blocks = s->dtype == HSV_C128 ? reduce_t<double, RD_CHUNK>(s, nullptr, 0, nullptr, nullptr)
                                    : reduce_t<float, RD_CHUNK>(s, nullptr, 0, nullptr, nullptr);
  if (blocks < 0) return blocks;
  const uint64_t chunk = s->len / blocks;
  std::vector<double> amps(2 * chunk);
  double cum = 0;
  int sp = 0;
  for (int c = 0; c < blocks && sp < num_samples; ++c) {
    const double mass = s->host_partial[2 * c];
    if (!(u[sp] < cum + mass)) { cum += mass; continue; }
    if (int rc = hsv_read(s, uint64_t(c) * chunk, chunk, amps.data())) return rc;
    for (uint64_t i = 0; i < chunk && sp < num_samples; ++i) {
      const double x = amps[2 * i], y = amps[2 * i + 1];
      const double p = x * x + y * y;
      const uint64_t pos = uint64_t(c) * chunk + i;
      while (sp < num_samples && u[sp] < cum + p) {
        for (int j = 0; j < num_qubits; ++j) out[sp * num_qubits + j] = (int32_t)((pos >> qubits[j]) & 1);
        ++sp;
      }
      cum += p;
    }
  }
  for (; sp < num_samples; ++sp)  // mass exhausted: the reference reports all ones (:973-980)
    for (int j = 0; j < num_qubits; ++j) out[sp * num_qubits + j] = 1;
  return HSV_OK;
}

int hsv_gate_matrix(const hsv_gate* g, double* out, int* dim) {
  if (!g || !out || !dim) return fail(HSV_ERR_ARG, "null argument");
  Lowered L;
  std::string err;
  if (int rc = lower(*g, &L, &err)) return fail(rc, err);
  cd m[16];
  dense_of(L, m);
  *dim = L.nt == 2 ? 4 : 2;
  for (int i = 0; i < 16; ++i) { out[2 * i] = m[i].real(); out[2 * i + 1] = m[i].imag(); }
  if (*dim == 2) {  // dense_of leaves a 2x2 in the first four entries
    for (int i = 4; i < 16; ++i) out[2 * i] = out[2 * i + 1] = 0.0;
  }
  return HSV_OK;
}

int hsv_fused_unitary(int num_qubits, const hsv_gate* gates, int num_gates, int* qubits_out, int* num_part_qubits,
                      double* u_out, double* unitarity_err) {
  if ((num_gates > 0 && !gates) || !qubits_out || !num_part_qubits || !u_out) return fail(HSV_ERR_ARG, "null argument");
  std::vector<Lowered> Ls(num_gates);
  for (int i = 0; i < num_gates; ++i) {
    std::string err;
    if (int rc = lower(gates[i], &Ls[i], &err)) return fail(rc, err);
    for (int j = 0; j < Ls[i].nt; ++j)
      if (Ls[i].t[j] >= num_qubits) return fail(HSV_ERR_ARG, "qubit out of range");
    for (int j = 0; j < Ls[i].nc; ++j)
      if (Ls[i].c[j] >= num_qubits) return fail(HSV_ERR_ARG, "qubit out of range");
  }
  std::vector<int> Q;
  std::vector<cd> U;
  std::string err;
  if (int rc = fused_unitary(num_qubits, Ls, &Q, &U, &err)) return fail(rc, err);
  *num_part_qubits = (int)Q.size();
  for (size_t i = 0; i < Q.size(); ++i) qubits_out[i] = Q[i];
  for (size_t i = 0; i < U.size(); ++i) { u_out[2 * i] = U[i].real(); u_out[2 * i + 1] = U[i].imag(); }
  if (unitarity_err) *unitarity_err = unitarity_error(U, 1 << Q.size());
  return HSV_OK;
}

int hsv_synchronize(hsv_state* s) {
  if (!s) return fail(HSV_ERR_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(s->stream));
  return HSV_OK;
}

int hsv_tune(hsv_state* s, const char* key, int value) {
  if (!s || !key) return fail(HSV_ERR_ARG, "null argument");
  const std::string k = key;
  if (k == "reg_waves" && value >= 1 && value <= 64) s->reg_waves = value;
  else if (k == "part_kernel" && value >= 0 && value <= 2) s->part_kernel = value;
  else if (k == "reg_bits" && value >= 0 && value <= 5) s->reg_bits = value;
  else if (k == "tile_bits" && (value == 0 || value == tile_part_default(s) || value == tile_part_default(s) - 1))
    s->tile_bits = value;
  else if (k == "low_slots" && value >= 0 && value <= 6) s->low_slots = value;
  else if (k == "io_bits" && value >= 0 && value <= 6) s->io_bits = value;
  else if (k == "fused_variant" && value >= 0 && value <= 2) s->bf.variant = value;
  else if (k == "fused_tiles_per_wg" && value >= 1 && value <= 64) s->bf.tile_tpw = value;
  else if (k == "part_tiles_per_wg" && value >= 0 && value <= 64) s->bf.part_tpw = value;
  else if (k == "tile_order" && value >= 0 && value <= 16) s->bf.tile_order = value;
  else if (k == "part13_u" && value >= 0 && value <= 2) s->bf.part13 = value;
  else if (k == "permute_in_place" && value >= 0 && value <= 2) s->permute_in_place = value;
  else return fail(HSV_ERR_ARG, "unknown tuning key or value out of range: " + k);
  return HSV_OK;
}

int hsv_set_timing(hsv_state* s, int on) {
  if (!s) return fail(HSV_ERR_ARG, "null handle");
  s->timing = on ? 1 : 0;
  s->events_used = 0;
  return HSV_OK;
}

int hsv_kernel_times(hsv_state* s, double* ms_out, int capacity, int* count) {
  if (!s || !count || (capacity > 0 && !ms_out)) return fail(HSV_ERR_ARG, "null argument");
  const int n = (int)s->events_used;
  *count = n;
  for (int i = 0; i < n && i < capacity; ++i) {
    HIPCHK(hipEventSynchronize(s->events[i].second));
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, s->events[i].first, s->events[i].second));
    ms_out[i] = f;
    g_part_us += (long long)(f * 1000.0f);
  }
  s->events_used = 0;
  return HSV_OK;
}

void hsv_set_opt_slots(int on) { g_opt_slots = on ? 1 : 0; }
int hsv_get_opt_slots(void) { return g_opt_slots; }

void hsv_get_counters(long long* apply_us, long long* gates, long long* parts, long long* part_us) {
  if (apply_us) *apply_us = g_apply_us;
  if (gates) *gates = g_gate_count;
  if (parts) *parts = g_part_count;
  if (part_us) *part_us = g_part_us;
}

}  // extern "C"
