// CPU oracle for the hierarchical state-vector hot path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU, the amplitude
// arithmetic of the reference HiSVSIM simulator so that the HIP product can be
// checked against it. Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may load the library built from it; the product never does.
//
// What it follows (reference file:line):
//   gate formulas (scalar, non-AVX512 branches) ... basic_gates.hpp:99-1408
//   gate name -> type table ........................ basic_gates.hpp:1487-1522
//   per-gate sweep StateVector::apply .............. state_vector.hpp:402-510
//   loop enumeration (offset = pdep(i, free mask)) . state_vector.hpp:267-283,
//                                                    loop.hpp:468-962
//   part driver gather -> execute -> scatter ....... execute.hpp:318-470
//     (needed qubits, fast slots 0..3, slot order) . execute.hpp:174-213
//   qubit-exchange permutation order ............... svsim-mpi.hpp:123-173
//   probability / measure / project ................ state_vector.hpp:829-981
//
// Arithmetic is written with std::complex<double> in the same operation order
// as the reference functors. Compiled with -ffp-contract=off (oracle/Makefile),
// it is bit-identical to a reference build with the same flag for every gate
// type and for the partitioned path. A reference build with GCC's default
// contraction on an FMA host differs from it by a few ulps (<= 2.5e-16 per gate).
//
// Parity pins (see tests/test_oracle_golden.py): golden vectors written by a
// build of the reference headers (oracle/_ref), the disabled doctest known
// answers of state_vector.hpp:1396-1739, closed forms (BV, QFT via FFT) and
// brute-force dense Kronecker products.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

using cd = std::complex<double>;
using u64 = uint64_t;

namespace {

// Numbering mirrors SvSim::GateType (basic_gates.hpp:12-45).
enum GT : int {
  G_INVALID = 0, G_ALLOC, G_FREE, G_I, G_X, G_Y, G_Z, G_H, G_S, G_T,
  G_RX, G_RY, G_RZ, G_RI, G_R1, G_RXFRAC, G_RYFRAC, G_RZFRAC, G_RIFRAC,
  G_R1FRAC, G_SWAP, G_C1, G_C2, G_U1, G_U2, G_U3, G_RXX, G_RYY, G_RZZ,
  G_TDG, G_SDG, G_COUNT
};

constexpr double kInvSqrt2 = 0.707106781186547524401;  // types.hpp:23
constexpr double kPi = 3.14159265358979323846;         // types.hpp:19
constexpr double kPiBy2 = 1.57079632679489661923;      // types.hpp:20

// ---- elementary 1-slot operations on (b[0], b[1<<s]) -------------------
// Each mirrors one reference functor's scalar branch.

struct RXc { cd c, is; };
RXc rx_make(cd theta) {  // basic_gates.hpp:343-350
  using namespace std::complex_literals;
  return {std::cos(theta / cd{2}), 1.0i * std::sin(theta / cd{2})};
}
void rx_apply(const RXc& g, cd* b, int s) {  // basic_gates.hpp:356-363
  cd v0 = b[0], v1 = b[1ul << s];
  b[0] = g.c * v0 - g.is * v1;
  b[1ul << s] = -g.is * v0 + g.c * v1;
}

struct RYc { cd c, s; };
RYc ry_make(cd theta) {  // basic_gates.hpp:401-407
  return {std::cos(theta / cd{2}), std::sin(theta / cd{2})};
}
void ry_apply(const RYc& g, cd* b, int s) {  // basic_gates.hpp:415-422
  cd v0 = b[0], v1 = b[1ul << s];
  b[0] = g.c * v0 - g.s * v1;
  b[1ul << s] = g.s * v0 + g.c * v1;
}

struct RZc { cd neg, pos; };
RZc rz_make(cd theta) {  // basic_gates.hpp:515-519
  using namespace std::complex_literals;
  return {std::exp(-1.0i * theta / cd{2}), std::exp(+1.0i * theta / cd{2})};
}
void rz_apply(const RZc& g, cd* b, int s) {  // basic_gates.hpp:530-535
  cd v0 = b[0], v1 = b[1ul << s];
  b[0] = v0 * g.neg;
  b[1ul << s] = v1 * g.pos;
}

struct RIc { cd neg; };
RIc ri_make(cd theta) {  // basic_gates.hpp:460-463
  using namespace std::complex_literals;
  return {std::exp(-1.0i * theta / cd{2})};
}
void ri_apply(const RIc& g, cd* b, int s) {  // basic_gates.hpp:474-478
  cd v0 = b[0], v1 = b[1ul << s];
  b[0] = v0 * g.neg;
  b[1ul << s] = v1 * g.neg;
}

void h_apply(cd* b, int s) {  // basic_gates.hpp:183-189
  cd v0 = b[0], v1 = b[1ul << s];
  b[0] = kInvSqrt2 * (v0 + v1);
  b[1ul << s] = kInvSqrt2 * (v0 - v1);
}

struct C1c { cd c00, c01, c10, c11; };
void c1_apply(const C1c& g, cd* b, int s) {  // basic_gates.hpp:852-858
  cd v0 = b[0], v1 = b[1ul << s];
  b[0] = g.c00 * v0 + g.c01 * v1;
  b[1ul << s] = g.c10 * v0 + g.c11 * v1;
}

C1c u3_make(cd theta, cd phi, cd lambda) {  // basic_gates.hpp:1042-1053
  using namespace std::complex_literals;
  cd e0 = std::cos(theta / 2.0) * std::cos((-phi - lambda) / 2.0) +
          1.0i * std::cos(theta / 2.0) * std::sin((-phi - lambda) / 2.0);
  cd e1 = -std::sin(theta / 2.0) * std::cos((-phi + lambda) / 2.0) +
          1.0i * -std::sin(theta / 2.0) * std::sin((-phi + lambda) / 2.0);
  cd e2 = std::sin(theta / 2.0) * std::cos((phi - lambda) / 2.0) +
          1.0i * std::sin(theta / 2.0) * std::sin((phi - lambda) / 2.0);
  cd e3 = std::cos(theta / 2.0) * std::cos((phi + lambda) / 2.0) +
          1.0i * std::cos(theta / 2.0) * std::sin((phi + lambda) / 2.0);
  return {e0, e1, e2, e3};
}

C1c u2_make(cd phi, cd lambda) {  // basic_gates.hpp:1015-1031
  using namespace std::complex_literals;
  cd e0 = kInvSqrt2 + 0.0i;
  cd e1 = -kInvSqrt2 * (std::cos(lambda) + 1.0i * std::sin(lambda));
  cd e2 = kInvSqrt2 * (std::cos(phi) + 1.0i * std::sin(phi));
  cd e3 = kInvSqrt2 * (std::cos(phi + lambda) + 1.0i * std::sin(phi + lambda));
  return {e0, e1, e2, e3};
}

// Angle of the *Frac gates: -pi * numerator / 2^(power-1)
// (basic_gates.hpp:733-780; unsigned arithmetic kept as in the reference).
double frac_angle(unsigned numerator, unsigned power) {
  return -kPi * numerator / std::pow(2, power - 1);
}

// ---- a prepared gate: constants computed once, as the reference ctors do --
struct Prepared {
  int type = G_INVALID;
  RXc rx{}; RYc ry{}; RZc rz{}; RIc ri{}; C1c c1{};
  RXc rx2{}; RXc rx3{};        // RYY uses RX(pi/2) and RX(-pi/2)
  cd c2row[4]{};               // C2: constructor keeps column 0 only
};

Prepared prepare(int type, const double* p) {
  auto P = [&](int i) { return cd{p[2 * i], p[2 * i + 1]}; };
  Prepared g;
  g.type = type;
  switch (type) {
    case G_RX: g.rx = rx_make(P(0)); break;
    case G_RY: g.ry = ry_make(P(0)); break;
    case G_RZ: case G_U1: g.rz = rz_make(P(0)); break;   // U1 = RZ (basic_gates.hpp:571)
    case G_RI: g.ri = ri_make(P(0)); break;
    case G_R1:  // R1(theta) = RZ(theta) then RI(-theta) (basic_gates.hpp:716-728)
      g.rz = rz_make(P(0)); g.ri = ri_make(-P(0)); break;
    case G_RXFRAC: g.rx = rx_make(cd{frac_angle((unsigned)p[0], (unsigned)p[2])}); break;
    case G_RYFRAC: g.ry = ry_make(cd{frac_angle((unsigned)p[0], (unsigned)p[2])}); break;
    case G_RZFRAC: g.rz = rz_make(cd{frac_angle((unsigned)p[0], (unsigned)p[2])}); break;
    case G_RIFRAC: g.ri = ri_make(cd{frac_angle((unsigned)p[0], (unsigned)p[2])}); break;
    case G_R1FRAC: {  // rzf_{-numerator, power+1}, rif_{numerator, power+1} (basic_gates.hpp:782-795)
      unsigned num = (unsigned)p[0], pw = (unsigned)p[2];
      g.rz = rz_make(cd{frac_angle(0u - num, pw + 1)});
      g.ri = ri_make(cd{frac_angle(num, pw + 1)});
      break;
    }
    case G_C1: g.c1 = {P(0), P(1), P(2), P(3)}; break;
    case G_C2:  // c0j = c00, c1j = c10, c2j = c20, c3j = c30 (basic_gates.hpp:972-987)
      g.c2row[0] = P(0); g.c2row[1] = P(4); g.c2row[2] = P(8); g.c2row[3] = P(12); break;
    case G_U2: g.c1 = u2_make(P(0), P(1)); break;
    case G_U3: g.c1 = u3_make(P(0), P(1), P(2)); break;
    case G_RXX: g.rz = rz_make(P(0)); break;
    case G_RZZ: g.rz = rz_make(P(0)); break;
    case G_RYY:  // basic_gates.hpp:1097
      g.rx2 = rx_make(cd{kPiBy2}); g.rx3 = rx_make(cd{-kPiBy2}); g.rz = rz_make(P(0)); break;
    default: break;
  }
  return g;
}

inline void swp(cd& a, cd& b) { std::swap(a, b); }

// Apply one prepared gate at `b` (already offset by base + control offset),
// `s` = target slots. Mirrors each functor's scalar branch.
void apply_at(const Prepared& g, cd* b, const int* s) {
  using namespace std::complex_literals;
  switch (g.type) {
    case G_I: break;
    case G_X: swp(b[0ul << s[0]], b[1ul << s[0]]); break;             // :115
    case G_Y: {                                                      // :140-146
      cd v0 = b[0ul << s[0]], v1 = b[1ul << s[0]];
      b[0ul << s[0]] = -1.0i * v1;
      b[1ul << s[0]] = 1.0i * v0;
      break;
    }
    case G_H: h_apply(b, s[0]); break;
    case G_Z: b[1ul << s[0]] *= -1.0; break;                         // :230
    case G_S: b[1ul << s[0]] *= 1.0i; break;                          // :262
    case G_T: b[1ul << s[0]] *= std::sqrt(1.0i); break;              // :297
    case G_RX: case G_RXFRAC: rx_apply(g.rx, b, s[0]); break;
    case G_RY: case G_RYFRAC: ry_apply(g.ry, b, s[0]); break;
    case G_RZ: case G_U1: case G_RZFRAC: rz_apply(g.rz, b, s[0]); break;
    case G_RI: case G_RIFRAC: ri_apply(g.ri, b, s[0]); break;
    case G_R1: case G_R1FRAC: rz_apply(g.rz, b, s[0]); ri_apply(g.ri, b, s[0]); break;
    case G_SWAP: swp(b[(1ul << s[0]) + (0ul << s[1])], b[(0ul << s[0]) + (1ul << s[1])]); break;  // :804-805
    case G_C1: case G_U2: case G_U3: c1_apply(g.c1, b, s[0]); break;
    case G_C2: {  // :991-1009, with the broadcast rows of the constructor
      const u64 o1 = 1ul << s[1], o2 = 1ul << s[0], o3 = o1 + o2;
      cd v0 = b[0], v1 = b[o1], v2 = b[o2], v3 = b[o3];
      const cd* r = g.c2row;
      b[0] = r[0] * v0 + r[0] * v1 + r[0] * v2 + r[0] * v3;
      b[o1] = r[1] * v0 + r[1] * v1 + r[1] * v2 + r[1] * v3;
      b[o2] = r[2] * v0 + r[2] * v1 + r[2] * v2 + r[2] * v3;
      b[o3] = r[3] * v0 + r[3] * v1 + r[3] * v2 + r[3] * v3;
      break;
    }
    case G_RXX: {  // :586-598 (note the second H on slots[1] of the base half, :591)
      const u64 p0 = 1ul << s[0], p1 = 1ul << s[1];
      h_apply(b, s[0]); h_apply(b + p1, s[0]);
      h_apply(b, s[1]); h_apply(b + p0, s[1]);
      h_apply(b, s[1]);
      swp(b[p0], b[p0 + p1]);
      rz_apply(g.rz, b, s[1]); rz_apply(g.rz, b + p0, s[1]);
      swp(b[p0], b[p0 + p1]);
      h_apply(b, s[1]); h_apply(b + p0, s[1]);
      h_apply(b, s[0]); h_apply(b + p1, s[0]);
      break;
    }
    case G_RYY: {  // :1109-1120
      const u64 p0 = 1ul << s[0], p1 = 1ul << s[1];
      rx_apply(g.rx2, b, s[0]); rx_apply(g.rx2, b + p1, s[0]);
      rx_apply(g.rx2, b, s[1]); rx_apply(g.rx2, b + p0, s[1]);
      swp(b[p0], b[p0 + p1]);
      rz_apply(g.rz, b, s[1]); rz_apply(g.rz, b + p0, s[1]);
      swp(b[p0], b[p0 + p1]);
      rx_apply(g.rx3, b, s[0]); rx_apply(g.rx3, b + p1, s[0]);
      rx_apply(g.rx3, b, s[1]); rx_apply(g.rx3, b + p0, s[1]);
      break;
    }
    case G_RZZ: {  // :1249-1252
      const u64 p0 = 1ul << s[0], p1 = 1ul << s[1];
      swp(b[p0], b[p0 + p1]);
      rz_apply(g.rz, b, s[1]); rz_apply(g.rz, b + p0, s[1]);
      swp(b[p0], b[p0 + p1]);
      break;
    }
    case G_TDG: {  // :1335-1338
      cd v1 = b[1ul << s[0]];
      b[1ul << s[0]] = kInvSqrt2 * (1.0 - 1.0i) * v1;
      break;
    }
    case G_SDG: b[1ul << s[0]] *= -1.0i; break;  // :1379
    default: break;
  }
}

int expected_targets(int type) {
  switch (type) {
    case G_SWAP: case G_C2: case G_RXX: case G_RYY: case G_RZZ: return 2;
    case G_I: case G_X: case G_Y: case G_Z: case G_H: case G_S: case G_T:
    case G_RX: case G_RY: case G_RZ: case G_RI: case G_R1: case G_RXFRAC:
    case G_RYFRAC: case G_RZFRAC: case G_RIFRAC: case G_R1FRAC: case G_C1:
    case G_U1: case G_U2: case G_U3: case G_TDG: case G_SDG: return 1;
    default: return -1;
  }
}

inline u64 pdep(u64 x, u64 mask) {
  u64 r = 0;
  for (u64 bit = 1; mask; bit <<= 1) {
    u64 low = mask & (~mask + 1);
    if (x & bit) r |= low;
    mask &= mask - 1;
  }
  return r;
}

// Per-gate sweep on a vector of `nslots` slots (slot == qubit at top level).
// state_vector.hpp:402-510: control_offset = sum 2^slot(c); loop over the
// slots that are neither targets nor controls.
int apply_sweep(cd* st, int nslots, const Prepared& g, int nc, const int* cs,
                int nt, const int* ts, bool parallel) {
  u64 used = 0, coff = 0;
  for (int i = 0; i < nc; ++i) {
    if (cs[i] < 0 || cs[i] >= nslots || (used >> cs[i] & 1)) return -2;
    used |= 1ull << cs[i];
    coff += 1ull << cs[i];
  }
  for (int i = 0; i < nt; ++i) {
    if (ts[i] < 0 || ts[i] >= nslots || (used >> ts[i] & 1)) return -2;
    used |= 1ull << ts[i];
  }
  const u64 all = (nslots == 64) ? ~0ull : ((1ull << nslots) - 1);
  const u64 freem = all & ~used;
  const long long iters = 1ll << (nslots - nc - nt);
  if (parallel) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < iters; ++i)
      apply_at(g, st + coff + pdep((u64)i, freem), ts);
  } else {
    for (long long i = 0; i < iters; ++i)
      apply_at(g, st + coff + pdep((u64)i, freem), ts);
  }
  return 0;
}

struct GateDesc {  // layout shared with oracle/oracle.py
  int32_t type, nc, nt;
  int32_t controls[5], targets[5];
  double params[32];
};

}  // namespace

extern "C" {

int orc_expected_targets(int type) { return expected_targets(type); }

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// Per-gate path: BoundGate::apply -> apply_gate -> StateVector::apply
// (state_vector.hpp:1349, 1202-1320, 402-510). Returns 0, or <0 on error.
int orc_apply_gate(double* state, int n, const GateDesc* gd) {
  if (gd->type <= G_FREE || gd->type >= G_COUNT) return -1;
  if (gd->nt != expected_targets(gd->type)) return -3;
  Prepared g = prepare(gd->type, gd->params);
  return apply_sweep(reinterpret_cast<cd*>(state), n, g, gd->nc, gd->controls,
                     gd->nt, gd->targets, true);
}

// Partitioned path: the live body of gather_and_execute_on_multi
// (execute.hpp:318-470) for one part, at top level where slot == qubit.
// opt_slots adds slots 0..3 to the part (execute.hpp:184-199).
int orc_run_part(double* state, int n, const GateDesc* gds, int ngates, int opt_slots) {
  cd* st = reinterpret_cast<cd*>(state);
  u64 part = 0;
  for (int i = 0; i < ngates; ++i) {
    const GateDesc& d = gds[i];
    if (d.type <= G_FREE || d.type >= G_COUNT) return -1;
    if (d.nt != expected_targets(d.type)) return -3;
    for (int j = 0; j < d.nc; ++j) { if (d.controls[j] < 0 || d.controls[j] >= n) return -2; part |= 1ull << d.controls[j]; }
    for (int j = 0; j < d.nt; ++j) { if (d.targets[j] < 0 || d.targets[j] >= n) return -2; part |= 1ull << d.targets[j]; }
  }
  if (opt_slots) for (int s = 0; s < 4 && s < n; ++s) part |= 1ull << s;
  // Inner slot of a qubit = its rank among the part's slots (qubits_ordered_by_slots).
  int inner_of[64];
  int k = 0;
  for (int s = 0; s < n; ++s) inner_of[s] = (part >> s & 1) ? k++ : -1;
  std::vector<Prepared> prep(ngates);
  std::vector<GateDesc> inner(ngates);
  for (int i = 0; i < ngates; ++i) {
    prep[i] = prepare(gds[i].type, gds[i].params);
    inner[i] = gds[i];
    for (int j = 0; j < gds[i].nc; ++j) inner[i].controls[j] = inner_of[gds[i].controls[j]];
    for (int j = 0; j < gds[i].nt; ++j) inner[i].targets[j] = inner_of[gds[i].targets[j]];
  }
  const u64 all = (n == 64) ? ~0ull : ((1ull << n) - 1);
  const u64 outer = all & ~part;
  const long long nbase = 1ll << (n - k);
  const u64 nin = 1ull << k;
  int err = 0;
#pragma omp parallel
  {
    std::vector<cd> buf(nin);
    std::vector<u64> off(nin);
    for (u64 j = 0; j < nin; ++j) off[j] = pdep(j, part);
#pragma omp for schedule(static)
    for (long long b = 0; b < nbase; ++b) {
      const u64 base = pdep((u64)b, outer);
      for (u64 j = 0; j < nin; ++j) buf[j] = st[base + off[j]];          // gather_qubits
      for (int i = 0; i < ngates; ++i) {                                   // execute_on
        int e = apply_sweep(buf.data(), k, prep[i], inner[i].nc, inner[i].controls,
                            inner[i].nt, inner[i].targets, false);
        if (e) err = e;
      }
      for (u64 j = 0; j < nin; ++j) st[base + off[j]] = buf[j];          // scatter_qubits
    }
  }
  return err;
}

// Bit permutation of the index: out[idx'] = in[idx] where bit new_slot[s] of
// idx' equals bit s of idx. The order used by the reference's qubit exchange
// (svsim-mpi.hpp:139-149) is produced by orc_exchange_order below.
int orc_permute(const double* in, double* out, int n, const int* new_slot) {
  const cd* a = reinterpret_cast<const cd*>(in);
  cd* o = reinterpret_cast<cd*>(out);
  u64 seen = 0;
  for (int s = 0; s < n; ++s) {
    if (new_slot[s] < 0 || new_slot[s] >= n || (seen >> new_slot[s] & 1)) return -2;
    seen |= 1ull << new_slot[s];
  }
  const long long N = 1ll << n;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < N; ++i) {
    u64 j = 0;
    for (int s = 0; s < n; ++s) j |= (((u64)i >> s) & 1ull) << new_slot[s];
    o[j] = a[i];
  }
  return 0;
}

// New slot order after moving `m` qubits to the top: the qubits not moved keep
// their relative order, then the moved ones in current slot order
// (svsim-mpi.hpp:139-149). slot_to_qubit/out_order have n entries.
void orc_exchange_order(int n, const int* slot_to_qubit, const int* moved, int m, int* out_order) {
  auto is_moved = [&](int q) { for (int i = 0; i < m; ++i) if (moved[i] == q) return true; return false; };
  int k = 0;
  for (int s = 0; s < n; ++s) if (!is_moved(slot_to_qubit[s])) out_order[k++] = slot_to_qubit[s];
  for (int s = 0; s < n; ++s) if (is_moved(slot_to_qubit[s])) out_order[k++] = slot_to_qubit[s];
}

double orc_norm2(const double* state, int n) {
  const cd* a = reinterpret_cast<const cd*>(state);
  const long long N = 1ll << n;
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long long i = 0; i < N; ++i) s += std::norm(a[i]);
  return s;
}

// Probability that `qubit` reads `value` (state_vector.hpp:897-923), with a
// proper reduction instead of the reference's shared accumulator.
double orc_probability(const double* state, int n, int qubit, int value) {
  const cd* a = reinterpret_cast<const cd*>(state);
  const long long N = 1ll << n;
  double p0 = 0;
#pragma omp parallel for reduction(+ : p0) schedule(static)
  for (long long i = 0; i < N; ++i)
    if (!((i >> qubit) & 1)) p0 += std::abs(a[i] * a[i]);
  return value == 0 ? p0 : 1 - p0;
}

// Sampling walk of StateVector::measure (state_vector.hpp:925-981) given the
// already sorted uniform samples. out[i*nq + j] = bit of qubits[j] for sample i.
void orc_measure_sorted(const double* state, int n, const int* qubits, int nq,
                        const double* sorted_u, int ns, int32_t* out) {
  const cd* a = reinterpret_cast<const cd*>(state);
  const u64 N = 1ull << n;
  double cum = 0;
  int sp = 0;
  for (u64 pos = 0; pos < N && sp < ns; ++pos) {
    const double p = std::abs(a[pos] * a[pos]);
    while (sp < ns && sorted_u[sp] < cum + p) {
      for (int j = 0; j < nq; ++j) out[sp * nq + j] = (int32_t)((pos >> qubits[j]) & 1);
      ++sp;
    }
    cum += p;
  }
  for (; sp < ns; ++sp)
    for (int j = 0; j < nq; ++j) out[sp * nq + j] = 1;
}

}  // extern "C"
