// Host-side gate lowering: reference gate -> FP64 matrix + kernel class.
// Formulas follow the reference functors' constructors and scalar bodies
// (basic_gates.hpp; line numbers at each case).
#include "gates.hpp"

#include <cmath>
#include <cstring>

namespace hsv {

namespace {

constexpr double kInvSqrt2 = 0.707106781186547524401;  // types.hpp:23
constexpr double kPi = 3.14159265358979323846;         // types.hpp:19
constexpr double kPiBy2 = 1.57079632679489661923;      // types.hpp:20
const cd kI{0.0, 1.0};

cd P(const hsv_gate& g, int i) { return cd{g.params[2 * i], g.params[2 * i + 1]}; }

// Constants of the rotation functors, computed as their constructors do
// (complex-valued angle, halved in complex arithmetic).
cd half_cos(cd th) { return std::cos(th / cd{2}); }
cd half_sin(cd th) { return std::sin(th / cd{2}); }
cd neg_exp(cd th) { return std::exp(-kI * th / cd{2}); }
cd pos_exp(cd th) { return std::exp(kI * th / cd{2}); }

// *Frac angle: -pi * numerator / 2^(power-1) (basic_gates.hpp:733-780).
double frac_angle(unsigned num, unsigned power) { return -kPi * num / std::pow(2, power - 1); }

void set2(Lowered* L, cd a, cd b, cd c, cd d) { L->m[0] = a; L->m[1] = b; L->m[2] = c; L->m[3] = d; }

void diag1(Lowered* L, cd d0, cd d1) {
  if (d0 == cd{1.0, 0.0}) { L->kind = K_PHASE1; L->m[0] = d1; }
  else { L->kind = K_DIAG1; L->m[0] = d0; L->m[1] = d1; }
}

// 4-amplitude helpers for the two-qubit functors that are written as a
// sequence of one-slot operations. v[2*b0 + b1], b0 = bit(slots[0]).
struct V4 {
  cd v[4];
  // one-slot op on slot 0 (pairs (0,2),(1,3)) or slot 1 (pairs (0,1),(2,3)),
  // restricted to one half when `half` >= 0 (half = value of the other bit).
  template <class F> void on(int slot, int half, F f) {
    for (int other = 0; other < 2; ++other) {
      if (half >= 0 && other != half) continue;
      int i0 = slot == 0 ? other : 2 * other;
      int i1 = slot == 0 ? 2 + other : 2 * other + 1;
      f(v[i0], v[i1]);
    }
  }
  void swap23() { std::swap(v[2], v[3]); }
};

auto H2 = [](cd& a, cd& b) { cd v0 = a, v1 = b; a = kInvSqrt2 * (v0 + v1); b = kInvSqrt2 * (v0 - v1); };

void rxx_matrix(cd theta, cd* M) {  // basic_gates.hpp:586-598
  const cd ng = neg_exp(theta), ps = pos_exp(theta);
  auto RZ = [&](cd& a, cd& b) { a = a * ng; b = b * ps; };
  for (int col = 0; col < 4; ++col) {
    V4 s{};
    s.v[col] = 1.0;
    s.on(0, -1, H2);
    s.on(1, -1, H2);
    s.on(1, 0, H2);  // the second H on slots[1] of the base half only (:591)
    s.swap23();
    s.on(1, -1, RZ);
    s.swap23();
    s.on(1, -1, H2);
    s.on(0, -1, H2);
    for (int r = 0; r < 4; ++r) M[r * 4 + col] = s.v[r];
  }
}

void ryy_matrix(cd theta, cd* M) {  // basic_gates.hpp:1097, 1109-1120
  const cd c1 = half_cos(cd{kPiBy2}), is1 = kI * half_sin(cd{kPiBy2});
  const cd c3 = half_cos(cd{-kPiBy2}), is3 = kI * half_sin(cd{-kPiBy2});
  const cd ng = neg_exp(theta), ps = pos_exp(theta);
  auto RXa = [&](cd& a, cd& b) { cd v0 = a, v1 = b; a = c1 * v0 - is1 * v1; b = -is1 * v0 + c1 * v1; };
  auto RXb = [&](cd& a, cd& b) { cd v0 = a, v1 = b; a = c3 * v0 - is3 * v1; b = -is3 * v0 + c3 * v1; };
  auto RZ = [&](cd& a, cd& b) { a = a * ng; b = b * ps; };
  for (int col = 0; col < 4; ++col) {
    V4 s{};
    s.v[col] = 1.0;
    s.on(0, -1, RXa);
    s.on(1, -1, RXa);
    s.swap23();
    s.on(1, -1, RZ);
    s.swap23();
    s.on(0, -1, RXb);
    s.on(1, -1, RXb);
    for (int r = 0; r < 4; ++r) M[r * 4 + col] = s.v[r];
  }
}

}  // namespace

int expected_targets(int type) {
  switch (type) {
    case HSV_G_SWAP: case HSV_G_C2: case HSV_G_RXX: case HSV_G_RYY: case HSV_G_RZZ: return 2;
    case HSV_G_I: case HSV_G_X: case HSV_G_Y: case HSV_G_Z: case HSV_G_H: case HSV_G_S:
    case HSV_G_T: case HSV_G_RX: case HSV_G_RY: case HSV_G_RZ: case HSV_G_RI: case HSV_G_R1:
    case HSV_G_RXFRAC: case HSV_G_RYFRAC: case HSV_G_RZFRAC: case HSV_G_RIFRAC:
    case HSV_G_R1FRAC: case HSV_G_C1: case HSV_G_U1: case HSV_G_U2: case HSV_G_U3:
    case HSV_G_TDG: case HSV_G_SDG: return 1;
    default: return -1;
  }
}

int lower(const hsv_gate& g, Lowered* L, std::string* err) {
  *L = Lowered{};
  int nt = g.type == HSV_G_MATRIX ? g.num_targets : expected_targets(g.type);
  if (nt < 0) { *err = "unknown gate type " + std::to_string(g.type); return HSV_ERR_GATE; }
  if (g.num_targets != nt || nt < 1 || nt > 2) {
    *err = "gate type " + std::to_string(g.type) + " takes " + std::to_string(nt) +
           " target(s), got " + std::to_string(g.num_targets);
    return HSV_ERR_GATE;
  }
  if (g.num_controls < 0 || g.num_controls > 5) { *err = "at most 5 controls"; return HSV_ERR_ARG; }
  L->nt = nt;
  L->nc = g.num_controls;
  for (int i = 0; i < nt; ++i) L->t[i] = g.targets[i];
  for (int i = 0; i < g.num_controls; ++i) L->c[i] = g.controls[i];
  // qubits must be distinct (assert at state_vector.hpp:426)
  int all[7], na = 0;
  for (int i = 0; i < nt; ++i) all[na++] = L->t[i];
  for (int i = 0; i < L->nc; ++i) all[na++] = L->c[i];
  for (int i = 0; i < na; ++i) {
    if (all[i] < 0 || all[i] > 63) { *err = "qubit index out of range"; return HSV_ERR_ARG; }
    for (int j = 0; j < i; ++j)
      if (all[i] == all[j]) { *err = "duplicate qubit in gate"; return HSV_ERR_ARG; }
  }

  switch (g.type) {
    case HSV_G_I: L->kind = K_IDENT; break;                                        // :99-106
    case HSV_G_X: L->kind = K_XSWAP1; break;                                       // :115
    case HSV_G_Y: L->kind = K_DENSE1; set2(L, 0.0, -kI, kI, 0.0); break;           // :143-146
    case HSV_G_H: L->kind = K_DENSE1; set2(L, kInvSqrt2, kInvSqrt2, kInvSqrt2, -kInvSqrt2); break;  // :186-189
    case HSV_G_Z: L->kind = K_PHASE1; L->m[0] = -1.0; break;                       // :230
    case HSV_G_S: L->kind = K_PHASE1; L->m[0] = kI; break;                         // :262
    case HSV_G_T: L->kind = K_PHASE1; L->m[0] = std::sqrt(kI); break;              // :297
    case HSV_G_SDG: L->kind = K_PHASE1; L->m[0] = -kI; break;                      // :1379
    case HSV_G_TDG: L->kind = K_PHASE1; L->m[0] = kInvSqrt2 * cd{1.0, -1.0}; break;  // :1338
    case HSV_G_RX: case HSV_G_RXFRAC: {                                            // :343-363
      cd th = g.type == HSV_G_RX ? P(g, 0) : cd{frac_angle((unsigned)g.params[0], (unsigned)g.params[2])};
      cd c = half_cos(th), is = kI * half_sin(th);
      L->kind = K_DENSE1; set2(L, c, -is, -is, c);
      break;
    }
    case HSV_G_RY: case HSV_G_RYFRAC: {                                            // :401-422
      cd th = g.type == HSV_G_RY ? P(g, 0) : cd{frac_angle((unsigned)g.params[0], (unsigned)g.params[2])};
      cd c = half_cos(th), s = half_sin(th);
      L->kind = K_DENSE1; set2(L, c, -s, s, c);
      break;
    }
    case HSV_G_RZ: case HSV_G_U1: case HSV_G_RZFRAC: {                             // :515-535, U1 = RZ :571
      cd th = g.type != HSV_G_RZFRAC ? P(g, 0) : cd{frac_angle((unsigned)g.params[0], (unsigned)g.params[2])};
      diag1(L, neg_exp(th), pos_exp(th));
      break;
    }
    case HSV_G_RI: case HSV_G_RIFRAC: {                                            // :460-478
      cd th = g.type == HSV_G_RI ? P(g, 0) : cd{frac_angle((unsigned)g.params[0], (unsigned)g.params[2])};
      cd e = neg_exp(th);
      diag1(L, e, e);
      break;
    }
    case HSV_G_R1: {  // RZ(theta) then RI(-theta) (:716-728)
      cd th = P(g, 0);
      cd ri = neg_exp(-th);
      diag1(L, neg_exp(th) * ri, pos_exp(th) * ri);
      break;
    }
    case HSV_G_R1FRAC: {  // RZFrac(-numerator, power+1) then RIFrac(numerator, power+1) (:782-795)
      unsigned num = (unsigned)g.params[0], pw = (unsigned)g.params[2];
      cd a = cd{frac_angle(0u - num, pw + 1)}, b = cd{frac_angle(num, pw + 1)};
      cd ri = neg_exp(b);
      diag1(L, neg_exp(a) * ri, pos_exp(a) * ri);
      break;
    }
    case HSV_G_SWAP: L->kind = K_SWAP2; break;                                     // :804-805
    case HSV_G_C1: L->kind = K_DENSE1; set2(L, P(g, 0), P(g, 1), P(g, 2), P(g, 3)); break;  // :852-858
    case HSV_G_C2: {  // constructor keeps column 0 of each row (:972-987)
      L->kind = K_DENSE2;
      const cd row[4] = {P(g, 0), P(g, 4), P(g, 8), P(g, 12)};
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) L->m[r * 4 + c] = row[r];
      break;
    }
    case HSV_G_U2: {  // :1015-1031
      cd phi = P(g, 0), lam = P(g, 1);
      L->kind = K_DENSE1;
      set2(L, cd{kInvSqrt2, 0.0}, -kInvSqrt2 * (std::cos(lam) + kI * std::sin(lam)),
           kInvSqrt2 * (std::cos(phi) + kI * std::sin(phi)),
           kInvSqrt2 * (std::cos(phi + lam) + kI * std::sin(phi + lam)));
      break;
    }
    case HSV_G_U3: {  // e^{-i(phi+lambda)/2} times the textbook U3 (:1042-1053)
      cd th = P(g, 0), phi = P(g, 1), lam = P(g, 2);
      cd ct = std::cos(th / 2.0), st = std::sin(th / 2.0);
      L->kind = K_DENSE1;
      set2(L, ct * std::cos((-phi - lam) / 2.0) + kI * ct * std::sin((-phi - lam) / 2.0),
           -st * std::cos((-phi + lam) / 2.0) + kI * -st * std::sin((-phi + lam) / 2.0),
           st * std::cos((phi - lam) / 2.0) + kI * st * std::sin((phi - lam) / 2.0),
           ct * std::cos((phi + lam) / 2.0) + kI * ct * std::sin((phi + lam) / 2.0));
      break;
    }
    case HSV_G_RZZ: {  // CNOT . RZ(theta)@slot1 . CNOT = diag(e-, e+, e+, e-) (:1249-1252)
      cd ng = neg_exp(P(g, 0)), ps = pos_exp(P(g, 0));
      L->kind = K_DIAG2; L->m[0] = ng; L->m[1] = ps; L->m[2] = ps; L->m[3] = ng;
      break;
    }
    case HSV_G_RXX: L->kind = K_DENSE2; rxx_matrix(P(g, 0), L->m); break;
    case HSV_G_RYY: L->kind = K_DENSE2; ryy_matrix(P(g, 0), L->m); break;
    case HSV_G_MATRIX: {
      const int d = 1 << nt;
      L->kind = nt == 1 ? K_DENSE1 : K_DENSE2;
      for (int i = 0; i < d * d; ++i) L->m[i] = P(g, i);
      break;
    }
    default: *err = "unsupported gate type " + std::to_string(g.type); return HSV_ERR_GATE;
  }
  return HSV_OK;
}

void dense_of(const Lowered& L, cd* m) {
  for (int i = 0; i < 16; ++i) m[i] = 0.0;
  switch (L.kind) {
    case K_IDENT: m[0] = m[3] = 1.0; break;
    case K_PHASE1: m[0] = 1.0; m[3] = L.m[0]; break;
    case K_DIAG1: m[0] = L.m[0]; m[3] = L.m[1]; break;
    case K_XSWAP1: m[1] = m[2] = 1.0; break;
    case K_DENSE1: for (int i = 0; i < 4; ++i) m[i] = L.m[i]; break;
    case K_DIAG2: for (int i = 0; i < 4; ++i) m[i * 5] = L.m[i]; break;
    case K_SWAP2: m[0] = m[15] = 1.0; m[1 * 4 + 2] = m[2 * 4 + 1] = 1.0; break;
    case K_DENSE2: for (int i = 0; i < 16; ++i) m[i] = L.m[i]; break;
  }
}

void build_fused(const std::vector<Lowered>& gates, const std::vector<int>& qubits,
                 std::vector<cd>* Uout) {
  const int k = (int)qubits.size();
  const size_t dim = size_t(1) << k;
  auto local = [&](int q) {
    for (int i = 0; i < k; ++i) if (qubits[i] == q) return i;
    return -1;
  };
  // Columns of U are the images of the basis states; propagate them all.
  std::vector<cd> S(dim * dim, 0.0);  // S[col * dim + amp]
  for (size_t c = 0; c < dim; ++c) S[c * dim + c] = 1.0;
  for (const Lowered& L : gates) {
    if (L.kind == K_IDENT) continue;
    cd m[16];
    dense_of(L, m);
    size_t cmask = 0;
    for (int i = 0; i < L.nc; ++i) cmask |= size_t(1) << local(L.c[i]);
    const size_t b0 = size_t(1) << local(L.t[0]);
    const size_t b1 = L.nt == 2 ? size_t(1) << local(L.t[1]) : 0;
    for (size_t c = 0; c < dim; ++c) {
      cd* v = &S[c * dim];
      for (size_t i = 0; i < dim; ++i) {
        if ((i & cmask) != cmask) continue;
        if (L.nt == 1) {
          if (i & b0) continue;
          cd x0 = v[i], x1 = v[i | b0];
          v[i] = m[0] * x0 + m[1] * x1;
          v[i | b0] = m[2] * x0 + m[3] * x1;
        } else {
          if ((i & b0) || (i & b1)) continue;
          cd x[4] = {v[i], v[i | b1], v[i | b0], v[i | b0 | b1]};
          cd y[4];
          for (int r = 0; r < 4; ++r)
            y[r] = m[r * 4] * x[0] + m[r * 4 + 1] * x[1] + m[r * 4 + 2] * x[2] + m[r * 4 + 3] * x[3];
          v[i] = y[0]; v[i | b1] = y[1]; v[i | b0] = y[2]; v[i | b0 | b1] = y[3];
        }
      }
    }
  }
  Uout->assign(dim * dim, 0.0);
  for (size_t c = 0; c < dim; ++c)
    for (size_t r = 0; r < dim; ++r) (*Uout)[r * dim + c] = S[c * dim + r];
}

double unitarity_error(const std::vector<cd>& U, int dim) {
  double worst = 0;
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) {
      cd s = 0.0;
      for (int r = 0; r < dim; ++r) s += std::conj(U[r * dim + i]) * U[r * dim + j];
      if (i == j) s -= 1.0;
      worst = std::max(worst, std::abs(s));
    }
  return worst;
}

}  // namespace hsv
