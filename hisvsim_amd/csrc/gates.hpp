// Host-side gate lowering for libhsv.
//
// Every reference gate functor (basic_gates.hpp:99-1408) is a fixed linear map
// on the 2 or 4 amplitudes it touches. Here each one is lowered once, in FP64
// on the host, to that map (a 2x2 or 4x4 complex matrix) plus the cheapest
// kernel class that applies it: a phase on the |1> half, a diagonal, a
// permutation, or a dense matrix. The device then only ever sees matrices.
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "../../include/hsv.h"

namespace hsv {

using cd = std::complex<double>;

enum Kind : int {
  K_IDENT = 0,   // I: nothing to do
  K_PHASE1 = 1,  // diag(1, d): only amplitudes with the target bit set
  K_DIAG1 = 2,   // diag(d0, d1)
  K_XSWAP1 = 3,  // X: swap the pair
  K_DENSE1 = 4,  // dense 2x2
  K_DIAG2 = 5,   // diag over 4 amplitudes
  K_SWAP2 = 6,   // SWAP: exchange |10> and |01>
  K_DENSE2 = 7   // dense 4x4
};

struct Lowered {
  int kind = K_IDENT;
  int nt = 0, nc = 0;
  int t[2] = {0, 0};   // target qubits; 4x4 index = 2*bit(t[0]) + bit(t[1])
  int c[5] = {0, 0, 0, 0, 0};
  cd m[16];            // row-major 2x2 (m[0..3]) or 4x4; diagonals in m[0..3]
};

// Number of targets the reference functor expects (its assert), or -1.
int expected_targets(int type);

// Lowers one gate. Returns 0 or a negative hsv_status with *err set.
int lower(const hsv_gate& g, Lowered* out, std::string* err);

// The dense matrix of a lowered gate (2x2 or 4x4, row-major).
void dense_of(const Lowered& L, cd* m /* 16 */);

// Builds the 2^k x 2^k unitary (row-major, FP64) of a sequence of lowered
// gates acting on `qubits` (sorted ascending; local bit i <-> qubits[i]).
void build_fused(const std::vector<Lowered>& gates, const std::vector<int>& qubits,
                 std::vector<cd>* U);

// max |(U^H U - I)_ij|.
double unitarity_error(const std::vector<cd>& U, int dim);

}  // namespace hsv
