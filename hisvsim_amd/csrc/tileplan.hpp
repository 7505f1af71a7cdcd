// The FP64 tile kernel's per-part plan, built on the host by libhsv (hsv.hip)
// and used by the part path and by the circuit batch (batch.hip).
#pragma once

#include <vector>

#include "gates.hpp"
#include "kernels.hpp"

namespace hsv {

struct TilePlan {
  std::vector<RegSeg> segs;        // layout segments; fast ones index rounds
  std::vector<TileRound> rounds;
  std::vector<RegGate> gates;      // m2 left unset: the uploader points it at mat2
  std::vector<int> m2idx;          // per gate: its matrix in mat2, or -1
  std::vector<double> mat2;        // 32 doubles per dense two-qubit gate
};

// Ls: the part's gates; inner_of[q]: tile bit of qubit q (-1 outside the
// tile); KT tile bits, RB >= 2 register bits per thread (a gate with two
// non-diagonal targets needs both in registers). Runs of uncontrolled
// one-qubit gates on one qubit are first multiplied into one gate.
// HSV_ERR_LIMIT when a segment's layout does not hold RB register bits.
int build_tile_plan(const std::vector<Lowered>& Ls, const int* inner_of, int KT, int RB, TilePlan* T);

// The fusion build_tile_plan applies: each run of uncontrolled one-qubit
// gates on one qubit (no other gate on that qubit in between) becomes one
// gate, the later matrix times the earlier (dense unless all are diagonal);
// identities are dropped.
std::vector<Lowered> fuse_one_qubit_runs(const std::vector<Lowered>& Ls);

}  // namespace hsv
