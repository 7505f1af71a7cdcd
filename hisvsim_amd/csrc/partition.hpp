// Host-side grouping of a part's gates into the fused steps of the wide
// BF16-split part kernel (libhsv, host code only).
#pragma once

#include <cstdint>
#include <vector>

namespace hsv {

// Splits a part's gates (qubit masks, in circuit order) into groups of at
// most k qubits that run in sequence: each group is the blocking scan of one
// k-qubit set Q (the gates, in order, inside Q and behind no skipped gate on
// a shared qubit), so every qubit keeps its gate order. Q ranges over every
// k-subset of the part's qubits; a beam of `width` partial groupings (those
// with the fewest gates left) searches for the fewest groups. Returns the
// groups' gate indices; empty when a gate has more than k qubits or the
// part has more than 65536 k-subsets.
std::vector<std::vector<int>> step_groups(const std::vector<uint64_t>& qm, int k, int width);

}  // namespace hsv
