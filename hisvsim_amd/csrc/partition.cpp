// Host-side planning for the moving-layout part chain (hsv_apply_fused_chain):
// the method of hisvsim_amd/partition.py chain_parts, on 64-bit qubit masks,
// so that the lookahead over candidate low-slot sets can be wider at the same
// planning time. Results equal the Python method's gate for gate.
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hsv.h"

namespace {

// The blocking scan for a fixed qubit set Q: the remaining gates, in order,
// that act inside Q and follow no skipped gate on a shared qubit.
int absorb(const std::vector<uint64_t>& qm, const std::vector<int>& rem, uint64_t Q, std::vector<int>* taken) {
  uint64_t blocked = 0;
  int count = 0;
  for (int i : rem) {
    const uint64_t g = qm[i];
    if ((g & blocked) || (g & ~Q)) {
      blocked |= g;
      if ((Q & ~blocked) == 0) break;
      continue;
    }
    ++count;
    if (taken) taken->push_back(i);
  }
  return count;
}

// Grows a qubit set from `base` up to k qubits: each step adds the candidate
// (the new qubits of one of the first `candidates` remaining gates outside Q
// and not behind a skipped gate) with the most absorbed gates per added qubit.
uint64_t grow(const std::vector<uint64_t>& qm, const std::vector<int>& rem, uint64_t base, int k, int candidates) {
  uint64_t Q = base;
  while (__builtin_popcountll(Q) < k) {
    uint64_t blocked = 0;
    std::vector<uint64_t> cands;
    for (int i : rem) {
      const uint64_t g = qm[i];
      if ((g & ~Q) == 0) continue;
      const uint64_t add = g & ~Q;
      if (!(g & blocked) && __builtin_popcountll(Q) + __builtin_popcountll(add) <= k &&
          std::find(cands.begin(), cands.end(), add) == cands.end())
        cands.push_back(add);
      blocked |= g;
      if ((int)cands.size() >= candidates) break;
    }
    if (cands.empty()) break;
    uint64_t best = 0;
    double best_score = -1.0;
    for (uint64_t add : cands) {
      const double score = double(absorb(qm, rem, Q | add, nullptr)) / __builtin_popcountll(add);
      if (score > best_score) {
        best_score = score;
        best = add;
      }
    }
    Q |= best;
  }
  return Q;
}

}  // namespace

extern "C" int hsv_chain_parts(const uint64_t* gate_masks, int num_gates, int k, int low, int lookahead,
                               int candidates, int32_t* part_of_gate, int32_t* lows_out, int* num_parts) {
  if (num_gates < 0 || (num_gates && (!gate_masks || !part_of_gate)) || !lows_out || !num_parts || low < 1 ||
      low > k || k > 63 || lookahead < low || candidates < 1)
    return HSV_ERR_ARG;
  std::vector<uint64_t> qm(gate_masks, gate_masks + num_gates);
  uint64_t all = 0;
  for (uint64_t g : qm) all |= g;
  if (64 - __builtin_clzll(all | 1) < low) return HSV_ERR_ARG;  // fewer qubits than low slots
  std::vector<int> rem(num_gates);
  for (int i = 0; i < num_gates; ++i) rem[i] = i;
  uint64_t lowset = (1ull << low) - 1;
  int p = 0;
  while (!rem.empty()) {
    std::vector<int> taken;
    absorb(qm, rem, grow(qm, rem, lowset, k, candidates), &taken);
    if (taken.empty()) return HSV_ERR_LIMIT;  // a gate wider than the part once its lows are in
    uint64_t tile = lowset;
    for (int i : taken) {
      part_of_gate[i] = p;
      tile |= qm[i];
    }
    for (int b = 0, j = 0; b < 64; ++b)
      if (lowset >> b & 1) lows_out[p * low + j++] = b;
    ++p;
    std::vector<int> keep;
    keep.reserve(rem.size());
    size_t ti = 0;
    for (int i : rem) {
      if (ti < taken.size() && taken[ti] == i) {
        ++ti;
        continue;
      }
      keep.push_back(i);
    }
    rem.swap(keep);
    if (rem.empty()) break;
    // the tile's qubits ordered by the position of their next gate (then index)
    std::vector<int> first(64, (int)rem.size());
    for (int pos = (int)rem.size() - 1; pos >= 0; --pos) {
      uint64_t g = qm[rem[pos]] & tile;
      while (g) {
        first[__builtin_ctzll(g)] = pos;
        g &= g - 1;
      }
    }
    std::vector<int> order;
    for (int b = 0; b < 64; ++b)
      if (tile >> b & 1) order.push_back(b);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return first[a] < first[b]; });
    const int m = std::min<int>(lookahead, (int)order.size());
    // combinations of `low` of the first m, in lexicographic order; the first
    // with the most gates absorbed by the next part wins
    std::vector<int> idx(low);
    for (int i = 0; i < low; ++i) idx[i] = i;
    int best_got = -1;
    uint64_t best = 0;
    while (true) {
      uint64_t cand = 0;
      for (int i : idx) cand |= 1ull << order[i];
      const int got = absorb(qm, rem, grow(qm, rem, cand, k, candidates), nullptr);
      if (got > best_got) {
        best_got = got;
        best = cand;
      }
      int i = low - 1;
      while (i >= 0 && idx[i] == m - low + i) --i;
      if (i < 0) break;
      ++idx[i];
      for (int j = i + 1; j < low; ++j) idx[j] = idx[j - 1] + 1;
    }
    lowset = best;
  }
  *num_parts = p;
  return HSV_OK;
}
