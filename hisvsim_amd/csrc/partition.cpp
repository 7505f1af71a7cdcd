// Host-side planning for the moving-layout part chain (hsv_apply_fused_chain):
// the method of hisvsim_amd/partition.py chain_parts, on 64-bit qubit masks,
// so that the lookahead over candidate low-slot sets can be wider at the same
// planning time (the greedy chain equals the Python method's gate for gate),
// optionally widened into a beam search; and the grouping of a wide part's
// gates into the 5-qubit fused steps of the BF16-split part kernel.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hsv.h"
#include "partition.hpp"

namespace hsv {
std::vector<std::vector<int>> step_groups(const std::vector<uint64_t>& qm, int k, int width);
}

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

// One part of the chain from state (rem, lowset): the part's gates (taken),
// the gates left (keep) and the part's tile.
void take_part(const std::vector<uint64_t>& qm, const std::vector<int>& rem, uint64_t lowset, int k, int candidates,
               std::vector<int>* taken, std::vector<int>* keep, uint64_t* tile) {
  taken->clear();
  absorb(qm, rem, grow(qm, rem, lowset, k, candidates), taken);
  *tile = lowset;
  for (int i : *taken) *tile |= qm[i];
  keep->clear();
  keep->reserve(rem.size());
  size_t ti = 0;
  for (int i : rem) {
    if (ti < taken->size() && (*taken)[ti] == i) {
      ++ti;
      continue;
    }
    keep->push_back(i);
  }
}

// The candidate low sets for the part after one with this tile: every
// `low`-subset of the first `lookahead` tile qubits ordered by the position
// of their next gate (then index), in lexicographic order, each with the
// gates the next part would absorb; sorted by that count, descending, ties
// in enumeration order.
std::vector<std::pair<int, uint64_t>> rank_lows(const std::vector<uint64_t>& qm, const std::vector<int>& rem,
                                                uint64_t tile, int k, int low, int lookahead, int candidates) {
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
  std::vector<std::pair<int, uint64_t>> out;
  std::vector<int> idx(low);
  for (int i = 0; i < low; ++i) idx[i] = i;
  while (true) {
    uint64_t cand = 0;
    for (int i : idx) cand |= 1ull << order[i];
    out.push_back({absorb(qm, rem, grow(qm, rem, cand, k, candidates), nullptr), cand});
    int i = low - 1;
    while (i >= 0 && idx[i] == m - low + i) --i;
    if (i < 0) break;
    ++idx[i];
    for (int j = i + 1; j < low; ++j) idx[j] = idx[j - 1] + 1;
  }
  std::stable_sort(out.begin(), out.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  return out;
}

struct ChainState {
  std::vector<int> rem;
  uint64_t lowset;
  std::vector<int32_t> part_of_gate;
  std::vector<uint64_t> lows;
  int steps = 0;  // fused 5-qubit steps of the parts so far (beam only)
};

// One beam state's next part, its fused steps and the ranking of the lows
// after it.
struct Expansion {
  std::vector<int> taken, keep;
  uint64_t tile = 0;
  int steps = 0;
  std::vector<std::pair<int, uint64_t>> ranked;
};

// The fused steps the wide part kernel cuts a part into (as hsv_part_steps).
int part_step_count(const std::vector<uint64_t>& qm, const std::vector<int>& taken) {
  std::vector<uint64_t> m;
  m.reserve(taken.size());
  for (int i : taken) m.push_back(qm[i]);
  const auto g = hsv::step_groups(m, 5, 16);
  return g.empty() ? (int)taken.size() : (int)g.size();
}

// Beam search over the low-set choices: every state of the beam takes its
// next part, each branches into its `branch` best-ranked low sets, and the
// `width` children with the fewest gates left survive (duplicates dropped).
// The round in which states run out of gates ends the search, with the one
// whose parts cut into the fewest fused steps (ranking the children by steps
// as well measured worse: it lets plans with one more part through). Width
// 1, branch 1 is the greedy chain (no step counts). The
// states of one round are expanded on host threads; the result does not
// depend on their number.
int chain_beam(const std::vector<uint64_t>& qm, int k, int low, int lookahead, int candidates, int width, int branch,
               ChainState* best) {
  const int num_gates = (int)qm.size();
  ChainState s0;
  s0.rem.resize(num_gates);
  for (int i = 0; i < num_gates; ++i) s0.rem[i] = i;
  s0.lowset = (1ull << low) - 1;
  s0.part_of_gate.assign(num_gates, 0);
  std::vector<ChainState> beam{s0};
  const int hw = std::max(1, (int)std::thread::hardware_concurrency());
  const bool steered = width > 1 || branch > 1;
  while (true) {
    std::vector<Expansion> ex(beam.size());
    std::atomic<int> next{0};
    auto work = [&]() {
      for (int i = next++; i < (int)beam.size(); i = next++) {
        Expansion& e = ex[i];
        take_part(qm, beam[i].rem, beam[i].lowset, k, candidates, &e.taken, &e.keep, &e.tile);
        if (!e.taken.empty() && !e.keep.empty()) e.ranked = rank_lows(qm, e.keep, e.tile, k, low, lookahead, candidates);
        if (steered && !e.taken.empty()) e.steps = part_step_count(qm, e.taken);
      }
    };
    const int nthreads = std::min<int>(hw, (int)beam.size());
    std::vector<std::thread> pool;
    try {
      for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
    } catch (...) {
      // no more threads: the calling thread takes what is left
    }
    work();
    for (std::thread& t : pool) t.join();
    std::vector<std::pair<int, ChainState>> kids;  // (gates left after the child's next part, child)
    int done = -1;  // the finished state of fewest steps, this round
    for (size_t b = 0; b < beam.size(); ++b)
      if (!ex[b].taken.empty() && ex[b].keep.empty() &&
          (done < 0 || beam[b].steps + ex[b].steps < beam[done].steps + ex[done].steps))
        done = (int)b;
    if (done >= 0) {
      const ChainState& s = beam[done];
      *best = s;
      for (int i : ex[done].taken) best->part_of_gate[i] = (int)s.lows.size();
      best->lows.push_back(s.lowset);
      best->steps = s.steps + ex[done].steps;
      best->rem.clear();
      return HSV_OK;
    }
    for (size_t b = 0; b < beam.size(); ++b) {
      const ChainState& s = beam[b];
      const std::vector<int>& taken = ex[b].taken;
      const std::vector<int>& keep = ex[b].keep;
      if (taken.empty()) continue;  // a gate wider than the part once its lows are in
      const int p = (int)s.lows.size();
      const auto& ranked = ex[b].ranked;
      for (int j = 0; j < branch && j < (int)ranked.size(); ++j) {
        if (ranked[j].first == 0) break;
        ChainState c;
        c.rem = keep;
        c.lowset = ranked[j].second;
        c.part_of_gate = s.part_of_gate;
        for (int i : taken) c.part_of_gate[i] = p;
        c.lows = s.lows;
        c.lows.push_back(s.lowset);
        c.steps = s.steps + ex[b].steps;
        kids.push_back({(int)keep.size() - ranked[j].first, std::move(c)});
      }
    }
    if (kids.empty()) return HSV_ERR_LIMIT;
    std::stable_sort(kids.begin(), kids.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    beam.clear();
    for (auto& kd : kids) {
      bool dup = false;
      for (const ChainState& b : beam) dup = dup || (b.lowset == kd.second.lowset && b.rem == kd.second.rem);
      if (!dup) beam.push_back(std::move(kd.second));
      if ((int)beam.size() >= width) break;
    }
  }
}

}  // namespace

namespace hsv {

std::vector<std::vector<int>> step_groups(const std::vector<uint64_t>& qm, int k, int width) {
  const int ng = (int)qm.size();
  std::vector<std::vector<int>> out;
  if (ng == 0) return out;
  uint64_t all = 0;
  for (uint64_t g : qm) {
    if (__builtin_popcountll(g) > k) return {};
    all |= g;
  }
  // every k-subset of the part's qubits (all of them when fewer)
  std::vector<int> qs;
  for (int b = 0; b < 64; ++b)
    if (all >> b & 1) qs.push_back(b);
  const int n = (int)qs.size(), kk = std::min(k, n);
  // C(n, kk) candidate sets: a part of up to 14 qubits has at most 3432
  double count = 1;
  for (int i = 0; i < kk; ++i) count = count * (n - i) / (i + 1);
  if (count > 65536) return {};
  std::vector<uint64_t> subsets;
  std::vector<int> idx(kk);
  for (int i = 0; i < kk; ++i) idx[i] = i;
  while (true) {
    uint64_t c = 0;
    for (int i : idx) c |= 1ull << qs[i];
    subsets.push_back(c);
    int i = kk - 1;
    while (i >= 0 && idx[i] == n - kk + i) --i;
    if (i < 0) break;
    ++idx[i];
    for (int j = i + 1; j < kk; ++j) idx[j] = idx[j - 1] + 1;
  }
  struct Grouping {
    std::vector<int> rem;
    std::vector<uint64_t> sets;  // the groups' qubit sets so far
  };
  Grouping g0;
  g0.rem.resize(ng);
  for (int i = 0; i < ng; ++i) g0.rem[i] = i;
  std::vector<Grouping> beam{g0};
  std::vector<int> taken, keep;
  while (true) {
    // (gates left, state, subset) for every extension that takes a gate
    std::vector<std::pair<int, std::pair<int, int>>> ext;
    for (int b = 0; b < (int)beam.size(); ++b)
      for (int q = 0; q < (int)subsets.size(); ++q) {
        const int got = absorb(qm, beam[b].rem, subsets[q], nullptr);
        if (got > 0) ext.push_back({(int)beam[b].rem.size() - got, {b, q}});
      }
    if (ext.empty()) return {};
    std::stable_sort(ext.begin(), ext.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<Grouping> next;
    for (const auto& e : ext) {
      const Grouping& s = beam[e.second.first];
      taken.clear();
      absorb(qm, s.rem, subsets[e.second.second], &taken);
      keep.clear();
      size_t ti = 0;
      for (int i : s.rem) {
        if (ti < taken.size() && taken[ti] == i) {
          ++ti;
          continue;
        }
        keep.push_back(i);
      }
      Grouping c;
      c.rem = keep;
      c.sets = s.sets;
      c.sets.push_back(subsets[e.second.second]);
      if (keep.empty()) {
        // replay the sets: the groups' gate lists
        std::vector<int> rem = g0.rem;
        for (uint64_t Q : c.sets) {
          std::vector<int> t;
          absorb(qm, rem, Q, &t);
          out.push_back(t);
          std::vector<int> kp;
          size_t j = 0;
          for (int i : rem) {
            if (j < t.size() && t[j] == i) {
              ++j;
              continue;
            }
            kp.push_back(i);
          }
          rem.swap(kp);
        }
        return out;
      }
      bool dup = false;
      for (const Grouping& x : next) dup = dup || x.rem == c.rem;
      if (!dup) next.push_back(std::move(c));
      if ((int)next.size() >= width) break;
    }
    beam.swap(next);
  }
}

}  // namespace hsv

extern "C" int hsv_part_steps(const uint64_t* gate_masks, int num_gates, int k, int width, int32_t* group_of_gate,
                              int* num_groups) {
  if (num_gates < 0 || (num_gates && (!gate_masks || !group_of_gate)) || !num_groups || k < 1 || k > 14 ||
      width < 1)
    return HSV_ERR_ARG;
  const std::vector<uint64_t> qm(gate_masks, gate_masks + num_gates);
  const auto groups = hsv::step_groups(qm, k, width);
  if (num_gates && groups.empty()) return HSV_ERR_LIMIT;
  for (size_t g = 0; g < groups.size(); ++g)
    for (int i : groups[g]) group_of_gate[i] = (int32_t)g;
  *num_groups = (int)groups.size();
  return HSV_OK;
}

extern "C" int hsv_chain_parts(const uint64_t* gate_masks, int num_gates, int k, int low, int lookahead,
                               int candidates, int beam_width, int beam_branch, int32_t* part_of_gate,
                               int32_t* lows_out, int* num_parts) {
  if (num_gates < 0 || (num_gates && (!gate_masks || !part_of_gate)) || !lows_out || !num_parts || low < 1 ||
      low > k || k > 63 || lookahead < low || candidates < 1 || beam_width < 1 || beam_branch < 1)
    return HSV_ERR_ARG;
  std::vector<uint64_t> qm(gate_masks, gate_masks + num_gates);
  uint64_t all = 0;
  for (uint64_t g : qm) all |= g;
  if (64 - __builtin_clzll(all | 1) < low) return HSV_ERR_ARG;  // fewer qubits than low slots
  if (num_gates == 0) {
    *num_parts = 0;
    return HSV_OK;
  }
  ChainState best;
  if (int rc = chain_beam(qm, k, low, lookahead, candidates, 1, 1, &best)) return rc;
  if (beam_width > 1 || beam_branch > 1) {
    ChainState wide;
    // the beam can end worse than the greedy chain: keep the shorter, or at
    // equal length the one of fewer fused steps
    if (chain_beam(qm, k, low, lookahead, candidates, beam_width, beam_branch, &wide) == HSV_OK) {
      std::vector<std::vector<int>> gp(best.lows.size());
      for (int i = 0; i < num_gates; ++i) gp[best.part_of_gate[i]].push_back(i);
      best.steps = 0;
      for (const auto& t : gp) best.steps += part_step_count(qm, t);
      if (wide.lows.size() < best.lows.size() || (wide.lows.size() == best.lows.size() && wide.steps < best.steps))
        best = std::move(wide);
    }
  }
  for (int i = 0; i < num_gates; ++i) part_of_gate[i] = best.part_of_gate[i];
  for (size_t p = 0; p < best.lows.size(); ++p)
    for (int b = 0, j = 0; b < 64; ++b)
      if (best.lows[p] >> b & 1) lows_out[p * low + j++] = b;
  *num_parts = (int)best.lows.size();
  return HSV_OK;
}
