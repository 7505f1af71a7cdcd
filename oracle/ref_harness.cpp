// Harness over the reference simulator's own headers.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile with -I pointing at the
// reference checkout; the output goes to oracle/_ref/ only. It exposes the
// reference's per-gate path (BoundGate::apply, state_vector.hpp:1349) and the
// live part driver gather_and_execute_on_multi (execute.hpp:318-470; the
// pybind-exported gather_and_execute_on has its body commented out at
// execute.hpp:270/276/283) through a C ABI, so the oracle restatement can be
// checked against it and so the reference can be timed as a CPU baseline.

#include <fcntl.h>
#include <unistd.h>

#include <cstdio>
#include <iostream>
#include <set>
#include <vector>

#include "execute.hpp"

using namespace SvSim;

namespace {

struct GateDesc {  // same layout as oracle/hsv_oracle.cpp
  int32_t type, nc, nt;
  int32_t controls[5], targets[5];
  double params[32];
};

Type P(const GateDesc& d, int i) { return Type{d.params[2 * i], d.params[2 * i + 1]}; }

bool make_factory(const GateDesc& d, AggregateGateFactory** out) {
  const GateType t = static_cast<GateType>(d.type);
  switch (t) {
    case GateType::I: case GateType::X: case GateType::Y: case GateType::Z:
    case GateType::H: case GateType::S: case GateType::T: case GateType::SWAP:
    case GateType::TDG: case GateType::SDG:
      *out = new AggregateGateFactory(AggregateGate_0(t)); return true;
    case GateType::RI: case GateType::RX: case GateType::RY: case GateType::RZ:
    case GateType::R1: case GateType::RXX: case GateType::RZZ: case GateType::RYY:
    case GateType::U1:
      *out = new AggregateGateFactory(AggregateGate_1(t, P(d, 0))); return true;
    case GateType::U2:
      *out = new AggregateGateFactory(AggregateGate_2(t, P(d, 0), P(d, 1))); return true;
    case GateType::U3:
      *out = new AggregateGateFactory(AggregateGate_3(t, P(d, 0), P(d, 1), P(d, 2))); return true;
    case GateType::C1:
      *out = new AggregateGateFactory(AggregateGate_C1(t, P(d, 0), P(d, 1), P(d, 2), P(d, 3))); return true;
    case GateType::C2:
      *out = new AggregateGateFactory(AggregateGate_C2(
          t, P(d, 0), P(d, 1), P(d, 2), P(d, 3), P(d, 4), P(d, 5), P(d, 6), P(d, 7),
          P(d, 8), P(d, 9), P(d, 10), P(d, 11), P(d, 12), P(d, 13), P(d, 14), P(d, 15)));
      return true;
    case GateType::RXFrac: case GateType::RYFrac: case GateType::RZFrac:
    case GateType::RIFrac: case GateType::R1Frac:
      *out = new AggregateGateFactory(
          AggregateGate_Frac(t, (unsigned)d.params[0], (unsigned)d.params[2]));
      return true;
    default: return false;
  }
}

BoundGate make_bound(const GateDesc& d, AggregateGateFactory* f) {
  std::vector<Qubit> c(d.controls, d.controls + d.nc), t(d.targets, d.targets + d.nt);
  return BoundGate(*f, c, t);
}

// Number of maximal runs of consecutive set bits: loop() takes the recursive
// N > 10 branch beyond 10 runs (loop.hpp:922-960), which indexes past its
// arrays, so such loops are refused here.
int runs(uint64_t m) {
  int r = 0;
  for (int s = 0; s < 64; ++s)
    if ((m >> s & 1) && (s == 0 || !(m >> (s - 1) & 1))) ++r;
  return r;
}

struct Quiet {  // the part driver prints planning info to stdout
  int saved = -1;
  Quiet() {
    std::cout.flush(); fflush(stdout);
    int dn = open("/dev/null", O_WRONLY);
    if (dn >= 0) { saved = dup(1); dup2(dn, 1); close(dn); }
  }
  ~Quiet() {
    std::cout.flush(); fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
  }
};

}  // namespace

extern "C" {

void ref_set_threads(int n) { set_num_threads(n); set_init_num_threads(n); }

void* ref_create(int n) { return new StateVector(static_cast<Qubit>(n), 0); }

void ref_destroy(void* p) {
  StateVector* sv = static_cast<StateVector*>(p);
  free(sv->state());
  delete sv;
}

void ref_write(void* p, const double* in, int n) {
  Type* s = static_cast<StateVector*>(p)->state();
  for (uint64_t i = 0; i < (1ull << n); ++i) s[i] = Type{in[2 * i], in[2 * i + 1]};
}

void ref_read(void* p, double* out, int n) {
  const Type* s = static_cast<StateVector*>(p)->state();
  for (uint64_t i = 0; i < (1ull << n); ++i) { out[2 * i] = s[i].real(); out[2 * i + 1] = s[i].imag(); }
}

int ref_apply_gate(void* p, const GateDesc* d) {
  AggregateGateFactory* f = nullptr;
  if (!make_factory(*d, &f)) return -1;
  BoundGate bg = make_bound(*d, f);
  bg.apply(*static_cast<StateVector*>(p));
  delete f;
  return 0;
}

// StateVector::probability (state_vector.hpp:897-923).
double ref_probability(void* p, int qubit, int value) {
  return static_cast<StateVector*>(p)->probability(static_cast<Qubit>(qubit), value);
}

// StateVector::project (state_vector.hpp:829-895): the state shrinks to the
// 2^(n - count) amplitudes of the remaining qubits, written to out.
void ref_project(void* p, const int* qubits, const int* values, int count, double* out, int n) {
  StateVector* sv = static_cast<StateVector*>(p);
  std::vector<Qubit> q(qubits, qubits + count);
  std::vector<int> v(values, values + count);
  sv->project(q, v, true);
  const Type* s = sv->state();
  for (uint64_t i = 0; i < (1ull << (n - count)); ++i) { out[2 * i] = s[i].real(); out[2 * i + 1] = s[i].imag(); }
}

int ref_run_part(void* p, const GateDesc* ds, int ngates, int opt_slots, int n) {
  uint64_t part = 0;
  for (int i = 0; i < ngates; ++i) {
    if (ds[i].type == static_cast<int>(GateType::I)) return -4;  // execute_on aborts on I
    for (int j = 0; j < ds[i].nc; ++j) part |= 1ull << ds[i].controls[j];
    for (int j = 0; j < ds[i].nt; ++j) part |= 1ull << ds[i].targets[j];
  }
  if (opt_slots) {
    if (n < 4) return -5;  // slot_to_qubit(3) asserts
    part |= 0xF;
  }
  const uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1);
  if (runs(part) > 10 || runs(all & ~part) > 10) return -6;
  std::vector<AggregateGateFactory*> fs(ngates);
  std::vector<BoundGate> bgs;
  for (int i = 0; i < ngates; ++i) {
    if (!make_factory(ds[i], &fs[i])) return -1;
    bgs.push_back(make_bound(ds[i], fs[i]));
  }
  set_opt_slots(opt_slots);
  {
    Quiet q;
    gather_and_execute_on_multi(*static_cast<StateVector*>(p), bgs, {}, 0);
  }
  for (auto* f : fs) delete f;
  return 0;
}

}  // extern "C"
