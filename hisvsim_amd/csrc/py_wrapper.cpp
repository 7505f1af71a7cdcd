// pybind11 module `svsimulator_py_wrapper`: the reference's Python API
// (py_wrapper.cpp:15-177 of HiSVSIM) over the libhsv C ABI.
//
// Same class and function names, constructor signatures and argument
// meanings as the reference, so a driver written against the reference
// module runs unchanged. Differences, all on paths where the reference
// aborts the interpreter or computes nothing:
//  * invalid gate names / wrong argument counts raise ValueError instead of
//    assert(0) (basic_gates.hpp:1521, types.hpp:34);
//  * gather_and_execute_on really gathers, executes and scatters (the
//    reference's exported body is commented out, execute.hpp:270/276/283);
//  * the *Frac names are matched case-insensitively (the reference compares
//    an upper-cased name with "RXFrac", so they never match, :1505-1509);
//  * an I gate inside a part is a no-op (execute_on aborts on it, :56, :69).
// Extensions are keyword-only or new methods: dtype, amplitudes(),
// set_amplitudes(), norm2(), probability(), execute_fused(), ...
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <bitset>
#include <cctype>
#include <complex>
#include <cstdio>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hsv.h"

namespace py = pybind11;
using cd = std::complex<double>;

namespace {

[[noreturn]] void raise_status(int rc, const char* what) {
  std::string msg = std::string(what) + ": " + hsv_last_error();
  if (rc == HSV_ERR_ALLOC) throw std::bad_alloc();
  if (rc == HSV_ERR_ARG || rc == HSV_ERR_GATE || rc == HSV_ERR_LIMIT || rc == HSV_ERR_UNITARY)
    throw py::value_error(msg);
  throw std::runtime_error(msg);
}

void check(int rc, const char* what) {
  if (rc != HSV_OK) raise_status(rc, what);
}

std::string upper(std::string s) {
  for (auto& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// Name -> GateType of basic_gates.hpp:1487-1522 (numbering in hsv.h).
int gate_type(const std::string& name) {
  static const std::pair<const char*, int> table[] = {
      {"I", HSV_G_I}, {"X", HSV_G_X}, {"Y", HSV_G_Y}, {"Z", HSV_G_Z}, {"H", HSV_G_H}, {"S", HSV_G_S},
      {"T", HSV_G_T}, {"RX", HSV_G_RX}, {"RY", HSV_G_RY}, {"RZ", HSV_G_RZ}, {"RI", HSV_G_RI},
      {"R1", HSV_G_R1}, {"RXFRAC", HSV_G_RXFRAC}, {"RYFRAC", HSV_G_RYFRAC}, {"RZFRAC", HSV_G_RZFRAC},
      {"RIFRAC", HSV_G_RIFRAC}, {"R1FRAC", HSV_G_R1FRAC}, {"SWAP", HSV_G_SWAP}, {"C1", HSV_G_C1},
      {"C2", HSV_G_C2}, {"U1", HSV_G_U1}, {"U2", HSV_G_U2}, {"U3", HSV_G_U3}, {"RXX", HSV_G_RXX},
      {"RYY", HSV_G_RYY}, {"RZZ", HSV_G_RZZ}, {"TDG", HSV_G_TDG}, {"SDG", HSV_G_SDG}};
  const std::string u = upper(name);
  for (const auto& e : table)
    if (u == e.first) return e.second;
  return 0;
}

// The gate kinds each aggregate accepts (constructors at basic_gates.hpp:1527-1776).
enum AG { AG0, AG1, AG2, AG3, AGC1, AGC2, AGFRAC };
bool accepts(AG ag, int t) {
  switch (ag) {
    case AG0: return t == HSV_G_I || t == HSV_G_X || t == HSV_G_Y || t == HSV_G_Z || t == HSV_G_H || t == HSV_G_S ||
                     t == HSV_G_T || t == HSV_G_SWAP || t == HSV_G_TDG || t == HSV_G_SDG;
    case AG1: return t == HSV_G_RI || t == HSV_G_RX || t == HSV_G_RY || t == HSV_G_RZ || t == HSV_G_R1 ||
                     t == HSV_G_RXX || t == HSV_G_RZZ || t == HSV_G_RYY || t == HSV_G_U1;
    case AG2: return t == HSV_G_U2;
    case AG3: return t == HSV_G_U3;
    case AGC1: return t == HSV_G_C1;
    case AGC2: return t == HSV_G_C2;
    case AGFRAC: return t == HSV_G_RXFRAC || t == HSV_G_RYFRAC || t == HSV_G_RZFRAC || t == HSV_G_RIFRAC ||
                        t == HSV_G_R1FRAC;
  }
  return false;
}

struct Aggregate {
  int type = 0;
  double params[32] = {0};
  Aggregate(AG ag, const std::string& name, const std::vector<cd>& ps) {
    type = gate_type(name);
    if (!type) throw py::value_error("unknown gate name '" + name + "'");
    if (!accepts(ag, type)) throw py::value_error("gate '" + name + "' does not take this parameter count");
    for (size_t i = 0; i < ps.size() && i < 16; ++i) {
      params[2 * i] = ps[i].real();
      params[2 * i + 1] = ps[i].imag();
    }
  }
  Aggregate(const std::string& name, unsigned numerator, unsigned power) {
    type = gate_type(name);
    if (!type || !accepts(AGFRAC, type)) throw py::value_error("gate '" + name + "' is not a *Frac gate");
    params[0] = numerator;
    params[2] = power;
  }
};

struct AggregateGate_0 : Aggregate { AggregateGate_0(const std::string& n) : Aggregate(AG0, n, {}) {} };
struct AggregateGate_1 : Aggregate { AggregateGate_1(const std::string& n, cd a) : Aggregate(AG1, n, {a}) {} };
struct AggregateGate_2 : Aggregate {
  AggregateGate_2(const std::string& n, cd a, cd b) : Aggregate(AG2, n, {a, b}) {}
};
struct AggregateGate_3 : Aggregate {
  AggregateGate_3(const std::string& n, cd a, cd b, cd c) : Aggregate(AG3, n, {a, b, c}) {}
};
struct AggregateGate_C1 : Aggregate {
  AggregateGate_C1(const std::string& n, cd a, cd b, cd c, cd d) : Aggregate(AGC1, n, {a, b, c, d}) {}
};
struct AggregateGate_C2 : Aggregate {
  AggregateGate_C2(const std::string& n, cd c00, cd c01, cd c02, cd c03, cd c10, cd c11, cd c12, cd c13, cd c20,
                   cd c21, cd c22, cd c23, cd c30, cd c31, cd c32, cd c33)
      : Aggregate(AGC2, n, {c00, c01, c02, c03, c10, c11, c12, c13, c20, c21, c22, c23, c30, c31, c32, c33}) {}
};
struct AggregateGate_Frac : Aggregate {
  AggregateGate_Frac(const std::string& n, unsigned num, unsigned power) : Aggregate(n, num, power) {}
};

struct AggregateGateFactory {
  int type = 0;
  double params[32] = {0};
  explicit AggregateGateFactory(const Aggregate& a) : type(a.type) { std::copy(a.params, a.params + 32, params); }
};

struct StateVector;

struct BoundGate {
  hsv_gate g{};
  BoundGate(const AggregateGateFactory& f, const std::vector<int>& controls, const std::vector<int>& targets) {
    if (controls.size() > 5 || targets.size() > 5) throw py::value_error("at most 5 controls and 5 targets");
    g.type = f.type;
    std::copy(f.params, f.params + 32, g.params);
    g.num_controls = (int)controls.size();
    g.num_targets = (int)targets.size();
    std::copy(controls.begin(), controls.end(), g.controls);
    std::copy(targets.begin(), targets.end(), g.targets);
  }
  void apply(StateVector& sv);
  std::vector<int> controls() const { return std::vector<int>(g.controls, g.controls + g.num_controls); }
  std::vector<int> targets() const { return std::vector<int>(g.targets, g.targets + g.num_targets); }
  int type() const { return g.type; }
  std::vector<cd> params() const {
    std::vector<cd> p;
    for (int i = 0; i < 16; ++i) p.emplace_back(g.params[2 * i], g.params[2 * i + 1]);
    return p;
  }
};

struct StateVector {
  hsv_state* h = nullptr;
  int n = 0;
  StateVector(int num_qubits, int alloc, const std::string& dtype) : n(num_qubits) {
    (void)alloc;  // NUMA placement (state_vector.hpp:189-245) has no device meaning
    int dt = dtype == "complex128" ? HSV_C128 : dtype == "complex64" ? HSV_C64 : -1;
    if (dt < 0) throw py::value_error("dtype must be 'complex128' or 'complex64'");
    check(hsv_create(num_qubits, dt, &h), "StateVector");
  }
  ~StateVector() { hsv_destroy(h); }
  StateVector(const StateVector&) = delete;
  StateVector& operator=(const StateVector&) = delete;

  std::vector<int> qubits() const {
    std::vector<int> q(n);
    for (int i = 0; i < n; ++i) q[i] = i;
    return q;
  }
  py::array_t<cd> amplitudes(uint64_t offset, int64_t count) const {
    uint64_t len = 1ull << n;
    uint64_t c = count < 0 ? len - std::min(offset, len) : (uint64_t)count;
    py::array_t<cd> out(c);
    {
      py::gil_scoped_release nogil;
      check(hsv_read(h, offset, c, reinterpret_cast<double*>(out.mutable_data())), "amplitudes");
    }
    return out;
  }
  void set_amplitudes(py::array_t<cd, py::array::c_style | py::array::forcecast> a, uint64_t offset) {
    py::gil_scoped_release nogil;
    check(hsv_write(h, offset, (uint64_t)a.size(), reinterpret_cast<const double*>(a.data())), "set_amplitudes");
  }
  double norm2() const {
    double v = 0;
    py::gil_scoped_release nogil;
    check(hsv_norm2(h, &v), "norm2");
    return v;
  }
  double probability(int qubit, int value) const {
    double v = 0;
    py::gil_scoped_release nogil;
    check(hsv_probability(h, qubit, value, &v), "probability");
    return v;
  }
  double max_abs_diff(const StateVector& o) const {
    double v = 0;
    py::gil_scoped_release nogil;
    check(hsv_max_abs_diff(h, o.h, &v), "max_abs_diff");
    return v;
  }
  cd inner(const StateVector& o) const {
    double v[2] = {0, 0};
    py::gil_scoped_release nogil;
    check(hsv_inner(h, o.h, v), "inner");
    return cd{v[0], v[1]};
  }
  // Unit norm again (FP64); returns the norm^2 before.
  double renormalize() {
    double v = 0;
    py::gil_scoped_release nogil;
    check(hsv_renormalize(h, &v), "renormalize");
    return v;
  }
  // state_vector.hpp:829-895 (values default to 0), renormalised by
  // 1/sqrt(p); the qubits stay in the state. Returns p.
  double project(const std::vector<int>& qubits, std::vector<int> values) {
    values.resize(qubits.size(), 0);
    uint64_t mask = 0, val = 0;
    for (size_t i = 0; i < qubits.size(); ++i) {
      if (qubits[i] < 0 || qubits[i] >= n || (mask >> qubits[i] & 1) || (values[i] != 0 && values[i] != 1))
        throw py::value_error("project: qubits must be unique state qubits and values 0 or 1");
      mask |= 1ull << qubits[i];
      val |= uint64_t(values[i]) << qubits[i];
    }
    double p = 0;
    py::gil_scoped_release nogil;
    check(hsv_project(h, mask, val, &p), "project");
    return p;
  }
  // <psi|P|psi>; X on x_mask only, Z on z_mask only, Y on both.
  double expectation(uint64_t x_mask, uint64_t z_mask) const {
    double v[2] = {0, 0};
    py::gil_scoped_release nogil;
    check(hsv_expect_pauli(h, x_mask, z_mask, v), "expectation");
    return v[0];
  }
  // state_vector.hpp:1003-1010: every amplitude within 1e-10 (types.hpp:10).
  bool state_equal(const StateVector& o) const {
    if (o.n != n) return false;
    return max_abs_diff(o) <= 1e-10;
  }
  // state_vector.hpp:925-981. seed < 0 draws from std::random_device as the
  // reference does; otherwise mt19937(seed) for reproducible samples.
  std::vector<std::vector<int>> measure(const std::vector<int>& qubits, unsigned num_samples, long long seed) {
    if (num_samples == 0) return {};
    std::mt19937 gen(seed < 0 ? std::random_device{}() : (unsigned)seed);
    std::uniform_real_distribution<> dis(0.0, 1.0);
    std::vector<double> u(num_samples);
    for (auto& x : u) x = dis(gen);
    std::vector<int32_t> out((size_t)num_samples * qubits.size());
    {
      py::gil_scoped_release nogil;
      check(hsv_measure(h, qubits.data(), (int)qubits.size(), u.data(), (int)num_samples, out.data()), "measure");
    }
    std::vector<std::vector<int>> ret(num_samples);
    for (unsigned i = 0; i < num_samples; ++i)
      ret[i].assign(out.begin() + (size_t)i * qubits.size(), out.begin() + (size_t)(i + 1) * qubits.size());
    return ret;
  }
  // state_vector.hpp:134-140: "<16-bit index>: (re,im)" per amplitude.
  void dump_state() const {
    auto a = amplitudes(0, -1);
    auto r = a.unchecked<1>();
    std::ostringstream os;
    for (py::ssize_t i = 0; i < r.shape(0); ++i) os << std::bitset<16>((unsigned long)i) << ": " << r(i) << "\n";
    py::print(os.str(), py::arg("end") = "");
  }
  void synchronize() {
    py::gil_scoped_release nogil;
    check(hsv_synchronize(h), "synchronize");
  }
};

void BoundGate::apply(StateVector& sv) {
  py::gil_scoped_release nogil;
  check(hsv_apply_gate(sv.h, &g), "BoundGate.apply");
}

std::vector<hsv_gate> collect(const std::vector<BoundGate*>& gates) {
  std::vector<hsv_gate> v;
  v.reserve(gates.size());
  for (auto* b : gates) v.push_back(b->g);
  return v;
}

int g_num_threads = 1, g_init_threads = 1;

}  // namespace

PYBIND11_MODULE(svsimulator_py_wrapper, m) {
  m.doc() = "MI355X hierarchical state-vector simulator (HiSVSIM svsimulator_py_wrapper API over libhsv)";

  py::class_<AggregateGate_0>(m, "AggregateGate_0").def(py::init<const std::string&>());
  py::class_<AggregateGate_1>(m, "AggregateGate_1").def(py::init<const std::string&, cd>());
  py::class_<AggregateGate_2>(m, "AggregateGate_2").def(py::init<const std::string&, cd, cd>());
  py::class_<AggregateGate_3>(m, "AggregateGate_3").def(py::init<const std::string&, cd, cd, cd>());
  py::class_<AggregateGate_C1>(m, "AggregateGate_C1").def(py::init<const std::string&, cd, cd, cd, cd>());
  py::class_<AggregateGate_C2>(m, "AggregateGate_C2")
      .def(py::init<const std::string&, cd, cd, cd, cd, cd, cd, cd, cd, cd, cd, cd, cd, cd, cd, cd, cd>());
  py::class_<AggregateGate_Frac>(m, "AggregateGate_Frac").def(py::init<const std::string&, unsigned, unsigned>());

  // Overload order of py_wrapper.cpp:132-140: 0, 1, 2, 3, Frac, C1, C2.
  py::class_<AggregateGateFactory>(m, "AggregateGateFactory")
      .def(py::init([](const AggregateGate_0& a) { return AggregateGateFactory(a); }))
      .def(py::init([](const AggregateGate_1& a) { return AggregateGateFactory(a); }))
      .def(py::init([](const AggregateGate_2& a) { return AggregateGateFactory(a); }))
      .def(py::init([](const AggregateGate_3& a) { return AggregateGateFactory(a); }))
      .def(py::init([](const AggregateGate_Frac& a) { return AggregateGateFactory(a); }))
      .def(py::init([](const AggregateGate_C1& a) { return AggregateGateFactory(a); }))
      .def(py::init([](const AggregateGate_C2& a) { return AggregateGateFactory(a); }));

  py::class_<StateVector>(m, "StateVector")
      .def(py::init<int, int, const std::string&>(), py::arg("num_qubits"), py::arg("alloc") = 0, py::kw_only(),
           py::arg("dtype") = "complex128")
      .def("qubits", &StateVector::qubits)
      .def("measure", &StateVector::measure, py::arg("qubits"), py::arg("num_samples") = 1, py::kw_only(),
           py::arg("seed") = -1)
      .def("state_equal", &StateVector::state_equal)
      .def("dump_state", &StateVector::dump_state)
      // operator== (state_vector.hpp:1053-1072): same size, every amplitude within 1e-10.
      .def("__eq__", [](const StateVector& a, const StateVector& b) { return a.state_equal(b); })
      .def("__ne__", [](const StateVector& a, const StateVector& b) { return !a.state_equal(b); })
      .def_property_readonly("num_qubits", [](const StateVector& s) { return s.n; })
      .def("amplitudes", &StateVector::amplitudes, py::arg("offset") = 0, py::arg("count") = -1)
      .def("set_amplitudes", &StateVector::set_amplitudes, py::arg("amps"), py::arg("offset") = 0)
      .def("norm2", &StateVector::norm2)
      .def("probability", &StateVector::probability, py::arg("qubit"), py::arg("value") = 0)
      .def("project", &StateVector::project, py::arg("qubits"), py::arg("values") = std::vector<int>{})
      .def("renormalize", &StateVector::renormalize)
      .def("expectation", &StateVector::expectation, py::arg("x_mask"), py::arg("z_mask"))
      .def("max_abs_diff", &StateVector::max_abs_diff)
      .def("inner", &StateVector::inner)
      .def("reset", [](StateVector& s) { check(hsv_reset(s.h), "reset"); })
      .def("synchronize", &StateVector::synchronize)
      .def("permute_qubits",
           [](StateVector& s, const std::vector<int>& new_slot) {
             if ((int)new_slot.size() != s.n) throw py::value_error("new_slot needs one entry per qubit");
             py::gil_scoped_release nogil;
             check(hsv_permute_qubits(s.h, new_slot.data()), "permute_qubits");
           })
      .def("set_timing", [](StateVector& s, bool on) { check(hsv_set_timing(s.h, on), "set_timing"); })
      .def("kernel_times", [](StateVector& s) {
        int cnt = 0;
        std::vector<double> ms(1 << 16);
        check(hsv_kernel_times(s.h, ms.data(), (int)ms.size(), &cnt), "kernel_times");
        ms.resize(std::min<int>(cnt, (int)ms.size()));
        return ms;
      });
  m.attr("StateVector").attr("__hash__") = py::none();

  py::class_<BoundGate>(m, "BoundGate")
      .def(py::init<const AggregateGateFactory&, const std::vector<int>&, const std::vector<int>&>())
      .def("apply", &BoundGate::apply)
      .def("controls", &BoundGate::controls)
      .def("targets", &BoundGate::targets)
      .def("gate_type", &BoundGate::type)
      .def("params", &BoundGate::params);

  // py_wrapper.cpp:156 -> execute.hpp:171-316 (live body 318-470).
  m.def("gather_and_execute_on", [](StateVector& sv, const std::vector<BoundGate*>& gates) {
    auto v = collect(gates);
    py::gil_scoped_release nogil;
    check(hsv_apply_part(sv.h, v.data(), (int)v.size(), hsv_get_opt_slots()), "gather_and_execute_on");
  });
  m.def(
      "execute_fused",
      [](StateVector& sv, const std::vector<BoundGate*>& gates, const std::string& precision) {
        int mode = precision == "f64" ? HSV_FUSED_F64
                   : precision == "f32split" ? HSV_FUSED_F32SPLIT
                   : precision == "bf16split" ? HSV_FUSED_BF16SPLIT
                   : precision == "bf16x3" ? HSV_FUSED_BF16X3 : -1;
        if (mode < 0) throw py::value_error("precision must be 'f64', 'f32split', 'bf16split' or 'bf16x3'");
        auto v = collect(gates);
        py::gil_scoped_release nogil;
        check(hsv_apply_fused(sv.h, v.data(), (int)v.size(), mode), "execute_fused");
      },
      py::arg("sv"), py::arg("gates"), py::arg("precision") = "f64");

  // Not in the reference: the bench's default path from this module. The
  // planner (hsv_chain_parts) cuts a gate list into parts of at most k
  // qubits with a moving set of 5 low qubits; execute_fused_chain runs them
  // on a BF16-split path (hsv_apply_fused_chain; precision "bf16split" or
  // the three-piece "bf16x3" with parts of at most 12 qubits) and
  // renormalises.
  m.def(
      "plan_chain",
      [](const std::vector<BoundGate*>& gates, int k, int beam_width, int beam_branch, int lookahead) {
        std::vector<uint64_t> masks(gates.size());
        for (size_t i = 0; i < gates.size(); ++i) {
          const hsv_gate& g = gates[i]->g;
          uint64_t m = 0;
          if (g.type != HSV_G_I) {   // the chain leaves identities out of a part's tile
            for (int j = 0; j < g.num_controls; ++j) m |= 1ull << g.controls[j];
            for (int j = 0; j < g.num_targets; ++j) m |= 1ull << g.targets[j];
          }
          masks[i] = m;
        }
        std::vector<int32_t> part(std::max<size_t>(1, gates.size()));
        std::vector<int32_t> lows(5 * (gates.size() + 1));
        int np = 0;
        {
          py::gil_scoped_release nogil;
          check(hsv_chain_parts(masks.data(), (int)gates.size(), k, 5, lookahead, 24, beam_width, beam_branch,
                                part.data(), lows.data(), &np),
                "plan_chain");
        }
        part.resize(gates.size());
        std::vector<std::vector<int>> lv(np);
        for (int p = 0; p < np; ++p) lv[p].assign(lows.begin() + 5 * p, lows.begin() + 5 * p + 5);
        return py::make_tuple(part, lv);
      },
      py::arg("gates"), py::arg("k") = 13, py::arg("beam_width") = 16, py::arg("beam_branch") = 8,
      py::arg("lookahead") = 13);
  m.def(
      "execute_fused_chain",
      [](StateVector& sv, const std::vector<std::vector<BoundGate*>>& parts, const std::vector<std::vector<int>>& lows,
         bool renormalize, const std::string& precision) {
        if (lows.size() != parts.size()) throw py::value_error("one low set per part");
        const int pmode = precision == "bf16split" ? HSV_FUSED_BF16SPLIT
                          : precision == "bf16x3" ? HSV_FUSED_BF16X3 : -1;
        if (pmode < 0) throw py::value_error("precision must be 'bf16split' or 'bf16x3'");
        std::vector<hsv_gate> v;
        std::vector<int> sizes, lw;
        for (size_t p = 0; p < parts.size(); ++p) {
          if (lows[p].size() != 5) throw py::value_error("each low set names 5 qubits");
          for (auto* b : parts[p]) v.push_back(b->g);
          sizes.push_back((int)parts[p].size());
          lw.insert(lw.end(), lows[p].begin(), lows[p].end());
        }
        const int mode = pmode | (renormalize ? HSV_CHAIN_RENORMALIZE : 0);
        py::gil_scoped_release nogil;
        check(hsv_apply_fused_chain(sv.h, v.data(), (int)v.size(), sizes.data(), (int)sizes.size(), lw.data(), mode),
              "execute_fused_chain");
      },
      py::arg("sv"), py::arg("parts"), py::arg("lows"), py::arg("renormalize") = true,
      py::arg("precision") = "bf16split");

  m.def("obtain_apply_time", []() { long long a; hsv_get_counters(&a, nullptr, nullptr, nullptr); return a; });
  m.def("obtain_gate_counter", []() { long long g; hsv_get_counters(nullptr, &g, nullptr, nullptr); return g; });
  m.def("obtain_move_counter", []() { return 0LL; });  // reset to 0 after every part (execute.hpp:306)
  m.def("obtain_gather_time", []() { long long u; hsv_get_counters(nullptr, nullptr, nullptr, &u); return u; });
  m.def("obtain_part_counter", []() { long long p; hsv_get_counters(nullptr, nullptr, &p, nullptr); return p; });
  m.def("set_num_threads", [](int n) { g_num_threads = n; });
  m.def("set_init_num_threads", [](int n) { g_init_threads = n; });
  m.def("get_num_threads", []() { return g_num_threads; });
  m.def("set_opt_slots", [](int on) { hsv_set_opt_slots(on); });
}
