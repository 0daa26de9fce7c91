// Python bindings of the C++ host layer (module cast_simulator_amd._host).
// Names follow the reference's Rust API (types, CircuitGraph, fusion,
// cost_model, profile, cuda::{CudaKernelManager, CudaStatevector} ->
// Hip*, simulator::Simulator).
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <optional>

#include "circuit.hpp"
#include "cost_model.hpp"
#include "fusion.hpp"
#include "hip_backend.hpp"
#include "qasm.hpp"
#include "serving.hpp"
#include "trajectory.hpp"

#include <cstring>
#include "profile.hpp"
#include "types.hpp"

namespace py = pybind11;
using namespace cast;
using namespace cast::hip;

namespace {

py::array_t<cplx> to_numpy(const ComplexSquareMatrix &m) {
  const ssize_t n = ssize_t(m.edge_size());
  py::array_t<cplx> a({n, n});
  auto r = a.mutable_unchecked<2>();
  for (ssize_t i = 0; i < n; ++i)
    for (ssize_t j = 0; j < n; ++j)
      r(i, j) = m.get(size_t(i), size_t(j));
  return a;
}

ComplexSquareMatrix from_numpy(py::array_t<cplx, py::array::c_style | py::array::forcecast> a) {
  if (a.ndim() != 2 || a.shape(0) != a.shape(1))
    throw Error("matrix must be a square 2-D array");
  const size_t n = size_t(a.shape(0));
  auto r = a.unchecked<2>();
  ComplexSquareMatrix m(n);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j)
      m.set(i, j, r(ssize_t(i), ssize_t(j)));
  return m;
}

py::array_t<double> vec_to_numpy(const std::vector<double> &v) {
  py::array_t<double> a(ssize_t(v.size()));
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}

std::vector<double> numpy_to_vec(py::array_t<double, py::array::c_style | py::array::forcecast> a) {
  return std::vector<double>(a.data(), a.data() + a.size());
}

std::vector<SweepEntry> entries_from(const std::vector<std::tuple<double, double, double>> &t) {
  std::vector<SweepEntry> v;
  for (const auto &e : t)
    v.push_back({std::get<0>(e), std::get<1>(e), std::get<2>(e)});
  return v;
}

} // namespace

PYBIND11_MODULE(_host, m) {
  m.doc() = "CAST MI355X host layer: gate types, circuit graph, fusion, cost model, "
            "profiler and the gfx950 kernel manager / simulator.";
  py::register_exception<Error>(m, "CastError", PyExc_RuntimeError);

  // ── types ──────────────────────────────────────────────────────────────
  py::class_<QuantumGate, std::shared_ptr<QuantumGate>>(m, "QuantumGate")
      .def(py::init([](py::array_t<cplx> mat, std::vector<uint32_t> qubits) {
             return QuantumGate(from_numpy(mat), std::move(qubits));
           }),
           py::arg("matrix"), py::arg("qubits"))
      .def_static("from_reals", &QuantumGate::from_reals)
      .def_static("x", &QuantumGate::x)
      .def_static("y", &QuantumGate::y)
      .def_static("z", &QuantumGate::z)
      .def_static("h", &QuantumGate::h)
      .def_static("s", &QuantumGate::s)
      .def_static("t", &QuantumGate::t)
      .def_static("rx", &QuantumGate::rx)
      .def_static("ry", &QuantumGate::ry)
      .def_static("rz", &QuantumGate::rz)
      .def_static("u3", &QuantumGate::u3)
      .def_static("cx", &QuantumGate::cx)
      .def_static("cz", &QuantumGate::cz)
      .def_static("cp", &QuantumGate::cp)
      .def_static("swap", &QuantumGate::swap)
      .def_static("ccx", &QuantumGate::ccx)
      .def_static("depolarizing", &QuantumGate::depolarizing)
      .def_static("random_unitary",
                  [](std::vector<uint32_t> q, uint64_t seed) {
                    Rng rng(seed);
                    return QuantumGate::random_unitary(q, rng);
                  })
      .def_static("random_sparse",
                  [](std::vector<uint32_t> q, double sparsity, uint64_t seed) {
                    Rng rng(seed);
                    return QuantumGate::random_sparse(q, sparsity, rng);
                  })
      .def_static("random_arithmetic_intensity",
                  [](uint32_t n_sv, double ai, uint64_t seed) {
                    Rng rng(seed);
                    return QuantumGate::random_arithmetic_intensity(n_sv, ai, rng);
                  })
      .def("matrix", [](const QuantumGate &g) { return to_numpy(g.matrix()); })
      .def("qubits", &QuantumGate::qubits)
      .def("n_qubits", &QuantumGate::n_qubits)
      .def("effective_n_qubits", &QuantumGate::effective_n_qubits)
      .def("is_unitary", &QuantumGate::is_unitary)
      .def("noise_branches",
           [](const QuantumGate &g) {
             py::list out;
             if (g.noise_model())
               for (const auto &b : g.noise_model()->branches)
                 out.append(py::make_tuple(b.first, to_numpy(b.second)));
             return out;
           })
      .def("with_noise",
           [](const QuantumGate &g, const std::vector<std::pair<double, py::array_t<cplx>>> &ops) {
             std::vector<std::pair<double, ComplexSquareMatrix>> v;
             for (const auto &o : ops)
               v.emplace_back(o.first, from_numpy(o.second));
             return g.with_noise(v);
           })
      .def("with_depolarizing", &QuantumGate::with_depolarizing)
      .def("scalar_nnz", &QuantumGate::scalar_nnz)
      .def("arithmetic_intensity", &QuantumGate::arithmetic_intensity)
      .def("matmul", &QuantumGate::matmul)
      .def("to_density_matrix_gate", &QuantumGate::to_density_matrix_gate)
      .def("origin", &QuantumGate::origin)
      .def("with_origin", &QuantumGate::with_origin)
      .def("__eq__", &QuantumGate::operator==)
      .def("__repr__", [](const QuantumGate &g) {
        std::string s = "QuantumGate(k=" + std::to_string(g.n_qubits()) + ", qubits=[";
        for (size_t i = 0; i < g.qubits().size(); ++i)
          s += (i ? "," : "") + std::to_string(g.qubits()[i]);
        return s + (g.is_unitary() ? "])" : "], noisy)");
      });

  m.def("expand_matrix", [](py::array_t<cplx> mat, std::vector<uint32_t> gq,
                            std::vector<uint32_t> uq) {
    return to_numpy(expand_matrix(from_numpy(mat), gq, uq));
  });
  m.def("superoperator_of", [](py::array_t<cplx> mat) {
    return to_numpy(superoperator_of(from_numpy(mat)));
  });
  m.def("haar_su2", [](uint64_t seed) {
    Rng rng(seed);
    return to_numpy(haar_su2(rng));
  });
  m.def("random_unitary_matrix", [](size_t edge, uint64_t seed) {
    Rng rng(seed);
    return to_numpy(ComplexSquareMatrix::random_unitary(edge, rng));
  });
  m.def("random_statevector", [](uint32_t n, uint64_t seed) {
    return vec_to_numpy(random_statevector(n, seed));
  });

  py::class_<QuantumCircuit>(m, "QuantumCircuit")
      .def(py::init<uint32_t>())
      .def("add", [](QuantumCircuit &c, const QuantumGate &g) { c.add(g); })
      .def("gates",
           [](const QuantumCircuit &c) {
             std::vector<std::shared_ptr<QuantumGate>> out;
             for (const auto &g : c.gates())
               out.push_back(std::make_shared<QuantumGate>(*g));
             return out;
           })
      .def("n_qubits", &QuantumCircuit::n_qubits)
      .def("__len__", &QuantumCircuit::len)
      .def_static("qft", &QuantumCircuit::qft, py::arg("n"), py::arg("swap") = true)
      .def_static("from_qasm",
                  [](const std::string &text) { return circuit_from_qasm(parse_qasm(text)); })
      .def_static("from_openqasm", &circuit_from_qasm);

  py::class_<QasmAngle>(m, "QasmAngle")
      .def_readonly("pi_multiple", &QasmAngle::pi_multiple)
      .def_property_readonly("numerator", [](const QasmAngle &a) { return a.frac.num; })
      .def_property_readonly("denominator", [](const QasmAngle &a) { return a.frac.den; })
      .def_readonly("number", &QasmAngle::number)
      .def("radians", &QasmAngle::radians)
      .def("to_qasm", &QasmAngle::to_qasm);
  py::class_<QasmGate>(m, "QasmGate")
      .def("name", &QasmGate::name)
      .def_readonly("params", &QasmGate::params)
      .def_readonly("qubits", &QasmGate::qubits)
      .def("to_qasm", &QasmGate::to_qasm);
  py::class_<QasmCircuit>(m, "QasmCircuit")
      .def_readonly("gates", &QasmCircuit::gates)
      .def("required_qreg_size", &QasmCircuit::required_qreg_size)
      .def("serialize", &QasmCircuit::serialize);
  m.def("parse_qasm", &parse_qasm);
  m.def("qasm_to_quantum_gate", &to_quantum_gate);
  m.def("random_layer_circuit", &random_layer_circuit, py::arg("n"), py::arg("depth"),
        py::arg("seed"));

  py::class_<CircuitGraph>(m, "CircuitGraph")
      .def(py::init<>())
      .def_static("from_circuit", &CircuitGraph::from_circuit)
      .def_static("from_qasm_circuit", &graph_from_qasm)
      .def("n_qubits", &CircuitGraph::n_qubits)
      .def("n_rows", &CircuitGraph::n_rows)
      .def("insert_gate",
           [](CircuitGraph &g, const QuantumGate &q) {
             return g.insert_gate(std::make_shared<QuantumGate>(q));
           })
      .def("gates_in_row_order",
           [](const CircuitGraph &g) {
             std::vector<std::shared_ptr<QuantumGate>> out;
             for (const auto &p : g.gates_in_row_order())
               out.push_back(std::make_shared<QuantumGate>(*p));
             return out;
           })
      .def("row_gate_ids", &CircuitGraph::row_gate_ids)
      .def("gate_id_at",
           [](const CircuitGraph &g, size_t r, size_t q) -> py::object {
             const GateId id = g.gate_id_at(r, q);
             if (id == kNoGate)
               return py::none();
             return py::int_(id);
           })
      .def("check_consistency", &CircuitGraph::check_consistency)
      .def("squeeze", &CircuitGraph::squeeze)
      .def("copy", [](const CircuitGraph &g) { return CircuitGraph(g); });

  // ── cost model / fusion ────────────────────────────────────────────────
  py::class_<CostModel, std::shared_ptr<CostModel>>(m, "CostModel")
      .def("cost_of", &CostModel::cost_of);
  py::class_<SizeOnlyCostModel, CostModel, std::shared_ptr<SizeOnlyCostModel>>(m,
                                                                               "SizeOnlyCostModel")
      .def(py::init<size_t, size_t, double, bool>(), py::arg("max_size"),
           py::arg("max_ai") = ~size_t(0), py::arg("zero_tol") = 0.0,
           py::arg("trajectory_mode") = false);

  py::class_<ProfileConfig>(m, "ProfileConfig")
      .def(py::init<>())
      .def_readwrite("backend", &ProfileConfig::backend)
      .def_readwrite("name", &ProfileConfig::name)
      .def_readwrite("arch", &ProfileConfig::arch)
      .def_readwrite("n_threads", &ProfileConfig::n_threads)
      .def_readwrite("precision", &ProfileConfig::precision)
      .def_readwrite("n_qubits", &ProfileConfig::n_qubits);

  py::class_<HardwareProfile>(m, "HardwareProfile")
      .def(py::init<>())
      .def_static("from_roofline", &HardwareProfile::from_roofline)
      .def_static("from_measurements", &HardwareProfile::from_measurements)
      .def_static("from_json", &HardwareProfile::from_json)
      .def_static("load", &HardwareProfile::load)
      .def("to_json", &HardwareProfile::to_json)
      .def("save", &HardwareProfile::save)
      .def_readwrite("config", &HardwareProfile::config)
      .def_readwrite("peak_bw_gib_s", &HardwareProfile::peak_bw_gib_s)
      .def_readwrite("peak_gflops_s", &HardwareProfile::peak_gflops_s)
      .def_readwrite("bw_slope", &HardwareProfile::bw_slope)
      .def_readwrite("crossover_ai", &HardwareProfile::crossover_ai)
      .def_property_readonly("raw", [](const HardwareProfile &p) {
        py::list out;
        for (const auto &e : p.raw)
          out.append(py::make_tuple(e.ai, e.gflops_s, e.gib_s));
        return out;
      });
  m.def("compute_bw_slope", &compute_bw_slope);

  py::class_<HardwareAdaptiveCostModel, CostModel, std::shared_ptr<HardwareAdaptiveCostModel>>(
      m, "HardwareAdaptiveCostModel")
      .def(py::init<const HardwareProfile &, size_t, bool, double>(), py::arg("profile"),
           py::arg("max_size"), py::arg("trajectory_mode") = false, py::arg("zero_tol") = 1e-12)
      .def_readonly("crossover_ai", &HardwareAdaptiveCostModel::crossover_ai);

  py::class_<FusionConfig>(m, "FusionConfig")
      .def_static("size_only", &FusionConfig::size_only)
      .def_static("size_only_trajectory", &FusionConfig::size_only_trajectory)
      .def_static("aggressive", &FusionConfig::aggressive)
      .def_static("hardware_adaptive", &FusionConfig::hardware_adaptive)
      .def_static("hardware_adaptive_trajectory", &FusionConfig::hardware_adaptive_trajectory)
      .def_readwrite("size_max", &FusionConfig::size_max)
      .def_readwrite("benefit_margin", &FusionConfig::benefit_margin)
      .def_readwrite("zero_tol", &FusionConfig::zero_tol)
      .def("cost_of", [](const FusionConfig &c, const QuantumGate &g) {
        return c.cost_model->cost_of(g);
      });

  py::class_<FusionDecision>(m, "FusionDecision")
      .def_readonly("cdd_size", &FusionDecision::cdd_size)
      .def_readonly("n_candidates", &FusionDecision::n_candidates)
      .def_readonly("candidates", &FusionDecision::candidates)
      .def_readonly("product_n_qubits", &FusionDecision::product_n_qubits)
      .def_readonly("product_ai", &FusionDecision::product_ai)
      .def_readonly("old_cost", &FusionDecision::old_cost)
      .def_readonly("new_cost", &FusionDecision::new_cost)
      .def_readonly("benefit", &FusionDecision::benefit)
      .def_readonly("accepted", &FusionDecision::accepted);
  py::class_<FusionLog>(m, "FusionLog")
      .def_readonly("decisions", &FusionLog::decisions)
      .def("n_accepted", &FusionLog::n_accepted)
      .def("n_rejected", &FusionLog::n_rejected)
      .def("summary", &FusionLog::summary);

  m.def("apply_size_two_fusion", &apply_size_two_fusion);
  m.def("optimize", &optimize);
  m.def("optimize_with_log", &optimize_with_log);

  // ── profile / timing ───────────────────────────────────────────────────
  m.def("fit_roofline", [](const std::vector<std::tuple<double, double, double>> &s) {
    auto v = entries_from(s);
    std::sort(v.begin(), v.end(), [](auto &a, auto &b) { return a.ai < b.ai; });
    const auto p = fit_roofline(v);
    return py::make_tuple(p.crossover_ai, p.bw_slope, roofline_r2(v, p));
  });
  m.def("fit_from_samples",
        [](ProfileConfig c, const std::vector<std::tuple<double, double, double>> &s) {
          return fit_from_samples(std::move(c), entries_from(s));
        });
  m.def("adaptive_sweep",
        [](std::function<std::tuple<double, double, double>(double, double)> probe,
           ProfileConfig c, double budget) {
          return adaptive_sweep(
              [&](double ai, double b) {
                const auto t = probe(ai, b);
                return SweepEntry{std::get<0>(t), std::get<1>(t), std::get<2>(t)};
              },
              std::move(c), budget);
        });

  // ── gfx950 backend ─────────────────────────────────────────────────────
  py::enum_<cast_precision_t>(m, "Precision")
      .value("F32", CAST_PRECISION_F32)
      .value("F64", CAST_PRECISION_F64);

  py::class_<HipKernelGenSpec>(m, "HipKernelGenSpec")
      .def(py::init<>())
      .def_static("f32", &HipKernelGenSpec::f32)
      .def_static("f64", &HipKernelGenSpec::f64)
      .def_readwrite("precision", &HipKernelGenSpec::precision)
      .def_readwrite("ztol", &HipKernelGenSpec::ztol)
      .def_readwrite("otol", &HipKernelGenSpec::otol)
      .def_readwrite("variant", &HipKernelGenSpec::variant);

  py::class_<HipStatevector>(m, "HipStatevector")
      .def(py::init<uint32_t, cast_precision_t>())
      .def("n_qubits", &HipStatevector::n_qubits)
      .def("precision", &HipStatevector::precision)
      .def("__len__", &HipStatevector::len)
      .def("dptr", &HipStatevector::dptr)
      .def("zero", &HipStatevector::zero)
      .def("upload", [](HipStatevector &s, py::array_t<double> a) { s.upload(numpy_to_vec(a)); })
      .def("download", [](const HipStatevector &s) { return vec_to_numpy(s.download()); })
      .def("norm_squared", &HipStatevector::norm_squared)
      .def("normalize", &HipStatevector::normalize)
      .def("clone_device", &HipStatevector::clone_device)
      .def("inner", &HipStatevector::inner)
      .def("fidelity", &HipStatevector::fidelity)
      .def("expect_zz", &HipStatevector::expect_zz)
      .def("marginal_probabilities",
           [](const HipStatevector &s, const std::vector<uint32_t> &measured) {
             return vec_to_numpy(marginal_probabilities(s, measured));
           });

  py::enum_<Representation>(m, "Representation")
      .value("StateVector", Representation::StateVector)
      .value("DensityMatrix", Representation::DensityMatrix);

  py::class_<QuantumState>(m, "QuantumState")
      .def("n_qubits", &QuantumState::n_qubits)
      .def("is_pure", &QuantumState::is_pure)
      .def("representation", &QuantumState::representation)
      .def("populations", [](const QuantumState &s) { return vec_to_numpy(s.populations()); })
      .def("trace", &QuantumState::trace)
      .def("download_amplitudes",
           [](const QuantumState &s) { return vec_to_numpy(s.download_amplitudes()); })
      .def("statevector", &QuantumState::statevector, py::return_value_policy::reference_internal);

  py::class_<TrajectoryOpts>(m, "TrajectoryOpts")
      .def(py::init([](std::vector<uint32_t> measured, uint64_t n_samples,
                       std::optional<uint64_t> seed, std::optional<size_t> max_ensemble) {
             TrajectoryOpts o;
             o.measured_qubits = std::move(measured);
             o.n_samples = n_samples;
             o.has_seed = seed.has_value();
             o.seed = seed.value_or(0);
             o.max_ensemble = max_ensemble.value_or(1);
             return o;
           }),
           py::arg("measured_qubits"), py::arg("n_samples"), py::arg("seed") = py::none(),
           py::arg("max_ensemble") = py::none())
      .def_readwrite("measured_qubits", &TrajectoryOpts::measured_qubits)
      .def_readwrite("n_samples", &TrajectoryOpts::n_samples)
      .def_readwrite("max_ensemble", &TrajectoryOpts::max_ensemble);

  py::class_<ExploredBranch>(m, "ExploredBranch")
      .def_readonly("weight", &ExploredBranch::weight)
      .def_readonly("noise_path", &ExploredBranch::noise_path)
      .def_readonly("n_samples", &ExploredBranch::n_samples);

  py::class_<TrajectoryResult>(m, "TrajectoryResult")
      .def_readonly("histogram", &TrajectoryResult::histogram)
      .def_readonly("n_samples", &TrajectoryResult::n_samples)
      .def_readonly("branches", &TrajectoryResult::branches)
      .def_readonly("explored_weight", &TrajectoryResult::explored_weight);

  m.def(
      "batch_sample",
      [](const std::vector<double> &probs, uint64_t n_samples, uint64_t seed) {
        Rng rng(seed);
        return batch_sample(probs, n_samples, rng);
      },
      py::arg("probs"), py::arg("n_samples"), py::arg("seed"));

  py::class_<PlanInfo>(m, "PlanInfo")
      .def_readonly("k", &PlanInfo::k)
      .def_readonly("precision", &PlanInfo::precision)
      .def_readonly("variant", &PlanInfo::variant)
      .def_readonly("nnz_rows", &PlanInfo::nnz_rows)
      .def_readonly("scalar_nnz", &PlanInfo::scalar_nnz)
      .def_readonly("qubits", &PlanInfo::qubits);

  py::class_<HipKernelManager>(m, "HipKernelManager")
      .def(py::init<HipKernelGenSpec>())
      .def("generate", &HipKernelManager::generate)
      .def("apply", &HipKernelManager::apply)
      .def("sync",
           [](HipKernelManager &mgr, bool timed) {
             SyncStats st;
             {
               py::gil_scoped_release rel;
               st = mgr.sync(timed);
             }
             py::list ks;
             for (const auto &k : st.kernels)
               ks.append(py::make_tuple(k.id, k.gpu_s));
             return py::make_tuple(ks, st.wall_s);
           },
           py::arg("timed") = true)
      .def("time_adaptive",
           [](HipKernelManager &mgr, KernelId id, HipStatevector &sv, double budget) {
             const auto t = mgr.time_adaptive(id, sv, budget);
             return py::dict(py::arg("n_iters") = t.n_iters, py::arg("mean_s") = t.mean_s,
                             py::arg("stddev_s") = t.stddev_s, py::arg("min_s") = t.min_s,
                             py::arg("max_s") = t.max_s);
           })
      .def("plan_info", &HipKernelManager::plan_info)
      .def("family", &HipKernelManager::family, py::arg("id"), py::arg("n_qubits"),
           "Kernel a launch of `id` on an n-qubit statevector runs: 1 direct, 2 monomial, "
           "3 wave tile, 4 matrix core (the CAST_HIP_VARIANT_* codes), or the autotuner's "
           "choice: 5 wave tile bit-reversed, 6 three-product wave tile, 7 both, 8 wave tile "
           "in XCD order, 9 three-product in XCD order, 10 matrix core with f32 products, "
           "11 matrix core in XCD order, 12 matrix core bit-reversed, 13 wave tile in XCD "
           "order bit-reversed, 14 three-product likewise, 15 matrix core likewise.")
      .def("n_kernels", &HipKernelManager::n_kernels)
      .def("set_autotune", &HipKernelManager::set_autotune, py::arg("on"),
           py::arg("min_qubits") = 22, py::arg("samples") = 2)
      .def("autotune", &HipKernelManager::autotune)
      .def("tune", &HipKernelManager::tune, py::arg("ids"), py::arg("sv"),
           py::arg("max_rounds") = 16)
      .def("tuning_report", [](const HipKernelManager &m) {
        py::list out;
        for (const auto &e : m.tuning_report())
          out.append(py::dict(py::arg("id") = e.id, py::arg("n_qubits") = e.n_qubits,
                              py::arg("chosen") = e.chosen_family,
                              py::arg("best_s") = e.best_s));
        return out;
      });

  py::class_<Simulator>(m, "Simulator")
      .def(py::init<HipKernelGenSpec>())
      .def_static("f32", [] { return std::make_unique<Simulator>(HipKernelGenSpec::f32()); })
      .def_static("f64", [] { return std::make_unique<Simulator>(HipKernelGenSpec::f64()); })
      .def("compile", &Simulator::compile)
      .def("run",
           [](Simulator &s, const CircuitGraph &g, HipStatevector &sv) {
             const auto st = s.run(g, sv);
             py::list ks;
             for (const auto &k : st.kernels)
               ks.append(py::make_tuple(k.id, k.gpu_s));
             return py::make_tuple(ks, st.wall_s);
           })
      .def("simulate", py::overload_cast<const CircuitGraph &>(&Simulator::simulate))
      .def("simulate",
           py::overload_cast<const CircuitGraph &, Representation>(&Simulator::simulate))
      .def("sample_trajectory",
           [](Simulator &s, const CircuitGraph &g, const TrajectoryOpts &o) {
             py::gil_scoped_release rel;
             return s.sample_trajectory(g, o);
           })
      .def("bench",
           [](Simulator &s, const CircuitGraph &g, double budget, Representation repr) {
             const auto r = s.bench(g, budget, repr);
             return py::dict(py::arg("compile_s") = r.compile_s,
                             py::arg("samples_s") = r.samples_s, py::arg("n_gates") = r.n_gates,
                             py::arg("mean_s") = r.exec.mean_s, py::arg("min_s") = r.exec.min_s);
           },
           py::arg("graph"), py::arg("budget_s"),
           py::arg("repr") = Representation::StateVector)
      .def("manager", &Simulator::manager, py::return_value_policy::reference_internal);

  m.def("measure_hip",
        [](HipKernelGenSpec spec, uint32_t n_qubits, double budget_s, uint64_t seed) {
          HipKernelManager mgr(spec);
          HipStatevector sv(n_qubits, spec.precision);
          sv.zero();
          ProfileConfig c;
          c.backend = "hip";
          cast_hip_device_info_t di{};
          char err[1024];
          if (cast_hip_device_info(&di, err, sizeof(err)) == 0) {
            c.name = di.name;
            c.arch = di.arch;
          }
          c.precision = spec.precision == CAST_PRECISION_F32 ? "f32" : "f64";
          c.n_qubits = n_qubits;
          return measure(c, n_qubits, spec.scalar_bytes(), budget_s, seed,
                         [&](const QuantumGate &g, double b) {
                           const KernelId id = mgr.generate(g);
                           // Settle the per-gate kernel choice first, so the
                           // timed samples all run the kernel bench runs.
                           mgr.tune({id}, sv, 12);
                           return mgr.time_adaptive(id, sv, b).mean_s;
                         });
        },
        py::arg("spec"), py::arg("n_qubits"), py::arg("budget_s"), py::arg("seed") = 1);

  // ── batched trajectories ───────────────────────────────────────────────
  py::class_<LayeredCircuit>(m, "LayeredCircuit")
      .def_static("random_layers", &LayeredCircuit::random_layers, py::arg("n"),
                  py::arg("depth"), py::arg("seed"))
      .def_static("from_circuit", &LayeredCircuit::from_circuit, py::arg("circuit"))
      .def_readonly("n_qubits", &LayeredCircuit::n_qubits)
      .def_readonly("layer_start", &LayeredCircuit::layer_start)
      .def("n_layers", &LayeredCircuit::n_layers)
      .def("gates", [](const LayeredCircuit &c) {
        std::vector<std::shared_ptr<QuantumGate>> out;
        for (const auto &g : c.gates)
          out.push_back(std::make_shared<QuantumGate>(*g));
        return out;
      });

  py::class_<NoiseSpec>(m, "NoiseSpec")
      .def(py::init([](double p, uint32_t min_depth, uint32_t max_depth, uint64_t seed) {
             NoiseSpec n;
             n.p = p;
             n.min_depth = min_depth;
             n.max_depth = max_depth;
             n.seed = seed;
             return n;
           }),
           py::arg("p"), py::arg("min_depth"), py::arg("max_depth"), py::arg("seed"))
      .def_readwrite("p", &NoiseSpec::p)
      .def_readwrite("min_depth", &NoiseSpec::min_depth)
      .def_readwrite("max_depth", &NoiseSpec::max_depth)
      .def_readwrite("seed", &NoiseSpec::seed);

  // Events as tuples (gate, qubit, pauli) with pauli 1..3 for a Pauli on
  // `qubit` after the gate, or (gate, branch, 0) for a noisy gate that took
  // noise branch `branch` instead of its most likely one.
  py::class_<TrajectorySample>(m, "TrajectorySample")
      .def(py::init([](uint32_t depth, const std::vector<std::tuple<uint32_t, uint32_t, uint8_t>> &e) {
             TrajectorySample s;
             s.depth = depth;
             for (const auto &[g, q, p] : e) {
               if (p == 0)
                 s.errors.push_back({g, 0, 0, q});
               else
                 s.errors.push_back({g, q, p, 0});
             }
             return s;
           }),
           py::arg("depth"), py::arg("errors"))
      .def_readonly("depth", &TrajectorySample::depth)
      .def_property_readonly("errors", [](const TrajectorySample &s) {
        std::vector<std::tuple<uint32_t, uint32_t, uint8_t>> out;
        for (const auto &e : s.errors)
          out.emplace_back(e.gate, e.pauli == 0 ? e.branch : e.qubit, e.pauli);
        return out;
      });
  m.def("clean_branch", [](const QuantumGate &g) {
    if (!g.noise_model())
      throw Error("clean_branch: the gate has no noise model");
    return clean_branch(*g.noise_model());
  });

  m.def("sample_trajectories", &sample_trajectories, py::arg("circuit"), py::arg("noise"),
        py::arg("count"), py::arg("first") = 0);
  m.def("trajectory_circuit", &trajectory_circuit);

  py::class_<TrajectoryTuning>(m, "TrajectoryTuning")
      .def(py::init([](uint32_t chunk_layers, uint32_t fusion_max, uint32_t block,
                       uint32_t slots_per_cu, uint32_t max_slots, uint32_t waves_per_simd,
                       uint32_t mfma_f32) {
             TrajectoryTuning t;
             t.chunk_layers = chunk_layers;
             t.fusion_max = fusion_max;
             t.block = block;
             t.slots_per_cu = slots_per_cu;
             t.max_slots = max_slots;
             t.waves_per_simd = waves_per_simd;
             t.mfma_f32 = mfma_f32;
             return t;
           }),
           py::arg("chunk_layers") = 4, py::arg("fusion_max") = 4, py::arg("block") = 256,
           py::arg("slots_per_cu") = 2, py::arg("max_slots") = 0, py::arg("waves_per_simd") = 2,
           py::arg("mfma_f32") = 0)
      .def_readwrite("waves_per_simd", &TrajectoryTuning::waves_per_simd)
      .def_readwrite("mfma_f32", &TrajectoryTuning::mfma_f32)
      .def_readwrite("chunk_layers", &TrajectoryTuning::chunk_layers)
      .def_readwrite("fusion_max", &TrajectoryTuning::fusion_max)
      .def_readwrite("block", &TrajectoryTuning::block)
      .def_readwrite("slots_per_cu", &TrajectoryTuning::slots_per_cu)
      .def_readwrite("max_slots", &TrajectoryTuning::max_slots)
      .def("__repr__", [](const TrajectoryTuning &t) {
        return "TrajectoryTuning(chunk_layers=" + std::to_string(t.chunk_layers) +
               ", fusion_max=" + std::to_string(t.fusion_max) +
               ", block=" + std::to_string(t.block) +
               ", slots_per_cu=" + std::to_string(t.slots_per_cu) +
               ", max_slots=" + std::to_string(t.max_slots) +
               ", waves_per_simd=" + std::to_string(t.waves_per_simd) +
               ", mfma_f32=" + std::to_string(t.mfma_f32) + ")";
      });

  py::class_<TrajectoryRun>(m, "TrajectoryRun")
      .def_readonly("n_pairs", &TrajectoryRun::n_pairs)
      .def_readonly("task_of", &TrajectoryRun::task_of)
      .def_readonly("task_ckpt_layer", &TrajectoryRun::task_ckpt_layer)
      .def_readonly("task_gates", &TrajectoryRun::task_gates)
      .def_readonly("norm_sq", &TrajectoryRun::norm_sq)
      .def_readonly("zz_sum", &TrajectoryRun::zz_sum)
      .def_readonly("traj_zz", &TrajectoryRun::traj_zz)
      .def_readonly("mean_zz", &TrajectoryRun::mean_zz)
      .def_readonly("min_norm_sq", &TrajectoryRun::min_norm_sq)
      .def_readonly("max_norm_sq", &TrajectoryRun::max_norm_sq)
      .def_readonly("plan_s", &TrajectoryRun::plan_s)
      .def_readonly("checkpoint_s", &TrajectoryRun::checkpoint_s)
      .def_readonly("gpu_s", &TrajectoryRun::gpu_s)
      .def_readonly("total_s", &TrajectoryRun::total_s)
      .def_readonly("n_tasks", &TrajectoryRun::n_tasks)
      .def_readonly("n_gate_passes", &TrajectoryRun::n_gate_passes)
      .def_readonly("n_shared_clean", &TrajectoryRun::n_shared_clean)
      .def_readonly("n_dirty_matrices", &TrajectoryRun::n_dirty_matrices)
      .def_readonly("n_slots", &TrajectoryRun::n_slots)
      .def_readonly("block", &TrajectoryRun::block)
      .def_readonly("chunk_layers", &TrajectoryRun::chunk_layers)
      .def_readonly("fusion_max", &TrajectoryRun::fusion_max)
      .def_readonly("waves_per_simd", &TrajectoryRun::waves_per_simd)
      .def_readonly("mfma_f32", &TrajectoryRun::mfma_f32)
      .def("dump", [](const TrajectoryRun &r) {
        return py::array_t<double>(py::ssize_t(r.dump.size()), r.dump.data());
      });

  py::class_<TrajectoryEngine>(m, "TrajectoryEngine")
      .def(py::init([](const LayeredCircuit &c, cast_precision_t prec,
                       std::vector<uint32_t> pairs) {
             return std::make_unique<TrajectoryEngine>(c, prec, std::move(pairs));
           }),
           py::arg("circuit"), py::arg("precision"), py::arg("pairs"))
      .def("set_initial_state", &TrajectoryEngine::set_initial_state)
      .def("plan",
           [](TrajectoryEngine &e, const std::vector<TrajectorySample> &samples,
              const TrajectoryTuning &t) {
             const TrajectoryPlan P = e.plan(samples, t);
             py::list tasks, gates;
             for (const auto &k : P.tasks)
               tasks.append(py::make_tuple(k.first_gate, k.n_gates, k.checkpoint));
             for (const auto &g : P.gates) {
               std::vector<uint32_t> q(g.q, g.q + g.k);
               gates.append(py::make_tuple(g.matrix_offset, q));
             }
             py::array_t<cplx> pool(py::ssize_t(P.pool.size() / 2));
             std::memcpy(pool.mutable_data(), P.pool.data(), P.pool.size() * sizeof(double));
             return py::dict(py::arg("tasks") = tasks, py::arg("gates") = gates,
                             py::arg("pool") = pool, py::arg("order") = P.order,
                             py::arg("task_of") = P.task_of,
                             py::arg("n_shared_clean") = P.n_shared_clean,
                             py::arg("n_dirty_matrices") = P.n_dirty_matrices,
                             py::arg("plan_s") = P.plan_s);
           },
           py::arg("samples"), py::arg("tuning"))
      .def("run", &TrajectoryEngine::run, py::arg("samples"), py::arg("tuning"),
           py::arg("dump") = false, py::call_guard<py::gil_scoped_release>())
      .def("autotune",
           [](TrajectoryEngine &e, const std::vector<TrajectorySample> &sample,
              const TrajectoryTuning &start) {
             std::vector<TrajectoryTuneRecord> log;
             TrajectoryTuning best;
             {
               py::gil_scoped_release rel;
               best = e.autotune(sample, start, &log);
             }
             py::list trials;
             for (const auto &r : log)
               trials.append(py::make_tuple(r.tuning, r.s_per_trajectory));
             return py::make_tuple(best, trials);
           },
           py::arg("sample"), py::arg("start"));

  // ── serving ──────────────────────────────────────────────────────────────
  py::class_<ServeRequest>(m, "ServeRequest")
      .def(py::init([](uint32_t id, uint32_t n, uint32_t depth, uint64_t seed,
                       cast_precision_t prec, double arrival_s) {
             ServeRequest r;
             r.id = id;
             r.n_qubits = n;
             r.depth = depth;
             r.seed = seed;
             r.precision = prec;
             r.arrival_s = arrival_s;
             return r;
           }),
           py::arg("id"), py::arg("n_qubits"), py::arg("depth"), py::arg("seed"),
           py::arg("precision"), py::arg("arrival_s") = 0.0)
      .def_static(
          "from_circuit",
          [](uint32_t id, const QuantumCircuit &c, cast_precision_t prec, double arrival_s) {
            return circuit_request(id, std::make_shared<const QuantumCircuit>(c), prec, arrival_s);
          },
          py::arg("id"), py::arg("circuit"), py::arg("precision"), py::arg("arrival_s") = 0.0)
      .def_static(
          "from_qasm",
          [](uint32_t id, const std::string &text, cast_precision_t prec, double arrival_s) {
            return circuit_request(
                id, std::make_shared<const QuantumCircuit>(circuit_from_qasm(parse_qasm(text))),
                prec, arrival_s);
          },
          py::arg("id"), py::arg("qasm"), py::arg("precision"), py::arg("arrival_s") = 0.0)
      .def_property_readonly("has_circuit",
                             [](const ServeRequest &r) { return bool(r.circuit); })
      .def("with_precision",
           [](const ServeRequest &r, cast_precision_t prec) {
             ServeRequest o = r;
             o.precision = prec;
             return o;
           })
      .def_readwrite("id", &ServeRequest::id)
      .def_readwrite("n_qubits", &ServeRequest::n_qubits)
      .def_readwrite("depth", &ServeRequest::depth)
      .def_readwrite("seed", &ServeRequest::seed)
      .def_readwrite("precision", &ServeRequest::precision)
      .def_readwrite("arrival_s", &ServeRequest::arrival_s);
  m.def("poisson_requests", &poisson_requests, py::arg("count"), py::arg("rate"),
        py::arg("n_min"), py::arg("n_max"), py::arg("depth_min"), py::arg("depth_max"),
        py::arg("f32_fraction"), py::arg("seed"));

  py::class_<ServeConfig>(m, "ServeConfig")
      .def(py::init<>())
      .def_readwrite("streams", &ServeConfig::streams)
      .def_readwrite("compile_threads", &ServeConfig::compile_threads)
      .def_readwrite("fusion_max", &ServeConfig::fusion_max)
      .def_readwrite("budget_bytes", &ServeConfig::budget_bytes)
      .def_readwrite("reserve_bytes", &ServeConfig::reserve_bytes)
      .def_readwrite("realtime", &ServeConfig::realtime)
      .def_readwrite("keep_states", &ServeConfig::keep_states)
      .def_readwrite("f64_shadow", &ServeConfig::f64_shadow)
      .def_readwrite("batch_max_qubits", &ServeConfig::batch_max_qubits)
      .def_readwrite("batch_size", &ServeConfig::batch_size);

  py::class_<ServeResult>(m, "ServeResult")
      .def_readonly("id", &ServeResult::id)
      .def_readonly("n_qubits", &ServeResult::n_qubits)
      .def_readonly("depth", &ServeResult::depth)
      .def_readonly("n_gates", &ServeResult::n_gates)
      .def_readonly("precision", &ServeResult::precision)
      .def_readonly("arrival_s", &ServeResult::arrival_s)
      .def_readonly("compiled_s", &ServeResult::compiled_s)
      .def_readonly("admit_s", &ServeResult::admit_s)
      .def_readonly("done_s", &ServeResult::done_s)
      .def_readonly("norm_sq", &ServeResult::norm_sq)
      .def_readonly("zz", &ServeResult::zz)
      .def_readonly("fidelity_vs_f64", &ServeResult::fidelity_vs_f64)
      .def("state", [](const ServeResult &r) {
        return py::array_t<double>(py::ssize_t(r.state.size()), r.state.data());
      });

  py::class_<Server>(m, "Server")
      .def(py::init([](const ServeConfig &c) { return std::make_unique<Server>(c); }))
      .def("run", [](Server &s, const std::vector<ServeRequest> &reqs) {
        ServeStats st;
        std::vector<ServeResult> res;
        {
          py::gil_scoped_release rel;
          res = s.run(reqs, &st);
        }
        py::dict d(py::arg("wall_s") = st.wall_s, py::arg("throughput") = st.throughput,
                   py::arg("p50_s") = st.p50_s, py::arg("p99_s") = st.p99_s,
                   py::arg("mean_s") = st.mean_s, py::arg("max_s") = st.max_s,
                   py::arg("queue_p50_s") = st.queue_p50_s,
                   py::arg("queue_p99_s") = st.queue_p99_s,
                   py::arg("max_in_flight") = st.max_in_flight,
                   py::arg("peak_sv_bytes") = st.peak_sv_bytes,
                   py::arg("budget_bytes") = st.budget_bytes, py::arg("sv_allocs") = st.sv_allocs,
                   py::arg("sv_reuses") = st.sv_reuses, py::arg("batches") = st.batches,
                   py::arg("batched_requests") = st.batched_requests,
                   py::arg("batch_launches") = st.batch_launches);
        return py::make_tuple(res, d);
      });
}
