// Circuits, the circuit graph (rows x qubits grid) and the seeded circuit
// generators of the benchmark configurations.
//
// QuantumCircuit follows src/types/circuit.rs (ordered gate list; qft at
// :151-171).  CircuitGraph follows src/circuit_graph.rs: each row holds at
// most one gate per qubit, gates are pushed to the earliest row after their
// qubits' last use (:466-486), squeeze() re-packs rows preserving causal
// order (:356-394), gates_in_row_order() deduplicates multi-qubit cells
// (:490-504).
#pragma once

#include "types.hpp"

#include <cstdint>
#include <optional>
#include <vector>

namespace cast {

class QuantumCircuit {
public:
  explicit QuantumCircuit(uint32_t n_qubits) : n_(n_qubits) {}
  uint32_t n_qubits() const { return n_; }
  const std::vector<GatePtr> &gates() const { return gates_; }
  size_t len() const { return gates_.size(); }
  void add(GatePtr g);
  void add(const QuantumGate &g) { add(std::make_shared<QuantumGate>(g)); }

  // src/types/circuit.rs:151-171
  static QuantumCircuit qft(uint32_t n, bool swap);

private:
  uint32_t n_;
  std::vector<GatePtr> gates_;
};

// ── Seeded generators for the benchmark workloads (owned by this project;
//    the reference has no seeded layer circuits) ──────────────────────────

// Haar U(2) by row Gram-Schmidt, divided by sqrt(det) to land in SU(2).
ComplexSquareMatrix haar_su2(Rng &rng);

// `depth` layers; each layer = Haar SU(2) on every qubit, then CZ on a
// seeded random perfect matching (Fisher-Yates shuffle, consecutive pairs;
// with odd n the last qubit of the shuffle idles).
QuantumCircuit random_layer_circuit(uint32_t n, uint32_t depth, uint64_t seed);

// Complex Gaussian statevector (N(0,1) + i N(0,1)), normalised in f64;
// returned interleaved (re, im), 2 * 2^n doubles.
std::vector<double> random_statevector(uint32_t n, uint64_t seed);

// ── CircuitGraph ─────────────────────────────────────────────────────────

using GateId = size_t;
constexpr GateId kNoGate = ~size_t(0);

class CircuitGraph {
public:
  CircuitGraph() = default;
  static CircuitGraph from_circuit(const QuantumCircuit &c);

  size_t n_qubits() const { return n_qubits_; }
  size_t n_rows() const { return rows_.size(); }
  const std::vector<GatePtr> &gate_slots() const { return gates_; }

  const QuantumGate *gate(GateId id) const;
  GatePtr gate_ptr(GateId id) const;
  GateId gate_id_at(size_t row, size_t qubit) const;

  void ensure_n_qubits(size_t required);
  size_t append_empty_row();
  bool is_row_vacant(size_t row, const std::vector<uint32_t> &qubits) const;
  GateId insert_gate(GatePtr g);
  GateId insert_gate_at_row(size_t row, GatePtr g);
  GatePtr remove_gate_at(size_t row, size_t qubit);
  void squeeze();
  // circuit_graph.rs:403-427 / 429-464
  std::optional<GateId> fuse_gates_in_same_row(size_t row, size_t qa, size_t qb);
  std::optional<std::pair<GateId, size_t>> fuse_gates_across_rows(size_t row_a, size_t qa,
                                                                  size_t row_b, size_t qb);
  std::vector<GatePtr> gates_in_row_order() const;
  std::vector<GateId> row_gate_ids(size_t row) const;
  // Throws cast::Error describing the first inconsistency (circuit_graph.rs:178-250).
  void check_consistency() const;

private:
  size_t find_or_create_row(const QuantumGate &g);
  std::vector<GateId> live_gate_ids_in_row_order() const;

  size_t n_qubits_ = 0;
  std::vector<std::vector<GateId>> rows_;
  std::vector<GatePtr> gates_;
};

} // namespace cast
