// Circuits, circuit graph and seeded generators; see circuit.hpp.
#include "circuit.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>

namespace cast {

void QuantumCircuit::add(GatePtr g) {
  for (uint32_t q : g->qubits())
    if (q >= n_)
      throw Error("qubit " + std::to_string(q) + " out of range for a " + std::to_string(n_) +
                  "-qubit circuit");
  gates_.push_back(std::move(g));
}

QuantumCircuit QuantumCircuit::qft(uint32_t n, bool swap) {
  QuantumCircuit c(n);
  for (uint32_t i = 0; i < n; ++i) {
    c.add(QuantumGate::h(i));
    for (uint32_t j = i + 1; j < n; ++j) {
      const uint32_t k = j - i;
      c.add(QuantumGate::cp(3.14159265358979323846 / double(uint64_t(1) << k), j, i));
    }
  }
  if (swap)
    for (uint32_t i = 0; i < n / 2; ++i)
      c.add(QuantumGate::swap(i, n - 1 - i));
  return c;
}

ComplexSquareMatrix haar_su2(Rng &rng) {
  auto u = ComplexSquareMatrix::random_unitary(2, rng);
  const cplx det = u.get(0, 0) * u.get(1, 1) - u.get(0, 1) * u.get(1, 0);
  const cplx r = std::sqrt(det);
  for (auto &v : u.data_mut())
    v /= r;
  return u;
}

QuantumCircuit random_layer_circuit(uint32_t n, uint32_t depth, uint64_t seed) {
  Rng rng(seed);
  QuantumCircuit c(n);
  std::vector<uint32_t> perm(n);
  for (uint32_t layer = 0; layer < depth; ++layer) {
    for (uint32_t q = 0; q < n; ++q)
      c.add(QuantumGate(haar_su2(rng), {q}));
    std::iota(perm.begin(), perm.end(), 0u);
    for (uint32_t i = n; i > 1; --i) {
      const uint32_t j = uint32_t(rng.below(i));
      std::swap(perm[i - 1], perm[j]);
    }
    for (uint32_t i = 0; i + 1 < n; i += 2)
      c.add(QuantumGate::cz(perm[i], perm[i + 1]));
  }
  return c;
}

std::vector<double> random_statevector(uint32_t n, uint64_t seed) {
  Rng rng(seed);
  const size_t N = size_t(1) << n;
  std::vector<double> v(2 * N);
  double s = 0.0;
  for (size_t i = 0; i < 2 * N; ++i) {
    v[i] = rng.normal();
    s += v[i] * v[i];
  }
  const double f = 1.0 / std::sqrt(s);
  for (auto &x : v)
    x *= f;
  return v;
}

// ── CircuitGraph ─────────────────────────────────────────────────────────

CircuitGraph CircuitGraph::from_circuit(const QuantumCircuit &c) {
  CircuitGraph g;
  g.ensure_n_qubits(c.n_qubits());
  for (const auto &gate : c.gates())
    g.insert_gate(gate);
  return g;
}

const QuantumGate *CircuitGraph::gate(GateId id) const {
  return id < gates_.size() ? gates_[id].get() : nullptr;
}

GatePtr CircuitGraph::gate_ptr(GateId id) const {
  return id < gates_.size() ? gates_[id] : nullptr;
}

GateId CircuitGraph::gate_id_at(size_t row, size_t qubit) const {
  if (row >= rows_.size() || qubit >= rows_[row].size())
    return kNoGate;
  return rows_[row][qubit];
}

void CircuitGraph::ensure_n_qubits(size_t required) {
  if (required <= n_qubits_)
    return;
  for (auto &r : rows_)
    r.resize(required, kNoGate);
  n_qubits_ = required;
}

size_t CircuitGraph::append_empty_row() {
  rows_.emplace_back(n_qubits_, kNoGate);
  return rows_.size() - 1;
}

bool CircuitGraph::is_row_vacant(size_t row, const std::vector<uint32_t> &qubits) const {
  for (uint32_t q : qubits)
    if (q < rows_[row].size() && rows_[row][q] != kNoGate)
      return false;
  return true;
}

size_t CircuitGraph::find_or_create_row(const QuantumGate &g) {
  size_t min_row = 0;
  for (uint32_t q : g.qubits())
    for (size_t r = rows_.size(); r-- > 0;)
      if (rows_[r][q] != kNoGate) {
        min_row = std::max(min_row, r + 1);
        break;
      }
  for (size_t r = min_row; r < rows_.size(); ++r)
    if (is_row_vacant(r, g.qubits()))
      return r;
  return append_empty_row();
}

GateId CircuitGraph::insert_gate(GatePtr g) {
  size_t req = 0;
  for (uint32_t q : g->qubits())
    req = std::max<size_t>(req, q + 1);
  ensure_n_qubits(req);
  const size_t row = find_or_create_row(*g);
  const GateId id = gates_.size();
  for (uint32_t q : g->qubits())
    rows_[row][q] = id;
  gates_.push_back(std::move(g));
  return id;
}

GateId CircuitGraph::insert_gate_at_row(size_t row, GatePtr g) {
  if (row >= rows_.size())
    throw Error("row index out of bounds");
  size_t req = 0;
  for (uint32_t q : g->qubits())
    req = std::max<size_t>(req, q + 1);
  ensure_n_qubits(req);
  if (!is_row_vacant(row, g->qubits()))
    throw Error("target row is not vacant for gate insertion");
  const GateId id = gates_.size();
  for (uint32_t q : g->qubits())
    rows_[row][q] = id;
  gates_.push_back(std::move(g));
  return id;
}

GatePtr CircuitGraph::remove_gate_at(size_t row, size_t qubit) {
  const GateId id = gate_id_at(row, qubit);
  if (id == kNoGate || !gates_[id])
    return nullptr;
  for (uint32_t q : gates_[id]->qubits())
    rows_[row][q] = kNoGate;
  GatePtr g = std::move(gates_[id]);
  gates_[id] = nullptr;
  return g;
}

std::vector<GateId> CircuitGraph::row_gate_ids(size_t row) const {
  std::vector<GateId> out;
  for (GateId id : rows_[row])
    if (id != kNoGate && std::find(out.begin(), out.end(), id) == out.end())
      out.push_back(id);
  return out;
}

std::vector<GateId> CircuitGraph::live_gate_ids_in_row_order() const {
  std::vector<GateId> out;
  for (size_t r = 0; r < rows_.size(); ++r) {
    auto ids = row_gate_ids(r);
    out.insert(out.end(), ids.begin(), ids.end());
  }
  return out;
}

void CircuitGraph::squeeze() {
  const auto live = live_gate_ids_in_row_order();
  std::vector<std::vector<GateId>> nr;
  std::vector<long> last(n_qubits_, -1);
  for (GateId id : live) {
    const QuantumGate &g = *gates_[id];
    long min_row = 0;
    for (uint32_t q : g.qubits())
      if (last[q] >= 0)
        min_row = std::max(min_row, last[q] + 1);
    size_t row = nr.size();
    for (size_t r = size_t(min_row); r < nr.size(); ++r) {
      bool vac = true;
      for (uint32_t q : g.qubits())
        vac = vac && nr[r][q] == kNoGate;
      if (vac) {
        row = r;
        break;
      }
    }
    if (row == nr.size())
      nr.emplace_back(n_qubits_, kNoGate);
    for (uint32_t q : g.qubits()) {
      nr[row][q] = id;
      last[q] = long(row);
    }
  }
  rows_ = std::move(nr);
}

std::optional<GateId> CircuitGraph::fuse_gates_in_same_row(size_t row, size_t qa, size_t qb) {
  const GateId a = gate_id_at(row, qa), b = gate_id_at(row, qb);
  if (a == kNoGate || b == kNoGate || a == b)
    return std::nullopt;
  GatePtr ga = gates_[a], gb = gates_[b];
  remove_gate_at(row, qa);
  remove_gate_at(row, qb);
  return insert_gate_at_row(row, std::make_shared<QuantumGate>(ga->matmul(*gb)));
}

std::optional<std::pair<GateId, size_t>>
CircuitGraph::fuse_gates_across_rows(size_t row_a, size_t qa, size_t row_b, size_t qb) {
  if (row_b != row_a + 1)
    throw Error("cross-row fusion requires consecutive rows");
  const GateId a = gate_id_at(row_a, qa), b = gate_id_at(row_b, qb);
  if (a == kNoGate || b == kNoGate)
    return std::nullopt;
  GatePtr ga = gates_[a], gb = gates_[b];
  remove_gate_at(row_a, qa);
  remove_gate_at(row_b, qb);
  auto fused = std::make_shared<QuantumGate>(gb->matmul(*ga));
  size_t target;
  if (is_row_vacant(row_b, fused->qubits())) {
    target = row_b;
  } else if (is_row_vacant(row_a, fused->qubits())) {
    target = row_a;
  } else {
    rows_.insert(rows_.begin() + long(row_b), std::vector<GateId>(n_qubits_, kNoGate));
    target = row_b;
  }
  const GateId id = insert_gate_at_row(target, fused);
  return std::make_pair(id, target);
}

std::vector<GatePtr> CircuitGraph::gates_in_row_order() const {
  std::vector<GatePtr> out;
  for (GateId id : live_gate_ids_in_row_order())
    if (gates_[id])
      out.push_back(gates_[id]);
  return out;
}

void CircuitGraph::check_consistency() const {
  for (size_t r = 0; r < rows_.size(); ++r)
    if (rows_[r].size() != n_qubits_)
      throw Error("row " + std::to_string(r) + " has width " + std::to_string(rows_[r].size()) +
                  ", expected " + std::to_string(n_qubits_));
  std::vector<long> seen_row(gates_.size(), -1);
  for (size_t r = 0; r < rows_.size(); ++r) {
    for (GateId id : row_gate_ids(r)) {
      if (id >= gates_.size() || !gates_[id])
        throw Error("row " + std::to_string(r) + " references dead or invalid gate id " +
                    std::to_string(id));
      if (seen_row[id] >= 0 && size_t(seen_row[id]) != r)
        throw Error("gate id " + std::to_string(id) + " appears on multiple rows");
      seen_row[id] = long(r);
      const auto &qs = gates_[id]->qubits();
      for (uint32_t q : qs) {
        if (q >= n_qubits_)
          throw Error("gate id " + std::to_string(id) + " targets qubit outside graph width");
        if (rows_[r][q] != id)
          throw Error("gate id " + std::to_string(id) + " is missing from row " +
                      std::to_string(r) + " at qubit " + std::to_string(q));
      }
      for (size_t q = 0; q < n_qubits_; ++q)
        if (rows_[r][q] == id && std::find(qs.begin(), qs.end(), uint32_t(q)) == qs.end())
          throw Error("row " + std::to_string(r) + " contains gate id " + std::to_string(id) +
                      " at unrelated qubit " + std::to_string(q));
    }
  }
  for (size_t id = 0; id < gates_.size(); ++id) {
    if (gates_[id] && seen_row[id] < 0)
      throw Error("live gate id " + std::to_string(id) + " is not placed in any row");
  }
}

} // namespace cast
