// Gate fusion; see fusion.hpp for the reference mapping.
#include "fusion.hpp"

#include <algorithm>
#include <unordered_set>

namespace cast {
namespace {

// fusion.rs:45-83
bool absorb_single_qubit_gate(CircuitGraph &cg, size_t row, size_t qubit) {
  const GateId id = cg.gate_id_at(row, qubit);
  if (id == kNoGate)
    return false;
  GatePtr g = cg.gate_ptr(id);
  if (!g || g->n_qubits() != 1)
    return false;
  for (size_t nr = row + 1; nr < cg.n_rows(); ++nr) {
    const GateId nid = cg.gate_id_at(nr, qubit);
    if (nid == kNoGate)
      continue;
    GatePtr ng = cg.gate_ptr(nid);
    if (!ng)
      continue;
    auto fused = std::make_shared<QuantumGate>(ng->matmul(*g));
    cg.remove_gate_at(row, qubit);
    cg.remove_gate_at(nr, qubit);
    cg.insert_gate_at_row(nr, fused);
    return true;
  }
  for (size_t pr = row; pr-- > 0;) {
    const GateId pid = cg.gate_id_at(pr, qubit);
    if (pid == kNoGate)
      continue;
    GatePtr pg = cg.gate_ptr(pid);
    if (!pg)
      continue;
    auto fused = std::make_shared<QuantumGate>(g->matmul(*pg));
    cg.remove_gate_at(row, qubit);
    cg.remove_gate_at(pr, qubit);
    cg.insert_gate_at_row(pr, fused);
    return true;
  }
  return false;
}

size_t absorb_single_qubit_gates(CircuitGraph &cg) {
  size_t n = 0;
  const size_t rows = cg.n_rows(), nq = cg.n_qubits();
  for (size_t r = 0; r < rows; ++r)
    for (size_t q = 0; q < nq; ++q)
      n += absorb_single_qubit_gate(cg, r, q);
  return n;
}

// fusion.rs:103-134
bool fuse_adjacent_two_qubit_gate(CircuitGraph &cg, size_t row, size_t qubit) {
  const size_t next = row + 1;
  const GateId lid = cg.gate_id_at(row, qubit);
  if (lid == kNoGate)
    return false;
  GatePtr lg = cg.gate_ptr(lid);
  if (!lg || lg->n_qubits() != 2 || lg->qubits()[0] != qubit)
    return false;
  const GateId rid = cg.gate_id_at(next, qubit);
  if (rid == kNoGate)
    return false;
  GatePtr rg = cg.gate_ptr(rid);
  if (!rg || rg->n_qubits() != 2 || rg->qubits() != lg->qubits())
    return false;
  auto fused = std::make_shared<QuantumGate>(rg->matmul(*lg));
  cg.remove_gate_at(row, qubit);
  cg.remove_gate_at(next, qubit);
  cg.insert_gate_at_row(next, fused);
  return true;
}

size_t fuse_adjacent_two_qubit_gates(CircuitGraph &cg) {
  size_t n = 0;
  if (cg.n_rows() < 2)
    return 0;
  for (size_t row = 0; row + 1 < cg.n_rows(); ++row)
    for (size_t q = 0; q < cg.n_qubits(); ++q)
      n += fuse_adjacent_two_qubit_gate(cg, row, q);
  return n;
}

size_t union_qubit_count(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
  std::vector<uint32_t> u;
  for (uint32_t q : a)
    if (std::find(u.begin(), u.end(), q) == u.end())
      u.push_back(q);
  for (uint32_t q : b)
    if (std::find(u.begin(), u.end(), q) == u.end())
      u.push_back(q);
  return u.size();
}

// fusion.rs:168-303
size_t start_fusion(CircuitGraph &cg, const FusionConfig &cfg, size_t cdd, size_t start_row,
                    size_t start_qubit, FusionLog *log) {
  const GateId seed_id = cg.gate_id_at(start_row, start_qubit);
  if (seed_id == kNoGate)
    return 0;
  GatePtr seed = cg.gate_ptr(seed_id);
  if (seed->qubits()[0] != start_qubit)
    return 0;

  std::vector<std::pair<GatePtr, size_t>> tentative{{seed, start_row}};
  std::unordered_set<GateId> absorbed{seed_id};
  GateId prod_id = seed_id;
  size_t prod_row = start_row;
  GatePtr prod = seed;

  // Same-row sweep.
  for (size_t q = start_qubit + 1; q < cg.n_qubits(); ++q) {
    const GateId cid = cg.gate_id_at(prod_row, q);
    if (cid == kNoGate || absorbed.count(cid))
      continue;
    GatePtr cand = cg.gate_ptr(cid);
    if (union_qubit_count(prod->qubits(), cand->qubits()) > cdd)
      continue;
    tentative.emplace_back(cand, prod_row);
    absorbed.insert(cid);
    const size_t anchor = prod->qubits()[0];
    prod_id = *cg.fuse_gates_in_same_row(prod_row, anchor, q);
    absorbed.insert(prod_id);
    prod = cg.gate_ptr(prod_id);
  }

  // Cross-row loop.
  for (;;) {
    const size_t next = prod_row + 1;
    if (next >= cg.n_rows())
      break;
    const std::vector<uint32_t> pq = prod->qubits();
    bool progressed = false;
    for (uint32_t q : pq) {
      const GateId cid = cg.gate_id_at(next, q);
      if (cid == kNoGate || absorbed.count(cid))
        continue;
      GatePtr cand = cg.gate_ptr(cid);
      if (union_qubit_count(prod->qubits(), cand->qubits()) > cdd)
        continue;
      tentative.emplace_back(cand, next);
      absorbed.insert(cid);
      const size_t anchor = prod->qubits()[0];
      const auto res = *cg.fuse_gates_across_rows(prod_row, anchor, next, q);
      prod_id = res.first;
      prod_row = res.second;
      absorbed.insert(prod_id);
      prod = cg.gate_ptr(prod_id);
      progressed = true;
      break;
    }
    if (!progressed)
      break;
  }

  if (tentative.size() == 1)
    return 0;

  double old_cost = 0.0;
  for (const auto &t : tentative)
    old_cost += cfg.cost_model->cost_of(*t.first);
  const double new_cost = cfg.cost_model->cost_of(*prod);
  const double benefit = old_cost / (new_cost + 1e-10) - 1.0;
  const bool accepted = benefit >= cfg.benefit_margin;

  if (log) {
    FusionDecision d;
    d.cdd_size = cdd;
    d.n_candidates = tentative.size();
    for (const auto &t : tentative)
      d.candidates.emplace_back(t.first->effective_n_qubits(),
                                t.first->arithmetic_intensity(cfg.zero_tol));
    d.product_n_qubits = prod->effective_n_qubits();
    d.product_ai = prod->arithmetic_intensity(cfg.zero_tol);
    d.old_cost = old_cost;
    d.new_cost = new_cost;
    d.benefit = benefit;
    d.accepted = accepted;
    log->decisions.push_back(std::move(d));
  }

  if (!accepted) {
    cg.remove_gate_at(prod_row, prod->qubits()[0]);
    for (const auto &t : tentative)
      cg.insert_gate_at_row(t.second, t.first);
    return 0;
  }
  return tentative.size() - 1;
}

} // namespace

void apply_size_two_fusion(CircuitGraph &cg) {
  if (absorb_single_qubit_gates(cg) > 0)
    cg.squeeze();
  if (fuse_adjacent_two_qubit_gates(cg) > 0)
    cg.squeeze();
}

size_t apply_gate_fusion(CircuitGraph &cg, const FusionConfig &cfg, size_t cdd_size,
                         FusionLog *log) {
  size_t n = 0;
  const size_t rows = cg.n_rows(), nq = cg.n_qubits();
  for (size_t r = 0; r < rows; ++r)
    for (size_t q = 0; q < nq; ++q)
      n += start_fusion(cg, cfg, cdd_size, r, q, log);
  if (n > 0)
    cg.squeeze();
  return n;
}

void optimize(CircuitGraph &cg, const FusionConfig &cfg) {
  apply_size_two_fusion(cg);
  for (size_t cdd = 3; cdd <= cfg.size_max; ++cdd)
    while (apply_gate_fusion(cg, cfg, cdd, nullptr) > 0) {
    }
}

FusionLog optimize_with_log(CircuitGraph &cg, const FusionConfig &cfg) {
  apply_size_two_fusion(cg);
  FusionLog log;
  for (size_t cdd = 3; cdd <= cfg.size_max; ++cdd)
    while (apply_gate_fusion(cg, cfg, cdd, &log) > 0) {
    }
  return log;
}

} // namespace cast
