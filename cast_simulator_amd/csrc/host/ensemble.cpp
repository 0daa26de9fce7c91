// Measurement, density-matrix simulation and ensemble trajectory sampling on
// the gfx950 backend.
//
//   marginal_probabilities  <- Backend::marginal_probabilities for Cuda
//                              (src/simulator/mod.rs:282-291), computed on the
//                              device: only 2^m probabilities reach the host.
//   QuantumState            <- src/simulator/mod.rs:320-414 (populations and
//                              trace of the density matrix from its diagonal,
//                              gathered on the device).
//   Simulator::simulate     <- src/simulator/mod.rs:523-547 + prepare_gates
//                              (:586-618).
//   batch_sample            <- src/simulator/measure.rs:45-78.
//   Simulator::sample_trajectory
//                           <- src/simulator/trajectory.rs:95-312.
//
// The reference sorts ensemble candidates with an unstable sort, so members
// of equal weight come out in an unspecified order; here the sort is stable
// (ties keep member-then-branch order), which makes the ensemble and its
// noise paths reproducible.  The reference draws samples from rand's StdRng;
// this project's seeded xoshiro256** generator (cast::Rng) takes its place,
// so histograms are reproducible for a seed but are not the reference's.
#include "hip_backend.hpp"

#include <algorithm>
#include <cmath>
#include <random>

namespace cast::hip {

namespace {
constexpr size_t kErr = 1024; // src/cuda/types.rs:72
} // namespace

std::vector<double> marginal_probabilities(const HipStatevector &sv,
                                           const std::vector<uint32_t> &measured) {
  if (measured.size() > sv.n_qubits())
    throw Error("marginal_probabilities: more measured qubits than the statevector has");
  std::vector<double> out(size_t(1) << measured.size());
  char err[kErr];
  throw_if(cast_hip_marginal_probabilities(sv.dptr(), sv.n_qubits(), uint8_t(sv.precision()),
                                           measured.data(), measured.size(), out.data(), nullptr,
                                           err, kErr),
           err);
  return out;
}

std::vector<double> QuantumState::populations() const {
  if (repr_ == Representation::StateVector) {
    std::vector<uint32_t> all(n_);
    for (uint32_t q = 0; q < n_; ++q)
      all[q] = q;
    return marginal_probabilities(sv_, all);
  }
  std::vector<double> out(size_t(1) << n_);
  char err[kErr];
  throw_if(cast_hip_dm_populations(sv_.dptr(), n_, uint8_t(sv_.precision()), out.data(), nullptr,
                                   err, kErr),
           err);
  return out;
}

double QuantumState::trace() const {
  if (repr_ == Representation::StateVector)
    return sv_.norm_squared();
  double t = 0.0;
  for (double p : populations())
    t += p;
  return t;
}

std::map<uint64_t, uint64_t> batch_sample(const std::vector<double> &probs, uint64_t n_samples,
                                          Rng &rng) {
  std::map<uint64_t, uint64_t> hist;
  if (n_samples == 0 || probs.empty())
    return hist;
  std::vector<double> cdf(probs.size());
  double acc = 0.0;
  for (size_t i = 0; i < probs.size(); ++i) {
    acc += probs[i];
    cdf[i] = acc;
  }
  const size_t last = probs.size() - 1;
  for (uint64_t s = 0; s < n_samples; ++s) {
    const double u = rng.uniform();
    const size_t k = size_t(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    ++hist[std::min(k, last)];
  }
  return hist;
}

QuantumState Simulator::simulate(const CircuitGraph &graph, Representation repr) {
  const uint32_t n_physical = uint32_t(graph.n_qubits());
  const auto gates = graph.gates_in_row_order();
  std::vector<KernelId> ids;
  ids.reserve(gates.size());
  uint32_t n_sv = n_physical;
  if (repr == Representation::StateVector) {
    for (size_t i = 0; i < gates.size(); ++i) {
      if (!gates[i]->is_unitary())
        throw Error("gate index " + std::to_string(i) +
                    " is non-unitary; Representation::StateVector requires unitary gates (use "
                    "DensityMatrix or sample_trajectory for noisy circuits)");
      ids.push_back(mgr_.generate(*gates[i]));
    }
  } else {
    n_sv = 2 * n_physical;
    for (const auto &g : gates)
      ids.push_back(mgr_.generate(g->to_density_matrix_gate(n_physical)));
  }
  HipStatevector sv(n_sv, spec_.precision);
  sv.zero(); // |0..0> on 2n qubits is rho = |0><0|
  for (KernelId id : ids)
    mgr_.apply(id, sv);
  mgr_.sync(false);
  return QuantumState(std::move(sv), n_physical, repr);
}

namespace {
struct Member {
  HipStatevector sv;
  double weight;
  std::vector<size_t> path;
};
struct Candidate {
  size_t member, branch;
  double weight;
};
} // namespace

TrajectoryResult Simulator::sample_trajectory(const CircuitGraph &graph,
                                              const TrajectoryOpts &opts) {
  const uint32_t n_physical = uint32_t(graph.n_qubits());
  for (uint32_t q : opts.measured_qubits)
    if (q >= n_physical)
      throw Error("measured qubit " + std::to_string(q) + " is out of range");
  const auto gates = graph.gates_in_row_order();

  // Base kernels for unitary gates; one kernel per noise branch otherwise
  // (trajectory.rs:131-182).
  std::vector<KernelId> base(gates.size(), 0);
  std::vector<std::vector<KernelId>> branch_ids(gates.size());
  std::vector<std::vector<double>> branch_w(gates.size());
  for (size_t i = 0; i < gates.size(); ++i) {
    const auto &g = gates[i];
    if (g->is_unitary()) {
      base[i] = mgr_.generate(*g);
      continue;
    }
    for (const auto &b : g->noise_model()->branches) {
      branch_ids[i].push_back(mgr_.generate(QuantumGate(b.second, g->qubits())));
      branch_w[i].push_back(b.first);
    }
  }

  // Ensemble (trajectory.rs:186-265).
  const size_t max_ensemble = std::max<size_t>(1, opts.max_ensemble);
  std::vector<Member> ens;
  {
    HipStatevector sv(n_physical, spec_.precision);
    sv.zero();
    ens.push_back(Member{std::move(sv), 1.0, {}});
  }
  for (size_t i = 0; i < gates.size(); ++i) {
    if (gates[i]->is_unitary()) {
      for (auto &m : ens)
        mgr_.apply(base[i], m.sv);
      continue;
    }
    const auto &kids = branch_ids[i];
    const auto &w = branch_w[i];
    std::vector<Candidate> cand;
    cand.reserve(ens.size() * w.size());
    for (size_t m = 0; m < ens.size(); ++m)
      for (size_t b = 0; b < w.size(); ++b)
        cand.push_back({m, b, ens[m].weight * w[b]});
    std::stable_sort(cand.begin(), cand.end(),
                     [](const Candidate &a, const Candidate &b) { return a.weight > b.weight; });
    const size_t kept = std::min(cand.size(), max_ensemble);
    cand.resize(kept);
    if (ens.size() == 1 && kept == 1) {
      ens[0].weight = cand[0].weight;
      ens[0].path.push_back(cand[0].branch);
      mgr_.apply(kids[cand[0].branch], ens[0].sv);
      continue;
    }
    mgr_.sync(false); // every queued gate lands before a member is copied
    std::vector<size_t> refs(ens.size(), 0);
    for (const auto &c : cand)
      ++refs[c.member];
    std::vector<Member> next;
    next.reserve(kept);
    for (const auto &c : cand) {
      Member &parent = ens[c.member];
      const bool last_ref = --refs[c.member] == 0;
      if (last_ref) {
        next.push_back(Member{std::move(parent.sv), c.weight, std::move(parent.path)});
      } else {
        next.push_back(Member{parent.sv.clone_device(), c.weight, parent.path});
      }
      next.back().path.push_back(c.branch);
      mgr_.apply(kids[c.branch], next.back().sv);
    }
    ens = std::move(next);
  }
  mgr_.sync(false);

  // Samples per member in proportion to weight (trajectory.rs:267-312).
  Rng rng(opts.has_seed ? opts.seed : (uint64_t(std::random_device{}()) << 32) ^
                                          uint64_t(std::random_device{}()));
  TrajectoryResult r;
  r.n_samples = opts.n_samples;
  double total = 0.0;
  for (const auto &m : ens)
    total += m.weight;
  uint64_t remaining = opts.n_samples;
  for (size_t idx = 0; idx < ens.size(); ++idx) {
    const Member &m = ens[idx];
    uint64_t ms = remaining;
    if (idx + 1 != ens.size())
      ms = std::min<uint64_t>(
          uint64_t(std::floor(m.weight / total * double(opts.n_samples))), remaining);
    remaining -= std::min(remaining, ms);
    if (ms > 0) {
      const auto probs = marginal_probabilities(m.sv, opts.measured_qubits);
      for (const auto &kv : batch_sample(probs, ms, rng))
        r.histogram[kv.first] += kv.second;
    }
    r.branches.push_back(ExploredBranch{m.weight, m.path, ms});
  }
  r.explored_weight = total;
  return r;
}

} // namespace cast::hip
