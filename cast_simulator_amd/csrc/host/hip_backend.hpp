// Host-side gfx950 backend: statevector handle, kernel manager and the
// simulator driver, all built on the C ABI of include/cast_hip.h.
//
// Mirrors the reference's Rust CUDA backend:
//   HipStatevector   <- src/cuda/statevector.rs:44-253 (alloc/zero/upload/
//                       download/norm_squared/normalize/clone_device)
//   HipKernelManager <- src/cuda/kernel.rs:705-1253: generate() builds (or
//                       reuses, keyed on matrix bytes + qubits, :213-232) a
//                       gate plan; apply() validates and queues (:894-924);
//                       sync() launches the queue on the manager's stream with
//                       events between launches and waits (:935-977), returning
//                       per-kernel GPU times.  There are no modules to cache:
//                       a plan is ready to launch once built.
//   Simulator        <- src/simulator/mod.rs:523-547 (simulate) and
//                       src/simulator/bench.rs:136-167 (bench; a sample is
//                       the sum of per-kernel GPU event times, mod.rs:295-307).
#pragma once

#include "cast_hip.h"
#include "circuit.hpp"
#include "profile.hpp"

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace cast::hip {

struct HipKernelGenSpec {
  cast_precision_t precision = CAST_PRECISION_F64;
  double ztol = 1e-12;
  double otol = 1e-12;
  uint32_t variant = CAST_HIP_VARIANT_AUTO;
  // src/cuda/types.rs:31-56 defaults.
  static HipKernelGenSpec f32() { return {CAST_PRECISION_F32, 1e-6, 1e-6, 0}; }
  static HipKernelGenSpec f64() { return {CAST_PRECISION_F64, 1e-12, 1e-12, 0}; }
  size_t scalar_bytes() const { return precision == CAST_PRECISION_F32 ? 4 : 8; }
};

// Throws cast::Error with the library's message when a C-ABI call fails.
void throw_if(int rc, const char *err);

class HipStatevector {
public:
  HipStatevector(uint32_t n_qubits, cast_precision_t precision);
  ~HipStatevector();
  HipStatevector(const HipStatevector &) = delete;
  HipStatevector &operator=(const HipStatevector &) = delete;
  HipStatevector(HipStatevector &&o) noexcept;

  uint32_t n_qubits() const { return n_; }
  cast_precision_t precision() const { return prec_; }
  size_t len() const { return size_t(1) << n_; }
  size_t n_elements() const { return n_elements_; }
  uint64_t dptr() const { return dptr_; }

  void zero();
  void upload(const std::vector<double> &interleaved);
  std::vector<double> download() const;
  double norm_squared() const;
  void normalize();
  HipStatevector clone_device() const;
  std::vector<double> expect_zz(const std::vector<uint32_t> &pairs) const;
  // <this|other> accumulated in f64 on the device (precisions may differ).
  cplx inner(const HipStatevector &other) const;
  // |<this|other>|^2 / (<this|this> <other|other>).
  double fidelity(const HipStatevector &other) const;

private:
  uint64_t dptr_ = 0;
  uint32_t n_ = 0;
  cast_precision_t prec_ = CAST_PRECISION_F64;
  size_t n_elements_ = 0;
};

using KernelId = uint64_t;

struct KernelExecTime {
  KernelId id = 0;
  double gpu_s = 0.0;
};

struct SyncStats {
  std::vector<KernelExecTime> kernels;
  double wall_s = 0.0;
};

struct PlanInfo {
  uint32_t k = 0, precision = 0, variant = 0, nnz_rows = 0;
  uint64_t scalar_nnz = 0;
  std::vector<uint32_t> qubits;
};

// Per-gate kernel autotuning: for a dense gate on a statevector of at least
// `min_qubits` qubits, every applicable kernel (wave tile in four orders and
// two product forms, matrix core in four orders and, at f32 k >= 5, two
// product forms, direct) is timed on the real gate, and the
// fastest (by minimum) is used from then on.  Keyed by (kernel id, n).
// Successive halving: every candidate is timed once; candidates slower than
// kTuneMargin x the best so far are dropped; the rest are timed until each
// has `samples` timings.
constexpr double kTuneMargin = 1.08;
// Candidate code of the wave-tile kernel visiting tiles in swizzled order
// (the other codes are the CAST_HIP_VARIANT_* families).
constexpr uint32_t kTileSwizzled = 5;
// ... and of the wave tiles in the three-product form (CAST_HIP_FLAG_TILE_GAUSS),
// in natural and swizzled order.
constexpr uint32_t kTileGauss = 6;
constexpr uint32_t kTileGaussSwizzled = 7;
// ... and both wave-tile forms in XCD-blocked order (CAST_HIP_FLAG_TILE_XCD).
constexpr uint32_t kTileXcd = 8;
constexpr uint32_t kTileGaussXcd = 9;
// ... and of the matrix-core kernel with f32 products for f32 k = 5, 6
// (CAST_HIP_FLAG_MFMA_F32), whose default there is the bf16 split products.
constexpr uint32_t kMfmaF32 = 10;
// ... and of the matrix-core kernel (default product form) in XCD-blocked
// order (CAST_HIP_FLAG_TILE_XCD) and in swizzled order
// (CAST_HIP_FLAG_TILE_SWIZZLE).
constexpr uint32_t kMfmaXcd = 11;
constexpr uint32_t kMfmaSwizzled = 12;
// ... and both wave-tile forms in XCD-blocked order with the low tile bits
// bit-reversed within each XCD's range (both flags).
constexpr uint32_t kTileXcdSwizzled = 13;
constexpr uint32_t kTileGaussXcdSwizzled = 14;
// ... and the matrix-core kernel with both orders.
constexpr uint32_t kMfmaXcdSwizzled = 15;

struct TuneEntry {
  KernelId id = 0;
  uint32_t n_qubits = 0;
  int chosen_family = -1; // CAST_HIP_VARIANT_* or kTileSwizzled, -1 while tuning
  std::vector<std::pair<uint32_t, double>> best_s; // (candidate code, min seconds)
};

class HipKernelManager {
public:
  explicit HipKernelManager(HipKernelGenSpec spec);
  ~HipKernelManager();
  HipKernelManager(const HipKernelManager &) = delete;
  HipKernelManager &operator=(const HipKernelManager &) = delete;

  const HipKernelGenSpec &spec() const { return spec_; }
  KernelId generate(const QuantumGate &gate);
  void apply(KernelId id, HipStatevector &sv);
  SyncStats sync(bool timed = true);
  TimingStats time_adaptive(KernelId id, HipStatevector &sv, double budget_s);
  PlanInfo plan_info(KernelId id) const;
  // Kernel family (CAST_HIP_VARIANT_*) launched for `id` on n_qubits; with
  // autotuning, the tuned candidate once decided (kTileSwizzled for the
  // swizzled wave tiles).
  uint32_t family(KernelId id, uint32_t n_qubits) const;
  size_t n_kernels() const;
  void *stream(); // created on first use
  const void *plan_handle(KernelId id) const;

  // Autotuning controls.
  void set_autotune(bool on, uint32_t min_qubits = 22, uint32_t samples = 2);
  bool autotune() const { return autotune_; }
  // Apply `ids` to `sv` repeatedly until every (id, n) is tuned (at most
  // max_rounds passes).  Modifies sv; returns the passes run.
  uint32_t tune(const std::vector<KernelId> &ids, HipStatevector &sv, uint32_t max_rounds = 16);
  std::vector<TuneEntry> tuning_report() const;

private:
  struct Pending {
    KernelId id;
    uint64_t dptr;
    uint32_t n_qubits;
    const void *plan;
    int cand; // tuning candidate index, -1 when not a tuning sample
  };
  struct TuneState {
    std::vector<uint32_t> fams;
    std::vector<void *> plans;
    std::vector<std::vector<double>> samples;
    std::vector<char> dropped;
    int chosen = -1;
  };
  void ensure_stream_locked();
  void ensure_events_locked(size_t n);
  TuneState &tune_state_locked(KernelId id, uint32_t n);

  HipKernelGenSpec spec_;
  mutable std::mutex mu_;
  std::unordered_map<KernelId, void *> plans_;
  std::unordered_map<KernelId, std::shared_ptr<const QuantumGate>> gates_;
  std::unordered_map<std::string, KernelId> dedup_;
  KernelId next_id_ = 0;
  std::vector<Pending> pending_;
  void *stream_ = nullptr;
  std::vector<void *> events_;
  bool autotune_ = true;
  uint32_t tune_min_qubits_ = 22;
  uint32_t tune_samples_ = 2;
  std::map<std::pair<KernelId, uint32_t>, TuneState> tune_;
};

// Marginal measurement probabilities over `measured` (bin bit i = qubit
// measured[i]), computed on the device in f64 (cast_hip_marginal_probabilities;
// the reference downloads and accumulates, src/simulator/mod.rs:282-291).
std::vector<double> marginal_probabilities(const HipStatevector &sv,
                                           const std::vector<uint32_t> &measured);

// src/simulator/mod.rs:418-432.
enum class Representation { StateVector, DensityMatrix };

// Final state of Simulator::simulate (src/simulator/mod.rs:320-414): a pure
// state on n qubits or a density matrix held as a 2n-qubit vectorised state
// (rho[i][j] at index i | (j << n)).
class QuantumState {
public:
  QuantumState(HipStatevector sv, uint32_t n_physical, Representation repr)
      : sv_(std::move(sv)), n_(n_physical), repr_(repr) {}
  uint32_t n_qubits() const { return n_; }
  Representation representation() const { return repr_; }
  bool is_pure() const { return repr_ == Representation::StateVector; }
  const HipStatevector &statevector() const { return sv_; }
  // |a_i|^2 (pure) or Re rho[i][i] (density matrix), from the device.
  std::vector<double> populations() const;
  // Squared norm (pure) or Tr rho.
  double trace() const;
  std::vector<double> download_amplitudes() const { return sv_.download(); }

private:
  HipStatevector sv_;
  uint32_t n_;
  Representation repr_;
};

// Ensemble trajectory sampling (src/simulator/trajectory.rs:17-60).
struct TrajectoryOpts {
  std::vector<uint32_t> measured_qubits;
  uint64_t n_samples = 0;
  bool has_seed = false;   // false: seeded from std::random_device
  uint64_t seed = 0;
  size_t max_ensemble = 1; // 0 is treated as 1
};

struct ExploredBranch {
  double weight = 0.0;
  std::vector<size_t> noise_path;
  uint64_t n_samples = 0;
};

struct TrajectoryResult {
  std::map<uint64_t, uint64_t> histogram;
  uint64_t n_samples = 0;
  std::vector<ExploredBranch> branches; // ensemble order: weight descending
  double explored_weight = 0.0;
};

// Inverse-CDF batch sampling (src/simulator/measure.rs:45-78): outcome =
// first index whose running sum exceeds u ~ U[0, 1), clamped to the last.
std::map<uint64_t, uint64_t> batch_sample(const std::vector<double> &probs, uint64_t n_samples,
                                          Rng &rng);

// Simulator<Hip>: state-vector simulation of a circuit graph.
class Simulator {
public:
  explicit Simulator(HipKernelGenSpec spec) : spec_(spec), mgr_(spec) {}
  static Simulator f32() { return Simulator(HipKernelGenSpec::f32()); }
  static Simulator f64() { return Simulator(HipKernelGenSpec::f64()); }

  HipKernelManager &manager() { return mgr_; }
  const HipKernelGenSpec &spec() const { return spec_; }

  // Gates of the graph in row order; noisy gates are rejected
  // (simulator/mod.rs:586-603).
  std::vector<KernelId> compile(const CircuitGraph &graph);
  // Applies every gate of `graph` to `sv` and waits.
  SyncStats run(const CircuitGraph &graph, HipStatevector &sv);
  // New |0...0> statevector of graph.n_qubits(), simulated.
  HipStatevector simulate(const CircuitGraph &graph);
  // src/simulator/mod.rs:523-547 with prepare_gates (:586-618): for
  // DensityMatrix every gate is lifted to its superoperator on 2n qubits.
  QuantumState simulate(const CircuitGraph &graph, Representation repr);

  // Deterministic top-M ensemble over noise branches, then measurement
  // samples from each surviving member (src/simulator/trajectory.rs:95-312).
  TrajectoryResult sample_trajectory(const CircuitGraph &graph, const TrajectoryOpts &opts);

  struct BenchResult {
    double compile_s = 0.0;
    std::vector<double> samples_s; // per exec: sum of kernel GPU times
    TimingStats exec;
    size_t n_gates = 0;
  };
  // src/simulator/bench.rs:136-167: compile once, then time repeated
  // executions (reset + every gate); DensityMatrix lifts the gates first.
  BenchResult bench(const CircuitGraph &graph, double budget_s,
                    Representation repr = Representation::StateVector);

private:
  HipKernelGenSpec spec_;
  HipKernelManager mgr_;
};

} // namespace cast::hip
