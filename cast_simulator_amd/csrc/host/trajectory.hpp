// Batched noisy trajectories on one GPU in one persistent launch.
//
// The reference simulates noise by deterministic branching on the host
// (src/simulator/trajectory.rs:180-265: top-M Kraus branches, each branch a
// cloned device statevector).  This engine instead samples stochastic
// trajectories of a layered circuit.  Two noise sources, both drawn from a
// per-trajectory seeded generator:
//  * a gate's own noise model (src/types/noise.rs:12-33: branches (p_b, K_b)
//    with full Kraus operators K_b = V_b U, src/types/gate.rs:129-151): the
//    trajectory applies branch b with probability p_b, the reference's
//    state-independent branch weights (trajectory.rs:214-219);
//  * optionally, a depolarizing error after every gate on each qubit it
//    touches with probability p: X, Y or Z with p/3 each (the Kraus set of
//    src/types/gate.rs:174-197).
// Each trajectory runs the first `depth` layers.  Branches need not be
// unitary: a trajectory's state is then not normalised, and its squared norm
// (f64, on the device) divides its expectations.
//
// Work sharing.  Layers are grouped into chunks of B.  The clean state (every
// noisy gate on its most likely branch, no depolarizing error) after every
// chunk boundary is computed once (a checkpoint).  A trajectory becomes a
// task that starts at the checkpoint before its first deviation and applies
// the fused gates of the remaining chunks; clean trajectories of equal depth
// share one task.  Each chunk is fused once (fusion with provenance: every
// fused gate knows its source gates in application order); a deviation in
// component i of fused gate F = C_{m-1} ... C_0 replaces C_i by the
// trajectory's operator (the sampled branch, then its Paulis) — the fused
// structure never changes, only the matrices of the fused gates that
// contain deviations.
//
// All tasks run in one cast_hip_traj_run launch: one workgroup per slot,
// tasks taken from an atomic counter longest first.  Each task returns its
// f64 squared norm and <Z_a Z_b> sums.
#pragma once

#include "hip_backend.hpp"

#include <map>
#include <tuple>

namespace cast::hip {

// One sampled deviation of a trajectory from the clean circuit.
//  pauli 1..3: the depolarizing model's X, Y or Z on `qubit` right after
//              source gate `gate`;
//  pauli 0:    source gate `gate` (a noisy gate) took branch `branch` of its
//              noise model instead of its most likely one.
// Events of one gate: its branch event first, then its Paulis.
struct NoiseEvent {
  uint32_t gate = 0;
  uint32_t qubit = 0;
  uint8_t pauli = 0;
  uint32_t branch = 0;
};

struct TrajectorySample {
  uint32_t depth = 0;             // layers applied
  std::vector<NoiseEvent> errors; // in circuit order
};

// A circuit as a list of layers: layer l holds gates
// [layer_start[l], layer_start[l + 1]).
struct LayeredCircuit {
  uint32_t n_qubits = 0;
  std::vector<GatePtr> gates;
  std::vector<uint32_t> layer_start{0};
  uint32_t n_layers() const { return uint32_t(layer_start.size() - 1); }
  // random_layer_circuit(n, depth, seed), split into its layers (n Haar
  // SU(2) gates then the CZ gates of one matching per layer).
  static LayeredCircuit random_layers(uint32_t n, uint32_t depth, uint64_t seed);
  // Any circuit (unitary or noisy gates on at most 4 qubits), in moments:
  // each gate joins the layer after the last one that touches any of its
  // qubits, and gates are listed layer by layer (gates of one layer act on
  // disjoint qubits, so the reordering leaves the circuit's operator as is).
  static LayeredCircuit from_circuit(const QuantumCircuit &qc);
};

// The branch a trajectory takes when nothing happens: the first one of
// highest probability.
uint32_t clean_branch(const NoiseModel &m);

struct NoiseSpec {
  double p = 1e-3; // depolarizing probability per touched qubit after every gate (0 = none)
  uint32_t min_depth = 1;
  uint32_t max_depth = 1;
  uint64_t seed = 0;
};

// Trajectories first .. first + count - 1; trajectory t draws only from its
// own generator (seeded by (seed, t)), so any subset samples identically.
std::vector<TrajectorySample> sample_trajectories(const LayeredCircuit &c, const NoiseSpec &ns,
                                                  uint32_t count, uint32_t first = 0);

// The explicit gate list of one trajectory: the source gates of its first
// `depth` layers, noisy ones replaced by the branch the trajectory took, with
// each sampled Pauli inserted right after its gate.
QuantumCircuit trajectory_circuit(const LayeredCircuit &c, const TrajectorySample &s);

struct TrajectoryTuning {
  uint32_t chunk_layers = 4; // B
  uint32_t fusion_max = 4;   // largest fused gate, 1..4
  uint32_t block = 256;      // threads per workgroup: 64, 128, 256 or 512
  uint32_t slots_per_cu = 2; // resident statevector slots per compute unit
  uint32_t max_slots = 0;    // 0 = no cap beyond memory and slots_per_cu
  uint32_t waves_per_simd = 2; // kernel build: 2 (no spills) or 4 (two workgroups per CU)
  uint32_t mfma_f32 = 0;       // f32 k = 5, 6 gates: 0 = bf16 split products, 1 = f32 products
};

struct TrajectoryRun {
  uint32_t n_pairs = 0;
  std::vector<uint32_t> task_of;  // per trajectory
  std::vector<uint32_t> task_ckpt_layer; // per task: layer of its start checkpoint
  std::vector<uint32_t> task_gates;      // per task: fused gates applied
  std::vector<double> norm_sq;    // per task, f64
  std::vector<double> zz_sum;     // per task x pair: sum |amp|^2 (-1)^(a xor b), f64
  std::vector<double> traj_zz;    // per trajectory: (1/P) sum_j zz_j / norm
  // Expectation over the batch: sum_t (1/P) sum_j zz_j / sum_t norm_t, in
  // f64.  Branch probabilities do not depend on the state, so the average of
  // the unnormalised trajectory states is the channel's output state; this
  // ratio equals the mean of traj_zz when every branch is unitary.
  double mean_zz = 0.0;
  double min_norm_sq = 0.0, max_norm_sq = 0.0;
  double plan_s = 0.0, checkpoint_s = 0.0, gpu_s = 0.0, total_s = 0.0;
  uint64_t n_tasks = 0, n_gate_passes = 0, n_shared_clean = 0, n_dirty_matrices = 0;
  uint32_t n_slots = 0, block = 0, chunk_layers = 0, fusion_max = 0, waves_per_simd = 0;
  uint32_t mfma_f32 = 0;
  std::vector<double> dump; // final states per task (interleaved f64), if asked for
};

// Host-side plan of one batch: the task table, the gate table and the matrix
// pool (complex entries, f64) that cast_hip_traj_run consumes.
struct TrajectoryPlan {
  std::vector<cast_hip_traj_task_t> tasks;
  std::vector<cast_hip_traj_gate_t> gates;
  std::vector<double> pool;      // k <= 3: interleaved complex K x K; k = 4: Re, Im, Re + Im
  std::vector<uint32_t> order;   // visit order: longest task first
  std::vector<uint32_t> task_of; // per trajectory
  uint64_t n_shared_clean = 0, n_dirty_matrices = 0;
  double plan_s = 0.0;
};

struct TrajectoryTuneRecord {
  TrajectoryTuning tuning;
  double s_per_trajectory = 0.0;
};

class TrajectoryEngine {
public:
  // pairs: flat (a0, b0, a1, b1, ...), at most 32 pairs.
  TrajectoryEngine(LayeredCircuit circuit, cast_precision_t precision,
                   std::vector<uint32_t> pairs);
  ~TrajectoryEngine();
  TrajectoryEngine(const TrajectoryEngine &) = delete;
  TrajectoryEngine &operator=(const TrajectoryEngine &) = delete;

  const LayeredCircuit &circuit() const { return circ_; }
  // Start state of every trajectory (interleaved f64, 2 * 2^n); default |0>.
  void set_initial_state(std::vector<double> interleaved);

  // Host only: tasks, fused gate lists and matrices for a batch.  Task k
  // starts from the clean state after tasks[k].checkpoint * chunk_layers
  // layers.
  TrajectoryPlan plan(const std::vector<TrajectorySample> &samples, const TrajectoryTuning &t);

  TrajectoryRun run(const std::vector<TrajectorySample> &samples, const TrajectoryTuning &t,
                    bool dump = false);

  // Project-written autotuner: coordinate search over the launch shape
  // (block, slots per CU) and then the planning knobs (chunk layers, fusion
  // size), each candidate timed end to end (plan + checkpoints + launch) on
  // `sample`.  Returns the best tuning; `log` receives every trial.
  TrajectoryTuning autotune(const std::vector<TrajectorySample> &sample,
                            const TrajectoryTuning &start,
                            std::vector<TrajectoryTuneRecord> *log = nullptr);

private:
  struct FusedGate {
    std::vector<uint32_t> qubits;
    ComplexSquareMatrix m;
    std::vector<uint32_t> origin;                      // source gates, application order
    mutable std::vector<ComplexSquareMatrix> prefix;   // products of origin[0..i]
    mutable std::vector<ComplexSquareMatrix> suffix;   // products of origin[i..m), from the end
  };
  struct Segment {
    std::vector<FusedGate> gates;
    std::map<uint32_t, std::pair<uint32_t, uint32_t>> where; // source gate -> (fused, pos)
  };
  struct DevBuf {
    uint64_t ptr = 0;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
  };

  const Segment &segment(uint32_t kmax, uint32_t first_layer, uint32_t end_layer);
  const ComplexSquareMatrix &prefix(const FusedGate &f, uint32_t pos) const;
  const ComplexSquareMatrix &suffix(const FusedGate &f, uint32_t pos) const;
  ComplexSquareMatrix component(const FusedGate &f, uint32_t pos) const;
  ComplexSquareMatrix dirty_matrix(const FusedGate &f,
                                   std::vector<std::pair<uint32_t, NoiseEvent>> errs) const;
  const ComplexSquareMatrix &clean_matrix(uint32_t g) const { return clean_[g]; }
  void ensure_checkpoints(uint32_t chunk_layers);
  size_t sv_bytes() const;

  LayeredCircuit circ_;
  std::vector<ComplexSquareMatrix> clean_; // per source gate: its clean-branch operator
  cast_precision_t prec_;
  std::vector<uint32_t> pairs_;
  std::vector<double> initial_;
  std::map<std::tuple<uint32_t, uint32_t, uint32_t>, Segment> segs_;
  HipKernelManager mgr_;
  DevBuf ckpt_, slots_, pool_, gates_, tasks_, order_, out_, dump_, work_;
  uint32_t ckpt_chunk_ = 0; // chunk size the checkpoints were built for
  uint32_t n_ckpt_ = 0;
  double last_checkpoint_s_ = 0.0;
};

} // namespace cast::hip
