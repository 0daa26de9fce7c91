// Continuous-batching serving loop for circuit simulation on one GPU.
//
// The reference runs one circuit at a time (src/simulator/mod.rs:523-547:
// compile, allocate, apply every kernel, synchronise).  Here requests arrive
// over time (seeded Poisson process), are compiled on a pool of host threads
// as they arrive (circuit -> fusion -> gate plans), and are admitted to the
// GPU while earlier requests are still running, as soon as a stream slot is
// free and their statevector fits the device-memory budget.  Each admitted
// request owns one non-blocking stream and a resident statevector for its
// lifetime: |0> fill, every fused gate, and a deterministic f64 reduction of
// the norm and <Z_i Z_i+1> are queued back to back, followed by an event
// that the scheduler polls.  Statevector blocks are recycled by size class,
// so steady-state serving never calls the device allocator or its free
// (which would wait for the whole device).
//
// Every request's result depends only on its own circuit: the kernels are
// deterministic, so it is the same alone and in any mix.
#pragma once

#include "hip_backend.hpp"

#include <cstdint>
#include <memory>
#include <vector>

namespace cast::hip {

struct ServeRequest {
  uint32_t id = 0;
  uint32_t n_qubits = 0;
  uint32_t depth = 0;
  uint64_t seed = 0; // random_layer_circuit(n_qubits, depth, seed) when circuit is null
  cast_precision_t precision = CAST_PRECISION_F64;
  double arrival_s = 0.0; // offset from the start of the run
  // The circuit to serve (any unitary circuit, e.g. parsed from OpenQASM);
  // null = the seeded random-layer circuit above.  n_qubits must equal
  // circuit->n_qubits().
  std::shared_ptr<const QuantumCircuit> circuit;
};

// A request that serves `circuit` (n_qubits taken from it).
ServeRequest circuit_request(uint32_t id, std::shared_ptr<const QuantumCircuit> circuit,
                             cast_precision_t precision, double arrival_s);

// `count` requests with exponential inter-arrival gaps of mean 1/rate,
// n ~ U{n_min..n_max}, depth ~ U{depth_min..depth_max}, f32 with
// probability f32_fraction, all drawn from one generator seeded by `seed`.
std::vector<ServeRequest> poisson_requests(uint32_t count, double rate, uint32_t n_min,
                                           uint32_t n_max, uint32_t depth_min,
                                           uint32_t depth_max, double f32_fraction,
                                           uint64_t seed);

struct ServeConfig {
  uint32_t streams = 4;          // concurrent requests on the device (one per
                                 // hardware queue of the MI355X process)
  uint32_t compile_threads = 8;  // host threads compiling arrived requests
  uint32_t fusion_max = 6;       // size-only fusion bound (1..6); k = 6 runs on
                                 // matrix cores and fewer passes serve more
                                 // circuits per second than k <= 5
  uint64_t budget_bytes = 0;     // statevector memory; 0 = free memory - reserve
  uint64_t reserve_bytes = uint64_t(8) << 30;
  bool realtime = true;          // release requests at their arrival times
  bool keep_states = false;      // copy every final statevector to the host
  // Checking mode: every F32 request also runs its fused circuit at F64 on a
  // second resident statevector, queued on the same stream, and the device
  // computes the fidelity |<f32|f64>|^2 / (<f32|f32><f64|f64>) in f64.  It
  // doubles the memory and adds the F64 work of each F32 request, so it is
  // for accuracy runs, not for latency figures.
  bool f64_shadow = false;
  // Batch lane: requests of at most batch_max_qubits qubits (0 = off) are
  // admitted up to batch_size (1..64) at a time into one stream slot, and
  // each fused-gate step of the batch is one launch per (precision, gate
  // size) through cast_hip_gate_apply_batch instead of one launch per gate
  // per request.  Such a request always runs this way (alone too), so its
  // result does not depend on the mix.
  uint32_t batch_max_qubits = 18;
  uint32_t batch_size = 32;
};

struct ServeResult {
  uint32_t id = 0, n_qubits = 0, depth = 0, n_gates = 0;
  cast_precision_t precision = CAST_PRECISION_F64;
  double arrival_s = 0.0, compiled_s = 0.0, admit_s = 0.0, done_s = 0.0;
  double norm_sq = 0.0;
  std::vector<double> zz;    // <Z_i Z_i+1> sums, i = 0 .. n-2 (unnormalised)
  double fidelity_vs_f64 = -1.0; // f64_shadow runs, F32 requests; -1 otherwise
  std::vector<double> state; // interleaved f64, when keep_states
};

struct ServeStats {
  double wall_s = 0.0;
  double throughput = 0.0;       // completed requests / wall
  double p50_s = 0.0, p99_s = 0.0, mean_s = 0.0, max_s = 0.0; // arrival -> result
  double queue_p50_s = 0.0, queue_p99_s = 0.0;                 // arrival -> admission
  uint32_t max_in_flight = 0;
  uint64_t peak_sv_bytes = 0;
  uint64_t budget_bytes = 0;
  uint64_t sv_allocs = 0, sv_reuses = 0;
  uint64_t batches = 0, batched_requests = 0, batch_launches = 0; // batch lane
};

class Server {
public:
  explicit Server(ServeConfig cfg);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  // Serves every request (in arrival order); returns results by request id
  // order of `requests` and the latency statistics.
  std::vector<ServeResult> run(const std::vector<ServeRequest> &requests, ServeStats *stats);

private:
  struct Slot;
  struct Pool;
  ServeConfig cfg_;
  std::vector<Slot *> slots_;
  Pool *pool_ = nullptr;
};

} // namespace cast::hip
