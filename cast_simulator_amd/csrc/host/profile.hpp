// Adaptive timing (src/timing.rs) and the roofline profiler (src/profile.rs).
//
// time_adaptive_samples_with: one probe, ~25% of the budget as warmup, the
// rest timed; a probe over 1.5x the budget is returned as the only sample
// (timing.rs:88-123).
//
// Profiler: probes at AI = 1, 2, 4, ..., 64, then up to 5 refinement rounds
// of 3 probes at {0.5, 1, 1.5} x the estimated crossover, until R^2 >= 0.95
// and changes by < 0.01 (profile.rs:229-344).  fit_roofline takes the median
// gflops/ai of the lower half of the sweep as bw_slope and the max gflops as
// peak (:137-156).  Throughputs of a probe: gib_s = 2 * sv_bytes / t / 2^30,
// gflops = AI * 2^n * 2 / t / 1e9 (:427-437).
#pragma once

#include "cost_model.hpp"

#include <functional>
#include <vector>

namespace cast {

struct TimingStats {
  size_t n_iters = 0;
  double mean_s = 0, stddev_s = 0, cv = 0, min_s = 0, max_s = 0;
};

TimingStats stats_from_samples(const std::vector<double> &samples);
// f returns the duration (seconds) it wants recorded.
std::vector<double> time_adaptive_samples_with(const std::function<double()> &f,
                                               double budget_s);

struct RooflineParams {
  double crossover_ai = 0.0;
  double bw_slope = 0.0;
  double peak_gflops() const { return bw_slope * crossover_ai; }
  double predicted_gflops(double ai) const;
};

RooflineParams fit_roofline(std::vector<SweepEntry> sweep); // sweep sorted by ai
double roofline_r2(const std::vector<SweepEntry> &sweep, const RooflineParams &p);

// probe(target_ai, per_point_budget_s) -> measured entry.
HardwareProfile adaptive_sweep(const std::function<SweepEntry(double, double)> &probe,
                               ProfileConfig config, double budget_s);
HardwareProfile fit_from_samples(ProfileConfig config, std::vector<SweepEntry> samples);

// time_kernel(gate, budget_s) -> mean execution time in seconds.
HardwareProfile measure(ProfileConfig config, uint32_t n_qubits, size_t scalar_bytes,
                        double budget_s, uint64_t seed,
                        const std::function<double(const QuantumGate &, double)> &time_kernel);

} // namespace cast
