// Cost models, the roofline hardware profile and the fusion configuration.
//
// Follows src/cost_model.rs:
//   SizeOnlyCostModel        :29-44   (1e-10 if it fits, else 1.0)
//   HardwareProfile          :128-225 (peak BW GiB/s, peak GFLOP/s, bw_slope,
//                                      crossover = peak / bw_slope, raw sweep)
//   compute_bw_slope         :157-159 (BW * 2^30 / (2 * scalar_bytes) / 1e9)
//   HardwareAdaptiveCostModel:263-275 (max(1, AI / crossover), inf > max_size)
//   FusionConfig             :286-369, FusionDecision/FusionLog :377-451
//
// The profile's "GFLOP/s" keeps the reference's unit (2 * AI * 2^n / t,
// src/profile.rs:436), which counts one complex multiply-add per nonzero
// scalar pair — half the true real-flop count.
#pragma once

#include "types.hpp"

#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace cast {

class CostModel {
public:
  virtual ~CostModel() = default;
  virtual double cost_of(const QuantumGate &g) const = 0;
};

class SizeOnlyCostModel : public CostModel {
public:
  SizeOnlyCostModel(size_t max_size, size_t max_ai, double zero_tol, bool trajectory_mode)
      : max_size(max_size), max_ai(max_ai), zero_tol(zero_tol), trajectory_mode(trajectory_mode) {}
  double cost_of(const QuantumGate &g) const override;
  size_t max_size;
  size_t max_ai;
  double zero_tol;
  bool trajectory_mode;
};

struct SweepEntry {
  double ai = 0.0;
  double gflops_s = 0.0;
  double gib_s = 0.0;
};

struct ProfileConfig {
  std::string backend = "hip"; // "cpu", "cuda" or "hip"
  std::string name;            // device name
  std::string arch;            // e.g. gfx950 (hip)
  uint32_t n_threads = 0;      // cpu only
  uint32_t sm_major = 0, sm_minor = 0; // cuda only
  std::string precision = "f64";
  uint32_t n_qubits = 0;
};

struct HardwareProfile {
  ProfileConfig config;
  double peak_bw_gib_s = 0.0;
  double peak_gflops_s = 0.0;
  double bw_slope = 0.0;
  double crossover_ai = std::numeric_limits<double>::infinity();
  std::vector<SweepEntry> raw;

  static HardwareProfile from_roofline(ProfileConfig c, double peak_bw_gib_s,
                                       double peak_gflops_s, size_t scalar_bytes);
  static HardwareProfile from_measurements(ProfileConfig c, double peak_bw_gib_s,
                                           double peak_gflops_s, double bw_slope);
  std::string to_json() const;
  static HardwareProfile from_json(const std::string &text);
  void save(const std::string &path) const;
  static HardwareProfile load(const std::string &path);
};

double compute_bw_slope(double peak_bw_gib_s, size_t scalar_bytes);
double compute_crossover(double peak_gflops_s, double bw_slope);

class HardwareAdaptiveCostModel : public CostModel {
public:
  HardwareAdaptiveCostModel(const HardwareProfile &p, size_t max_size, bool trajectory_mode,
                            double zero_tol)
      : crossover_ai(p.crossover_ai), max_size(max_size), zero_tol(zero_tol),
        trajectory_mode(trajectory_mode) {}
  double cost_of(const QuantumGate &g) const override;
  double crossover_ai;
  size_t max_size;
  double zero_tol;
  bool trajectory_mode;
};

struct FusionConfig {
  size_t size_max = 6;
  double benefit_margin = 0.0;
  std::shared_ptr<const CostModel> cost_model;
  bool trajectory_mode = false;
  double zero_tol = 0.0;

  static FusionConfig size_only(size_t max_size);
  static FusionConfig size_only_trajectory(size_t max_size);
  static FusionConfig aggressive() { return size_only(6); }
  static FusionConfig hardware_adaptive(const HardwareProfile &p, size_t max_size,
                                        double zero_tol);
  static FusionConfig hardware_adaptive_trajectory(const HardwareProfile &p, size_t max_size,
                                                   double zero_tol);
};

struct FusionDecision {
  size_t cdd_size = 0;
  size_t n_candidates = 0;
  std::vector<std::pair<size_t, double>> candidates; // (effective qubits, AI)
  size_t product_n_qubits = 0;
  double product_ai = 0.0;
  double old_cost = 0.0;
  double new_cost = 0.0;
  double benefit = 0.0;
  bool accepted = false;
};

struct FusionLog {
  std::vector<FusionDecision> decisions;
  size_t n_accepted() const;
  size_t n_rejected() const;
  std::string summary() const;
};

} // namespace cast
