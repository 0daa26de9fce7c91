// Adaptive timing and roofline profiler; see profile.hpp.
#include "profile.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

namespace cast {

TimingStats stats_from_samples(const std::vector<double> &s) {
  TimingStats t;
  if (s.empty())
    return t;
  const double n = double(s.size());
  double mean = 0;
  for (double x : s)
    mean += x;
  mean /= n;
  double var = 0;
  for (double x : s)
    var += (x - mean) * (x - mean);
  var /= n;
  t.n_iters = s.size();
  t.mean_s = mean;
  t.stddev_s = std::sqrt(var);
  t.cv = t.stddev_s / mean;
  t.min_s = *std::min_element(s.begin(), s.end());
  t.max_s = *std::max_element(s.begin(), s.end());
  return t;
}

std::vector<double> time_adaptive_samples_with(const std::function<double()> &f,
                                               double budget_s) {
  using clk = std::chrono::steady_clock;
  constexpr double kWarmupFraction = 0.25, kOverBudget = 1.5;
  const auto wall_start = clk::now();
  const auto before = clk::now();
  const double probe = f();
  const double probe_wall = std::chrono::duration<double>(clk::now() - before).count();
  const double est = std::max(probe_wall, 1e-9);
  if (probe_wall > budget_s * kOverBudget)
    return {probe};
  const double remaining_warmup = std::max(0.0, budget_s * kWarmupFraction - probe_wall);
  const auto n_warm = uint64_t(remaining_warmup / est);
  for (uint64_t i = 0; i < n_warm; ++i)
    f();
  const double elapsed = std::chrono::duration<double>(clk::now() - wall_start).count();
  const double remaining = std::max(0.0, budget_s - elapsed);
  const uint64_t n_timed = std::max<uint64_t>(1, uint64_t(remaining / est));
  std::vector<double> samples;
  samples.reserve(n_timed);
  for (uint64_t i = 0; i < n_timed; ++i)
    samples.push_back(f());
  return samples;
}

double RooflineParams::predicted_gflops(double ai) const {
  return std::min(bw_slope * ai, peak_gflops());
}

RooflineParams fit_roofline(std::vector<SweepEntry> sweep) {
  if (sweep.empty())
    throw Error("fit_roofline: empty sweep");
  const size_t n = sweep.size();
  const size_t lower = (n + 1) / 2;
  std::vector<double> ratios;
  for (size_t i = 0; i < lower; ++i)
    ratios.push_back(sweep[i].gflops_s / sweep[i].ai);
  std::sort(ratios.begin(), ratios.end());
  RooflineParams p;
  p.bw_slope = ratios[ratios.size() / 2];
  double peak = 0.0;
  for (const auto &e : sweep)
    peak = std::max(peak, e.gflops_s);
  p.crossover_ai = p.bw_slope > 0.0 ? peak / p.bw_slope : INFINITY;
  return p;
}

double roofline_r2(const std::vector<SweepEntry> &sweep, const RooflineParams &p) {
  double mean = 0;
  for (const auto &e : sweep)
    mean += e.gflops_s;
  mean /= double(sweep.size());
  double ss_tot = 0, ss_res = 0;
  for (const auto &e : sweep) {
    ss_tot += (e.gflops_s - mean) * (e.gflops_s - mean);
    const double d = e.gflops_s - p.predicted_gflops(e.ai);
    ss_res += d * d;
  }
  if (ss_tot < 1e-30)
    return 1.0;
  return 1.0 - ss_res / ss_tot;
}

namespace {
constexpr uint32_t kMaxGateQubits = 6;
constexpr double kR2Stable = 0.95, kR2Delta = 0.01, kR2Warn = 0.90;
constexpr uint32_t kMaxRefineRounds = 5;
constexpr size_t kMaxCandidates = 3;

void sort_by_ai(std::vector<SweepEntry> &s) {
  std::sort(s.begin(), s.end(), [](const SweepEntry &a, const SweepEntry &b) { return a.ai < b.ai; });
}

HardwareProfile finish(ProfileConfig config, const std::vector<SweepEntry> &sweep,
                       const char *label) {
  const auto params = fit_roofline(sweep);
  const double r2 = roofline_r2(sweep, params);
  if (r2 < kR2Warn)
    std::fprintf(stderr,
                 "  %s: Warning: roofline fit R^2 = %.3f (< %.2f); crossover estimate may be "
                 "unreliable.\n",
                 label, r2, kR2Warn);
  double bw = 0, gf = 0;
  for (const auto &e : sweep) {
    bw = std::max(bw, e.gib_s);
    gf = std::max(gf, e.gflops_s);
  }
  auto p = HardwareProfile::from_measurements(std::move(config), bw, gf, params.bw_slope);
  p.raw = sweep;
  return p;
}
} // namespace

HardwareProfile adaptive_sweep(const std::function<SweepEntry(double, double)> &probe,
                               ProfileConfig config, double budget_s) {
  using clk = std::chrono::steady_clock;
  const size_t max_points = (kMaxGateQubits + 1) + kMaxRefineRounds * kMaxCandidates;
  const double per_point = budget_s / double(max_points);
  const auto t0 = clk::now();
  std::vector<SweepEntry> sweep;
  for (uint32_t e = 0; e <= kMaxGateQubits; ++e)
    sweep.push_back(probe(double(uint64_t(1) << e), per_point));
  sort_by_ai(sweep);
  double prev_r2 = 0.0;
  for (uint32_t round = 0; round < kMaxRefineRounds; ++round) {
    const auto params = fit_roofline(sweep);
    const double r2 = roofline_r2(sweep, params);
    if (r2 >= kR2Stable && std::fabs(r2 - prev_r2) < kR2Delta)
      break;
    prev_r2 = r2;
    if (std::chrono::duration<double>(clk::now() - t0).count() >= budget_s)
      break;
    const double cx = params.crossover_ai;
    std::vector<double> cands;
    for (double f : {0.5, 1.0, 1.5}) {
      const double ai = std::max(1.0, std::round(cx * f));
      bool dup = false;
      for (const auto &e : sweep)
        dup = dup || std::fabs(e.ai - ai) < 0.5;
      if (!dup && std::isfinite(ai) && cands.size() < kMaxCandidates)
        cands.push_back(ai);
    }
    if (cands.empty())
      break;
    for (double ai : cands)
      sweep.push_back(probe(ai, per_point));
    sort_by_ai(sweep);
  }
  return finish(std::move(config), sweep, "adaptive_sweep");
}

HardwareProfile fit_from_samples(ProfileConfig config, std::vector<SweepEntry> samples) {
  if (samples.empty())
    throw Error("fit_from_samples: need at least one SweepEntry");
  sort_by_ai(samples);
  return finish(std::move(config), samples, "fit_from_samples");
}

HardwareProfile measure(ProfileConfig config, uint32_t n_qubits, size_t scalar_bytes,
                        double budget_s, uint64_t seed,
                        const std::function<double(const QuantumGate &, double)> &time_kernel) {
  const double ztol = 1e-12;
  const double n_amp = double(uint64_t(1) << n_qubits);
  const double sv_bytes = n_amp * 2.0 * double(scalar_bytes);
  Rng rng(seed);
  return adaptive_sweep(
      [&](double target_ai, double budget) {
        // Refinement probes can ask for more than a dense 6-qubit gate
        // reaches (AI 2 * 2^6); the backend's kernels stop at k = 6.
        target_ai = std::min(target_ai, 128.0);
        const auto g = QuantumGate::random_arithmetic_intensity(n_qubits, target_ai, rng);
        const double ai = g.arithmetic_intensity(ztol);
        const double mean_s = time_kernel(g, budget);
        SweepEntry e;
        e.ai = ai;
        e.gib_s = 2.0 * sv_bytes / mean_s / double(uint64_t(1) << 30);
        e.gflops_s = ai * n_amp * 2.0 / mean_s / 1e9;
        return e;
      },
      std::move(config), budget_s);
}

} // namespace cast
