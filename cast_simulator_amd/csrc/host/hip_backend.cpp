// Host-side gfx950 backend over the C ABI; see hip_backend.hpp.
#include "hip_backend.hpp"

#include <chrono>
#include <cmath>
#include <cstring>

namespace cast::hip {

namespace {
constexpr size_t kErr = 1024; // src/cuda/types.rs:72
}

void throw_if(int rc, const char *err) {
  if (rc != 0)
    throw Error(err && err[0] ? err : "cast_hip call failed");
}

// ── HipStatevector ─────────────────────────────────────────────────────────

HipStatevector::HipStatevector(uint32_t n_qubits, cast_precision_t precision)
    : n_(n_qubits), prec_(precision), n_elements_(size_t(2) << n_qubits) {
  if (n_qubits > 40)
    throw Error("statevector qubit count out of range");
  char err[kErr];
  dptr_ = cast_hip_sv_alloc(n_elements_, uint8_t(prec_), err, kErr);
  if (dptr_ == 0)
    throw Error(err);
}

HipStatevector::~HipStatevector() {
  if (dptr_)
    cast_hip_sv_free(dptr_);
}

HipStatevector::HipStatevector(HipStatevector &&o) noexcept
    : dptr_(o.dptr_), n_(o.n_), prec_(o.prec_), n_elements_(o.n_elements_) {
  o.dptr_ = 0;
}

void HipStatevector::zero() {
  char err[kErr];
  throw_if(cast_hip_sv_zero(dptr_, n_elements_, uint8_t(prec_), err, kErr), err);
}

void HipStatevector::upload(const std::vector<double> &v) {
  if (v.size() != n_elements_)
    throw Error("upload: expected " + std::to_string(len()) + " amplitudes, got " +
                std::to_string(v.size() / 2));
  char err[kErr];
  throw_if(cast_hip_sv_upload(dptr_, v.data(), n_elements_, uint8_t(prec_), err, kErr), err);
}

std::vector<double> HipStatevector::download() const {
  std::vector<double> v(n_elements_);
  char err[kErr];
  throw_if(cast_hip_sv_download(dptr_, v.data(), n_elements_, uint8_t(prec_), err, kErr), err);
  return v;
}

double HipStatevector::norm_squared() const {
  double out = 0.0;
  char err[kErr];
  throw_if(cast_hip_norm_squared(dptr_, n_elements_, uint8_t(prec_), &out, err, kErr), err);
  return out;
}

void HipStatevector::normalize() {
  const double ns = norm_squared();
  if (ns <= 0.0)
    throw Error("cannot normalize zero-norm statevector");
  char err[kErr];
  throw_if(cast_hip_scale(dptr_, n_elements_, uint8_t(prec_), 1.0 / std::sqrt(ns), err, kErr),
           err);
}

HipStatevector HipStatevector::clone_device() const {
  HipStatevector copy(n_, prec_);
  char err[kErr];
  const size_t bytes = n_elements_ * (prec_ == CAST_PRECISION_F32 ? 4 : 8);
  throw_if(cast_hip_memcpy_dtod_async(copy.dptr_, dptr_, bytes, nullptr, err, kErr), err);
  return copy;
}

std::vector<double> HipStatevector::expect_zz(const std::vector<uint32_t> &pairs) const {
  if (pairs.size() % 2)
    throw Error("expect_zz: pairs must hold an even number of qubit indices");
  std::vector<double> out(pairs.size() / 2);
  char err[kErr];
  throw_if(cast_hip_expect_zz(dptr_, n_, uint8_t(prec_), pairs.data(), pairs.size() / 2,
                              out.data(), nullptr, err, kErr),
           err);
  return out;
}

cplx HipStatevector::inner(const HipStatevector &o) const {
  if (o.n_ != n_)
    throw Error("inner: statevectors have different qubit counts");
  double out[2] = {0.0, 0.0};
  char err[kErr];
  throw_if(cast_hip_inner_product(dptr_, uint8_t(prec_), o.dptr_, uint8_t(o.prec_), n_, out,
                                  nullptr, err, kErr),
           err);
  return {out[0], out[1]};
}

double HipStatevector::fidelity(const HipStatevector &o) const {
  const cplx ab = inner(o);
  const double aa = inner(*this).real(), bb = o.inner(o).real();
  return std::norm(ab) / (aa * bb);
}

// ── HipKernelManager ──────────────────────────────────────────────────────

namespace {
std::string dedup_key(const QuantumGate &g) {
  const auto &d = g.matrix().data();
  std::string key(reinterpret_cast<const char *>(d.data()), d.size() * sizeof(cplx));
  key.append(reinterpret_cast<const char *>(g.qubits().data()),
             g.qubits().size() * sizeof(uint32_t));
  return key;
}
} // namespace

HipKernelManager::HipKernelManager(HipKernelGenSpec spec) : spec_(spec) {}

HipKernelManager::~HipKernelManager() {
  std::lock_guard<std::mutex> lock(mu_);
  if (stream_) {
    char err[kErr];
    (void)cast_hip_stream_sync(stream_, err, kErr);
  }
  for (void *e : events_)
    cast_hip_event_destroy(e);
  if (stream_)
    cast_hip_stream_destroy(stream_);
  for (auto &kv : plans_)
    cast_hip_gate_plan_destroy(kv.second);
  for (auto &kv : tune_)
    for (void *p : kv.second.plans)
      cast_hip_gate_plan_destroy(p);
}

void HipKernelManager::set_autotune(bool on, uint32_t min_qubits, uint32_t samples) {
  std::lock_guard<std::mutex> lock(mu_);
  autotune_ = on;
  tune_min_qubits_ = min_qubits;
  tune_samples_ = std::max<uint32_t>(1, samples);
}

HipKernelManager::TuneState &HipKernelManager::tune_state_locked(KernelId id, uint32_t n) {
  auto key = std::make_pair(id, n);
  auto it = tune_.find(key);
  if (it != tune_.end())
    return it->second;
  TuneState st;
  const QuantumGate &g = *gates_.at(id);
  const auto &d = g.matrix().data();
  std::vector<cast_complex64_t> m(d.size());
  for (size_t i = 0; i < d.size(); ++i)
    m[i] = {d[i].real(), d[i].imag()};
  char err[kErr];
  const std::pair<uint32_t, uint32_t> cands[] = {
      {CAST_HIP_VARIANT_TILE, 0u},
      {CAST_HIP_VARIANT_TILE, CAST_HIP_FLAG_TILE_SWIZZLE},
      {CAST_HIP_VARIANT_TILE, CAST_HIP_FLAG_TILE_GAUSS},
      {CAST_HIP_VARIANT_TILE, CAST_HIP_FLAG_TILE_GAUSS | CAST_HIP_FLAG_TILE_SWIZZLE},
      {CAST_HIP_VARIANT_TILE, CAST_HIP_FLAG_TILE_XCD},
      {CAST_HIP_VARIANT_TILE, CAST_HIP_FLAG_TILE_GAUSS | CAST_HIP_FLAG_TILE_XCD},
      {CAST_HIP_VARIANT_TILE, CAST_HIP_FLAG_TILE_XCD | CAST_HIP_FLAG_TILE_SWIZZLE},
      {CAST_HIP_VARIANT_TILE,
       CAST_HIP_FLAG_TILE_GAUSS | CAST_HIP_FLAG_TILE_XCD | CAST_HIP_FLAG_TILE_SWIZZLE},
      {CAST_HIP_VARIANT_MFMA, 0u},
      {CAST_HIP_VARIANT_MFMA, CAST_HIP_FLAG_MFMA_F32},
      {CAST_HIP_VARIANT_MFMA, CAST_HIP_FLAG_TILE_XCD},
      {CAST_HIP_VARIANT_MFMA, CAST_HIP_FLAG_TILE_SWIZZLE},
      {CAST_HIP_VARIANT_MFMA, CAST_HIP_FLAG_TILE_XCD | CAST_HIP_FLAG_TILE_SWIZZLE},
      {CAST_HIP_VARIANT_DENSE, 0u}};
  for (const auto &c : cands) {
    cast_hip_kernel_gen_spec_t s{};
    s.precision = spec_.precision;
    s.ztol = spec_.ztol;
    s.otol = spec_.otol;
    s.variant = c.first;
    s.flags = c.second;
    void *plan = nullptr;
    throw_if(cast_hip_gate_plan_create(&s, m.data(), m.size(), g.qubits().data(),
                                       g.qubits().size(), &plan, err, kErr),
             err);
    uint32_t fam = 0;
    const bool ok = cast_hip_gate_plan_family(plan, n, &fam, err, kErr) == 0;
    // Candidate code: the family, or one of the wave-tile variants.
    const bool gauss = c.second & CAST_HIP_FLAG_TILE_GAUSS;
    const bool swz = c.second & CAST_HIP_FLAG_TILE_SWIZZLE;
    const bool xcd = c.second & CAST_HIP_FLAG_TILE_XCD;
    const bool f32_products = c.second & CAST_HIP_FLAG_MFMA_F32;
    uint32_t code = fam;
    if (fam == CAST_HIP_VARIANT_TILE && c.second)
      code = xcd ? (swz ? (gauss ? kTileGaussXcdSwizzled : kTileXcdSwizzled)
                        : (gauss ? kTileGaussXcd : kTileXcd))
                 : gauss ? (swz ? kTileGaussSwizzled : kTileGauss) : kTileSwizzled;
    if (fam == CAST_HIP_VARIANT_MFMA && f32_products)
      code = kMfmaF32;
    if (fam == CAST_HIP_VARIANT_MFMA && (xcd || swz))
      code = xcd ? (swz ? kMfmaXcdSwizzled : kMfmaXcd) : kMfmaSwizzled;
    const bool gauss_ok = g.n_qubits() == 4 || g.n_qubits() == 5;
    // f32 products differ from the default matrix-core kernel only at f32, k >= 5
    const bool f32_products_ok = spec_.precision == CAST_PRECISION_F32 && g.n_qubits() >= 5 &&
                                 fam == CAST_HIP_VARIANT_MFMA;
    const bool flags_ok = !c.second || (fam == CAST_HIP_VARIANT_TILE && !f32_products) ||
                          (f32_products && f32_products_ok) ||
                          (fam == CAST_HIP_VARIANT_MFMA &&
                           (c.second & ~(CAST_HIP_FLAG_TILE_XCD | CAST_HIP_FLAG_TILE_SWIZZLE)) == 0);
    if (!ok || !flags_ok || (gauss && !gauss_ok) ||
        std::find(st.fams.begin(), st.fams.end(), code) != st.fams.end()) {
      cast_hip_gate_plan_destroy(plan); // inapplicable here, or same kernel as another
      continue;
    }
    st.fams.push_back(code);
    st.plans.push_back(plan);
  }
  st.samples.resize(st.fams.size());
  st.dropped.assign(st.fams.size(), 0);
  if (st.fams.size() <= 1)
    st.chosen = 0;
  return tune_.emplace(key, std::move(st)).first->second;
}

KernelId HipKernelManager::generate(const QuantumGate &gate) {
  if (!gate.is_unitary())
    throw Error("noisy gates cannot be compiled as a single kernel");
  const std::string key = dedup_key(gate);
  {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = dedup_.find(key);
    if (it != dedup_.end())
      return it->second;
  }
  cast_hip_kernel_gen_spec_t s{};
  s.precision = spec_.precision;
  s.ztol = spec_.ztol;
  s.otol = spec_.otol;
  s.variant = spec_.variant;
  const auto &d = gate.matrix().data();
  std::vector<cast_complex64_t> m(d.size());
  for (size_t i = 0; i < d.size(); ++i)
    m[i] = {d[i].real(), d[i].imag()};
  void *plan = nullptr;
  char err[kErr];
  throw_if(cast_hip_gate_plan_create(&s, m.data(), m.size(), gate.qubits().data(),
                                     gate.qubits().size(), &plan, err, kErr),
           err);
  std::lock_guard<std::mutex> lock(mu_);
  auto it = dedup_.find(key);
  if (it != dedup_.end()) { // another thread won the race
    cast_hip_gate_plan_destroy(plan);
    return it->second;
  }
  const KernelId id = next_id_++;
  plans_[id] = plan;
  gates_[id] = std::make_shared<QuantumGate>(gate);
  dedup_[key] = id;
  return id;
}

void HipKernelManager::apply(KernelId id, HipStatevector &sv) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = plans_.find(id);
  if (it == plans_.end())
    throw Error("kernel " + std::to_string(id) + " not found");
  cast_hip_gate_plan_info_t info{};
  char err[kErr];
  throw_if(cast_hip_gate_plan_info(it->second, &info, err, kErr), err);
  if (sv.n_qubits() < info.n_qubits)
    throw Error("statevector has fewer qubits than the gate kernel requires");
  if (uint32_t(sv.precision()) != info.precision)
    throw Error("statevector precision does not match kernel precision");
  if (info.qubits[info.n_qubits - 1] >= sv.n_qubits())
    throw Error("gate acts on a qubit outside the statevector");
  const void *plan = it->second;
  int cand = -1;
  // Tuning applies to dense plans of the automatic family on large states.
  if (autotune_ && spec_.variant == CAST_HIP_VARIANT_AUTO && sv.n_qubits() >= tune_min_qubits_ &&
      info.variant != CAST_HIP_VARIANT_MONOMIAL) {
    TuneState &st = tune_state_locked(id, sv.n_qubits());
    if (st.chosen >= 0) {
      plan = st.plans[size_t(st.chosen)];
    } else {
      // the surviving candidate with the fewest timings
      size_t best = st.samples.size();
      for (size_t c = 0; c < st.samples.size(); ++c)
        if (!st.dropped[c] &&
            (best == st.samples.size() || st.samples[c].size() < st.samples[best].size()))
          best = c;
      cand = int(best);
      plan = st.plans[best];
    }
  }
  pending_.push_back({id, sv.dptr(), sv.n_qubits(), plan, cand});
}

uint32_t HipKernelManager::tune(const std::vector<KernelId> &ids, HipStatevector &sv,
                                uint32_t max_rounds) {
  uint32_t rounds = 0;
  for (; rounds < max_rounds; ++rounds) {
    bool pending_tune = false;
    {
      std::lock_guard<std::mutex> lock(mu_);
      if (!autotune_ || sv.n_qubits() < tune_min_qubits_)
        return rounds;
      for (KernelId id : ids) {
        auto it = tune_.find(std::make_pair(id, sv.n_qubits()));
        if (it == tune_.end() || it->second.chosen < 0)
          pending_tune = true;
      }
    }
    if (!pending_tune)
      break;
    for (KernelId id : ids)
      apply(id, sv);
    sync(true);
  }
  return rounds;
}

std::vector<TuneEntry> HipKernelManager::tuning_report() const {
  std::lock_guard<std::mutex> lock(mu_);
  std::vector<TuneEntry> out;
  for (const auto &kv : tune_) {
    TuneEntry e;
    e.id = kv.first.first;
    e.n_qubits = kv.first.second;
    e.chosen_family = kv.second.chosen >= 0 ? int(kv.second.fams[size_t(kv.second.chosen)]) : -1;
    for (size_t c = 0; c < kv.second.fams.size(); ++c) {
      const auto &s = kv.second.samples[c];
      e.best_s.emplace_back(kv.second.fams[c],
                            s.empty() ? -1.0 : *std::min_element(s.begin(), s.end()));
    }
    out.push_back(std::move(e));
  }
  return out;
}

void HipKernelManager::ensure_stream_locked() {
  if (stream_)
    return;
  char err[kErr];
  stream_ = cast_hip_stream_create(err, kErr);
  if (!stream_)
    throw Error(err);
}

void HipKernelManager::ensure_events_locked(size_t n) {
  char err[kErr];
  while (events_.size() < n) {
    void *e = cast_hip_event_create(err, kErr);
    if (!e)
      throw Error(err);
    events_.push_back(e);
  }
}

void *HipKernelManager::stream() {
  std::lock_guard<std::mutex> lock(mu_);
  ensure_stream_locked();
  return stream_;
}

SyncStats HipKernelManager::sync(bool timed) {
  const auto w0 = std::chrono::steady_clock::now();
  std::vector<Pending> work;
  void *stream = nullptr;
  std::vector<void *> ev;
  {
    std::lock_guard<std::mutex> lock(mu_);
    SyncStats empty;
    if (pending_.empty() && !stream_)
      return empty;
    ensure_stream_locked();
    stream = stream_;
    work.swap(pending_);
    char err[kErr];
    for (const auto &w : work)
      timed = timed || w.cand >= 0; // tuning samples are always timed
    if (timed) {
      ensure_events_locked(work.size() + 1);
      ev.assign(events_.begin(), events_.begin() + long(work.size() + 1));
      if (!work.empty())
        throw_if(cast_hip_event_record(ev[0], stream, err, kErr), err);
    }
    for (size_t i = 0; i < work.size(); ++i) {
      throw_if(cast_hip_gate_apply(work[i].plan, stream, work[i].dptr, work[i].n_qubits, err,
                                   kErr),
               err);
      if (timed)
        throw_if(cast_hip_event_record(ev[i + 1], stream, err, kErr), err);
    }
  }
  char err[kErr];
  throw_if(cast_hip_stream_sync(stream, err, kErr), err);
  SyncStats st;
  if (timed) {
    for (size_t i = 0; i < work.size(); ++i) {
      float ms = 0.f;
      throw_if(cast_hip_event_elapsed_ms(ev[i], ev[i + 1], &ms, err, kErr), err);
      st.kernels.push_back({work[i].id, double(ms) * 1e-3});
    }
    std::lock_guard<std::mutex> lock(mu_);
    for (size_t i = 0; i < work.size(); ++i) {
      if (work[i].cand < 0)
        continue;
      auto it = tune_.find(std::make_pair(work[i].id, work[i].n_qubits));
      if (it == tune_.end() || it->second.chosen >= 0)
        continue;
      TuneState &ts = it->second;
      ts.samples[size_t(work[i].cand)].push_back(st.kernels[i].gpu_s);
      // best minimum so far; drop candidates beyond the margin
      double best = 1e300;
      size_t arg = 0;
      for (size_t c = 0; c < ts.samples.size(); ++c)
        if (!ts.dropped[c] && !ts.samples[c].empty()) {
          const double m = *std::min_element(ts.samples[c].begin(), ts.samples[c].end());
          if (m < best) {
            best = m;
            arg = c;
          }
        }
      bool done = true;
      for (size_t c = 0; c < ts.samples.size(); ++c) {
        if (ts.dropped[c])
          continue;
        if (!ts.samples[c].empty() && c != arg &&
            *std::min_element(ts.samples[c].begin(), ts.samples[c].end()) > kTuneMargin * best) {
          ts.dropped[c] = 1;
          continue;
        }
        done = done && ts.samples[c].size() >= tune_samples_;
      }
      if (done)
        ts.chosen = int(arg);
    }
  }
  st.wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
  return st;
}

TimingStats HipKernelManager::time_adaptive(KernelId id, HipStatevector &sv, double budget_s) {
  const auto samples = time_adaptive_samples_with(
      [&] {
        apply(id, sv);
        return sync(true).kernels.at(0).gpu_s;
      },
      budget_s);
  return stats_from_samples(samples);
}

PlanInfo HipKernelManager::plan_info(KernelId id) const {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = plans_.find(id);
  if (it == plans_.end())
    throw Error("kernel " + std::to_string(id) + " not found");
  cast_hip_gate_plan_info_t info{};
  char err[kErr];
  throw_if(cast_hip_gate_plan_info(it->second, &info, err, kErr), err);
  PlanInfo p;
  p.k = info.n_qubits;
  p.precision = info.precision;
  p.variant = info.variant;
  p.nnz_rows = info.nnz_rows;
  p.scalar_nnz = info.scalar_nnz;
  p.qubits.assign(info.qubits, info.qubits + info.n_qubits);
  return p;
}

uint32_t HipKernelManager::family(KernelId id, uint32_t n_qubits) const {
  std::lock_guard<std::mutex> lock(mu_);
  auto tit = tune_.find(std::make_pair(id, n_qubits));
  if (tit != tune_.end() && tit->second.chosen >= 0)
    return tit->second.fams[size_t(tit->second.chosen)];
  auto it = plans_.find(id);
  if (it == plans_.end())
    throw Error("kernel " + std::to_string(id) + " not found");
  uint32_t v = 0;
  char err[kErr];
  throw_if(cast_hip_gate_plan_family(it->second, n_qubits, &v, err, kErr), err);
  return v;
}

size_t HipKernelManager::n_kernels() const {
  std::lock_guard<std::mutex> lock(mu_);
  return plans_.size();
}

const void *HipKernelManager::plan_handle(KernelId id) const {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = plans_.find(id);
  if (it == plans_.end())
    throw Error("kernel " + std::to_string(id) + " not found");
  return it->second;
}

// ── Simulator ─────────────────────────────────────────────────────────────

std::vector<KernelId> Simulator::compile(const CircuitGraph &graph) {
  std::vector<KernelId> ids;
  size_t idx = 0;
  for (const auto &g : graph.gates_in_row_order()) {
    if (!g->is_unitary())
      throw Error("gate " + std::to_string(idx) +
                  " is noisy: state-vector simulation accepts unitary gates only");
    ids.push_back(mgr_.generate(*g));
    ++idx;
  }
  return ids;
}

SyncStats Simulator::run(const CircuitGraph &graph, HipStatevector &sv) {
  const auto ids = compile(graph);
  for (KernelId id : ids)
    mgr_.apply(id, sv);
  return mgr_.sync(true);
}

HipStatevector Simulator::simulate(const CircuitGraph &graph) {
  HipStatevector sv(uint32_t(graph.n_qubits()), spec_.precision);
  sv.zero();
  run(graph, sv);
  return sv;
}

Simulator::BenchResult Simulator::bench(const CircuitGraph &graph, double budget_s,
                                        Representation repr) {
  BenchResult r;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<KernelId> ids;
  uint32_t n_sv = uint32_t(graph.n_qubits());
  if (repr == Representation::StateVector) {
    ids = compile(graph);
  } else { // prepare_gates, src/simulator/mod.rs:607-615
    const uint32_t n = n_sv;
    n_sv = 2 * n;
    for (const auto &g : graph.gates_in_row_order())
      ids.push_back(mgr_.generate(g->to_density_matrix_gate(n)));
  }
  r.compile_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  r.n_gates = ids.size();
  HipStatevector sv(n_sv, spec_.precision);
  r.samples_s = time_adaptive_samples_with(
      [&] {
        sv.zero();
        for (KernelId id : ids)
          mgr_.apply(id, sv);
        const auto st = mgr_.sync(true);
        double s = 0.0;
        for (const auto &k : st.kernels)
          s += k.gpu_s;
        return s;
      },
      budget_s);
  r.exec = stats_from_samples(r.samples_s);
  return r;
}

} // namespace cast::hip
