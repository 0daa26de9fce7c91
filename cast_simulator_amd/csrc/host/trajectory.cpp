// Batched noisy trajectories; see trajectory.hpp.
#include "trajectory.hpp"

#include "fusion.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <thread>

namespace cast::hip {

namespace {

constexpr size_t kErr = 1024;

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

ComplexSquareMatrix pauli_matrix(uint8_t p) {
  switch (p) {
  case 1:
    return ComplexSquareMatrix::x();
  case 2:
    return ComplexSquareMatrix::y();
  case 3:
    return ComplexSquareMatrix::z();
  default:
    throw Error("Pauli index must be 1 (X), 2 (Y) or 3 (Z)");
  }
}

} // namespace

// ── Circuit and sampling ───────────────────────────────────────────────────

LayeredCircuit LayeredCircuit::random_layers(uint32_t n, uint32_t depth, uint64_t seed) {
  const QuantumCircuit qc = random_layer_circuit(n, depth, seed);
  LayeredCircuit c;
  c.n_qubits = n;
  c.gates = qc.gates();
  const uint32_t per_layer = n + n / 2;
  if (c.gates.size() != size_t(per_layer) * depth)
    throw Error("random_layers: unexpected gate count");
  for (uint32_t l = 1; l <= depth; ++l)
    c.layer_start.push_back(l * per_layer);
  return c;
}

LayeredCircuit LayeredCircuit::from_circuit(const QuantumCircuit &qc) {
  LayeredCircuit c;
  c.n_qubits = qc.n_qubits();
  std::vector<uint32_t> last(c.n_qubits, 0); // per qubit: layers already holding it
  std::vector<std::vector<GatePtr>> layers;
  for (const GatePtr &g : qc.gates()) {
    uint32_t l = 0;
    for (uint32_t q : g->qubits()) {
      if (q >= c.n_qubits)
        throw Error("from_circuit: gate qubit outside the circuit");
      l = std::max(l, last[q]);
    }
    for (uint32_t q : g->qubits())
      last[q] = l + 1;
    if (layers.size() <= l)
      layers.resize(l + 1);
    layers[l].push_back(g);
  }
  for (const auto &layer : layers) {
    for (const GatePtr &g : layer)
      c.gates.push_back(g);
    c.layer_start.push_back(uint32_t(c.gates.size()));
  }
  return c;
}

uint32_t clean_branch(const NoiseModel &m) {
  if (m.branches.empty())
    throw Error("noise model without branches");
  uint32_t best = 0;
  for (uint32_t b = 1; b < m.branches.size(); ++b)
    if (m.branches[b].first > m.branches[best].first)
      best = b;
  return best;
}

std::vector<TrajectorySample> sample_trajectories(const LayeredCircuit &c, const NoiseSpec &ns,
                                                  uint32_t count, uint32_t first) {
  if (ns.min_depth < 1 || ns.min_depth > ns.max_depth || ns.max_depth > c.n_layers())
    throw Error("sample_trajectories: need 1 <= min_depth <= max_depth <= circuit layers (" +
                std::to_string(c.n_layers()) + ")");
  if (!(ns.p >= 0.0 && ns.p <= 1.0))
    throw Error("sample_trajectories: p must be in [0, 1]");
  std::vector<TrajectorySample> out(count);
  for (uint32_t i = 0; i < count; ++i) {
    const uint64_t t = uint64_t(first) + i;
    Rng rng(mix64(ns.seed ^ mix64(t)));
    TrajectorySample &s = out[i];
    s.depth = ns.min_depth + uint32_t(rng.below(ns.max_depth - ns.min_depth + 1));
    const uint32_t end = c.layer_start[s.depth];
    for (uint32_t g = 0; g < end; ++g) {
      if (const auto &nm = c.gates[g]->noise_model()) {
        // Inverse CDF over the branch probabilities (the last branch takes
        // any rounding remainder).
        const double u = rng.uniform();
        const uint32_t nb = uint32_t(nm->branches.size());
        uint32_t b = 0;
        double acc = nm->branches[0].first;
        while (b + 1 < nb && u >= acc)
          acc += nm->branches[++b].first;
        if (b != clean_branch(*nm))
          s.errors.push_back({g, 0, 0, b});
      }
      if (ns.p > 0.0)
        for (uint32_t q : c.gates[g]->qubits())
          if (rng.uniform() < ns.p)
            s.errors.push_back({g, q, uint8_t(1 + rng.below(3)), 0});
    }
  }
  return out;
}

namespace {
// Operator source gate g applies in a trajectory whose events at g are
// ev[0..n): the clean or sampled branch (or the unitary), then the Paulis.
ComplexSquareMatrix trajectory_operator(const QuantumGate &g, const NoiseEvent *ev, size_t n) {
  ComplexSquareMatrix m = g.matrix();
  size_t i = 0;
  if (const auto &nm = g.noise_model()) {
    uint32_t b = clean_branch(*nm);
    if (n > 0 && ev[0].pauli == 0) {
      b = ev[0].branch;
      i = 1;
    }
    if (b >= nm->branches.size())
      throw Error("noise event names branch " + std::to_string(b) + " of a gate with " +
                  std::to_string(nm->branches.size()));
    m = nm->branches[b].second;
  }
  for (; i < n; ++i) {
    if (ev[i].pauli < 1 || ev[i].pauli > 3)
      throw Error("a gate's branch event must come before its Pauli events");
    m = expand_matrix(pauli_matrix(ev[i].pauli), {ev[i].qubit}, g.qubits()).matmul(m);
  }
  return m;
}
} // namespace

QuantumCircuit trajectory_circuit(const LayeredCircuit &c, const TrajectorySample &s) {
  QuantumCircuit qc(c.n_qubits);
  size_t e = 0;
  const uint32_t end = c.layer_start.at(s.depth);
  for (uint32_t g = 0; g < end; ++g) {
    const QuantumGate &src = *c.gates[g];
    size_t e1 = e;
    while (e1 < s.errors.size() && s.errors[e1].gate == g)
      ++e1;
    if (src.noise_model()) {
      // the branch as a plain gate, then the Paulis
      const bool has_branch = e1 > e && s.errors[e].pauli == 0;
      qc.add(QuantumGate(trajectory_operator(src, s.errors.data() + e, has_branch ? 1 : 0),
                         src.qubits()));
      if (has_branch)
        ++e;
    } else {
      qc.add(c.gates[g]);
    }
    for (; e < e1; ++e) {
      if (s.errors[e].pauli < 1 || s.errors[e].pauli > 3)
        throw Error("trajectory_circuit: a branch event on a unitary gate or after a Pauli");
      qc.add(QuantumGate(pauli_matrix(s.errors[e].pauli), {s.errors[e].qubit}));
    }
  }
  if (e != s.errors.size())
    throw Error("trajectory_circuit: errors out of order or beyond the trajectory's depth");
  return qc;
}

// ── Engine ─────────────────────────────────────────────────────────────────

void TrajectoryEngine::DevBuf::ensure(size_t n) {
  if (n <= bytes)
    return;
  release();
  char err[kErr];
  ptr = cast_hip_malloc(n, err, kErr);
  if (!ptr)
    throw Error(err);
  bytes = n;
}

void TrajectoryEngine::DevBuf::release() {
  if (ptr)
    cast_hip_free(ptr);
  ptr = 0;
  bytes = 0;
}

TrajectoryEngine::TrajectoryEngine(LayeredCircuit circuit, cast_precision_t precision,
                                   std::vector<uint32_t> pairs)
    : circ_(std::move(circuit)), prec_(precision), pairs_(std::move(pairs)),
      mgr_(precision == CAST_PRECISION_F32 ? HipKernelGenSpec::f32() : HipKernelGenSpec::f64()) {
  if (circ_.n_qubits < 1 || circ_.n_qubits > 34)
    throw Error("TrajectoryEngine: n_qubits must be in 1..34");
  if (pairs_.size() % 2 || pairs_.size() > 64)
    throw Error("TrajectoryEngine: pairs must be at most 32 (a, b) pairs");
  for (size_t i = 0; i < pairs_.size(); i += 2)
    if (pairs_[i] >= circ_.n_qubits || pairs_[i + 1] >= circ_.n_qubits ||
        pairs_[i] == pairs_[i + 1])
      throw Error("TrajectoryEngine: pair qubits must be distinct and below n");
  clean_.reserve(circ_.gates.size());
  for (const auto &g : circ_.gates) {
    if (g->n_qubits() > 4)
      throw Error("TrajectoryEngine: source gates must act on <= 4 qubits");
    clean_.push_back(trajectory_operator(*g, nullptr, 0));
  }
  mgr_.set_autotune(false);
}

TrajectoryEngine::~TrajectoryEngine() {
  for (DevBuf *b : {&ckpt_, &slots_, &pool_, &gates_, &tasks_, &order_, &out_, &dump_, &work_})
    b->release();
}

void TrajectoryEngine::set_initial_state(std::vector<double> v) {
  if (!v.empty() && v.size() != (size_t(2) << circ_.n_qubits))
    throw Error("set_initial_state: expected 2 * 2^n doubles");
  initial_ = std::move(v);
  ckpt_chunk_ = 0; // checkpoints depend on the start state
}

size_t TrajectoryEngine::sv_bytes() const {
  return (size_t(2) << circ_.n_qubits) * (prec_ == CAST_PRECISION_F32 ? 4 : 8);
}

const TrajectoryEngine::Segment &TrajectoryEngine::segment(uint32_t kmax, uint32_t l0,
                                                          uint32_t l1) {
  const auto key = std::make_tuple(kmax, l0, l1);
  auto it = segs_.find(key);
  if (it != segs_.end())
    return it->second;
  QuantumCircuit qc(circ_.n_qubits);
  for (uint32_t g = circ_.layer_start[l0]; g < circ_.layer_start[l1]; ++g)
    qc.add(QuantumGate(clean_[g], circ_.gates[g]->qubits()).with_origin({g}));
  CircuitGraph cg = CircuitGraph::from_circuit(qc);
  optimize(cg, FusionConfig::size_only(kmax));
  Segment seg;
  for (const GatePtr &g : cg.gates_in_row_order()) {
    if (g->n_qubits() > kmax && g->n_qubits() > 2)
      throw Error("trajectory fusion produced a gate on more than " + std::to_string(kmax) +
                  " qubits");
    FusedGate f;
    f.qubits = g->qubits();
    f.m = g->matrix();
    f.origin = g->origin();
    const uint32_t fi = uint32_t(seg.gates.size());
    for (uint32_t i = 0; i < f.origin.size(); ++i)
      seg.where[f.origin[i]] = {fi, i};
    seg.gates.push_back(std::move(f));
  }
  // Every source gate of the layers must appear exactly once.
  size_t n_origin = 0;
  for (const FusedGate &f : seg.gates)
    n_origin += f.origin.size();
  const size_t n_src = circ_.layer_start[l1] - circ_.layer_start[l0];
  if (seg.where.size() != n_src || n_origin != n_src)
    throw Error("trajectory fusion lost track of its source gates");
  return segs_.emplace(key, std::move(seg)).first->second;
}

ComplexSquareMatrix TrajectoryEngine::component(const FusedGate &f, uint32_t pos) const {
  const uint32_t g = f.origin[pos];
  return expand_matrix(clean_[g], circ_.gates[g]->qubits(), f.qubits);
}

// prefix(f, i) = C_i ... C_0 and suffix(f, i) = C_{m-1} ... C_i for the
// clean components C of F (filled on demand; callers on several threads only
// read entries filled before they start).
const ComplexSquareMatrix &TrajectoryEngine::prefix(const FusedGate &f, uint32_t pos) const {
  while (f.prefix.size() <= pos) {
    const size_t i = f.prefix.size();
    ComplexSquareMatrix e = component(f, uint32_t(i));
    f.prefix.push_back(i == 0 ? std::move(e) : e.matmul(f.prefix[i - 1]));
  }
  return f.prefix[pos];
}

const ComplexSquareMatrix &TrajectoryEngine::suffix(const FusedGate &f, uint32_t pos) const {
  const uint32_t m = uint32_t(f.origin.size());
  // f.suffix[j] = C_{m-1} ... C_{m-1-j}
  while (f.suffix.size() < m - pos) {
    const size_t j = f.suffix.size();
    ComplexSquareMatrix e = component(f, uint32_t(m - 1 - j));
    f.suffix.push_back(j == 0 ? std::move(e) : f.suffix[j - 1].matmul(e));
  }
  return f.suffix[m - 1 - pos];
}

// F with deviating components: F' = S X_last ... X_first P, where P =
// prefix(first - 1), S = suffix(last + 1), and X_i is the trajectory's
// operator for a component with events (its sampled branch, then its Paulis,
// expanded to F's qubits) or the clean component between them.  `errs` holds
// (component position, event) pairs in circuit order.
ComplexSquareMatrix
TrajectoryEngine::dirty_matrix(const FusedGate &f,
                               std::vector<std::pair<uint32_t, NoiseEvent>> errs) const {
  std::stable_sort(errs.begin(), errs.end(),
                   [](const auto &a, const auto &b) { return a.first < b.first; });
  const uint32_t first = errs.front().first, last = errs.back().first;
  const uint32_t m = uint32_t(f.origin.size());
  ComplexSquareMatrix acc;
  bool have = false;
  if (first > 0) {
    acc = prefix(f, first - 1);
    have = true;
  }
  std::vector<NoiseEvent> ev;
  size_t e = 0;
  for (uint32_t pos = first; pos <= last; ++pos) {
    ev.clear();
    for (; e < errs.size() && errs[e].first == pos; ++e)
      ev.push_back(errs[e].second);
    ComplexSquareMatrix x;
    if (ev.empty()) {
      x = component(f, pos);
    } else {
      const uint32_t g = f.origin[pos];
      const QuantumGate &src = *circ_.gates[g];
      x = expand_matrix(trajectory_operator(src, ev.data(), ev.size()), src.qubits(), f.qubits);
    }
    acc = have ? x.matmul(acc) : std::move(x);
    have = true;
  }
  if (last + 1 < m)
    acc = suffix(f, last + 1).matmul(acc);
  return acc;
}

void TrajectoryEngine::ensure_checkpoints(uint32_t B) {
  if (ckpt_chunk_ == B)
    return;
  const double t0 = now_s();
  const uint32_t D = circ_.n_layers();
  n_ckpt_ = D / B + 1; // states after 0, B, 2B, ... layers
  const size_t bytes = sv_bytes();
  ckpt_.ensure(bytes * n_ckpt_);
  HipStatevector sv(circ_.n_qubits, prec_);
  if (initial_.empty())
    sv.zero();
  else
    sv.upload(initial_);
  char err[kErr];
  void *s = mgr_.stream();
  throw_if(cast_hip_memcpy_dtod_async(ckpt_.ptr, sv.dptr(), bytes, s, err, kErr), err);
  for (uint32_t c = 1; c < n_ckpt_; ++c) {
    const Segment &seg = segment(4, (c - 1) * B, c * B);
    for (const FusedGate &f : seg.gates)
      mgr_.apply(mgr_.generate(QuantumGate(f.m, f.qubits)), sv);
    mgr_.sync(false);
    throw_if(cast_hip_memcpy_dtod_async(ckpt_.ptr + bytes * c, sv.dptr(), bytes, s, err, kErr),
             err);
  }
  throw_if(cast_hip_stream_sync(s, err, kErr), err);
  ckpt_chunk_ = B;
  last_checkpoint_s_ = now_s() - t0;
}

TrajectoryPlan TrajectoryEngine::plan(const std::vector<TrajectorySample> &samples,
                                      const TrajectoryTuning &tn) {
  const double t_start = now_s();
  const uint32_t B = tn.chunk_layers, kmax = tn.fusion_max, D = circ_.n_layers();
  if (B < 1 || B > D)
    throw Error("chunk_layers must be in 1..circuit layers");
  // Gate sizes the kernel takes: k <= 4 at f64; k <= 6 at f32, k = 6 from 12
  // qubits on (its matrix-core path works on 64 tiles of 16 combos or more).
  const uint32_t kcap =
      prec_ != CAST_PRECISION_F32 ? 4 : (circ_.n_qubits >= 12 ? 6 : 5);
  if (kmax < 1 || kmax > kcap)
    throw Error("fusion_max must be in 1.." + std::to_string(kcap) +
                " at this precision and size");
  if (samples.empty())
    throw Error("no trajectories to run");
  TrajectoryPlan P;
  auto &gtab = P.gates;
  auto &tasks = P.tasks;
  auto &pool = P.pool;
  std::map<const Segment *, uint32_t> seg_pool; // clean matrices of a segment in this plan
  // k <= 3: interleaved complex; k >= 4: three real matrices Re, Im,
  // Re + Im (the kernel's three-product form); both in whole complex slots.
  auto reserve_matrix = [&](const ComplexSquareMatrix &m) {
    const uint64_t off = pool.size() / 2;
    if (off + m.data().size() * 2 > 0xffffffffull)
      throw Error("trajectory matrix pool exceeds 2^32 complex entries");
    pool.resize(pool.size() + (m.edge_size() >= 16 ? 3 : 2) * m.data().size());
    return uint32_t(off);
  };
  auto write_matrix = [&pool](const ComplexSquareMatrix &m, uint32_t off) {
    double *d = pool.data() + 2 * size_t(off);
    const size_t sz = m.data().size();
    if (m.edge_size() >= 16) {
      for (size_t i = 0; i < sz; ++i) {
        const cplx z = m.data()[i];
        d[i] = z.real();
        d[sz + i] = z.imag();
        d[2 * sz + i] = z.real() + z.imag();
      }
    } else {
      for (size_t i = 0; i < sz; ++i) {
        d[2 * i] = m.data()[i].real();
        d[2 * i + 1] = m.data()[i].imag();
      }
    }
  };
  auto push_matrix = [&](const ComplexSquareMatrix &m) {
    const uint32_t off = reserve_matrix(m);
    write_matrix(m, off);
    return off;
  };
  // Matrices of fused gates that carry errors: their pool slots are reserved
  // in plan order and filled afterwards on several threads.
  struct DirtyJob {
    uint32_t off;
    const FusedGate *f;
    std::vector<std::pair<uint32_t, NoiseEvent>> errs;
  };
  std::vector<DirtyJob> jobs;
  auto slots_of = [](const ComplexSquareMatrix &m) {
    return uint32_t(m.edge_size() >= 16 ? 3 * m.data().size() / 2 : m.data().size());
  };
  auto seg_offsets = [&](const Segment &seg) {
    auto it = seg_pool.find(&seg);
    if (it != seg_pool.end())
      return it->second;
    uint32_t first = 0;
    for (size_t i = 0; i < seg.gates.size(); ++i) {
      const uint32_t off = push_matrix(seg.gates[i].m);
      if (i == 0)
        first = off;
    }
    seg_pool[&seg] = first;
    return first;
  };
  auto push_gate = [&](const FusedGate &f, uint32_t off) {
    cast_hip_traj_gate_t d{};
    d.matrix_offset = off;
    d.k = uint8_t(f.qubits.size());
    for (size_t b = 0; b < f.qubits.size(); ++b)
      d.q[b] = uint8_t(f.qubits[b]);
    gtab.push_back(d);
  };
  // Appends the fused gates of layers [l0, l1) with the given errors; clean
  // matrices of a segment sit consecutively in the pool.
  auto push_segment = [&](uint32_t l0, uint32_t l1, const NoiseEvent *e, size_t n_e) {
    const Segment &seg = segment(kmax, l0, l1);
    const uint32_t base = seg_offsets(seg);
    std::map<uint32_t, std::vector<std::pair<uint32_t, NoiseEvent>>> dirty;
    for (size_t i = 0; i < n_e; ++i) {
      const auto w = seg.where.at(e[i].gate);
      dirty[w.first].push_back({w.second, e[i]});
    }
    uint32_t off = base;
    for (uint32_t fi = 0; fi < seg.gates.size(); ++fi) {
      const FusedGate &f = seg.gates[fi];
      auto it = dirty.find(fi);
      if (it == dirty.end()) {
        push_gate(f, off);
      } else {
        // Fill the component-product caches here: the threads only read them.
        uint32_t lo = ~0u, hi = 0;
        for (const auto &pe : it->second) {
          lo = std::min(lo, pe.first);
          hi = std::max(hi, pe.first);
        }
        if (lo > 0)
          prefix(f, lo - 1);
        if (hi + 1 < f.origin.size())
          suffix(f, hi + 1);
        const uint32_t doff = reserve_matrix(f.m);
        jobs.push_back({doff, &f, std::move(it->second)});
        push_gate(f, doff);
        ++P.n_dirty_matrices;
      }
      off += slots_of(f.m);
    }
  };

  P.task_of.resize(samples.size());
  std::map<uint32_t, uint32_t> clean_task; // depth -> task
  for (size_t t = 0; t < samples.size(); ++t) {
    const TrajectorySample &s = samples[t];
    if (s.depth < 1 || s.depth > D)
      throw Error("trajectory depth out of range");
    if (s.errors.empty()) {
      auto it = clean_task.find(s.depth);
      if (it != clean_task.end()) {
        P.task_of[t] = it->second;
        ++P.n_shared_clean;
        continue;
      }
    }
    // Start at the checkpoint before the first error (or the last one <= depth).
    uint32_t c0 = s.depth / B;
    if (!s.errors.empty()) {
      const uint32_t g0 = s.errors.front().gate;
      const uint32_t l0 = uint32_t(std::upper_bound(circ_.layer_start.begin(),
                                                    circ_.layer_start.end(), g0) -
                                   circ_.layer_start.begin()) - 1;
      c0 = std::min(l0 / B, c0);
    }
    cast_hip_traj_task_t task{};
    task.first_gate = uint32_t(gtab.size());
    task.checkpoint = c0;
    size_t e = 0;
    for (uint32_t c = c0; c * B < s.depth; ++c) {
      const uint32_t l0 = c * B, l1 = std::min((c + 1) * B, s.depth);
      const uint32_t g1 = circ_.layer_start[l1];
      size_t e1 = e;
      while (e1 < s.errors.size() && s.errors[e1].gate < g1)
        ++e1;
      for (size_t i = e + 1; i < e1; ++i)
        if (s.errors[i].gate < s.errors[i - 1].gate)
          throw Error("trajectory errors must be in circuit order");
      push_segment(l0, l1, s.errors.data() + e, e1 - e);
      e = e1;
    }
    if (e != s.errors.size())
      throw Error("trajectory errors beyond its depth or out of order");
    task.n_gates = uint32_t(gtab.size()) - task.first_gate;
    P.task_of[t] = uint32_t(tasks.size());
    if (s.errors.empty())
      clean_task[s.depth] = uint32_t(tasks.size());
    tasks.push_back(task);
  }
  {
    const size_t n_threads = std::min<size_t>(
        {size_t(16), size_t(std::max(1u, std::thread::hardware_concurrency())),
         jobs.size() / 16 + 1});
    std::atomic<size_t> next{0};
    std::exception_ptr failure;
    std::mutex failure_mu;
    auto work = [&] {
      try {
        for (size_t i = next.fetch_add(1); i < jobs.size(); i = next.fetch_add(1))
          write_matrix(dirty_matrix(*jobs[i].f, jobs[i].errs), jobs[i].off);
      } catch (...) {
        std::lock_guard<std::mutex> lk(failure_mu);
        if (!failure)
          failure = std::current_exception();
        next.store(jobs.size());
      }
    };
    std::vector<std::thread> threads;
    for (size_t i = 1; i < n_threads; ++i)
      threads.emplace_back(work);
    work();
    for (auto &t : threads)
      t.join();
    if (failure)
      std::rethrow_exception(failure);
  }
  P.order.resize(tasks.size());
  std::iota(P.order.begin(), P.order.end(), 0u);
  std::stable_sort(P.order.begin(), P.order.end(), [&](uint32_t a, uint32_t b) {
    return tasks[a].n_gates > tasks[b].n_gates;
  });
  P.plan_s = now_s() - t_start;
  return P;
}

TrajectoryRun TrajectoryEngine::run(const std::vector<TrajectorySample> &samples,
                                    const TrajectoryTuning &tn, bool dump) {
  const double t_start = now_s();
  const uint32_t B = tn.chunk_layers;
  TrajectoryPlan P = plan(samples, tn);
  const auto &gtab = P.gates;
  const auto &tasks = P.tasks;
  const auto &pool = P.pool;
  const auto &order = P.order;
  const bool f32 = prec_ == CAST_PRECISION_F32;
  TrajectoryRun R;
  R.n_pairs = uint32_t(pairs_.size() / 2);
  R.chunk_layers = B;
  R.fusion_max = tn.fusion_max;
  R.task_of = P.task_of;
  for (const auto &t : tasks) {
    R.task_ckpt_layer.push_back(t.checkpoint * B);
    R.task_gates.push_back(t.n_gates);
  }
  R.n_shared_clean = P.n_shared_clean;
  R.n_dirty_matrices = P.n_dirty_matrices;
  R.n_tasks = tasks.size();
  R.n_gate_passes = gtab.size();
  R.plan_s = P.plan_s;

  // ── Checkpoints ─────────────────────────────────────────────────────────
  const bool rebuilt = ckpt_chunk_ != B;
  ensure_checkpoints(B);
  R.checkpoint_s = rebuilt ? last_checkpoint_s_ : 0.0;

  // ── Device tables ───────────────────────────────────────────────────────
  const size_t svb = sv_bytes();
  cast_hip_device_info_t info{};
  char err[kErr];
  throw_if(cast_hip_device_info(&info, err, kErr), err);
  uint64_t free_b = 0, total_b = 0;
  throw_if(cast_hip_free_memory(&free_b, &total_b, err, kErr), err);
  const size_t elem = f32 ? 4 : 8;
  const size_t out_bytes = tasks.size() * (1 + R.n_pairs) * sizeof(double);
  const size_t other = pool.size() * elem + gtab.size() * sizeof(cast_hip_traj_gate_t) +
                       tasks.size() * (sizeof(cast_hip_traj_task_t) + 4) + out_bytes +
                       (dump ? tasks.size() * svb : 0);
  const uint64_t reserve = uint64_t(2) << 30;
  const uint64_t avail = free_b + slots_.bytes > other + reserve ? free_b + slots_.bytes - other - reserve
                                                                 : 0;
  uint64_t n_slots = uint64_t(std::max(1u, info.n_cu)) * std::max(1u, tn.slots_per_cu);
  n_slots = std::min<uint64_t>(n_slots, tasks.size());
  if (tn.max_slots)
    n_slots = std::min<uint64_t>(n_slots, tn.max_slots);
  n_slots = std::min<uint64_t>(n_slots, avail / svb);
  if (n_slots == 0)
    throw Error("not enough device memory for one trajectory slot");
  R.n_slots = uint32_t(n_slots);
  R.block = tn.block;
  R.waves_per_simd = tn.waves_per_simd;
  R.mfma_f32 = tn.mfma_f32;

  slots_.ensure(n_slots * svb);
  pool_.ensure(std::max<size_t>(16, pool.size() * elem));
  gates_.ensure(std::max<size_t>(16, gtab.size() * sizeof(cast_hip_traj_gate_t)));
  tasks_.ensure(tasks.size() * sizeof(cast_hip_traj_task_t));
  order_.ensure(order.size() * sizeof(uint32_t));
  out_.ensure(out_bytes);
  work_.ensure(64);
  if (dump)
    dump_.ensure(tasks.size() * svb);

  void *s = mgr_.stream();
  std::vector<float> pool32;
  const void *pool_host = pool.data();
  if (f32) {
    pool32.assign(pool.begin(), pool.end());
    pool_host = pool32.data();
  }
  if (!pool.empty())
    throw_if(cast_hip_memcpy_htod_async(pool_.ptr, pool_host, pool.size() * elem, s, err, kErr),
             err);
  if (!gtab.empty())
    throw_if(cast_hip_memcpy_htod_async(gates_.ptr, gtab.data(),
                                        gtab.size() * sizeof(cast_hip_traj_gate_t), s, err, kErr),
             err);
  throw_if(cast_hip_memcpy_htod_async(tasks_.ptr, tasks.data(),
                                      tasks.size() * sizeof(cast_hip_traj_task_t), s, err, kErr),
           err);
  throw_if(cast_hip_memcpy_htod_async(order_.ptr, order.data(), order.size() * sizeof(uint32_t),
                                      s, err, kErr),
           err);

  cast_hip_traj_launch_t L{};
  L.precision = f32 ? 0 : 1;
  L.n_qubits = circ_.n_qubits;
  L.tasks = tasks_.ptr;
  L.order = order_.ptr;
  L.n_tasks = uint32_t(tasks.size());
  L.n_gates = uint32_t(gtab.size());
  L.gates = gates_.ptr;
  L.matrices = pool_.ptr;
  L.n_matrix_elems = pool.size() / 2;
  L.checkpoints = ckpt_.ptr;
  L.n_checkpoints = n_ckpt_;
  L.n_slots = uint32_t(n_slots);
  L.slots = slots_.ptr;
  L.block = tn.block;
  L.n_pairs = R.n_pairs;
  for (size_t i = 0; i < pairs_.size(); ++i)
    L.pairs[i] = pairs_[i];
  L.out = out_.ptr;
  L.dump = dump ? dump_.ptr : 0;
  L.work = work_.ptr;
  L.waves_per_simd = tn.waves_per_simd;
  L.flags = tn.mfma_f32 ? CAST_HIP_TRAJ_FLAG_MFMA_F32 : 0u;

  void *e0 = cast_hip_event_create(err, kErr);
  if (!e0)
    throw Error(err);
  void *e1 = cast_hip_event_create(err, kErr);
  if (!e1) {
    cast_hip_event_destroy(e0);
    throw Error(err);
  }
  float ms = 0.f;
  int rc = cast_hip_event_record(e0, s, err, kErr);
  if (rc == 0)
    rc = cast_hip_traj_run(&L, s, err, kErr);
  if (rc == 0)
    rc = cast_hip_event_record(e1, s, err, kErr);
  if (rc == 0)
    rc = cast_hip_event_synchronize(e1, err, kErr);
  if (rc == 0)
    rc = cast_hip_event_elapsed_ms(e0, e1, &ms, err, kErr);
  if (rc != 0) {
    // The host tables above must outlive the queued copies.
    char err2[kErr];
    (void)cast_hip_stream_sync(s, err2, kErr);
  }
  cast_hip_event_destroy(e0);
  cast_hip_event_destroy(e1);
  throw_if(rc, err);
  R.gpu_s = ms * 1e-3;

  uint32_t work[2] = {0, 0};
  throw_if(cast_hip_memcpy_dtoh(work, work_.ptr, sizeof(work), s, err, kErr), err);
  if (work[1] != 0)
    throw Error("trajectory kernel rejected a task or gate descriptor");
  std::vector<double> out(tasks.size() * (1 + R.n_pairs));
  throw_if(cast_hip_memcpy_dtoh(out.data(), out_.ptr, out_bytes, s, err, kErr), err);
  if (dump) {
    const size_t n_el = (size_t(2) << circ_.n_qubits) * tasks.size();
    R.dump.resize(n_el);
    if (f32) {
      std::vector<float> tmp(n_el);
      throw_if(cast_hip_memcpy_dtoh(tmp.data(), dump_.ptr, n_el * 4, s, err, kErr), err);
      std::copy(tmp.begin(), tmp.end(), R.dump.begin());
    } else {
      throw_if(cast_hip_memcpy_dtoh(R.dump.data(), dump_.ptr, n_el * 8, s, err, kErr), err);
    }
  }

  // ── Results (f64) ───────────────────────────────────────────────────────
  const uint32_t NP = R.n_pairs;
  R.norm_sq.resize(tasks.size());
  R.zz_sum.resize(tasks.size() * NP);
  for (size_t k = 0; k < tasks.size(); ++k) {
    R.norm_sq[k] = out[k * (1 + NP)];
    for (uint32_t j = 0; j < NP; ++j)
      R.zz_sum[k * NP + j] = out[k * (1 + NP) + 1 + j];
  }
  R.min_norm_sq = *std::min_element(R.norm_sq.begin(), R.norm_sq.end());
  R.max_norm_sq = *std::max_element(R.norm_sq.begin(), R.norm_sq.end());
  R.traj_zz.resize(samples.size());
  double acc_zz = 0.0, acc_norm = 0.0;
  for (size_t t = 0; t < samples.size(); ++t) {
    const uint32_t k = R.task_of[t];
    double z = 0.0;
    for (uint32_t j = 0; j < NP; ++j)
      z += R.zz_sum[size_t(k) * NP + j];
    R.traj_zz[t] = NP ? z / (double(NP) * R.norm_sq[k]) : 0.0;
    acc_zz += NP ? z / double(NP) : 0.0;
    acc_norm += R.norm_sq[k];
  }
  R.mean_zz = acc_zz / acc_norm;
  R.total_s = now_s() - t_start;
  return R;
}

TrajectoryTuning TrajectoryEngine::autotune(const std::vector<TrajectorySample> &sample,
                                            const TrajectoryTuning &start,
                                            std::vector<TrajectoryTuneRecord> *log) {
  TrajectoryTuning best = start, second = start;
  double best_s = 1e300, second_s = 1e300;
  auto trial = [&](const TrajectoryTuning &t) {
    // Checkpoints are part of the cost of a chunk size: time them too.
    ckpt_chunk_ = 0;
    const TrajectoryRun r = run(sample, t);
    const double s = r.total_s / double(sample.size());
    if (log)
      log->push_back({t, s});
    if (s < best_s) {
      second_s = best_s;
      second = best;
      best_s = s;
      best = t;
    } else if (s < second_s) {
      second_s = s;
      second = t;
    }
  };
  trial(best);
  // Launch shape at the start tuning's occupancy: the resident workgroups
  // per CU are bounded by it, so only shapes that can be resident are tried
  // (slots_per_cu x block <= 256 x waves_per_simd).  The 4-waves build is
  // not searched: it spills registers and was slower in every measured
  // shape; a caller can still ask for it.
  const TrajectoryTuning start_shape = best;
  for (uint32_t block : {128u, 256u, 512u})
    for (uint32_t spc : {1u, 2u, 4u}) {
      if (spc * block > 256u * start_shape.waves_per_simd)
        continue;
      TrajectoryTuning t = start_shape;
      t.block = block;
      t.slots_per_cu = spc;
      if (t.block == start_shape.block && t.slots_per_cu == start_shape.slots_per_cu)
        continue;
      trial(t);
    }
  const TrajectoryTuning shape = best;
  // Fusion sizes: the two largest the kernel takes at this precision and
  // size (smaller ones never won a measurement).
  const uint32_t ktop =
      prec_ != CAST_PRECISION_F32 ? 4u : (circ_.n_qubits >= 12 ? 6u : 5u);
  for (uint32_t kmax : {ktop - 1, ktop})
    for (uint32_t B : {1u, 2u, 4u, 8u}) {
      if (B > circ_.n_layers())
        continue;
      TrajectoryTuning t = shape;
      t.fusion_max = kmax;
      t.chunk_layers = B;
      if (t.fusion_max == best.fusion_max && t.chunk_layers == best.chunk_layers)
        continue;
      trial(t);
    }
  // f32 k = 5, 6 gates: the other matrix-core product form (bf16 split or
  // f32 products) at the two best settings so far.
  if (prec_ == CAST_PRECISION_F32) {
    auto same = [](const TrajectoryTuning &a, const TrajectoryTuning &b) {
      return a.chunk_layers == b.chunk_layers && a.fusion_max == b.fusion_max &&
             a.block == b.block && a.slots_per_cu == b.slots_per_cu &&
             a.waves_per_simd == b.waves_per_simd && a.mfma_f32 == b.mfma_f32;
    };
    const TrajectoryTuning top[2] = {best, second};
    for (int i = 0; i < 2; ++i) {
      const TrajectoryTuning &b = top[i];
      if (b.fusion_max < 5 || (i == 1 && same(top[0], top[1])))
        continue;
      TrajectoryTuning t = b;
      t.mfma_f32 = b.mfma_f32 ? 0u : 1u;
      trial(t);
    }
  }
  return best;
}

} // namespace cast::hip
