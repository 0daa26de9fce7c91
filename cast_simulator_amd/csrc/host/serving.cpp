// Continuous-batching serving loop; see serving.hpp.
#include "serving.hpp"

#include "fusion.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>

namespace cast::hip {

namespace {

constexpr size_t kErr = 1024;
// One observe call reduces at most kPairs pairs; a request of n qubits has
// n - 1 neighbour pairs, reduced in chunks of kPairs (two chunks up to 40
// qubits).  Device result layout per slot (doubles): chunk c at
// c * (1 + kPairs) (norm, then its pairs), then the shadow's inner product
// (re, im) and the shadow's norm.
constexpr uint32_t kPairs = 32;
constexpr uint32_t kChunks = 2;
constexpr uint32_t kOutInner = kChunks * (1 + kPairs);
constexpr uint32_t kOutShadowNorm = kOutInner + 2;
constexpr uint32_t kOut = kOutShadowNorm + 1;
constexpr uint32_t kMaxQubits = 40;

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

size_t sv_bytes_of(uint32_t n, cast_precision_t p) {
  return (size_t(2) << n) * (p == CAST_PRECISION_F32 ? 4 : 8);
}

size_t round_up_256(size_t b) { return (b + 255) & ~size_t(255); }

// Device bytes a request holds while admitted: its statevector, plus the F64
// shadow right after it (256-byte aligned) when it has one.
size_t request_bytes(const ServeRequest &r, bool shadow) {
  const size_t b = sv_bytes_of(r.n_qubits, r.precision);
  if (!shadow)
    return b;
  return round_up_256(b) + sv_bytes_of(r.n_qubits, CAST_PRECISION_F64);
}

double percentile(std::vector<double> v, double q) {
  if (v.empty())
    return 0.0;
  std::sort(v.begin(), v.end());
  const size_t k = size_t(std::ceil(q * double(v.size()))) - (q > 0 ? 1 : 0);
  return v[std::min(k, v.size() - 1)];
}

} // namespace

ServeRequest circuit_request(uint32_t id, std::shared_ptr<const QuantumCircuit> circuit,
                             cast_precision_t precision, double arrival_s) {
  if (!circuit)
    throw Error("circuit_request: null circuit");
  ServeRequest r;
  r.id = id;
  r.n_qubits = circuit->n_qubits();
  r.depth = uint32_t(circuit->len());
  r.precision = precision;
  r.arrival_s = arrival_s;
  r.circuit = std::move(circuit);
  return r;
}

std::vector<ServeRequest> poisson_requests(uint32_t count, double rate, uint32_t n_min,
                                           uint32_t n_max, uint32_t depth_min,
                                           uint32_t depth_max, double f32_fraction,
                                           uint64_t seed) {
  if (!(rate > 0.0) || n_min < 2 || n_min > n_max || n_max > 34 || depth_min < 1 ||
      depth_min > depth_max || !(f32_fraction >= 0.0 && f32_fraction <= 1.0))
    throw Error("poisson_requests: invalid parameters");
  Rng rng(seed);
  std::vector<ServeRequest> out(count);
  double t = 0.0;
  for (uint32_t i = 0; i < count; ++i) {
    t += rng.exponential(rate);
    ServeRequest &r = out[i];
    r.id = i;
    r.arrival_s = t;
    r.n_qubits = n_min + uint32_t(rng.below(n_max - n_min + 1));
    r.depth = depth_min + uint32_t(rng.below(depth_max - depth_min + 1));
    r.precision = rng.uniform() < f32_fraction ? CAST_PRECISION_F32 : CAST_PRECISION_F64;
    r.seed = mix64(seed ^ mix64(i + 1));
  }
  return out;
}

// One admitted request: its statevector (and F64 shadow), gate plans and,
// with keep_states, a pinned host copy of its final state.
struct Member {
  size_t req = 0;
  uint64_t sv = 0, shadow = 0; // shadow: F64 replay of an F32 request, or 0
  size_t sv_bytes = 0;
  std::vector<void *> plans, shadow_plans;
};

// One stream slot: a non-blocking stream with its completion event, the
// reduction scratch and result buffers of the request (or, in the batch
// lane, the requests) it is running.
struct Server::Slot {
  void *stream = nullptr, *event = nullptr;
  uint64_t scratch = 0, out = 0; // out: kOut doubles per member
  double *h_out = nullptr;
  std::vector<std::pair<void *, size_t>> h_states; // keep_states: per member, reused
  bool busy = false;
  bool batched = false;
  std::vector<Member> members;
};

// Statevector blocks recycled by exact size under a byte budget.
struct Server::Pool {
  std::map<size_t, std::vector<uint64_t>> free_blocks;
  uint64_t budget = 0, in_use = 0, cached = 0, peak = 0, allocs = 0, reuses = 0;

  bool fits(size_t b) const {
    auto it = free_blocks.find(b);
    if (it != free_blocks.end() && !it->second.empty())
      return true;
    return in_use + b <= budget;
  }
  uint64_t acquire(size_t b) {
    auto it = free_blocks.find(b);
    if (it != free_blocks.end() && !it->second.empty()) {
      const uint64_t p = it->second.back();
      it->second.pop_back();
      cached -= b;
      in_use += b;
      ++reuses;
      peak = std::max(peak, in_use);
      return p;
    }
    // Trim cached blocks of other sizes until the new one fits the budget.
    // Freeing waits for the device, so it only happens when the mix of sizes
    // changes.
    for (auto &kv : free_blocks) {
      while (in_use + cached + b > budget && !kv.second.empty()) {
        cast_hip_free(kv.second.back());
        kv.second.pop_back();
        cached -= kv.first;
      }
    }
    if (in_use + cached + b > budget)
      return 0;
    char err[kErr];
    const uint64_t p = cast_hip_malloc(b, err, kErr);
    if (!p)
      throw Error(err);
    in_use += b;
    ++allocs;
    peak = std::max(peak, in_use);
    return p;
  }
  void release(uint64_t p, size_t b) {
    free_blocks[b].push_back(p);
    in_use -= b;
    cached += b;
  }
  void clear() {
    for (auto &kv : free_blocks)
      for (uint64_t p : kv.second)
        cast_hip_free(p);
    free_blocks.clear();
    cached = 0;
  }
};

Server::Server(ServeConfig cfg) : cfg_(cfg), pool_(new Pool) {
  if (cfg_.streams < 1 || cfg_.streams > 64)
    throw Error("serving: streams must be in 1..64");
  if (cfg_.compile_threads < 1 || cfg_.compile_threads > 64)
    throw Error("serving: compile_threads must be in 1..64");
  if (cfg_.fusion_max < 1 || cfg_.fusion_max > 6)
    throw Error("serving: fusion_max must be in 1..6");
  if (cfg_.batch_size < 1 || cfg_.batch_size > 64)
    throw Error("serving: batch_size must be in 1..64");
  if (cfg_.batch_max_qubits > 24)
    throw Error("serving: batch_max_qubits must be at most 24 (0 = no batch lane)");
  char err[kErr];
  const size_t scratch = cast_hip_observe_scratch_bytes(kPairs);
  const size_t out_doubles = size_t(kOut) * (cfg_.batch_max_qubits ? cfg_.batch_size : 1);
  for (uint32_t i = 0; i < cfg_.streams; ++i) {
    auto *s = new Slot;
    slots_.push_back(s);
    s->stream = cast_hip_stream_create_nonblocking(err, kErr);
    if (!s->stream)
      throw Error(err);
    s->event = cast_hip_event_create(err, kErr);
    if (!s->event)
      throw Error(err);
    s->scratch = cast_hip_malloc(std::max(scratch, cast_hip_inner_product_scratch_bytes()), err,
                                 kErr);
    s->out = s->scratch ? cast_hip_malloc(out_doubles * sizeof(double), err, kErr) : 0;
    s->h_out = static_cast<double *>(
        s->out ? cast_hip_host_alloc(out_doubles * sizeof(double), err, kErr) : nullptr);
    if (!s->h_out)
      throw Error(err);
  }
}

Server::~Server() {
  char err[kErr];
  for (Slot *s : slots_) {
    if (s->stream)
      (void)cast_hip_stream_sync(s->stream, err, kErr);
    for (Member &m : s->members) {
      for (void *p : m.plans)
        cast_hip_gate_plan_destroy(p);
      for (void *p : m.shadow_plans)
        cast_hip_gate_plan_destroy(p);
      if (s->busy && m.sv)
        cast_hip_free(m.sv);
    }
    cast_hip_free(s->scratch);
    cast_hip_free(s->out);
    cast_hip_host_free(s->h_out);
    for (auto &h : s->h_states)
      cast_hip_host_free(h.first);
    if (s->event)
      cast_hip_event_destroy(s->event);
    if (s->stream)
      cast_hip_stream_destroy(s->stream);
    delete s;
  }
  if (pool_)
    pool_->clear();
  delete pool_;
}

std::vector<ServeResult> Server::run(const std::vector<ServeRequest> &reqs, ServeStats *stats) {
  const size_t N = reqs.size();
  std::vector<ServeResult> res(N);
  if (N == 0)
    return res;
  // Requests in arrival order.
  std::vector<size_t> order(N);
  for (size_t i = 0; i < N; ++i)
    order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });

  char err[kErr];
  uint64_t free_b = 0, total_b = 0;
  throw_if(cast_hip_free_memory(&free_b, &total_b, err, kErr), err);
  Pool &pool = *pool_;
  const uint64_t avail = free_b + pool.cached > cfg_.reserve_bytes
                             ? free_b + pool.cached - cfg_.reserve_bytes
                             : 0;
  pool.budget = cfg_.budget_bytes ? std::min<uint64_t>(cfg_.budget_bytes, avail + pool.in_use)
                                  : avail + pool.in_use;
  pool.peak = pool.in_use;
  for (const ServeRequest &r : reqs) {
    const std::string who = "serving: request " + std::to_string(r.id);
    if (r.circuit) {
      if (r.n_qubits != r.circuit->n_qubits())
        throw Error(who + " has n_qubits " + std::to_string(r.n_qubits) + " but its circuit has " +
                    std::to_string(r.circuit->n_qubits()));
      if (r.n_qubits < 1 || r.n_qubits > kMaxQubits)
        throw Error(who + " has an invalid qubit count (1.." + std::to_string(kMaxQubits) + ")");
      for (const GatePtr &g : r.circuit->gates())
        if (!g->is_unitary())
          throw Error(who + " contains a noisy gate; serving takes unitary circuits");
    } else if (r.n_qubits < 2 || r.n_qubits > kMaxQubits || r.depth < 1) {
      throw Error(who + " has an invalid shape");
    }
    if (request_bytes(r, cfg_.f64_shadow && r.precision == CAST_PRECISION_F32) > pool.budget)
      throw Error(who + " needs more memory than the statevector budget");
  }

  // ── compile workers ──────────────────────────────────────────────────────
  struct Compiled {
    std::vector<void *> plans, shadow_plans;
    std::string error;
  };
  std::vector<Compiled> compiled(N);
  std::mutex mu;
  std::condition_variable cv;
  std::map<size_t, size_t> ready; // arrival rank -> request index
  std::atomic<size_t> next{0};
  std::atomic<bool> abort{false};
  const double t0 = now_s();
  auto worker = [&] {
    for (;;) {
      const size_t rank = next.fetch_add(1);
      if (rank >= N || abort.load())
        return;
      const size_t i = order[rank];
      const ServeRequest &r = reqs[i];
      if (cfg_.realtime) {
        const double wait = t0 + r.arrival_s - now_s();
        if (wait > 0)
          std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      }
      Compiled c;
      try {
        CircuitGraph g = CircuitGraph::from_circuit(
            r.circuit ? *r.circuit : random_layer_circuit(r.n_qubits, r.depth, r.seed));
        optimize(g, FusionConfig::size_only(cfg_.fusion_max));
        const std::vector<GatePtr> fused = g.gates_in_row_order();
        auto compile = [&](cast_precision_t prec, std::vector<void *> &out) {
          const HipKernelGenSpec spec =
              prec == CAST_PRECISION_F32 ? HipKernelGenSpec::f32() : HipKernelGenSpec::f64();
          cast_hip_kernel_gen_spec_t cs{};
          cs.precision = spec.precision;
          cs.ztol = spec.ztol;
          cs.otol = spec.otol;
          cs.variant = CAST_HIP_VARIANT_AUTO;
          char e[kErr];
          for (const GatePtr &gp : fused) {
            const auto &m = gp->matrix().data();
            std::vector<cast_complex64_t> cm(m.size());
            for (size_t k = 0; k < m.size(); ++k)
              cm[k] = {m[k].real(), m[k].imag()};
            void *plan = nullptr;
            if (cast_hip_gate_plan_create(&cs, cm.data(), cm.size(), gp->qubits().data(),
                                          gp->qubits().size(), &plan, e, kErr) != 0)
              throw Error(e);
            out.push_back(plan);
          }
        };
        compile(r.precision, c.plans);
        if (cfg_.f64_shadow && r.precision == CAST_PRECISION_F32)
          compile(CAST_PRECISION_F64, c.shadow_plans);
      } catch (const std::exception &ex) {
        for (void *p : c.plans)
          cast_hip_gate_plan_destroy(p);
        for (void *p : c.shadow_plans)
          cast_hip_gate_plan_destroy(p);
        c.plans.clear();
        c.shadow_plans.clear();
        c.error = ex.what();
      }
      const double tc = now_s() - t0;
      {
        std::lock_guard<std::mutex> lock(mu);
        compiled[i] = std::move(c);
        res[i].compiled_s = tc;
        ready[rank] = i;
      }
      cv.notify_one();
    }
  };
  std::vector<std::thread> threads;
  for (uint32_t k = 0; k < cfg_.compile_threads; ++k)
    threads.emplace_back(worker);

  // ── scheduler ───────────────────────────────────────────────────────────
  size_t done = 0;
  uint32_t in_flight = 0, max_in_flight = 0;
  uint64_t n_batches = 0, n_batched = 0, n_batch_launches = 0;
  std::string failure;
  auto is_small = [&](size_t i) {
    return cfg_.batch_max_qubits != 0 && reqs[i].n_qubits <= cfg_.batch_max_qubits;
  };
  auto release_plans = [](Member &m) {
    for (void *p : m.plans)
      cast_hip_gate_plan_destroy(p);
    for (void *p : m.shadow_plans)
      cast_hip_gate_plan_destroy(p);
    m.plans.clear();
    m.shadow_plans.clear();
  };
  auto finish = [&](Slot &s) {
    const double t_done = now_s() - t0;
    for (size_t mi = 0; mi < s.members.size(); ++mi) {
      Member &m = s.members[mi];
      ServeResult &o = res[m.req];
      const ServeRequest &r = reqs[m.req];
      const double *h = s.h_out + mi * kOut;
      o.done_s = t_done;
      o.norm_sq = h[0];
      const uint32_t np = r.n_qubits - 1;
      o.zz.clear();
      for (uint32_t c = 0; c * kPairs < np; ++c) {
        const double *z = h + c * (1 + kPairs) + 1;
        o.zz.insert(o.zz.end(), z, z + std::min(kPairs, np - c * kPairs));
      }
      if (m.shadow) {
        const double re = h[kOutInner], im = h[kOutInner + 1];
        o.fidelity_vs_f64 = (re * re + im * im) / (h[0] * h[kOutShadowNorm]);
      }
      if (cfg_.keep_states) {
        const size_t n_el = size_t(2) << r.n_qubits;
        o.state.resize(n_el);
        const void *hs = s.h_states[mi].first;
        if (r.precision == CAST_PRECISION_F32) {
          const float *f = static_cast<const float *>(hs);
          std::copy(f, f + n_el, o.state.begin());
        } else {
          std::memcpy(o.state.data(), hs, n_el * sizeof(double));
        }
      }
      release_plans(m);
      pool.release(m.sv, m.sv_bytes);
      --in_flight;
      ++done;
    }
    s.members.clear();
    s.busy = false;
  };
  // Queue a slot's members (statevectors already acquired): |0> fill, the
  // gates (one launch per gate, or in the batch lane one launch per step and
  // (precision, k) group), the reductions, the result copy and the event.
  auto admit = [&](Slot &s, bool batched) {
    s.busy = true;
    s.batched = batched;
    const double t_admit = now_s() - t0;
    for (Member &m : s.members) {
      const ServeRequest &r = reqs[m.req];
      ServeResult &o = res[m.req];
      o.id = r.id;
      o.n_qubits = r.n_qubits;
      o.depth = r.depth;
      o.precision = r.precision;
      o.arrival_s = r.arrival_s;
      o.n_gates = uint32_t(m.plans.size());
      o.admit_s = t_admit;
      ++in_flight;
      const size_t n_el = size_t(2) << r.n_qubits;
      throw_if(cast_hip_sv_zero_async(m.sv, n_el, uint8_t(r.precision), s.stream, err, kErr), err);
      if (m.shadow)
        throw_if(cast_hip_sv_zero_async(m.shadow, n_el, 1, s.stream, err, kErr), err);
    }
    max_in_flight = std::max(max_in_flight, in_flight);
    if (!batched) {
      for (Member &m : s.members) {
        const uint32_t n = reqs[m.req].n_qubits;
        for (void *p : m.plans)
          throw_if(cast_hip_gate_apply(p, s.stream, m.sv, n, err, kErr), err);
        for (void *p : m.shadow_plans)
          throw_if(cast_hip_gate_apply(p, s.stream, m.shadow, n, err, kErr), err);
      }
    } else {
      ++n_batches;
      n_batched += s.members.size();
      size_t steps = 0;
      for (const Member &m : s.members)
        steps = std::max({steps, m.plans.size(), m.shadow_plans.size()});
      struct Group {
        std::vector<const void *> plans;
        std::vector<uint64_t> svs;
        std::vector<uint32_t> ns;
      };
      auto plan_key = [&](void *p) {
        cast_hip_gate_plan_info_t info{};
        throw_if(cast_hip_gate_plan_info(p, &info, err, kErr), err);
        return std::make_pair(uint32_t(info.precision), info.n_qubits);
      };
      for (size_t step = 0; step < steps; ++step) {
        std::map<std::pair<uint32_t, uint32_t>, Group> groups; // (precision, k)
        for (const Member &m : s.members) {
          const uint32_t n = reqs[m.req].n_qubits;
          if (step < m.plans.size()) {
            Group &g = groups[plan_key(m.plans[step])];
            g.plans.push_back(m.plans[step]);
            g.svs.push_back(m.sv);
            g.ns.push_back(n);
          }
          if (step < m.shadow_plans.size()) {
            Group &g = groups[plan_key(m.shadow_plans[step])];
            g.plans.push_back(m.shadow_plans[step]);
            g.svs.push_back(m.shadow);
            g.ns.push_back(n);
          }
        }
        for (auto &kv : groups) {
          Group &g = kv.second;
          throw_if(cast_hip_gate_apply_batch(g.plans.data(), g.svs.data(), g.ns.data(),
                                             g.plans.size(), s.stream, err, kErr),
                   err);
          n_batch_launches += (g.plans.size() + 63) / 64;
        }
      }
    }
    // Norm and neighbour pairs in chunks of kPairs, per member (the scratch
    // is reused in stream order).
    for (size_t mi = 0; mi < s.members.size(); ++mi) {
      const Member &m = s.members[mi];
      const ServeRequest &r = reqs[m.req];
      const uint8_t prec = uint8_t(r.precision);
      const uint64_t out = s.out + mi * kOut * sizeof(double);
      const uint32_t np = r.n_qubits - 1;
      uint32_t pairs[2 * kPairs];
      for (uint32_t c = 0; c == 0 || c * kPairs < np; ++c) {
        const uint32_t cnt = std::min(kPairs, np - c * kPairs);
        for (uint32_t j = 0; j < cnt; ++j) {
          pairs[2 * j] = c * kPairs + j;
          pairs[2 * j + 1] = c * kPairs + j + 1;
        }
        throw_if(cast_hip_observe_async(m.sv, r.n_qubits, prec, pairs, cnt,
                                        out + c * (1 + kPairs) * sizeof(double), s.scratch,
                                        s.stream, err, kErr),
                 err);
      }
      if (m.shadow) {
        throw_if(cast_hip_inner_product_async(m.sv, prec, m.shadow, 1, r.n_qubits,
                                              out + kOutInner * sizeof(double), s.scratch,
                                              s.stream, err, kErr),
                 err);
        throw_if(cast_hip_observe_async(m.shadow, r.n_qubits, 1, pairs, 0,
                                        out + kOutShadowNorm * sizeof(double), s.scratch,
                                        s.stream, err, kErr),
                 err);
      }
    }
    throw_if(cast_hip_memcpy_dtoh_async(s.h_out, s.out, s.members.size() * kOut * sizeof(double),
                                        s.stream, err, kErr),
             err);
    if (cfg_.keep_states) {
      if (s.h_states.size() < s.members.size())
        s.h_states.resize(s.members.size(), {nullptr, 0});
      for (size_t mi = 0; mi < s.members.size(); ++mi) {
        const ServeRequest &r = reqs[s.members[mi].req];
        const size_t b = sv_bytes_of(r.n_qubits, r.precision);
        auto &h = s.h_states[mi];
        if (h.second < b) {
          cast_hip_host_free(h.first);
          h.first = cast_hip_host_alloc(b, err, kErr);
          if (!h.first) {
            h.second = 0;
            throw Error(err);
          }
          h.second = b;
        }
        throw_if(cast_hip_memcpy_dtoh_async(h.first, s.members[mi].sv, b, s.stream, err, kErr),
                 err);
      }
    }
    throw_if(cast_hip_event_record(s.event, s.stream, err, kErr), err);
  };
  // Statevector memory for request i (and its shadow), or 0 if it does not
  // fit the budget yet.
  auto take = [&](size_t i, Compiled &c, Member &m) {
    const ServeRequest &r = reqs[i];
    const bool shadow = !c.shadow_plans.empty();
    const size_t held = request_bytes(r, shadow);
    if (!pool.fits(held))
      return false;
    m.sv = pool.acquire(held);
    if (!m.sv)
      return false;
    m.req = i;
    m.sv_bytes = held;
    m.shadow = shadow ? m.sv + round_up_256(sv_bytes_of(r.n_qubits, r.precision)) : 0;
    m.plans = std::move(c.plans);
    m.shadow_plans = std::move(c.shadow_plans);
    return true;
  };

  try {
    while (done < N) {
      bool progress = false;
      for (Slot *s : slots_) {
        if (!s->busy)
          continue;
        int fin = 0;
        throw_if(cast_hip_event_query(s->event, &fin, err, kErr), err);
        if (fin) {
          finish(*s);
          progress = true;
        }
      }
      for (Slot *s : slots_) {
        if (s->busy)
          continue;
        // Oldest compiled request that fits; younger ones may pass an older
        // one only when the older one's statevector does not fit yet.  A
        // small request takes up to batch_size - 1 further small ones along
        // (oldest first, each only if it fits).
        bool batched = false;
        {
          std::lock_guard<std::mutex> lock(mu);
          for (auto it = ready.begin(); it != ready.end();) {
            const size_t i = it->second;
            if (!compiled[i].error.empty())
              throw Error("serving: compiling request " + std::to_string(reqs[i].id) + ": " +
                          compiled[i].error);
            const bool small = is_small(i);
            if (!s->members.empty() && (!batched || !small)) {
              ++it;
              continue;
            }
            Member m;
            if (!take(i, compiled[i], m)) {
              ++it;
              continue;
            }
            s->members.push_back(std::move(m));
            it = ready.erase(it);
            if (!small)
              break;
            batched = true;
            if (s->members.size() >= cfg_.batch_size)
              break;
          }
        }
        if (s->members.empty())
          break;
        admit(*s, batched);
        progress = true;
      }
      if (!progress) {
        std::unique_lock<std::mutex> lock(mu);
        cv.wait_for(lock, std::chrono::microseconds(in_flight ? 20 : 500));
      }
    }
  } catch (const std::exception &ex) {
    failure = ex.what();
  }
  abort = true;
  for (auto &t : threads)
    t.join();
  if (!failure.empty()) {
    // Drain the device, then release everything this run still holds.
    for (Slot *s : slots_) {
      (void)cast_hip_stream_sync(s->stream, err, kErr);
      for (Member &m : s->members) {
        release_plans(m);
        pool.release(m.sv, m.sv_bytes);
      }
      s->members.clear();
      s->busy = false;
    }
    for (auto &c : compiled) {
      for (void *p : c.plans)
        cast_hip_gate_plan_destroy(p);
      for (void *p : c.shadow_plans)
        cast_hip_gate_plan_destroy(p);
    }
    throw Error(failure);
  }
  const double wall = now_s() - t0;
  if (stats) {
    std::vector<double> lat, queue;
    for (const ServeResult &o : res) {
      lat.push_back(o.done_s - o.arrival_s);
      queue.push_back(o.admit_s - o.arrival_s);
    }
    ServeStats &st = *stats;
    st.wall_s = wall;
    st.throughput = double(N) / wall;
    st.p50_s = percentile(lat, 0.50);
    st.p99_s = percentile(lat, 0.99);
    st.max_s = *std::max_element(lat.begin(), lat.end());
    double m = 0.0;
    for (double x : lat)
      m += x;
    st.mean_s = m / double(N);
    st.queue_p50_s = percentile(queue, 0.50);
    st.queue_p99_s = percentile(queue, 0.99);
    st.max_in_flight = max_in_flight;
    st.peak_sv_bytes = pool.peak;
    st.budget_bytes = pool.budget;
    st.sv_allocs = pool.allocs;
    st.sv_reuses = pool.reuses;
    st.batches = n_batches;
    st.batched_requests = n_batched;
    st.batch_launches = n_batch_launches;
  }
  return res;
}

} // namespace cast::hip
