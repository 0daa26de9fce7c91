// Gate and matrix types; see types.hpp for the reference mapping.
#include "types.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace cast {

// ── Rng ────────────────────────────────────────────────────────────────────

namespace {
uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
} // namespace

Rng::Rng(uint64_t seed) {
  uint64_t x = seed;
  for (auto &w : s_)
    w = splitmix64(x);
}

uint64_t Rng::next_u64() {
  const uint64_t result = rotl(s_[1] * 5, 7) * 9;
  const uint64_t t = s_[1] << 17;
  s_[2] ^= s_[0];
  s_[3] ^= s_[1];
  s_[1] ^= s_[2];
  s_[0] ^= s_[3];
  s_[2] ^= t;
  s_[3] = rotl(s_[3], 45);
  return result;
}

double Rng::uniform() { return double(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

double Rng::normal() {
  if (have_spare_) {
    have_spare_ = false;
    return spare_;
  }
  double u, v, s;
  do {
    u = 2.0 * uniform() - 1.0;
    v = 2.0 * uniform() - 1.0;
    s = u * u + v * v;
  } while (s >= 1.0 || s == 0.0);
  const double m = std::sqrt(-2.0 * std::log(s) / s);
  spare_ = v * m;
  have_spare_ = true;
  return u * m;
}

uint64_t Rng::below(uint64_t n) {
  if (n == 0)
    throw Error("Rng::below(0)");
  const uint64_t limit = ~uint64_t(0) - (~uint64_t(0) % n);
  uint64_t x;
  do {
    x = next_u64();
  } while (x >= limit);
  return x % n;
}

double Rng::exponential(double rate) { return -std::log(1.0 - uniform()) / rate; }

// ── ComplexSquareMatrix ─────────────────────────────────────────────────────

ComplexSquareMatrix::ComplexSquareMatrix(size_t edge, std::vector<cplx> data)
    : n_(edge), d_(std::move(data)) {
  if (d_.size() != n_ * n_)
    throw Error("matrix data size does not match edge_size");
}

ComplexSquareMatrix ComplexSquareMatrix::eye(size_t edge) {
  ComplexSquareMatrix m(edge);
  for (size_t i = 0; i < edge; ++i)
    m.set(i, i, 1.0);
  return m;
}

ComplexSquareMatrix ComplexSquareMatrix::from_reals(size_t edge, const std::vector<double> &v) {
  std::vector<cplx> d(v.size());
  for (size_t i = 0; i < v.size(); ++i)
    d[i] = cplx(v[i], 0.0);
  return ComplexSquareMatrix(edge, std::move(d));
}

ComplexSquareMatrix ComplexSquareMatrix::x() { return from_reals(2, {0, 1, 1, 0}); }
ComplexSquareMatrix ComplexSquareMatrix::y() {
  return ComplexSquareMatrix(2, {cplx(0, 0), cplx(0, -1), cplx(0, 1), cplx(0, 0)});
}
ComplexSquareMatrix ComplexSquareMatrix::z() { return from_reals(2, {1, 0, 0, -1}); }
ComplexSquareMatrix ComplexSquareMatrix::h() {
  const double s = 0.70710678118654752440; // FRAC_1_SQRT_2
  return from_reals(2, {s, s, s, -s});
}
ComplexSquareMatrix ComplexSquareMatrix::s() {
  return ComplexSquareMatrix(2, {cplx(1, 0), cplx(0, 0), cplx(0, 0), cplx(0, 1)});
}
ComplexSquareMatrix ComplexSquareMatrix::t() {
  const double ph = 0.78539816339744830962; // FRAC_PI_4
  return ComplexSquareMatrix(2, {cplx(1, 0), cplx(0, 0), cplx(0, 0), cplx(std::cos(ph), std::sin(ph))});
}
ComplexSquareMatrix ComplexSquareMatrix::rx(double theta) {
  const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
  return ComplexSquareMatrix(2, {cplx(c, 0), cplx(0, -s), cplx(0, -s), cplx(c, 0)});
}
ComplexSquareMatrix ComplexSquareMatrix::ry(double theta) {
  const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
  return from_reals(2, {c, -s, s, c});
}
ComplexSquareMatrix ComplexSquareMatrix::rz(double theta) {
  const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
  return ComplexSquareMatrix(2, {cplx(c, -s), cplx(0, 0), cplx(0, 0), cplx(c, s)});
}
ComplexSquareMatrix ComplexSquareMatrix::u3(double theta, double phi, double lambda) {
  const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
  const cplx eiphi(std::cos(phi), std::sin(phi));
  const cplx eil(std::cos(lambda), std::sin(lambda));
  const cplx eipl(std::cos(phi + lambda), std::sin(phi + lambda));
  return ComplexSquareMatrix(2, {cplx(c, 0), -eil * s, eiphi * s, eipl * c});
}
ComplexSquareMatrix ComplexSquareMatrix::cx() {
  return from_reals(4, {1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0});
}
ComplexSquareMatrix ComplexSquareMatrix::cz() {
  return from_reals(4, {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1});
}
ComplexSquareMatrix ComplexSquareMatrix::cp(double theta) {
  auto m = eye(4);
  m.set(3, 3, cplx(std::cos(theta), std::sin(theta)));
  return m;
}
ComplexSquareMatrix ComplexSquareMatrix::swap() {
  return from_reals(4, {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1});
}
ComplexSquareMatrix ComplexSquareMatrix::ccx() {
  auto m = eye(8);
  m.set(6, 6, 0.0);
  m.set(7, 7, 0.0);
  m.set(6, 7, 1.0);
  m.set(7, 6, 1.0);
  return m;
}
ComplexSquareMatrix ComplexSquareMatrix::pauli(int index) {
  switch (index) {
  case 0:
    return eye(2);
  case 1:
    return x();
  case 2:
    return y();
  case 3:
    return z();
  default:
    throw Error("Pauli index must be 0-3, got " + std::to_string(index));
  }
}

ComplexSquareMatrix ComplexSquareMatrix::random_unitary(size_t edge, Rng &rng) {
  ComplexSquareMatrix m(edge);
  std::vector<cplx> cur(edge);
  for (size_t row = 0; row < edge; ++row) {
    for (;;) {
      for (auto &v : cur) {
        const double re = rng.normal();
        const double im = rng.normal();
        v = cplx(re, im);
      }
      for (size_t prev = 0; prev < row; ++prev) {
        cplx proj(0, 0);
        for (size_t c = 0; c < edge; ++c)
          proj += std::conj(m.get(prev, c)) * cur[c];
        for (size_t c = 0; c < edge; ++c)
          cur[c] -= proj * m.get(prev, c);
      }
      double nrm = 0.0;
      for (const auto &v : cur)
        nrm += std::norm(v);
      nrm = std::sqrt(nrm);
      if (nrm > 1e-8) {
        for (auto &v : cur)
          v /= nrm;
        break;
      }
    }
    for (size_t c = 0; c < edge; ++c)
      m.set(row, c, cur[c]);
  }
  return m;
}

ComplexSquareMatrix ComplexSquareMatrix::random_sparse(size_t edge, double sparsity, Rng &rng) {
  ComplexSquareMatrix m(edge);
  if (sparsity < 1e-8)
    return m;
  for (auto &v : m.d_) {
    if (rng.uniform() < sparsity) {
      const double re = rng.normal();
      const double im = rng.normal();
      v = cplx(re, im);
    }
  }
  return m;
}

ComplexSquareMatrix ComplexSquareMatrix::matmul(const ComplexSquareMatrix &o) const {
  if (n_ != o.n_)
    throw Error("matrix size mismatch");
  // Row r of the product accumulates its terms in k order (as a dot
  // product would), with rows of `o` read contiguously.
  ComplexSquareMatrix out(n_);
  const cplx *a = d_.data(), *b = o.d_.data();
  cplx *d = out.d_.data();
  for (size_t r = 0; r < n_; ++r)
    for (size_t k = 0; k < n_; ++k) {
      const cplx x = a[r * n_ + k];
      const cplx *brow = b + k * n_;
      cplx *drow = d + r * n_;
      for (size_t c = 0; c < n_; ++c)
        drow[c] += x * brow[c];
    }
  return out;
}

ComplexSquareMatrix ComplexSquareMatrix::adjoint() const {
  ComplexSquareMatrix out(n_);
  for (size_t r = 0; r < n_; ++r)
    for (size_t c = 0; c < n_; ++c)
      out.set(c, r, std::conj(get(r, c)));
  return out;
}

ComplexSquareMatrix ComplexSquareMatrix::kron(const ComplexSquareMatrix &o) const {
  const size_t m = n_, n = o.n_, mn = m * n;
  ComplexSquareMatrix out(mn);
  for (size_t i = 0; i < m; ++i)
    for (size_t k = 0; k < n; ++k)
      for (size_t j = 0; j < m; ++j)
        for (size_t l = 0; l < n; ++l)
          out.set(i * n + k, j * n + l, get(i, j) * o.get(k, l));
  return out;
}

ComplexSquareMatrix ComplexSquareMatrix::scaled(double s) const {
  ComplexSquareMatrix out = *this;
  for (auto &v : out.d_)
    v *= s;
  return out;
}

ComplexSquareMatrix &ComplexSquareMatrix::operator+=(const ComplexSquareMatrix &o) {
  if (n_ != o.n_)
    throw Error("matrix size mismatch");
  for (size_t i = 0; i < d_.size(); ++i)
    d_[i] += o.d_[i];
  return *this;
}

double ComplexSquareMatrix::max_abs_diff(const ComplexSquareMatrix &o) const {
  if (n_ != o.n_)
    return INFINITY;
  double m = 0.0;
  for (size_t i = 0; i < d_.size(); ++i)
    m = std::max(m, std::abs(d_[i] - o.d_[i]));
  return m;
}

// ── helpers (gate.rs:436-661) ───────────────────────────────────────────────

size_t compress_bits(size_t index, const std::vector<size_t> &positions) {
  size_t out = 0;
  for (size_t b = 0; b < positions.size(); ++b)
    out |= ((index >> positions[b]) & 1u) << b;
  return out;
}

static size_t permute_basis_index(size_t index, const std::vector<size_t> &new_to_old) {
  size_t out = 0;
  for (size_t np = 0; np < new_to_old.size(); ++np)
    out |= ((index >> np) & 1u) << new_to_old[np];
  return out;
}

std::vector<uint32_t> union_qubits(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
  std::vector<uint32_t> out;
  out.reserve(a.size() + b.size());
  std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(out));
  return out;
}

ComplexSquareMatrix expand_matrix(const ComplexSquareMatrix &m,
                                  const std::vector<uint32_t> &gate_qubits,
                                  const std::vector<uint32_t> &union_q) {
  const size_t ue = size_t(1) << union_q.size();
  std::vector<size_t> pos;
  for (uint32_t q : gate_qubits) {
    auto it = std::find(union_q.begin(), union_q.end(), q);
    if (it == union_q.end())
      throw Error("gate qubit must exist in union qubits");
    pos.push_back(size_t(it - union_q.begin()));
  }
  size_t tmask = 0;
  for (size_t p : pos)
    tmask |= size_t(1) << p;
  ComplexSquareMatrix out(ue);
  for (size_t r = 0; r < ue; ++r)
    for (size_t c = 0; c < ue; ++c) {
      if (((r ^ c) & ~tmask) != 0)
        continue;
      out.set(r, c, m.get(compress_bits(r, pos), compress_bits(c, pos)));
    }
  return out;
}

ComplexSquareMatrix superoperator_of(const ComplexSquareMatrix &u) {
  const size_t dim = u.edge_size(), sd = dim * dim;
  size_t shift = 0;
  while ((size_t(1) << shift) < dim)
    ++shift;
  ComplexSquareMatrix s(sd);
  for (size_t r = 0; r < sd; ++r) {
    const size_t ik = r & (dim - 1), ib = r >> shift;
    for (size_t c = 0; c < sd; ++c) {
      const size_t jk = c & (dim - 1), jb = c >> shift;
      s.set(r, c, u.get(ik, jk) * std::conj(u.get(ib, jb)));
    }
  }
  return s;
}

// ── QuantumGate ─────────────────────────────────────────────────────────────

QuantumGate::QuantumGate(ComplexSquareMatrix m, std::vector<uint32_t> qubits) {
  if (qubits.size() > 20 || m.edge_size() != (size_t(1) << qubits.size()))
    throw Error("matrix dimension does not match number of target qubits");
  // Canonicalise (gate.rs:483-521): sort, then result[i][j] =
  // input[index_map[i]][index_map[j]] with index_map = permute_basis_index.
  std::vector<std::pair<uint32_t, size_t>> perm;
  for (size_t i = 0; i < qubits.size(); ++i)
    perm.emplace_back(qubits[i], i);
  std::sort(perm.begin(), perm.end());
  std::vector<uint32_t> sorted;
  std::vector<size_t> new_to_old;
  for (auto &p : perm) {
    sorted.push_back(p.first);
    new_to_old.push_back(p.second);
  }
  for (size_t i = 1; i < sorted.size(); ++i)
    if (sorted[i] == sorted[i - 1])
      throw Error("gate qubits must be distinct");
  bool ident = true;
  for (size_t i = 0; i < new_to_old.size(); ++i)
    ident = ident && new_to_old[i] == i;
  if (!ident) {
    const size_t n = m.edge_size();
    std::vector<size_t> index_map(n);
    for (size_t i = 0; i < n; ++i)
      index_map[i] = permute_basis_index(i, new_to_old);
    ComplexSquareMatrix pm(n);
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j)
        pm.set(i, j, m.get(index_map[i], index_map[j]));
    m = std::move(pm);
  }
  m_ = std::move(m);
  q_ = std::move(sorted);
}

QuantumGate QuantumGate::from_reals(std::vector<uint32_t> qubits, const std::vector<double> &d) {
  const size_t e = size_t(1) << qubits.size();
  return QuantumGate(ComplexSquareMatrix::from_reals(e, d), std::move(qubits));
}

QuantumGate QuantumGate::x(uint32_t q) { return {ComplexSquareMatrix::x(), {q}}; }
QuantumGate QuantumGate::y(uint32_t q) { return {ComplexSquareMatrix::y(), {q}}; }
QuantumGate QuantumGate::z(uint32_t q) { return {ComplexSquareMatrix::z(), {q}}; }
QuantumGate QuantumGate::h(uint32_t q) { return {ComplexSquareMatrix::h(), {q}}; }
QuantumGate QuantumGate::s(uint32_t q) { return {ComplexSquareMatrix::s(), {q}}; }
QuantumGate QuantumGate::t(uint32_t q) { return {ComplexSquareMatrix::t(), {q}}; }
QuantumGate QuantumGate::rx(double th, uint32_t q) { return {ComplexSquareMatrix::rx(th), {q}}; }
QuantumGate QuantumGate::ry(double th, uint32_t q) { return {ComplexSquareMatrix::ry(th), {q}}; }
QuantumGate QuantumGate::rz(double th, uint32_t q) { return {ComplexSquareMatrix::rz(th), {q}}; }
QuantumGate QuantumGate::u3(double th, double ph, double la, uint32_t q) {
  return {ComplexSquareMatrix::u3(th, ph, la), {q}};
}
QuantumGate QuantumGate::cx(uint32_t c, uint32_t t) { return {ComplexSquareMatrix::cx(), {c, t}}; }
QuantumGate QuantumGate::cz(uint32_t c, uint32_t t) { return {ComplexSquareMatrix::cz(), {c, t}}; }
QuantumGate QuantumGate::cp(double th, uint32_t c, uint32_t t) {
  return {ComplexSquareMatrix::cp(th), {c, t}};
}
QuantumGate QuantumGate::swap(uint32_t a, uint32_t b) {
  return {ComplexSquareMatrix::swap(), {a, b}};
}
QuantumGate QuantumGate::ccx(uint32_t c0, uint32_t c1, uint32_t t) {
  return {ComplexSquareMatrix::ccx(), {c0, c1, t}};
}

QuantumGate QuantumGate::depolarizing(uint32_t q, double p) {
  return QuantumGate(ComplexSquareMatrix::eye(2), {q}).with_depolarizing(p);
}

QuantumGate QuantumGate::with_noise(
    const std::vector<std::pair<double, ComplexSquareMatrix>> &ops) const {
  auto nm = std::make_shared<NoiseModel>();
  for (const auto &op : ops) {
    if (op.second.edge_size() != m_.edge_size())
      throw Error("noise operator dimension must match gate dimension");
    nm->branches.emplace_back(op.first, op.second.matmul(m_));
  }
  QuantumGate g = *this;
  g.noise_ = nm;
  return g;
}

QuantumGate QuantumGate::with_depolarizing(double p) const {
  if (n_qubits() != 1)
    throw Error("depolarizing noise is single-qubit");
  return with_noise({{1.0 - p, ComplexSquareMatrix::eye(2)},
                     {p / 3.0, ComplexSquareMatrix::x()},
                     {p / 3.0, ComplexSquareMatrix::y()},
                     {p / 3.0, ComplexSquareMatrix::z()}});
}

QuantumGate QuantumGate::random_unitary(const std::vector<uint32_t> &qubits, Rng &rng) {
  auto s = qubits;
  std::sort(s.begin(), s.end());
  return QuantumGate(ComplexSquareMatrix::random_unitary(size_t(1) << s.size(), rng), s);
}

QuantumGate QuantumGate::random_sparse(const std::vector<uint32_t> &qubits, double sparsity,
                                       Rng &rng) {
  auto s = qubits;
  std::sort(s.begin(), s.end());
  const size_t e = size_t(1) << s.size();
  auto m = ComplexSquareMatrix::random_sparse(e, sparsity, rng);
  // Every row keeps at least one nonzero entry (gate.rs:380-388).
  for (size_t r = 0; r < e; ++r) {
    bool any = false;
    for (size_t c = 0; c < e; ++c)
      any = any || std::abs(m.get(r, c)) > 1e-8;
    if (!any)
      m.set(r, size_t(rng.below(e)), cplx(0.25 + 0.75 * rng.uniform(), 0.0));
  }
  return QuantumGate(std::move(m), s);
}

QuantumGate QuantumGate::random_arithmetic_intensity(uint32_t n_sv, double ai, Rng &rng) {
  const uint32_t kMin = 4;
  uint32_t k = uint32_t(std::ceil(std::log2(std::max(1.0, std::ceil(ai / 2.0)))));
  k = std::max(k, kMin);
  if (n_sv < k)
    throw Error("n_qubits_sv too small for the requested arithmetic intensity");
  const double max_ai = 2.0 * double(size_t(1) << k);
  const double sparsity = std::min(1.0, std::max(0.0, ai / max_ai));
  std::vector<uint32_t> pool(n_sv);
  std::iota(pool.begin(), pool.end(), 0u);
  for (uint32_t i = 0; i < k; ++i) {
    const uint64_t j = i + rng.below(pool.size() - i);
    std::swap(pool[i], pool[j]);
  }
  std::vector<uint32_t> q(pool.begin(), pool.begin() + k);
  return random_sparse(q, sparsity, rng);
}

size_t QuantumGate::scalar_nnz(double ztol) const {
  size_t n = 0;
  for (const auto &z : m_.data())
    n += (std::fabs(z.real()) > ztol) + (std::fabs(z.imag()) > ztol);
  return n;
}

double QuantumGate::arithmetic_intensity(double ztol) const {
  return double(scalar_nnz(ztol)) / double(m_.edge_size());
}

QuantumGate QuantumGate::matmul(const QuantumGate &other) const {
  const auto uq = union_qubits(q_, other.q_);
  const auto lhs = expand_matrix(m_, q_, uq);
  const auto rhs = expand_matrix(other.m_, other.q_, uq);
  QuantumGate g(lhs.matmul(rhs), uq);
  if (!origin_.empty() || !other.origin_.empty()) {
    g.origin_ = other.origin_;
    g.origin_.insert(g.origin_.end(), origin_.begin(), origin_.end());
  }
  if (noise_ || other.noise_) {
    auto nm = std::make_shared<NoiseModel>();
    if (noise_ && !other.noise_) {
      for (const auto &b : noise_->branches)
        nm->branches.emplace_back(b.first, expand_matrix(b.second, q_, uq).matmul(rhs));
    } else if (!noise_ && other.noise_) {
      for (const auto &b : other.noise_->branches)
        nm->branches.emplace_back(b.first, lhs.matmul(expand_matrix(b.second, other.q_, uq)));
    } else {
      for (const auto &bi : noise_->branches) {
        const auto ei = expand_matrix(bi.second, q_, uq);
        for (const auto &bj : other.noise_->branches)
          nm->branches.emplace_back(bi.first * bj.first,
                                    ei.matmul(expand_matrix(bj.second, other.q_, uq)));
      }
    }
    g.noise_ = nm;
  }
  return g;
}

QuantumGate QuantumGate::to_density_matrix_gate(size_t n_total) const {
  for (uint32_t q : q_)
    if (q >= n_total)
      throw Error("qubit " + std::to_string(q) + " is out of range for n_total=" +
                  std::to_string(n_total));
  ComplexSquareMatrix s;
  if (is_unitary()) {
    s = superoperator_of(m_);
  } else {
    const size_t sd = m_.edge_size() * m_.edge_size();
    s = ComplexSquareMatrix(sd);
    for (const auto &b : noise_->branches)
      s += superoperator_of(b.second).scaled(b.first);
  }
  std::vector<uint32_t> vq = q_;
  for (uint32_t q : q_)
    vq.push_back(q + uint32_t(n_total));
  return QuantumGate(std::move(s), vq);
}

bool QuantumGate::operator==(const QuantumGate &o) const {
  if (!(m_ == o.m_) || q_ != o.q_)
    return false;
  if (bool(noise_) != bool(o.noise_))
    return false;
  if (!noise_)
    return true;
  if (noise_->branches.size() != o.noise_->branches.size())
    return false;
  for (size_t i = 0; i < noise_->branches.size(); ++i)
    if (noise_->branches[i].first != o.noise_->branches[i].first ||
        !(noise_->branches[i].second == o.noise_->branches[i].second))
      return false;
  return true;
}

} // namespace cast
