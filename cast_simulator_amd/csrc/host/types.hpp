// Gate and matrix types of the host layer.
//
// Semantics follow the reference's src/types/{matrix,gate,noise}.rs:
//   - ComplexSquareMatrix: dense row-major n x n of complex<double>
//     (matrix.rs:12-15).
//   - QuantumGate: matrix + strictly ascending qubits; the constructor sorts
//     the qubits and permutes rows/cols so that qubits[b] is local bit b
//     (gate.rs:32-45, 483-521).
//   - Noise: list of (probability, full Kraus operator K_i = V_i * U)
//     (gate.rs:129-151).
#pragma once

#include <complex>
#include <cstdint>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cast {

using cplx = std::complex<double>;

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Rng;

class ComplexSquareMatrix {
public:
  ComplexSquareMatrix() = default;
  explicit ComplexSquareMatrix(size_t edge) : n_(edge), d_(edge * edge) {}
  ComplexSquareMatrix(size_t edge, std::vector<cplx> data);

  static ComplexSquareMatrix zeros(size_t edge) { return ComplexSquareMatrix(edge); }
  static ComplexSquareMatrix eye(size_t edge);
  static ComplexSquareMatrix from_reals(size_t edge, const std::vector<double> &data);
  // Standard matrices, src/types/matrix.rs:59-210.
  static ComplexSquareMatrix x();
  static ComplexSquareMatrix y();
  static ComplexSquareMatrix z();
  static ComplexSquareMatrix h();
  static ComplexSquareMatrix s();
  static ComplexSquareMatrix t();
  static ComplexSquareMatrix rx(double theta);
  static ComplexSquareMatrix ry(double theta);
  static ComplexSquareMatrix rz(double theta);
  static ComplexSquareMatrix u3(double theta, double phi, double lambda);
  static ComplexSquareMatrix cx();
  static ComplexSquareMatrix cz();
  static ComplexSquareMatrix cp(double theta);
  static ComplexSquareMatrix swap();
  static ComplexSquareMatrix ccx();
  static ComplexSquareMatrix pauli(int index);
  // Haar-random unitary by row Gram-Schmidt of complex Gaussians
  // (matrix.rs:270-308).
  static ComplexSquareMatrix random_unitary(size_t edge, Rng &rng);
  // Random sparse (non-unitary) matrix with the given fraction of nonzero
  // complex entries.
  static ComplexSquareMatrix random_sparse(size_t edge, double sparsity, Rng &rng);

  size_t edge_size() const { return n_; }
  const cplx &get(size_t r, size_t c) const { return d_[r * n_ + c]; }
  void set(size_t r, size_t c, cplx v) { d_[r * n_ + c] = v; }
  const std::vector<cplx> &data() const { return d_; }
  std::vector<cplx> &data_mut() { return d_; }

  ComplexSquareMatrix matmul(const ComplexSquareMatrix &o) const;
  ComplexSquareMatrix adjoint() const;
  ComplexSquareMatrix kron(const ComplexSquareMatrix &o) const;
  ComplexSquareMatrix scaled(double s) const;
  ComplexSquareMatrix &operator+=(const ComplexSquareMatrix &o);
  bool operator==(const ComplexSquareMatrix &o) const { return n_ == o.n_ && d_ == o.d_; }
  // max |a_ij - b_ij|
  double max_abs_diff(const ComplexSquareMatrix &o) const;

private:
  size_t n_ = 0;
  std::vector<cplx> d_;
};

struct NoiseModel {
  std::vector<std::pair<double, ComplexSquareMatrix>> branches;
  size_t len() const { return branches.size(); }
};

class QuantumGate;
using GatePtr = std::shared_ptr<const QuantumGate>;

class QuantumGate {
public:
  QuantumGate(ComplexSquareMatrix m, std::vector<uint32_t> qubits);

  static QuantumGate from_reals(std::vector<uint32_t> qubits, const std::vector<double> &data);
  static QuantumGate x(uint32_t q);
  static QuantumGate y(uint32_t q);
  static QuantumGate z(uint32_t q);
  static QuantumGate h(uint32_t q);
  static QuantumGate s(uint32_t q);
  static QuantumGate t(uint32_t q);
  static QuantumGate rx(double theta, uint32_t q);
  static QuantumGate ry(double theta, uint32_t q);
  static QuantumGate rz(double theta, uint32_t q);
  static QuantumGate u3(double theta, double phi, double lambda, uint32_t q);
  static QuantumGate cx(uint32_t c, uint32_t t);
  static QuantumGate cz(uint32_t c, uint32_t t);
  static QuantumGate cp(double theta, uint32_t c, uint32_t t);
  static QuantumGate swap(uint32_t a, uint32_t b);
  static QuantumGate ccx(uint32_t c0, uint32_t c1, uint32_t t);
  static QuantumGate depolarizing(uint32_t q, double p);
  static QuantumGate random_unitary(const std::vector<uint32_t> &qubits, Rng &rng);
  static QuantumGate random_sparse(const std::vector<uint32_t> &qubits, double sparsity, Rng &rng);
  // gate.rs:395-431: smallest k >= 4 whose dense AI can reach `ai`.
  static QuantumGate random_arithmetic_intensity(uint32_t n_qubits_sv, double ai, Rng &rng);

  const ComplexSquareMatrix &matrix() const { return m_; }
  const std::vector<uint32_t> &qubits() const { return q_; }
  size_t n_qubits() const { return q_.size(); }
  size_t effective_n_qubits() const { return is_unitary() ? q_.size() : 2 * q_.size(); }
  bool is_unitary() const { return !noise_; }
  const std::shared_ptr<const NoiseModel> &noise_model() const { return noise_; }

  // Attach noise: each (p, V) becomes the branch (p, V * U) (gate.rs:129-151).
  QuantumGate with_noise(const std::vector<std::pair<double, ComplexSquareMatrix>> &ops) const;
  QuantumGate with_depolarizing(double p) const;

  // gate.rs:261-272
  size_t scalar_nnz(double ztol) const;
  double arithmetic_intensity(double ztol) const;

  // self * other on the union of qubits (gate.rs:280-322), with the
  // Cartesian Kraus product when either side is noisy.
  QuantumGate matmul(const QuantumGate &other) const;

  // Superoperator lift for density-matrix mode (gate.rs:236-258).
  QuantumGate to_density_matrix_gate(size_t n_total) const;

  // Provenance: ids of the source gates this gate is the product of, in
  // application order.  Empty unless set by the caller; matmul concatenates
  // (other's ids, then this gate's), so a fused gate lists its components.
  const std::vector<uint32_t> &origin() const { return origin_; }
  QuantumGate with_origin(std::vector<uint32_t> ids) const {
    QuantumGate g = *this;
    g.origin_ = std::move(ids);
    return g;
  }

  bool operator==(const QuantumGate &o) const;

private:
  QuantumGate() = default;
  ComplexSquareMatrix m_;
  std::vector<uint32_t> q_;
  std::shared_ptr<const NoiseModel> noise_;
  std::vector<uint32_t> origin_;
};

// gate.rs:436-468: expand `m` acting on `gate_qubits` to act on `union_qubits`.
ComplexSquareMatrix expand_matrix(const ComplexSquareMatrix &m,
                                  const std::vector<uint32_t> &gate_qubits,
                                  const std::vector<uint32_t> &union_qubits);
std::vector<uint32_t> union_qubits(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b);
size_t compress_bits(size_t index, const std::vector<size_t> &positions);
ComplexSquareMatrix superoperator_of(const ComplexSquareMatrix &u);

// Seeded generator owned by this project: xoshiro256** seeded by splitmix64,
// normals by the Marsaglia polar method.  Deterministic across platforms.
class Rng {
public:
  explicit Rng(uint64_t seed);
  uint64_t next_u64();
  double uniform();             // [0, 1), 53 random bits
  double normal();              // standard normal
  uint64_t below(uint64_t n);   // uniform integer in [0, n)
  double exponential(double rate);

private:
  uint64_t s_[4];
  bool have_spare_ = false;
  double spare_ = 0.0;
};

} // namespace cast
