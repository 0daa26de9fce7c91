// OpenQASM 2.0 ingest for the gate set the simulator supports.
//
// Behaviour follows the reference's openqasm module:
//   parse_qasm           <- src/openqasm/parser.rs:58-62 (statements :116-300,
//                           operands :310-370, angle expressions :372-444,
//                           literals and comments :474-641, exact multiples of
//                           pi :643-787)
//   QasmCircuit          <- src/openqasm/circuit.rs:1-105 (gate list,
//                           required_qreg_size, serialize)
//   to_quantum_gate      <- src/circuit_graph.rs:506-541
//   circuit_from_qasm    <- src/types/circuit.rs:49-70 (n_qubits = highest
//                           qubit used + 1)
//   graph_from_qasm      <- src/circuit_graph.rs:130-138
//
// Accepted: an optional `OPENQASM x;` header and `include ...;` lines
// (skipped), at most one `qreg name[size];` (bounds are then checked),
// `//` comments, and the gates x y z h s t rx ry rz u3 cx cz cp swap ccx with
// operands `name[i]`, `name<i>` or a bare integer.  Angle expressions use
// + - * / ^, parentheses, integers, floats (with exponents) and pi; one that
// is exactly a rational multiple of pi is kept exact, so `2*pi/3` becomes
// (2/3) * pi in radians rather than the float evaluation of the expression.
#pragma once

#include "circuit.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace cast {

// An exact fraction in lowest terms with a positive denominator.
struct Fraction {
  int32_t num = 0;
  int32_t den = 1;
  static Fraction make(int64_t n, int64_t d); // reduces; throws on d == 0 or overflow
  bool operator==(const Fraction &o) const { return num == o.num && den == o.den; }
};

struct QasmAngle {
  bool pi_multiple = false; // true: value is frac * pi exactly
  Fraction frac;
  double number = 0.0;      // when !pi_multiple
  double radians() const;
  std::string to_qasm() const;
};

enum class QasmOp { X, Y, Z, H, S, T, RX, RY, RZ, U3, CX, CZ, CP, SWAP, CCX };

struct QasmGate {
  QasmOp op = QasmOp::X;
  std::vector<QasmAngle> params;
  std::vector<uint32_t> qubits; // operand order as written
  std::string name() const;
  std::string to_qasm() const;
};

struct QasmCircuit {
  std::vector<QasmGate> gates;
  uint32_t required_qreg_size() const; // highest qubit used + 1 (0 if none)
  std::string serialize() const;       // header, qreg q[required], one gate a line
};

// Throws cast::Error with a message naming the problem.
QasmCircuit parse_qasm(const std::string &text);
QuantumGate to_quantum_gate(const QasmGate &g);
QuantumCircuit circuit_from_qasm(const QasmCircuit &c);
CircuitGraph graph_from_qasm(const QasmCircuit &c);

} // namespace cast
