"""OpenQASM 2.0 ingest on the CPU: the reference's own parser, serialiser and
graph-conversion tests (src/openqasm/parser.rs:790-929,
src/openqasm/circuit.rs:185-262, src/circuit_graph.rs:633-690,
src/simulator/tests.rs:57-68), plus round trips and exact angles."""
from __future__ import annotations

import math

import numpy as np
import pytest

import cast_simulator_amd as cs


def kinds(c):
    return [g.name() for g in c.gates]


def test_parses_program_without_qreg_declaration():
    c = cs.parse_qasm("x q0; // comment\nrz(2*pi/3) q1;\ncx q0,q1;")
    assert kinds(c) == ["x", "rz", "cx"]


def test_parses_openqasm_include_and_qreg_preamble():
    src = ('OPENQASM 2.0; include "qelib1.inc"; qreg q[3]; h q[0]; rx(pi/2) q[1]; '
           'ccx q[0],q[1],q[2];')
    assert kinds(cs.parse_qasm(src)) == ["h", "rx", "ccx"]


def test_parses_parametric_gates_and_angles():
    c = cs.parse_qasm("qreg q[2]; rx(2*pi/3) q[0]; u3(pi/2,0.25,1) q[1];")
    assert len(c.gates) == 2
    a = c.gates[0].params[0]
    assert c.gates[0].qubits == [0]
    assert a.pi_multiple and (a.numerator, a.denominator) == (2, 3)
    assert c.gates[1].name() == "u3" and c.gates[1].qubits == [1]


def test_parses_scientific_notation_angles():
    c = cs.parse_qasm("qreg q[2]; rx(1e-6) q[0]; rz(0.3e-1) q[1];")
    a, b = c.gates[0].params[0], c.gates[1].params[0]
    assert not a.pi_multiple and abs(a.number - 1e-6) < 1e-16
    assert not b.pi_multiple and abs(b.number - 3e-2) < 1e-16
    assert c.gates[1].qubits == [1]


@pytest.mark.parametrize("src,msg", [
    ("qreg q[2]; qreg r[3]; x q[0];", "multiple qreg declarations"),
    ("qreg q[2]; x q[2];", "out of bounds"),
    ("qreg q[2]; x r[0];", "unknown qreg"),
    ("qreg q[2]; foo q[0];", "unsupported gate"),
    ("x q[0]", "expected `;`"),
    ("rx(1e) q[0];", "invalid scientific literal"),
    ("rx(pi/0) q[0];", "division by zero"),
    ("rx(pi/(pi-pi)) q[0];", "division by zero"),
    ("x foo;", "expected qubit like"),
    ("OPENQASM 2.0", "terminate OPENQASM header"),
])
def test_parse_errors(src, msg):
    with pytest.raises(cs.CastError, match=msg):
        cs.parse_qasm(src)


@pytest.mark.parametrize("expr,exact,value", [
    ("pi", (1, 1), math.pi),
    ("-pi/4", (-1, 4), -math.pi / 4),
    ("pi*3/4", (3, 4), 0.75 * math.pi),
    ("(1+1)*pi/6", (1, 3), math.pi / 3),
    ("pi/2 - pi/3", (1, 6), math.pi / 6),
    ("2*pi/pi*pi", (2, 1), 2 * math.pi),     # (2 pi / pi) = 2, times pi
    ("0", None, 0.0),                         # rational, not a multiple of pi
    ("pi*pi", None, math.pi * math.pi),
    ("pi^2", None, math.pi ** 2),
    ("0.5*pi", None, 0.5 * math.pi),          # float literals are never exact
    ("1/pi", None, 1 / math.pi),
    ("0.5 + 1/0", None, math.inf),            # the symbolic pass stops at 0.5
    ("-(pi)", (-1, 1), -math.pi),
    ("2^3", None, 8.0),
])
def test_angle_expressions(expr, exact, value):
    a = cs.parse_qasm(f"rz({expr}) q[0];").gates[0].params[0]
    if exact is None:
        assert not a.pi_multiple and a.number == value
    else:
        assert a.pi_multiple and (a.numerator, a.denominator) == exact
        # exact multiples are converted as (n / d) * pi (circuit_graph.rs:550-552)
        assert a.radians() == exact[0] / exact[1] * math.pi


def test_serializes_with_header_and_inferred_qreg_size():
    c = cs.parse_qasm("x q[3]; h q[0];")
    q = c.serialize()
    assert q.startswith("OPENQASM 2.0;\nqreg q[4];\n")
    assert "x q[3];" in q and "h q[0];" in q


def test_serializes_and_round_trips_through_parser():
    c = cs.parse_qasm("rx(2*pi/3) q[0]; cx q[0],q[1]; u3(pi/2,0.25,1e-6) q[1];")
    s = c.serialize()
    assert "rx(2*pi/3) q[0];" in s and "cx q[0],q[1];" in s and "u3(pi/2,0.25,0.000001) q[1];" in s
    back = cs.parse_qasm(s)
    assert kinds(back) == ["rx", "cx", "u3"]
    assert back.gates[1].qubits == [0, 1] and back.gates[2].qubits == [1]
    assert back.serialize() == s


def test_number_formatting_matches_positional_shortest():
    s = cs.parse_qasm("rz(1e-6) q[2]; rz(0.1) q[0]; rz(2.0) q[0]; rz(-2.5e3) q[1];").serialize()
    assert s.startswith("OPENQASM 2.0;\nqreg q[3];\n")
    assert "rz(0.000001) q[2];\n" in s and "rz(0.1) q[0];" in s and "rz(2) q[0];" in s
    assert "rz(-2500) q[1];" in s


def same_gate(a, b):
    return list(a.qubits()) == list(b.qubits()) and np.array_equal(a.matrix(), b.matrix())


def test_graph_from_qasm_builds_packed_rows():
    g = cs.CircuitGraph.from_qasm_circuit(cs.parse_qasm("x q[0]; h q[1]; cx q[0],q[1];"))
    assert g.n_qubits() == 2 and g.n_rows() == 2
    gates = g.gates_in_row_order()
    for got, want in zip(gates, [cs.QuantumGate.x(0), cs.QuantumGate.h(1),
                                 cs.QuantumGate.cx(0, 1)]):
        assert same_gate(got, want)


def test_graph_from_qasm_sorts_operands_and_converts_angles():
    g = cs.CircuitGraph.from_qasm_circuit(cs.parse_qasm("cx q[2],q[0];"))
    assert g.n_qubits() == 3 and list(g.gates_in_row_order()[0].qubits()) == [0, 2]
    assert same_gate(g.gates_in_row_order()[0], cs.QuantumGate.cx(2, 0))
    g = cs.CircuitGraph.from_qasm_circuit(cs.parse_qasm("rz(0.25) q[1];"))
    assert same_gate(g.gates_in_row_order()[0], cs.QuantumGate.rz(0.25, 1))


def test_circuit_from_qasm_reference_case():
    c = cs.QuantumCircuit.from_qasm("OPENQASM 2.0; qreg q[6]; h q[0]; cx q[0],q[1]; cx q[4],q[5];")
    assert c.n_qubits() == 6 and len(c) == 3


def test_every_gate_kind_converts():
    src = ("x q[0]; y q[1]; z q[2]; h q[0]; s q[1]; t q[2]; rx(0.1) q[0]; ry(pi/5) q[1]; "
           "rz(-0.3) q[2]; u3(0.1,0.2,0.3) q[0]; cx q[0],q[1]; cz q[1],q[2]; cp(pi/8) q[2],q[0]; "
           "swap q[0],q[2]; ccx q[2],q[0],q[1];")
    c = cs.parse_qasm(src)
    want = [cs.QuantumGate.x(0), cs.QuantumGate.y(1), cs.QuantumGate.z(2), cs.QuantumGate.h(0),
            cs.QuantumGate.s(1), cs.QuantumGate.t(2), cs.QuantumGate.rx(0.1, 0),
            cs.QuantumGate.ry(math.pi / 5, 1), cs.QuantumGate.rz(-0.3, 2),
            cs.QuantumGate.u3(0.1, 0.2, 0.3, 0), cs.QuantumGate.cx(0, 1), cs.QuantumGate.cz(1, 2),
            cs.QuantumGate.cp(math.pi / 8, 2, 0), cs.QuantumGate.swap(0, 2),
            cs.QuantumGate.ccx(2, 0, 1)]
    for g, w in zip(c.gates, want):
        assert same_gate(cs.qasm_to_quantum_gate(g), w), g.to_qasm()


def test_qft_round_trips_through_qasm_exactly():
    n = 7
    qft = cs.QuantumCircuit.qft(n, True)
    lines = ["OPENQASM 2.0;", f"qreg q[{n}];"]
    for g in qft.gates():
        q = list(g.qubits())
        m = np.asarray(g.matrix())
        if len(q) == 1:
            lines.append(f"h q[{q[0]}];")
        elif abs(m[1, 2]) == 1:   # swap
            lines.append(f"swap q[{q[0]}],q[{q[1]}];")
        else:                      # cp(pi / 2^k)
            k = round(math.log2(math.pi / np.angle(m[3, 3])))
            lines.append(f"cp(pi/{2 ** k}) q[{q[1]}],q[{q[0]}];")
    c = cs.QuantumCircuit.from_qasm("\n".join(lines))
    assert len(c) == len(qft)
    for a, b in zip(c.gates(), qft.gates()):
        assert same_gate(a, b)
