"""QASM-ingested circuits on the GPU: the parsed, fused circuit's final state
against the oracle, and the QASM benchmark tool end to end."""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs
import oracle
from cast_simulator_amd import bench_qasm
from helpers import rel_l2

pytestmark = pytest.mark.gpu

SRC = """OPENQASM 2.0;
include "qelib1.inc";
qreg q[12];
// a mixed circuit touching every supported gate kind
h q[0]; x q[1]; y q[2]; z q[3]; s q[4]; t q[5];
rx(pi/3) q[6]; ry(-0.7) q[7]; rz(2*pi/5) q[8]; u3(0.1,pi/7,-0.3) q[9];
cx q[0],q[11]; cz q[1],q[10]; cp(pi/8) q[2],q[9]; swap q[3],q[8]; ccx q[4],q[7],q[5];
h q[10]; h q[11]; cx q[11],q[6]; rz(1e-3) q[6]; cp(-3*pi/4) q[6],q[0];
"""


@pytest.mark.parametrize("prec", [cs.Precision.F64, cs.Precision.F32])
@pytest.mark.parametrize("fusion", ["none", "size-only", "hw-adaptive"])
def test_qasm_circuit_matches_oracle(fusion, prec):
    print(f"[qasm] fusion={fusion} prec={prec}", flush=True)
    parsed = cs.parse_qasm(SRC)
    g = cs.CircuitGraph.from_qasm_circuit(parsed)
    pname = "f64" if prec == cs.Precision.F64 else "f32"
    prof = cs.HardwareProfile.load(str(bench_qasm.PROFILE_DIR / f"mi355x_{pname}.json"))
    cs.optimize(g, bench_qasm.fusion_config(cs, fusion, prof, 5, 1e-12))
    spec = cs.HipKernelGenSpec.f64() if prec == cs.Precision.F64 else cs.HipKernelGenSpec.f32()
    sv = cs.Simulator(spec).simulate(g)
    want = np.zeros(2 << 12)
    want[0] = 1.0
    for gate in parsed.gates:
        qg = cs.qasm_to_quantum_gate(gate)
        oracle.apply_gate(want, np.asarray(qg.matrix()), list(qg.qubits()))
    tol = 1e-12 if prec == cs.Precision.F64 else 1e-5
    assert rel_l2(want, sv.download()) < tol


def test_bench_tool_runs_every_mode(tmp_path, capsys):
    f = tmp_path / "mixed.qasm"
    f.write_text(SRC)
    print("[qasm] bench tool", flush=True)
    assert bench_qasm.main([str(f), "--bench-budget", "0.2", "--max-size", "4"]) == 0
    out = capsys.readouterr().out
    for mode in bench_qasm.MODES:
        assert f"mixed            {mode}" in out
