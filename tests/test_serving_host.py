"""Serving-loop request generator (CPU): seeded Poisson arrivals and the
uniform qubit-count, depth and precision draws."""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs


def test_poisson_requests_are_seeded_and_in_range():
    a = cs.poisson_requests(3000, 200.0, 12, 28, 10, 50, 0.5, 99)
    b = cs.poisson_requests(3000, 200.0, 12, 28, 10, 50, 0.5, 99)
    assert [(r.arrival_s, r.n_qubits, r.depth, r.seed, r.precision) for r in a] == \
           [(r.arrival_s, r.n_qubits, r.depth, r.seed, r.precision) for r in b]
    t = np.array([r.arrival_s for r in a])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    # exponential gaps of mean 1/200 s: mean within 5 standard errors
    assert abs(gaps.mean() - 1 / 200) < 5 * (1 / 200) / np.sqrt(len(gaps))
    n = np.array([r.n_qubits for r in a])
    d = np.array([r.depth for r in a])
    assert n.min() == 12 and n.max() == 28 and d.min() == 10 and d.max() == 50
    f32 = np.mean([r.precision == cs.Precision.F32 for r in a])
    assert abs(f32 - 0.5) < 5 * 0.5 / np.sqrt(len(a))
    assert [r.id for r in a] == list(range(3000))
    assert len({r.seed for r in a}) == 3000
    c = cs.poisson_requests(10, 200.0, 12, 28, 10, 50, 0.5, 100)
    assert [r.seed for r in c] != [r.seed for r in a[:10]]


@pytest.mark.parametrize("args", [(5, 0.0, 12, 28, 10, 50, 0.5, 1), (5, 1.0, 1, 28, 10, 50, 0.5, 1),
                                  (5, 1.0, 12, 40, 10, 50, 0.5, 1), (5, 1.0, 12, 28, 0, 50, 0.5, 1),
                                  (5, 1.0, 12, 28, 10, 5, 0.5, 1), (5, 1.0, 12, 28, 10, 50, 1.5, 1)])
def test_poisson_requests_reject_bad_parameters(args):
    with pytest.raises(cs.CastError):
        cs.poisson_requests(*args)


def test_circuit_requests_take_their_qubit_count_from_the_circuit():
    qasm = 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[9];\nh q[0];\ncx q[0],q[5];\n'
    a = cs.ServeRequest.from_qasm(3, qasm, cs.Precision.F32, 0.25)
    assert (a.id, a.n_qubits, a.depth, a.precision, a.arrival_s) == (3, 6, 2, cs.Precision.F32, 0.25)
    assert a.has_circuit
    b = cs.ServeRequest.from_circuit(4, cs.QuantumCircuit.qft(7), cs.Precision.F64)
    assert b.n_qubits == 7 and b.depth == 7 + 21 + 3 and b.has_circuit
    c = b.with_precision(cs.Precision.F32)
    assert c.precision == cs.Precision.F32 and c.has_circuit and b.precision == cs.Precision.F64
    g = cs.ServeRequest(0, 12, 10, 1, cs.Precision.F64)
    assert not g.has_circuit
    with pytest.raises(cs.CastError):
        cs.ServeRequest.from_qasm(0, "qreg q[2];\nfoo q[0];\n", cs.Precision.F64)


def test_serve_config_defaults_to_one_stream_per_hardware_queue():
    cfg = cs.ServeConfig()
    assert cfg.streams == 4 and cfg.f64_shadow is False
    # small requests in batches of up to 32 by default
    assert cfg.batch_max_qubits == 18 and cfg.batch_size == 32
