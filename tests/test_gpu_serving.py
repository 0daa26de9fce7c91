"""Continuous-batching serving loop on the GPU.

Each request's final state and f64 reductions must be the same alone (one
stream, one request at a time) and in the mix (many streams, overlapping
requests), and match the oracle; admission must respect the statevector
memory budget."""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs
import oracle
from helpers import fidelity, rel_l2

pytestmark = pytest.mark.gpu


def serve(reqs, streams, threads=4, realtime=False, keep=True, budget=0, shadow=False,
          batch_size=None, batch_max_qubits=None):
    cfg = cs.ServeConfig()
    cfg.streams, cfg.compile_threads, cfg.realtime, cfg.keep_states = streams, threads, realtime, keep
    cfg.budget_bytes = budget
    cfg.f64_shadow = shadow
    if batch_size is not None:
        cfg.batch_size = batch_size
    if batch_max_qubits is not None:
        cfg.batch_max_qubits = batch_max_qubits
    print(f"[serve] {len(reqs)} requests, streams={streams}, realtime={realtime}, "
          f"batch {cfg.batch_size} up to {cfg.batch_max_qubits} qubits", flush=True)
    res, st = cs.Server(cfg).run(reqs)
    return {r.id: r for r in res}, st


def oracle_state(r, circuit=None):
    """The oracle's state of a request: its unfused circuit on |0...0>."""
    if circuit is None:
        circuit = cs.random_layer_circuit(r.n_qubits, r.depth, r.seed)
    sv = np.zeros(2 << r.n_qubits)
    sv[0] = 1.0
    for gate in circuit.gates():
        oracle.apply_gate(sv, np.asarray(gate.matrix()), list(gate.qubits()))
    return sv


@pytest.mark.parametrize("batch_max", [0, 18], ids=["per-request", "batch-lane"])
def test_results_are_the_same_alone_and_in_the_mix_and_match_the_oracle(batch_max):
    """Alone: one stream, one request at a time (batches of one in the batch
    lane); mixed: 8 streams, and batches of up to 32 small requests."""
    reqs = cs.poisson_requests(48, 1000.0, 6, 16, 3, 20, 0.5, 7)
    alone, st1 = serve(reqs, streams=1, threads=1, batch_size=1, batch_max_qubits=batch_max)
    mixed, st8 = serve(reqs, streams=8, threads=8, batch_max_qubits=batch_max)
    assert st1["max_in_flight"] == 1
    if batch_max:
        assert st1["batched_requests"] == st8["batched_requests"] == len(reqs)
        assert st8["batches"] < len(reqs)  # several requests per batch in the mix
        # one launch per step and gate size, not one per gate per request
        assert st8["batch_launches"] < sum(r.n_gates for r in mixed.values())
    else:
        assert st1["batches"] == st8["batches"] == 0
    for r in reqs:
        a, m = alone[r.id], mixed[r.id]
        assert a.n_gates == m.n_gates and a.n_gates > 0
        assert np.array_equal(a.state(), m.state()), f"request {r.id}"
        assert a.norm_sq == m.norm_sq and list(a.zz) == list(m.zz)
        tol = 1e-12 if r.precision == cs.Precision.F64 else 1e-5
        want = oracle_state(r)
        assert rel_l2(want, m.state()) < tol, f"request {r.id}"
        assert abs(m.norm_sq - oracle.norm_squared(np.array(m.state()))) < 1e-12
        zz = oracle.expect_zz(np.array(m.state()), [(i, i + 1) for i in range(r.n_qubits - 1)])
        assert np.allclose(m.zz, zz, rtol=0, atol=1e-12)


def test_realtime_arrivals_overlap_and_latencies_are_consistent():
    reqs = cs.poisson_requests(60, 400.0, 14, 22, 10, 30, 0.5, 11)
    res, st = serve(reqs, streams=8, threads=8, realtime=True, keep=False)
    assert st["max_in_flight"] > 1
    for r in reqs:
        o = res[r.id]
        assert o.arrival_s <= o.compiled_s + 1e-6
        assert o.compiled_s <= o.admit_s + 1e-6 and o.admit_s <= o.done_s
        assert abs(o.norm_sq - 1.0) < (1e-10 if r.precision == cs.Precision.F64 else 1e-4)
    lat = sorted(o.done_s - o.arrival_s for o in res.values())
    assert st["p50_s"] <= st["p99_s"] <= st["max_s"] + 1e-12
    assert st["max_s"] == pytest.approx(lat[-1])
    assert st["throughput"] == pytest.approx(len(reqs) / st["wall_s"])


def test_admission_respects_the_memory_budget():
    reqs = cs.poisson_requests(40, 1000.0, 14, 16, 4, 8, 0.0, 5)  # all f64
    budget = 2 * (2 << 16) * 8  # two 16-qubit f64 statevectors
    res, st = serve(reqs, streams=8, threads=8, keep=False, budget=budget)
    assert st["peak_sv_bytes"] <= budget
    assert st["budget_bytes"] == budget
    assert len(res) == len(reqs)
    assert st["sv_reuses"] > 0


def test_requests_larger_than_the_budget_are_rejected():
    reqs = [cs.ServeRequest(0, 18, 4, 1, cs.Precision.F64, 0.0)]
    with pytest.raises(cs.CastError, match="budget"):
        serve(reqs, streams=2, keep=False, budget=(2 << 17) * 8)


QASM = """OPENQASM 2.0;
include "qelib1.inc";
qreg q[7];
h q[0];
cx q[0],q[1];
rz(pi/4) q[1];
ccx q[0],q[1],q[2];
u3(0.3,0.2,0.1) q[5];
swap q[2],q[6];
cp(pi/8) q[3],q[5];
ry(-0.7) q[4];
cz q[4],q[6];
"""


def test_circuit_and_qasm_requests_match_the_oracle():
    """Requests carry any unitary circuit (a QFT, an OpenQASM program) beside
    the generated random-layer ones; each is fused, served in the mix and
    matches the oracle run of its unfused circuit."""
    qft = cs.QuantumCircuit.qft(10)
    prog = cs.QuantumCircuit.from_qasm(QASM)
    reqs = [cs.ServeRequest.from_circuit(0, qft, cs.Precision.F64),
            cs.ServeRequest.from_circuit(1, qft, cs.Precision.F32, 0.001),
            cs.ServeRequest.from_qasm(2, QASM, cs.Precision.F64, 0.002),
            cs.ServeRequest.from_qasm(3, QASM, cs.Precision.F32, 0.003)]
    extra = cs.poisson_requests(12, 1000.0, 6, 14, 3, 12, 0.5, 21)
    for r in extra:
        r.id += 10
    circuits = {0: qft, 1: qft, 2: prog, 3: prog}
    assert reqs[2].n_qubits == 7 and reqs[0].has_circuit and not extra[0].has_circuit
    res, st = serve(reqs + extra, streams=4, threads=4)
    assert st["max_in_flight"] <= 4 * 32  # 4 streams, batches of up to 32 small requests
    for r in reqs + extra:
        o = res[r.id]
        want = oracle_state(r, circuits.get(r.id))
        tol = 1e-12 if r.precision == cs.Precision.F64 else 1e-5
        assert rel_l2(want, o.state()) < tol, f"request {r.id}"
        assert len(o.zz) == r.n_qubits - 1


def test_requests_with_noise_or_a_wrong_qubit_count_are_rejected():
    c = cs.QuantumCircuit(3)
    c.add(cs.QuantumGate.h(0))
    c.add(cs.QuantumGate.depolarizing(1, 0.1))
    with pytest.raises(cs.CastError, match="noisy"):
        serve([cs.ServeRequest.from_circuit(0, c, cs.Precision.F64)], streams=1)
    r = cs.ServeRequest.from_qasm(0, QASM, cs.Precision.F64)
    r.n_qubits = 9
    with pytest.raises(cs.CastError, match="n_qubits"):
        serve([r], streams=1)


def test_f64_shadow_gives_the_fidelity_of_every_f32_request():
    """Checking mode: each F32 request also runs at F64 on the device; the
    device fidelity equals the host fidelity of the F32 result against the
    oracle's f64 state, is >= 1 - 1e-6, and the F32 results are bit-identical
    to a run without the shadow."""
    reqs = cs.poisson_requests(24, 1000.0, 8, 16, 3, 20, 0.5, 13)
    plain, _ = serve(reqs, streams=4)
    res, st = serve(reqs, streams=4, shadow=True)
    n32 = 0
    for r in reqs:
        o = res[r.id]
        assert np.array_equal(o.state(), plain[r.id].state()), f"request {r.id}"
        if r.precision == cs.Precision.F32:
            n32 += 1
            f_host = fidelity(oracle_state(r), o.state())
            assert abs(o.fidelity_vs_f64 - f_host) < 1e-9, (o.fidelity_vs_f64, f_host)
            assert o.fidelity_vs_f64 >= 1.0 - 1e-6
        else:
            assert o.fidelity_vs_f64 == -1.0
    assert n32 >= 6


def test_34_qubit_f32_request_reduces_every_neighbour_pair():
    """A 34-qubit f32 request (128 GiB) has 33 neighbour pairs, reduced in two
    chunks.  RY(theta_i) on every qubit gives <Z_i Z_i+1> =
    cos(theta_i) cos(theta_i+1) exactly."""
    n = 34
    from cast_simulator_amd import abi

    free = abi.free_memory()[0]
    if free < (2 << n) * 4 + (10 << 30):
        pytest.skip(f"needs 138 GiB of device memory, {free / 2**30:.0f} GiB free")
    theta = 0.1 + 0.045 * np.arange(n)
    c = cs.QuantumCircuit(n)
    for q in range(n):
        c.add(cs.QuantumGate.ry(float(theta[q]), q))
    res, st = serve([cs.ServeRequest.from_circuit(0, c, cs.Precision.F32)], streams=1, keep=False)
    o = res[0]
    want = np.cos(theta[:-1]) * np.cos(theta[1:])
    assert len(o.zz) == n - 1
    assert abs(o.norm_sq - 1.0) < 1e-5
    assert np.allclose(np.asarray(o.zz) / o.norm_sq, want, rtol=0, atol=1e-5)


def test_batch_lane_with_f64_shadow_and_large_requests_beside_it():
    """Small requests (n <= 18) in batches, larger ones one per stream, in one
    run with the F64 replay of every F32 request: each state matches the
    oracle and every F32 request has a fidelity >= 1 - 1e-6; the batch lane
    and the per-request path give the same large-request states bit for bit."""
    reqs = cs.poisson_requests(40, 1000.0, 10, 20, 4, 12, 0.5, 23)
    res, st = serve(reqs, streams=4, threads=8, shadow=True)
    small = [r for r in reqs if r.n_qubits <= 18]
    assert st["batched_requests"] == len(small) > 0 and st["batches"] >= 1
    ref, _ = serve([r for r in reqs if r.n_qubits > 18], streams=1, threads=4,
                   batch_max_qubits=0)
    for r in reqs:
        o = res[r.id]
        tol = 1e-12 if r.precision == cs.Precision.F64 else 1e-5
        assert rel_l2(oracle_state(r), o.state()) < tol, f"request {r.id}"
        if r.precision == cs.Precision.F32:
            assert o.fidelity_vs_f64 >= 1 - 1e-6, (r.id, o.fidelity_vs_f64)
        else:
            assert o.fidelity_vs_f64 == -1.0
        if r.n_qubits > 18:
            assert np.array_equal(o.state(), ref[r.id].state()), f"request {r.id}"
