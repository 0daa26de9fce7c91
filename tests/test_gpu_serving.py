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
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def serve(reqs, streams, threads=4, realtime=False, keep=True, budget=0):
    cfg = cs.ServeConfig()
    cfg.streams, cfg.compile_threads, cfg.realtime, cfg.keep_states = streams, threads, realtime, keep
    cfg.budget_bytes = budget
    print(f"[serve] {len(reqs)} requests, streams={streams}, realtime={realtime}", flush=True)
    res, st = cs.Server(cfg).run(reqs)
    return {r.id: r for r in res}, st


def oracle_state(r):
    g = cs.CircuitGraph.from_circuit(cs.random_layer_circuit(r.n_qubits, r.depth, r.seed))
    sv = np.zeros(2 << r.n_qubits)
    sv[0] = 1.0
    for gate in g.gates_in_row_order():
        oracle.apply_gate(sv, np.asarray(gate.matrix()), list(gate.qubits()))
    return sv


def test_results_are_the_same_alone_and_in_the_mix_and_match_the_oracle():
    reqs = cs.poisson_requests(48, 1000.0, 6, 16, 3, 20, 0.5, 7)
    alone, st1 = serve(reqs, streams=1, threads=1)
    mixed, st8 = serve(reqs, streams=8, threads=8)
    assert st1["max_in_flight"] == 1
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
