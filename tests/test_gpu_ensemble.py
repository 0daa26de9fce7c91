"""Device marginals, density-matrix simulation and ensemble trajectory
sampling against the oracle, plus the reference's own simulator tests
(src/simulator/tests.rs:22-58, 88-305) run on the GPU backend."""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs
import oracle
from helpers import rel_l2

pytestmark = pytest.mark.gpu

F64, F32 = cs.Precision.F64, cs.Precision.F32
TEST_QUBITS = 6


def graph_of(n, gates):
    c = cs.QuantumCircuit(n)
    for g in gates:
        c.add(g)
    return cs.CircuitGraph.from_circuit(c)


def uploaded(n, prec, seed):
    x = np.asarray(cs.random_statevector(n, seed))
    sv = cs.HipStatevector(n, prec)
    sv.upload(x)
    if prec == F32:
        x = x.astype(np.float32).astype(np.float64)
    return sv, x


# ── marginals ───────────────────────────────────────────────────────────────

# (n, measured): rest-major with several chunks, with one chunk, bin-major,
# every qubit, none.
CASES = [(12, [0]), (12, [11, 3]), (16, [5, 0, 9]), (20, [19]), (20, [2, 17, 4, 11, 0]),
         (14, list(range(5, 14))), (14, [13, 0, 7, 1, 12, 2, 11, 3]), (10, list(range(10))),
         (10, list(reversed(range(10)))), (11, []), (22, [21, 20, 0])]


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("n,measured", CASES)
def test_marginal_matches_oracle(n, measured, prec):
    print(f"[marginal] n={n} measured={measured} prec={prec}", flush=True)
    sv, x = uploaded(n, prec, 100 + n + len(measured))
    got = sv.marginal_probabilities(measured)
    want = oracle.marginal(x, measured)
    assert got.shape == want.shape
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15)


def test_marginal_at_26_qubits_sums_to_norm_and_matches_oracle():
    n = 26
    print("[marginal] n=26 f32 measured [25, 0, 13]", flush=True)
    sv, x = uploaded(n, F32, 7)
    got = sv.marginal_probabilities([25, 0, 13])
    assert abs(got.sum() - sv.norm_squared()) < 1e-12
    assert np.allclose(got, oracle.marginal(x, [25, 0, 13]), rtol=1e-11, atol=0)


def test_marginal_rejects_bad_qubits():
    sv = cs.HipStatevector(8, F64)
    sv.zero()
    with pytest.raises(cs.CastError, match="out of range"):
        sv.marginal_probabilities([8])
    with pytest.raises(cs.CastError, match="repeated"):
        sv.marginal_probabilities([1, 1])


# ── density matrix ──────────────────────────────────────────────────────────

def noisy_circuit(n, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for layer in range(4):
        for q in range(n):
            gates.append(cs.QuantumGate.random_unitary([q], int(rng.integers(1 << 30))))
            if rng.random() < 0.5:
                gates.append(cs.QuantumGate.depolarizing(q, float(rng.uniform(0.01, 0.2))))
        for q in range(layer % 2, n - 1, 2):
            gates.append(cs.QuantumGate.cx(q, q + 1))
    return gates


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("n", [3, 5, 7])
def test_density_matrix_matches_oracle(n, prec):
    print(f"[dm] n={n} prec={prec}", flush=True)
    gates = noisy_circuit(n, 40 + n)
    sim = cs.Simulator(cs.HipKernelGenSpec.f64() if prec == F64 else cs.HipKernelGenSpec.f32())
    st = sim.simulate(graph_of(n, gates), cs.Representation.DensityMatrix)
    assert not st.is_pure() and st.n_qubits() == n
    lifted = []
    for g in gates:
        d = g.to_density_matrix_gate(n)
        lifted.append((np.asarray(d.matrix()), list(d.qubits())))
    rho = oracle.simulate_dm(n, lifted)
    tol = 1e-12 if prec == F64 else 1e-5
    assert rel_l2(rho, st.download_amplitudes()) < tol
    pops = st.populations()
    assert np.allclose(pops, oracle.dm_populations(rho, n), rtol=0, atol=tol)
    assert abs(st.trace() - 1.0) < tol


def test_reference_statevector_cases():
    sim = cs.Simulator.f64()
    st = sim.simulate(graph_of(TEST_QUBITS, [cs.QuantumGate.h(0)]), cs.Representation.StateVector)
    assert st.is_pure()                                                  # tests.rs:22-32
    pops = st.populations()
    assert abs(pops[0] - 0.5) < 1e-10 and abs(pops[1] - 0.5) < 1e-10
    with pytest.raises(cs.CastError, match="non-unitary"):              # tests.rs:34-39
        sim.simulate(graph_of(TEST_QUBITS, [cs.QuantumGate.h(0),
                                            cs.QuantumGate.depolarizing(0, 0.1)]),
                     cs.Representation.StateVector)
    st = sim.simulate(graph_of(TEST_QUBITS, [cs.QuantumGate.h(0), cs.QuantumGate.cx(0, 1),
                                             cs.QuantumGate.depolarizing(0, 0.05)]),
                      cs.Representation.DensityMatrix)                   # tests.rs:41-58
    assert not st.is_pure() and abs(st.trace() - 1.0) < 1e-10


# ── ensemble trajectories ───────────────────────────────────────────────────

def opts(measured, n_samples, M, seed=42):
    return cs.TrajectoryOpts(measured, n_samples, seed, M)


def ens_gates(gates):
    out = []
    for g in gates:
        if g.is_unitary():
            out.append(("u", np.asarray(g.matrix()), list(g.qubits())))
        else:
            out.append(("noise", [(p, np.asarray(k)) for p, k in g.noise_branches()],
                        list(g.qubits())))
    return out


def test_reference_trajectory_cases():
    sim = cs.Simulator.f64()
    H, D, CX = cs.QuantumGate.h, cs.QuantumGate.depolarizing, cs.QuantumGate.cx
    t = sim.sample_trajectory(graph_of(TEST_QUBITS, [H(0)]), opts([0], 10000, 1))
    assert t.n_samples == 10000 and sum(t.histogram.values()) == 10000  # tests.rs:88-108
    assert abs(t.histogram.get(0, 0) - 5000) < 500 and abs(t.histogram.get(1, 0) - 5000) < 500
    assert abs(t.explored_weight - 1.0) < 1e-10

    g = graph_of(TEST_QUBITS, [H(0), D(0, 0.1)])
    t = sim.sample_trajectory(g, opts([0], 100, 1))                     # tests.rs:110-123
    assert len(t.branches) == 1 and t.branches[0].noise_path == [0]
    assert abs(t.branches[0].weight - 0.9) < 1e-10 and abs(t.explored_weight - 0.9) < 1e-10
    t1 = sim.sample_trajectory(g, opts([0], 1000, 1))                   # tests.rs:125-150
    t4 = sim.sample_trajectory(g, opts([0], 1000, 4))
    assert len(t1.branches) == 1 and len(t4.branches) == 4
    assert t4.explored_weight > t1.explored_weight and abs(t4.explored_weight - 1) < 1e-10

    t = sim.sample_trajectory(graph_of(TEST_QUBITS, [H(0), D(0, 0.1), CX(0, 1), D(1, 0.1)]),
                              opts([0, 1], 10000, 2))                    # tests.rs:189-215
    assert len(t.branches) == 2 and t.branches[0].weight >= t.branches[1].weight
    assert 0.7 < t.explored_weight < 1.0 and sum(t.histogram.values()) == 10000

    t = sim.sample_trajectory(graph_of(TEST_QUBITS, [H(0), CX(0, 1)]), opts([0, 1], 1000, 10))
    assert len(t.branches) == 1 and abs(t.branches[0].weight - 1) < 1e-10  # tests.rs:253-270
    assert sorted(t.histogram) == [0, 3]


@pytest.mark.parametrize("gates,measured,N,seed,tol", [
    ("h d cx", [0, 1], 100_000, 99, 0.02),                               # tests.rs:152-187
    ("h d cx d d", [0, 1], 200_000, 123, 0.01),                          # tests.rs:272-305
])
def test_reference_ensemble_matches_density_matrix(gates, measured, N, seed, tol):
    H, D, CX = cs.QuantumGate.h, cs.QuantumGate.depolarizing, cs.QuantumGate.cx
    p = 0.1 if gates == "h d cx" else 0.05
    gl = ([H(0), D(0, p), CX(0, 1)] if gates == "h d cx"
          else [H(0), D(0, p), CX(0, 1), D(0, p), D(1, p)])
    sim = cs.Simulator.f64()
    g = graph_of(2, gl)
    pops = sim.simulate(g, cs.Representation.DensityMatrix).populations()
    t = sim.sample_trajectory(g, opts(measured, N, 64, seed))
    assert abs(t.explored_weight - 1.0) < 1e-10
    for k, want in enumerate(pops):
        assert abs(t.histogram.get(k, 0) / N - want) < tol


@pytest.mark.parametrize("M", [1, 3, 8, 64])
def test_ensemble_matches_oracle_ensemble(M):
    n = 8
    rng = np.random.default_rng(M)
    gl = []
    for layer in range(3):
        for q in range(n):
            gl.append(cs.QuantumGate.random_unitary([q], int(rng.integers(1 << 30))))
        for q in range(layer % 2, n - 1, 2):
            gl.append(cs.QuantumGate.cz(q, q + 1))
        gl.append(cs.QuantumGate.depolarizing(int(rng.integers(n)), 0.08))
    measured = [0, 5, 3]
    N = 50_000
    print(f"[ensemble] n={n} M={M}", flush=True)
    t = cs.Simulator.f64().sample_trajectory(graph_of(n, gl), opts(measured, N, M, 7))
    want = oracle.ensemble(n, ens_gates(gl), M)
    assert len(t.branches) == len(want)
    for b, (w, path, _) in zip(t.branches, want):
        assert abs(b.weight - w) < 1e-12 and b.noise_path == path
    counts = oracle.samples_per_member([w for w, _, _ in want], N)
    assert [b.n_samples for b in t.branches] == counts
    assert sum(t.histogram.values()) == N
    # frequencies against the sample-weighted mixture of the members' marginals
    mix = sum(c * oracle.marginal(x, measured) for c, (_, _, x) in zip(counts, want)) / N
    obs = np.array([t.histogram.get(k, 0) for k in range(8)]) / N
    assert np.abs(obs - mix).max() < 5 * np.sqrt(0.25 / N)
    again = cs.Simulator.f64().sample_trajectory(graph_of(n, gl), opts(measured, N, M, 7))
    assert again.histogram == t.histogram


def test_bench_times_density_matrix_runs():
    g = graph_of(5, noisy_circuit(5, 3))
    sim = cs.Simulator.f64()
    print("[dm] bench n=5", flush=True)
    r = sim.bench(g, 0.05, cs.Representation.DensityMatrix)
    assert r["n_gates"] == len(g.gates_in_row_order()) and len(r["samples_s"]) >= 1
    assert all(t > 0 for t in r["samples_s"])
    r2 = sim.bench(graph_of(5, [cs.QuantumGate.h(0)]), 0.02)
    assert r2["n_gates"] == 1


@pytest.mark.parametrize("n", [1, 3, 5])
def test_marginal_small_statevectors(n):
    sv, x = uploaded(n, F64, 500 + n)
    print(f"[marginal] small n={n}", flush=True)
    for measured in ([], [0], list(range(n)), list(reversed(range(n)))):
        got = sv.marginal_probabilities(measured)
        assert np.allclose(got, oracle.marginal(x, measured), rtol=1e-13, atol=1e-16)


def test_density_matrix_of_one_qubit():
    g = graph_of(1, [cs.QuantumGate.h(0), cs.QuantumGate.depolarizing(0, 0.3)])
    st = cs.Simulator.f64().simulate(g, cs.Representation.DensityMatrix)
    assert np.allclose(st.populations(), [0.5, 0.5], atol=1e-14)
    amps = st.download_amplitudes()
    # rho = 0.5 I + (1 - 4p/3) / 2 X for H|0> under depolarizing p
    rho01 = amps[2 * (0 | (1 << 1))]
    assert abs(rho01 - 0.5 * (1 - 4 * 0.3 / 3)) < 1e-14


def test_trajectory_edge_cases():
    sim = cs.Simulator.f64()
    g = graph_of(3, [cs.QuantumGate.h(0), cs.QuantumGate.depolarizing(0, 0.2)])
    t = sim.sample_trajectory(g, opts([], 500, 4))
    assert t.histogram == {0: 500}                        # no measured qubits: one outcome
    t = sim.sample_trajectory(g, opts([0, 2], 0, 4))
    assert t.histogram == {} and [b.n_samples for b in t.branches] == [0, 0, 0, 0]
    with pytest.raises(cs.CastError, match="out of range"):
        sim.sample_trajectory(g, opts([3], 10, 1))
