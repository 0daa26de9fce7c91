"""Host layer (C++ gate types, circuit graph, fusion, cost model, profiler)
against the reference's own unit tests and golden values.  CPU only."""
from __future__ import annotations

import json

import numpy as np
import pytest

import cast_simulator_amd as cs
import oracle
from helpers import S2, brute_force_apply, gate_list, seeded_state, to_complex


def mat(g):
    return np.asarray(g.matrix())


# ── gate.rs tests :542-660 ─────────────────────────────────────────────────

def test_canonicalizes_qubit_order():
    got = cs.QuantumGate.from_reals([1, 0], [0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0])
    want = cs.QuantumGate.from_reals([0, 1], [0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 1, 0, 0])
    assert got == want


def test_rejects_duplicate_qubits():
    with pytest.raises(cs.CastError, match="gate qubits must be distinct"):
        cs.QuantumGate(np.eye(4), [0, 0])


def test_rejects_wrong_matrix_size():
    with pytest.raises(cs.CastError, match="matrix dimension does not match"):
        cs.QuantumGate(np.eye(2), [0, 1])


def x0():
    return cs.QuantumGate.from_reals([0], [0, 1, 1, 0])


def test_matmul_involution():
    assert np.array_equal(mat(x0().matmul(x0())), np.eye(2))


def test_matmul_disjoint_qubits():
    prod = cs.QuantumGate.from_reals([1], [0, 1, 1, 0]).matmul(x0())
    want = cs.QuantumGate.from_reals([0, 1], [0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 0])
    assert prod == want


def test_matmul_non_adjacent_qubits():
    prod = x0().matmul(cs.QuantumGate.from_reals([2], [0, 1, 1, 0]))
    want = cs.QuantumGate.from_reals([0, 2], [0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 0])
    assert prod == want


def test_matmul_overlapping_qubits():
    x1 = cs.QuantumGate.from_reals([1], [0, 1, 1, 0])
    prod = x1.matmul(cs.QuantumGate(np.eye(4), [0, 1]))
    want = cs.QuantumGate.from_reals([0, 1], [0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 1, 0, 0])
    assert prod == want


def test_matmul_identity_neutral():
    h = cs.QuantumGate.h(0)
    eye = cs.QuantumGate(np.eye(2), [0])
    assert h.matmul(eye) == h and eye.matmul(h) == h


def test_standard_matrices():
    # src/types/matrix.rs:59-210
    assert np.allclose(mat(cs.QuantumGate.h(0)), [[S2, S2], [S2, -S2]])
    assert np.allclose(mat(cs.QuantumGate.y(0)), [[0, -1j], [1j, 0]])
    assert np.allclose(mat(cs.QuantumGate.s(0)), np.diag([1, 1j]))
    assert np.allclose(mat(cs.QuantumGate.t(0)), np.diag([1, np.exp(1j * np.pi / 4)]))
    assert np.allclose(mat(cs.QuantumGate.cp(0.3, 0, 1)), np.diag([1, 1, 1, np.exp(0.3j)]))
    # cx(c, t): qubits=[c, t], local bit 0 = control
    assert np.array_equal(mat(cs.QuantumGate.cx(0, 1)).real,
                          [[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]])
    # reversed control/target canonicalises to a different matrix on [1, 3]
    g = cs.QuantumGate.cx(3, 1)
    assert list(g.qubits()) == [1, 3]
    assert np.array_equal(mat(g).real, [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])
    th, ph, la = 0.4, 1.1, -0.7
    u3 = mat(cs.QuantumGate.u3(th, ph, la, 0))
    want = np.array([[np.cos(th / 2), -np.exp(1j * la) * np.sin(th / 2)],
                     [np.exp(1j * ph) * np.sin(th / 2), np.exp(1j * (ph + la)) * np.cos(th / 2)]])
    assert np.allclose(u3, want)


def test_random_unitary_is_unitary_and_seeded():
    a = np.asarray(cs.random_unitary_matrix(16, 5))
    assert np.allclose(a.conj().T @ a, np.eye(16), atol=1e-12)
    assert np.array_equal(a, np.asarray(cs.random_unitary_matrix(16, 5)))
    assert not np.array_equal(a, np.asarray(cs.random_unitary_matrix(16, 6)))


def test_haar_su2_has_unit_determinant():
    for s in range(20):
        u = np.asarray(cs.haar_su2(s))
        assert np.allclose(u.conj().T @ u, np.eye(2), atol=1e-13)
        assert abs(np.linalg.det(u) - 1.0) < 1e-13


def test_arithmetic_intensity_and_nnz():
    # gate.rs:261-272: AI = nnz scalars / K
    assert cs.QuantumGate.h(0).scalar_nnz(1e-12) == 4
    assert cs.QuantumGate.h(0).arithmetic_intensity(1e-12) == 2.0
    assert cs.QuantumGate.cx(0, 1).arithmetic_intensity(1e-12) == 1.0
    dense = cs.QuantumGate.random_unitary([0, 1, 2], 3)
    assert dense.arithmetic_intensity(1e-12) == 16.0  # 2K


def test_random_arithmetic_intensity_size():
    # gate.rs:395-431: k = max(4, ceil(log2(ceil(ai/2))))
    assert cs.QuantumGate.random_arithmetic_intensity(10, 1, 1).n_qubits() == 4
    assert cs.QuantumGate.random_arithmetic_intensity(10, 64, 1).n_qubits() == 5
    assert cs.QuantumGate.random_arithmetic_intensity(10, 100, 1).n_qubits() == 6


def test_noise_branches_and_density_matrix_lift():
    d = cs.QuantumGate.depolarizing(0, 0.1)
    br = d.noise_branches()
    assert [p for p, _ in br] == pytest.approx([0.9, 0.1 / 3, 0.1 / 3, 0.1 / 3])
    assert d.effective_n_qubits() == 2 and not d.is_unitary()
    # Superoperator of a channel preserves the trace of rho (gate.rs:236-258)
    s = np.asarray(d.to_density_matrix_gate(1).matrix())
    rho = np.array([[0.7, 0.2 - 0.1j], [0.2 + 0.1j, 0.3]])
    vec = rho.reshape(-1, order="F")  # rho[ket | bra << n] -> ket fastest
    out = (s @ vec).reshape(2, 2, order="F")
    want = sum(p * k @ rho @ k.conj().T for p, k in br)
    assert np.allclose(out, want)
    # Cartesian product of noisy gates: 4 x 4 = 16 branches (fusion.rs:1096-1139)
    assert len(d.matmul(d).noise_branches()) == 16


# ── circuit.rs qft ─────────────────────────────────────────────────────────

def test_qft_gate_count():
    assert len(cs.QuantumCircuit.qft(3, False)) == 6
    assert len(cs.QuantumCircuit.qft(3, True)) == 7
    assert len(cs.QuantumCircuit.qft(4, True)) == 12
    assert len(cs.QuantumCircuit.qft(20, True)) == 220  # docs/precision_analysis.md:36


def test_qft_circuit_matches_dft_via_oracle():
    n = 8
    c = cs.QuantumCircuit.qft(n, True)
    x = to_complex(seeded_state(n))
    got = to_complex(oracle.apply_gates(seeded_state(n), gate_list(c.gates())))
    r = np.array([int(format(j, f"0{n}b")[::-1], 2) for j in range(1 << n)])
    assert np.max(np.abs(got - (np.fft.ifft(x[r]) * np.sqrt(1 << n))[r])) < 1e-12


def test_add_rejects_out_of_range_qubit():
    c = cs.QuantumCircuit(2)
    with pytest.raises(cs.CastError, match="out of range"):
        c.add(cs.QuantumGate.h(2))


def test_random_layer_circuit_shape_and_seed():
    c = cs.random_layer_circuit(30, 20, 7)
    assert len(c) == 20 * (30 + 15)
    ks = [g.n_qubits() for g in c.gates()]
    assert ks.count(1) == 600 and ks.count(2) == 300
    # every layer's CZs form a perfect matching
    gates = c.gates()
    for layer in range(20):
        czs = gates[layer * 45 + 30: layer * 45 + 45]
        qs = sorted(q for g in czs for q in g.qubits())
        assert qs == list(range(30))
    c2 = cs.random_layer_circuit(30, 20, 7)
    assert all(a == b for a, b in zip(c.gates(), c2.gates()))


def test_random_statevector_normalised_and_seeded():
    v = np.asarray(cs.random_statevector(10, 3))
    assert abs(np.sum(v * v) - 1.0) < 1e-12
    assert np.array_equal(v, np.asarray(cs.random_statevector(10, 3)))


# ── circuit graph (circuit_graph.rs tests) ─────────────────────────────────

def test_graph_packs_rows_and_preserves_order():
    c = cs.QuantumCircuit(3)
    for g in (cs.QuantumGate.h(0), cs.QuantumGate.h(1), cs.QuantumGate.cx(0, 1),
              cs.QuantumGate.h(2), cs.QuantumGate.x(1)):
        c.add(g)
    g = cs.CircuitGraph.from_circuit(c)
    g.check_consistency()
    assert g.n_rows() == 3  # [h0 h1 h2] [cx01] [x1]
    order = [list(q.qubits()) for q in g.gates_in_row_order()]
    assert order == [[0], [1], [2], [0, 1], [1]]


# ── cost model (cost_model.rs tests :457-600) ──────────────────────────────

def test_crossover_ai_formula():
    cfg = cs.ProfileConfig()
    p = cs.HardwareProfile.from_roofline(cfg, 50.0, 200.0, 8)
    assert abs(p.crossover_ai - 59.6) < 1.0


def test_size_only_costs():
    m = cs.SizeOnlyCostModel(3)
    assert m.cost_of(cs.QuantumGate.ccx(0, 1, 2)) < 1e-9
    g4 = cs.QuantumGate.cx(0, 1).matmul(cs.QuantumGate.cx(2, 3)).matmul(cs.QuantumGate.cx(0, 2))
    assert g4.n_qubits() == 4 and m.cost_of(g4) == 1.0
    ch2 = cs.QuantumGate(np.eye(4), [0, 1]).with_noise([(1.0, np.eye(4))])
    assert ch2.effective_n_qubits() == 4 and m.cost_of(ch2) == 1.0


def test_adaptive_costs():
    cfg = cs.ProfileConfig()
    prof = cs.HardwareProfile.from_roofline(cfg, 50.0, 200.0, 8)
    m = cs.HardwareAdaptiveCostModel(prof, 4)
    assert abs(m.cost_of(cs.QuantumGate.x(0)) - 1.0) < 1e-9
    tiny = cs.HardwareAdaptiveCostModel(cs.HardwareProfile.from_roofline(cfg, 1000.0, 1.0, 8), 5)
    g4 = cs.QuantumGate.cx(0, 1).matmul(cs.QuantumGate.cx(2, 3)).matmul(cs.QuantumGate.cx(0, 2))
    assert tiny.cost_of(g4) > 1.0
    assert cs.HardwareAdaptiveCostModel(prof, 2).cost_of(cs.QuantumGate.ccx(0, 1, 2)) == float("inf")


def test_profile_json_roundtrip(tmp_path):
    cfg = cs.ProfileConfig()
    cfg.name, cfg.arch, cfg.precision, cfg.n_qubits = "MI355X", "gfx950", "f64", 30
    p = cs.fit_from_samples(cfg, [(1, 500, 6000), (2, 1000, 6000), (4, 2000, 6000),
                                  (64, 30000, 5000), (128, 31000, 2500)])
    f = tmp_path / "p.json"
    p.save(str(f))
    q = cs.HardwareProfile.load(str(f))
    assert q.crossover_ai == pytest.approx(p.crossover_ai)
    assert q.config.arch == "gfx950" and len(q.raw) == 5
    d = json.loads(f.read_text())
    assert d["config"]["device"]["backend"] == "hip"


def test_fit_roofline_recovers_synthetic_roofline():
    # slope 500 GFLOP/s per AI unit, plateau 20000 -> crossover 40
    pts = [(a, min(500 * a, 20000), 0) for a in (1, 2, 4, 8, 16, 32, 64, 128)]
    cx, slope, r2 = cs.fit_roofline(pts)
    assert slope == pytest.approx(500) and cx == pytest.approx(40) and r2 == pytest.approx(1)


def test_adaptive_sweep_drives_probes():
    seen = []

    def probe(ai, budget):
        seen.append(ai)
        return (ai, min(500 * ai, 20000), 1000.0)

    p = cs.adaptive_sweep(probe, cs.ProfileConfig(), 1.0)
    assert seen[:7] == [1, 2, 4, 8, 16, 32, 64]
    assert p.crossover_ai == pytest.approx(40)


# ── fusion (fusion.rs tests; tests/fusion_correctness.rs) ──────────────────

def circuit_unitary(gates, n):
    u = np.eye(1 << n, dtype=complex)
    return brute_force_apply(u, [(np.asarray(g.matrix()), list(g.qubits())) for g in gates], n)


@pytest.mark.parametrize("seed", [0, 1, 2, 7, 17, 42])
@pytest.mark.parametrize("size", [2, 3, 4, 5])
def test_fusion_preserves_unitary(seed, size):
    n = 6
    c = cs.random_layer_circuit(n, 4, seed)
    g = cs.CircuitGraph.from_circuit(c)
    u0 = circuit_unitary(g.gates_in_row_order(), n)
    cs.optimize(g, cs.FusionConfig.size_only(size))
    g.check_consistency()
    fused = g.gates_in_row_order()
    assert max(x.n_qubits() for x in fused) <= max(size, 2)
    assert np.max(np.abs(circuit_unitary(fused, n) - u0)) < 1e-10


def test_qft_fusion_two_is_phase_one_only():
    n = 10
    g = cs.CircuitGraph.from_circuit(cs.QuantumCircuit.qft(n, True))
    x = seeded_state(n)
    want = oracle.apply_gates(x.copy(), gate_list(g.gates_in_row_order()))
    cs.optimize(g, cs.FusionConfig.size_only(2))
    fused = g.gates_in_row_order()
    assert all(q.n_qubits() <= 2 for q in fused) and len(fused) < 220
    got = oracle.apply_gates(x.copy(), gate_list(fused))
    assert np.max(np.abs(got - want)) < 1e-12


def test_fusion_log_and_adaptive_rejects_past_crossover():
    n = 8
    c = cs.random_layer_circuit(n, 6, 3)
    cfg = cs.ProfileConfig()
    # crossover ~ 9: k=3 dense (AI 16) costs > 1, so large fusions get rejected
    prof = cs.HardwareProfile.from_measurements(cfg, 1000.0, 9.0 * 500.0, 500.0)
    g = cs.CircuitGraph.from_circuit(c)
    log = cs.optimize_with_log(g, cs.FusionConfig.hardware_adaptive(prof, 5, 1e-12))
    assert log.n_accepted() + log.n_rejected() == len(log.decisions)
    assert log.n_rejected() > 0
    assert "Fusion log" in log.summary()
    g2 = cs.CircuitGraph.from_circuit(c)
    cs.optimize(g2, cs.FusionConfig.size_only(5))
    assert len(g2.gates_in_row_order()) <= len(g.gates_in_row_order())


def test_r2_warning_below_threshold():
    """profile_hw reports an unreliable fit below R^2 0.90 with the
    reference's wording (src/profile.rs:69, 323-328): a sweep with a dip
    between the memory slope and the plateau fits worse than that, the ideal
    roofline does not."""
    from cast_simulator_amd import profile_hw

    ideal = [(a, min(500 * a, 20000), 0) for a in (1, 2, 4, 8, 16, 32, 64, 128)]
    dip = [(a, min(500 * a, 20000) * (0.45 if 24 <= a <= 64 else 1.0), 0)
           for a in (1, 2, 4, 8, 16, 24, 32, 48, 64, 96, 128)]
    r2_ideal, r2_dip = cs.fit_roofline(ideal)[2], cs.fit_roofline(dip)[2]
    assert r2_dip < profile_hw.R2_WARN <= r2_ideal
    assert profile_hw.r2_warning("hip f32", r2_ideal) is None
    msg = profile_hw.r2_warning("hip f32", r2_dip)
    assert msg == (f"  hip f32: Warning: roofline fit R^2 = {r2_dip:.3f} (< 0.9); "
                   "crossover estimate may be unreliable.")
