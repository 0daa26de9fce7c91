"""Measurement, density-matrix and ensemble pieces on the CPU.

Pins the oracle's marginal, density-matrix and ensemble restatements against
the reference's own known answers (src/simulator/measure.rs:103-137,
src/simulator/tests.rs:22-58 and :88-305), against closed forms, and against
independent numpy computations; and tests the product's host-side inverse-CDF
sampler (src/simulator/measure.rs:45-78), which has no GPU part.
"""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs
import oracle
from helpers import to_complex

TEST_QUBITS = 6  # src/simulator/tests.rs:4


def bell_sv(n=6):
    sv = np.zeros(2 << n)
    sv[0] = 1.0
    oracle.apply_gate(sv, np.asarray(cs.QuantumGate.h(0).matrix()), [0])
    cx = cs.QuantumGate.cx(0, 1)
    oracle.apply_gate(sv, np.asarray(cx.matrix()), list(cx.qubits()))
    return sv


def numpy_marginal(sv, measured):
    """Independent marginal: |amp|^2 as an n-axis tensor (axis a = qubit n-1-a),
    summed over unmeasured axes, axes ordered so bin bit i = measured[i]."""
    n = int(sv.size // 2).bit_length() - 1
    p = np.abs(to_complex(sv)) ** 2
    t = p.reshape([2] * n)
    keep = [n - 1 - q for q in measured]
    drop = tuple(a for a in range(n) if a not in keep)
    r = t.sum(axis=drop) if drop else t
    # remaining axes are in ascending axis order; want bin bit i <- measured[i],
    # i.e. axis order (most significant first) measured[m-1], ..., measured[0].
    cur = sorted(keep)
    want = [n - 1 - q for q in reversed(measured)]
    r = np.transpose(r, [cur.index(a) for a in want]) if len(cur) > 1 else r
    return np.asarray(r).reshape(-1)


# ── oracle marginal ─────────────────────────────────────────────────────────

def test_marginal_bell_known_answers():
    sv = bell_sv()
    assert np.allclose(oracle.marginal(sv, [0]), [0.5, 0.5], atol=1e-10)     # measure.rs:103-109
    assert np.allclose(oracle.marginal(sv, [0, 1]), [0.5, 0, 0, 0.5], atol=1e-10)  # :112-121


@pytest.mark.parametrize("measured", [[0], [3], [1, 4], [4, 1], [5, 0, 2], [2, 5, 0, 3],
                                      [0, 1, 2, 3, 4, 5], [5, 4, 3, 2, 1, 0], []])
def test_marginal_matches_numpy_tensor_sum(measured):
    n = 6
    sv = np.asarray(cs.random_statevector(n, 17))
    got = oracle.marginal(sv, measured)
    assert np.allclose(got, numpy_marginal(sv, measured), rtol=0, atol=1e-14)
    assert abs(got.sum() - oracle.norm_squared(sv)) < 1e-13


def test_marginal_index_order_is_not_symmetric():
    # |q0=1, q1=0>: measured [0, 1] -> bin 1; measured [1, 0] -> bin 2.
    sv = np.zeros(2 << 3)
    sv[2 * 1] = 1.0
    assert list(oracle.marginal(sv, [0, 1])) == [0, 1, 0, 0]
    assert list(oracle.marginal(sv, [1, 0])) == [0, 0, 1, 0]


# ── host sampler ────────────────────────────────────────────────────────────

def test_batch_sample_uniform_reference_case():
    hist = cs.batch_sample([0.25] * 4, 10000, 42)       # measure.rs:123-137
    assert sum(hist.values()) == 10000
    for k in range(4):
        assert abs(hist.get(k, 0) - 2500) < 300


def test_batch_sample_edges_and_determinism():
    assert cs.batch_sample([], 10, 1) == {}
    assert cs.batch_sample([0.5, 0.5], 0, 1) == {}
    assert cs.batch_sample([0, 0, 1.0, 0], 1000, 3) == {2: 1000}
    # Mass below one: draws beyond the last cumulative sum clamp to the last bin.
    h = cs.batch_sample([0.1, 0.1], 5000, 4)
    assert sum(h.values()) == 5000 and h[1] > 4000
    p = [1 / 16] * 16
    assert cs.batch_sample(p, 5000, 9) == cs.batch_sample(p, 5000, 9)
    assert cs.batch_sample(p, 5000, 9) != cs.batch_sample(p, 5000, 10)


def test_batch_sample_frequencies_chi_square():
    rng = np.random.default_rng(5)
    p = rng.random(16)
    p /= p.sum()
    N = 200_000
    h = cs.batch_sample(list(p), N, 11)
    obs = np.array([h.get(k, 0) for k in range(16)], dtype=float)
    chi2 = float(((obs - N * p) ** 2 / (N * p)).sum())
    assert chi2 < 50.0  # 15 dof: P(chi2 > 50) ~ 1e-5


# ── oracle density matrix ───────────────────────────────────────────────────

def lifted(g, n):
    d = g.to_density_matrix_gate(n)
    return np.asarray(d.matrix()), list(d.qubits())


@pytest.mark.parametrize("p", [0.0, 0.1, 0.3])
def test_dm_depolarizing_closed_form(p):
    n = 2
    rho = oracle.simulate_dm(n, [lifted(cs.QuantumGate.depolarizing(0, p), n)])
    pops = oracle.dm_populations(rho, n)
    # (1 - p) |0><0| + p/3 (X|0><0|X + Y|0><0|Y + Z|0><0|Z)
    assert np.allclose(pops, [1 - 2 * p / 3, 2 * p / 3, 0, 0], atol=1e-14)


def test_dm_h_depolarizing_cx_populations():
    n = 3
    gates = [cs.QuantumGate.h(0), cs.QuantumGate.depolarizing(0, 0.1), cs.QuantumGate.cx(0, 1)]
    rho = oracle.simulate_dm(n, [lifted(g, n) for g in gates])
    pops = oracle.dm_populations(rho, n)
    want = np.zeros(8)
    want[0] = want[3] = 0.5
    assert np.allclose(pops, want, atol=1e-14)


def test_dm_of_unitary_circuit_matches_statevector():
    n = 4
    gates = [cs.QuantumGate.random_unitary([0, 2], 1), cs.QuantumGate.random_unitary([1], 2),
             cs.QuantumGate.random_unitary([1, 3], 3), cs.QuantumGate.cx(3, 0)]
    sv = np.zeros(2 << n)
    sv[0] = 1.0
    for g in gates:
        oracle.apply_gate(sv, np.asarray(g.matrix()), list(g.qubits()))
    rho = oracle.simulate_dm(n, [lifted(g, n) for g in gates])
    psi = to_complex(sv)
    # rho[i][j] = psi_i conj(psi_j) at index i | (j << n)
    r = to_complex(rho).reshape(1 << n, 1 << n)   # r[j, i]
    assert np.allclose(r.T, np.outer(psi, psi.conj()), atol=1e-13)
    assert np.allclose(oracle.dm_populations(rho, n), np.abs(psi) ** 2, atol=1e-14)


def test_dm_trace_preserved_reference_case():
    n = TEST_QUBITS   # tests.rs:41-58 on 6 qubits (12 virtual)
    gates = [cs.QuantumGate.h(0), cs.QuantumGate.cx(0, 1), cs.QuantumGate.depolarizing(0, 0.05)]
    rho = oracle.simulate_dm(n, [lifted(g, n) for g in gates])
    assert abs(oracle.dm_populations(rho, n).sum() - 1.0) < 1e-12


# ── oracle ensemble ─────────────────────────────────────────────────────────

def ens_gates(gates):
    out = []
    for g in gates:
        if g.is_unitary():
            out.append(("u", np.asarray(g.matrix()), list(g.qubits())))
        else:
            out.append(("noise", [(p, np.asarray(k)) for p, k in g.noise_branches()],
                        list(g.qubits())))
    return out


def test_ensemble_deterministic_branch_reference_case():
    e = oracle.ensemble(TEST_QUBITS, ens_gates([cs.QuantumGate.h(0),
                                                cs.QuantumGate.depolarizing(0, 0.1)]), 1)
    assert len(e) == 1 and e[0][1] == [0] and abs(e[0][0] - 0.9) < 1e-10  # tests.rs:111-123


def test_ensemble_m4_explores_all_weight_reference_case():
    g = ens_gates([cs.QuantumGate.h(0), cs.QuantumGate.depolarizing(0, 0.1)])
    e1, e4 = oracle.ensemble(TEST_QUBITS, g, 1), oracle.ensemble(TEST_QUBITS, g, 4)
    assert len(e1) == 1 and len(e4) == 4                                   # tests.rs:125-150
    assert abs(sum(w for w, _, _ in e4) - 1.0) < 1e-10
    assert [p for _, p, _ in e4] == [[0], [1], [2], [3]]   # ties in branch order


def test_ensemble_pruning_reference_case():
    g = ens_gates([cs.QuantumGate.h(0), cs.QuantumGate.depolarizing(0, 0.1),
                   cs.QuantumGate.cx(0, 1), cs.QuantumGate.depolarizing(1, 0.1)])
    e = oracle.ensemble(TEST_QUBITS, g, 2)
    assert len(e) == 2 and e[0][0] >= e[1][0]                              # tests.rs:189-215
    w = sum(x[0] for x in e)
    assert 0.7 < w < 1.0
    assert e[0][1] == [0, 0] and abs(e[0][0] - 0.81) < 1e-12


def test_ensemble_noiseless_reference_case():
    e = oracle.ensemble(TEST_QUBITS, ens_gates([cs.QuantumGate.h(0), cs.QuantumGate.cx(0, 1)]), 10)
    assert len(e) == 1 and abs(e[0][0] - 1.0) < 1e-10                    # tests.rs:253-270
    m = oracle.marginal(e[0][2], [0, 1])
    assert np.allclose(m, [0.5, 0, 0, 0.5], atol=1e-14)


@pytest.mark.parametrize("gates", [
    [cs.QuantumGate.h(0), cs.QuantumGate.depolarizing(0, 0.1), cs.QuantumGate.cx(0, 1)],
    [cs.QuantumGate.h(0), cs.QuantumGate.depolarizing(0, 0.05), cs.QuantumGate.cx(0, 1),
     cs.QuantumGate.depolarizing(0, 0.05), cs.QuantumGate.depolarizing(1, 0.05)],
])
def test_full_ensemble_mixture_equals_density_matrix(gates):
    # With every branch kept, sum_b w_b |psi_b><psi_b| is the channel's output,
    # so the weighted marginals equal the density matrix's populations
    # (the exact form of tests.rs:152-187 and :272-305).
    n = 3
    e = oracle.ensemble(n, ens_gates(gates), 64)
    assert abs(sum(w for w, _, _ in e) - 1.0) < 1e-12
    mix = sum(w * oracle.marginal(x, [0, 1]) for w, _, x in e)
    rho = oracle.simulate_dm(n, [lifted(g, n) for g in gates])
    pops = oracle.dm_populations(rho, n)
    want = np.array([pops[i::4].sum() for i in range(4)])
    assert np.allclose(mix, want, atol=1e-13)


def test_samples_per_member():
    assert oracle.samples_per_member([0.5, 0.25, 0.25], 10) == [5, 2, 3]
    assert sum(oracle.samples_per_member([0.9, 0.1 / 3, 0.1 / 3, 0.1 / 3], 1001)) == 1001
    assert oracle.samples_per_member([1.0], 7) == [7]
