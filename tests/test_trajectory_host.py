"""Batched-trajectory host logic on the CPU: the sampler, the trajectory gate
lists, and the planner's fused task gate lists (checkpoints, chunk fusion,
Pauli errors folded into fused matrices), each checked against the oracle
statevector simulator; the sampler's statistics against an exact density
matrix evolution of the depolarizing channel (src/types/gate.rs:174-197:
identity with 1 - p, X, Y, Z with p/3 each), computed two ways."""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs
import oracle
from helpers import rel_l2, to_complex

PAULI = {
    1: np.array([[0, 1], [1, 0]], dtype=complex),
    2: np.array([[0, -1j], [1j, 0]], dtype=complex),
    3: np.array([[1, 0], [0, -1]], dtype=complex),
}


def err_tuples(s):
    return [tuple(e) for e in s.errors]


def oracle_run(circ_gates, init):
    sv = np.array(init, dtype=np.float64)
    for g in circ_gates:
        oracle.apply_gate(sv, np.asarray(g.matrix()), list(g.qubits()))
    return sv


def clean_prefix(lc, n_layers, init):
    gates = lc.gates()[: lc.layer_start[n_layers]]
    return oracle_run(gates, init)


# ── sampler ─────────────────────────────────────────────────────────────────

def test_sampler_is_deterministic_and_subset_stable():
    lc = cs.LayeredCircuit.random_layers(10, 12, 3)
    ns = cs.NoiseSpec(0.02, 4, 12, 77)
    a = cs.sample_trajectories(lc, ns, 30)
    b = cs.sample_trajectories(lc, ns, 30)
    c = cs.sample_trajectories(lc, ns, 10, first=13)
    assert [(s.depth, err_tuples(s)) for s in a] == [(s.depth, err_tuples(s)) for s in b]
    assert [(s.depth, err_tuples(s)) for s in a[13:23]] == [(s.depth, err_tuples(s)) for s in c]
    ns2 = cs.NoiseSpec(0.02, 4, 12, 78)
    d = cs.sample_trajectories(lc, ns2, 30)
    assert [err_tuples(s) for s in a] != [err_tuples(s) for s in d]


def test_sampler_rates_depths_and_pauli_mix():
    n, depth, p, count = 16, 10, 0.02, 400
    lc = cs.LayeredCircuit.random_layers(n, depth, 5)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(p, depth, depth, 11), count)
    events = count * depth * (n + 2 * (n // 2))  # qubits touched per layer
    errs = [e for s in S for e in s.errors]
    mu, sd = events * p, np.sqrt(events * p * (1 - p))
    assert abs(len(errs) - mu) < 5 * sd
    counts = np.bincount([e[2] for e in errs], minlength=4)[1:]
    third = len(errs) / 3
    assert np.all(np.abs(counts - third) < 5 * np.sqrt(third))
    # each error names one of its gate's qubits, in circuit order
    gates = lc.gates()
    for s in S:
        g = [e[0] for e in s.errors]
        assert g == sorted(g)
        for gi, q, _ in s.errors:
            assert q in gates[gi].qubits()
    S2 = cs.sample_trajectories(lc, cs.NoiseSpec(0.0, 1, depth, 3), 500)
    ds = np.array([s.depth for s in S2])
    assert ds.min() == 1 and ds.max() == depth and all(not s.errors for s in S2)
    S3 = cs.sample_trajectories(lc, cs.NoiseSpec(1.0, 2, 2, 3), 3)
    assert all(len(s.errors) == 2 * (n + 2 * (n // 2)) for s in S3)


def test_sampler_rejects_bad_specs():
    lc = cs.LayeredCircuit.random_layers(4, 5, 1)
    with pytest.raises(cs.CastError):
        cs.sample_trajectories(lc, cs.NoiseSpec(0.1, 0, 5, 1), 1)
    with pytest.raises(cs.CastError):
        cs.sample_trajectories(lc, cs.NoiseSpec(0.1, 2, 6, 1), 1)
    with pytest.raises(cs.CastError):
        cs.sample_trajectories(lc, cs.NoiseSpec(1.5, 1, 5, 1), 1)


def test_layered_circuit_shape():
    lc = cs.LayeredCircuit.random_layers(7, 4, 2)
    assert lc.n_layers() == 4
    assert list(lc.layer_start) == [0, 10, 20, 30, 40]  # 7 SU(2) + 3 CZ per layer
    ref = cs.random_layer_circuit(7, 4, 2).gates()
    got = lc.gates()
    assert len(ref) == len(got)
    assert all(a == b for a, b in zip(ref, got))


def test_trajectory_circuit_inserts_paulis_after_their_gate():
    lc = cs.LayeredCircuit.random_layers(5, 3, 9)
    s = cs.TrajectorySample(2, [(0, 0, 1), (5, 3, 3), (5, 4, 2)])
    qc = cs.trajectory_circuit(lc, s)
    gates = qc.gates()
    assert len(gates) == lc.layer_start[2] + 3
    assert np.allclose(np.asarray(gates[1].matrix()), PAULI[1]) and list(gates[1].qubits()) == [0]
    assert np.allclose(np.asarray(gates[7].matrix()), PAULI[3]) and list(gates[7].qubits()) == [3]
    assert np.allclose(np.asarray(gates[8].matrix()), PAULI[2]) and list(gates[8].qubits()) == [4]
    with pytest.raises(cs.CastError):  # error beyond the trajectory's depth
        cs.trajectory_circuit(lc, cs.TrajectorySample(1, [(lc.layer_start[1], 0, 1)]))


def test_origin_concatenates_in_application_order():
    a = cs.QuantumGate(np.asarray(cs.haar_su2(1)), [0]).with_origin([4])
    b = cs.QuantumGate(np.asarray(cs.haar_su2(2)), [1]).with_origin([7])
    assert list(b.matmul(a).origin()) == [4, 7]  # b * a: a first
    assert list(a.matmul(b).origin()) == [7, 4]


# ── planner vs oracle ───────────────────────────────────────────────────────

@pytest.mark.parametrize("B,kmax", [(1, 2), (3, 3), (4, 4), (9, 4), (2, 1), (4, 5), (9, 5)])
def test_plan_task_gate_lists_match_explicit_trajectories(B, kmax):
    n, D = 7, 9
    lc = cs.LayeredCircuit.random_layers(n, D, 21)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.04, 1, D, 1234 + B), 40)
    assert sum(len(s.errors) for s in S) > 20
    init = np.asarray(cs.random_statevector(n, 8))
    # k = 5 gates exist for f32 runs only; the host plan holds f64 matrices either way
    prec = cs.Precision.F32 if kmax == 5 else cs.Precision.F64
    eng = cs.TrajectoryEngine(lc, prec, [0, 1])
    eng.set_initial_state(list(init))
    P = eng.plan(S, cs.TrajectoryTuning(B, kmax, 256, 2, 0))
    pool = P["pool"]
    ck = {}
    for t, s in enumerate(S):
        want = oracle_run(cs.trajectory_circuit(lc, s).gates(), init)
        first, cnt, c = P["tasks"][P["task_of"][t]]
        if c not in ck:
            ck[c] = clean_prefix(lc, c * B, init)
        got = ck[c].copy()
        for off, q in P["gates"][first:first + cnt]:
            assert len(q) <= max(kmax, 2) and list(q) == sorted(q)  # CZ stays 2-qubit
            K = 1 << len(q)
            if K >= 16:  # three real matrices: Re, Im, Re + Im
                reals = pool.view(np.float64)[2 * off:2 * off + 3 * K * K]
                re, im, rs = reals[:K * K], reals[K * K:2 * K * K], reals[2 * K * K:]
                assert np.array_equal(rs, re + im)
                m = (re + 1j * im).reshape(K, K)
            else:
                m = pool[off:off + K * K].reshape(K, K)
            oracle.apply_gate(got, m, list(q))
        assert rel_l2(want, got) < 1e-12, f"trajectory {t}"
        assert c * B <= s.depth
        if s.errors:
            first_layer = max(l for l in range(D + 1) if lc.layer_start[l] <= s.errors[0][0])
            assert c == min(first_layer, s.depth) // B


def test_plan_shares_clean_tasks_and_orders_longest_first():
    lc = cs.LayeredCircuit.random_layers(6, 12, 4)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.0, 1, 12, 5), 200)
    eng = cs.TrajectoryEngine(lc, cs.Precision.F32, [0, 1])
    P = eng.plan(S, cs.TrajectoryTuning(4, 4, 256, 2, 0))
    depths = sorted({s.depth for s in S})
    assert len(P["tasks"]) == len(depths)
    assert P["n_shared_clean"] == len(S) - len(depths)
    assert P["n_dirty_matrices"] == 0
    for t, s in enumerate(S):
        first, cnt, c = P["tasks"][P["task_of"][t]]
        assert c == s.depth // 4
        assert (cnt == 0) == (s.depth % 4 == 0)
    lens = [P["tasks"][k][1] for k in P["order"]]
    assert lens == sorted(lens, reverse=True)
    assert sorted(P["order"]) == list(range(len(P["tasks"])))


def test_plan_rejects_bad_tuning_and_samples():
    lc = cs.LayeredCircuit.random_layers(5, 6, 4)
    eng = cs.TrajectoryEngine(lc, cs.Precision.F64, [0, 1])
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.1, 1, 6, 5), 4)
    for bad in (cs.TrajectoryTuning(0, 4, 256, 2, 0), cs.TrajectoryTuning(7, 4, 256, 2, 0),
                cs.TrajectoryTuning(2, 5, 256, 2, 0), cs.TrajectoryTuning(2, 0, 256, 2, 0)):
        with pytest.raises(cs.CastError):
            eng.plan(S, bad)
    with pytest.raises(cs.CastError):
        eng.plan([], cs.TrajectoryTuning())
    with pytest.raises(cs.CastError):  # out of circuit order
        eng.plan([cs.TrajectorySample(3, [(9, 4, 1), (2, 2, 1)])], cs.TrajectoryTuning(1, 4))
    with pytest.raises(cs.CastError):
        eng.plan([cs.TrajectorySample(7, [])], cs.TrajectoryTuning(1, 4))
    with pytest.raises(cs.CastError):
        cs.TrajectoryEngine(lc, cs.Precision.F64, [0, 0])
    e32 = cs.TrajectoryEngine(lc, cs.Precision.F32, [0, 1])
    e32.plan(S, cs.TrajectoryTuning(2, 5, 256, 2, 0))
    with pytest.raises(cs.CastError, match="1..5"):
        e32.plan(S, cs.TrajectoryTuning(2, 6, 256, 2, 0))
    with pytest.raises(cs.CastError):
        cs.TrajectoryEngine(lc, cs.Precision.F64, [0, 5])


# ── sampler statistics vs the exact channel ─────────────────────────────────

def embed(m, qubits, n):
    """Dense operator of m on `qubits` (local bit b = qubits[b])."""
    from helpers import full_operator
    return full_operator(np.asarray(m), list(qubits), n)


def dm_zz_by_depth(lc, p, pairs):
    """<(1/P) sum ZZ> after d layers of the noisy circuit, for every d, by
    evolving rho with U rho U^dag and the depolarizing channel on every
    touched qubit after every gate."""
    n = lc.n_qubits
    N = 1 << n
    rho = np.zeros((N, N), dtype=complex)
    rho[0, 0] = 1
    idx = np.arange(N)
    zz = np.mean([1 - 2 * (((idx >> a) ^ (idx >> b)) & 1) for a, b in pairs], axis=0)
    paulis = {q: [embed(PAULI[k], [q], n) for k in (1, 2, 3)] for q in range(n)}
    out = {}
    gates = lc.gates()
    for d in range(1, lc.n_layers() + 1):
        for g in gates[lc.layer_start[d - 1]:lc.layer_start[d]]:
            U = embed(g.matrix(), g.qubits(), n)
            rho = U @ rho @ U.conj().T
            for q in g.qubits():
                rho = (1 - p) * rho + (p / 3) * sum(P @ rho @ P for P in paulis[q])
        out[d] = float(np.real(np.sum(np.diag(rho) * zz)))
    return out


def dm_zz_superoperator(lc, p, pairs, depth):
    """The same quantity through the reference's density-matrix method: each
    gate and channel lifted to a superoperator on 2n virtual qubits
    (src/types/gate.rs:236-258), applied by the oracle to vec(rho)."""
    n = lc.n_qubits
    v = np.zeros(2 << (2 * n))
    v[0] = 1.0
    gates = lc.gates()
    for g in gates[:lc.layer_start[depth]]:
        sg = g.to_density_matrix_gate(n)
        oracle.apply_gate(v, np.asarray(sg.matrix()), list(sg.qubits()))
        for q in g.qubits():
            ch = cs.QuantumGate(np.eye(2), [q]).with_depolarizing(p).to_density_matrix_gate(n)
            oracle.apply_gate(v, np.asarray(ch.matrix()), list(ch.qubits()))
    rho = to_complex(v).reshape(1 << n, 1 << n).T  # column-major vec
    N = 1 << n
    idx = np.arange(N)
    zz = np.mean([1 - 2 * (((idx >> a) ^ (idx >> b)) & 1) for a, b in pairs], axis=0)
    return float(np.real(np.sum(np.diag(rho) * zz)))


def test_two_density_matrix_methods_agree():
    lc = cs.LayeredCircuit.random_layers(3, 3, 8)
    pairs = [(0, 1), (1, 2)]
    exact = dm_zz_by_depth(lc, 0.1, pairs)
    for d in (1, 3):
        assert abs(dm_zz_superoperator(lc, 0.1, pairs, d) - exact[d]) < 1e-12


def test_sampled_trajectories_converge_to_density_matrix():
    n, D, p, count = 4, 6, 0.06, 3000
    lc = cs.LayeredCircuit.random_layers(n, D, 31)
    pairs = [(i, i + 1) for i in range(n - 1)]
    exact = dm_zz_by_depth(lc, p, pairs)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(p, 2, D, 2024), count)
    vals, want = [], []
    z0 = np.zeros(2 << n)
    z0[0] = 1.0
    idx = np.arange(1 << n)
    zzv = np.mean([1 - 2 * (((idx >> a) ^ (idx >> b)) & 1) for a, b in pairs], axis=0)
    for s in S:
        sv = oracle_run(cs.trajectory_circuit(lc, s).gates(), z0)
        pr = np.abs(to_complex(sv)) ** 2
        vals.append(float(np.sum(pr * zzv)))
        want.append(exact[s.depth])
    vals, want = np.array(vals), np.array(want)
    err = abs(vals.mean() - want.mean())
    sigma = vals.std(ddof=1) / np.sqrt(count)
    assert err < 3 * sigma + 1e-12, (err, sigma)
    # with p = 0 the trajectories are the noiseless circuit: zero spread
    S0 = cs.sample_trajectories(lc, cs.NoiseSpec(0.0, D, D, 1), 3)
    sv = oracle_run(cs.trajectory_circuit(lc, S0[0]).gates(), z0)
    assert abs(float(np.sum(np.abs(to_complex(sv)) ** 2 * zzv)) - dm_zz_by_depth(lc, 0.0, pairs)[D]) < 1e-12


# ── circuits with their own noise models (src/types/noise.rs:12-33) ─────────

def amplitude_damping(q, gamma, p1):
    """Amplitude damping as the reference's noise model: branches (p_b, A_b /
    sqrt(p_b)) with Kraus operators A_0 = diag(1, sqrt(1 - gamma)), A_1 =
    sqrt(gamma) |0><1|, so that sum_b p_b K_b rho K_b^dag is the channel.
    Neither branch is unitary."""
    a0 = np.array([[1, 0], [0, np.sqrt(1 - gamma)]], dtype=complex)
    a1 = np.array([[0, np.sqrt(gamma)], [0, 0]], dtype=complex)
    p0 = 1.0 - p1
    return cs.QuantumGate(np.eye(2), [q]).with_noise(
        [(p0, a0 / np.sqrt(p0)), (p1, a1 / np.sqrt(p1))])


def noisy_circuit(n, depth, seed, gamma=0.2, pxx=0.1):
    """Random layers with a correlated XX error on every CZ (branch 1, p =
    pxx) and amplitude damping on qubit 0 and n - 1 after every layer."""
    XX = np.kron(PAULI[1], PAULI[1])
    base = cs.random_layer_circuit(n, depth, seed).gates()
    qc = cs.QuantumCircuit(n)
    per_layer = n + n // 2
    for i, g in enumerate(base):
        if g.n_qubits() == 2:
            g = g.with_noise([(1 - pxx, np.eye(4)), (pxx, XX)])
        qc.add(g)
        if i % per_layer == per_layer - 1:
            qc.add(amplitude_damping(0, gamma, 0.3))
            qc.add(amplitude_damping(n - 1, gamma, 0.3))
    return qc


def test_from_circuit_layers_are_moments_and_keep_the_operator():
    qc = noisy_circuit(5, 3, 12)
    lc = cs.LayeredCircuit.from_circuit(qc)
    gates = lc.gates()
    assert len(gates) == len(qc.gates())
    for l in range(lc.n_layers()):
        qs = [q for g in gates[lc.layer_start[l]:lc.layer_start[l + 1]] for q in g.qubits()]
        assert len(qs) == len(set(qs))  # one layer: disjoint qubits
    # unitary parts in layer order equal the circuit in its own order
    init = np.asarray(cs.random_statevector(5, 3))
    a, b = init.copy(), init.copy()
    for g in qc.gates():
        oracle.apply_gate(a, np.asarray(g.matrix()), list(g.qubits()))
    for g in gates:
        oracle.apply_gate(b, np.asarray(g.matrix()), list(g.qubits()))
    assert rel_l2(a, b) < 1e-13


def test_sampler_draws_noise_branches_at_their_probabilities():
    qc = noisy_circuit(6, 4, 2, pxx=0.1)
    lc = cs.LayeredCircuit.from_circuit(qc)
    gates = lc.gates()
    noisy = [i for i, g in enumerate(gates) if not g.is_unitary()]
    count = 3000
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.0, lc.n_layers(), lc.n_layers(), 5), count)
    hits = {}
    for s in S:
        for g, b, kind in s.errors:
            assert kind == 0 and g in noisy
            hits[(g, b)] = hits.get((g, b), 0) + 1
    for g in noisy:
        clean = cs.clean_branch(gates[g])
        probs = [p for p, _ in gates[g].noise_branches()]
        assert clean == int(np.argmax(probs))
        for b, p in enumerate(probs):
            if b == clean:
                assert (g, b) not in hits
                continue
            mu, sd = count * p, np.sqrt(count * p * (1 - p))
            assert abs(hits.get((g, b), 0) - mu) < 5 * sd, (g, b)
    # branch events and depolarizing Paulis together: branch first per gate
    S2 = cs.sample_trajectories(lc, cs.NoiseSpec(0.2, 1, lc.n_layers(), 9), 200)
    for s in S2:
        for e0, e1 in zip(s.errors, s.errors[1:]):
            assert e0[0] <= e1[0]
            if e0[0] == e1[0]:
                assert e1[2] != 0


def test_trajectory_circuit_takes_the_sampled_branch():
    qc = noisy_circuit(4, 2, 6)
    lc = cs.LayeredCircuit.from_circuit(qc)
    gates = lc.gates()
    g = next(i for i, x in enumerate(gates) if x.n_qubits() == 2 and not x.is_unitary())
    s = cs.TrajectorySample(lc.n_layers(), [(g, 1, 0), (g, gates[g].qubits()[1], 3)])
    out = cs.trajectory_circuit(lc, s).gates()
    k1 = np.asarray(gates[g].noise_branches()[1][1])
    assert np.allclose(np.asarray(out[g].matrix()), k1)
    assert np.allclose(np.asarray(out[g + 1].matrix()), PAULI[3])
    k0 = np.asarray(gates[g].noise_branches()[0][1])
    clean = cs.trajectory_circuit(lc, cs.TrajectorySample(lc.n_layers(), [])).gates()
    assert np.allclose(np.asarray(clean[g].matrix()), k0)
    with pytest.raises(cs.CastError):  # a branch event on a unitary gate
        cs.trajectory_circuit(lc, cs.TrajectorySample(lc.n_layers(), [(0, 1, 0)]))


@pytest.mark.parametrize("B,kmax", [(1, 2), (2, 4), (3, 3), (2, 5)])
def test_plan_with_noise_models_matches_explicit_trajectories(B, kmax):
    qc = noisy_circuit(6, 4, 19)
    lc = cs.LayeredCircuit.from_circuit(qc)
    D = lc.n_layers()
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.05, 1, D, 300 + B), 40)
    assert sum(1 for s in S for e in s.errors if e[2] == 0) > 20
    init = np.asarray(cs.random_statevector(6, 4))
    prec = cs.Precision.F32 if kmax == 5 else cs.Precision.F64
    eng = cs.TrajectoryEngine(lc, prec, [0, 1])
    eng.set_initial_state(list(init))
    P = eng.plan(S, cs.TrajectoryTuning(B, kmax, 256, 2, 0))
    pool = P["pool"]
    clean = cs.trajectory_circuit(lc, cs.TrajectorySample(D, [])).gates()
    for t, s in enumerate(S):
        want = oracle_run(cs.trajectory_circuit(lc, s).gates(), init)
        first, cnt, c = P["tasks"][P["task_of"][t]]
        got = oracle_run(clean[:lc.layer_start[c * B]], init)
        for off, q in P["gates"][first:first + cnt]:
            K = 1 << len(q)
            if K >= 16:
                reals = pool.view(np.float64)[2 * off:2 * off + 3 * K * K]
                m = (reals[:K * K] + 1j * reals[K * K:2 * K * K]).reshape(K, K)
            else:
                m = pool[off:off + K * K].reshape(K, K)
            oracle.apply_gate(got, m, list(q))
        assert rel_l2(want, got) < 1e-12, f"trajectory {t}"


def test_noise_model_trajectories_converge_to_density_matrix():
    """Non-unitary branches: the ratio of summed unnormalised expectations to
    summed squared norms converges to the density-matrix value (each noisy
    gate lifted to sum_b p_b K_b (x) conj(K_b), src/types/gate.rs:236-258),
    within 3 sigma."""
    n, count = 4, 4000
    qc = noisy_circuit(n, 3, 44, gamma=0.5, pxx=0.2)
    lc = cs.LayeredCircuit.from_circuit(qc)
    D = lc.n_layers()
    pairs = [(i, i + 1) for i in range(n - 1)]
    idx = np.arange(1 << n)
    zzv = np.mean([1 - 2 * (((idx >> a) ^ (idx >> b)) & 1) for a, b in pairs], axis=0)
    v = np.zeros(2 << (2 * n))
    v[0] = 1.0
    for g in lc.gates():
        sg = g.to_density_matrix_gate(n)
        oracle.apply_gate(v, np.asarray(sg.matrix()), list(sg.qubits()))
    rho = to_complex(v).reshape(1 << n, 1 << n).T
    assert abs(np.trace(rho) - 1) < 1e-12
    exact = float(np.real(np.sum(np.diag(rho) * zzv)))
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.0, D, D, 77), count)
    z0 = np.zeros(2 << n)
    z0[0] = 1.0
    zz, nrm = [], []
    for s in S:
        pr = np.abs(to_complex(oracle_run(cs.trajectory_circuit(lc, s).gates(), z0))) ** 2
        zz.append(float(np.sum(pr * zzv)))
        nrm.append(float(pr.sum()))
    zz, nrm = np.array(zz), np.array(nrm)
    assert nrm.std() > 0.05  # the branches really change the norm
    ratio = zz.sum() / nrm.sum()
    sigma = np.std(zz - ratio * nrm, ddof=1) / np.sqrt(count) / nrm.mean()
    assert abs(ratio - exact) < 3 * sigma + 1e-12, (ratio, exact, sigma)
