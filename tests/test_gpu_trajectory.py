"""Persistent batched-trajectory kernel on the GPU against the oracle.

Each task's final state (dumped) is compared with the oracle's simulation of
the trajectory's explicit gate list (source gates with the sampled Paulis
inserted); the per-task f64 norm and <Z_a Z_b> sums with the oracle's own
reductions of that state.  Sizes cover the one-combo-per-thread path
(n < k + 6), the wave-tile path with one tile per step (n < 10) and with
16/K tiles per step, every launch shape, few slots (many tasks per slot) and
more slots than tasks."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import cast_simulator_amd as cs
import oracle
from cast_simulator_amd import abi
from helpers import rel_l2, to_complex

pytestmark = pytest.mark.gpu

F64, F32 = cs.Precision.F64, cs.Precision.F32


def oracle_final(lc, s, init):
    sv = np.array(init, dtype=np.float64)
    for g in cs.trajectory_circuit(lc, s).gates():
        oracle.apply_gate(sv, np.asarray(g.matrix()), list(g.qubits()))
    return sv


def check_run(lc, S, prec, tuning, pairs, init=None, tol=None):
    n = lc.n_qubits
    eng = cs.TrajectoryEngine(lc, prec, [q for p in pairs for q in p])
    if init is None:
        init = np.zeros(2 << n)
        init[0] = 1.0
    else:
        eng.set_initial_state(list(init))
    print(f"[traj] n={n} prec={prec} tuning={tuning} trajectories={len(S)}", flush=True)
    r = eng.run(S, tuning, True)
    dump = r.dump().reshape(r.n_tasks, 2 << n)
    tol = tol if tol is not None else (1e-12 if prec == F64 else 1e-5)
    seen = set()
    for t, s in enumerate(S):
        k = r.task_of[t]
        want = oracle_final(lc, s, init)
        got = dump[k]
        assert rel_l2(want, got) < tol, f"trajectory {t} task {k}"
        if k in seen:
            continue
        seen.add(k)
        ns = oracle.norm_squared(np.array(got))
        zz = oracle.expect_zz(np.array(got), pairs)
        P = len(pairs)
        assert abs(r.norm_sq[k] - ns) < 1e-12 * max(1.0, ns)
        assert np.allclose(r.zz_sum[k * P:(k + 1) * P], zz, rtol=0, atol=1e-12)
    # host-side f64 results: per-trajectory normalised values and the batch
    # ratio sum_t (1/P) sum_j zz / sum_t norm
    P = len(pairs)
    ks = [r.task_of[t] for t in range(len(S))]
    zz_t = np.array([sum(r.zz_sum[k * P:(k + 1) * P]) / P for k in ks])
    nrm_t = np.array([r.norm_sq[k] for k in ks])
    assert np.allclose(r.traj_zz, zz_t / nrm_t, rtol=0, atol=1e-12)
    assert abs(r.mean_zz - zz_t.sum() / nrm_t.sum()) < 1e-12
    return r


@pytest.mark.parametrize("n", [5, 8, 9, 11, 13])
@pytest.mark.parametrize("prec", [F64, F32])
def test_tasks_match_oracle(n, prec):
    D = 8
    lc = cs.LayeredCircuit.random_layers(n, D, 100 + n)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.05, 1, D, 7 * n), 24)
    init = np.asarray(cs.random_statevector(n, n))
    pairs = [(i, i + 1) for i in range(n - 1)]
    r = check_run(lc, S, prec, cs.TrajectoryTuning(3, 4, 256, 2, 0), pairs, init)
    assert r.n_dirty_matrices > 0


@pytest.mark.parametrize("waves", [2, 4])
@pytest.mark.parametrize("block", [64, 128, 256, 512])
@pytest.mark.parametrize("max_slots", [1, 3, 0])
def test_launch_shapes(block, max_slots, waves):
    n, D = 12, 6
    lc = cs.LayeredCircuit.random_layers(n, D, 5)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.03, 1, D, block + max_slots), 20)
    pairs = [(0, 11), (3, 4), (10, 11), (5, 9)]
    r = check_run(lc, S, F32, cs.TrajectoryTuning(2, 3, block, 2, max_slots, waves), pairs)
    if max_slots:
        assert r.n_slots == min(max_slots, r.n_tasks)
    assert r.block == block and r.waves_per_simd == waves


@pytest.mark.parametrize("prec", [F64, F32])
def test_high_occupancy_build_matches_oracle(prec):
    n, D = 13, 8
    lc = cs.LayeredCircuit.random_layers(n, D, 44)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.05, 1, D, 45), 24)
    init = np.asarray(cs.random_statevector(n, 46))
    pairs = [(i, i + 1) for i in range(n - 1)]
    check_run(lc, S, prec, cs.TrajectoryTuning(3, 4, 512, 2, 0, 4), pairs, init)


@pytest.mark.parametrize("B,kmax", [(1, 2), (4, 4), (6, 3)])
def test_chunking_and_fusion_sizes(B, kmax):
    n, D = 10, 12
    lc = cs.LayeredCircuit.random_layers(n, D, 8)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.02, 1, D, B * 10 + kmax), 30)
    check_run(lc, S, F64, cs.TrajectoryTuning(B, kmax, 128, 2, 0), [(0, 1), (8, 9)])


def test_noiseless_limit_equals_statevector_expectations():
    n, D = 14, 10
    lc = cs.LayeredCircuit.random_layers(n, D, 9)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.0, 1, D, 3), 64)
    pairs = [(i, i + 1) for i in range(n - 1)]
    eng = cs.TrajectoryEngine(lc, F64, [q for p in pairs for q in p])
    r = eng.run(S, cs.TrajectoryTuning(4, 4, 256, 2, 0))
    assert r.n_tasks == len({s.depth for s in S})
    z0 = np.zeros(2 << n)
    z0[0] = 1.0
    for t, s in enumerate(S[:10]):
        sv = oracle_final(lc, s, z0)
        want = np.mean(oracle.expect_zz(sv, pairs))
        assert abs(r.traj_zz[t] - want) < 1e-12
    assert abs(r.min_norm_sq - 1.0) < 1e-12 and abs(r.max_norm_sq - 1.0) < 1e-12


@pytest.mark.parametrize("kmax32,mf32", [(4, 0), (6, 0), (6, 1)],
                         ids=["k4", "k6-split", "k6-f32products"])
def test_f32_matches_f64_at_24_qubits(kmax32, mf32):
    """The benchmark's statevector size: f32 trajectories (k <= 4 on wave
    tiles, or k <= 6 with the k = 5, 6 gates on matrix cores, as bf16 split
    or f32 products) against the f64 trajectories of the same samples
    (k <= 4), state by state."""
    n, D = 24, 12
    lc = cs.LayeredCircuit.random_layers(n, D, 2026)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(1e-2, 10, D, 4), 6)
    pairs = [(i, i + 1) for i in range(n - 1)]
    flat = [q for p in pairs for q in p]
    print(f"[traj] n=24 f32 (k<={kmax32}) vs f64 dump, 6 trajectories", flush=True)
    a = cs.TrajectoryEngine(lc, F64, flat).run(S, cs.TrajectoryTuning(4, 4, 256, 2, 0), True)
    b = cs.TrajectoryEngine(lc, F32, flat).run(
        S, cs.TrajectoryTuning(4, kmax32, 512, 1, 0, 2, mf32), True)
    assert b.n_dirty_matrices > 0
    da, db = a.dump().reshape(a.n_tasks, -1), b.dump().reshape(b.n_tasks, -1)
    for t in range(len(S)):
        x, y = to_complex(da[a.task_of[t]]), to_complex(db[b.task_of[t]])
        fid = abs(np.vdot(x, y)) ** 2 / (np.vdot(x, x).real * np.vdot(y, y).real)
        assert fid > 1 - 1e-6, f"trajectory {t}: fidelity {fid}"
        assert rel_l2(da[a.task_of[t]], db[b.task_of[t]].astype(np.float64)) <= 1e-5
        assert abs(a.traj_zz[t] - b.traj_zz[t]) < 1e-5
    assert abs(a.min_norm_sq - 1) < 1e-10 and abs(b.max_norm_sq - 1) < 1e-5


# ── C ABI level: invalid descriptors are skipped and flagged ────────────────

def test_abi_flags_invalid_descriptors_and_keeps_valid_tasks():
    n = 8
    svb = (2 << n) * 8
    h = np.zeros(2 << n)
    h[0] = 1.0
    x = np.array([[0, 1], [1, 0]], dtype=complex)
    pool = np.ascontiguousarray(x.reshape(-1).view(np.float64))
    bufs = []

    def dev(arr):
        a = np.ascontiguousarray(arr)
        p = abi.malloc(max(a.nbytes, 16))
        bufs.append(p)
        abi.memcpy_htod(p, a)
        return p

    try:
        gates = (abi.TrajGate * 3)()
        gates[0].matrix_offset, gates[0].k, gates[0].q[0] = 0, 1, 3      # X on qubit 3
        gates[1].matrix_offset, gates[1].k, gates[1].q[0] = 0, 1, n      # qubit out of range
        gates[2].matrix_offset, gates[2].k, gates[2].q[0] = 5, 1, 0      # matrix outside the pool
        tasks = (abi.TrajTask * 3)()
        tasks[0].first_gate, tasks[0].n_gates, tasks[0].checkpoint = 0, 1, 0
        tasks[1].first_gate, tasks[1].n_gates, tasks[1].checkpoint = 0, 3, 0
        tasks[2].first_gate, tasks[2].n_gates, tasks[2].checkpoint = 0, 1, 4  # bad checkpoint
        g_d = dev(np.frombuffer(bytes(gates), dtype=np.uint8))
        t_d = dev(np.frombuffer(bytes(tasks), dtype=np.uint8))
        o_d = dev(np.array([0, 1, 2], dtype=np.uint32))
        m_d = dev(pool)
        c_d = dev(h)
        s_d = abi.malloc(2 * svb)
        bufs.append(s_d)
        out_d = abi.malloc(3 * 2 * 8)
        bufs.append(out_d)
        w_d = abi.malloc(64)
        bufs.append(w_d)
        L = abi.TrajLaunch()
        L.precision, L.n_qubits = 1, n
        L.tasks, L.order, L.n_tasks = t_d, o_d, 3
        L.n_gates, L.gates, L.matrices, L.n_matrix_elems = 3, g_d, m_d, 4
        L.checkpoints, L.n_checkpoints = c_d, 1
        L.n_slots, L.slots, L.block = 2, s_d, 64
        L.n_pairs = 1
        L.pairs[0], L.pairs[1] = 3, 4
        L.out, L.dump, L.work = out_d, 0, w_d
        print("[traj] ABI invalid-descriptor test", flush=True)
        abi.traj_run(L)
        abi.stream_sync(None)
        work = np.zeros(2, dtype=np.uint32)
        abi.memcpy_dtoh(work, w_d)
        assert work[0] >= 3 and work[1] == 1
        out = np.zeros(6)
        abi.memcpy_dtoh(out, out_d)
        # task 0: X on qubit 3 of |0>: norm 1, <Z3 Z4> = -1
        assert out[0] == pytest.approx(1.0, abs=1e-15) and out[1] == pytest.approx(-1.0, abs=1e-15)
        # task 1: only its valid gate applied
        assert out[2] == pytest.approx(1.0, abs=1e-15) and out[3] == pytest.approx(-1.0, abs=1e-15)
    finally:
        for p in bufs:
            abi.free(p)


def test_abi_rejects_bad_launch_shapes():
    L = abi.TrajLaunch()
    L.precision, L.n_qubits, L.n_slots, L.block, L.work = 1, 8, 1, 96, 1
    with pytest.raises(abi.AbiError, match="block"):
        abi.traj_run(L)
    L.block, L.n_pairs = 64, 33
    with pytest.raises(abi.AbiError, match="32 qubit pairs"):
        abi.traj_run(L)
    L.n_pairs = 1
    L.pairs[0], L.pairs[1] = 2, 2
    with pytest.raises(abi.AbiError, match="pair 0"):
        abi.traj_run(L)
    L.pairs[1] = 3
    L.waves_per_simd = 3
    with pytest.raises(abi.AbiError, match="waves_per_simd"):
        abi.traj_run(L)
    L.waves_per_simd = 4
    L.precision = 2
    with pytest.raises(abi.AbiError, match="precision"):
        abi.traj_run(L)
    L.precision, L.n_tasks = 1, 4
    with pytest.raises(abi.AbiError, match="null task"):
        abi.traj_run(L)


def test_autotuner_returns_a_tried_tuning_that_matches_oracle():
    n, D = 12, 8
    lc = cs.LayeredCircuit.random_layers(n, D, 61)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.05, 1, D, 62), 40)
    pairs = [(i, i + 1) for i in range(n - 1)]
    eng = cs.TrajectoryEngine(lc, F32, [q for p in pairs for q in p])
    print("[traj] autotune n=12, 40 trajectories", flush=True)
    best, trials = eng.autotune(S, cs.TrajectoryTuning(2, 4, 256, 2, 0))
    assert len(trials) >= 10
    assert repr(best) in {repr(t) for t, _ in trials}
    assert all(t.waves_per_simd == 2 and t.slots_per_cu * t.block <= 512 for t, _ in trials)
    check_run(lc, S, F32, best, pairs)


@pytest.mark.parametrize("n", [9, 11, 13, 16])
def test_k5_gates_at_f32_match_oracle(n):
    """Fusion up to 5 qubits (f32 runs): one-combo-per-thread (n < 11), wave
    tiles with at most two targets below bit 6, and the one-combo path for
    k = 5 gates with more."""
    D = 8
    lc = cs.LayeredCircuit.random_layers(n, D, 700 + n)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.05, 1, D, 71 * n), 24)
    pairs = [(i, i + 1) for i in range(n - 1)]
    tn = cs.TrajectoryTuning(4, 5, 256, 2, 0)
    eng = cs.TrajectoryEngine(lc, F32, [q for p in pairs for q in p])
    P = eng.plan(S, tn)
    k5 = [q for _, q in P["gates"] if len(q) == 5]
    assert k5, "no 5-qubit fused gates to test"
    if n == 13:
        assert any(sum(b < 6 for b in q) > 2 for q in k5)   # the one-combo path
        assert any(sum(b < 6 for b in q) <= 2 for q in k5)  # the wave-tile path
    init = np.asarray(cs.random_statevector(n, n + 5))
    check_run(lc, S, F32, tn, pairs, init)


def test_k5_rejected_at_f64_by_planner_and_kernel():
    lc = cs.LayeredCircuit.random_layers(8, 4, 3)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.05, 1, 4, 4), 4)
    with pytest.raises(cs.CastError, match="1..4"):
        cs.TrajectoryEngine(lc, F64, [0, 1]).run(S, cs.TrajectoryTuning(2, 5, 256, 2, 0))


@pytest.mark.parametrize("n", [12, 13, 16])
def test_k6_gates_at_f32_match_oracle(n):
    """Fusion up to 6 qubits (f32 runs, n >= 12): the LDS-staged matrix-core
    path, alongside k <= 5 gates of every kind."""
    D = 8
    lc = cs.LayeredCircuit.random_layers(n, D, 900 + n)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.05, 1, D, 37 * n), 24)
    pairs = [(i, i + 1) for i in range(n - 1)]
    tn = cs.TrajectoryTuning(8, 6, 256, 2, 0)
    eng = cs.TrajectoryEngine(lc, F32, [q for p in pairs for q in p])
    P = eng.plan(S, tn)
    assert any(len(q) == 6 for _, q in P["gates"]), "no 6-qubit fused gates to test"
    init = np.asarray(cs.random_statevector(n, n + 9))
    check_run(lc, S, F32, tn, pairs, init)


def test_k6_needs_twelve_qubits():
    lc = cs.LayeredCircuit.random_layers(11, 4, 3)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.05, 1, 4, 4), 4)
    with pytest.raises(cs.CastError, match="1..5"):
        cs.TrajectoryEngine(lc, F32, [0, 1]).run(S, cs.TrajectoryTuning(2, 6, 256, 2, 0))


# ── circuits with their own noise models ────────────────────────────────────

def noisy_layers(n, depth, seed, gamma, pxx):
    """Random layers, a correlated XX branch (p = pxx) on every CZ, and
    amplitude damping (non-unitary branches) on qubits 0 and n - 1 after
    every layer; in moments."""
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    a0 = np.array([[1, 0], [0, np.sqrt(1 - gamma)]], dtype=complex)
    a1 = np.array([[0, np.sqrt(gamma)], [0, 0]], dtype=complex)
    p1 = 0.3
    qc = cs.QuantumCircuit(n)
    per_layer = n + n // 2
    for i, g in enumerate(cs.random_layer_circuit(n, depth, seed).gates()):
        if g.n_qubits() == 2:
            g = g.with_noise([(1 - pxx, np.eye(4)), (pxx, np.kron(X, X))])
        qc.add(g)
        if i % per_layer == per_layer - 1:
            for q in (0, n - 1):
                qc.add(cs.QuantumGate(np.eye(2), [q]).with_noise(
                    [(1 - p1, a0 / np.sqrt(1 - p1)), (p1, a1 / np.sqrt(p1))]))
    return cs.LayeredCircuit.from_circuit(qc)


@pytest.mark.parametrize("prec,kmax,mf32", [(F64, 4, 0), (F32, 4, 0), (F32, 6, 0), (F32, 6, 1)],
                         ids=["f64-k4", "f32-k4", "f32-k6-split", "f32-k6-f32products"])
def test_noise_model_tasks_match_oracle(prec, kmax, mf32):
    """Non-unitary and correlated branches sampled per trajectory (plus
    depolarizing Paulis): every dumped state against the oracle run of the
    trajectory's explicit gate list; norms differ from 1 and are divided out
    in f64.  f32 k = 5, 6 gates run as bf16 split products or f32 products."""
    n = 12
    lc = noisy_layers(n, 5, 61, gamma=0.3, pxx=0.1)
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.01, 1, lc.n_layers(), 62), 32)
    assert sum(1 for s in S for e in s.errors if e[2] == 0) > 10
    init = np.asarray(cs.random_statevector(n, 63))
    pairs = [(i, i + 1) for i in range(n - 1)]
    r = check_run(lc, S, prec, cs.TrajectoryTuning(2, kmax, 256, 2, 0, 2, mf32), pairs, init)
    assert r.min_norm_sq < 0.99 and r.mfma_f32 == mf32


def test_noise_model_mean_matches_density_matrix():
    """<(1/P) sum Z_i Z_i+1> over 4096 f64 trajectories of a 6-qubit noisy
    circuit against the exact density-matrix value (every noisy gate lifted
    to sum_b p_b K_b (x) conj(K_b), src/types/gate.rs:236-258, applied by the
    oracle), within 3 sigma of the ratio estimator."""
    n, count = 6, 4096
    lc = noisy_layers(n, 4, 71, gamma=0.5, pxx=0.2)
    D = lc.n_layers()
    pairs = [(i, i + 1) for i in range(n - 1)]
    v = np.zeros(2 << (2 * n))
    v[0] = 1.0
    for g in lc.gates():
        sg = g.to_density_matrix_gate(n)
        oracle.apply_gate(v, np.asarray(sg.matrix()), list(sg.qubits()))
    rho = to_complex(v).reshape(1 << n, 1 << n).T
    idx = np.arange(1 << n)
    zzv = np.mean([1 - 2 * (((idx >> a) ^ (idx >> b)) & 1) for a, b in pairs], axis=0)
    exact = float(np.real(np.sum(np.diag(rho) * zzv)))
    S = cs.sample_trajectories(lc, cs.NoiseSpec(0.0, D, D, 72), count)
    eng = cs.TrajectoryEngine(lc, F64, [q for p in pairs for q in p])
    r = eng.run(S, cs.TrajectoryTuning(2, 4, 256, 2, 0))
    P = len(pairs)
    zz = np.array([sum(r.zz_sum[r.task_of[t] * P:(r.task_of[t] + 1) * P]) / P
                   for t in range(count)])
    nrm = np.array([r.norm_sq[r.task_of[t]] for t in range(count)])
    sigma = np.std(zz - r.mean_zz * nrm, ddof=1) / np.sqrt(count) / nrm.mean()
    print(f"[traj] noise-model mean {r.mean_zz:.6f} exact {exact:.6f} sigma {sigma:.2e}",
          flush=True)
    assert abs(r.mean_zz - exact) < 3 * sigma, (r.mean_zz, exact, sigma)
