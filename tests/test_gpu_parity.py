"""GPU parity: the gfx950 kernels, called through the C ABI (directly via
cast_simulator_amd.abi, or through the host kernel manager which binds the
same symbols), against the oracle on identical seeded inputs.

Tolerances (from the reference's tests): per-amplitude |delta| < 1e-10 for
f64 and < 1e-5 for f32 (tests/cpu_cuda_compare.rs:27, 397); circuit-level
relative L2 error <= 1e-12 (f64) and <= 1e-5 (f32).
"""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs
import oracle
from cast_simulator_amd import abi
from helpers import (
    S2,
    assert_amps_close,
    fidelity,
    gate_list,
    rel_l2,
    seeded_state,
)

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-10, 1e-5


def spec(f32=False):
    return cs.HipKernelGenSpec.f32() if f32 else cs.HipKernelGenSpec.f64()


def prec(f32=False):
    return cs.Precision.F32 if f32 else cs.Precision.F64


def apply_hip(gates, init, f32=False):
    n = (init.size // 2).bit_length() - 1
    mgr = cs.HipKernelManager(spec(f32))
    sv = cs.HipStatevector(n, prec(f32))
    sv.upload(init)
    for g in gates:
        mgr.apply(mgr.generate(g), sv)
    mgr.sync()
    return np.asarray(sv.download())


def apply_oracle(gates, init, f32=False):
    x = init.astype(np.float32) if f32 else init.copy()
    oracle.apply_gates(x, gate_list(gates))
    return x.astype(np.float64)


def compare(gate, n, label, f32=False):
    init = seeded_state(n)
    got = apply_hip([gate], init, f32)
    want = apply_oracle([gate], init, f32)
    assert_amps_close(want, got, TOL32 if f32 else TOL64, label)


# ── src/cuda/tests.rs:79-187 known answers ────────────────────────────────

def test_statevector_zero_state():
    sv = cs.HipStatevector(3, cs.Precision.F64)
    sv.zero()
    a = np.asarray(sv.download())
    assert a[0] == 1.0 and np.all(a[1:] == 0.0)


def test_statevector_upload_download_roundtrip():
    data = np.array([v for i in range(4) for v in (i * 0.1, i * -0.05)])
    sv = cs.HipStatevector(2, cs.Precision.F64)
    sv.upload(data)
    assert np.array_equal(np.asarray(sv.download()), data)


@pytest.mark.parametrize("f32", [False, True])
def test_h_x_known_answers(f32):
    mgr = cs.HipKernelManager(spec(f32))
    h, x = mgr.generate(cs.QuantumGate.h(0)), mgr.generate(cs.QuantumGate.x(0))
    sv = cs.HipStatevector(1, prec(f32))
    sv.zero()
    mgr.apply(h, sv)
    mgr.sync()
    assert np.allclose(sv.download(), [S2, 0, S2, 0], atol=1e-7 if f32 else 1e-15)
    sv.zero()
    mgr.apply(x, sv)
    mgr.apply(h, sv)
    mgr.sync()
    assert np.allclose(sv.download(), [S2, 0, -S2, 0], atol=1e-7 if f32 else 1e-15)
    big = cs.HipStatevector(4, prec(f32))
    big.zero()
    mgr.apply(h, big)
    mgr.sync()
    want = np.zeros(32)
    want[0] = want[2] = S2
    assert np.allclose(big.download(), want, atol=1e-7 if f32 else 1e-15)


# ── tests/cpu_cuda_compare.rs gate list (Haar cases re-seeded) ─────────────

T_GATE = cs.QuantumGate(np.array([[1, 0], [0, S2 + 1j * S2]]), [0])
COMPARE = [
    (cs.QuantumGate.h(0), 2), (cs.QuantumGate.x(0), 2), (cs.QuantumGate.h(0), 8),
    (cs.QuantumGate.h(3), 6), (T_GATE, 5), (cs.QuantumGate.rx(np.pi / 3, 0), 5),
    (cs.QuantumGate.rz(np.pi / 5, 0), 5), (cs.QuantumGate.cx(0, 1), 5),
    (cs.QuantumGate.cx(0, 4), 6), (cs.QuantumGate.cx(3, 1), 5),
    (cs.QuantumGate.random_unitary([0, 1], 11), 5), (cs.QuantumGate.random_unitary([1, 4], 12), 6),
    (cs.QuantumGate.ccx(0, 1, 2), 5), (cs.QuantumGate.ccx(0, 2, 5), 6),
    (cs.QuantumGate.random_unitary([0, 1, 2], 13), 5), (cs.QuantumGate.h(0), 20),
    (cs.QuantumGate.cx(0, 19), 20), (cs.QuantumGate.y(0), 4),
    (cs.QuantumGate.ry(np.pi / 4, 0), 5), (cs.QuantumGate.swap(0, 1), 5),
    (cs.QuantumGate.swap(1, 4), 6), (cs.QuantumGate.cz(0, 1), 5),
    (cs.QuantumGate.random_unitary([0, 1, 2, 3], 14), 6),
]


@pytest.mark.parametrize("gate,n", COMPARE, ids=[f"{i}" for i in range(len(COMPARE))])
def test_compare_f64(gate, n):
    compare(gate, n, f"{gate!r} on {n}q")


F32_COMPARE = [(cs.QuantumGate.h(0), 4), (cs.QuantumGate.cx(0, 1), 5),
               (cs.QuantumGate.rx(np.pi / 7, 0), 5), (cs.QuantumGate.random_unitary([0, 1], 15), 5)]


@pytest.mark.parametrize("gate,n", F32_COMPARE, ids=[f"{i}" for i in range(len(F32_COMPARE))])
def test_compare_f32(gate, n):
    compare(gate, n, f"{gate!r} on {n}q (f32)", f32=True)


def test_sequential_h_then_x_and_h_squared():
    init = seeded_state(4)
    gates = [cs.QuantumGate.h(0), cs.QuantumGate.x(0)]
    assert_amps_close(apply_oracle(gates, init), apply_hip(gates, init), TOL64, "H then X")
    hh = apply_hip([cs.QuantumGate.h(0), cs.QuantumGate.h(0)], seeded_state(3))
    assert_amps_close(seeded_state(3), hh, TOL64, "H^2 == I")


def test_multi_kernel_same_manager_and_dedup():
    mgr = cs.HipKernelManager(spec())
    a = mgr.generate(cs.QuantumGate.h(0))
    b = mgr.generate(cs.QuantumGate.h(0))
    c = mgr.generate(cs.QuantumGate.cx(0, 1))
    assert a == b and a != c and mgr.n_kernels() == 2
    init = seeded_state(5)
    sv = cs.HipStatevector(5, cs.Precision.F64)
    sv.upload(init)
    mgr.apply(a, sv)
    mgr.apply(c, sv)
    kernels, wall = mgr.sync()
    assert len(kernels) == 2 and all(t >= 0 for _, t in kernels)
    want = apply_oracle([cs.QuantumGate.h(0), cs.QuantumGate.cx(0, 1)], init)
    assert_amps_close(want, sv.download(), TOL64, "multi-kernel")


# ── every gate size and target shape, both precisions, every variant ──────

SHAPES = [
    [0], [7], [13], [0, 1], [0, 9], [5, 12], [0, 1, 2], [1, 4, 11], [8, 9, 10],
    [0, 1, 2, 3], [0, 5, 6, 13], [9, 10, 11, 12], [0, 1, 2, 3, 4], [1, 3, 5, 7, 9],
    [2, 8, 10, 12, 13], [3, 6, 8, 11, 13], [0, 1, 2, 3, 4, 5], [3, 4, 7, 9, 12, 13],
]


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("qubits", SHAPES, ids=["-".join(map(str, q)) for q in SHAPES])
def test_dense_every_shape(qubits, f32):
    n = 14
    g = cs.QuantumGate.random_unitary(qubits, 100 + sum(qubits))
    compare(g, n, f"dense {qubits}", f32)


def apply_abi(matrix, qubits, init, f32, variant, n, flags=0):
    p = abi.F32 if f32 else abi.F64
    plan = abi.plan_create(np.asarray(matrix), list(qubits), p, variant=variant, flags=flags)
    d = abi.sv_alloc(n, p)
    try:
        abi.sv_upload(d, init, p)
        abi.gate_apply(plan, d, n)
        abi.stream_sync(None)
        return abi.sv_download(d, n, p)
    finally:
        abi.sv_free(d)
        abi.plan_destroy(plan)


VARIANTS = [(abi.VARIANT_AUTO, 0), (abi.VARIANT_DENSE, 0), (abi.VARIANT_TILE, 0),
            (abi.VARIANT_TILE, abi.FLAG_TILE_SWIZZLE), (abi.VARIANT_TILE, abi.FLAG_TILE_GAUSS),
            (abi.VARIANT_TILE, abi.FLAG_TILE_GAUSS | abi.FLAG_TILE_SWIZZLE),
            (abi.VARIANT_TILE, abi.FLAG_TILE_XCD),
            (abi.VARIANT_TILE, abi.FLAG_TILE_GAUSS | abi.FLAG_TILE_XCD),
            (abi.VARIANT_MFMA, 0), (abi.VARIANT_MFMA, abi.FLAG_MFMA_F32),
            (abi.VARIANT_MFMA, abi.FLAG_TILE_XCD), (abi.VARIANT_MFMA, abi.FLAG_TILE_SWIZZLE),
            (abi.VARIANT_TILE, abi.FLAG_TILE_XCD | abi.FLAG_TILE_SWIZZLE),
            (abi.VARIANT_MFMA, abi.FLAG_TILE_XCD | abi.FLAG_TILE_SWIZZLE)]


@pytest.mark.parametrize("variant,flags", VARIANTS,
                         ids=["auto", "direct", "tile", "tile-swizzled", "tile-gauss",
                              "tile-gauss-swizzled", "tile-xcd", "tile-gauss-xcd", "mfma",
                              "mfma-f32-products", "mfma-xcd", "mfma-swizzled",
                              "tile-xcd-swizzled", "mfma-xcd-swizzled"])
@pytest.mark.parametrize("f32", [False, True], ids=["f64", "f32"])
def test_every_kernel_family_every_shape(f32, variant, flags):
    """Each dense kernel family, forced through the C ABI, on every target
    shape (low/high/mixed targets, k = 1..6) of a 14-qubit statevector (16
    for the XCD-blocked order, which needs at least 32 tiles) and of the
    smallest statevector each family accepts; the XCD-blocked order of the
    matrix-core kernels needs a grid of at least 8 workgroups)."""
    big = 16 if flags & abi.FLAG_TILE_XCD else 14
    for qubits in SHAPES:
        g = cs.QuantumGate.random_unitary(qubits, 300 + sum(qubits))
        for n in (big, len(qubits) + 6, max(qubits) + 1):
            if max(qubits) >= n:
                continue
            init = seeded_state(n)
            want = apply_oracle([g], init, f32)
            got = apply_abi(g.matrix(), qubits, init, f32, variant, n, flags)
            assert_amps_close(want, got, TOL32 if f32 else TOL64, f"{qubits} n={n} v={variant}")


@pytest.mark.parametrize("f32", [False, True])
def test_dense_nonunitary_sparse(f32):
    for seed, q in enumerate([[0, 3], [1, 2, 6], [0, 4, 5, 9]]):
        g = cs.QuantumGate.random_sparse(q, 0.4, seed)
        init = seeded_state(10)
        want = apply_oracle([g], init, f32)
        got = apply_hip([g], init, f32)
        tol = 1e-5 * max(1.0, np.max(np.abs(want))) if f32 else 1e-10
        assert_amps_close(want, got, tol, f"sparse {q}")


def test_monomial_variant_and_identity_rows():
    mgr = cs.HipKernelManager(spec())
    for g, nnz in [(cs.QuantumGate.cz(2, 5), 1), (cs.QuantumGate.cp(0.7, 0, 3), 1),
                   (cs.QuantumGate.swap(1, 4), 2), (cs.QuantumGate.cx(0, 6), 2),
                   (cs.QuantumGate.ccx(1, 2, 3), 2), (cs.QuantumGate.x(3), 2)]:
        info = mgr.plan_info(mgr.generate(g))
        assert info.variant == abi.VARIANT_MONOMIAL and info.nnz_rows == nnz
        compare(g, 9, f"monomial {g!r}")
    dense = mgr.plan_info(mgr.generate(cs.QuantumGate.h(0)))
    assert dense.variant == abi.VARIANT_DENSE


def test_forced_dense_equals_monomial():
    g = cs.QuantumGate.cp(0.3, 1, 6).matmul(cs.QuantumGate.swap(2, 6))
    init = seeded_state(9)
    outs = []
    for variant in (abi.VARIANT_DENSE, abi.VARIANT_MONOMIAL):
        p = abi.plan_create(np.asarray(g.matrix()), list(g.qubits()), abi.F64, variant=variant)
        d = abi.sv_alloc(9, abi.F64)
        try:
            abi.sv_upload(d, init, abi.F64)
            abi.gate_apply(p, d, 9)
            abi.stream_sync(None)
            outs.append(abi.sv_download(d, 9, abi.F64))
        finally:
            abi.sv_free(d)
            abi.plan_destroy(p)
    assert_amps_close(outs[0], outs[1], 1e-14, "dense vs monomial")


# ── config 1: 20q QFT, fusion <= 2, seeded Gaussian state ─────────────────

def test_qft20_fusion2_matches_oracle():
    n = 20
    g = cs.CircuitGraph.from_circuit(cs.QuantumCircuit.qft(n, True))
    cs.optimize(g, cs.FusionConfig.size_only(2))
    gates = g.gates_in_row_order()
    init = np.asarray(cs.random_statevector(n, 20260101))
    got = apply_hip(gates, init)
    want = apply_oracle(gates, init)
    assert_amps_close(want, got, TOL64, "QFT20")
    assert rel_l2(want, got) <= 1e-12


# ── random layer circuits, fused k=3..5, f64 and f32 ──────────────────────

@pytest.mark.parametrize("n,depth,size", [(12, 8, 3), (14, 10, 4), (16, 6, 5)])
def test_random_layers_fused(n, depth, size):
    c = cs.random_layer_circuit(n, depth, 1000 + n)
    g = cs.CircuitGraph.from_circuit(c)
    unfused = g.gates_in_row_order()
    cs.optimize(g, cs.FusionConfig.size_only(size))
    fused = g.gates_in_row_order()
    init = np.asarray(cs.random_statevector(n, 7 + n))
    want = apply_oracle(unfused, init)
    got64 = apply_hip(fused, init)
    assert rel_l2(want, got64) <= 1e-12
    got32 = apply_hip(fused, init, f32=True)
    assert rel_l2(want, got32) <= 1e-5
    assert fidelity(want, got32) >= 1 - 1e-6


# ── reductions ────────────────────────────────────────────────────────────

@pytest.mark.parametrize("f32", [False, True])
def test_norm_known_answer_and_normalize(f32):
    sv = cs.HipStatevector(0, prec(f32))
    sv.upload(np.array([3.0, 4.0]))
    assert sv.norm_squared() == 25.0
    sv.normalize()
    assert abs(sv.norm_squared() - 1.0) < (1e-7 if f32 else 1e-15)
    z = cs.HipStatevector(12, prec(f32))
    z.zero()
    assert z.norm_squared() == 1.0


def test_norm_and_scale_match_oracle():
    init = seeded_state(15) * 3.0
    for f32 in (False, True):
        sv = cs.HipStatevector(15, prec(f32))
        sv.upload(init)
        host = init.astype(np.float32) if f32 else init.copy()
        assert sv.norm_squared() == pytest.approx(oracle.norm_squared(host), rel=1e-13)
        d = sv.dptr()
        abi.scale(d, 15, abi.F32 if f32 else abi.F64, 0.25)
        oracle.scale(host, 0.25)
        assert np.max(np.abs(np.asarray(sv.download()) - host.astype(np.float64))) < 1e-15


def test_expect_zz_matches_oracle():
    n = 12
    init = np.asarray(cs.random_statevector(n, 5))
    pairs = [(i, i + 1) for i in range(n - 1)] + [(0, 11), (3, 7)]
    sv = cs.HipStatevector(n, cs.Precision.F64)
    sv.upload(init)
    got = sv.expect_zz([q for p in pairs for q in p])
    want = oracle.expect_zz(init, pairs)
    assert np.max(np.abs(np.asarray(got) - want)) < 1e-14


# ── batched launch == individual launches ─────────────────────────────────

@pytest.mark.parametrize("f32", [False, True])
def test_batch_apply_matches_individual(f32):
    p = abi.F32 if f32 else abi.F64
    items = [(10, [0, 4, 7]), (12, [2, 3, 11]), (11, [1, 5, 10]), (10, [7, 8, 9])]
    plans, svs, ns, inits = [], [], [], []
    try:
        for i, (n, q) in enumerate(items):
            m = np.asarray(cs.QuantumGate.random_unitary(q, 50 + i).matrix())
            plans.append(abi.plan_create(m, q, p, variant=abi.VARIANT_DENSE))
            d = abi.sv_alloc(n, p)
            svs.append(d)
            ns.append(n)
            inits.append(seeded_state(n))
            abi.sv_upload(d, inits[-1], p)
        abi.gate_apply_batch(plans, svs, ns)
        abi.stream_sync(None)
        for i, (n, q) in enumerate(items):
            got = abi.sv_download(svs[i], n, p)
            m = np.asarray(cs.QuantumGate.random_unitary(q, 50 + i).matrix())
            want = inits[i].astype(np.float32) if f32 else inits[i].copy()
            oracle.apply_gate(want, m, q)
            assert_amps_close(want.astype(np.float64), got, TOL32 if f32 else TOL64, f"item {i}")
    finally:
        for d in svs:
            abi.sv_free(d)
        for pl in plans:
            abi.plan_destroy(pl)


def test_batch_rejects_mixed_k():
    a = abi.plan_create(np.eye(2), [0], abi.F64)
    b = abi.plan_create(np.eye(4), [0, 1], abi.F64)
    d = abi.sv_alloc(4, abi.F64)
    try:
        with pytest.raises(abi.AbiError, match="share k and precision"):
            abi.gate_apply_batch([a, b], [d, d], [4, 4])
    finally:
        abi.sv_free(d)
        abi.plan_destroy(a)
        abi.plan_destroy(b)


# ── copies, events, memory, host-forceable errors ─────────────────────────

def test_clone_device_and_events():
    init = seeded_state(10)
    sv = cs.HipStatevector(10, cs.Precision.F64)
    sv.upload(init)
    cp = sv.clone_device()
    assert np.array_equal(np.asarray(cp.download()), init)
    s = abi.stream_create()
    e0, e1 = abi.event_create(), abi.event_create()
    try:
        abi.event_record(e0, s)
        abi.event_record(e1, s)
        abi.event_synchronize(e1)
        assert abi.event_elapsed_ms(e0, e1) >= 0.0
    finally:
        abi.event_destroy(e0)
        abi.event_destroy(e1)
        abi.stream_destroy(s)
    free, total = abi.free_memory()
    assert 0 < free <= total
    info = abi.device_info()
    assert info.arch.startswith(b"gfx950") and info.n_cu > 0 and info.wavefront_size == 64


def test_apply_errors():
    mgr = cs.HipKernelManager(spec())
    kid = mgr.generate(cs.QuantumGate.random_unitary([0, 1, 2], 1))
    small = cs.HipStatevector(2, cs.Precision.F64)
    with pytest.raises(cs.CastError, match="fewer qubits"):
        mgr.apply(kid, small)
    wrong = cs.HipStatevector(4, cs.Precision.F32)
    with pytest.raises(cs.CastError, match="precision"):
        mgr.apply(kid, wrong)
    with pytest.raises(cs.CastError, match="not found"):
        mgr.apply(9999, wrong)
    hi = mgr.generate(cs.QuantumGate.h(7))
    with pytest.raises(cs.CastError, match="outside"):
        mgr.apply(hi, cs.HipStatevector(5, cs.Precision.F64))
    p = abi.plan_create(np.eye(2), [6], abi.F64)
    d = abi.sv_alloc(4, abi.F64)
    try:
        with pytest.raises(abi.AbiError, match="outside"):
            abi.gate_apply(p, d, 4)
    finally:
        abi.sv_free(d)
        abi.plan_destroy(p)


def test_noisy_gate_rejected_by_simulator():
    c = cs.QuantumCircuit(2)
    c.add(cs.QuantumGate.depolarizing(0, 0.1))
    with pytest.raises(cs.CastError, match="noisy"):
        cs.Simulator.f64().simulate(cs.CircuitGraph.from_circuit(c))


# ── per-gate kernel autotuning ────────────────────────────────────────────

@pytest.mark.parametrize("f32", [False, True])
def test_autotuned_kernels_match_oracle(f32):
    n = 22
    gates = [cs.QuantumGate.random_unitary(q, 40 + i) for i, q in
             enumerate([[0, 1, 2, 3], [3, 9, 14, 20, 21], [5, 17], [1, 2, 8, 16, 19]])]
    mgr = cs.HipKernelManager(spec(f32))
    mgr.set_autotune(True, min_qubits=22, samples=1)
    ids = [mgr.generate(g) for g in gates]
    scratch = cs.HipStatevector(n, prec(f32))
    scratch.zero()
    assert mgr.tune(ids, scratch) >= 1
    report = mgr.tuning_report()
    # chosen: direct (1), wave tile (3), matrix core (4), wave tile swizzled (5),
    # three products (6), both (7), XCD order (8), three products in XCD order
    # (9), matrix core with f32 products at f32 k >= 5 (10), matrix core in XCD
    # order (11) and in swizzled order (12), wave tiles in XCD order with the
    # low tile bits bit-reversed (13) and the same in three products (14), matrix
    # core in both orders (15)
    assert len(report) == len(ids) and all(e["chosen"] in range(1, 16) and e["chosen"] != 2
                                           for e in report)
    for e in report:
        codes = [c for c, _ in e["best_s"]]
        k = len(gates[e["id"]].qubits())
        assert (4 in codes and 11 in codes and 12 in codes) == (k >= 4)  # matrix-core orders
        if f32:  # both matrix-core product forms are candidates for the k = 5 gates
            assert (4 in codes and 10 in codes) == (k == 5)
    init = np.asarray(cs.random_statevector(n, 3))
    sv = cs.HipStatevector(n, prec(f32))
    sv.upload(init)
    for i in ids:
        mgr.apply(i, sv)
    mgr.sync()
    want = apply_oracle(gates, init, f32)
    assert rel_l2(want, sv.download()) <= (1e-5 if f32 else 1e-12)


# ── full-size properties (no host transfer of the 16 GiB state) ───────────

@pytest.mark.parametrize("f32", [False, True])
def test_roundtrip_at_30_qubits(f32):
    """U then U^dagger on the 30-qubit statevector returns |0...0>: norm 1 and
    every <Z_i Z_j> = 1, checked on the device."""
    n = 30
    rng_q = [[0, 1, 2, 3, 4], [5, 11, 17, 23, 29], [1, 9, 15, 26], [0, 29], [14]]
    mgr = cs.HipKernelManager(spec(f32))
    sv = cs.HipStatevector(n, prec(f32))
    sv.zero()
    fwd = [cs.QuantumGate.random_unitary(q, 900 + i) for i, q in enumerate(rng_q)]
    inv = [cs.QuantumGate(np.asarray(g.matrix()).conj().T, list(g.qubits())) for g in fwd]
    for g in fwd + inv[::-1]:
        mgr.apply(mgr.generate(g), sv)
    mgr.sync()
    tol = 1e-5 if f32 else 1e-12
    assert abs(sv.norm_squared() - 1.0) < tol
    zz = np.asarray(sv.expect_zz([0, 1, 4, 5, 28, 29, 0, 29]))
    assert np.all(np.abs(zz - 1.0) < tol)


ROUNDTRIP_SHAPES = [[0, 1, 2, 3, 4], [5, 11, 17, 23, 29], [1, 9, 15, 26], [3, 6, 8, 11, 13, 27]]


@pytest.mark.parametrize("variant,flags", VARIANTS[1:],
                         ids=["direct", "tile", "tile-swizzled", "tile-gauss",
                              "tile-gauss-swizzled", "tile-xcd", "tile-gauss-xcd", "mfma",
                              "mfma-f32-products", "mfma-xcd", "mfma-swizzled",
                              "tile-xcd-swizzled", "mfma-xcd-swizzled"])
@pytest.mark.parametrize("f32", [False, True], ids=["f64", "f32"])
def test_roundtrip_at_30_qubits_every_family(f32, variant, flags):
    """Each kernel family forced through the C ABI at the headline size: the
    gates (k = 4..6, low, high and mixed targets) and then their adjoints in
    reverse order return |0...0>; norm and <Z_i Z_j> checked on the device,
    so the 16 GiB state never crosses PCIe."""
    n, p = 30, (abi.F32 if f32 else abi.F64)
    print(f"[parity] 30q roundtrip variant={variant} flags={flags} f32={f32}", flush=True)
    fwd = [cs.QuantumGate.random_unitary(q, 950 + i) for i, q in enumerate(ROUNDTRIP_SHAPES)]
    mats = [np.asarray(g.matrix()) for g in fwd] + [np.asarray(g.matrix()).conj().T
                                                    for g in fwd[::-1]]
    qs = [list(g.qubits()) for g in fwd] + [list(g.qubits()) for g in fwd[::-1]]
    d = abi.sv_alloc(n, p)
    plans = []
    try:
        abi.sv_zero(d, n, p)
        for m, q in zip(mats, qs):
            plans.append(abi.plan_create(m, q, p, variant=variant, flags=flags))
            abi.gate_apply(plans[-1], d, n)
        abi.stream_sync(None)
        tol = 1e-5 if f32 else 1e-12
        assert abs(abi.norm_squared(d, n, p) - 1.0) < tol
        zz = abi.expect_zz(d, n, p, [(0, 1), (4, 5), (28, 29), (0, 29), (3, 27)])
        assert np.all(np.abs(zz - 1.0) < tol), zz
    finally:
        for pl in plans:
            abi.plan_destroy(pl)
        abi.sv_free(d)


@pytest.mark.parametrize("f32", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("qubits", [[0, 1, 2, 3, 4, 5], [2, 9, 14, 17, 20, 23], [18, 19, 20, 21, 22, 23]],
                         ids=["low", "mixed", "high"])
def test_mfma_k6_at_24_qubits_matches_direct_kernel(qubits, f32):
    """k = 6 on matrix cores with several 16-combo tiles per wave (2^24
    amplitudes) against the direct kernel (itself pinned to the oracle on the
    same shapes above)."""
    n = 24
    g = cs.QuantumGate.random_unitary(qubits, 77 + sum(qubits))
    init = np.asarray(cs.random_statevector(n, 5))
    print(f"[parity] k=6 {qubits} n={n} f32={f32}", flush=True)
    a = apply_abi(g.matrix(), qubits, init, f32, abi.VARIANT_MFMA, n)
    b = apply_abi(g.matrix(), qubits, init, f32, abi.VARIANT_DENSE, n)
    assert_amps_close(b, a, TOL32 if f32 else TOL64, f"k6 {qubits}")
    plan = abi.plan_create(np.asarray(g.matrix()), qubits, abi.F32 if f32 else abi.F64,
                           variant=abi.VARIANT_AUTO)
    try:
        assert abi.plan_family(plan, n) == abi.VARIANT_MFMA
    finally:
        abi.plan_destroy(plan)


@pytest.mark.parametrize("pa,pb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_inner_product_and_fidelity_match_numpy(pa, pb):
    n = 13
    x = np.asarray(cs.random_statevector(n, 3))
    y = np.asarray(cs.random_statevector(n, 4))
    prec = {0: cs.Precision.F32, 1: cs.Precision.F64}
    a, b = cs.HipStatevector(n, prec[pa]), cs.HipStatevector(n, prec[pb])
    a.upload(x)
    b.upload(y)
    xa, yb = np.asarray(a.download()), np.asarray(b.download())  # rounded as stored
    za, zb = xa[0::2] + 1j * xa[1::2], yb[0::2] + 1j * yb[1::2]
    want = np.vdot(za, zb)
    got = a.inner(b)
    assert abs(got - want) < 1e-12
    f = abs(want) ** 2 / (np.vdot(za, za).real * np.vdot(zb, zb).real)
    assert abs(a.fidelity(b) - f) < 1e-12
    assert abs(a.fidelity(a) - 1.0) < 1e-12
    with pytest.raises(cs.CastError):
        a.inner(cs.HipStatevector(n - 1, prec[pb]))


@pytest.mark.parametrize("qubits", [[0, 1, 2, 3, 4], [2, 5, 9, 13, 15], [11, 12, 13, 14, 15],
                                    [0, 1, 2, 3, 4, 5], [2, 5, 9, 11, 13, 15],
                                    [10, 11, 12, 13, 14, 15], [0, 3, 7, 8, 12, 13]])
def test_f32_split_products_accuracy(qubits):
    """f32 k = 5, 6 on the matrix cores: the bf16 split products (every f32
    operand in three bf16 parts, six partial products) against the f64 oracle
    on the same f32-rounded input: relative L2 error within 1.5e-7 (f32
    output rounding is about 6e-8), and no larger than the f32 products'."""
    n = 16
    g = cs.QuantumGate.random_unitary(qubits, 900 + sum(qubits))
    init = seeded_state(n).astype(np.float32).astype(np.float64)
    want = apply_oracle([g], init, False)
    errs = {}
    for name, fl in (("split", abi.FLAG_MFMA_SPLIT), ("f32", abi.FLAG_MFMA_F32)):
        got = apply_abi(g.matrix(), qubits, init, True, abi.VARIANT_MFMA, n, fl)
        errs[name] = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    print(f"[parity] split products {qubits}: {errs}", flush=True)
    assert errs["split"] < 1.5e-7, errs
    assert errs["split"] <= errs["f32"], errs
