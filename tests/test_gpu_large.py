"""GPU checks at the large dense-fusion sizes: 33-qubit f64 (128 GiB) and
34-qubit f32 (128 GiB) statevectors under the random-layer circuit of depth
12 fused to k <= 6 (size-only fusion, automatic kernel families: k = 5, 6 on
the matrix cores), and k = 6 matrix-core launches above 30 qubits, where the
k = 6 grid rule gives more than two workgroups per CU.

No state crosses PCIe: norms, <Z_a Z_b> and f32-against-f64 fidelities are
reduced on the device in f64.  Bars: norm within 1e-12 (f64) / 1e-5 (f32) of
1, fidelity >= 1 - 1e-6, and a circuit followed by its adjoint in reverse
order returns |0...0> (every <Z_a Z_b> = 1 within the same bounds).
"""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs
from cast_simulator_amd import abi

pytestmark = pytest.mark.gpu

SEED = 20260101
PAIRS = [(0, 1), (5, 6), (15, 16), (0, 31), (30, 31), (3, 27)]


def fused_gates(n, depth, kmax=6):
    g = cs.CircuitGraph.from_circuit(cs.random_layer_circuit(n, depth, SEED))
    cs.optimize(g, cs.FusionConfig.size_only(kmax))
    return [(np.asarray(x.matrix()), list(x.qubits())) for x in g.gates_in_row_order()]


class Plans:
    """Gate plans of one precision through the C ABI, destroyed on exit."""

    def __init__(self, gates, precision, variant=abi.VARIANT_AUTO):
        self.ids = []
        try:
            for m, q in gates:
                self.ids.append(abi.plan_create(m, q, precision, variant=variant))
        except BaseException:
            self.close()
            raise

    def apply(self, d, n, order=None):
        for i in (range(len(self.ids)) if order is None else order):
            abi.gate_apply(self.ids[i], d, n)

    def close(self):
        for p in self.ids:
            abi.plan_destroy(p)
        self.ids = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def adjoints(gates):
    return [(m.conj().T.copy(), q) for m, q in gates[::-1]]


def fidelity(a, pa, b, pb, n):
    ab = abi.inner_product(a, pa, b, pb, n)
    return abs(ab) ** 2 / (abi.norm_squared(a, n, pa) * abi.norm_squared(b, n, pb))


def free_gib():
    return abi.free_memory()[0] / 2**30


def test_33_qubits_f64_and_f32_resident_fidelity_and_roundtrip():
    """33 qubits, depth 12, k <= 6: the f64 (128 GiB) and f32 (64 GiB) states
    of the same fused circuit are resident together; their device fidelity
    is >= 1 - 1e-6 and both norms hold.  Then the f64 state runs the adjoint
    circuit and returns to |0...0>."""
    n = 33
    if free_gib() < 200:
        pytest.skip(f"needs 192 GiB of device memory, {free_gib():.0f} GiB free")
    gates = fused_gates(n, 12)
    ks = [len(q) for _, q in gates]
    assert max(ks) == 6 and ks.count(6) > 0 and ks.count(5) > 0
    print(f"[large] 33q: {len(gates)} fused gates, k histogram "
          f"{ {k: ks.count(k) for k in sorted(set(ks))} }", flush=True)
    d64 = abi.sv_alloc(n, abi.F64)
    d32 = 0
    try:
        d32 = abi.sv_alloc(n, abi.F32)
        with Plans(gates, abi.F64) as p64, Plans(gates, abi.F32) as p32:
            fams = {abi.plan_family(p, n) for p in p64.ids}
            assert abi.VARIANT_MFMA in fams
            abi.sv_zero(d64, n, abi.F64)
            abi.sv_zero(d32, n, abi.F32)
            p64.apply(d64, n)
            p32.apply(d32, n)
            abi.stream_sync(None)
            print("[large] 33q f64 and f32 passes done", flush=True)
            assert abs(abi.norm_squared(d64, n, abi.F64) - 1.0) < 1e-12
            assert abs(abi.norm_squared(d32, n, abi.F32) - 1.0) < 1e-5
            fid = fidelity(d32, abi.F32, d64, abi.F64, n)
            print(f"[large] 33q f32 vs f64 fidelity {fid:.15f}", flush=True)
            assert fid >= 1.0 - 1e-6
        abi.sv_free(d32)
        d32 = 0
        with Plans(adjoints(gates), abi.F64) as inv:
            inv.apply(d64, n)
            abi.stream_sync(None)
        assert abs(abi.norm_squared(d64, n, abi.F64) - 1.0) < 1e-12
        zz = abi.expect_zz(d64, n, abi.F64, PAIRS + [(31, 32), (0, 32)])
        assert np.all(np.abs(zz - 1.0) < 1e-12), zz
    finally:
        if d32:
            abi.sv_free(d32)
        abi.sv_free(d64)


def test_34_qubits_f32_roundtrip():
    """34 qubits f32 (128 GiB), depth 12, k <= 6 on the matrix cores: the
    fused circuit and its adjoint return |0...0> (norm and <Z_a Z_b> within
    1e-5 on the device)."""
    n = 34
    if free_gib() < 136:
        pytest.skip(f"needs 128 GiB of device memory, {free_gib():.0f} GiB free")
    gates = fused_gates(n, 12)
    print(f"[large] 34q f32: {len(gates)} fused gates", flush=True)
    d = abi.sv_alloc(n, abi.F32)
    try:
        abi.sv_zero(d, n, abi.F32)
        with Plans(gates, abi.F32) as fwd:
            assert abi.VARIANT_MFMA in {abi.plan_family(p, n) for p in fwd.ids}
            fwd.apply(d, n)
            abi.stream_sync(None)
            assert abs(abi.norm_squared(d, n, abi.F32) - 1.0) < 1e-5
            # a random-layer state is far from |0...0>
            assert np.all(np.abs(abi.expect_zz(d, n, abi.F32, [(0, 1), (32, 33)])) < 0.5)
        with Plans(adjoints(gates), abi.F32) as inv:
            inv.apply(d, n)
            abi.stream_sync(None)
        print("[large] 34q f32 roundtrip done", flush=True)
        assert abs(abi.norm_squared(d, n, abi.F32) - 1.0) < 1e-5
        zz = abi.expect_zz(d, n, abi.F32, PAIRS + [(32, 33), (0, 33)])
        assert np.all(np.abs(zz - 1.0) < 1e-5), zz
    finally:
        abi.sv_free(d)


K6_SHAPES = [[0, 1, 2, 3, 4, 5], [2, 9, 14, 17, 20, 23], [25, 26, 27, 28, 29, 30]]


@pytest.mark.parametrize("n,f32", [(31, True), (32, True), (31, False)],
                         ids=["31q-f32", "32q-f32", "31q-f64"])
def test_k6_matrix_cores_above_30_qubits_match_direct_kernel(n, f32):
    """k = 6 gates (low, mixed and high targets) forced onto the matrix cores
    and onto the direct kernel from the same product state: device fidelity
    of the two results >= 1 - 1e-9 (f32) / 1 - 1e-13 (f64); then the
    matrix-core state runs the adjoints and returns to |0...0>.  Above 30
    qubits the k = 6 grid holds up to 32 workgroups per CU."""
    p = abi.F32 if f32 else abi.F64
    tol = 1e-5 if f32 else 1e-12
    layer = [(np.asarray(cs.QuantumGate.random_unitary([q], 300 + q).matrix()), [q])
             for q in range(n)]
    k6 = [(np.asarray(cs.QuantumGate.random_unitary(q, 700 + i).matrix()), q)
          for i, q in enumerate(K6_SHAPES)]
    a = abi.sv_alloc(n, p)
    b = 0
    try:
        b = abi.sv_alloc(n, p)
        with Plans(layer, p) as lp, Plans(k6, p, abi.VARIANT_MFMA) as mf, \
                Plans(k6, p, abi.VARIANT_DENSE) as dn:
            assert all(abi.plan_family(x, n) == abi.VARIANT_MFMA for x in mf.ids)
            for d in (a, b):
                abi.sv_zero(d, n, p)
                lp.apply(d, n)
            mf.apply(a, n)
            dn.apply(b, n)
            abi.stream_sync(None)
            print(f"[large] k=6 at {n}q f32={f32}: families applied", flush=True)
            fid = fidelity(a, p, b, p, n)
            assert fid >= 1.0 - (1e-9 if f32 else 1e-13), fid
            with Plans(adjoints(layer + k6), p, abi.VARIANT_MFMA) as inv:
                inv.apply(a, n)
                abi.stream_sync(None)
        assert abs(abi.norm_squared(a, n, p) - 1.0) < tol
        zz = abi.expect_zz(a, n, p, [(0, 1), (5, 6), (29, 30), (0, n - 1)])
        assert np.all(np.abs(zz - 1.0) < tol), zz
    finally:
        if b:
            abi.sv_free(b)
        abi.sv_free(a)
