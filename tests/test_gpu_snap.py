"""Matrix scalar classification at its thresholds (src/cpp/cuda/cuda_gen.cpp:67-112).

A plan snaps each real and imaginary scalar v of a K x K matrix: |v| < ztol/K
becomes exactly 0 (tested first), |v - 1| < otol/K exactly +1, |v + 1| <
otol/K exactly -1; every other value is kept (rounded through float for an
f32 plan).  Entries are placed at 0.5x and 2x of each threshold, with tolerances
large enough (1e-3) that a wrong decision moves the applied result far
beyond the parity tolerance.  The snapped matrix is written out by hand, the
plan's nonzero count must match it, and the applied result must match the
oracle run on the hand-snapped matrix and NOT the one run on the raw matrix.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from cast_simulator_amd import abi
from helpers import assert_amps_close, seeded_state

pytestmark = pytest.mark.gpu

TOL = 1e-3
K = 4
Z_IN, Z_OUT = 0.5 * TOL / K, 2.0 * TOL / K  # 1.25e-4 (snapped), 5e-4 (kept)
O_IN, O_OUT = 0.5 * TOL / K, 2.0 * TOL / K


def matrices():
    rng = np.random.default_rng(515)
    # ordinary entries, far from 0 and from +-1 in both parts
    raw = (rng.uniform(0.2, 0.7, (K, K)) * rng.choice([-1, 1], (K, K))
           + 1j * rng.uniform(0.2, 0.7, (K, K)) * rng.choice([-1, 1], (K, K)))
    snapped = raw.copy()

    def put(r, c, v, s):
        raw[r, c], snapped[r, c] = v, s

    put(0, 0, (1 + O_IN) + 1j * Z_OUT, 1 + 1j * Z_OUT)          # re -> +1, im kept
    put(0, 1, Z_IN + 1j * Z_OUT, 0 + 1j * Z_OUT)                # re -> 0
    put(1, 1, (-1 - O_IN) - 1j * Z_IN, -1 + 0j)                 # re -> -1, im -> 0
    put(1, 2, (1 - O_OUT) + 1j * (1 + O_IN), (1 - O_OUT) + 1j)  # re kept, im -> +1
    put(2, 3, (-1 + O_OUT) - 1j * Z_OUT, (-1 + O_OUT) - 1j * Z_OUT)  # both kept
    put(3, 2, -Z_IN + 1j * (-1 + O_IN), 0 - 1j)                # re -> 0, im -> -1
    put(3, 3, O_IN + 1j * Z_IN, 0j)                             # 1.25e-4 is a zero first
    put(2, 0, Z_OUT - 1j * Z_OUT, Z_OUT - 1j * Z_OUT)           # kept
    return raw, snapped


def nnz(m):
    return int(np.count_nonzero(m.real) + np.count_nonzero(m.imag))


@pytest.mark.parametrize("f32", [False, True], ids=["f64", "f32"])
def test_scalars_at_half_and_twice_the_thresholds(f32):
    raw, snapped = matrices()
    p = abi.F32 if f32 else abi.F64
    qubits, n = [1, 3], 8
    plan = abi.plan_create(raw, qubits, p, ztol=TOL, otol=TOL)
    d = abi.sv_alloc(n, p)
    try:
        info = abi.plan_info(plan)
        assert info.scalar_nnz == nnz(snapped) == 2 * K * K - 5
        assert info.variant == abi.VARIANT_DENSE
        init = seeded_state(n)
        abi.sv_upload(d, init, p)
        abi.gate_apply(plan, d, n)
        abi.stream_sync(None)
        got = abi.sv_download(d, n, p)
    finally:
        abi.sv_free(d)
        abi.plan_destroy(plan)
    tol = 1e-5 if f32 else 1e-10
    x = init.astype(np.float32) if f32 else init.copy()
    want = oracle.apply_gate(x.copy(), snapped, qubits).astype(np.float64)
    assert_amps_close(want, got, tol, "snapped")
    unsnapped = oracle.apply_gate(x.copy(), raw, qubits).astype(np.float64)
    # The test can tell them apart: a wrong decision moves the result by about
    # 1e-4 relative (2e-5 absolute here), far above the rounding difference
    # between the device and the oracle on the snapped matrix at f64 or f32.
    err_snap = float(np.max(np.abs(want - got)))
    err_raw = float(np.max(np.abs(unsnapped - got)))
    assert err_raw > 1e-5 and err_raw > 100 * max(err_snap, 1e-9), (err_raw, err_snap)


@pytest.mark.parametrize("f32", [False, True], ids=["f64", "f32"])
def test_near_identity_snaps_to_identity_and_moves_nothing(f32):
    """Identity plus off-diagonal entries at 0.5 ztol/K and diagonal entries at
    1 +- 0.5 otol/K snaps to the exact identity: a monomial plan with no
    active rows, whose apply leaves the statevector bit-identical; at 2x the
    thresholds the plan is dense and the state changes."""
    p = abi.F32 if f32 else abi.F64
    n, qubits = 6, [0, 4]
    near = np.eye(K, dtype=complex) + Z_IN * (np.ones((K, K)) - np.eye(K)) * (1 - 1j)
    near[np.diag_indices(K)] = [1 + O_IN, 1 - O_IN, 1 + 1j * Z_IN, 1 - O_IN - 1j * Z_IN]
    far = np.eye(K, dtype=complex) + Z_OUT * (np.ones((K, K)) - np.eye(K))
    init = seeded_state(n)
    for m, identity in ((near, True), (far, False)):
        plan = abi.plan_create(m, qubits, p, ztol=TOL, otol=TOL)
        d = abi.sv_alloc(n, p)
        try:
            info = abi.plan_info(plan)
            abi.sv_upload(d, init, p)
            stored = abi.sv_download(d, n, p)
            abi.gate_apply(plan, d, n)
            abi.stream_sync(None)
            got = abi.sv_download(d, n, p)
        finally:
            abi.sv_free(d)
            abi.plan_destroy(plan)
        if identity:
            assert info.variant == abi.VARIANT_MONOMIAL and info.nnz_rows == 0
            assert info.scalar_nnz == K
            assert np.array_equal(got, stored)
        else:
            assert info.variant == abi.VARIANT_DENSE
            assert info.scalar_nnz == K + K * (K - 1)
            assert np.max(np.abs(got - stored)) > 1e-5
