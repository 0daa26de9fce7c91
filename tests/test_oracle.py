"""Pins the oracle (oracle/) before it is trusted as the parity checker.

Each pin is chosen so that a plausible slip in the oracle — a dropped or
transposed matrix term, a wrong bit order, a sign error, an f32 path that
skips the per-gate rounding — fails at least one of:
  * the reference's own known answers (src/cuda/tests.rs:79-187,
    src/simulator/tests.rs:320-339, tests/cpu_cuda_compare.rs gates);
  * an independent brute-force propagator (dense 2^n x 2^n operator built
    entry-by-entry from the definition, no bit-deposit);
  * closed forms (QFT == unitary DFT; |0..0> correlations; norm invariance).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import (
    S2,
    brute_force_apply,
    full_operator,
    seeded_state,
    to_complex,
    to_interleaved,
)

H = np.array([[S2, S2], [S2, -S2]], dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
CX = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]], dtype=complex)


def rand_unitary(k: int, rng: np.random.Generator) -> np.ndarray:
    z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def zero_state(n: int) -> np.ndarray:
    v = np.zeros(2 << n)
    v[0] = 1.0
    return v


# ── pdep (src/cpu/statevector.rs:592-606) ──────────────────────────────────

@pytest.mark.parametrize("src,mask,want", [
    (0b1, 0b100, 0b100),
    (0b11, 0b1010, 0b1010),
    (0b101, 0b111000, 0b101000),
    (0b11001, 0b1011_0110, 0b1010_0010),
    # combo 0b11001 with targets [2, 4, 5] (the setting of cuda_gen.cpp:185-206):
    # deposit into the non-target positions 0, 1, 3, 6, 7
    (0b11001, 0b1100_1011, 0b1100_0001),
    (0, 0xFF, 0),
    (0xF, 0, 0),
])
def test_pdep_known_values(src, mask, want):
    assert oracle.pdep(src, mask) == want


def test_pdep_matches_definition():
    rng = np.random.default_rng(5)
    for _ in range(200):
        mask = int(rng.integers(0, 1 << 20))
        src = int(rng.integers(0, 1 << 20))
        out, b = 0, 0
        for pos in range(20):
            if mask >> pos & 1:
                out |= (src >> b & 1) << pos
                b += 1
        assert oracle.pdep(src, mask) == out


# ── known answers (src/cuda/tests.rs:79-187) ───────────────────────────────

def test_h_on_zero_state():
    sv = oracle.apply_gate(zero_state(1), H, [0])
    assert np.allclose(sv, [S2, 0, S2, 0], atol=1e-14)


def test_x_on_zero_state():
    sv = oracle.apply_gate(zero_state(1), X, [0])
    assert np.allclose(sv, [0, 0, 1, 0], atol=1e-14)


def test_h_on_larger_statevector():
    sv = oracle.apply_gate(zero_state(4), H, [0])
    want = np.zeros(32)
    want[0] = want[2] = S2
    assert np.allclose(sv, want, atol=1e-14)


def test_x_then_h_is_minus_state():
    sv = oracle.apply_gate(oracle.apply_gate(zero_state(1), X, [0]), H, [0])
    assert np.allclose(sv, [S2, 0, -S2, 0], atol=1e-14)


def test_h_squared_is_identity_on_ramp():
    init = seeded_state(3)
    sv = oracle.apply_gate(oracle.apply_gate(init.copy(), H, [0]), H, [0])
    assert np.allclose(sv, init, atol=1e-14)


# ── brute force on the reference's compare-test gate shapes ───────────────

SHAPES = [
    ([0], 2), ([0], 8), ([3], 6), ([0, 1], 5), ([0, 4], 6), ([1, 3], 5), ([1, 4], 6),
    ([0, 1, 2], 5), ([0, 2, 5], 6), ([0, 1, 2, 3], 6), ([2, 5, 7], 8), ([0, 3, 4, 8], 9),
    ([1, 2, 4, 6, 7], 8), ([0, 1, 2, 3, 4, 5], 7), ([3, 4, 5, 6, 7, 8], 9),
]


@pytest.mark.parametrize("qubits,n", SHAPES)
def test_scalar_apply_matches_brute_force(qubits, n):
    rng = np.random.default_rng(1000 + 31 * n + len(qubits))
    m = rand_unitary(len(qubits), rng)
    init = seeded_state(n)
    got = oracle.apply_gate(init.copy(), m, qubits)
    want = full_operator(m, qubits, n) @ to_complex(init)
    assert np.max(np.abs(to_complex(got) - want)) < 1e-13


def test_scalar_apply_nonunitary_sparse_matches_brute_force():
    # A transposed or conjugated matrix would pass unitarity-based checks on
    # some inputs; a random non-unitary complex matrix does not.
    rng = np.random.default_rng(77)
    m = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    m[rng.random((8, 8)) < 0.5] = 0
    init = seeded_state(7)
    got = oracle.apply_gate(init.copy(), m, [1, 3, 6])
    want = full_operator(m, [1, 3, 6], 7) @ to_complex(init)
    assert np.max(np.abs(to_complex(got) - want)) < 1e-12


def test_sequence_matches_brute_force():
    rng = np.random.default_rng(9)
    n = 7
    gates = []
    for _ in range(12):
        k = int(rng.integers(1, 4))
        q = sorted(rng.choice(n, size=k, replace=False).tolist())
        gates.append((rand_unitary(k, rng), q))
    init = seeded_state(n)
    got = oracle.apply_gates(init.copy(), gates)
    want = brute_force_apply(to_complex(init), gates, n)
    assert np.max(np.abs(to_complex(got) - want)) < 1e-13


def test_f32_storage_rounds_once_per_gate():
    rng = np.random.default_rng(3)
    m = rand_unitary(2, rng)
    init = seeded_state(6)
    f64 = oracle.apply_gate(init.copy(), m, [1, 4])
    f32 = oracle.apply_gate(init.astype(np.float32), m, [1, 4])
    # f64 math on f32-rounded inputs, one rounding on output: well inside
    # a few f32 ulps of the f64 result.
    assert np.max(np.abs(f32.astype(np.float64) - f64)) < 4e-7
    assert f32.dtype == np.float32


# ── QFT closed form (src/types/circuit.rs:151-171; circuit tests :201-306) ──

def qft_gates(n: int, swap: bool = True):
    gates = []
    for i in range(n):
        gates.append((H, [i]))
        for j in range(i + 1, n):
            th = np.pi / (1 << (j - i))
            cp = np.diag([1, 1, 1, np.exp(1j * th)])
            gates.append((cp, [i, j]))  # cp is symmetric in its two qubits
    if swap:
        sw = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)
        for i in range(n // 2):
            gates.append((sw, [i, n - 1 - i]))
    return gates


def bitrev(n: int) -> np.ndarray:
    return np.array([int(format(j, f"0{n}b")[::-1], 2) for j in range(1 << n)])


@pytest.mark.parametrize("n", [3, 6, 10])
def test_qft_equals_unitary_dft(n):
    # The circuit starts its H/CP ladder at qubit 0, i.e. qubit 0 is the most
    # significant bit of the textbook QFT.  In index terms the swap variant is
    # the unitary DFT  y_k = sum_j e^{+2 pi i jk/N} x_j / sqrt(N)  conjugated by
    # the bit reversal R of the n-bit index: QFT = R F R.
    x = to_complex(seeded_state(n))
    got = to_complex(oracle.apply_gates(to_interleaved(x), qft_gates(n)))
    r = bitrev(n)
    want = (np.fft.ifft(x[r]) * np.sqrt(1 << n))[r]
    assert np.max(np.abs(got - want)) < 1e-12
    # The un-conjugated DFT is a different operator: a bit-order slip fails.
    assert np.max(np.abs(got - np.fft.ifft(x) * np.sqrt(1 << n))) > 1e-3


def test_qft_on_zero_is_uniform():
    n = 8
    got = to_complex(oracle.apply_gates(zero_state(n), qft_gates(n)))
    assert np.allclose(np.abs(got) ** 2, 1.0 / (1 << n), atol=1e-12)


def test_qft_swap_is_bit_reversal_of_noswap():
    n = 6
    one = zero_state(n)
    one[0], one[2] = 0.0, 1.0  # |1>
    a = to_complex(oracle.apply_gates(one.copy(), qft_gates(n, True)))
    b = to_complex(oracle.apply_gates(one.copy(), qft_gates(n, False)))
    rev = [int(format(j, f"0{n}b")[::-1], 2) for j in range(1 << n)]
    assert np.max(np.abs(a - b[rev])) < 1e-12


# ── norm / scale (src/simulator/tests.rs:320-339) ──────────────────────────

def test_norm_known_answer_and_normalize():
    v = np.array([3.0, 4.0])
    assert oracle.norm_squared(v) == 25.0
    oracle.scale(v, 1.0 / np.sqrt(25.0))
    assert abs(oracle.norm_squared(v) - 1.0) < 1e-15
    v32 = np.array([3.0, 4.0], dtype=np.float32)
    assert oracle.norm_squared(v32) == 25.0


def test_norm_of_zero_state_is_one():
    assert oracle.norm_squared(zero_state(5)) == 1.0


# ── <Z_a Z_b> ─────────────────────────────────────────────────────────────

def test_expect_zz_closed_forms():
    n = 5
    assert np.allclose(oracle.expect_zz(zero_state(n), [(0, 1), (2, 4)]), [1, 1])
    # X on qubit 1: |00010> -> Z0Z1 = -1, Z1Z2 = -1, Z2Z3 = +1
    sv = oracle.apply_gate(zero_state(n), X, [1])
    assert np.allclose(oracle.expect_zz(sv, [(0, 1), (1, 2), (2, 3)]), [-1, -1, 1])
    # Bell pair (H then CX): <Z0 Z1> = 1, and H on one side of |0>: <Z0 Z1> = 0
    bell = oracle.apply_gate(oracle.apply_gate(zero_state(2), H, [0]), CX, [0, 1])
    assert np.allclose(oracle.expect_zz(bell, [(0, 1)]), [1.0])
    plus = oracle.apply_gate(zero_state(2), H, [0])
    assert np.allclose(oracle.expect_zz(plus, [(0, 1)]), [0.0], atol=1e-15)


def test_expect_zz_matches_dense_operator():
    rng = np.random.default_rng(2)
    n = 6
    z = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    z /= np.linalg.norm(z)
    idx = np.arange(1 << n)
    pairs = [(i, i + 1) for i in range(n - 1)] + [(0, 5)]
    want = [np.sum(np.abs(z) ** 2 * (1 - 2 * (((idx >> a) ^ (idx >> b)) & 1))) for a, b in pairs]
    assert np.allclose(oracle.expect_zz(to_interleaved(z), pairs), want, atol=1e-14)


# ── SIMD split-layout baseline == scalar oracle (src/cpu/tests.rs:895-1357
#    layout coverage: lo-only, hi-only, mixed, straddling) ─────────────────

SIMD_SHAPES = [
    [0], [5], [0, 1], [3, 4], [0, 1, 2, 3, 4], [0, 1, 2, 3, 5], [0, 1, 3, 4, 6], [4, 5, 6, 7],
    [1, 6], [2, 3, 9], [0, 7, 8, 10, 11], [5, 6, 7, 8, 9, 10], [0, 1, 2, 3, 4, 5],
]


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("qubits", SIMD_SHAPES)
def test_simd_baseline_matches_scalar(qubits, f32):
    n = 12
    rng = np.random.default_rng(sum(qubits) + 17 * len(qubits))
    m = rand_unitary(len(qubits), rng)
    init = seeded_state(n)
    want = oracle.apply_gate(init.copy(), m, qubits)
    sv = oracle.SimdStatevector(n, f32=f32)
    sv.load(init)
    sv.apply(m, qubits, n_threads=3)
    got = sv.interleaved()
    tol = 5e-6 if f32 else 1e-13
    assert np.max(np.abs(got - want)) < tol


def test_simd_pack_roundtrip():
    init = seeded_state(9)
    for f32 in (False, True):
        sv = oracle.SimdStatevector(9, f32=f32)
        sv.load(init)
        back = sv.interleaved()
        assert np.max(np.abs(back - init)) < (1e-7 if f32 else 0.0) + 1e-300


def test_simd_rejects_too_small_statevector():
    sv = oracle.SimdStatevector(6, f32=False)
    with pytest.raises(ValueError):
        sv.apply(np.eye(16), [0, 1, 2, 3], 1)  # needs n >= k + s = 7
