"""Shared test helpers: reference fixtures and an independent brute-force
statevector propagator used to pin the oracle."""
from __future__ import annotations

import numpy as np

S2 = 1.0 / np.sqrt(2.0)


def seeded_state(n: int) -> np.ndarray:
    """The reference's deterministic ramp state ((i+1), 0.5 i - 0.25) / norm,
    returned interleaved (tests/cpu_cuda_compare.rs:54-69)."""
    i = np.arange(1 << n, dtype=np.float64)
    re, im = i + 1.0, 0.5 * i - 0.25
    nrm = np.sqrt(np.sum(re * re + im * im))
    out = np.empty(2 << n)
    out[0::2], out[1::2] = re / nrm, im / nrm
    return out


def to_complex(inter: np.ndarray) -> np.ndarray:
    x = np.asarray(inter, dtype=np.float64)
    return x[0::2] + 1j * x[1::2]


def to_interleaved(z: np.ndarray) -> np.ndarray:
    out = np.empty(2 * z.size)
    out[0::2], out[1::2] = z.real, z.imag
    return out


def local_index(x: np.ndarray, qubits) -> np.ndarray:
    """Gate-local basis index of global indices x: bit b <- bit qubits[b]."""
    loc = np.zeros_like(x)
    for b, q in enumerate(qubits):
        loc |= ((x >> q) & 1) << b
    return loc


def full_operator(matrix: np.ndarray, qubits, n: int) -> np.ndarray:
    """Dense 2^n x 2^n operator of a k-qubit gate, built entry by entry from
    the definition U[i, j] = M[loc(i), loc(j)] if i and j agree on every
    non-target bit, else 0 (no bit-deposit, no strides)."""
    N = 1 << n
    idx = np.arange(N, dtype=np.int64)
    tmask = 0
    for q in qubits:
        tmask |= 1 << q
    li = local_index(idx, qubits)
    same = ((idx[:, None] ^ idx[None, :]) & ~tmask) == 0
    U = np.where(same, matrix[li[:, None], li[None, :]], 0.0)
    return U


def brute_force_apply(z: np.ndarray, gates, n: int) -> np.ndarray:
    for m, q in gates:
        z = full_operator(np.asarray(m), q, n) @ z
    return z


def assert_amps_close(want, got, tol: float, label: str = "") -> None:
    """Per-amplitude |delta| < tol (tests/cpu_cuda_compare.rs:113-134)."""
    a, b = to_complex(want), to_complex(got)
    assert a.shape == b.shape, f"{label}: {a.shape} vs {b.shape}"
    d = np.abs(a - b)
    i = int(np.argmax(d))
    assert d[i] < tol, f"{label}: amp[{i}] |delta|={d[i]:.3e} > tol={tol:.1e}"


def rel_l2(want, got) -> float:
    a, b = to_complex(want), to_complex(got)
    return float(np.linalg.norm(a - b) / np.linalg.norm(a))


def fidelity(a, b) -> float:
    za, zb = to_complex(a), to_complex(b)
    ov = np.vdot(za, zb)
    return float(abs(ov) ** 2 / (np.vdot(za, za).real * np.vdot(zb, zb).real))


def gate_list(gates):
    """[(matrix ndarray, qubits list)] from host QuantumGate objects."""
    return [(g.matrix(), list(g.qubits())) for g in gates]
